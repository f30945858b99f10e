/*
 * virial_lj.h - C ABI of the MI355X (gfx950) Lennard-Jones library.
 *
 * Drop-in for the host API of virial-engine's `virial_particles::lennard_jones`
 * (crates/virial-particles/src/lennard_jones.rs). Each entry point names the
 * reference item it replaces. Plain C types only: the caller owns every host
 * array and pre-sizes outputs; the library owns device memory.
 *
 * Errors: the reference panics (assert!/expect); here every call returns a
 * status (VLJ_OK = 0) and vlj_last_error() holds the message, worded as the
 * reference's panic where there is one.
 *
 * Threading: a handle is not thread-safe (the reference takes &mut self).
 * Each handle owns one HIP stream on its device; vlj_step and vlj_measure
 * only enqueue; vlj_poll never blocks; the *_now calls, vlj_run and the
 * read calls synchronise the handle's stream.
 *
 * One handle holds one or more independent periodic systems ("replicas").
 * vlj_create makes a handle with one; vlj_batch_create one with many, run by
 * the same kernels in the same launches over a replica-offset table. The
 * calls that act on one system (measure, poll, measure_now, box_size,
 * target_temperature, set_temperature, the heat exchange and heat current,
 * read_profile and read_profile_raw) return VLJ_E_STATE for a handle of
 * several replicas, which has its own vlj_batch_* calls; on such a handle
 * vlj_count is the total over the replicas and vlj_volume and vlj_time are
 * NaN.
 */
#ifndef VIRIAL_LJ_H
#define VIRIAL_LJ_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VLJ_OK 0
#define VLJ_E_INVALID 1     /* argument the reference asserts against */
#define VLJ_E_HIP 2         /* HIP runtime error (message in vlj_last_error) */
#define VLJ_E_STATE 3       /* call not valid in this state (reference panics) */
#define VLJ_E_UNSUPPORTED 4 /* feature not built */
#define VLJ_E_NO_DEVICE 5   /* no usable gfx950 device */

/* Particle (lennard_jones.rs:25-30). Mass 1, reduced units. */
typedef struct {
    double pos[3];
    double vel[3];
} vlj_particle;

#define VLJ_THERMOSTAT_NONE 0
#define VLJ_THERMOSTAT_LANGEVIN 1
#define VLJ_THERMOSTAT_NOSE_HOOVER 2

/* Thermostat (lennard_jones.rs:32-61): Langevin {temperature, friction, seed}
 * or NoseHoover {temperature, period}. The seed keys a counter-based
 * generator (Philox4x32-10), not the reference's per-particle PCG. */
typedef struct {
    uint32_t kind;
    double temperature;
    double friction;
    double period;
    uint64_t seed;
} vlj_thermostat;

#define VLJ_PRECISION_FP64 0 /* FP64 state and FP64 pair forces */
#define VLJ_PRECISION_FP32 1 /* FP32 pair forces on FP64 displacements */

/* Settings (lennard_jones.rs:63-82) plus the pair-force precision. */
typedef struct {
    double box_size[3];
    double cutoff;
    double dt;
    double skin;
    vlj_thermostat thermostat;
    int32_t shifted_force;
    uint32_t profile_bins;  /* slabs along z for the profiles (0: none), at most 2^20 */
    uint32_t precision;
} vlj_settings;

/* Measurement (lennard_jones.rs:84-131). raw[] holds the 64-bit fixed-point
 * totals the doubles were decoded from: KE, PE, W, momentum x/y/z (energy
 * or momentum scale times 2^32), condensed count, and twice the pairs within
 * the cutoff. */
typedef struct {
    uint64_t step;
    double time;
    double kinetic;
    double potential;
    double virial;
    double momentum[3];
    uint32_t count;
    double volume;
    uint32_t list_rebuilds;
    uint32_t list_overflow;
    double thermostat_energy;
    uint32_t condensed;
    uint64_t pairs_within_cutoff;
    int64_t raw[8];
} vlj_measurement;

typedef struct vlj_sim vlj_sim;
typedef struct vlj_sim vlj_batch;

/* LennardJones::new (lennard_jones.rs:443-801): copies the particles
 * (positions wrapped into the box), rounds the box to f32, derives cells,
 * list capacity, fixed-point scales and pair constants. VLJ_E_INVALID with
 * "the box must be at least two (cutoff + skin) wide on every axis" where
 * the reference asserts three (rs:458-461). No kernel runs until the first step or
 * measurement (lazy priming, rs:1080-1089). */
int vlj_create(const vlj_particle *particles, size_t count, const vlj_settings *settings,
               vlj_sim **out);
/* Drop (the reference frees its wgpu buffers on drop). */
void vlj_destroy(vlj_sim *sim);

/* LennardJones::step (rs:1092-1138): enqueue `steps` steps; returns at once. */
int vlj_step(vlj_sim *sim, uint32_t steps);
/* LennardJones::run (rs:1274-1284): advance and wait. */
int vlj_run(vlj_sim *sim, uint64_t steps);
/* LennardJones::measure (rs:1141-1163): enqueue a measurement into a free
 * slot of the 3-slot read-back ring; *accepted = 0 when all are busy. */
int vlj_measure(vlj_sim *sim, int32_t *accepted);
/* LennardJones::poll (rs:1209-1246): finished measurements, oldest first;
 * never blocks. *written <= capacity. */
int vlj_poll(vlj_sim *sim, vlj_measurement *out, size_t capacity, size_t *written);
/* LennardJones::measure_now (rs:1250-1260): VLJ_E_STATE with "a measurement
 * is already in flight" when the ring is full, as the reference panics. */
int vlj_measure_now(vlj_sim *sim, vlj_measurement *out);
/* LennardJones::read_particles (rs:1354-1374): full precision, original order. */
int vlj_read_particles(vlj_sim *sim, vlj_particle *out, size_t count);

/* Accessors and setters (rs:870-1030). count: particles of the handle. */
uint32_t vlj_count(const vlj_sim *sim);
int vlj_box_size(const vlj_sim *sim, double out[3]);
double vlj_volume(const vlj_sim *sim);
uint64_t vlj_steps(const vlj_sim *sim);
double vlj_time(const vlj_sim *sim);
/* target_temperature (rs:875-881): *has = 0 without a thermostat. */
int vlj_target_temperature(const vlj_sim *sim, double *temperature, int32_t *has);
/* set_temperature (rs:905-925): VLJ_E_STATE "set_temperature needs a
 * thermostat" without one. */
int vlj_set_temperature(vlj_sim *sim, double temperature);
/* set_condensed_neighbours (rs:895-899). */
int vlj_set_condensed_neighbours(vlj_sim *sim, uint32_t neighbours);

/* Profiles along z (lennard_jones.wgsl:628-730). Layout of the
 * profile_bins * 8 + 3 values: per slab b from b * 8 the particle count,
 * m v_a^2 for a = x, y, z, the Irving-Kirkwood virial diagonal W_aa and the
 * potential energy; then the whole-box W_aa summed pair by pair.
 *
 * vlj_sample_profile (rs:1167-1183) adds one sample to the running sums of
 * every replica with profile slabs, asynchronously; VLJ_E_STATE "profiles
 * need Settings::profile_bins above zero" when none has any.
 * vlj_clear_profile (rs:1186-1191) starts the sums afresh.
 * vlj_read_profile (rs:1297-1325) writes the sums divided by the sample
 * count (NaN before the first sample), as the reference's Profile holds
 * them; capacity must be at least profile_bins * 8 + 3.
 * vlj_read_profile_raw writes the 64-bit fixed-point sums (value * 2^32)
 * themselves, and vlj_batch_read_profile_raw those of one replica. */
int vlj_sample_profile(vlj_sim *sim);
int vlj_clear_profile(vlj_sim *sim);
int vlj_read_profile(vlj_sim *sim, double *out, size_t capacity, uint64_t *samples);
int vlj_read_profile_raw(vlj_sim *sim, int64_t *raw, size_t capacity, uint64_t *samples);
int vlj_batch_read_profile_raw(vlj_batch *batch, size_t replica, int64_t *raw, size_t capacity,
                               uint64_t *samples);

/* Heat exchange, eHEX/a (lennard_jones.rs:927-943, lennard_jones.wgsl:732-870):
 * after every step the slab hot[0] <= z < hot[1] gains `energy` and the
 * slab cold[0] <= z < cold[1] loses it, by scaling velocities about each
 * slab's centre-of-mass velocity with the O(dt^3) position correction of
 * Wirnsberger, Frenkel and Dellago (2015). on = 0 stops it. VLJ_E_STATE
 * "heat exchange runs at constant energy, without a thermostat" with one. */
int vlj_set_heat_exchange(vlj_sim *sim, int32_t on, const double hot[2], const double cold[2],
                          double energy);
/* heat_flux (rs:945-951): energy / (2 Lx Ly dt); *has = 0 when off. */
int vlj_heat_flux(const vlj_sim *sim, double *flux, int32_t *has);
/* Heat-current series for Green-Kubo (rs:953-1012, lennard_jones.wgsl:872-947):
 * J = sum_i e_i v_i + 1/2 sum_ij (F_ij . v_i) r_ij. record makes room for
 * `capacity` samples (dropping any); sample adds one (VLJ_E_STATE "record_heat_current
 * first", "heat current series full"); read writes 3 values (x, y, z) per
 * sample, capacity counted in samples (out NULL: only the number of
 * samples); read_raw the fixed-point sums (value * 2^32). */
int vlj_record_heat_current(vlj_sim *sim, uint32_t capacity);
int vlj_sample_heat_current(vlj_sim *sim);
int vlj_read_heat_current(vlj_sim *sim, double *out, size_t capacity, uint32_t *samples);
int vlj_read_heat_current_raw(vlj_sim *sim, int64_t *out, size_t capacity, uint32_t *samples);

/* Last error message of the handle (or of the last failed vlj_create when
 * sim is NULL). Never NULL. */
const char *vlj_last_error(const vlj_sim *sim);

/* ---- batched replicas (no counterpart in the reference) ----
 * `offsets[r]` .. `offsets[r + 1]` index replica r's particles in `all`
 * (replicas + 1 entries); `settings[r]` are its settings. All replicas share
 * the handle's launches: one kernel per phase over a replica-offset table. */
int vlj_batch_create(const vlj_particle *all, const uint64_t *offsets,
                     const vlj_settings *settings, size_t replicas, vlj_batch **out);
int vlj_batch_run(vlj_batch *batch, uint64_t steps);
/* set_temperature (rs:905-925) of one replica: VLJ_E_STATE "set_temperature
 * needs a thermostat" when it has none, VLJ_E_INVALID for no such replica. */
int vlj_batch_set_temperature(vlj_batch *batch, size_t replica, double temperature);
/* One measurement per replica, in replica order. */
int vlj_batch_measure_now(vlj_batch *batch, vlj_measurement *per_replica, size_t replicas);

/* ---- Renderer bindings: LennardJones::positions / velocities / set_drawn
 * (rs:1041-1051, :885-889; lennard_jones.wgsl:596-603). */
/* Device pointers to the render arrays, `count` float4 each in slot order
 * (the device's order of the same particles): the position wrapped into the
 * box (w 0), and the velocity with in w the value chosen by vlj_set_drawn.
 * Allocated on the first call, valid until the handle is destroyed; filled
 * by vlj_render and by every measurement. */
int vlj_render_buffers(vlj_sim *sim, void **positions, void **velocities);
/* Fills the render arrays from the current state, ordered after the steps
 * enqueued so far (asynchronous). */
int vlj_render(vlj_sim *sim);
/* vlj_render, then copies both arrays (count float4 each) to host memory. */
int vlj_read_render(vlj_sim *sim, float *positions, float *velocities, size_t count);
/* set_drawn: 0 neighbours within 1.5 sigma (the default), 1 local
 * temperature, m v^2 / 3 averaged over the molecule and its neighbours
 * within the cutoff. VLJ_E_INVALID for another value. */
int vlj_set_drawn(vlj_sim *sim, uint32_t drawn);

/* ---- Checkpoint and exact resume (no counterpart in the reference, whose
 * runs start again from particles; the on-disk formats of its app crate,
 * crates/virial/src/data.rs, hold measurements and snapshots only). */
/* Writes the complete state of a handle (single or batch) after the steps
 * enqueued so far: positions, velocities, accelerations, positions at the
 * last list build, original indices, cell tables, lists and bricks, the
 * control words (step counter of the noise, rebuild count, Nose-Hoover xi
 * and eta), the parameters as last set, profile sums and the heat-current
 * series. Measurements in flight are not kept. */
int vlj_save(vlj_sim *sim, const char *path);
/* A new handle from a file written by vlj_save; stepping it gives the same
 * bits as stepping the handle that saved it. VLJ_E_INVALID for a file that
 * is not a state file, is truncated, or was written by another version. */
int vlj_load(const char *path, vlj_sim **out);
/* The number of replicas of a handle and, for replica `replica`, the
 * settings it runs with (thermostat temperature as last set) and its
 * particle count. Any output may be NULL. */
int vlj_replica_info(const vlj_sim *sim, size_t replica, vlj_settings *settings, uint64_t *count,
                     size_t *replicas);
size_t vlj_batch_replicas(const vlj_batch *batch);
/* Particles of every replica, concatenated as given to vlj_batch_create. */
int vlj_batch_read_particles(vlj_batch *batch, vlj_particle *out, size_t count);

/* ---- brick planner and its tuner (no counterpart in the reference) ----
 * The kernels stage each brick's neighbourhood (gx x gy x gz home cells of
 * one or two z layers and the cells around them) in LDS. Brick shapes are
 * planned per replica and z layer from the cell occupancy, the largest whose
 * every brick stages at most `fill` of the LDS staging limit (a window's
 * count plus four of its Poisson standard deviations); two layers share
 * bricks when that stages fewer cells per home cell.
 * The layout (workgroup size) follows from the staged count. Plans never change
 * results: lists are built in the same order whatever the bricks.
 *
 * vlj_tune measures: for each candidate fill (0.8, 0.9, 1.0) it plans
 * every replica's bricks, rebuilds the lists from the positions of their
 * last build for that plan, runs `steps_per_candidate` real steps timing
 * the pair-force launches with HIP events, and keeps the plan with the
 * least force time plus list-build time per step. It blocks; *steps_taken
 * gets the steps it ran (they count as steps of the run). A handle of at
 * least 65,536 atoms tunes itself (6 steps per candidate) on the first
 * vlj_step or vlj_run that asks for at least 64 steps. Every 256 steps the
 * control words are copied back without a wait; a replica whose bricks
 * were cut into parts since the check before is planned again from the
 * occupancy of its last list build (its density rose since), held to 10%
 * less of the staging limit each time (down to 60%), with its lists rebuilt
 * for the new bricks. VLJ_E_STATE when vlj_debug_plan fixes the plan.
 * Trials time the state they run on: a handle started from a lattice that
 * melts is best tuned again once it has (FCC 4 x 64^3: fills 0.9 and 1.0
 * tie on the lattice, and 1.0 runs the liquid's force launch ~5% slower). */
typedef struct {
    double fill;     /* the candidate's fill */
    double force_us; /* mean pair-force time per step */
    double build_us; /* list build over the whole brick table */
    uint32_t bricks; /* bricks in the table */
} vlj_tune_trial;
int vlj_tune(vlj_sim *sim, uint32_t steps_per_candidate, uint64_t *steps_taken);
/* The planner's state: the fill of the plan in use, the re-plans since the
 * handle was created (tuning aside), and the trials of the last tuning
 * (*n_trials of them; up to `capacity` written). */
int vlj_plan_info(const vlj_sim *sim, double *fill, uint32_t *replans, vlj_tune_trial *trials,
                  size_t capacity, size_t *n_trials);

/* ---- lattices (rs:1377-1441), host only ---- */
/* lattice: `count` particles on a simple cubic lattice filling a cube. */
int vlj_lattice(size_t count, double box_size, double temperature, uint64_t seed,
                vlj_particle *out);
/* lattice_box: the same in a rectangular box. */
int vlj_lattice_box(size_t count, const double size[3], double temperature, uint64_t seed,
                    vlj_particle *out);
/* FCC, 4 n^3 particles at `density` (no counterpart in the reference).
 * *side = the cube's side, already rounded to f32. */
int vlj_fcc(uint32_t cells_per_side, double density, double temperature, uint64_t seed,
            vlj_particle *out, double *side);
/* FCC sites for any count (no counterpart in the reference): the smallest
 * 4 n^3 lattice holding `count` atoms, filled in (cz, cy, cx, basis) order,
 * in a cube of side cbrt(count / density) rounded to f32 (written to *side). */
int vlj_fcc_count(size_t count, double density, double temperature, uint64_t seed, vlj_particle *out,
                  double *side);
/* Random non-overlapping insertion (no counterpart in the reference):
 * uniform trials from the xorshift stream, rejected within min_distance
 * (nearest image) of an earlier particle, 1000 trials per particle at most
 * (VLJ_E_INVALID when they do not suffice); Gaussian velocities from the
 * same stream at `temperature`, mean removed. */
int vlj_random_insertion(size_t count, const double box[3], double min_distance, double temperature,
                         uint64_t seed, vlj_particle *out);

/* ---- fixed point (crates/virial-precision/src/fixed.rs:9-57) ----
 * Device reduction used by every measurement: sum of round-ties-even(values[i]
 * * fixed_scale) as a 64-bit integer, by per-workgroup pre-reduction and one
 * 64-bit atomic per workgroup (fixed.wgsl:71-105; KAT tests/gpu.rs:89-149). */
int vlj_fixed_sum(const double *values, size_t count, double fixed_scale, int64_t *out);

/* ---- fixed-point sum without allocation (no counterpart in the reference) ----
 * For callers that cannot allocate device memory: code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc. The same
 * sum as vlj_fixed_sum, but `values` (count doubles) and `out` (one int64)
 * are device pointers the caller owns, `stream` is a hipStream_t (NULL: the
 * default stream), and the call only enqueues: *out is written when the
 * stream gets there. It allocates nothing: its partial sums (one per
 * workgroup, at most 1024) and a ticket live in a device array of about
 * 8.5 KB inside the library, one copy per GPU, held until the process ends.
 * At most 2^36 values (VLJ_E_INVALID beyond). Calls that overlap on one GPU
 * (from other threads or streams) race on that array. */
int vlj_fixed_sum_device(const double *values, size_t count, double fixed_scale, int64_t *out,
                         void *stream);

/* ---- inspection, for parity tests (no counterpart in the reference) ----
 * State in slot order (the device keeps particles in cell order). Each
 * pointer may be NULL. */
int vlj_debug_state(vlj_sim *sim, double *pos, double *vel, double *acc, uint32_t *orig);
/* Cell grid of replica 0 after the last rebuild: cell of each slot,
 * total_cells + 1 starts (relative to the replica's first slot), slots in
 * cell order. */
int vlj_debug_cells(vlj_sim *sim, uint32_t *cell_of, uint32_t *cell_start, uint32_t *sorted);
/* Neighbour lists of replica 0: counts[n] and entries[n * capacity], row per
 * slot, relative slot indices. */
int vlj_debug_lists(vlj_sim *sim, uint32_t *counts, uint32_t *entries);
/* Derived integers of replica 0: cells[3], total_cells, capacity, rebuilds. */
int vlj_debug_grid(const vlj_sim *sim, uint32_t out[6]);
/* Totals over the handle's replicas: list rebuilds, brick list builds that
 * cut the brick into parts (its neighbourhood outgrew LDS), the largest list
 * overflow, the number of bricks, list builds of single home cells that
 * read global memory (not even one cell's neighbourhood fits), and the
 * number of bricks with the sparse layout. No counterpart in the reference. */
int vlj_debug_counters(const vlj_sim *sim, uint64_t out[6]);
/* The same for one replica: list rebuilds, brick list builds that cut the
 * brick into parts, list overflow, global-memory single-cell builds. */
int vlj_debug_replica_counters(const vlj_sim *sim, size_t replica, uint64_t out[4]);
/* Derived reals of replica 0: energy_scale, momentum_scale, shift,
 * force_shift, c1, c2. */
int vlj_debug_params(const vlj_sim *sim, double out[6]);

/* Overrides of the planning rules, for tests (no counterpart in the
 * reference); each value 0 keeps its rule. list_capacity: list entries per
 * atom (at least 8) in place of the capacity rule, so that lists overflow.
 * stage_fill: dense bricks are planned to stage at most this share of the
 * LDS staging limit (in place of 0.9; above 1 plans bricks that are cut
 * into parts). sparse_fill: a z layer whose bricks stage at most this share
 * of the sparse layout's limit takes that layout (in place of 0.9; negative:
 * every layer dense). They apply to handles created (or loaded) afterwards,
 * in every thread of the process. */
int vlj_debug_plan(uint32_t list_capacity, double stage_fill, double sparse_fill);

/* Whether handles created afterwards run the PARTS launches, for tests (no
 * counterpart in the reference): -1 (the rule) when their first list build
 * would cut a brick, else only once the periodic check sees a brick outgrow
 * LDS; 0 not from the start (a brick that outgrows LDS runs whole, reading
 * positions through the cell tables); 1 from the start. */
int vlj_debug_parts(int32_t mode);

/* Workgroup-to-table mapping (no counterpart in the reference). With
 * blocks != 0: the table entry workgroup `block` works on -- workgroups go
 * round-robin to the 8 XCDs, and the table is dealt to them in chunks of 4
 * consecutive entries, chunk c to XCD c % 8. With blocks == 0: the grid
 * launched for a table of `block` entries (rounded up to whole rounds of
 * 32 workgroups; those past the end return at once). */
uint32_t vlj_debug_swizzle(uint32_t block, uint32_t blocks);

/* The force loop's reciprocal 1 / x for x in [1e-4, 1e37] (n values), for
 * the check that it rounds as IEEE division does. No counterpart in the
 * reference. */
int vlj_debug_rcp(const double *x, double *out, size_t n);
/* The FP32 force loop's reciprocal 1 / x and square root for positive
 * normal x (n values each), for the check that they round as IEEE division
 * and square root do (R5f's 1.0f / q and sqrtf). No counterpart in the
 * reference. */
int vlj_debug_rcp32(const float *x, float *rcp, float *root, size_t n);
/* Plain streaming copy of `bytes` (read once, written once) between two
 * device buffers, `repeats` times; *ms = average kernel time from HIP
 * events. Measures the bandwidth a streaming kernel reaches and calibrates
 * rocprof byte counters (no counterpart in the reference). */
int vlj_copy_bandwidth(size_t bytes, uint32_t repeats, double *ms);
/* FP64 vector rate the device reaches: chains of dependent fused
 * multiply-adds, eight independent chains a thread over a grid that fills
 * every CU, `repeats` launches timed with HIP events; *tflops = 2 flops per
 * FMA / time. The measured peak beside the published one for the force
 * kernel's VALU roofline (no counterpart in the reference). */
int vlj_fp64_fma_rate(uint32_t repeats, double *tflops);

/* Kernel timing with HIP events on the handle's stream: runs `steps` steps
 * launching the kernels directly (not from the captured graph) with an event
 * pair around every pair-force launch; returns the summed force-kernel time
 * and the whole elapsed time in milliseconds. */
int vlj_timed_run(vlj_sim *sim, uint64_t steps, double *force_ms, double *total_ms);

#ifdef __cplusplus
}
#endif
#endif
