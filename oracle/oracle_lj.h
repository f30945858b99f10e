/*
 * CPU f64 oracle for the Lennard-Jones hot path of virial-engine
 * (crates/virial-particles/src/lennard_jones.{rs,wgsl}, grid.wgsl,
 * crates/virial-precision/src/fixed.{rs,wgsl}).
 *
 * TEST INFRASTRUCTURE ONLY. The test suite, the smoke check and the
 * CPU-baseline leg of the benchmark call this library to produce the expected
 * answers. The product library never links, loads or calls it.
 *
 * The reference is Rust + WGSL and has no CPU Lennard-Jones path of its own
 * (crates/virial-reference/src/lib.rs holds gravity only), so this file is a
 * restatement of the reference's GPU algorithm in plain C with FP64 state.
 * The arithmetic rules (R1-R12) are written out in oracle_lj.c next to each
 * function; the HIP kernels follow the same rules so that cell orders and
 * neighbour lists are bit-exact and FP64 trajectories can be compared bit for
 * bit on the same inputs.
 */
#ifndef VIRIAL_ORACLE_LJ_H
#define VIRIAL_ORACLE_LJ_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Derived parameters (lennard_jones.rs:443-560). */
typedef struct {
    double box[3];          /* requested sides rounded to f32 (rs:450) */
    double half_box[3];
    uint32_t cells[3];      /* floor(L / (rc + skin)), >= 2 (rs:457-465 asks 3) */
    uint32_t total_cells;
    float cell_size[3];     /* f32 cell sides used by the binning rule (rs:533) */
    uint32_t count;
    uint32_t capacity;      /* list entries per atom */
    uint32_t capacity_reference; /* the reference's own rule (rs:471-474) */
    double rc, rc2, list_r2, half_skin2;
    double shift, force_shift;
    double dt, half_dt;
    double energy_scale, momentum_scale;   /* powers of two (rs:435-440, 491-505) */
    double energy_fixed, momentum_fixed;   /* scale * 2^32 */
    int32_t thermostat;     /* 0 none, 1 Langevin (BAOAB) */
    double c1, c2;          /* Langevin constants (rs:518-529) */
    uint64_t seed;
    int32_t precision;      /* 0 FP64 pair forces, 1 FP32 pair forces */
    int32_t shifted_force;
    uint32_t near_count;    /* condensed criterion (rs:557) */
    double nh_gkt, nh_q;    /* Nose-Hoover g k T and mass Q (rs:516-528) */
    double nh_period;
} vo_params;

typedef struct {
    uint64_t step;
    double kinetic, potential, virial, momentum[3];
    uint32_t count;
    double volume;
    uint32_t list_rebuilds, list_overflow;
    uint32_t condensed;
    uint64_t pairs_within_cutoff;
    int64_t raw[8];         /* fixed-point sums: KE, PE, W, Px, Py, Pz, condensed, 2*pairs */
    double thermostat_energy; /* Nose-Hoover: Q xi^2 / 2 + g k T eta (rs:1235-1241) */
} vo_measurement;

typedef struct vo_sim vo_sim;

/* ---- fixed point (fixed.rs:9-57, fixed.wgsl:20-55) ---- */
void vo_fixed_split(float x, uint32_t *low, uint32_t *high);
int64_t vo_fixed_to_raw(float x);
int64_t vo_fixed_from_words(uint32_t low, uint32_t high);
double vo_fixed_to_f64(int64_t raw);
int64_t vo_fixed_f64(double x, double fixed_scale);
double vo_scale_for(double magnitude);

/* ---- initial states (lennard_jones.rs:1377-1441; tests) ---- */
uint64_t vo_xorshift(uint64_t *s);
void vo_lattice(size_t count, double box_size, double temperature, uint64_t seed,
                double *pos, double *vel);
void vo_lattice_box(size_t count, const double size[3], double temperature, uint64_t seed,
                    double *pos, double *vel);
void vo_jitter(size_t count, uint64_t seed, double amplitude, double *pos);
double vo_fcc(uint32_t cells_per_side, double density, double temperature, uint64_t seed,
              double *pos, double *vel);
int vo_random_insertion(size_t count, const double box[3], double min_distance,
                        double temperature, uint64_t seed, double *pos, double *vel);

/* ---- simulation ---- */
/* thermostat: 0 none, 1 Langevin (friction, seed), 2 Nose-Hoover (period).
 * precision: 0 FP64, 1 FP32 pair forces. */
int vo_derive_params(const double *pos, size_t n, const double box[3], double cutoff,
                     double skin, double dt, int shifted_force, int precision,
                     int thermostat, double temperature, double friction, double period,
                     uint64_t seed, const double *vel, vo_params *out);
vo_sim *vo_create(const double *pos, const double *vel, size_t n, const double box[3],
                  double cutoff, double skin, double dt, int shifted_force, int precision,
                  int thermostat, double temperature, double friction, double period,
                  uint64_t seed);
/* set_temperature (lennard_jones.rs:905-925); no effect without a thermostat. */
void vo_set_temperature(vo_sim *s, double temperature);
void vo_destroy(vo_sim *s);
void vo_get_params(const vo_sim *s, vo_params *out);
void vo_set_near_count(vo_sim *s, uint32_t near_count);
void vo_prime(vo_sim *s);
void vo_run(vo_sim *s, uint64_t steps);
void vo_measure(vo_sim *s, vo_measurement *out);
void vo_read_state(const vo_sim *s, double *pos, double *vel, double *acc, uint32_t *orig);
void vo_read_particles(const vo_sim *s, double *pos, double *vel);
void vo_read_cells(const vo_sim *s, uint32_t *cell_of, uint32_t *cell_start, uint32_t *sorted);
void vo_read_lists(const vo_sim *s, uint32_t *counts, uint32_t *entries);
void vo_read_energies(const vo_sim *s, double *pe, double *w);
uint32_t vo_rebuilds(const vo_sim *s);
/* Profiles along z (lennard_jones.wgsl:628-730): bins * 8 + 3 int64 sums. */
void vo_set_profile_bins(vo_sim *s, uint32_t bins);
void vo_clear_profile(vo_sim *s);
void vo_sample_profile(vo_sim *s);
void vo_read_profile(const vo_sim *s, int64_t *raw, uint64_t *samples);

/* Heat exchange, eHEX/a (lennard_jones.wgsl:732-870; rs:927-950): -1 with a
 * thermostat. on = 0 stops it. */
int vo_set_heat_exchange(vo_sim *s, int on, const double hot[2], const double cold[2], double energy);
/* Heat-current series (lennard_jones.wgsl:872-947; rs:952-1012): raw holds
 * 3 int64 (value * 2^32) per sample. sample returns -1 when full or not
 * recorded. */
void vo_record_heat_current(vo_sim *s, uint32_t capacity);
int vo_sample_heat_current(vo_sim *s);
void vo_read_heat_current(const vo_sim *s, int64_t *raw, uint32_t *samples);

/* Brute-force all pairs, nearest image (tests/thermodynamics.rs:20-53,
 * tests/coexistence.rs:38-76). Any output pointer may be NULL. */
void vo_all_pairs(const double *pos, size_t n, const double box[3], double cutoff,
                  int shifted_force, double *pe, double *w_tensor, double *forces);

/* Counter-based random numbers for the Langevin noise. */
void vo_philox4x32(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
void vo_langevin_noise(uint64_t seed, uint32_t stream, uint32_t atom, uint64_t step,
                       double out[3]);
double vo_log(double x);
void vo_sincos_turns(double u, double *s, double *c);

/* List capacity in place of the rule for sims created afterwards (0: the
 * rule; at least 8), so that tests reach list overflow. */
void vo_set_capacity_override(uint32_t capacity);

void vo_set_threads(int threads);
int vo_max_threads(void);

#ifdef __cplusplus
}
#endif
#endif
