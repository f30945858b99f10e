// Host runtime and C ABI of the Lennard-Jones library (include/virial_lj.h).
//
// The handle owns one HIP stream, every device buffer, a 3-slot pinned
// read-back ring for measurements and the step sequence captured as
// hipGraphs (one step and kGraphSteps steps). It replaces the wgpu-side host
// of crates/virial-particles/src/lennard_jones.rs:443-1375.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/virial_lj.h"
#include "vlj_device.h"
#include "vlj_launch.h"

using namespace vlj;

namespace {

constexpr int kRing = 3;         // read-back slots (lennard_jones.rs:320)
constexpr uint32_t kGraphSteps = 16;
// Brick planner. Bricks are planned to stage at most kDefaultFill of the
// LDS staging limit until the tuner has measured the candidates kTuneFills
// (kTuneSteps steps each) on the handle's own state; a handle of at least
// kTuneMinAtoms atoms tunes itself on the first call that asks for at least
// kTuneAutoSteps steps. Every kCheckSteps steps the control words are copied
// back without a wait, and a replica whose bricks were cut since the
// check before is planned again from its present cell occupancy.
constexpr double kDefaultFill = 0.9;
constexpr double kTuneFills[] = {0.8, 0.9, 1.0};
constexpr uint32_t kTuneSteps = 6;
constexpr uint32_t kTuneMinAtoms = 65536;
constexpr uint64_t kTuneAutoSteps = 64;
constexpr uint64_t kCheckSteps = 256;
constexpr double kTwo32 = 4294967296.0;
constexpr double kPi = 3.14159265358979323846;  // the f64 nearest pi (std::f64::consts::PI)

thread_local std::string g_error = "no error";

// Overrides of the planning rules for tests (vlj_debug_plan); zero keeps a
// rule. Read when a handle is created.
struct PlanOverrides {
  uint32_t list_capacity = 0;  // list entries per atom (at least 8)
  double stage_fill = 0.0;     // dense-layout bricks stage at most this share of kStageMax
  double sparse_fill = 0.0;    // a layer takes the sparse layout under this share of its stage max
  int parts = -1;              // PARTS launches from the start: -1 when the first build would cut, 0 no, 1 yes
};
PlanOverrides g_plan;

struct Ring {
  long long* slots = nullptr;  // pinned
  Control* ctl = nullptr;      // pinned
  hipEvent_t done = nullptr;
  uint64_t step = 0;
  bool in_flight = false;
};

struct HostReplica {
  vlj_settings settings;
  double energy_scale, momentum_scale;
  std::vector<uint32_t> brick_x, brick_y;  // brick shape in cells, per z layer
  std::vector<uint8_t> brick_z;            // per z layer: home layers of the bricks from it (0: the layer
                                           // below's bricks hold it)
  std::vector<uint8_t> sparse;             // per z layer: its bricks use the sparse layout
  double headroom = 1.0;  // share of the plan's fill this replica's bricks are held to
  uint64_t prof_samples;      // profile samples since the last clear
};

}  // namespace

struct vlj_sim {
  int device = 0;
  hipStream_t stream = nullptr;
  vlj::Side side;  // the sparse-layout brick launches, beside the dense ones
  std::string error = "no error";
  std::vector<Replica> rep;
  std::vector<HostReplica> host;
  std::vector<Range> atom_blocks, scan_blocks, sort_blocks;
  std::vector<Brick> bricks;  // n_dense dense bricks, then the sparse ones
  uint32_t n_dense = 0;
  uint32_t n_total = 0, cells_total = 0;
  uint64_t list_total = 0;
  uint64_t prof_total = 0;  // profile sums over all replicas
  int precision = 0;
  bool shifted = false;
  bool any_langevin = false;
  bool any_nose_hoover = false;
  bool any_heat_exchange = false;
  uint32_t hc_samples = 0;  // heat-current samples recorded (replica 0)
  Buffers B{};
  std::vector<void*> allocations;
  Ring ring[kRing];
  // captured steps, for each orientation of the position sets
  hipGraphExec_t graph1[2] = {nullptr, nullptr}, graphN[2] = {nullptr, nullptr};
  // The drift-ahead form (every replica NVE): a step's force launch also
  // drifts the positions of the next step into the second position set
  // (B.x2); `ahead` says it holds them. Each step that uses them swaps the
  // sets (and the rebuild words), so between calls B.x always holds the
  // positions of the present state. `orient`: whether the sets are swapped
  // from their allocation.
  bool ahead = false;
  int orient = 0;
  bool primed = false;
  uint64_t steps = 0;
  long long* fixed_scratch = nullptr;
  // brick planner: the fill the plans are made with, whether it was tuned,
  // the trials of the tuning, and the periodic check for replicas whose
  // bricks are cut (their density rose since they were planned)
  double fill = kDefaultFill;
  bool tuned = false;
  std::vector<vlj::TuneTrial> trials;
  std::vector<uint32_t> rep_bricks;            // bricks of each replica in the table
  std::vector<uint32_t> cut_seen, rebuilds_seen;  // counters at the last check
  Control* check_ctl = nullptr;                // pinned copy of the control words
  hipEvent_t check_done = nullptr;
  bool check_in_flight = false;
  uint64_t last_check = 0;                     // steps at the last check
  uint32_t replans = 0;                        // re-plans after the first plan
};

namespace {

int fail(vlj_sim* s, int code, const std::string& msg) {
  if (s) s->error = msg;
  g_error = msg;
  return code;
}

// Calls that act on one system refuse a handle of several replicas, which
// has its own vlj_batch_* calls.
int single_only(vlj_sim* s, const char* what) {
  if (s->rep.size() <= 1) return VLJ_OK;
  return fail(s, VLJ_E_STATE,
              std::string(what) + " acts on one system; this handle holds " + std::to_string(s->rep.size()) +
                  " replicas (use the vlj_batch_* calls)");
}

#define SINGLE_ONLY(sim, what)                  \
  do {                                          \
    int rc_ = single_only((sim), (what));       \
    if (rc_ != VLJ_OK) return rc_;              \
  } while (0)

#define HIP_TRY(sim, expr)                                                             \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess)                                                              \
      return fail((sim), VLJ_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

// lennard_jones.rs:435-440
double scale_for(double magnitude) {
  if (magnitude <= 0.0 || !std::isfinite(magnitude)) return 1.0;
  return std::ldexp(1.0, (int)std::floor(20.0 - std::log2(magnitude)));
}

double wrap_initial(double x, double side) {  // Rust rem_euclid (lennard_jones.rs:477)
  double r = std::fmod(x, side);
  return r < 0.0 ? r + side : r;
}

uint32_t host_cell_axis(double x, float cs, uint32_t n) {
  float q = (float)x / cs;
  int c = (int)std::floor(q);
  int m = (int)n - 1;
  return (uint32_t)(c < 0 ? 0 : (c > m ? m : c));
}

// lennard_jones.rs:443-560 for one replica. Deliberate changes: list
// capacity from the densest cell of the initial state (not only the box
// average), and dt kept in FP64.
int derive(const vlj_particle* p, size_t n, const vlj_settings& st, Replica& R, HostReplica& H,
           std::vector<uint32_t>& occ, std::string& err) {
  std::memset(&R, 0, sizeof(R));
  H.settings = st;
  if (!(st.cutoff > 0.0) || !(st.skin >= 0.0) || !(st.dt > 0.0)) {
    err = "cutoff and dt must be positive and skin non-negative";
    return VLJ_E_INVALID;
  }
  if (st.thermostat.kind > VLJ_THERMOSTAT_NOSE_HOOVER) {
    err = "unknown thermostat";
    return VLJ_E_INVALID;
  }
  if (st.precision > VLJ_PRECISION_FP32) {
    err = "unknown precision";
    return VLJ_E_INVALID;
  }
  const double list_radius = st.cutoff + st.skin;
  uint32_t cells[3];
  for (int k = 0; k < 3; k++) {
    if (!(st.box_size[k] > 0.0) || !std::isfinite(st.box_size[k])) {
      err = "box sides must be positive";
      return VLJ_E_INVALID;
    }
    R.box[k] = (double)(float)st.box_size[k];
    R.half_box[k] = 0.5 * R.box[k];
    R.inv_box[k] = 1.0 / R.box[k];
    double c = std::floor(R.box[k] / list_radius);
    cells[k] = c < 0 ? 0 : (c > 4.0e9 ? 4000000000u : (uint32_t)c);
  }
  // The reference asks for 3 cells per axis (lennard_jones.rs:458-461); 2
  // suffice: with L >= 2 (rc + skin) at most one image of a pair is within
  // rc + skin, so a stencil meeting a cell twice lists each pair once.
  for (int k = 0; k < 3; k++)
    if (cells[k] < 2) {
      err = "the box must be at least two (cutoff + skin) wide on every axis";
      return VLJ_E_INVALID;
    }
  while ((uint64_t)cells[0] * cells[1] * cells[2] > (1u << 20)) {
    int longest = 0;
    for (int k = 1; k < 3; k++)
      if (cells[k] >= cells[longest]) longest = k;
    cells[longest] -= 1;
  }
  double cell_size[3];
  for (int k = 0; k < 3; k++) {
    R.cells[k] = cells[k];
    cell_size[k] = R.box[k] / (double)cells[k];
    R.cell_size[k] = (float)cell_size[k];
  }
  R.total_cells = cells[0] * cells[1] * cells[2];
  R.count = (uint32_t)n;
  const double volume = R.box[0] * R.box[1] * R.box[2];
  const double density = (double)n / volume;
  const double sphere = 4.0 / 3.0 * 3.14159265358979323846 * list_radius * list_radius * list_radius;
  occ.assign(R.total_cells, 0);
  uint32_t occ_max = 0;
  for (size_t i = 0; i < n; i++) {
    double w[3];
    for (int k = 0; k < 3; k++) {
      if (!std::isfinite(p[i].pos[k]) || !std::isfinite(p[i].vel[k])) {
        err = "particle positions and velocities must be finite";
        return VLJ_E_INVALID;
      }
      w[k] = wrap_initial(p[i].pos[k], R.box[k]);
    }
    uint32_t c = (host_cell_axis(w[2], R.cell_size[2], cells[2]) * cells[1] +
                  host_cell_axis(w[1], R.cell_size[1], cells[1])) *
                     cells[0] +
                 host_cell_axis(w[0], R.cell_size[0], cells[0]);
    occ_max = std::max(occ_max, ++occ[c]);
  }
  const double cell_volume = cell_size[0] * cell_size[1] * cell_size[2];
  const double use = std::max((double)occ_max / cell_volume, density);
  R.capacity = (uint32_t)std::ceil(2.0 * sphere * use + 32.0);
  // vlj_debug_plan may replace the rule (at least 8): tests use a small
  // capacity to reach list overflow, which the rule itself avoids for the
  // initial state
  if (g_plan.list_capacity) R.capacity = std::max(8u, g_plan.list_capacity);

  double kinetic = 0.0, speed = 0.0;
  for (size_t i = 0; i < n; i++) {
    double v2 = 0.0;
    for (int k = 0; k < 3; k++) v2 += p[i].vel[k] * p[i].vel[k];
    kinetic += 0.5 * v2;
    speed += std::sqrt(v2);
  }
  const bool langevin = st.thermostat.kind == VLJ_THERMOSTAT_LANGEVIN;
  const bool nose_hoover = st.thermostat.kind == VLJ_THERMOSTAT_NOSE_HOOVER;
  const double thermal = langevin || nose_hoover ? 1.5 * st.thermostat.temperature * (double)n : 0.0;
  H.energy_scale = scale_for(std::max(std::max(kinetic, thermal), (double)n));
  H.momentum_scale = scale_for(std::max(speed, (double)n));
  R.energy_fixed = H.energy_scale * kTwo32;
  R.momentum_fixed = H.momentum_scale * kTwo32;
  R.rc = st.cutoff;
  R.rc2 = st.cutoff * st.cutoff;
  R.list_r2 = list_radius * list_radius;
  R.half_skin2 = 0.25 * st.skin * st.skin;
  const double inv6 = 1.0 / (R.rc2 * R.rc2 * R.rc2);
  R.shift = 4.0 * (inv6 * inv6 - inv6);
  R.force_shift = st.shifted_force ? 24.0 * (2.0 * inv6 * inv6 - inv6) / st.cutoff : 0.0;
  R.force_shift24 = R.force_shift / 24.0;
  R.dt = st.dt;
  R.half_dt = 0.5 * st.dt;
  R.thermostat = langevin ? 1u : (nose_hoover ? 2u : 0u);
  R.c1 = 1.0;
  R.c2 = 0.0;
  R.nh_q = 1.0;
  if (nose_hoover) {
    // lennard_jones.rs:516-528: zero total momentum, g = 3N - 3
    if (!(st.thermostat.temperature >= 0.0) || !(st.thermostat.period > 0.0)) {
      err = "Nose-Hoover needs a non-negative temperature and a positive period";
      return VLJ_E_INVALID;
    }
    const double dof = 3.0 * (double)n - 3.0;
    R.nh_period = st.thermostat.period;
    R.nh_gkt = dof * st.thermostat.temperature;
    R.nh_q = dof * st.thermostat.temperature * R.nh_period * R.nh_period / (4.0 * (kPi * kPi));
  }
  if (langevin) {
    if (!(st.thermostat.temperature >= 0.0) || !(st.thermostat.friction >= 0.0)) {
      err = "thermostat temperature and friction must be non-negative";
      return VLJ_E_INVALID;
    }
    R.c1 = std::exp(-st.thermostat.friction * st.dt);
    R.c2 = std::sqrt((1.0 - R.c1 * R.c1) * st.thermostat.temperature);
  }
  R.seed = st.thermostat.seed;
  R.near_count = 4;
  if (st.profile_bins > kMaxProfileBins) {
    err = "profile_bins must be at most " + std::to_string(kMaxProfileBins);
    return VLJ_E_INVALID;
  }
  R.prof_bins = st.profile_bins;
  R.bin_width = R.box[2] / (double)(st.profile_bins ? st.profile_bins : 1u);
  H.prof_samples = 0;
  return VLJ_OK;
}

template <typename T>
int dev_alloc(vlj_sim* s, T** out, size_t count) {
  void* p = nullptr;
  size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
  hipError_t e = hipMalloc(&p, bytes);
  if (e != hipSuccess)
    return fail(s, VLJ_E_HIP, "hipMalloc(" + std::to_string(bytes) + "): " + hipGetErrorString(e));
  s->allocations.push_back(p);
  *out = (T*)p;
  return VLJ_OK;
}

void free_device(vlj_sim* s, void* p) {
  if (!p) return;
  auto it = std::find(s->allocations.begin(), s->allocations.end(), p);
  if (it != s->allocations.end()) s->allocations.erase(it);
  (void)hipFree(p);
}

void free_all(vlj_sim* s) {
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  for (int k = 0; k < 2; k++) {
    if (s->graph1[k]) (void)hipGraphExecDestroy(s->graph1[k]);
    if (s->graphN[k]) (void)hipGraphExecDestroy(s->graphN[k]);
  }
  for (auto& r : s->ring) {
    if (r.slots) (void)hipHostFree(r.slots);
    if (r.ctl) (void)hipHostFree(r.ctl);
    if (r.done) (void)hipEventDestroy(r.done);
  }
  if (s->check_ctl) (void)hipHostFree(s->check_ctl);
  s->check_ctl = nullptr;
  if (s->check_done) (void)hipEventDestroy(s->check_done);
  s->check_done = nullptr;
  for (void* p : s->allocations) (void)hipFree(p);
  s->allocations.clear();
  if (s->B.hc) (void)hipFree(s->B.hc);
  s->B.hc = nullptr;
  if (s->side.stream) (void)hipStreamSynchronize(s->side.stream);
  if (s->side.fork) (void)hipEventDestroy(s->side.fork);
  if (s->side.join) (void)hipEventDestroy(s->side.join);
  if (s->side.stream) (void)hipStreamDestroy(s->side.stream);
  s->side = vlj::Side{};
  if (s->stream) (void)hipStreamDestroy(s->stream);
  s->stream = nullptr;
}

int check_device(vlj_sim* s) {
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n == 0)
    return fail(s, VLJ_E_NO_DEVICE, "no HIP device available (this library runs on gfx950)");
  HIP_TRY(s, hipGetDevice(&s->device));
  hipDeviceProp_t prop;
  HIP_TRY(s, hipGetDeviceProperties(&prop, s->device));
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return fail(s, VLJ_E_NO_DEVICE,
                std::string("this library is built for gfx950; the device is ") + prop.gcnArchName);
  return VLJ_OK;
}

// Enqueue one step: drift, rebuild chain, forces + kick.
// One step (lennard_jones.rs:1092-1138): with Nose-Hoover replicas a
// thermostat half step before the drift (its scaling is applied there) and
// one after the force kick.
// `around_force`, when given, is a pair of events recorded before and after
// the pair-force launches.
// Whether the steps take the drift-ahead form: every replica NVE (a
// thermostat or heat exchange acts between the half steps).
bool drift_ahead(const vlj_sim* s) { return !s->any_langevin && !s->any_nose_hoover && !s->any_heat_exchange; }

// Swaps the two position sets and the two rebuild words.
void flip_positions(vlj_sim* s) {
  Buffers& B = s->B;
  std::swap(B.x, B.x2);
  std::swap(B.y, B.y2);
  std::swap(B.z, B.z2);
  std::swap(B.rebuilding, B.rebuilding_next);
  std::swap(B.amax, B.amax_next);
  s->orient ^= 1;
}

int enqueue_step(vlj_sim* s, hipEvent_t* around_force = nullptr) {
  const bool fused = drift_ahead(s);
  if (s->any_nose_hoover) HIP_TRY(s, launch_nh_half(s->B, false, s->stream));
  if (fused && s->ahead) {
    flip_positions(s);  // the positions the last force launch drifted
    Buffers B = s->B;
    B.bin_decides = 1u;  // the rebuild decision in k_bin, from the force launch's maxima
    HIP_TRY(s, launch_rebuild(B, s->stream, &s->side));
  } else {
    HIP_TRY(s, launch_drift(s->B, s->any_langevin, fused, s->stream));
    HIP_TRY(s, launch_rebuild(s->B, s->stream, &s->side));
  }
  if (around_force) HIP_TRY(s, hipEventRecord(around_force[0], s->stream));
  HIP_TRY(s, launch_force(s->B, s->precision, s->shifted, fused ? 2 : 1, s->stream, &s->side));
  if (around_force) HIP_TRY(s, hipEventRecord(around_force[1], s->stream));
  s->ahead = fused;
  if (s->any_nose_hoover) HIP_TRY(s, launch_nh_half(s->B, true, s->stream));
  if (s->any_heat_exchange) HIP_TRY(s, launch_heat_exchange(s->B, s->stream));
  return VLJ_OK;
}

// The captured steps change shape (a kernel more or less): capture again
// on the next step.
void drop_graphs(vlj_sim* s) {
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  for (int k = 0; k < 2; k++) {
    if (s->graph1[k]) (void)hipGraphExecDestroy(s->graph1[k]);
    if (s->graphN[k]) (void)hipGraphExecDestroy(s->graphN[k]);
    s->graph1[k] = s->graphN[k] = nullptr;
  }
}

// Captures `steps` steps from the present state; the handle's host state
// (position sets, drift-ahead flag) is left as before, since nothing ran.
int capture(vlj_sim* s, uint32_t steps, hipGraphExec_t* out) {
  hipGraph_t g = nullptr;
  const Buffers saved = s->B;
  const bool ahead = s->ahead;
  const int orient = s->orient;
  HIP_TRY(s, hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
  int rc = VLJ_OK;
  for (uint32_t t = 0; t < steps && rc == VLJ_OK; t++) rc = enqueue_step(s);
  s->B = saved;
  s->ahead = ahead;
  s->orient = orient;
  hipError_t e = hipStreamEndCapture(s->stream, &g);
  if (rc != VLJ_OK) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) return fail(s, VLJ_E_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
  e = hipGraphInstantiateWithFlags(out, g, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess)
    return fail(s, VLJ_E_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
  return VLJ_OK;
}

// Brick shape of each z layer: the largest gx x gy (home cells, up to
// kBrickMax each way) whose every brick stages at most `fill` of kStageMax
// atoms. The count a brick is held to is the larger of what it stages at
// its own layer and at the layers above and below (a density profile may
// drift by a cell, as a liquid-vapour interface does) and what it would
// stage at the replica's mean occupancy. The occupancy is the initial
// state's when a handle is created and that of the last list build when a
// replica is planned again. Window sums come
// from a 2D prefix sum of the three layers' counts over the periodically
// extended plane. A brick that outgrows LDS later on is cut into parts on
// the device. A layer whose bricks are held to at most `sparse_fill` of the
// sparse layout's 1024 atoms takes that layout.
void choose_bricks(const Replica& R, const std::vector<uint32_t>& occ, double fill, double sparse_fill,
                   std::vector<uint32_t>& bx, std::vector<uint32_t>& by, std::vector<uint8_t>& bz,
                   std::vector<uint8_t>& sparse) {
  const int nx = (int)R.cells[0], ny = (int)R.cells[1], nz = (int)R.cells[2];
  const double limit = fill * kStageMax;
  const double sparse_limit = sparse_fill * Sparse::kStageMax;
  // Two floors for a brick's count: the replica's mean occupancy (a
  // lattice filled from the bottom, with empty layers above, melts into
  // them; a gas condenses into droplets anywhere) sets the brick shape; the
  // mean of the layers it stages (a slab's vapour stays a vapour) sets the
  // layout.
  const double whole = (double)R.count / (double)R.total_cells;
  std::vector<uint64_t> layer(nz, 0);
  for (int z = 0; z < nz; z++)
    for (size_t c = 0; c < (size_t)nx * ny; c++) layer[z] += occ[(size_t)z * nx * ny + c];
  auto wrap_z = [nz](int z) { return ((z % nz) + nz) % nz; };
  // mean occupancy of the layers z - 1 .. z + k (what a brick of k home
  // layers from z stages)
  auto layers_mean = [&](int z, int k) {
    uint64_t n = 0;
    for (int i = -1; i <= k; i++) n += layer[wrap_z(z + i)];
    return (double)n / ((double)(k + 2) * nx * ny);
  };
  // per layer, prefix sums of its counts over the periodically extended plane
  const int ex = nx + 2, ey = ny + 2;  // plane extended by one cell each side
  const size_t plane = (size_t)(ex + 1) * (ey + 1);
  std::vector<uint64_t> pre(plane * nz);
  for (int z = 0; z < nz; z++) {
    uint64_t* P = pre.data() + plane * z;
    for (int y = 0; y <= ey; y++)
      for (int x = 0; x <= ex; x++) {
        uint64_t v = 0;
        if (x > 0 && y > 0) {
          const int cx = (x - 2 + nx) % nx, cy = (y - 2 + ny) % ny;  // extended x - 1 is cell x - 2
          v = occ[(size_t)(z * ny + cy) * nx + cx] + P[(size_t)y * (ex + 1) + (x - 1)] +
              P[(size_t)(y - 1) * (ex + 1) + x] - P[(size_t)(y - 1) * (ex + 1) + (x - 1)];
        }
        P[(size_t)y * (ex + 1) + x] = v;
      }
  }
  // atoms in the cells x0 - 1 .. x0 + w, y0 - 1 .. y0 + h of layer z (the
  // extended columns x0 .. x0 + w + 1)
  auto window1 = [&](int z, int x0, int y0, int w, int h) {
    const uint64_t* P = pre.data() + plane * wrap_z(z);
    const int xa = x0, xb = x0 + w + 2, ya = y0, yb = y0 + h + 2;
    return P[(size_t)yb * (ex + 1) + xb] - P[(size_t)ya * (ex + 1) + xb] - P[(size_t)yb * (ex + 1) + xa] +
           P[(size_t)ya * (ex + 1) + xa];
  };
  // what a brick of k home layers from layer z stages: layers z - 1 .. z + k
  auto staged_count = [&](int z, int k, int x0, int y0, int w, int h) {
    uint64_t n = 0;
    for (int i = -1; i <= k; i++) n += window1(z + i, x0, y0, w, h);
    return n;
  };
  // A count held to the limit is the count plus four of its Poisson
  // standard deviations: a liquid's windows fluctuate about what the plan
  // saw (a lattice at the start), and a brick that outgrows LDS is cut.
  auto margin = [](uint64_t n) { return (double)n + 4.0 * std::sqrt((double)n); };
  struct Shape {
    uint32_t gx = 1, gy = 1;
    double worst = 0.0, local = 0.0;  // held count; held count by the staged layers' mean
    bool fits = false;
  };
  // The largest gx x gy of k home layers from layer z whose every brick
  // stages at most the limit, there and one layer up or down (a density
  // profile may drift by a cell, as a liquid-vapour interface does), and at
  // the replica's mean occupancy.
  auto best_shape = [&](int z, int k) {
    Shape b;
    uint32_t best = 0;
    for (int gy = 1; gy <= std::min(kBrickMax, ny); gy++)
      for (int gx = gy; gx <= std::min(kBrickMax, nx); gx++) {
        const uint32_t home = (uint32_t)(gx * gy * k);
        if (home <= best || home > (uint32_t)kHomeCells || (k + 2) * (gx + 2) * (gy + 2) > kStageCells) continue;
        const double staged_cells = (double)(k + 2) * (gx + 2) * (gy + 2);
        double windows = 0.0;  // the largest window, with its margin
        const double floor_count = margin((uint64_t)(whole * staged_cells));
        for (int dz = -1; dz <= 1 && windows <= limit; dz++)
          for (int y0 = 0; y0 < ny && windows <= limit; y0 += gy)
            for (int x0 = 0; x0 < nx && windows <= limit; x0 += gx)
              windows = std::max(
                  windows, margin(staged_count(z + dz, k, x0, y0, std::min(gx, nx - x0), std::min(gy, ny - y0))));
        const double worst = std::max(windows, floor_count);
        if (worst <= limit) {
          best = home;
          b.gx = (uint32_t)gx;
          b.gy = (uint32_t)gy;
          b.worst = worst;
          b.local = std::max(windows, margin((uint64_t)(layers_mean(z, k) * staged_cells)));
          b.fits = true;
        }
      }
    return b;
  };
  auto staged_per_home = [](const Shape& b, int k) {
    return (double)(k + 2) * (b.gx + 2) * (b.gy + 2) / ((double)b.gx * b.gy * k);
  };
  // a brick that stages few atoms takes the sparse layout: few by its own
  // layers' counts, and no more than twice the sparse limit by the
  // replica's mean
  auto sparse_of = [&](const Shape& b) {
    return b.fits && b.local <= sparse_limit && b.worst <= 2.0 * sparse_limit ? 1 : 0;
  };
  bx.assign(nz, 1);
  by.assign(nz, 1);
  bz.assign(nz, 1);
  sparse.assign(nz, 0);
  std::vector<Shape> one(nz);
  for (int z = 0; z < nz; z++) {
    one[z] = best_shape(z, 1);
    bx[z] = one[z].gx;
    by[z] = one[z].gy;
    sparse[z] = (uint8_t)sparse_of(one[z]);
  }
  // Layers z, z + 1 (z even) share two-layer bricks when these stage fewer
  // cells per home cell than the two layers' own bricks: a brick stages one
  // layer below and one above its home layers, so two home layers stage
  // four where two one-layer bricks stage six.
  if (nz >= 4)
    for (int z = 0; z + 1 < nz; z += 2) {
      const Shape two = best_shape(z, 2);
      if (!two.fits) continue;
      const double apart = 0.5 * (staged_per_home(one[z], 1) + staged_per_home(one[z + 1], 1));
      if (!(staged_per_home(two, 2) < apart)) continue;
      bx[z] = bx[z + 1] = two.gx;
      by[z] = by[z + 1] = two.gy;
      bz[z] = 2;
      bz[z + 1] = 0;
      sparse[z] = sparse[z + 1] = (uint8_t)sparse_of(two);
    }
}

// The fills a plan is made with: the handle's (tuned) fill unless
// vlj_debug_plan overrides it; the sparse layout's threshold follows the
// dense one unless overridden (a negative override keeps every layer dense).
double plan_fill(const vlj_sim* s) { return g_plan.stage_fill != 0.0 ? g_plan.stage_fill : s->fill; }
double plan_sparse_fill(const vlj_sim* s) {
  return g_plan.sparse_fill != 0.0 ? g_plan.sparse_fill : plan_fill(s);
}

// The brick table from the replicas' brick shapes: every replica's dense
// bricks, then every sparse one, each part with its larger bricks (more
// home cells) first and otherwise in layer, row and column order, so that
// the last workgroups of a launch, which run while the rest of the GPU has
// nothing left to start, are the smallest, and neighbouring bricks still
// run side by side (sharing their halos in L2). The 1M-atom force kernel:
// 154 -> 146 us; the 512-replica batch +3.7%. Ordering by each brick's
// estimated pair work instead scattered neighbours: 149 us, +2.6%.
void build_table(vlj_sim* s) {
  s->bricks.clear();
  s->rep_bricks.assign(s->rep.size(), 0u);
  for (int pass = 0; pass < 2; pass++) {
    const size_t pass_start = s->bricks.size();
    for (size_t r = 0; r < s->rep.size(); r++) {
      const Replica& R = s->rep[r];
      const HostReplica& H = s->host[r];
      for (uint32_t z = 0; z < R.cells[2]; z++) {
        if (H.brick_z[z] == 0 || (H.sparse[z] != 0) != (pass == 1)) continue;  // layer held by the one below
        const uint32_t gx = H.brick_x[z], gy = H.brick_y[z];
        for (uint32_t y0 = 0; y0 < R.cells[1]; y0 += gy)
          for (uint32_t x0 = 0; x0 < R.cells[0]; x0 += gx) {
            s->bricks.push_back({(uint32_t)r, (z * R.cells[1] + y0) * R.cells[0] + x0,
                                 std::min(gx, R.cells[0] - x0), std::min(gy, R.cells[1] - y0), H.brick_z[z]});
            s->rep_bricks[r]++;
          }
      }
    }
    std::stable_sort(s->bricks.begin() + (std::ptrdiff_t)pass_start, s->bricks.end(),
                     [](const Brick& a, const Brick& b) { return home_cells(a) > home_cells(b); });
    if (pass == 0) s->n_dense = (uint32_t)s->bricks.size();
  }
}

// lennard_jones.rs:1080-1089: first lists and accelerations.
int prime_if_needed(vlj_sim* s) {
  if (s->primed) return VLJ_OK;
  HIP_TRY(s, launch_decide(s->B, s->stream));
  HIP_TRY(s, launch_rebuild(s->B, s->stream, &s->side));
  HIP_TRY(s, launch_force(s->B, s->precision, s->shifted, 0, s->stream, &s->side));
  s->primed = true;
  return VLJ_OK;
}

// The side stream for the sparse-layout launches, once both layouts have
// bricks (default priority).
int ensure_side(vlj_sim* s) {
  if (s->side.stream || s->n_dense == 0 || s->n_dense == s->bricks.size()) return VLJ_OK;
  HIP_TRY(s, hipStreamCreateWithFlags(&s->side.stream, hipStreamNonBlocking));
  HIP_TRY(s, hipEventCreateWithFlags(&s->side.fork, hipEventDisableTiming));
  HIP_TRY(s, hipEventCreateWithFlags(&s->side.join, hipEventDisableTiming));
  return VLJ_OK;
}

// Puts the brick table s->bricks on the device (new buffers when its size
// changed) with every cut mark cleared; the captured steps hold the old
// table's size and pointers, so they are captured again.
int upload_table(vlj_sim* s) {
  Buffers& B = s->B;
  const size_t nb = s->bricks.size(), ng = (nb + kCutGroup - 1) / kCutGroup;
  drop_graphs(s);
  if (nb != B.n_bricks) {
    free_device(s, (void*)B.bricks);
    free_device(s, B.cut);
    free_device(s, B.parts);
    free_device(s, B.cut_groups);
    B.bricks = nullptr;
    B.cut = nullptr;
    B.parts = nullptr;
    B.cut_groups = nullptr;
    Brick* gb = nullptr;
    int rc = dev_alloc(s, &gb, nb);
    if (rc == VLJ_OK) rc = dev_alloc(s, &B.cut, nb);
    if (rc == VLJ_OK) rc = dev_alloc(s, &B.parts, nb * kHomeCells);
    if (rc == VLJ_OK) rc = dev_alloc(s, &B.cut_groups, ng);
    if (rc != VLJ_OK) return rc;
    B.bricks = gb;
  }
  B.n_bricks = (uint32_t)nb;
  B.n_dense = s->n_dense;
  HIP_TRY(s, hipMemcpy((void*)B.bricks, s->bricks.data(), nb * sizeof(Brick), hipMemcpyHostToDevice));
  HIP_TRY(s, hipMemset(B.cut, 0, std::max<size_t>(nb, 1) * 4));
  HIP_TRY(s, hipMemset(B.cut_groups, 0, std::max<size_t>(ng, 1) * 4));
  HIP_TRY(s, hipMemset(B.flags + 1, 0, 4));
  return ensure_side(s);
}

// Plans the bricks of the replicas marked in `which` again with `fill`,
// from the cell occupancy of their last list build, and builds every list
// again from the positions of that build (launch_relist): the same lists,
// in the same order, staged for the new bricks, so the trajectory does not
// change. *build_ms, when given, gets the time of that list build.
int replan(vlj_sim* s, const std::vector<char>& which, double fill, double* build_ms = nullptr) {
  HIP_TRY(s, hipStreamSynchronize(s->stream));
  std::vector<uint32_t> cn(std::max<size_t>(s->cells_total, 1));
  if (s->cells_total) HIP_TRY(s, hipMemcpy(cn.data(), s->B.cell_n, s->cells_total * 4, hipMemcpyDeviceToHost));
  const double saved = s->fill;
  s->fill = fill;
  for (size_t r = 0; r < s->rep.size(); r++) {
    if (!which[r]) continue;
    const Replica& R = s->rep[r];
    std::vector<uint32_t> occ(cn.begin() + R.cell_off, cn.begin() + R.cell_off + R.total_cells);
    HostReplica& H = s->host[r];
    choose_bricks(R, occ, plan_fill(s) * H.headroom, plan_sparse_fill(s) * H.headroom, H.brick_x, H.brick_y,
                  H.brick_z, H.sparse);
  }
  if (g_plan.stage_fill != 0.0) s->fill = saved;
  build_table(s);
  int rc = upload_table(s);
  if (rc != VLJ_OK) return rc;
  hipEvent_t ev[2] = {nullptr, nullptr};
  if (build_ms) {
    HIP_TRY(s, hipEventCreate(&ev[0]));
    HIP_TRY(s, hipEventCreate(&ev[1]));
    HIP_TRY(s, hipEventRecord(ev[0], s->stream));
  }
  HIP_TRY(s, launch_relist(s->B, s->stream));
  if (build_ms) HIP_TRY(s, hipEventRecord(ev[1], s->stream));
  HIP_TRY(s, hipStreamSynchronize(s->stream));
  if (build_ms) {
    float ms = 0.0f;
    HIP_TRY(s, hipEventElapsedTime(&ms, ev[0], ev[1]));
    *build_ms = ms;
    (void)hipEventDestroy(ev[0]);
    (void)hipEventDestroy(ev[1]);
  }
  return VLJ_OK;
}

// The planner's tuner: for each candidate fill, plan every replica's bricks
// with it (replan), then run `per` steps launching the kernels directly
// with HIP events around the pair-force launches; the score is the mean
// force time per step plus the list build's time times the rebuilds per
// step seen over the trials. The fastest plan is kept. The steps are real
// steps of the run, and bricks do not change results, so tuning changes no
// bit of the trajectory. Returns the steps taken.
int tune(vlj_sim* s, uint32_t per, uint64_t* taken) {
  *taken = 0;
  int rc = prime_if_needed(s);
  if (rc != VLJ_OK) return rc;
  const std::vector<char> all(s->rep.size(), 1);
  std::vector<hipEvent_t> ev(2 * per, nullptr);
  for (auto& e : ev) HIP_TRY(s, hipEventCreate(&e));
  auto rebuilds = [&](uint64_t* out) {
    std::vector<Control> c(s->rep.size());
    HIP_TRY(s, hipMemcpy(c.data(), s->B.ctl, c.size() * sizeof(Control), hipMemcpyDeviceToHost));
    *out = 0;
    for (const Control& x : c) *out += x.rebuilds;
    return VLJ_OK;
  };
  uint64_t r0 = 0, r1 = 0;
  rc = rebuilds(&r0);
  s->trials.clear();
  for (double f : kTuneFills) {
    double build_ms = 0.0;
    if (rc == VLJ_OK) rc = replan(s, all, f, &build_ms);
    double force_ms = 0.0;
    for (uint32_t t = 0; t < per && rc == VLJ_OK; t++) rc = enqueue_step(s, &ev[2 * t]);
    if (rc != VLJ_OK) break;
    HIP_TRY(s, hipStreamSynchronize(s->stream));
    for (uint32_t t = 0; t < per; t++) {
      float ms = 0.0f;
      HIP_TRY(s, hipEventElapsedTime(&ms, ev[2 * t], ev[2 * t + 1]));
      force_ms += ms;
    }
    s->steps += per;
    *taken += per;
    s->trials.push_back({f, 1e3 * force_ms / per, 1e3 * build_ms, (uint32_t)s->bricks.size()});
  }
  for (auto& e : ev) (void)hipEventDestroy(e);
  if (rc == VLJ_OK) rc = rebuilds(&r1);
  if (rc != VLJ_OK) return rc;
  const double per_step = std::max(1.0 / 16.0, (double)(r1 - r0) / std::max<double>(1.0, (double)*taken)) /
                          (double)std::max<size_t>(s->rep.size(), 1);
  size_t best = 0;
  for (size_t k = 1; k < s->trials.size(); k++) {
    const TuneTrial& a = s->trials[k];
    const TuneTrial& b = s->trials[best];
    if (a.force_us + per_step * a.build_us < b.force_us + per_step * b.build_us) best = k;
  }
  s->tuned = true;
  if (s->trials.empty()) return VLJ_OK;
  if (s->trials[best].fill != s->trials.back().fill) return replan(s, all, s->trials[best].fill);
  s->fill = s->trials.back().fill;
  return VLJ_OK;
}

bool tune_wanted(const vlj_sim* s) {
  return !s->tuned && g_plan.stage_fill == 0.0 && s->n_total >= kTuneMinAtoms && !s->bricks.empty();
}

// The periodic check: takes the control words copied back at the last
// check, if they have arrived, and plans again every replica whose bricks
// were cut since the check before (when that changes its plan); then
// starts the next copy
// once kCheckSteps steps have passed. Never waits for the device except
// to plan again.
int check_bricks(vlj_sim* s) {
  if (g_plan.stage_fill != 0.0 || s->bricks.empty()) return VLJ_OK;
  const size_t R = s->rep.size();
  if (!s->check_ctl) {
    HIP_TRY(s, hipHostMalloc((void**)&s->check_ctl, R * sizeof(Control), hipHostMallocDefault));
    HIP_TRY(s, hipEventCreateWithFlags(&s->check_done, hipEventDisableTiming));
    s->cut_seen.assign(R, 0u);
    s->rebuilds_seen.assign(R, 0u);
  }
  if (s->check_in_flight) {
    hipError_t e = hipEventQuery(s->check_done);
    if (e == hipErrorNotReady) return VLJ_OK;
    if (e != hipSuccess) return fail(s, VLJ_E_HIP, std::string("hipEventQuery: ") + hipGetErrorString(e));
    s->check_in_flight = false;
    std::vector<char> which(R, 0);
    bool any = false;
    for (size_t r = 0; r < R; r++) {
      const Control& c = s->check_ctl[r];
      const uint32_t dcut = c.stage_overflow - s->cut_seen[r], dreb = c.rebuilds - s->rebuilds_seen[r];
      s->cut_seen[r] = c.stage_overflow;
      s->rebuilds_seen[r] = c.rebuilds;
      if (dreb == 0 || dcut == 0) continue;
      // plan it from the present occupancy only if that gives other bricks
      which[r] = 1;
      any = true;
    }
    if (any) {
      std::vector<uint32_t> cn(s->cells_total);
      HIP_TRY(s, hipStreamSynchronize(s->stream));
      HIP_TRY(s, hipMemcpy(cn.data(), s->B.cell_n, s->cells_total * 4, hipMemcpyDeviceToHost));
      // A handle without the PARTS launches ran its outgrown bricks whole
      // from global memory; from now on they are cut into parts, which needs
      // the lists of those replicas built again.
      const bool parts_now = s->B.parts_on == 0u;
      s->B.parts_on = 1u;
      any = false;
      for (size_t r = 0; r < R; r++) {
        if (!which[r]) continue;
        const Replica& rp = s->rep[r];
        std::vector<uint32_t> occ(cn.begin() + rp.cell_off, cn.begin() + rp.cell_off + rp.total_cells);
        // a replica whose bricks keep being cut is held to 10% less of the
        // staging limit each time (down to 60%)
        HostReplica& H = s->host[r];
        H.headroom = std::max(0.6, 0.9 * H.headroom);
        std::vector<uint32_t> bx, by;
        std::vector<uint8_t> bz, sp;
        choose_bricks(rp, occ, plan_fill(s) * H.headroom, plan_sparse_fill(s) * H.headroom, bx, by, bz, sp);
        which[r] = parts_now || bx != H.brick_x || by != H.brick_y || bz != H.brick_z || sp != H.sparse;
        any = any || which[r];
      }
      if (any) {
        int rc = replan(s, which, s->fill);
        if (rc != VLJ_OK) return rc;
        s->replans++;
      }
    }
  }
  if (s->steps >= s->last_check + kCheckSteps) {
    HIP_TRY(s, hipMemcpyAsync(s->check_ctl, s->B.ctl, R * sizeof(Control), hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(s, hipEventRecord(s->check_done, s->stream));
    s->check_in_flight = true;
    s->last_check = s->steps;
  }
  return VLJ_OK;
}

// Whether a brick of the table stages more atoms than its layout's LDS holds
// at the given cell occupancy: a handle whose first list build would cut
// bricks runs the PARTS launches from the start. Otherwise it starts without
// them (a brick outgrowing LDS later runs whole from global memory until the
// periodic check turns them on).
bool any_overflow(const vlj_sim* s, const std::vector<std::vector<uint32_t>>& occs) {
  for (size_t b = 0; b < s->bricks.size(); b++) {
    const Brick& g = s->bricks[b];
    if (home_cells(g) < 2u) continue;
    const Replica& R = s->rep[g.rep];
    const std::vector<uint32_t>& occ = occs[g.rep];
    const int nx = (int)R.cells[0], ny = (int)R.cells[1], nz = (int)R.cells[2];
    const int x0 = (int)(g.cell0 % R.cells[0]), y0 = (int)((g.cell0 / R.cells[0]) % R.cells[1]);
    const int z0 = (int)(g.cell0 / (R.cells[0] * R.cells[1]));
    uint64_t total = 0;
    for (int dz = 0; dz < (int)g.gz + 2; dz++)
      for (int dy = 0; dy < (int)g.gy + 2; dy++)
        for (int dx = 0; dx < (int)g.gx + 2; dx++) {
          const int cx = (x0 - 1 + dx + nx) % nx, cy = (y0 - 1 + dy + ny) % ny, cz = (z0 - 1 + dz + nz) % nz;
          total += occ[(size_t)(cz * ny + cy) * nx + cx];
        }
    if (total > (b < s->n_dense ? (uint64_t)kStageMax : (uint64_t)Sparse::kStageMax)) return true;
  }
  return false;
}

// Builds a handle over `replicas` systems.
// `saved` (from a state file) fixes what derive() takes from the initial
// particles -- list capacity and fixed-point scales -- to the saved values.
int create(const vlj_particle* all, const uint64_t* offsets, const vlj_settings* settings,
           size_t replicas, vlj_sim** out, const Replica* saved = nullptr) {
  if (!out) return fail(nullptr, VLJ_E_INVALID, "out must not be NULL");
  *out = nullptr;
  if (replicas == 0 || !settings || !offsets)
    return fail(nullptr, VLJ_E_INVALID, "at least one replica with settings is needed");
  uint64_t total = offsets[replicas];
  if (offsets[0] != 0) return fail(nullptr, VLJ_E_INVALID, "offsets must start at 0");
  for (size_t r = 0; r < replicas; r++)
    if (offsets[r + 1] < offsets[r]) return fail(nullptr, VLJ_E_INVALID, "offsets must not decrease");
  if (total >= (1ull << 31)) return fail(nullptr, VLJ_E_INVALID, "at most 2^31 - 1 particles");
  if (total && !all) return fail(nullptr, VLJ_E_INVALID, "particles must not be NULL");
  for (size_t r = 1; r < replicas; r++)
    if (settings[r].precision != settings[0].precision ||
        (settings[r].shifted_force != 0) != (settings[0].shifted_force != 0))
      return fail(nullptr, VLJ_E_INVALID,
                  "replicas of one batch share the pair-force precision and cutoff style");

  vlj_sim* s = new vlj_sim();
  int rc = VLJ_OK;
  s->rep.resize(replicas);
  s->host.resize(replicas);
  s->precision = (int)settings[0].precision;
  s->shifted = settings[0].shifted_force != 0;
  uint32_t atom_off = 0, cell_off = 0, scan_off = 0;
  uint64_t list_off = 0, prof_off = 0;
  std::vector<std::vector<uint32_t>> occs(replicas);  // each replica's initial cell occupancy
  for (size_t r = 0; r < replicas; r++) {
    Replica& R = s->rep[r];
    std::string err;
    size_t n = (size_t)(offsets[r + 1] - offsets[r]);
    std::vector<uint32_t>& occ = occs[r];
    rc = derive(all + offsets[r], n, settings[r], R, s->host[r], occ, err);
    if (rc != VLJ_OK) {
      delete s;
      return fail(nullptr, rc, replicas > 1 ? "replica " + std::to_string(r) + ": " + err : err);
    }
    if (saved) {
      R.capacity = saved[r].capacity;
      R.energy_fixed = saved[r].energy_fixed;
      R.momentum_fixed = saved[r].momentum_fixed;
      s->host[r].energy_scale = R.energy_fixed / kTwo32;
      s->host[r].momentum_scale = R.momentum_fixed / kTwo32;
    }
    R.index = (uint32_t)r;
    R.atom_off = atom_off;
    R.cell_off = cell_off;
    R.scan_block_off = scan_off;
    const bool dense = (double)R.count >= kSortDense * (double)R.total_cells;
    R.scan_cells = dense ? kScanCells : kScanCellsSparse;
    R.scan_blocks = (R.total_cells + R.scan_cells - 1) / R.scan_cells;
    R.list_off = list_off;
    R.prof_off = (uint32_t)prof_off;
    if (R.prof_bins) prof_off += (uint64_t)R.prof_bins * kProfileSums + 3u;
    if (prof_off >= (1ull << 31)) {
      delete s;
      return fail(nullptr, VLJ_E_INVALID, "the profiles of the batch need too many sums");
    }
    R.atom_block_off = (uint32_t)s->atom_blocks.size();
    R.atom_blocks = (R.count + kBlock - 1) / kBlock;
    choose_bricks(R, occ, plan_fill(s), plan_sparse_fill(s), s->host[r].brick_x, s->host[r].brick_y,
                  s->host[r].brick_z, s->host[r].sparse);
    for (uint32_t b = 0; b < R.count; b += kBlock)
      s->atom_blocks.push_back({(uint32_t)r, atom_off + b, atom_off + std::min(R.count, b + kBlock), 0});
    for (uint32_t b = 0; b < R.total_cells; b += R.scan_cells)
      s->scan_blocks.push_back(
          {(uint32_t)r, cell_off + b, cell_off + std::min(R.total_cells, b + R.scan_cells), 0});
    const uint32_t per = dense ? kSortCells : kSortCellsSparse;
    for (uint32_t b = 0; b < R.total_cells; b += per)
      s->sort_blocks.push_back({(uint32_t)r, cell_off + b, cell_off + std::min(R.total_cells, b + per), 0});
    atom_off += R.count;
    cell_off += R.total_cells;
    scan_off += R.scan_blocks;
    list_off += (uint64_t)list_octets(R.capacity) * 8u * ((R.count + kListStride - 1) / kListStride) * kListStride;
    if (R.thermostat == 1) s->any_langevin = true;
    if (R.thermostat == 2) s->any_nose_hoover = true;
  }
  build_table(s);
  s->n_total = atom_off;
  s->cells_total = cell_off;
  s->list_total = list_off;
  s->prof_total = prof_off;
  if (list_off >= (1ull << 40)) {
    delete s;
    return fail(nullptr, VLJ_E_INVALID, "neighbour lists would not fit in device memory");
  }
  // Arguments are validated first, so their errors do not depend on a device.
  rc = check_device(s);
  if (rc != VLJ_OK) {
    std::string msg = s->error;
    delete s;
    return fail(nullptr, rc, msg);
  }

  rc = VLJ_OK;
  auto A = [&](auto** p, size_t count) {
    if (rc == VLJ_OK) rc = dev_alloc(s, p, count);
  };
  if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess) {
    delete s;
    return fail(nullptr, VLJ_E_HIP, "hipStreamCreate failed");
  }
  if (s->n_dense > 0 && s->n_dense < s->bricks.size()) {  // both layouts: a side stream (default priority)
    if (hipStreamCreateWithFlags(&s->side.stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&s->side.fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&s->side.join, hipEventDisableTiming) != hipSuccess) {
      free_all(s);
      delete s;
      return fail(nullptr, VLJ_E_HIP, "hipStreamCreate failed");
    }
  }
  Buffers& B = s->B;
  const size_t N = s->n_total, C = s->cells_total;
  Replica* rep_d = nullptr;
  Range *ab = nullptr, *sb = nullptr, *ob = nullptr;
  Brick* gb = nullptr;
  A(&rep_d, replicas);
  A(&B.ctl, replicas);
  A(&ab, s->atom_blocks.size());
  A(&sb, s->scan_blocks.size());
  A(&ob, s->sort_blocks.size());
  A(&gb, s->bricks.size());
  A(&B.cut, s->bricks.size());
  A(&B.parts, s->bricks.size() * kHomeCells);
  A(&B.cut_groups, (s->bricks.size() + kCutGroup - 1) / kCutGroup);
  A(&B.bmax, s->atom_blocks.size());
  A(&B.amax, 2 * replicas * kAheadSlots);
  A(&B.kepart, s->atom_blocks.size());
  A(&B.hexpart, s->atom_blocks.size() * kHexSums);
  A(&B.hexstate, replicas * kHexState);
  double** d[] = {&B.x, &B.y, &B.z, &B.vx, &B.vy, &B.vz, &B.ax, &B.ay, &B.az,
                  &B.x2, &B.y2, &B.z2, &B.vx2, &B.vy2, &B.vz2, &B.ax2, &B.ay2, &B.az2,
                  &B.xb, &B.yb, &B.zb};
  for (double** p : d) A(p, N);
  uint32_t** u[] = {&B.orig, &B.orig2, &B.cell_of, &B.cell_of2, &B.unsorted, &B.sorted, &B.nbr_count};
  for (uint32_t** p : u) A(p, N);
  uint32_t** c[] = {&B.cell_count, &B.cell_cursor, &B.cell_scan, &B.cell_start, &B.cell_n};
  for (uint32_t** p : c) A(p, C);
  A(&B.bsum, s->scan_blocks.size());
  A(&B.bprefix, s->scan_blocks.size());
  A(&B.nbr, s->list_total);
  A(&B.slots, replicas * kSlots);
  A(&B.profile, std::max<uint64_t>(s->prof_total, 1));
  A(&B.tickets, 4);
  A(&B.flags, 3);
  A(&s->fixed_scratch, kSlots);
  if (rc != VLJ_OK) {
    std::string msg = s->error;
    free_all(s);
    delete s;
    return fail(nullptr, rc, msg);
  }
  B.rep = rep_d;
  B.atom_blocks = ab;
  B.scan_blocks = sb;
  B.sort_blocks = ob;
  B.bricks = gb;
  B.n_bricks = (uint32_t)s->bricks.size();
  B.n_dense = s->n_dense;
  B.parts_on = g_plan.parts >= 0 ? (uint32_t)g_plan.parts : (any_overflow(s, occs) ? 1u : 0u);
  B.n_atom_blocks = (uint32_t)s->atom_blocks.size();
  B.n_scan_blocks = (uint32_t)s->scan_blocks.size();
  B.n_sort_blocks = (uint32_t)s->sort_blocks.size();
  B.n_replicas = (uint32_t)replicas;

  // Host-side initial state: wrapped positions, slot = original index.
  std::vector<double> h(N);
  std::vector<uint32_t> hu(N);
  auto upload = [&](double* dst) { return hipMemcpy(dst, h.data(), N * sizeof(double), hipMemcpyHostToDevice); };
  hipError_t e = hipSuccess;
  for (int k = 0; k < 3 && e == hipSuccess; k++) {
    for (size_t r = 0; r < replicas; r++)
      for (uint32_t i = 0; i < s->rep[r].count; i++)
        h[s->rep[r].atom_off + i] = wrap_initial(all[offsets[r] + i].pos[k], s->rep[r].box[k]);
    e = upload(k == 0 ? B.x : (k == 1 ? B.y : B.z));
    if (e != hipSuccess) break;
    for (size_t r = 0; r < replicas; r++)
      for (uint32_t i = 0; i < s->rep[r].count; i++) h[s->rep[r].atom_off + i] = all[offsets[r] + i].vel[k];
    e = upload(k == 0 ? B.vx : (k == 1 ? B.vy : B.vz));
  }
  for (size_t r = 0; r < replicas; r++)
    for (uint32_t i = 0; i < s->rep[r].count; i++) hu[s->rep[r].atom_off + i] = i;
  std::vector<Control> ctl(replicas);
  std::memset(ctl.data(), 0, ctl.size() * sizeof(Control));
  if (e == hipSuccess) e = hipMemcpy(B.orig, hu.data(), N * 4, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(rep_d, s->rep.data(), replicas * sizeof(Replica), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(B.ctl, ctl.data(), replicas * sizeof(Control), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(ab, s->atom_blocks.data(), s->atom_blocks.size() * sizeof(Range), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(sb, s->scan_blocks.data(), s->scan_blocks.size() * sizeof(Range), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(ob, s->sort_blocks.data(), s->sort_blocks.size() * sizeof(Range), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(gb, s->bricks.data(), s->bricks.size() * sizeof(Brick), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemset(B.bmax, 0, std::max<size_t>(s->atom_blocks.size(), 1) * 8);
  if (e == hipSuccess) e = hipMemset(B.amax, 0, 2 * replicas * kAheadSlots * 8);
  B.amax_next = B.amax + replicas * kAheadSlots;
  if (e == hipSuccess) e = hipMemset(B.cut, 0, std::max<size_t>(s->bricks.size(), 1) * 4);
  if (e == hipSuccess) e = hipMemset(B.cut_groups, 0, std::max<size_t>((s->bricks.size() + kCutGroup - 1) / kCutGroup, 1) * 4);
  // Zeroed once here; afterwards the sort kernel leaves them zero.
  if (e == hipSuccess) e = hipMemset(B.cell_count, 0, std::max<size_t>(C, 1) * 4);
  if (e == hipSuccess) e = hipMemset(B.cell_cursor, 0, std::max<size_t>(C, 1) * 4);
  if (e == hipSuccess) e = hipMemset(B.tickets, 0, 16);
  if (e == hipSuccess) e = hipMemset(B.flags, 0, 12);
  B.rebuilding = B.flags;
  B.rebuilding_next = B.flags + 2;
  if (e == hipSuccess) e = hipMemset(B.ax, 0, std::max<size_t>(N, 1) * 8);
  if (e == hipSuccess) e = hipMemset(B.ay, 0, std::max<size_t>(N, 1) * 8);
  if (e == hipSuccess) e = hipMemset(B.az, 0, std::max<size_t>(N, 1) * 8);
  if (e == hipSuccess) e = hipMemset(B.nbr_count, 0, std::max<size_t>(N, 1) * 4);
  if (e == hipSuccess) e = hipMemset(B.profile, 0, std::max<uint64_t>(s->prof_total, 1) * 8);
  for (auto& r : s->ring) {
    if (e == hipSuccess) e = hipHostMalloc((void**)&r.slots, replicas * kSlots * sizeof(long long), hipHostMallocDefault);
    if (e == hipSuccess) e = hipHostMalloc((void**)&r.ctl, replicas * sizeof(Control), hipHostMallocDefault);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&r.done, hipEventDisableTiming);
  }
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e != hipSuccess) {
    std::string msg = std::string("initial upload: ") + hipGetErrorString(e);
    free_all(s);
    delete s;
    return fail(nullptr, VLJ_E_HIP, msg);
  }
  *out = s;
  return VLJ_OK;
}

// The captured steps for the present orientation of the position sets (the
// drift-ahead form swaps them every step: one pair of graphs each way).
int ensure_graphs(vlj_sim* s) {
  const int o = s->orient;
  int rc = VLJ_OK;
  if (!s->graph1[o]) rc = capture(s, 1, &s->graph1[o]);
  if (rc == VLJ_OK && !s->graphN[o]) rc = capture(s, kGraphSteps, &s->graphN[o]);
  return rc;
}

int enqueue_steps(vlj_sim* s, uint64_t steps) {
  int rc = prime_if_needed(s);
  if (rc != VLJ_OK) return rc;
  if (steps == 0 || s->n_total == 0) {
    s->steps += steps;
    return VLJ_OK;
  }
  if (tune_wanted(s) && steps >= kTuneAutoSteps) {
    uint64_t taken = 0;
    rc = tune(s, kTuneSteps, &taken);
    if (rc != VLJ_OK) return rc;
    steps -= taken;
  }
  rc = check_bricks(s);
  if (rc != VLJ_OK) return rc;
  uint64_t left = steps;
  // The captured drift-ahead steps start from drifted positions: the first
  // step after a change of form drifts on its own launch.
  if (drift_ahead(s) && !s->ahead && left > 0) {
    rc = enqueue_step(s);
    if (rc != VLJ_OK) return rc;
    left--;
  }
  while (left >= kGraphSteps) {
    rc = ensure_graphs(s);
    if (rc != VLJ_OK) return rc;
    HIP_TRY(s, hipGraphLaunch(s->graphN[s->orient], s->stream));  // an even count: the sets end as they began
    left -= kGraphSteps;
  }
  while (left > 0) {
    rc = ensure_graphs(s);
    if (rc != VLJ_OK) return rc;
    HIP_TRY(s, hipGraphLaunch(s->graph1[s->orient], s->stream));
    if (drift_ahead(s)) flip_positions(s);
    left--;
  }
  // what the replayed steps leave: drifted positions in B.x2 after
  // drift-ahead steps, none after the others (which use B.x2 as scratch)
  if (steps > 0) s->ahead = drift_ahead(s);
  s->steps += steps;
  return check_bricks(s);
}

void decode(const vlj_sim* s, uint32_t r, const long long* slots, const Control& c, uint64_t step,
            vlj_measurement* m) {
  const Replica& R = s->rep[r];
  std::memset(m, 0, sizeof(*m));
  m->step = step;
  m->time = (double)step * R.dt;
  m->kinetic = (double)slots[0] / R.energy_fixed;
  m->potential = (double)slots[1] / R.energy_fixed;
  m->virial = (double)slots[2] / R.energy_fixed;
  for (int k = 0; k < 3; k++) m->momentum[k] = (double)slots[3 + k] / R.momentum_fixed;
  m->count = R.count;
  m->volume = R.box[0] * R.box[1] * R.box[2];
  m->list_rebuilds = c.rebuilds;
  m->list_overflow = c.overflow;
  // lennard_jones.rs:1235-1241
  m->thermostat_energy = R.thermostat == 2 ? 0.5 * R.nh_q * c.xi * c.xi + R.nh_gkt * c.eta : 0.0;
  m->condensed = (uint32_t)slots[6];
  m->pairs_within_cutoff = (uint64_t)slots[7] / 2u;
  for (int k = 0; k < kSlots; k++) m->raw[k] = slots[k];
}

int enqueue_measure(vlj_sim* s, int slot) {
  int rc = prime_if_needed(s);
  if (rc != VLJ_OK) return rc;
  Ring& r = s->ring[slot];
  const size_t R = s->rep.size();
  HIP_TRY(s, launch_measure(s->B, s->precision, s->shifted, s->stream));
  HIP_TRY(s, launch_render(s->B, s->stream));  // the drawn values, as a measurement leaves them
  HIP_TRY(s, hipMemcpyAsync(r.slots, s->B.slots, R * kSlots * sizeof(long long), hipMemcpyDeviceToHost, s->stream));
  HIP_TRY(s, hipMemcpyAsync(r.ctl, s->B.ctl, R * sizeof(Control), hipMemcpyDeviceToHost, s->stream));
  HIP_TRY(s, hipEventRecord(r.done, s->stream));
  r.step = s->steps;
  r.in_flight = true;
  return VLJ_OK;
}

int free_slot(const vlj_sim* s) {
  for (int k = 0; k < kRing; k++)
    if (!s->ring[k].in_flight) return k;
  return -1;
}

int read_dev(vlj_sim* s, void* dst, const void* src, size_t bytes) {
  HIP_TRY(s, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s->stream));
  HIP_TRY(s, hipStreamSynchronize(s->stream));
  return VLJ_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------

extern "C" {

int vlj_create(const vlj_particle* particles, size_t count, const vlj_settings* settings, vlj_sim** out) {
  if (!settings) return fail(nullptr, VLJ_E_INVALID, "settings must not be NULL");
  uint64_t offsets[2] = {0, (uint64_t)count};
  return create(particles, offsets, settings, 1, out);
}

void vlj_destroy(vlj_sim* sim) {
  if (!sim) return;
  free_all(sim);
  delete sim;
}

int vlj_step(vlj_sim* sim, uint32_t steps) {
  if (!sim) return fail(nullptr, VLJ_E_INVALID, "sim must not be NULL");
  return enqueue_steps(sim, steps);
}

int vlj_run(vlj_sim* sim, uint64_t steps) {
  if (!sim) return fail(nullptr, VLJ_E_INVALID, "sim must not be NULL");
  int rc = enqueue_steps(sim, steps);
  if (rc != VLJ_OK) return rc;
  HIP_TRY(sim, hipStreamSynchronize(sim->stream));
  return VLJ_OK;
}

int vlj_measure(vlj_sim* sim, int32_t* accepted) {
  if (!sim || !accepted) return fail(sim, VLJ_E_INVALID, "sim and accepted must not be NULL");
  SINGLE_ONLY(sim, "measure");
  int slot = free_slot(sim);
  if (slot < 0) {
    *accepted = 0;
    return VLJ_OK;
  }
  int rc = enqueue_measure(sim, slot);
  *accepted = rc == VLJ_OK ? 1 : 0;
  return rc;
}

int vlj_poll(vlj_sim* sim, vlj_measurement* out, size_t capacity, size_t* written) {
  if (!sim || !written || (capacity && !out)) return fail(sim, VLJ_E_INVALID, "bad arguments to vlj_poll");
  SINGLE_ONLY(sim, "poll");
  *written = 0;
  std::vector<std::pair<uint64_t, int>> done;
  for (int k = 0; k < kRing; k++) {
    Ring& r = sim->ring[k];
    if (!r.in_flight) continue;
    hipError_t e = hipEventQuery(r.done);
    if (e == hipErrorNotReady) continue;
    if (e != hipSuccess) return fail(sim, VLJ_E_HIP, std::string("hipEventQuery: ") + hipGetErrorString(e));
    done.push_back({r.step, k});
  }
  std::sort(done.begin(), done.end());
  for (auto& d : done) {
    if (*written >= capacity) break;
    Ring& r = sim->ring[d.second];
    decode(sim, 0, r.slots, r.ctl[0], r.step, out + *written);
    r.in_flight = false;
    (*written)++;
  }
  return VLJ_OK;
}

int vlj_measure_now(vlj_sim* sim, vlj_measurement* out) {
  if (!sim || !out) return fail(sim, VLJ_E_INVALID, "sim and out must not be NULL");
  SINGLE_ONLY(sim, "measure_now");
  int slot = free_slot(sim);
  if (slot < 0) return fail(sim, VLJ_E_STATE, "a measurement is already in flight");
  int rc = enqueue_measure(sim, slot);
  if (rc != VLJ_OK) return rc;
  HIP_TRY(sim, hipEventSynchronize(sim->ring[slot].done));
  Ring& r = sim->ring[slot];
  decode(sim, 0, r.slots, r.ctl[0], r.step, out);
  r.in_flight = false;
  return VLJ_OK;
}

static int read_state(vlj_sim* sim, std::vector<double>* pos, std::vector<double>* vel,
                      std::vector<double>* acc, std::vector<uint32_t>* orig) {
  const size_t N = sim->n_total;
  const Buffers& B = sim->B;
  int rc = VLJ_OK;
  double* const src[3][3] = {{B.x, B.y, B.z}, {B.vx, B.vy, B.vz}, {B.ax, B.ay, B.az}};
  std::vector<double>* dst[3] = {pos, vel, acc};
  std::vector<double> tmp(N);
  for (int g = 0; g < 3; g++) {
    if (!dst[g]) continue;
    dst[g]->assign(3 * N, 0.0);
    for (int k = 0; k < 3 && rc == VLJ_OK; k++) {
      rc = read_dev(sim, tmp.data(), src[g][k], N * sizeof(double));
      for (size_t i = 0; i < N; i++) (*dst[g])[3 * i + k] = tmp[i];
    }
  }
  if (orig && rc == VLJ_OK) {
    orig->assign(N, 0);
    rc = read_dev(sim, orig->data(), B.orig, N * 4);
  }
  return rc;
}

int vlj_batch_read_particles(vlj_batch* sim, vlj_particle* out, size_t count) {
  if (!sim || (count && !out)) return fail(sim, VLJ_E_INVALID, "bad arguments to read_particles");
  if (count != sim->n_total) return fail(sim, VLJ_E_INVALID, "count must equal the number of particles");
  std::vector<double> pos, vel;
  std::vector<uint32_t> orig;
  int rc = read_state(sim, &pos, &vel, nullptr, &orig);
  if (rc != VLJ_OK) return rc;
  for (size_t r = 0; r < sim->rep.size(); r++) {
    const Replica& R = sim->rep[r];
    for (uint32_t t = 0; t < R.count; t++) {
      const size_t slot = R.atom_off + t;
      vlj_particle& p = out[R.atom_off + orig[slot]];
      for (int k = 0; k < 3; k++) {
        p.pos[k] = wrap_one(pos[3 * slot + k], R.box[k]);  // R1: wrapped on the way out
        p.vel[k] = vel[3 * slot + k];
      }
    }
  }
  return VLJ_OK;
}

int vlj_read_particles(vlj_sim* sim, vlj_particle* out, size_t count) {
  return vlj_batch_read_particles(sim, out, count);
}

uint32_t vlj_count(const vlj_sim* sim) { return sim ? sim->n_total : 0u; }

int vlj_box_size(const vlj_sim* sim, double out[3]) {
  if (!sim || !out) return VLJ_E_INVALID;
  SINGLE_ONLY(const_cast<vlj_sim*>(sim), "box_size");
  for (int k = 0; k < 3; k++) out[k] = sim->rep[0].box[k];
  return VLJ_OK;
}

double vlj_volume(const vlj_sim* sim) {
  if (!sim) return 0.0;
  if (sim->rep.size() > 1) return std::nan("");
  return sim->rep[0].box[0] * sim->rep[0].box[1] * sim->rep[0].box[2];
}

uint64_t vlj_steps(const vlj_sim* sim) { return sim ? sim->steps : 0u; }

double vlj_time(const vlj_sim* sim) {
  if (!sim) return 0.0;
  if (sim->rep.size() > 1) return std::nan("");
  return (double)sim->steps * sim->rep[0].dt;
}

int vlj_target_temperature(const vlj_sim* sim, double* temperature, int32_t* has) {
  if (!sim || !temperature || !has) return VLJ_E_INVALID;
  SINGLE_ONLY(const_cast<vlj_sim*>(sim), "target_temperature");
  const vlj_thermostat& t = sim->host[0].settings.thermostat;
  *has = t.kind != VLJ_THERMOSTAT_NONE ? 1 : 0;
  *temperature = *has ? t.temperature : 0.0;
  return VLJ_OK;
}

static int set_temperature(vlj_sim* sim, size_t replica, double temperature) {
  if (replica >= sim->rep.size()) return fail(sim, VLJ_E_INVALID, "no such replica");
  vlj_thermostat& t = sim->host[replica].settings.thermostat;
  if (t.kind == VLJ_THERMOSTAT_NONE) return fail(sim, VLJ_E_STATE, "set_temperature needs a thermostat");
  if (!(temperature >= 0.0)) return fail(sim, VLJ_E_INVALID, "temperature must be non-negative");
  t.temperature = temperature;
  Replica& R = sim->rep[replica];
  if (t.kind == VLJ_THERMOSTAT_LANGEVIN) {
    const double c1 = std::exp(-t.friction * R.dt);
    R.c2 = std::sqrt((1.0 - c1 * c1) * temperature);
  } else {
    // Nose-Hoover keeps its period, so Q scales with the temperature (rs:912-920)
    const double dof = 3.0 * (double)R.count - 3.0;
    R.nh_gkt = dof * temperature;
    R.nh_q = R.nh_gkt * R.nh_period * R.nh_period / (4.0 * (kPi * kPi));
  }
  HIP_TRY(sim, hipMemcpyAsync((void*)&sim->B.rep[replica], &R, sizeof(Replica), hipMemcpyHostToDevice,
                              sim->stream));
  HIP_TRY(sim, hipStreamSynchronize(sim->stream));
  sim->ahead = false;  // a parameter the next drift uses changed
  return VLJ_OK;
}

int vlj_set_temperature(vlj_sim* sim, double temperature) {
  if (!sim) return fail(nullptr, VLJ_E_INVALID, "sim must not be NULL");
  SINGLE_ONLY(sim, "set_temperature");
  return set_temperature(sim, 0, temperature);
}

int vlj_batch_set_temperature(vlj_batch* batch, size_t replica, double temperature) {
  if (!batch) return fail(nullptr, VLJ_E_INVALID, "batch must not be NULL");
  return set_temperature(batch, replica, temperature);
}

int vlj_set_condensed_neighbours(vlj_sim* sim, uint32_t neighbours) {
  if (!sim) return fail(nullptr, VLJ_E_INVALID, "sim must not be NULL");
  for (size_t r = 0; r < sim->rep.size(); r++) sim->rep[r].near_count = neighbours;
  HIP_TRY(sim, hipMemcpyAsync((void*)sim->B.rep, sim->rep.data(), sim->rep.size() * sizeof(Replica),
                              hipMemcpyHostToDevice, sim->stream));
  HIP_TRY(sim, hipStreamSynchronize(sim->stream));
  return VLJ_OK;
}

static bool any_profiles(const vlj_sim* s) { return s->prof_total != 0; }

// lennard_jones.rs:1167-1183: prime if needed, then one sample added to the
// running sums of every replica with profile slabs. Asynchronous.
int vlj_sample_profile(vlj_sim* sim) {
  if (!sim) return fail(nullptr, VLJ_E_INVALID, "sim must not be NULL");
  if (!any_profiles(sim)) return fail(sim, VLJ_E_STATE, "profiles need Settings::profile_bins above zero");
  int rc = prime_if_needed(sim);
  if (rc != VLJ_OK) return rc;
  bool wide = true;
  for (const Replica& R : sim->rep) wide = wide && (R.prof_bins == 0 || R.bin_width >= R.rc);
  HIP_TRY(sim, launch_profile(sim->B, sim->shifted, wide, sim->stream));
  for (auto& h : sim->host)
    if (h.settings.profile_bins) h.prof_samples++;
  return VLJ_OK;
}

// lennard_jones.rs:1186-1191. Without profile slabs there is nothing to clear.
int vlj_clear_profile(vlj_sim* sim) {
  if (!sim) return fail(nullptr, VLJ_E_INVALID, "sim must not be NULL");
  if (!any_profiles(sim)) return VLJ_OK;
  HIP_TRY(sim, launch_zero(sim->B.profile, (uint32_t)sim->prof_total, sim->stream));
  for (auto& h : sim->host) h.prof_samples = 0;
  return VLJ_OK;
}

static int read_profile_sums(vlj_sim* sim, size_t replica, std::vector<long long>& raw, uint64_t* samples) {
  if (!sim) return fail(nullptr, VLJ_E_INVALID, "sim must not be NULL");
  if (replica >= sim->rep.size()) return fail(sim, VLJ_E_INVALID, "no such replica");
  const Replica& R = sim->rep[replica];
  if (R.prof_bins == 0) return fail(sim, VLJ_E_STATE, "profiles need Settings::profile_bins above zero");
  raw.assign((size_t)R.prof_bins * kProfileSums + 3u, 0);
  int rc = read_dev(sim, raw.data(), sim->B.profile + R.prof_off, raw.size() * sizeof(long long));
  if (rc != VLJ_OK) return rc;
  if (samples) *samples = sim->host[replica].prof_samples;
  return VLJ_OK;
}

static int copy_raw(vlj_sim* sim, const std::vector<long long>& raw, int64_t* out, size_t capacity) {
  if (!out || capacity < raw.size())
    return fail(sim, VLJ_E_INVALID, "the profile needs profile_bins * 8 + 3 values");
  std::memcpy(out, raw.data(), raw.size() * sizeof(long long));
  return VLJ_OK;
}

// lennard_jones.rs:1297-1325: the sums decoded from fixed point and divided
// by the sample count (NaN with no samples, as 0 / 0 there).
int vlj_read_profile(vlj_sim* sim, double* out, size_t capacity, uint64_t* samples) {
  if (!sim) return fail(nullptr, VLJ_E_INVALID, "sim must not be NULL");
  SINGLE_ONLY(sim, "read_profile");
  std::vector<long long> raw;
  uint64_t n = 0;
  int rc = read_profile_sums(sim, 0, raw, &n);
  if (rc != VLJ_OK) return rc;
  if (!out || capacity < raw.size())
    return fail(sim, VLJ_E_INVALID, "the profile needs profile_bins * 8 + 3 values");
  for (size_t k = 0; k < raw.size(); k++) out[k] = ((double)raw[k] / kTwo32) / (double)n;
  if (samples) *samples = n;
  return VLJ_OK;
}

int vlj_read_profile_raw(vlj_sim* sim, int64_t* raw_out, size_t capacity, uint64_t* samples) {
  if (!sim) return fail(nullptr, VLJ_E_INVALID, "sim must not be NULL");
  SINGLE_ONLY(sim, "read_profile_raw");
  std::vector<long long> raw;
  int rc = read_profile_sums(sim, 0, raw, samples);
  return rc != VLJ_OK ? rc : copy_raw(sim, raw, raw_out, capacity);
}

int vlj_batch_read_profile_raw(vlj_batch* sim, size_t replica, int64_t* raw_out, size_t capacity,
                               uint64_t* samples) {
  std::vector<long long> raw;
  int rc = read_profile_sums(sim, replica, raw, samples);
  return rc != VLJ_OK ? rc : copy_raw(sim, raw, raw_out, capacity);
}

// lennard_jones.rs:927-943: constant energy only; the captured step gains
// or loses the exchange kernels.
int vlj_set_heat_exchange(vlj_sim* sim, int32_t on, const double hot[2], const double cold[2], double energy) {
  if (!sim) return fail(nullptr, VLJ_E_INVALID, "sim must not be NULL");
  SINGLE_ONLY(sim, "set_heat_exchange");
  Replica& R = sim->rep[0];
  if (R.thermostat != 0u)
    return fail(sim, VLJ_E_STATE, "heat exchange runs at constant energy, without a thermostat");
  if (on && (!hot || !cold || !std::isfinite(energy)))
    return fail(sim, VLJ_E_INVALID, "heat exchange needs hot and cold slabs and a finite energy");
  R.hex_on = on ? 1u : 0u;
  for (int k = 0; k < 2; k++) {
    R.hex_slabs[k] = on ? hot[k] : 0.0;
    R.hex_slabs[2 + k] = on ? cold[k] : 0.0;
  }
  R.hex_energy = on ? energy : 0.0;
  bool any = false;
  for (const Replica& r : sim->rep) any = any || r.hex_on;
  HIP_TRY(sim, hipMemcpyAsync((void*)&sim->B.rep[0], &R, sizeof(Replica), hipMemcpyHostToDevice, sim->stream));
  HIP_TRY(sim, hipStreamSynchronize(sim->stream));
  if (any != sim->any_heat_exchange) {
    drop_graphs(sim);
    sim->any_heat_exchange = any;
  }
  sim->ahead = false;
  return VLJ_OK;
}

// lennard_jones.rs:945-951: energy moved per step over 2 A dt.
int vlj_heat_flux(const vlj_sim* sim, double* flux, int32_t* has) {
  if (!sim || !flux || !has) return fail(nullptr, VLJ_E_INVALID, "bad arguments to vlj_heat_flux");
  SINGLE_ONLY(const_cast<vlj_sim*>(sim), "heat_flux");
  const Replica& R = sim->rep[0];
  *has = R.hex_on ? 1 : 0;
  *flux = R.hex_on ? R.hex_energy / (2.0 * R.box[0] * R.box[1] * R.dt) : 0.0;
  return VLJ_OK;
}

// lennard_jones.rs:953-973: room for `capacity` samples, any recorded dropped.
int vlj_record_heat_current(vlj_sim* sim, uint32_t capacity) {
  if (!sim) return fail(nullptr, VLJ_E_INVALID, "sim must not be NULL");
  SINGLE_ONLY(sim, "record_heat_current");
  HIP_TRY(sim, hipStreamSynchronize(sim->stream));
  if (sim->B.hc) HIP_TRY(sim, hipFree(sim->B.hc));
  sim->B.hc = nullptr;
  if (capacity) {
    const size_t words = (size_t)capacity * 3;
    HIP_TRY(sim, hipMalloc((void**)&sim->B.hc, words * sizeof(long long)));
    HIP_TRY(sim, launch_zero(sim->B.hc, (uint32_t)words, sim->stream));
  }
  sim->rep[0].hc_cap = capacity;
  sim->hc_samples = 0;
  HIP_TRY(sim, hipMemcpyAsync((void*)&sim->B.rep[0], &sim->rep[0], sizeof(Replica), hipMemcpyHostToDevice,
                              sim->stream));
  HIP_TRY(sim, hipStreamSynchronize(sim->stream));
  drop_graphs(sim);  // the graphs hold the old buffer pointers
  return VLJ_OK;
}

// lennard_jones.rs:975-996: one sample, asynchronous.
int vlj_sample_heat_current(vlj_sim* sim) {
  if (!sim) return fail(nullptr, VLJ_E_INVALID, "sim must not be NULL");
  SINGLE_ONLY(sim, "sample_heat_current");
  if (!sim->B.hc) return fail(sim, VLJ_E_STATE, "record_heat_current first");
  if (sim->hc_samples >= sim->rep[0].hc_cap) return fail(sim, VLJ_E_STATE, "heat current series full");
  int rc = prime_if_needed(sim);
  if (rc != VLJ_OK) return rc;
  HIP_TRY(sim, launch_heat_current(sim->B, sim->shifted, sim->hc_samples, sim->stream));
  sim->hc_samples++;
  return VLJ_OK;
}

static int read_series(vlj_sim* sim, std::vector<long long>& raw, uint32_t* samples) {
  if (!sim) return fail(nullptr, VLJ_E_INVALID, "sim must not be NULL");
  SINGLE_ONLY(sim, "read_heat_current");
  if (!sim->B.hc) return fail(sim, VLJ_E_STATE, "record_heat_current first");
  raw.assign((size_t)sim->hc_samples * 3, 0);
  if (!raw.empty()) {
    int rc = read_dev(sim, raw.data(), sim->B.hc, raw.size() * sizeof(long long));
    if (rc != VLJ_OK) return rc;
  }
  if (samples) *samples = sim->hc_samples;
  return VLJ_OK;
}

// lennard_jones.rs:998-1012: the recorded samples, x y z each.
int vlj_read_heat_current(vlj_sim* sim, double* out, size_t capacity, uint32_t* samples) {
  std::vector<long long> raw;
  int rc = read_series(sim, raw, samples);
  if (rc != VLJ_OK || !out) return rc;  // out NULL: the number of samples only
  if (raw.size() > 3 * capacity)
    return fail(sim, VLJ_E_INVALID, "the output holds fewer samples than recorded");
  for (size_t k = 0; k < raw.size(); k++) out[k] = (double)raw[k] / kTwo32;
  return VLJ_OK;
}

int vlj_read_heat_current_raw(vlj_sim* sim, int64_t* out, size_t capacity, uint32_t* samples) {
  std::vector<long long> raw;
  int rc = read_series(sim, raw, samples);
  if (rc != VLJ_OK || !out) return rc;  // out NULL: the number of samples only
  if (raw.size() > 3 * capacity)
    return fail(sim, VLJ_E_INVALID, "the output holds fewer samples than recorded");
  if (!raw.empty()) std::memcpy(out, raw.data(), raw.size() * sizeof(long long));
  return VLJ_OK;
}

const char* vlj_last_error(const vlj_sim* sim) { return sim ? sim->error.c_str() : g_error.c_str(); }

int vlj_batch_create(const vlj_particle* all, const uint64_t* offsets, const vlj_settings* settings,
                     size_t replicas, vlj_batch** out) {
  return create(all, offsets, settings, replicas, out);
}

int vlj_batch_run(vlj_batch* batch, uint64_t steps) { return vlj_run(batch, steps); }

int vlj_batch_measure_now(vlj_batch* sim, vlj_measurement* per_replica, size_t replicas) {
  if (!sim || !per_replica) return fail(sim, VLJ_E_INVALID, "bad arguments to batch_measure_now");
  if (replicas != sim->rep.size()) return fail(sim, VLJ_E_INVALID, "replicas must equal the batch size");
  int slot = free_slot(sim);
  if (slot < 0) return fail(sim, VLJ_E_STATE, "a measurement is already in flight");
  int rc = enqueue_measure(sim, slot);
  if (rc != VLJ_OK) return rc;
  Ring& r = sim->ring[slot];
  HIP_TRY(sim, hipEventSynchronize(r.done));
  for (size_t k = 0; k < replicas; k++)
    decode(sim, (uint32_t)k, r.slots + k * kSlots, r.ctl[k], r.step, per_replica + k);
  r.in_flight = false;
  return VLJ_OK;
}

size_t vlj_batch_replicas(const vlj_batch* batch) { return batch ? batch->rep.size() : 0; }

int vlj_fixed_sum(const double* values, size_t count, double fixed_scale, int64_t* out) {
  if (!out || (count && !values)) return fail(nullptr, VLJ_E_INVALID, "bad arguments to vlj_fixed_sum");
  vlj_sim tmp;
  int rc = check_device(&tmp);
  if (rc != VLJ_OK) return rc;
  double* d = nullptr;
  long long* acc = nullptr;
  hipStream_t st = nullptr;
  hipError_t e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipMalloc(&d, std::max<size_t>(count, 1) * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&acc, kSlots * sizeof(long long));
  if (e == hipSuccess && count) e = hipMemcpyAsync(d, values, count * sizeof(double), hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = launch_fixed_sum(d, count, fixed_scale, acc, st);
  long long host[kSlots] = {0};
  if (e == hipSuccess) e = hipMemcpyAsync(host, acc, sizeof(host), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (d) (void)hipFree(d);
  if (acc) (void)hipFree(acc);
  if (st) (void)hipStreamDestroy(st);
  if (e != hipSuccess) return fail(nullptr, VLJ_E_HIP, std::string("vlj_fixed_sum: ") + hipGetErrorString(e));
  *out = host[0];
  return VLJ_OK;
}

int vlj_fixed_sum_device(const double* values, size_t count, double fixed_scale, int64_t* out, void* stream) {
  if (!out || (count && !values)) return fail(nullptr, VLJ_E_INVALID, "bad arguments to vlj_fixed_sum_device");
  if ((uint64_t)count > kFixedSumMaxCount)
    return fail(nullptr, VLJ_E_INVALID, "vlj_fixed_sum_device takes at most 2^36 values");
  hipError_t e = launch_fixed_sum_noalloc(values, count, fixed_scale, (long long*)out, (hipStream_t)stream);
  if (e != hipSuccess) return fail(nullptr, VLJ_E_HIP, std::string("vlj_fixed_sum_device: ") + hipGetErrorString(e));
  return VLJ_OK;
}

int vlj_debug_state(vlj_sim* sim, double* pos, double* vel, double* acc, uint32_t* orig) {
  if (!sim) return fail(nullptr, VLJ_E_INVALID, "sim must not be NULL");
  std::vector<double> p, v, a;
  std::vector<uint32_t> o;
  int rc = read_state(sim, pos ? &p : nullptr, vel ? &v : nullptr, acc ? &a : nullptr, orig ? &o : nullptr);
  if (rc != VLJ_OK) return rc;
  if (pos) std::memcpy(pos, p.data(), p.size() * sizeof(double));
  if (vel) std::memcpy(vel, v.data(), v.size() * sizeof(double));
  if (acc) std::memcpy(acc, a.data(), a.size() * sizeof(double));
  if (orig) std::memcpy(orig, o.data(), o.size() * 4);
  return VLJ_OK;
}

int vlj_debug_cells(vlj_sim* sim, uint32_t* cell_of, uint32_t* cell_start, uint32_t* sorted) {
  if (!sim) return fail(nullptr, VLJ_E_INVALID, "sim must not be NULL");
  const Replica& R = sim->rep[0];
  const Buffers& B = sim->B;
  int rc = VLJ_OK;
  if (cell_of) {
    rc = read_dev(sim, cell_of, B.cell_of + R.atom_off, (size_t)R.count * 4);
    for (uint32_t i = 0; i < R.count && rc == VLJ_OK; i++) cell_of[i] -= R.cell_off;
  }
  if (cell_start && rc == VLJ_OK) {
    rc = read_dev(sim, cell_start, B.cell_start + R.cell_off, (size_t)R.total_cells * 4);
    for (uint32_t c = 0; c < R.total_cells; c++) cell_start[c] -= R.atom_off;
    cell_start[R.total_cells] = R.count;
  }
  if (sorted && rc == VLJ_OK) {
    rc = read_dev(sim, sorted, B.sorted + R.atom_off, (size_t)R.count * 4);
    for (uint32_t i = 0; i < R.count && rc == VLJ_OK; i++) sorted[i] -= R.atom_off;
  }
  return rc;
}

int vlj_debug_lists(vlj_sim* sim, uint32_t* counts, uint32_t* entries) {
  if (!sim) return fail(nullptr, VLJ_E_INVALID, "sim must not be NULL");
  const Replica& R = sim->rep[0];
  const Buffers& B = sim->B;
  std::vector<uint32_t> cnt(R.count);
  int rc = read_dev(sim, cnt.data(), B.nbr_count + R.atom_off, (size_t)R.count * 4);
  if (rc != VLJ_OK) return rc;
  if (counts) std::memcpy(counts, cnt.data(), cnt.size() * 4);
  if (entries) {
    // Entries are 8 x the staged indices of the atom's brick (the indices
    // themselves where the brick reads global memory); rebuild each brick's
    // staging order from the cell tables to turn them into slots.
    const size_t padded = (size_t)((R.count + kListStride - 1) / kListStride) * kListStride;
    std::vector<uint16_t> t((size_t)list_octets(R.capacity) * 8u * padded);
    std::vector<uint32_t> cell_of(R.count), start(R.total_cells), ncell(R.total_cells), sorted(R.count);
    rc = read_dev(sim, t.data(), B.nbr + R.list_off, t.size() * 2);
    if (rc == VLJ_OK) rc = read_dev(sim, cell_of.data(), B.cell_of + R.atom_off, (size_t)R.count * 4);
    if (rc == VLJ_OK) rc = read_dev(sim, start.data(), B.cell_start + R.cell_off, (size_t)R.total_cells * 4);
    if (rc == VLJ_OK) rc = read_dev(sim, ncell.data(), B.cell_n + R.cell_off, (size_t)R.total_cells * 4);
    if (rc == VLJ_OK) rc = read_dev(sim, sorted.data(), B.sorted + R.atom_off, (size_t)R.count * 4);
    if (rc != VLJ_OK) return rc;
    const int nx = (int)R.cells[0], ny = (int)R.cells[1], nz = (int)R.cells[2];
    for (uint32_t i = 0; i < R.count; i++) {
      const int c = (int)(cell_of[i] - R.cell_off);
      const int cx = c % nx, cy = (c / nx) % ny, cz = c / (nx * ny);
      const HostReplica& H = sim->host[0];
      const int z0 = H.brick_z[cz] == 0 ? cz - 1 : cz, gz = H.brick_z[z0];  // the brick's lowest layer
      const int bx = (int)H.brick_x[z0], by = (int)H.brick_y[z0];
      const int x0 = cx / bx * bx, y0 = cy / by * by;
      const int gx = std::min(bx, nx - x0), gy = std::min(by, ny - y0);
      std::vector<uint32_t> off, first;  // staged offset and first slot of each staged cell
      uint32_t run = 0;
      for (int dz = 0; dz < gz + 2; dz++)
        for (int yy = 0; yy < gy + 2; yy++)
          for (int xx = 0; xx < gx + 2; xx++) {
            const int ccx = (x0 - 1 + xx + nx) % nx, ccy = (y0 - 1 + yy + ny) % ny, ccz = (z0 - 1 + dz + nz) % nz;
            const int cell = (ccz * ny + ccy) * nx + ccx;
            off.push_back(run);
            first.push_back(start[cell] - R.atom_off);
            run += ncell[cell];
          }
      off.push_back(run);
      const bool staged = run <= (uint32_t)(H.sparse[z0] ? Sparse::kStageMax : Dense::kStageMax);
      const uint64_t base = list_base(R, i) - R.list_off;
      for (uint32_t k = 0; k < R.capacity; k++) {
        uint32_t slot = 0;
        if (k < cnt[i]) {
          const uint32_t e = list_index(t[base + list_entry(k)], staged);
          size_t sc = std::upper_bound(off.begin(), off.end(), e) - off.begin() - 1;
          while (sc + 1 < off.size() && off[sc + 1] <= e) sc++;
          slot = sorted[first[sc] + (e - off[sc])] - R.atom_off;
        }
        entries[(size_t)i * R.capacity + k] = slot;
      }
    }
  }
  return VLJ_OK;
}

int vlj_debug_grid(const vlj_sim* sim, uint32_t out[6]) {
  if (!sim || !out) return VLJ_E_INVALID;
  const Replica& R = sim->rep[0];
  out[0] = R.cells[0];
  out[1] = R.cells[1];
  out[2] = R.cells[2];
  out[3] = R.total_cells;
  out[4] = R.capacity;
  Control c;
  if (hipMemcpy(&c, sim->B.ctl, sizeof(Control), hipMemcpyDeviceToHost) != hipSuccess) return VLJ_E_HIP;
  out[5] = c.rebuilds;
  return VLJ_OK;
}

int vlj_debug_counters(const vlj_sim* sim, uint64_t out[6]) {
  if (!sim || !out) return VLJ_E_INVALID;
  std::vector<Control> c(sim->rep.size());
  if (!c.empty() && hipMemcpy(c.data(), sim->B.ctl, c.size() * sizeof(Control), hipMemcpyDeviceToHost) != hipSuccess)
    return VLJ_E_HIP;
  out[0] = out[1] = out[2] = out[4] = 0;
  for (const Control& x : c) {
    out[0] += x.rebuilds;
    out[1] += x.stage_overflow;
    out[2] = std::max<uint64_t>(out[2], x.overflow);
    out[4] += x.stage_global;
  }
  out[3] = sim->bricks.size();
  out[5] = sim->bricks.size() - sim->n_dense;
  return VLJ_OK;
}

int vlj_debug_replica_counters(const vlj_sim* sim, size_t replica, uint64_t out[4]) {
  if (!sim || !out || replica >= sim->rep.size()) return VLJ_E_INVALID;
  Control c;
  if (hipMemcpy(&c, sim->B.ctl + replica, sizeof(Control), hipMemcpyDeviceToHost) != hipSuccess) return VLJ_E_HIP;
  out[0] = c.rebuilds;
  out[1] = c.stage_overflow;
  out[2] = c.overflow;
  out[3] = c.stage_global;
  return VLJ_OK;
}

int vlj_debug_params(const vlj_sim* sim, double out[6]) {
  if (!sim || !out) return VLJ_E_INVALID;
  out[0] = sim->host[0].energy_scale;
  out[1] = sim->host[0].momentum_scale;
  out[2] = sim->rep[0].shift;
  out[3] = sim->rep[0].force_shift;
  out[4] = sim->rep[0].c1;
  out[5] = sim->rep[0].c2;
  return VLJ_OK;
}

int vlj_tune(vlj_sim* sim, uint32_t steps_per_candidate, uint64_t* steps_taken) {
  if (!sim || steps_per_candidate == 0) return fail(sim, VLJ_E_INVALID, "bad arguments to vlj_tune");
  if (g_plan.stage_fill != 0.0) return fail(sim, VLJ_E_STATE, "the brick plan is fixed by vlj_debug_plan");
  uint64_t taken = 0;
  int rc = VLJ_OK;
  if (sim->n_total && !sim->bricks.empty()) rc = tune(sim, steps_per_candidate, &taken);
  if (steps_taken) *steps_taken = taken;
  return rc;
}

int vlj_plan_info(const vlj_sim* sim, double* fill, uint32_t* replans, vlj_tune_trial* trials, size_t capacity,
                  size_t* n_trials) {
  if (!sim || (capacity && !trials)) return fail(nullptr, VLJ_E_INVALID, "bad arguments to vlj_plan_info");
  if (fill) *fill = plan_fill(sim);
  if (replans) *replans = sim->replans;
  for (size_t k = 0; k < std::min(capacity, sim->trials.size()); k++) {
    const TuneTrial& t = sim->trials[k];
    trials[k] = vlj_tune_trial{t.fill, t.force_us, t.build_us, t.bricks};
  }
  if (n_trials) *n_trials = sim->trials.size();
  return VLJ_OK;
}

int vlj_debug_plan(uint32_t list_capacity, double stage_fill, double sparse_fill) {
  if (!std::isfinite(stage_fill) || !std::isfinite(sparse_fill) || stage_fill < 0.0)
    return fail(nullptr, VLJ_E_INVALID, "bad arguments to vlj_debug_plan");
  g_plan.list_capacity = list_capacity;
  g_plan.stage_fill = stage_fill;
  g_plan.sparse_fill = sparse_fill;
  return VLJ_OK;
}

int vlj_debug_parts(int32_t mode) {
  if (mode < -1 || mode > 1) return fail(nullptr, VLJ_E_INVALID, "bad argument to vlj_debug_parts");
  g_plan.parts = mode;
  return VLJ_OK;
}

uint32_t vlj_debug_swizzle(uint32_t block, uint32_t blocks) {
  return blocks == 0 ? dealt_grid(block) : dealt_index(block);
}

int vlj_copy_bandwidth(size_t bytes, uint32_t repeats, double* ms) {
  if (!ms || bytes < 16 || repeats == 0) return fail(nullptr, VLJ_E_INVALID, "bad arguments to vlj_copy_bandwidth");
  vlj_sim tmp;
  int rc = check_device(&tmp);
  if (rc != VLJ_OK) return rc;
  const size_t words = bytes / 16;
  void *a = nullptr, *b = nullptr;
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipMalloc(&a, words * 16);
  if (e == hipSuccess) e = hipMalloc(&b, words * 16);
  if (e == hipSuccess) e = hipMemsetAsync(a, 0, words * 16, st);
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  if (e == hipSuccess) e = launch_copy(a, b, words, st);  // warm-up
  if (e == hipSuccess) e = hipEventRecord(e0, st);
  for (uint32_t k = 0; k < repeats && e == hipSuccess; k++) e = launch_copy(a, b, words, st);
  if (e == hipSuccess) e = hipEventRecord(e1, st);
  if (e == hipSuccess) e = hipEventSynchronize(e1);
  float t = 0.0f;
  if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
  if (a) (void)hipFree(a);
  if (b) (void)hipFree(b);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (st) (void)hipStreamDestroy(st);
  if (e != hipSuccess) return fail(nullptr, VLJ_E_HIP, std::string("vlj_copy_bandwidth: ") + hipGetErrorString(e));
  *ms = (double)t / repeats;
  return VLJ_OK;
}

int vlj_fp64_fma_rate(uint32_t repeats, double* tflops) {
  if (!tflops || repeats == 0) return fail(nullptr, VLJ_E_INVALID, "bad arguments to vlj_fp64_fma_rate");
  vlj_sim tmp;
  int rc = check_device(&tmp);
  if (rc != VLJ_OK) return rc;
  const uint32_t blocks = 256u * 16u, iters = 256u;  // 1M threads x 8 chains x 256 FMAs
  double* out = nullptr;
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipMalloc(&out, (size_t)blocks * kBlock * sizeof(double));
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  if (e == hipSuccess) e = launch_fma(out, blocks, iters, st);  // warm-up
  if (e == hipSuccess) e = hipEventRecord(e0, st);
  for (uint32_t k = 0; k < repeats && e == hipSuccess; k++) e = launch_fma(out, blocks, iters, st);
  if (e == hipSuccess) e = hipEventRecord(e1, st);
  if (e == hipSuccess) e = hipEventSynchronize(e1);
  float t = 0.0f;
  if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
  if (out) (void)hipFree(out);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (st) (void)hipStreamDestroy(st);
  if (e != hipSuccess) return fail(nullptr, VLJ_E_HIP, std::string("vlj_fp64_fma_rate: ") + hipGetErrorString(e));
  const double flops = 2.0 * 8.0 * iters * (double)blocks * kBlock * repeats;
  *tflops = t > 0.0f ? flops / ((double)t * 1e-3) / 1e12 : 0.0;
  return VLJ_OK;
}

int vlj_debug_rcp(const double* x, double* out, size_t n) {
  if (n && (!x || !out)) return fail(nullptr, VLJ_E_INVALID, "bad arguments to vlj_debug_rcp");
  vlj_sim tmp;
  int rc = check_device(&tmp);
  if (rc != VLJ_OK) return rc;
  double *dx = nullptr, *dout = nullptr;
  const size_t bytes = std::max<size_t>(n, 1) * sizeof(double);
  hipError_t e = hipMalloc(&dx, bytes);
  if (e == hipSuccess) e = hipMalloc(&dout, bytes);
  if (e == hipSuccess) e = hipMemcpy(dx, x, n * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = launch_rcp(dx, dout, n, nullptr);
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e == hipSuccess) e = hipMemcpy(out, dout, n * sizeof(double), hipMemcpyDeviceToHost);
  if (dx) (void)hipFree(dx);
  if (dout) (void)hipFree(dout);
  if (e != hipSuccess) return fail(nullptr, VLJ_E_HIP, std::string("vlj_debug_rcp: ") + hipGetErrorString(e));
  return VLJ_OK;
}

int vlj_debug_rcp32(const float* x, float* rcp, float* root, size_t n) {
  if (n && (!x || !rcp || !root)) return fail(nullptr, VLJ_E_INVALID, "bad arguments to vlj_debug_rcp32");
  vlj_sim tmp;
  int rc = check_device(&tmp);
  if (rc != VLJ_OK) return rc;
  float *dx = nullptr, *dr = nullptr, *ds = nullptr;
  const size_t bytes = std::max<size_t>(n, 1) * sizeof(float);
  hipError_t e = hipMalloc(&dx, bytes);
  if (e == hipSuccess) e = hipMalloc(&dr, bytes);
  if (e == hipSuccess) e = hipMalloc(&ds, bytes);
  if (e == hipSuccess) e = hipMemcpy(dx, x, n * sizeof(float), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = launch_rcp32(dx, dr, ds, n, nullptr);
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e == hipSuccess) e = hipMemcpy(rcp, dr, n * sizeof(float), hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(root, ds, n * sizeof(float), hipMemcpyDeviceToHost);
  if (dx) (void)hipFree(dx);
  if (dr) (void)hipFree(dr);
  if (ds) (void)hipFree(ds);
  if (e != hipSuccess) return fail(nullptr, VLJ_E_HIP, std::string("vlj_debug_rcp32: ") + hipGetErrorString(e));
  return VLJ_OK;
}

int vlj_timed_run(vlj_sim* sim, uint64_t steps, double* force_ms, double* total_ms) {
  if (!sim || !force_ms || !total_ms) return fail(sim, VLJ_E_INVALID, "bad arguments to vlj_timed_run");
  int rc = prime_if_needed(sim);
  if (rc != VLJ_OK) return rc;
  std::vector<hipEvent_t> ev(2 * steps + 2, nullptr);
  for (auto& e : ev) HIP_TRY(sim, hipEventCreate(&e));
  HIP_TRY(sim, hipEventRecord(ev[2 * steps], sim->stream));
  for (uint64_t t = 0; t < steps && rc == VLJ_OK; t++) rc = enqueue_step(sim, &ev[2 * t]);
  if (rc != VLJ_OK) {
    for (auto& e : ev) (void)hipEventDestroy(e);
    return rc;
  }
  HIP_TRY(sim, hipEventRecord(ev[2 * steps + 1], sim->stream));
  HIP_TRY(sim, hipEventSynchronize(ev[2 * steps + 1]));
  double f = 0.0;
  for (uint64_t t = 0; t < steps; t++) {
    float ms = 0.0f;
    HIP_TRY(sim, hipEventElapsedTime(&ms, ev[2 * t], ev[2 * t + 1]));
    f += ms;
  }
  float total = 0.0f;
  HIP_TRY(sim, hipEventElapsedTime(&total, ev[2 * steps], ev[2 * steps + 1]));
  for (auto& e : ev) (void)hipEventDestroy(e);
  sim->steps += steps;
  *force_ms = f;
  *total_ms = total;
  return VLJ_OK;
}

// ---- lattices (lennard_jones.rs:1377-1441), host only ----

static uint64_t xorshift(uint64_t* s) {
  *s ^= *s << 13;
  *s ^= *s >> 7;
  *s ^= *s << 17;
  return *s;
}

static double lattice_random(uint64_t* s) {
  xorshift(s);
  return ((double)(*s >> 11) + 0.5) / 9007199254740992.0;
}

static double lattice_gaussian(uint64_t* s) {
  double r = std::sqrt(-2.0 * std::log(lattice_random(s)));
  double u = lattice_random(s);
  return r * std::cos(6.283185307179586 * u);
}

static void zero_momentum(size_t count, vlj_particle* out) {
  for (int k = 0; k < 3; k++) {
    double sum = 0.0;
    for (size_t i = 0; i < count; i++) sum += out[i].vel[k];
    double mean = sum / (double)count;
    for (size_t i = 0; i < count; i++) out[i].vel[k] -= mean;
  }
}

static void lattice_sites(size_t count, const double size[3], const size_t sites[3], double temperature,
                          uint64_t seed, vlj_particle* out) {
  double gap[3];
  for (int k = 0; k < 3; k++) gap[k] = size[k] / (double)sites[k];
  uint64_t s = seed | 1u;
  double st = std::sqrt(temperature);
  for (size_t i = 0; i < count; i++) {
    size_t c[3] = {i % sites[0], (i / sites[0]) % sites[1], i / (sites[0] * sites[1])};
    for (int k = 0; k < 3; k++) out[i].pos[k] = ((double)c[k] + 0.5) * gap[k];
    for (int k = 0; k < 3; k++) out[i].vel[k] = lattice_gaussian(&s) * st;
  }
  if (count) zero_momentum(count, out);
}

int vlj_lattice(size_t count, double box_size, double temperature, uint64_t seed, vlj_particle* out) {
  if (count && !out) return fail(nullptr, VLJ_E_INVALID, "out must not be NULL");
  size_t per_side = (size_t)std::ceil(std::cbrt((double)count));
  double size[3] = {box_size, box_size, box_size};
  size_t sites[3] = {per_side, per_side, per_side};
  lattice_sites(count, size, sites, temperature, seed, out);
  return VLJ_OK;
}

int vlj_lattice_box(size_t count, const double size[3], double temperature, uint64_t seed, vlj_particle* out) {
  if ((count && !out) || !size) return fail(nullptr, VLJ_E_INVALID, "bad arguments to lattice_box");
  if (count == 0) return VLJ_OK;
  double spacing = std::cbrt(size[0] * size[1] * size[2] / (double)count);
  size_t sites[3];
  for (int k = 0; k < 3; k++) {
    double r = std::round(size[k] / spacing);
    sites[k] = r < 1.0 ? 1 : (size_t)r;
  }
  while (sites[0] * sites[1] * sites[2] < count) {
    int widest = 0;  // Rust's max_by keeps the last of equal maxima
    for (int k = 1; k < 3; k++)
      if (size[k] / (double)sites[k] >= size[widest] / (double)sites[widest]) widest = k;
    sites[widest] += 1;
  }
  lattice_sites(count, size, sites, temperature, seed, out);
  return VLJ_OK;
}

int vlj_fcc(uint32_t n, double density, double temperature, uint64_t seed, vlj_particle* out, double* side) {
  if (!out || !side || n == 0 || !(density > 0.0)) return fail(nullptr, VLJ_E_INVALID, "bad arguments to fcc");
  const double a0 = std::cbrt(4.0 / density);
  const double L = (double)(float)(a0 * (double)n);
  const double a = L / (double)n;
  static const double basis[4][3] = {{0, 0, 0}, {0.5, 0.5, 0}, {0.5, 0, 0.5}, {0, 0.5, 0.5}};
  uint64_t s = seed | 1u;
  const double st = std::sqrt(temperature);
  size_t i = 0;
  for (uint32_t cz = 0; cz < n; cz++)
    for (uint32_t cy = 0; cy < n; cy++)
      for (uint32_t cx = 0; cx < n; cx++)
        for (int b = 0; b < 4; b++, i++) {
          const double c[3] = {(double)cx, (double)cy, (double)cz};
          for (int k = 0; k < 3; k++) out[i].pos[k] = (c[k] + basis[b][k] + 0.25) * a;
          for (int k = 0; k < 3; k++) out[i].vel[k] = lattice_gaussian(&s) * st;
        }
  zero_momentum(i, out);
  *side = L;
  return VLJ_OK;
}

// FCC sites for any count: the smallest 4 n^3 lattice that holds `count`
// atoms, filled in (cz, cy, cx, basis) order, in a cube of side
// cbrt(count / density) rounded to f32; sites offset by a/4 as in vlj_fcc.
int vlj_fcc_count(size_t count, double density, double temperature, uint64_t seed, vlj_particle* out,
                  double* side) {
  if ((count && !out) || !side || !(density > 0.0)) return fail(nullptr, VLJ_E_INVALID, "bad arguments to fcc_count");
  uint32_t n = 1;
  while (4ull * n * n * n < count) n++;
  const double L = (double)(float)std::cbrt((double)count / density);
  const double a = L / (double)n;
  static const double basis[4][3] = {{0, 0, 0}, {0.5, 0.5, 0}, {0.5, 0, 0.5}, {0, 0.5, 0.5}};
  uint64_t s = seed | 1u;
  const double st = std::sqrt(temperature);
  for (size_t i = 0; i < count; i++) {
    const size_t cell = i / 4, b = i % 4;
    const double c[3] = {(double)(cell % n), (double)((cell / n) % n), (double)(cell / ((size_t)n * n))};
    for (int k = 0; k < 3; k++) out[i].pos[k] = (c[k] + basis[b][k] + 0.25) * a;
    for (int k = 0; k < 3; k++) out[i].vel[k] = lattice_gaussian(&s) * st;
  }
  if (count) zero_momentum(count, out);
  *side = L;
  return VLJ_OK;
}

// Random non-overlapping insertion: uniform trial positions from the
// xorshift stream (seed | 1), rejected when closer than min_distance
// (nearest image) to an earlier particle, at most 1000 trials per particle;
// then Gaussian velocities from the same stream, mean removed. A cell grid
// of cells at least min_distance wide finds the earlier particles that can
// be that close, so the result is the same as checking every earlier one.
int vlj_random_insertion(size_t count, const double box[3], double min_distance, double temperature,
                         uint64_t seed, vlj_particle* out) {
  if ((count && !out) || !box || !(min_distance > 0.0))
    return fail(nullptr, VLJ_E_INVALID, "bad arguments to random_insertion");
  for (int k = 0; k < 3; k++)
    if (!(box[k] > 0.0)) return fail(nullptr, VLJ_E_INVALID, "box sides must be positive");
  int nc[3];
  for (int k = 0; k < 3; k++) {
    // one cell fewer than fit: cells then exceed min_distance by a margin
    // far above any rounding of the bucket index
    const double c = std::floor(box[k] / min_distance) - 1.0;
    nc[k] = c < 1.0 ? 1 : (c > 1024.0 ? 1024 : (int)c);
  }
  std::vector<std::vector<uint32_t>> grid((size_t)nc[0] * nc[1] * nc[2]);
  auto cell_axis = [&](double x, int k) {
    int c = (int)std::floor(x / box[k] * nc[k]);
    return c < 0 ? 0 : (c >= nc[k] ? nc[k] - 1 : c);
  };
  uint64_t s = seed | 1u;
  const double md2 = min_distance * min_distance;
  std::vector<size_t> near;
  for (size_t i = 0; i < count; i++) {
    bool placed = false;
    for (int trial = 0; trial < 1000 && !placed; trial++) {
      double p[3];
      for (int k = 0; k < 3; k++) p[k] = lattice_random(&s) * box[k];
      int c[3];
      for (int k = 0; k < 3; k++) c[k] = cell_axis(p[k], k);
      near.clear();
      for (int dz = -1; dz <= 1; dz++)
        for (int dy = -1; dy <= 1; dy++)
          for (int dx = -1; dx <= 1; dx++) {
            const int x = (c[0] + dx + nc[0]) % nc[0], y = (c[1] + dy + nc[1]) % nc[1],
                      z = (c[2] + dz + nc[2]) % nc[2];
            near.push_back(((size_t)z * nc[1] + y) * nc[0] + x);
          }
      std::sort(near.begin(), near.end());
      near.erase(std::unique(near.begin(), near.end()), near.end());
      placed = true;
      for (size_t cell : near) {
        for (uint32_t j : grid[cell]) {
          double r2 = 0;
          for (int k = 0; k < 3; k++) {
            double d = out[j].pos[k] - p[k];
            d -= box[k] * std::round(d / box[k]);
            r2 += d * d;
          }
          if (r2 < md2) {
            placed = false;
            break;
          }
        }
        if (!placed) break;
      }
      if (placed) {
        for (int k = 0; k < 3; k++) out[i].pos[k] = p[k];
        grid[((size_t)c[2] * nc[1] + c[1]) * nc[0] + c[0]].push_back((uint32_t)i);
      }
    }
    if (!placed)
      return fail(nullptr, VLJ_E_INVALID,
                  "random insertion found no free place for particle " + std::to_string(i) + " in 1000 trials");
  }
  const double st = std::sqrt(temperature);
  for (size_t i = 0; i < count; i++)
    for (int k = 0; k < 3; k++) out[i].vel[k] = lattice_gaussian(&s) * st;
  if (count) zero_momentum(count, out);
  return VLJ_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Checkpoint and exact resume. A state file holds everything that carries
// from one step to the next, as the device holds it (slot order), so a
// handle loaded from it steps on to the same bits as the handle that saved
// it: positions, velocities and accelerations; positions at the last list
// build; original indices; cell tables, lists and the bricks (with the parts
// of cut bricks) they refer to; the control words (the step counter that
// keys the Langevin noise, the rebuild count, Nose-Hoover xi and eta); the
// replicas' parameters as last set (target temperature, heat exchange);
// profile sums and the heat-current series. Measurements in flight are not
// kept. Layout: a StateHeader, then the sections in the order written below,
// all little-endian, each array as raw elements.
// ---------------------------------------------------------------------------

namespace {

constexpr char kStateMagic[8] = {'V', 'L', 'J', 'S', 'T', 'A', 'T', 'E'};
constexpr uint32_t kStateVersion = 2;

struct StateHeader {
  char magic[8];
  uint32_t version, replicas;
  uint64_t n_total, cells_total, list_total, prof_total, n_bricks, steps;
  uint32_t primed, hc_samples, hc_cap;
  uint32_t sizeof_replica, sizeof_control, sizeof_settings, sizeof_brick, n_dense;
};

struct FileOut {
  FILE* f = nullptr;
  bool ok = true;
  void put(const void* p, size_t n) {
    if (ok && n) ok = std::fwrite(p, 1, n, f) == n;
  }
};

struct FileIn {
  FILE* f = nullptr;
  bool ok = true;
  void get(void* p, size_t n) {
    if (ok && n) ok = std::fread(p, 1, n, f) == n;
  }
};

template <typename T>
int save_array(vlj_sim* s, FileOut& w, const T* dev, size_t n) {
  std::vector<T> h(n);
  if (n) HIP_TRY(s, hipMemcpy(h.data(), dev, n * sizeof(T), hipMemcpyDeviceToHost));
  w.put(h.data(), n * sizeof(T));
  return VLJ_OK;
}

template <typename T>
bool read_vec(FileIn& r, std::vector<T>& v, size_t n) {
  v.resize(n);
  r.get(v.data(), n * sizeof(T));
  return r.ok;
}

template <typename T>
int upload(vlj_sim* s, T* dev, const std::vector<T>& h) {
  if (!h.empty()) HIP_TRY(s, hipMemcpy(dev, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
  return VLJ_OK;
}

// The state arrays of a file, in file order.
struct StateArrays {
  std::vector<double> f64[12];  // x y z vx vy vz ax ay az xb yb zb
  std::vector<uint32_t> orig, cell_of, nbr_count, cell_start, cell_n, cut, cut_groups;
  std::vector<uint16_t> nbr, parts;
  std::vector<long long> profile, hc;
};

// Everything in a state file that indexes device memory, checked on the
// host before any device work: a file with valid sizes but bad contents
// gets VLJ_E_INVALID instead of driving out-of-bounds accesses. Returns ""
// or what is wrong.
std::string check_state(const std::vector<Replica>& rep, const std::vector<Brick>& bricks, const StateHeader& h,
                        const std::vector<Control>& ctl, const StateArrays& a) {
  const size_t R = rep.size(), nb = bricks.size();
  for (size_t k = 0; k < R; k++) {
    const Replica& x = rep[k];
    if (x.cells[0] < 2 || x.cells[1] < 2 || x.cells[2] < 2 ||
        (uint64_t)x.cells[0] * x.cells[1] * x.cells[2] != x.total_cells || x.total_cells > (1u << 20))
      return "replica " + std::to_string(k) + ": bad cell grid";
    if (x.capacity == 0 || (k > 0 && x.hc_cap != 0) || (k == 0 && x.hc_cap != h.hc_cap))
      return "replica " + std::to_string(k) + ": bad capacity or heat-current series";
    if (x.prof_bins > kMaxProfileBins) return "replica " + std::to_string(k) + ": bad profile slabs";
    if (ctl[k].rebuild > 1u || ctl[k].reorder > 1u) return "replica " + std::to_string(k) + ": bad control words";
  }
  // bricks: inside their replica's rows, tiling every replica's cells once
  std::vector<uint32_t> owner(h.cells_total, 0xffffffffu);
  for (size_t b = 0; b < nb; b++) {
    const Brick& g = bricks[b];
    if (g.rep >= R || g.gx < 1 || g.gy < 1 || g.gz < 1 || g.gz > 2u || g.gx > (uint32_t)kBrickMax ||
        g.gy > (uint32_t)kBrickMax || home_cells(g) > (uint32_t)kHomeCells ||
        (g.gz + 2u) * (g.gx + 2u) * (g.gy + 2u) > (uint32_t)kStageCells)
      return "bad brick table";
    const Replica& x = rep[g.rep];
    if (g.cell0 >= x.total_cells) return "bad brick table";
    const uint32_t cx = g.cell0 % x.cells[0], cy = (g.cell0 / x.cells[0]) % x.cells[1];
    const uint32_t cz = g.cell0 / (x.cells[0] * x.cells[1]);
    if (cx + g.gx > x.cells[0] || cy + g.gy > x.cells[1] || cz + g.gz > x.cells[2]) return "a brick runs past its row";
    for (uint32_t zz = 0; zz < g.gz; zz++)
      for (uint32_t yy = 0; yy < g.gy; yy++)
        for (uint32_t xx = 0; xx < g.gx; xx++) {
          uint32_t& o = owner[x.cell_off + g.cell0 + (zz * x.cells[1] + yy) * x.cells[0] + xx];
          if (o != 0xffffffffu) return "bricks overlap";
          o = (uint32_t)b;
        }
  }
  for (uint32_t o : owner)
    if (o == 0xffffffffu) return "bricks do not cover every cell";
  // parts of cut bricks
  const size_t ng = (nb + kCutGroup - 1) / kCutGroup;
  std::vector<uint32_t> group_cut(ng, 0);
  for (size_t b = 0; b < nb; b++) {
    const uint32_t np = a.cut[b];
    if (np > (uint32_t)kHomeCells) return "bad part count";
    if (np) group_cut[b / kCutGroup]++;
    for (uint32_t u = 0; u < np; u++) {
      const uint32_t c = a.parts[b * kHomeCells + u];
      const uint32_t x0 = c & 7u, y0 = (c >> 3) & 7u, w = ((c >> 6) & 7u) + 1u, hh = ((c >> 9) & 7u) + 1u,
                     z0 = c >> 12;
      if (x0 + w > bricks[b].gx || y0 + hh > bricks[b].gy || z0 >= bricks[b].gz) return "bad part of a cut brick";
    }
  }
  for (size_t g = 0; g < ng; g++)
    if (a.cut_groups[g] != group_cut[g]) return "cut counts do not match the cut bricks";
  // cells in cell order inside their replica; atoms and lists inside theirs
  // (before the first list build the cell tables and lists hold nothing)
  if (!h.primed) return "";
  for (size_t k = 0; k < R; k++) {
    const Replica& x = rep[k];
    uint32_t next = x.atom_off;
    for (uint32_t c = x.cell_off; c < x.cell_off + x.total_cells; c++) {
      if (a.cell_start[c] != next || a.cell_n[c] > x.count) return "bad cell table";
      next += a.cell_n[c];
    }
    if (next != x.atom_off + x.count) return "bad cell table";
    std::vector<char> seen(x.count, 0);
    for (uint32_t i = x.atom_off; i < x.atom_off + x.count; i++) {
      if (a.orig[i] >= x.count || seen[a.orig[i]]) return "bad original index";
      seen[a.orig[i]] = 1;
      if (a.cell_of[i] < x.cell_off || a.cell_of[i] >= x.cell_off + x.total_cells) return "bad cell of an atom";
      if (a.nbr_count[i] > x.capacity) return "a list longer than the capacity";
    }
  }
  // list entries: staged indices below the staged count of the atom's brick
  // (or part), from the same cell counts the kernels stage
  for (size_t b = 0; b < nb; b++) {
    const Brick& whole = bricks[b];
    const Replica& x = rep[whole.rep];
    const uint32_t np = a.cut[b];
    const uint32_t stage_max = b < h.n_dense ? (uint32_t)kStageMax : (uint32_t)Sparse::kStageMax;
    for (uint32_t u = 0; u < (np ? np : 1u); u++) {
      Brick g = whole;
      if (np) {
        const uint32_t c = a.parts[b * kHomeCells + u];
        const uint32_t x0 = c & 7u, y0 = (c >> 3) & 7u, z0 = (c >> 12) & 1u;
        g = Brick{whole.rep, whole.cell0 + (z0 * x.cells[1] + y0) * x.cells[0] + x0, ((c >> 6) & 7u) + 1u,
                  ((c >> 9) & 7u) + 1u, 1u};
      }
      const int nx = (int)x.cells[0], ny = (int)x.cells[1], nz = (int)x.cells[2];
      const int cx0 = (int)(g.cell0 % x.cells[0]), cy0 = (int)((g.cell0 / x.cells[0]) % x.cells[1]);
      const int cz = (int)(g.cell0 / (x.cells[0] * x.cells[1]));
      uint64_t total = 0;
      for (int dz = 0; dz < (int)g.gz + 2; dz++)
        for (int yy = 0; yy < (int)g.gy + 2; yy++)
          for (int xx = 0; xx < (int)g.gx + 2; xx++) {
            const int ccx = (cx0 - 1 + xx + nx) % nx, ccy = (cy0 - 1 + yy + ny) % ny, ccz = (cz - 1 + dz + nz) % nz;
            total += a.cell_n[x.cell_off + (uint32_t)((ccz * ny + ccy) * nx + ccx)];
          }
      const bool staged = total <= stage_max;
      for (uint32_t zz = 0; zz < g.gz; zz++)
      for (uint32_t yy = 0; yy < g.gy; yy++)
        for (uint32_t xx = 0; xx < g.gx; xx++) {
          const uint32_t c = x.cell_off + g.cell0 + (zz * x.cells[1] + yy) * x.cells[0] + xx;
          for (uint32_t i = a.cell_start[c]; i < a.cell_start[c] + a.cell_n[c]; i++) {
            const uint64_t base = list_base(x, i - x.atom_off);
            for (uint32_t k = 0; k < a.nbr_count[i]; k++) {
              const uint64_t at = base + list_entry(k);
              if (at >= h.list_total) return "a list outside the list memory";
              if (list_index(a.nbr[at], staged) >= total) return "a list entry outside its staged neighbourhood";
            }
          }
        }
    }
  }
  return "";
}

double* const* state_f64(Buffers& B) {
  static thread_local double* p[12];
  double* q[12] = {B.x, B.y, B.z, B.vx, B.vy, B.vz, B.ax, B.ay, B.az, B.xb, B.yb, B.zb};
  for (int k = 0; k < 12; k++) p[k] = q[k];
  return p;
}

}  // namespace

extern "C" {

int vlj_save(vlj_sim* sim, const char* path) {
  if (!sim || !path) return fail(sim, VLJ_E_INVALID, "sim and path must not be NULL");
  HIP_TRY(sim, hipStreamSynchronize(sim->stream));
  FileOut w;
  w.f = std::fopen(path, "wb");
  if (!w.f) return fail(sim, VLJ_E_INVALID, std::string("could not open ") + path + " for writing");
  const size_t R = sim->rep.size(), N = sim->n_total, C = sim->cells_total, nb = sim->bricks.size();
  const size_t ng = (nb + kCutGroup - 1) / kCutGroup;
  StateHeader h{};
  std::memcpy(h.magic, kStateMagic, 8);
  h.version = kStateVersion;
  h.replicas = (uint32_t)R;
  h.n_total = N;
  h.cells_total = C;
  h.list_total = sim->list_total;
  h.prof_total = sim->prof_total;
  h.n_bricks = nb;
  h.steps = sim->steps;
  h.primed = sim->primed ? 1u : 0u;
  h.hc_samples = sim->hc_samples;
  h.hc_cap = sim->rep[0].hc_cap;
  h.sizeof_replica = sizeof(Replica);
  h.sizeof_control = sizeof(Control);
  h.sizeof_settings = sizeof(vlj_settings);
  h.sizeof_brick = sizeof(Brick);
  h.n_dense = sim->n_dense;
  w.put(&h, sizeof h);
  for (const HostReplica& x : sim->host) w.put(&x.settings, sizeof(vlj_settings));
  w.put(sim->rep.data(), R * sizeof(Replica));
  for (const HostReplica& x : sim->host) {
    w.put(&x.prof_samples, sizeof(uint64_t));
    w.put(x.brick_x.data(), x.brick_x.size() * sizeof(uint32_t));
    w.put(x.brick_y.data(), x.brick_y.size() * sizeof(uint32_t));
    w.put(x.brick_z.data(), x.brick_z.size());
    w.put(x.sparse.data(), x.sparse.size());
  }
  w.put(sim->bricks.data(), nb * sizeof(Brick));
  int rc = save_array(sim, w, sim->B.ctl, R);
  double* const* f = state_f64(sim->B);
  for (int k = 0; k < 12 && rc == VLJ_OK; k++) rc = save_array(sim, w, f[k], N);
  if (rc == VLJ_OK) rc = save_array(sim, w, sim->B.orig, N);
  if (rc == VLJ_OK) rc = save_array(sim, w, sim->B.cell_of, N);
  if (rc == VLJ_OK) rc = save_array(sim, w, sim->B.nbr_count, N);
  if (rc == VLJ_OK) rc = save_array(sim, w, sim->B.cell_start, C);
  if (rc == VLJ_OK) rc = save_array(sim, w, sim->B.cell_n, C);
  if (rc == VLJ_OK) rc = save_array(sim, w, sim->B.nbr, sim->list_total);
  if (rc == VLJ_OK) rc = save_array(sim, w, sim->B.cut, nb);
  if (rc == VLJ_OK) rc = save_array(sim, w, sim->B.parts, nb * kHomeCells);
  if (rc == VLJ_OK) rc = save_array(sim, w, sim->B.cut_groups, ng);
  if (rc == VLJ_OK) rc = save_array(sim, w, sim->B.profile, sim->prof_total);
  if (rc == VLJ_OK) rc = save_array(sim, w, sim->B.hc, (size_t)h.hc_cap * 3);
  const bool closed = std::fclose(w.f) == 0;
  if (rc != VLJ_OK) return rc;
  if (!w.ok || !closed) return fail(sim, VLJ_E_INVALID, std::string("could not write ") + path);
  return VLJ_OK;
}

int vlj_render_buffers(vlj_sim* sim, void** positions, void** velocities) {
  if (!sim) return fail(nullptr, VLJ_E_INVALID, "sim must not be NULL");
  if (!sim->B.render_pos) {
    int rc = dev_alloc(sim, &sim->B.render_pos, (size_t)sim->n_total * 4);
    if (rc == VLJ_OK) rc = dev_alloc(sim, &sim->B.render_vel, (size_t)sim->n_total * 4);
    if (rc != VLJ_OK) return rc;
    HIP_TRY(sim, hipMemsetAsync(sim->B.render_pos, 0, std::max<size_t>(sim->n_total, 1) * 16, sim->stream));
    HIP_TRY(sim, hipMemsetAsync(sim->B.render_vel, 0, std::max<size_t>(sim->n_total, 1) * 16, sim->stream));
    HIP_TRY(sim, hipStreamSynchronize(sim->stream));
  }
  if (positions) *positions = sim->B.render_pos;
  if (velocities) *velocities = sim->B.render_vel;
  return VLJ_OK;
}

int vlj_render(vlj_sim* sim) {
  if (!sim) return fail(nullptr, VLJ_E_INVALID, "sim must not be NULL");
  int rc = vlj_render_buffers(sim, nullptr, nullptr);
  if (rc == VLJ_OK) rc = prime_if_needed(sim);
  if (rc != VLJ_OK) return rc;
  HIP_TRY(sim, launch_render(sim->B, sim->stream));
  return VLJ_OK;
}

int vlj_read_render(vlj_sim* sim, float* positions, float* velocities, size_t count) {
  if (!sim) return fail(nullptr, VLJ_E_INVALID, "sim must not be NULL");
  if (count < sim->n_total) return fail(sim, VLJ_E_INVALID, "the output holds fewer particles than the handle");
  int rc = vlj_render(sim);
  if (rc != VLJ_OK) return rc;
  const size_t bytes = (size_t)sim->n_total * 16;
  if (positions && bytes) HIP_TRY(sim, hipMemcpyAsync(positions, sim->B.render_pos, bytes, hipMemcpyDeviceToHost, sim->stream));
  if (velocities && bytes) HIP_TRY(sim, hipMemcpyAsync(velocities, sim->B.render_vel, bytes, hipMemcpyDeviceToHost, sim->stream));
  HIP_TRY(sim, hipStreamSynchronize(sim->stream));
  return VLJ_OK;
}

int vlj_set_drawn(vlj_sim* sim, uint32_t drawn) {
  if (!sim) return fail(nullptr, VLJ_E_INVALID, "sim must not be NULL");
  if (drawn > 1u) return fail(sim, VLJ_E_INVALID, "drawn is 0 (neighbours) or 1 (temperature)");
  HIP_TRY(sim, hipStreamSynchronize(sim->stream));
  for (size_t r = 0; r < sim->rep.size(); r++) {
    sim->rep[r].drawn = drawn;
    HIP_TRY(sim, hipMemcpyAsync((void*)&sim->B.rep[r], &sim->rep[r], sizeof(Replica), hipMemcpyHostToDevice,
                                sim->stream));
  }
  HIP_TRY(sim, hipStreamSynchronize(sim->stream));
  return VLJ_OK;
}

int vlj_replica_info(const vlj_sim* sim, size_t replica, vlj_settings* settings, uint64_t* count,
                     size_t* replicas) {
  if (!sim) return fail(nullptr, VLJ_E_INVALID, "sim must not be NULL");
  if (replicas) *replicas = sim->rep.size();
  if (replica >= sim->rep.size()) return fail(nullptr, VLJ_E_INVALID, "replica out of range");
  if (settings) *settings = sim->host[replica].settings;
  if (count) *count = sim->rep[replica].count;
  return VLJ_OK;
}

int vlj_load(const char* path, vlj_sim** out) {
  if (!path || !out) return fail(nullptr, VLJ_E_INVALID, "path and out must not be NULL");
  *out = nullptr;
  FileIn r;
  r.f = std::fopen(path, "rb");
  if (!r.f) return fail(nullptr, VLJ_E_INVALID, std::string("could not open ") + path);
  auto bad = [&](const std::string& why) {
    std::fclose(r.f);
    return fail(nullptr, VLJ_E_INVALID, std::string(path) + ": " + why);
  };
  StateHeader h{};
  r.get(&h, sizeof h);
  if (!r.ok || std::memcmp(h.magic, kStateMagic, 8) != 0) return bad("not a state file");
  if (h.version != kStateVersion || h.sizeof_replica != sizeof(Replica) || h.sizeof_control != sizeof(Control) ||
      h.sizeof_settings != sizeof(vlj_settings) || h.sizeof_brick != sizeof(Brick))
    return bad("state file written by another version of the library");
  if (h.replicas == 0 || h.replicas > (1u << 24)) return bad("bad replica count");
  const size_t R = h.replicas;
  std::vector<vlj_settings> settings(R);
  std::vector<Replica> rep(R);
  r.get(settings.data(), R * sizeof(vlj_settings));
  r.get(rep.data(), R * sizeof(Replica));
  if (!r.ok) return bad("truncated");
  uint64_t atoms = 0, cells = 0;
  for (const Replica& x : rep) {
    if (x.atom_off != atoms || x.cell_off != cells) return bad("inconsistent replica table");
    atoms += x.count;
    cells += x.total_cells;
  }
  if (atoms != h.n_total || cells != h.cells_total) return bad("inconsistent sizes");
  std::vector<uint64_t> prof_samples(R);
  std::vector<std::vector<uint32_t>> bx(R), by(R);
  std::vector<std::vector<uint8_t>> bz(R), sp(R);
  for (size_t k = 0; k < R; k++) {
    r.get(&prof_samples[k], sizeof(uint64_t));
    if (!read_vec(r, bx[k], rep[k].cells[2]) || !read_vec(r, by[k], rep[k].cells[2]) ||
        !read_vec(r, bz[k], rep[k].cells[2]) || !read_vec(r, sp[k], rep[k].cells[2]))
      return bad("truncated");
    // Depths: 1, or 2 followed by the 0 of the layer that brick also holds.
    const std::vector<uint8_t>& d = bz[k];
    for (size_t z = 0; z < d.size(); z++)
      if (d[z] > 2u || (d[z] == 2u && (z + 1 == d.size() || d[z + 1] != 0u)) || (d[z] == 0u && (z == 0 || d[z - 1] != 2u)))
        return bad("replica " + std::to_string(k) + ": bad brick depth of layer " + std::to_string(z));
  }
  if (h.n_dense > h.n_bricks) return bad("bad brick table");
  std::vector<Brick> bricks;
  std::vector<Control> ctl;
  StateArrays a;
  const size_t N = h.n_total, C = h.cells_total, nb = h.n_bricks, ng = (nb + kCutGroup - 1) / kCutGroup;
  bool ok = read_vec(r, bricks, nb) && read_vec(r, ctl, R);
  for (int k = 0; k < 12 && ok; k++) ok = read_vec(r, a.f64[k], N);
  ok = ok && read_vec(r, a.orig, N) && read_vec(r, a.cell_of, N) && read_vec(r, a.nbr_count, N) &&
       read_vec(r, a.cell_start, C) && read_vec(r, a.cell_n, C) && read_vec(r, a.nbr, h.list_total) &&
       read_vec(r, a.cut, nb) && read_vec(r, a.parts, nb * kHomeCells) && read_vec(r, a.cut_groups, ng) &&
       read_vec(r, a.profile, h.prof_total) && read_vec(r, a.hc, (size_t)h.hc_cap * 3);
  if (!ok) return bad("truncated");
  std::fclose(r.f);
  {
    std::string why = check_state(rep, bricks, h, ctl, a);
    if (!why.empty()) return fail(nullptr, VLJ_E_INVALID, std::string(path) + ": " + why);
  }
  for (const Brick& g : bricks) {  // each brick as its layer's shape gives it
    const Replica& x = rep[g.rep];
    const uint32_t nx = x.cells[0], ny = x.cells[1], z = g.cell0 / (nx * ny);
    if (z >= bz[g.rep].size()) return fail(nullptr, VLJ_E_INVALID, std::string(path) + ": bad brick table");
    const uint32_t y0 = (g.cell0 / nx) % ny, x0 = g.cell0 % nx, bw = bx[g.rep][z], bh = by[g.rep][z];
    if (bw == 0 || bh == 0 || g.gz != bz[g.rep][z] || x0 % bw != 0 || y0 % bh != 0 || g.gx != std::min(bw, nx - x0) ||
        g.gy != std::min(bh, ny - y0))
      return fail(nullptr, VLJ_E_INVALID, std::string(path) + ": a brick does not match its layer's shape");
  }
  // The particles in original order, to build the handle from.
  std::vector<vlj_particle> parts(N);
  std::vector<uint64_t> offsets(R + 1, 0);
  for (size_t k = 0; k < R; k++) {
    offsets[k + 1] = offsets[k] + rep[k].count;
    for (uint32_t i = rep[k].atom_off; i < rep[k].atom_off + rep[k].count; i++) {
      const uint32_t o = a.orig[i];
      if (o >= rep[k].count) return fail(nullptr, VLJ_E_INVALID, std::string(path) + ": bad original index");
      vlj_particle& p = parts[offsets[k] + o];
      p.pos[0] = a.f64[0][i];
      p.pos[1] = a.f64[1][i];
      p.pos[2] = a.f64[2][i];
      p.vel[0] = a.f64[3][i];
      p.vel[1] = a.f64[4][i];
      p.vel[2] = a.f64[5][i];
    }
  }
  vlj_sim* s = nullptr;
  int rc = create(parts.data(), offsets.data(), settings.data(), R, &s, rep.data());
  if (rc != VLJ_OK) return rc;
  auto reject = [&](int code, const std::string& why) {
    vlj_destroy(s);
    return fail(nullptr, code, std::string(path) + ": " + why);
  };
  if (s->n_total != N || s->cells_total != C || s->list_total != h.list_total || s->prof_total != h.prof_total)
    return reject(VLJ_E_INVALID, "the settings give another layout than the saved one");
  for (size_t k = 0; k < R; k++) {
    const Replica& x = s->rep[k];
    const Replica& y = rep[k];
    if (x.count != y.count || x.total_cells != y.total_cells || x.capacity != y.capacity ||
        x.list_off != y.list_off || x.prof_off != y.prof_off || x.cells[0] != y.cells[0] ||
        x.cells[1] != y.cells[1] || x.cells[2] != y.cells[2] || x.atom_block_off != y.atom_block_off ||
        x.scan_block_off != y.scan_block_off || x.atom_off != y.atom_off || x.cell_off != y.cell_off ||
        x.scan_blocks != y.scan_blocks || x.scan_cells != y.scan_cells || x.atom_blocks != y.atom_blocks ||
        x.prof_bins != y.prof_bins)
      return reject(VLJ_E_INVALID, "replica " + std::to_string(k) + ": the settings give another layout than the saved one");
  }
  // The saved bricks replace the ones planned from the particles.
  Buffers& B = s->B;
  if (nb != s->bricks.size()) {
    free_device(s, (void*)B.bricks);
    free_device(s, B.cut);
    free_device(s, B.parts);
    free_device(s, B.cut_groups);
    Brick* gb = nullptr;
    rc = dev_alloc(s, &gb, nb);
    if (rc == VLJ_OK) rc = dev_alloc(s, &B.cut, nb);
    if (rc == VLJ_OK) rc = dev_alloc(s, &B.parts, nb * kHomeCells);
    if (rc == VLJ_OK) rc = dev_alloc(s, &B.cut_groups, ng);
    if (rc != VLJ_OK) return reject(rc, s->error);
    B.bricks = gb;
    B.n_bricks = (uint32_t)nb;
  }
  s->bricks = bricks;
  s->n_dense = h.n_dense;
  B.n_dense = h.n_dense;
  B.parts_on = std::any_of(a.cut.begin(), a.cut.end(), [](uint32_t np) { return np != 0u; }) ? 1u : 0u;
  s->rep_bricks.assign(R, 0u);
  for (const Brick& g : bricks) s->rep_bricks[g.rep]++;
  s->tuned = true;  // a loaded handle keeps the saved plan
  s->rep = rep;
  for (size_t k = 0; k < R; k++) {
    s->host[k].prof_samples = prof_samples[k];
    s->host[k].brick_x = bx[k];
    s->host[k].brick_y = by[k];
    s->host[k].brick_z = bz[k];
    s->host[k].sparse = sp[k];
  }
  s->any_heat_exchange = false;
  for (const Replica& x : rep) s->any_heat_exchange = s->any_heat_exchange || x.hex_on;
  s->steps = h.steps;
  s->primed = h.primed != 0;
  s->hc_samples = h.hc_samples;
  if (h.hc_cap) {
    hipError_t e = hipMalloc((void**)&B.hc, (size_t)h.hc_cap * 3 * sizeof(long long));
    if (e != hipSuccess) return reject(VLJ_E_HIP, std::string("hipMalloc: ") + hipGetErrorString(e));
  }
  double* const* f = state_f64(B);
  rc = upload(s, (Brick*)B.bricks, bricks);
  if (rc == VLJ_OK) rc = upload(s, (Replica*)B.rep, rep);
  if (rc == VLJ_OK) rc = upload(s, B.ctl, ctl);
  for (int k = 0; k < 12 && rc == VLJ_OK; k++) rc = upload(s, f[k], a.f64[k]);
  if (rc == VLJ_OK) rc = upload(s, B.orig, a.orig);
  if (rc == VLJ_OK) rc = upload(s, B.cell_of, a.cell_of);
  if (rc == VLJ_OK) rc = upload(s, B.nbr_count, a.nbr_count);
  if (rc == VLJ_OK) rc = upload(s, B.cell_start, a.cell_start);
  if (rc == VLJ_OK) rc = upload(s, B.cell_n, a.cell_n);
  if (rc == VLJ_OK) rc = upload(s, B.nbr, a.nbr);
  if (rc == VLJ_OK) rc = upload(s, B.cut, a.cut);
  if (rc == VLJ_OK) rc = upload(s, B.parts, a.parts);
  if (rc == VLJ_OK) rc = upload(s, B.cut_groups, a.cut_groups);
  if (rc == VLJ_OK) {
    uint32_t n_cut = 0;
    for (uint32_t c : a.cut) n_cut += c != 0u ? 1u : 0u;
    rc = upload(s, B.flags + 1, std::vector<uint32_t>{n_cut});
  }
  if (rc == VLJ_OK) rc = upload(s, B.profile, a.profile);
  if (rc == VLJ_OK) rc = upload(s, B.hc, a.hc);
  if (rc != VLJ_OK) return reject(rc, s->error);
  *out = s;
  return VLJ_OK;
}

}  // extern "C"
