// Host-side launchers of the kernels in vlj_kernels.hip.
#pragma once

#include <hip/hip_runtime.h>

#include "vlj_device.h"

namespace vlj {

// A second stream for the sparse-layout brick launches, forked from and
// joined back into the main stream by events. The two layouts' bricks hold
// disjoint atoms, so their launches may overlap: the sparse bricks fill the
// GPU while the dense launch's last workgroups run.
struct Side {
  hipStream_t stream = nullptr;
  hipEvent_t fork = nullptr, join = nullptr;
};

// One candidate the brick planner's tuner measured: the fill its bricks were
// planned with, the mean pair-force time per step and the time of the list
// build over the whole table, from HIP events on the handle's stream.
struct TuneTrial {
  double fill, force_us, build_us;
  uint32_t bricks;
};

hipError_t launch_decide(const Buffers& B, hipStream_t s);
// The drift and the step's rebuild decision; keep_v: positions only (the
// drift-ahead form's first step).
hipError_t launch_drift(const Buffers& B, bool langevin, bool keep_v, hipStream_t s);
hipError_t launch_rebuild(const Buffers& B, hipStream_t s, const Side* side = nullptr);
// The lists of every replica again from the positions of the last build
// (xb), for a new brick table; the cut marks of the table must start at
// zero. Leaves the state as it is.
hipError_t launch_relist(const Buffers& B, hipStream_t s);
// Nose-Hoover half step: kinetic energy, xi / eta / scale; with `scale` the
// velocities are scaled too (the second half step of a step).
hipError_t launch_nh_half(const Buffers& B, bool scale, hipStream_t s);
// Heat exchange after the step (sums, scales, apply).
hipError_t launch_heat_exchange(const Buffers& B, hipStream_t s);
// One heat-current sample into sample `sample` of the series.
hipError_t launch_heat_current(const Buffers& B, bool shifted, uint32_t sample, hipStream_t s);
// kick 0: accelerations only; 1: and the second half kick; 2: the
// drift-ahead form (NVE): the whole velocity step and the next step's drift
// into the second position set.
hipError_t launch_force(const Buffers& B, int precision, bool shifted, int kick, hipStream_t s,
                        const Side* side = nullptr);
hipError_t launch_measure(const Buffers& B, int precision, bool shifted, hipStream_t s);
// wide: every replica's profile slabs are at least a cutoff wide.
hipError_t launch_profile(const Buffers& B, bool shifted, bool wide, hipStream_t s);
// The float4 render arrays (positions; velocities with the drawn value in w).
hipError_t launch_render(const Buffers& B, hipStream_t s);
hipError_t launch_zero(long long* p, uint32_t n, hipStream_t s);
hipError_t launch_rcp(const double* x, double* out, size_t n, hipStream_t s);
hipError_t launch_rcp32(const float* x, float* rcp, float* root, size_t n, hipStream_t s);
hipError_t launch_fixed_sum(const double* values, size_t n, double scale, long long* out,
                            hipStream_t s);
// The same sum with no allocation: partial sums in a file-scope device
// array of kFixedSumBlocks entries (looked up once per device); *out is
// device memory. Calls that overlap on one GPU race.
hipError_t launch_fixed_sum_noalloc(const double* values, size_t n, double scale, long long* out,
                                    hipStream_t s);
hipError_t launch_copy(const void* src, void* dst, size_t words, hipStream_t s);
hipError_t launch_fma(double* out, uint32_t blocks, uint32_t iters, hipStream_t s);

}  // namespace vlj
