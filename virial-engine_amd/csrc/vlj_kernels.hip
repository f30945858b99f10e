// HIP kernels of the Lennard-Jones path, written for CDNA4 (gfx950, wave64).
//
// One step is (lennard_jones.rs:1092-1138):
//   drift       half kick + drift (or BAOAB), displacement since the last
//               build, one maximum per workgroup (lennard_jones.wgsl:204-276)
//   decide      per replica: rebuild or not (lennard_jones.wgsl:278-307)
//   bin         wrap into the box, cell of each atom, count   } rebuild chain,
//   scan        exclusive prefix sum of counts                } each kernel
//   scatter     atoms into their cell's range                 } returns at once
//   sort        each cell's slots ascending (wave per cell)   } when its replica
//   reorder     atoms into cell order                         } does not rebuild
//   build       brick-staged Verlet lists                     } (grid.wgsl, wgsl:419-499)
//   force       brick-staged pair forces + second half kick (wgsl:501-574)
// The sequence is static, so it is captured once into a hipGraph.
#include <mutex>

#include "vlj_device.h"
#include "vlj_launch.h"

namespace vlj {

// ---------------------------------------------------------------------------
// Wave64 / workgroup helpers
// ---------------------------------------------------------------------------

__device__ __forceinline__ unsigned long long wave_max_u64(unsigned long long v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    unsigned long long o = __shfl_xor(v, off, 64);
    v = o > v ? o : v;
  }
  return v;
}

__device__ __forceinline__ long long wave_sum_i64(long long v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

__device__ __forceinline__ uint32_t wave_incl_scan_u32(uint32_t v, int lane) {
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    uint32_t o = __shfl_up(v, off, 64);
    if (lane >= off) v += o;
  }
  return v;
}

// Reads a word another workgroup wrote, past this CU's L1.
__device__ __forceinline__ uint32_t load_agent(const uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// True in exactly one workgroup: the last of the grid to get here. All
// earlier workgroups' writes are visible to it afterwards.
__device__ __forceinline__ bool last_block(uint32_t* ticket) {
  __shared__ bool is_last;
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    uint32_t t = atomicAdd(ticket, 1u);
    is_last = (t == gridDim.x - 1);
  }
  __syncthreads();
  if (is_last) __threadfence();
  return is_last;
}

// ---------------------------------------------------------------------------
// Rebuild decision (lennard_jones.wgsl:278-307), one thread per replica.
// ---------------------------------------------------------------------------

__device__ void decide_from(const Buffers& B, uint32_t r, double moved2, bool advance_step) {
  Control* c = B.ctl + r;
  const Replica R = B.rep[r];
  uint32_t built = c->built;
  uint32_t rebuilds = c->rebuilds;
  bool rebuild = !built || moved2 > R.half_skin2;
  c->rebuild = rebuild ? 1u : 0u;
  c->reorder = rebuild ? 1u : 0u;  // R8: every rebuild moves atoms into cell order
  if (rebuild) {
    c->rebuilds = rebuilds + 1u;
    c->built = 1u;
    atomicAdd(B.rebuilding, 1u);
  }
  if (advance_step) c->step = c->step + 1ull;
}

// One workgroup per replica: the largest moved^2 over the replica's drift
// workgroups (bmax), then the decision. `advance` also counts the step and
// zeroes the replica's drift-ahead maxima (the force launch of a drift-ahead
// step takes them); the first workgroup zeroes the next step's rebuild word,
// which no kernel of this step reads.
__global__ void __launch_bounds__(kBlock) k_decide(Buffers B, int advance) {
  const uint32_t r = blockIdx.x;
  const Replica R = B.rep[r];
  unsigned long long m = 0ull;
  if (advance) {
    // 16 loads in flight per thread (4096 workgroup maxima of a 1M-atom
    // replica in one round trip)
    const uint32_t end = R.atom_block_off + R.atom_blocks;
    for (uint32_t b0 = R.atom_block_off + threadIdx.x; b0 < end; b0 += 16u * kBlock) {
      unsigned long long v[16];
#pragma unroll
      for (uint32_t j = 0; j < 16u; j++) v[j] = b0 + j * kBlock < end ? B.bmax[b0 + j * kBlock] : 0ull;
#pragma unroll
      for (uint32_t j = 0; j < 16u; j++) m = v[j] > m ? v[j] : m;
    }
    if (threadIdx.x < kAheadSlots) {
      B.amax[(size_t)r * kAheadSlots + threadIdx.x] = 0ull;
      B.amax_next[(size_t)r * kAheadSlots + threadIdx.x] = 0ull;
    }
    if (r == 0 && threadIdx.x == 0) *B.rebuilding_next = 0u;
  }
  __shared__ unsigned long long wmax[kBlock / 64];
  m = wave_max_u64(m);
  if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int w = 1; w < kBlock / 64; w++) m = wmax[w] > m ? wmax[w] : m;
    decide_from(B, r, __longlong_as_double((long long)m), advance != 0);
  }
}

// ---------------------------------------------------------------------------
// Drift: v += a dt/2; x += v dt (or BAOAB); wrap; max moved^2 since build.
// ---------------------------------------------------------------------------

// LANGEVIN: some replica of the launch uses BAOAB (others stay NVE).
// keep_v: the velocities stay v(t) (the drift-ahead force forms the half
// step again, as here, before its kick); only the positions move.
template <bool LANGEVIN>
__global__ void __launch_bounds__(kBlock) k_drift(Buffers B, int keep_v) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *B.rebuilding = 0u;  // counted again by k_decide
  const uint32_t b = dealt_index(blockIdx.x);
  if (b >= B.n_atom_blocks) return;
  const Range blk = B.atom_blocks[b];
  const Replica R = B.rep[blk.rep];
  const uint32_t i = blk.first + threadIdx.x;
  unsigned long long moved = 0ull;
  if (i < blk.end) {
    const bool baoab = LANGEVIN && R.thermostat == 1u;
    double n[3] = {0.0, 0.0, 0.0};
    if (baoab) langevin_noise(R.seed, 0u, B.orig[i], B.ctl[blk.rep].step, n);
    double* X[3] = {B.x, B.y, B.z};
    double* V[3] = {B.vx, B.vy, B.vz};
    const double* A[3] = {B.ax, B.ay, B.az};
    const double* XB[3] = {B.xb, B.yb, B.zb};
    double d[3];
    const bool nh = R.thermostat == 2u;
    const double sm1 = nh ? B.ctl[blk.rep].sm1 : 0.0;
#pragma unroll
    for (int k = 0; k < 3; k++) {
      double v0 = V[k][i];
      if (nh) v0 = __builtin_fma(v0, sm1, v0);  // R13: the first half step's scaling
      double v = __builtin_fma(A[k][i], R.half_dt, v0);
      double x = X[k][i];
      if (baoab) {
        x = __builtin_fma(v, R.half_dt, x);
        v = __builtin_fma(R.c2, n[k], R.c1 * v);
        x = __builtin_fma(v, R.half_dt, x);
      } else {
        x = __builtin_fma(v, R.dt, x);
      }
      if (!keep_v) V[k][i] = v;
      X[k][i] = x;  // wrapped at the next rebuild (R1)
      d[k] = x - XB[k][i];
    }
    moved = (unsigned long long)__double_as_longlong(dist2(d[0], d[1], d[2]));
  }
  // One value per workgroup, no atomics (lennard_jones.wgsl:215-220 uses
  // one global atomicMax per workgroup); k_decide reduces them. (A ticket
  // per workgroup, so that the replica's last workgroup could decide, is
  // itself an atomic on one address per replica: from the 4096 workgroups
  // of a 1M-atom replica it made a step 130 us longer.)
  __shared__ unsigned long long wmax[kBlock / 64];
  moved = wave_max_u64(moved);
  if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = moved;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long m = wmax[0];
#pragma unroll
    for (int w = 1; w < kBlock / 64; w++) m = wmax[w] > m ? wmax[w] : m;
    B.bmax[b] = m;
  }
}

// ---------------------------------------------------------------------------
// Nose-Hoover half step (lennard_jones.wgsl:309-397, R13): the kinetic
// energy in fixed point per atom block, then per replica xi, eta and the
// velocity scale s - 1 = exp(-xi dt / 2) - 1 by its series, in FP64; the
// second half step's scaling after the force kick.
// ---------------------------------------------------------------------------

__global__ void __launch_bounds__(kBlock) k_ke_sum(Buffers B) {
  const uint32_t b = dealt_index(blockIdx.x);
  if (b >= B.n_atom_blocks) return;
  const Range blk = B.atom_blocks[b];
  const Replica R = B.rep[blk.rep];
  if (R.thermostat != 2u) return;
  const uint32_t i = blk.first + threadIdx.x;
  long long ke = 0;
  if (i < blk.end) ke = to_fixed(0.5 * dist2(B.vx[i], B.vy[i], B.vz[i]), R.energy_fixed);
  __shared__ long long part[kBlock / 64];
  ke = wave_sum_i64(ke);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = ke;
  __syncthreads();
  if (threadIdx.x == 0) {
    long long t = 0;
#pragma unroll
    for (int w = 0; w < kBlock / 64; w++) t += part[w];
    B.kepart[b] = t;
  }
}

__global__ void __launch_bounds__(kBlock) k_nh_update(Buffers B) {
  const uint32_t r = blockIdx.x;
  const Replica R = B.rep[r];
  if (R.thermostat != 2u) return;
  long long t = 0;
  for (uint32_t b = R.atom_block_off + threadIdx.x; b < R.atom_block_off + R.atom_blocks; b += kBlock)
    t += B.kepart[b];
  __shared__ long long part[kBlock / 64];
  t = wave_sum_i64(t);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = t;
  __syncthreads();
  if (threadIdx.x == 0) {
    long long raw = 0;
#pragma unroll
    for (int w = 0; w < kBlock / 64; w++) raw += part[w];
    Control* c = B.ctl + r;
    const double ke = (double)raw / R.energy_fixed;
    const double quarter = 0.25 * R.dt;
    double xi = __builtin_fma((2.0 * ke - R.nh_gkt) / R.nh_q, quarter, c->xi);
    const double x = xi * R.half_dt;
    const double sm1 = -x * (1.0 - x * (0.5 - x * (1.0 / 6.0 - x / 24.0)));
    c->eta = __builtin_fma(xi, R.half_dt, c->eta);
    const double ke2 = __builtin_fma(ke, sm1 * (2.0 + sm1), ke);
    c->xi = __builtin_fma((2.0 * ke2 - R.nh_gkt) / R.nh_q, quarter, xi);
    c->sm1 = sm1;
  }
}

__global__ void __launch_bounds__(kBlock) k_nh_scale(Buffers B) {
  const uint32_t ab = dealt_index(blockIdx.x);
  if (ab >= B.n_atom_blocks) return;
  const Range blk = B.atom_blocks[ab];
  const Replica R = B.rep[blk.rep];
  if (R.thermostat != 2u) return;
  const uint32_t i = blk.first + threadIdx.x;
  if (i >= blk.end) return;
  const double sm1 = B.ctl[blk.rep].sm1;
  B.vx[i] = __builtin_fma(B.vx[i], sm1, B.vx[i]);
  B.vy[i] = __builtin_fma(B.vy[i], sm1, B.vy[i]);
  B.vz[i] = __builtin_fma(B.vz[i], sm1, B.vz[i]);
}

// ---------------------------------------------------------------------------
// Heat exchange, eHEX/a (Wirnsberger, Frenkel and Dellago 2015, eq. 20;
// lennard_jones.wgsl:732-870, R15): after the step the hot slab gains
// hex_energy and the cold slab loses it, by scaling velocities about each
// slab's centre-of-mass velocity, with the O(dt^3) position correction.
// Slab sums in 64-bit fixed point per atom block; per replica the scales.
// ---------------------------------------------------------------------------

__device__ __forceinline__ int hex_slab(const Replica& R, double z) {
  const double zw = wrap_one(z, R.box[2]);
  if (zw >= R.hex_slabs[0] && zw < R.hex_slabs[1]) return 0;
  if (zw >= R.hex_slabs[2] && zw < R.hex_slabs[3]) return 1;
  return 2;
}

__global__ void __launch_bounds__(kBlock) k_hex_sum(Buffers B) {
  const uint32_t b = dealt_index(blockIdx.x);
  if (b >= B.n_atom_blocks) return;
  const Range blk = B.atom_blocks[b];
  const Replica R = B.rep[blk.rep];
  if (!R.hex_on) return;
  const uint32_t i = blk.first + threadIdx.x;
  long long q[kHexSums];
#pragma unroll
  for (uint32_t k = 0; k < kHexSums; k++) q[k] = 0;
  if (i < blk.end) {
    const int slab = hex_slab(R, B.z[i]);
    if (slab < 2) {
      const double v[3] = {B.vx[i], B.vy[i], B.vz[i]};
      const double f[3] = {B.ax[i], B.ay[i], B.az[i]};
      const long long c = 1;
      long long t[9];
      t[0] = c;
#pragma unroll
      for (int k = 0; k < 3; k++) {
        t[1 + k] = to_fixed(v[k], R.momentum_fixed);
        t[6 + k] = to_fixed(f[k], R.momentum_fixed);
      }
      t[4] = to_fixed(dist2(v[0], v[1], v[2]), R.energy_fixed);
      t[5] = to_fixed(__builtin_fma(f[2], v[2], __builtin_fma(f[1], v[1], f[0] * v[0])), R.energy_fixed);
#pragma unroll
      for (int k = 0; k < 9; k++) {
        q[k] = slab == 0 ? t[k] : 0;
        q[9 + k] = slab == 1 ? t[k] : 0;
      }
    }
  }
  __shared__ long long part[kBlock / 64][kHexSums];
#pragma unroll
  for (uint32_t k = 0; k < kHexSums; k++) {
    const long long w = wave_sum_i64(q[k]);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6][k] = w;
  }
  __syncthreads();
  if (threadIdx.x < kHexSums) {
    long long t = 0;
#pragma unroll
    for (int w = 0; w < kBlock / 64; w++) t += part[w][threadIdx.x];
    B.hexpart[(size_t)b * kHexSums + threadIdx.x] = t;
  }
}

__global__ void __launch_bounds__(kBlock) k_hex_update(Buffers B) {
  const uint32_t r = blockIdx.x;
  const Replica R = B.rep[r];
  if (!R.hex_on) return;
  __shared__ long long sums[kHexSums];
  if (threadIdx.x < kHexSums) {
    long long t = 0;
    for (uint32_t b = R.atom_block_off; b < R.atom_block_off + R.atom_blocks; b++)
      t += B.hexpart[(size_t)b * kHexSums + threadIdx.x];
    sums[threadIdx.x] = t;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  double* out = B.hexstate + (size_t)r * kHexState;
  const double dt3 = R.dt * R.dt * R.dt;
  for (int slab = 0; slab < 2; slab++) {
    const long long* q = sums + 9 * slab;
    double* o = out + 9 * slab;
    const double m = (double)q[0];
#pragma unroll
    for (int k = 0; k < 9; k++) o[k] = 0.0;
    if (!(m > 0.0)) continue;
    double pm[3], fm[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
      pm[k] = (double)q[1 + k] / R.momentum_fixed;
      fm[k] = (double)q[6 + k] / R.momentum_fixed;
      o[k] = pm[k] / m;      // centre-of-mass velocity
      o[3 + k] = fm[k] / m;  // mean force
    }
    const double v2 = (double)q[4] / R.energy_fixed;
    const double fv = (double)q[5] / R.energy_fixed;
    const double pp = __builtin_fma(pm[2], pm[2], __builtin_fma(pm[1], pm[1], pm[0] * pm[0]));
    const double kinetic = 0.5 * v2 - 0.5 * pp / m;
    const double energy = slab == 0 ? R.hex_energy : -R.hex_energy;
    const double flux = energy / R.dt;
    const double x = energy / kinetic;
    o[6] = x / (1.0 + __builtin_sqrt(1.0 + x));
    const double power = fv - __builtin_fma(fm[2], o[2], __builtin_fma(fm[1], o[1], fm[0] * o[0]));
    o[7] = dt3 * flux / (2.0 * kinetic * kinetic) * (flux / 48.0 + power / 6.0);
    o[8] = dt3 * flux / (12.0 * kinetic);
  }
}

__global__ void __launch_bounds__(kBlock) k_hex_apply(Buffers B) {
  const uint32_t ab = dealt_index(blockIdx.x);
  if (ab >= B.n_atom_blocks) return;
  const Range blk = B.atom_blocks[ab];
  const Replica R = B.rep[blk.rep];
  if (!R.hex_on) return;
  const uint32_t i = blk.first + threadIdx.x;
  if (i >= blk.end) return;
  const int slab = hex_slab(R, B.z[i]);
  if (slab == 2) return;
  const double* o = B.hexstate + (size_t)blk.rep * kHexState + 9 * slab;
  double* X[3] = {B.x, B.y, B.z};
  double* V[3] = {B.vx, B.vy, B.vz};
  const double* A[3] = {B.ax, B.ay, B.az};
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const double v = V[k][i];
    const double rel = v - o[k];
    const double shift = o[7] * rel - o[8] * (A[k][i] - o[3 + k]);
    X[k][i] = X[k][i] - shift;
    V[k][i] = __builtin_fma(rel, o[6], v);
  }
}

// ---------------------------------------------------------------------------
// Cell grid (grid.wgsl:22-148) with a deterministic order: counts, scan,
// scatter in any order, then a rank sort inside each cell by slot index.
// ---------------------------------------------------------------------------

// Adds each active lane's 1 to counter[key] and returns the counter's value
// before this lane's add. Lanes with the same key side by side (atoms are
// in cell order after every rebuild, so runs of ~20 are the rule) share one
// atomic: the run's first lane adds the run length for all of them.
__device__ __forceinline__ uint32_t run_atomic_add(uint32_t* counter, uint32_t key, bool active) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t k = active ? key : 0xffffffffu;
  const uint32_t prev = __shfl_up(k, 1, 64);
  const bool head = lane == 0u || prev != k;
  const unsigned long long heads = __builtin_amdgcn_ballot_w64(head);
  const unsigned long long upto = lane == 63u ? ~0ull : ((2ull << lane) - 1ull);  // lanes <= this one
  const uint32_t leader = 63u - (uint32_t)__builtin_clzll(heads & upto);
  const unsigned long long above = heads & ~upto;
  const uint32_t next = above ? (uint32_t)__builtin_ctzll(above) : 64u;
  uint32_t base = 0;
  if (head && active) base = atomicAdd(counter + key, next - lane);
  return (uint32_t)__shfl(base, (int)leader, 64) + (lane - leader);
}

// With bin_decides (drift-ahead steps) every workgroup first takes the rebuild
// decision of its replica from the maxima the last force launch left (the
// same rule as k_decide's), and the replica's first atom block records it;
// such steps come after a k_decide step, so `built` is already set and no
// workgroup reads a word the recording one writes.
__global__ void __launch_bounds__(kBlock) k_bin(Buffers B) {
  const uint32_t ab = dealt_index(blockIdx.x);
  if (B.bin_decides) {
    if (ab == 0 && threadIdx.x == 0) *B.rebuilding_next = 0u;
    if (ab >= B.n_atom_blocks) return;
    const uint32_t rep = B.atom_blocks[ab].rep;
    __shared__ unsigned long long moved;
    if (threadIdx.x < kAheadSlots) {
      const unsigned long long v = wave_max_u64(B.amax[(size_t)rep * kAheadSlots + threadIdx.x]);
      if (threadIdx.x == 0) moved = v;
    }
    __syncthreads();
    const double moved2 = __longlong_as_double((long long)moved);
    const bool rebuild = !B.ctl[rep].built || moved2 > B.rep[rep].half_skin2;
    if (ab == B.rep[rep].atom_block_off && threadIdx.x == 0) decide_from(B, rep, moved2, true);
    if (!rebuild) return;
  } else {
    if (*B.rebuilding == 0u) return;  // no replica rebuilds this step
    if (ab >= B.n_atom_blocks) return;
    if (!B.ctl[B.atom_blocks[ab].rep].rebuild) return;
  }
  const Range blk = B.atom_blocks[ab];
  const Replica R = B.rep[blk.rep];
  const uint32_t i = blk.first + threadIdx.x;
  const bool active = i < blk.end;
  uint32_t c = 0;
  if (active) {
    // R1: positions go back into the box when the lists are rebuilt
    const double x = wrap_one(B.x[i], R.box[0]), y = wrap_one(B.y[i], R.box[1]), z = wrap_one(B.z[i], R.box[2]);
    B.x[i] = x;
    B.y[i] = y;
    B.z[i] = z;
    c = R.cell_off + cell_local(R, x, y, z);
    B.cell_of[i] = c;
  }
  run_atomic_add(B.cell_count, c, active);
}

// Exclusive scan of counts inside each block of 1024 cells (4 per thread);
// the last workgroup scans the block totals of each rebuilding replica.
__global__ void __launch_bounds__(kBlock) k_scan(Buffers B) {
  if (*B.rebuilding == 0u) return;  // no replica rebuilds this step
  const Range blk = B.scan_blocks[blockIdx.x];
  const bool active = B.ctl[blk.rep].rebuild != 0u;
  __shared__ uint32_t wsum[kBlock / 64];
  if (active) {
    // each thread scans `per` consecutive cells (4 or 16, by the block's size)
    const uint32_t per = (blk.end - blk.first + kBlock - 1u) / kBlock;
    const uint32_t base = blk.first + threadIdx.x * per;
    constexpr int kMaxPer = kScanCellsSparse / kBlock;
    uint32_t v[kMaxPer];
    uint32_t local = 0;
#pragma unroll
    for (int k = 0; k < kMaxPer; k++) {
      uint32_t c = base + (uint32_t)k;
      v[k] = (uint32_t)k < per && c < blk.end ? B.cell_count[c] : 0u;
      local += v[k];
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t incl = wave_incl_scan_u32(local, lane);
    if (lane == 63) wsum[wave] = incl;
    __syncthreads();
    uint32_t before = 0, total = 0;
#pragma unroll
    for (int w = 0; w < kBlock / 64; w++) {
      uint32_t s = wsum[w];
      if (w < wave) before += s;
      total += s;
    }
    uint32_t run = before + incl - local;
#pragma unroll
    for (int k = 0; k < kMaxPer; k++) {
      uint32_t c = base + (uint32_t)k;
      if ((uint32_t)k < per && c < blk.end) B.cell_scan[c] = run;
      run += v[k];
    }
    if (threadIdx.x == 0) B.bsum[blockIdx.x] = total;
  }
  if (last_block(B.tickets + 1)) {
    for (uint32_t r = threadIdx.x; r < B.n_replicas; r += blockDim.x) {
      if (!B.ctl[r].rebuild) continue;
      const Replica R = B.rep[r];
      uint32_t acc = 0;
      // 16 block sums in flight at a time (a dilute replica has tens of blocks)
      for (uint32_t b0 = 0; b0 < R.scan_blocks; b0 += 16u) {
        uint32_t v[16];
#pragma unroll
        for (uint32_t k = 0; k < 16u; k++)
          v[k] = b0 + k < R.scan_blocks ? load_agent(B.bsum + R.scan_block_off + b0 + k) : 0u;
#pragma unroll
        for (uint32_t k = 0; k < 16u; k++) {
          if (b0 + k < R.scan_blocks) B.bprefix[R.scan_block_off + b0 + k] = acc;
          acc += v[k];
        }
      }
    }
    if (threadIdx.x == 0) atomicExch(B.tickets + 1, 0u);
  }
}

__device__ __forceinline__ uint32_t cell_first_slot(const Buffers& B, const Replica& R, uint32_t c) {
  uint32_t sb = R.scan_block_off + (c - R.cell_off) / R.scan_cells;
  return R.atom_off + B.cell_scan[c] + B.bprefix[sb];
}

__global__ void __launch_bounds__(kBlock) k_scatter(Buffers B) {
  if (*B.rebuilding == 0u) return;  // no replica rebuilds this step
  const uint32_t ab = dealt_index(blockIdx.x);
  if (ab >= B.n_atom_blocks) return;
  const Range blk = B.atom_blocks[ab];
  if (!B.ctl[blk.rep].rebuild) return;
  const Replica R = B.rep[blk.rep];
  const uint32_t i = blk.first + threadIdx.x;
  const bool active = i < blk.end;
  const uint32_t c = active ? B.cell_of[i] : 0u;
  const uint32_t rank = run_atomic_add(B.cell_cursor, c, active);
  if (active) B.unsorted[cell_first_slot(B, R, c) + rank] = i;
}

// One wave64 per cell: every entry's rank is the number of smaller entries
// in the cell, which puts the cell in ascending slot order whatever order
// the scatter produced (the result of grid.wgsl:129-148's insertion sort).
// Per-cell order (R3/R8): each cell's slots in ascending original slot,
// by rank (the number of smaller entries). A sort block covers kSortCells
// cells of a dense replica (a wave per cell) or kSortCellsSparse of a dilute
// one: there a thread ranks a cell of at most kSmallCell atoms on its own,
// and the waves take the larger cells of the block in turn.
__global__ void __launch_bounds__(kBlock) k_sort(Buffers B) {
  if (*B.rebuilding == 0u) return;  // no replica rebuilds this step
  const uint32_t sb = dealt_index(blockIdx.x);
  if (sb >= B.n_sort_blocks) return;
  const Range blk = B.sort_blocks[sb];
  if (!B.ctl[blk.rep].rebuild) return;
  const Replica R = B.rep[blk.rep];
  __shared__ uint32_t big_c[kBlock], big_start[kBlock], big_n[kBlock];
  __shared__ uint32_t n_big;
  if (threadIdx.x == 0) n_big = 0u;
  __syncthreads();
  const uint32_t c = blk.first + threadIdx.x;
  if (c < blk.end) {
    const uint32_t start = cell_first_slot(B, R, c);
    const uint32_t n = B.cell_count[c];
    B.cell_start[c] = start;
    B.cell_n[c] = n;
    B.cell_count[c] = 0u;
    B.cell_cursor[c] = 0u;
    if (n <= kSmallCell) {
      uint32_t e[kSmallCell];
#pragma unroll
      for (uint32_t m = 0; m < kSmallCell; m++) e[m] = m < n ? B.unsorted[start + m] : 0xffffffffu;
#pragma unroll
      for (uint32_t m = 0; m < kSmallCell; m++) {
        uint32_t rank = 0;
#pragma unroll
        for (uint32_t j = 0; j < kSmallCell; j++) rank += e[j] < e[m] ? 1u : 0u;
        if (m < n) B.sorted[start + rank] = e[m];
      }
    } else {
      const uint32_t k = atomicAdd(&n_big, 1u);
      big_c[k] = c;
      big_start[k] = start;
      big_n[k] = n;
    }
  }
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const uint32_t nb = n_big;
  for (uint32_t k = threadIdx.x >> 6; k < nb; k += kBlock / 64) {
    const uint32_t start = big_start[k], n = big_n[k];
    if (n <= 64u) {
      uint32_t e = lane < (int)n ? B.unsorted[start + lane] : 0xffffffffu;
      uint32_t rank = 0;
      for (uint32_t m = 0; m < n; m++) rank += (__shfl(e, (int)m, 64) < e) ? 1u : 0u;
      if (lane < (int)n) B.sorted[start + rank] = e;
    } else {
      for (uint32_t j = (uint32_t)lane; j < n; j += 64u) {
        uint32_t e = B.unsorted[start + j];
        uint32_t rank = 0;
        for (uint32_t m = 0; m < n; m++) rank += (B.unsorted[start + m] < e) ? 1u : 0u;
        B.sorted[start + rank] = e;
      }
    }
  }
}

// Slot t takes the atom at sorted[t] (lennard_jones.wgsl:419-452); the copy
// goes to the second state set and `build` copies it back.
__global__ void __launch_bounds__(kBlock) k_reorder(Buffers B) {
  if (*B.rebuilding == 0u) return;  // no replica rebuilds this step
  const uint32_t ab = dealt_index(blockIdx.x);
  if (ab >= B.n_atom_blocks) return;
  const Range blk = B.atom_blocks[ab];
  if (!B.ctl[blk.rep].reorder) return;
  const uint32_t t = blk.first + threadIdx.x;
  if (t >= blk.end) return;
  const uint32_t s = B.sorted[t];
  B.x2[t] = B.x[s];
  B.y2[t] = B.y[s];
  B.z2[t] = B.z[s];
  B.vx2[t] = B.vx[s];
  B.vy2[t] = B.vy[s];
  B.vz2[t] = B.vz[s];
  B.ax2[t] = B.ax[s];
  B.ay2[t] = B.ay[s];
  B.az2[t] = B.az[s];
  B.orig2[t] = B.orig[s];
  B.cell_of2[t] = B.cell_of[s];
  B.sorted[t] = t;
}

// ---------------------------------------------------------------------------
// Bricks: staging of a brick's neighbourhood in LDS
// ---------------------------------------------------------------------------

// LDS of a brick workgroup: the staged cells (first slot, staged offset,
// periodic image) and, when they fit, the staged positions with the image
// added, so a displacement is one subtraction (R4, R5). Every rebuild puts
// the atoms in cell order (R8), so staged cell c holds the slots
// start[c] .. start[c] + (off[c + 1] - off[c]).
constexpr uint32_t kCoarse = 16;  // staged atoms per entry of Stage::coarse

template <class L>
struct Stage {
  // Staged positions and, at L::kListPad, the sentinel. They open the
  // structure and are a multiple of 64 doubles long, so that x[e] and y[e]
  // are one ds_read2st64_b64 and z[e] one ds_read_b64, all three addressed
  // by 8 e with no further offset.
  double x[L::kStageMax + 64], y[L::kStageMax + 64], z[L::kStageMax + 64];
  uint32_t base[kStageCells];     // slot of staged atom e of cell c: base[c] + e (mod 2^32)
  uint32_t off[kStageCells + 1];
  uint32_t hoff[kHomeCells + 1];  // home atoms numbered home cell by home cell
  uint8_t image[kStageCells];     // (ix + 1) | (iy + 1) << 2 | (iz + 1) << 4, i in {-1, 0, 1}
  uint32_t ns;                    // staged cells
  uint32_t total;                 // staged atoms
  // the staged cell of every kCoarse-th staged atom (up to L::kStageMax)
  uint16_t coarse[L::kStageMax / kCoarse];
};

// Staged index of the cell at offset (dz, yy, xx) from the brick's corner
// (-1, -1, -1): dz in 0..gz+1, yy in 0..gy+1, xx in 0..gx+1.
__device__ __forceinline__ uint32_t staged_cell(const Brick& g, uint32_t dz, uint32_t yy, uint32_t xx) {
  return (dz * (g.gy + 2u) + yy) * (g.gx + 2u) + xx;
}

// Staged cell of staged atom e: the last cell whose first staged atom is at
// most e, by binary search over the staged offsets between the cells of
// the kCoarse-th atoms at or below e and after it (a dense brick's cells
// hold more than kCoarse atoms, so the search is over one or two cells).
template <class L>
__device__ __forceinline__ uint32_t stage_cell_of(const Stage<L>& S, uint32_t e) {
  uint32_t lo = 0, hi = S.ns - 1u;
  const uint32_t known = S.total < (uint32_t)L::kStageMax ? S.total : (uint32_t)L::kStageMax;
  if (e < known) {
    const uint32_t m = e / kCoarse;
    lo = S.coarse[m];
    if ((m + 1u) * kCoarse < known) hi = S.coarse[m + 1u];
  }
  while (lo < hi) {
    const uint32_t mid = (lo + hi + 1u) >> 1;
    if (S.off[mid] <= e) lo = mid; else hi = mid - 1u;
  }
  return lo;
}

constexpr uint32_t kNoImage = 1u | 1u << 2 | 1u << 4;  // the image code of no shift

__device__ __forceinline__ void image_shift(const Replica& R, uint32_t code, double& sx, double& sy, double& sz) {
  sx = (double)((int)(code & 3u) - 1) * R.box[0];
  sy = (double)((int)((code >> 2) & 3u) - 1) * R.box[1];
  sz = (double)((int)(code >> 4) - 1) * R.box[2];
}

// Wave priorities (s_setprio). While a brick is staged: with two workgroups
// on a CU, the one staging its next brick issues its loads and LDS stores
// ahead of the one summing pairs, so the CU spends less time with both
// waiting on memory (FP64 force kernel at N = 1M: 165 -> 153 us; 1 and 3
// alike). While the FP64 pair loop issues an octet's gathers: 1, below
// staging. Again on the present step (FCC-1M, atom-steps/s, one box):
// stage 2 / gathers 1 5.52e9-5.54e9, 1 / 1 5.53e9, 2 / 0 5.49e9, 2 / 2
// 5.51e9, 3 / 1 5.41e9.
#ifndef VLJ_STAGE_PRIO
#define VLJ_STAGE_PRIO 2
#endif
#ifndef VLJ_READ_PRIO
#define VLJ_READ_PRIO 1
#endif

// Fills S for brick g from the positions X, Y, Z. Every thread of the
// workgroup calls it. Returns whether the positions were staged (the
// neighbourhood holds at most L::kStageMax atoms).
template <class L>
__device__ bool stage_brick(const Buffers& B, const Replica& R, const Brick& g, const double* X, const double* Y,
                            const double* Z, Stage<L>& S) {
#if VLJ_STAGE_PRIO
  __builtin_amdgcn_s_setprio(VLJ_STAGE_PRIO);
#endif
  const int nx = (int)R.cells[0], ny = (int)R.cells[1], nz = (int)R.cells[2];
  const int c0 = (int)g.cell0;
  const int cx0 = c0 % nx, cy0 = (c0 / nx) % ny, cz = c0 / (nx * ny);
  const uint32_t sx = g.gx + 2u, sy = g.gy + 2u;
  const uint32_t ns = (g.gz + 2u) * sy * sx;  // at most kStageCells
  const uint32_t t = threadIdx.x;
  // thread t stages the kCellsPerThread consecutive cells from t * kCellsPerThread
  constexpr uint32_t kCPT = (uint32_t)L::kCellsPerThread;
  uint32_t n_c[kCPT], st_c[kCPT];
  bool shifted_cell = false;  // a cell of this thread is a periodic image
#pragma unroll
  for (uint32_t h = 0; h < kCPT; h++) {
    const uint32_t c = kCPT * t + h;
    n_c[h] = 0u;
    st_c[h] = 0u;
    if (c < ns) {
      const int xx = (int)(c % sx), yy = (int)((c / sx) % sy), dz = (int)(c / (sx * sy));
      const int ux = cx0 - 1 + xx, uy = cy0 - 1 + yy, uz = cz - 1 + dz;  // unwrapped
      const int cx = (ux + nx) % nx, cy = (uy + ny) % ny, ccz = (uz + nz) % nz;
      const uint32_t cell = R.cell_off + (uint32_t)((ccz * ny + cy) * nx + cx);
      st_c[h] = B.cell_start[cell];
      n_c[h] = B.cell_n[cell];
      // image of the cell seen from the brick (R4)
      const int ix = ux < 0 ? -1 : (ux >= nx ? 1 : 0);
      const int iy = uy < 0 ? -1 : (uy >= ny ? 1 : 0);
      const int iz = uz < 0 ? -1 : (uz >= nz ? 1 : 0);
      S.image[c] = (uint8_t)((ix + 1) | (iy + 1) << 2 | (iz + 1) << 4);
      shifted_cell = shifted_cell || ix != 0 || iy != 0 || iz != 0;
    }
  }
  // exclusive scan of the staged counts over the workgroup; each wave's
  // total carries in its top bit whether the wave staged an image cell
  __shared__ uint32_t wtot[L::kThreads / 64];
  uint32_t n_t = 0;
#pragma unroll
  for (uint32_t h = 0; h < kCPT; h++) n_t += n_c[h];
  const int lane = t & 63, wave = t >> 6;
  const uint32_t incl = wave_incl_scan_u32(n_t, lane);
  const bool wave_shifted = __builtin_amdgcn_ballot_w64(shifted_cell) != 0ull;
  if (lane == 63) wtot[wave] = incl | (wave_shifted ? 0x80000000u : 0u);
  __syncthreads();
  uint32_t before = 0, shifted_any = 0;
#pragma unroll
  for (int w = 0; w < L::kThreads / 64; w++) {
    const uint32_t x = wtot[w];
    before += w < wave ? x & 0x7fffffffu : 0u;
    shifted_any |= x >> 31;
  }
  uint32_t run = before + incl - n_t;
#pragma unroll
  for (uint32_t h = 0; h < kCPT; h++) {
    const uint32_t c = kCPT * t + h;
    if (c < ns) {
      S.off[c] = run;
      S.base[c] = st_c[h] - run;
      // the coarse index: cell c holds staged atoms run .. run + n - 1 (those
      // below kStageMax are indexed), known here without reading S.off back
      const uint32_t end = run + n_c[h] < (uint32_t)L::kStageMax ? run + n_c[h] : (uint32_t)L::kStageMax;
      for (uint32_t m = (run + kCoarse - 1u) / kCoarse; m * kCoarse < end; m++) S.coarse[m] = (uint16_t)c;
    }
    run += n_c[h];
  }
  if (t == (uint32_t)L::kThreads - 1u) {
    S.off[ns] = before + incl;
    S.total = before + incl;
    S.ns = ns;
  }
  __syncthreads();
  if (t < 64u) {  // home atoms, home cell by home cell (hz, hy outer, hx inner): a scan over wave 0
    const uint32_t nh = home_cells(g);  // at most kHomeCells = 64
    uint32_t n = 0;
    if (t < nh) {
      const uint32_t hs = staged_cell(g, t / (g.gx * g.gy) + 1u, (t / g.gx) % g.gy + 1u, t % g.gx + 1u);
      n = S.off[hs + 1u] - S.off[hs];
    }
    const uint32_t incl = wave_incl_scan_u32(n, lane);
    if (t < nh) S.hoff[t] = incl - n;
    if (t + 1u == nh) S.hoff[nh] = incl;
  }
  const uint32_t total = S.total;
  const bool staged = total <= (uint32_t)L::kStageMax;
  if (staged) {
    // every load of this thread's entries first, then the LDS stores
    constexpr int kPer = (L::kStageMax + L::kThreads - 1) / L::kThreads;
    uint32_t code[kPer];
    double px[kPer], py[kPer], pz[kPer];
#pragma unroll
    for (int k = 0; k < kPer; k++) {
      const uint32_t e = t + (uint32_t)k * L::kThreads;
      code[k] = kNoImage;
      px[k] = py[k] = pz[k] = 0.0;
      if (e < total) {
        const uint32_t c = stage_cell_of(S, e);
        const uint32_t j = S.base[c] + e;
        code[k] = S.image[c];
        px[k] = X[j];
        py[k] = Y[j];
        pz[k] = Z[j];
      }
    }
    if (shifted_any) {
#pragma unroll
      for (int k = 0; k < kPer; k++) {
        const uint32_t e = t + (uint32_t)k * L::kThreads;
        if (e < total) {
          double shx, shy, shz;
          image_shift(R, code[k], shx, shy, shz);
          S.x[e] = px[k] + shx;
          S.y[e] = py[k] + shy;
          S.z[e] = pz[k] + shz;
        }
      }
    } else {  // no cell of the brick is an image: every shift is +0
#pragma unroll
      for (int k = 0; k < kPer; k++) {
        const uint32_t e = t + (uint32_t)k * L::kThreads;
        if (e < total) {
          S.x[e] = px[k] + 0.0;
          S.y[e] = py[k] + 0.0;
          S.z[e] = pz[k] + 0.0;
        }
      }
    }
    if (t == 0) S.x[L::kListPad] = S.y[L::kListPad] = S.z[L::kListPad] = kFar;
  }
  __syncthreads();
#if VLJ_STAGE_PRIO
  __builtin_amdgcn_s_setprio(0);
#endif
  return staged;
}

// Home atom h of the brick: its home cell, staged index and slot.
struct Home {
  uint32_t hx, hy, hz, e, slot;
};

template <class L>
__device__ __forceinline__ Home home_atom(const Brick& g, const Stage<L>& S, uint32_t h) {
  // the last home cell whose first home atom is at most h (binary search)
  uint32_t hc = 0, hi = home_cells(g) - 1u;
  while (hc < hi) {
    const uint32_t mid = (hc + hi + 1u) >> 1;
    if (S.hoff[mid] <= h) hc = mid; else hi = mid - 1u;
  }
  Home a;
  a.hx = hc % g.gx;
  a.hy = (hc / g.gx) % g.gy;
  a.hz = hc / (g.gx * g.gy);
  const uint32_t hs = staged_cell(g, a.hz + 1u, a.hy + 1u, a.hx + 1u);
  const uint32_t idx = h - S.hoff[hc];
  a.e = S.off[hs] + idx;
  a.slot = S.base[hs] + a.e;
  return a;
}

// Position of staged atom e, with its image: from LDS when staged, else
// looked up through the cell tables in global memory.
template <class L>
__device__ __forceinline__ void staged_pos(const Replica& R, const Stage<L>& S, bool staged, uint32_t e,
                                           const double* X, const double* Y, const double* Z, double& x,
                                           double& y, double& z) {
  if (staged) {
    x = S.x[e];
    y = S.y[e];
    z = S.z[e];
  } else {
    const uint32_t c = stage_cell_of(S, e);
    const uint32_t j = S.base[c] + e;
    double shx, shy, shz;
    image_shift(R, S.image[c], shx, shy, shz);
    x = X[j] + shx;
    y = Y[j] + shy;
    z = Z[j] + shz;
  }
}

// A value that is the same in every lane, as a scalar.
__device__ __forceinline__ uint32_t uniform(uint32_t v) { return (uint32_t)__builtin_amdgcn_readfirstlane((int)v); }

// A brick whose neighbourhood holds more than kStageMax atoms (the density
// rose since the bricks were planned) is cut into parts: each of its cell
// layers on its own, each layer into runs of whole home rows, each run as
// long as still stages at most kStageMax atoms, and a home row that does
// not fit alone cut the same way into runs of home cells. A part of a
// single home cell that still does not fit reads its positions from global
// memory. Every kernel that reads the lists cuts a brick the same way, from
// the same cell counts, so the staged indices of a list refer to the
// staging of its part. A part is packed as x0 | y0 << 3 | (w - 1) << 6 |
// (h - 1) << 9 | z0 << 12 (home-cell offsets and sizes within the brick;
// one layer).
template <class L>
__device__ __forceinline__ uint32_t part_rows(const Brick& g, const Stage<L>& S, uint32_t z0, uint32_t y0,
                                              uint32_t h) {
  const uint32_t sx = g.gx + 2u, sy = g.gy + 2u;
  uint32_t n = 0;
  for (uint32_t dz = z0; dz < z0 + 3u; dz++)
    n += S.off[(dz * sy + y0 + h + 2u) * sx] - S.off[(dz * sy + y0) * sx];
  return n;
}

template <class L>
__device__ __forceinline__ uint32_t part_cells(const Brick& g, const Stage<L>& S, uint32_t z0, uint32_t y0,
                                               uint32_t x0, uint32_t w) {
  const uint32_t sx = g.gx + 2u, sy = g.gy + 2u;
  uint32_t n = 0;
  for (uint32_t dz = z0; dz < z0 + 3u; dz++)
    for (uint32_t yy = y0; yy < y0 + 3u; yy++) {
      const uint32_t row = (dz * sy + yy) * sx;
      n += S.off[row + x0 + w + 2u] - S.off[row + x0];
    }
  return n;
}

__device__ __forceinline__ uint16_t part_code(uint32_t z0, uint32_t y0, uint32_t x0, uint32_t w, uint32_t h) {
  return (uint16_t)(x0 | y0 << 3 | (w - 1u) << 6 | (h - 1u) << 9 | z0 << 12);
}

// The parts of brick g from its staged cell offsets (one thread).
template <class L>
__device__ uint32_t cut_brick(const Brick& g, const Stage<L>& S, uint16_t* parts) {
  const uint32_t lim = (uint32_t)L::kStageMax;
  uint32_t np = 0;
  for (uint32_t z0 = 0; z0 < g.gz; z0++)
    for (uint32_t y0 = 0; y0 < g.gy;) {
      uint32_t h = 1;
      while (y0 + h < g.gy && part_rows(g, S, z0, y0, h + 1u) <= lim) h++;
      if (h > 1u || g.gx == 1u || part_rows(g, S, z0, y0, 1u) <= lim) {
        parts[np++] = part_code(z0, y0, 0u, g.gx, h);
      } else {
        for (uint32_t x0 = 0; x0 < g.gx;) {
          uint32_t w = 1;
          while (x0 + w < g.gx && part_cells(g, S, z0, y0, x0, w + 1u) <= lim) w++;
          parts[np++] = part_code(z0, y0, x0, w, 1u);
          x0 += w;
        }
      }
      y0 += h;
    }
  return np;
}

__device__ __forceinline__ Brick part_brick(const Brick& g, uint32_t c, const Replica& R) {
  const uint32_t x0 = c & 7u, y0 = (c >> 3) & 7u, z0 = (c >> 12) & 1u;
  return Brick{g.rep, g.cell0 + (z0 * R.cells[1] + y0) * R.cells[0] + x0, ((c >> 6) & 7u) + 1u, ((c >> 9) & 7u) + 1u,
               1u};
}

// Workgroup-wide exclusive scan of one value per thread (THREADS threads);
// *total gets the sum. Every thread calls it.
template <int THREADS>
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* total) {
  __shared__ uint32_t wsum[THREADS / 64];
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  const uint32_t incl = wave_incl_scan_u32(v, (int)lane);
  __syncthreads();  // earlier readers of wsum are done
  if (lane == 63u) wsum[wave] = incl;
  __syncthreads();
  uint32_t before = 0, all = 0;
  for (uint32_t w = 0; w < (uint32_t)(THREADS / 64); w++) {
    const uint32_t x = wsum[w];
    before += w < wave ? x : 0u;
    all += x;
  }
  *total = all;
  return before + incl - v;
}

// Calls f(gb, code) for every part (code packed as above) of every brick
// marked cut; every thread of the workgroup calls it, and f may hold
// barriers. The parts of all cut bricks, in table order, are numbered 0,
// 1, ...; workgroup w takes the parts w, w + gridDim.x, ..., so the parts
// of one brick run side by side. Every workgroup finds the cut bricks the
// same way: the table comes in groups of kCutGroup bricks, each group with
// a count of its cut bricks, and only groups that hold some are read.
// The range of the brick table that layout L covers: the dense bricks
// first, then the sparse ones.
template <class L>
__device__ __forceinline__ uint32_t layout_first(const Buffers& B) {
  return L::kSparse ? B.n_dense : 0u;
}
template <class L>
__device__ __forceinline__ uint32_t layout_end(const Buffers& B) {
  return L::kSparse ? B.n_bricks : B.n_dense;
}

template <class L, class F>
__device__ __forceinline__ void each_cut_part(const Buffers& B, F&& f) {
  __shared__ uint32_t glist[L::kThreads];  // groups with cut bricks, of this round
  __shared__ uint32_t blist[kCutGroup];    // cut bricks of the current group
  __shared__ uint32_t first[kCutGroup];    // number of each one's first part within the group
  if (B.flags[1] == 0u) return;  // no brick is cut
  const uint32_t t = threadIdx.x, G = gridDim.x;
  const uint32_t b0 = layout_first<L>(B), b1 = layout_end<L>(B);
  const uint32_t groups = (B.n_bricks + kCutGroup - 1u) / kCutGroup;
  uint32_t base = 0;  // parts of the groups before
  for (uint32_t g0 = 0; g0 < groups; g0 += (uint32_t)L::kThreads) {
    const uint32_t grp = g0 + t;
    const uint32_t has = grp < groups && B.cut_groups[grp] != 0u ? 1u : 0u;
    uint32_t n_groups;
    const uint32_t gpos = block_excl_scan<L::kThreads>(has, &n_groups);  // past every earlier use of the lists
    if (has) glist[gpos] = grp;
    __syncthreads();
    for (uint32_t k = 0; k < n_groups; k++) {
      // a thread per brick of the group; bricks of the other layout count as uncut
      const uint32_t gb = uniform(glist[k]) * kCutGroup + t;
      const uint32_t np = t < kCutGroup && gb >= b0 && gb < b1 ? B.cut[gb] : 0u;
      uint32_t n_cut, n_parts;
      const uint32_t bpos = block_excl_scan<L::kThreads>(np != 0u ? 1u : 0u, &n_cut);
      const uint32_t upos = block_excl_scan<L::kThreads>(np, &n_parts);
      if (np != 0u) {
        blist[bpos] = gb;
        first[bpos] = upos;
      }
      __syncthreads();
      for (uint32_t u = (blockIdx.x + G - base % G) % G; u < n_parts; u += G) {
        uint32_t lo = 0, hi = n_cut - 1u;  // the last cut brick whose first part is at most u
        while (lo < hi) {
          const uint32_t mid = (lo + hi + 1u) >> 1;
          if (uniform(first[mid]) <= u) lo = mid; else hi = mid - 1u;
        }
        const uint32_t cb = uniform(blist[lo]);
        f(cb, uniform(B.parts[(size_t)cb * kHomeCells + (u - uniform(first[lo]))]));
      }
      base += n_parts;
    }
  }
}

// ---------------------------------------------------------------------------
// Verlet lists (lennard_jones.wgsl:454-499)
// ---------------------------------------------------------------------------

// Lane rank among the set bits of a wave mask (bits of lower lanes).
__device__ __forceinline__ uint32_t lane_rank(unsigned long long mask) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// A lane's value read from lane `l` (the same in every lane), as a scalar.
__device__ __forceinline__ uint32_t readlane_u32(uint32_t v, uint32_t l) {
  return (uint32_t)__builtin_amdgcn_readlane((int)v, (int)l);
}

// Atoms of a home cell per work item of the list build, and chunks of 64
// candidates a wave holds in registers at once.
// Measured on the 1M-atom liquid (ms per step, lists every ~10 steps):
// piece 16 / 4 chunks 0.302, 64 / 4 0.298, 16 / 8 0.298, 32 / 10 0.290
// (a dense cell's ~520 candidates all in registers, so each home atom's
// position and list length are read once), 64 / 12 0.292. Again with
// two-layer bricks and the drift-ahead step (atom-steps/s, one box):
// 32 / 10 5.510e9 and 5.512e9, 32 / 9 5.523e9, 16 / 10 5.410e9, 12 / 10
// 5.391e9, 32 / 12 4.79e9.
#ifndef VLJ_BUILD_PIECE
#define VLJ_BUILD_PIECE 32
#endif
#ifndef VLJ_BUILD_GROUP
#define VLJ_BUILD_GROUP 10
#endif
constexpr uint32_t kPiece = VLJ_BUILD_PIECE;
constexpr int kGroup = VLJ_BUILD_GROUP;
static_assert(kPiece >= 1 && kPiece <= 64, "a lane per home atom of a work item");
static_assert(kListStride == 64, "entry pos of a list lies (pos / 8) KB + 2 (pos % 8) bytes from its base");

// One workgroup per brick; its waves build the lists of one work item (a
// piece of a home cell's atoms) at a time. A home cell's candidates are the
// atoms of its 27 neighbour cells in list order: the nine runs (dz, dy) one
// after the other, each run the three staged cells along x, which are
// staged one after the other, so a run is one range of staged atoms. The
// 64 lanes take 64 consecutive candidates each (up to kGroup such chunks in
// registers). Then, home atom by home atom, a ballot per chunk marks the
// candidates within rc + skin (not the atom itself); each hit is stored at
// the atom's list length plus its rank in the ballot, and the length grows
// by the ballot's count. So the entries come out in stencil order dz, dy,
// dx, then slot (R4), and a sparse cell's 27 neighbours share chunks.
// On a reorder the workgroup also copies its home atoms' state back from
// the second set.
// Sparse bricks (dilute layers): a lane per home atom, walking its own
// candidates in list order, the nine runs (dz, dy), each the three staged
// cells along x, one range of staged atoms. A dilute cell holds a few atoms,
// so the 64 candidates of a chunk (below) would be mostly empty lanes and
// every home atom would cost a pass over its chunks; here every lane tests
// only its own candidates. The lists are the same, entry for entry.
template <class L>
__device__ __forceinline__ void build_lanes(const Buffers& B, const Replica& R, const Stage<L>& S, const Brick& g,
                                            bool staged, const double* X, const double* Y, const double* Z) {
  const uint32_t cap = R.capacity;
  const uint32_t H = S.hoff[home_cells(g)];
  for (uint32_t h = threadIdx.x; h < H; h += L::kThreads) {
    const Home a = home_atom(g, S, h);
    double xi, yi, zi;
    staged_pos(R, S, staged, a.e, X, Y, Z, xi, yi, zi);
    const uint32_t li = a.slot - R.atom_off;
    char* lst = (char*)(B.nbr + list_base(R, li));
    uint32_t found = 0u;
    for (uint32_t r = 0; r < 9u; r++) {
      const uint32_t sc = staged_cell(g, a.hz + r / 3u, a.hy + r % 3u, a.hx);
      const uint32_t s1 = S.off[sc + 3u];
      for (uint32_t q = S.off[sc]; q < s1; q++) {
        double xj, yj, zj;
        staged_pos(R, S, staged, q, X, Y, Z, xj, yj, zj);
        const double d2 = dist2(xj - xi, yj - yi, zj - zi);
        if (d2 < R.list_r2 && q != a.e) {
          // byte offset (pos / 8) KB + 2 (pos % 8) = 2 pos + 1008 (pos / 8)
          if (found < cap)
            *(uint16_t*)(lst + (size_t)(__umul24(found >> 3, 1008u) + 2u * found)) =
                (uint16_t)list_entry_of(q, staged);
          found++;
        }
      }
    }
    // past the capacity the list keeps its first `cap` entries and the
    // overflow is reported (lennard_jones.wgsl:494-497); the last octet is
    // filled up with pads
    const uint32_t len = found < cap ? found : cap;
    uint16_t* l16 = B.nbr + list_base(R, li);
    for (uint32_t p = len; (p & 7u) != 0u; p++) l16[list_entry(p)] = (uint16_t)entry_of(L::kListPad);
    B.nbr_count[a.slot] = len;
    if (found > cap) atomicMax(&B.ctl[g.rep].overflow, found);
  }
}

// The lists of the home atoms of brick (or part) g, staged in S.
// With `relist` the positions are those of the last build (xb) and the
// state is left as it is: the same lists again, for another brick plan.
template <class L>
__device__ __forceinline__ void build_part(const Buffers& B, const Replica& R, const Stage<L>& S, const Brick& g,
                                           bool staged, bool moved, bool relist, const double* X, const double* Y,
                                           const double* Z, uint32_t& next_item) {
  const uint32_t cap = R.capacity;
  const uint32_t total = S.total;
  if (!staged && !relist && threadIdx.x == 0) atomicAdd(&B.ctl[g.rep].stage_global, 1u);
  const uint32_t nh = home_cells(g);
  const uint32_t H = S.hoff[nh];
  for (uint32_t h = threadIdx.x; h < H; h += L::kThreads) {
    const Home a = home_atom(g, S, h);
    const uint32_t i = a.slot;
    double xi, yi, zi;
    staged_pos(R, S, staged, a.e, X, Y, Z, xi, yi, zi);
    if (moved) {  // copy the reordered state back into the main set
      B.x[i] = xi;
      B.y[i] = yi;
      B.z[i] = zi;
      B.vx[i] = B.vx2[i];
      B.vy[i] = B.vy2[i];
      B.vz[i] = B.vz2[i];
      B.ax[i] = B.ax2[i];
      B.ay[i] = B.ay2[i];
      B.az[i] = B.az2[i];
      B.orig[i] = B.orig2[i];
      B.cell_of[i] = B.cell_of2[i];
    }
    if (!relist) {
      B.xb[i] = xi;
      B.yb[i] = yi;
      B.zb[i] = zi;
    }
    // A neighbourhood too large for 16-bit staged indices keeps no list;
    // the overflow is reported whatever the capacity, with the staged count
    // as the length (lennard_jones.wgsl:494-497).
    if (total > 0xffffu) B.nbr_count[i] = 0u;
  }
  if (total > 0xffffu) {
    if (threadIdx.x == 0 && H > 0u) atomicMax(&B.ctl[g.rep].overflow, total);
    return;
  }
  if constexpr (L::kSparse) {
    build_lanes(B, R, S, g, staged, X, Y, Z);
    return;
  }
  // Work items: the atoms of each home cell cut into pieces of at most
  // kPiece, taken by the waves one at a time from a counter in LDS. Lane c
  // of every wave holds home cell c's first home atom, atom count, first
  // item and item count.
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t hv = lane < nh ? S.hoff[lane] : H;
  const uint32_t hn = lane < nh ? S.hoff[lane + 1u] - hv : 0u;
  const uint32_t pieces = (hn + kPiece - 1u) / kPiece;
  const uint32_t pincl = wave_incl_scan_u32(pieces, (int)lane);
  const uint32_t pv = pincl - pieces;  // lanes past the home cells: the item count
  const uint32_t n_items = readlane_u32(pincl, 63);
  for (;;) {
    uint32_t item = 0u;
    if (lane == 0u) item = atomicAdd(&next_item, 1u);
    item = readlane_u32(item, 0);
    if (item >= n_items) break;
    const uint32_t hc = (uint32_t)__builtin_popcountll(__builtin_amdgcn_ballot_w64(pv <= item)) - 1u;
    const uint32_t h0 = readlane_u32(hv, hc), n = readlane_u32(hn, hc);
    const uint32_t p = item - readlane_u32(pv, hc), np = readlane_u32(pieces, hc);
    const uint32_t k0 = h0 + p * n / np, k1 = h0 + (p + 1u) * n / np;  // home atoms k0 .. k1 - 1
    const uint32_t hx = hc % g.gx, hy = (hc / g.gx) % g.gy, hz = hc / (g.gx * g.gy);
    const uint32_t hs = staged_cell(g, hz + 1u, hy + 1u, hx + 1u);
    const uint32_t ebase = uniform(S.off[hs]) - h0;                 // staged index of home atom k: ebase + k
    const uint32_t lbase = uniform(S.base[hs]) + ebase - R.atom_off;  // its replica slot: lbase + k
    // The candidates in list order: the nine runs (dz, dy), each the three
    // staged cells hx .. hx + 2 of row hy + dy of layer hz + dz, one after
    // the other. Run r (lane r) holds sequence numbers rend[r - 1] .. rend[r] - 1;
    // sequence number s of run r is staged atom s + rdelta[r].
    uint32_t rs = 0u, rl = 0u;
    if (lane < 9u) {
      const uint32_t sc = staged_cell(g, hz + lane / 3u, hy + lane % 3u, hx);
      rs = S.off[sc];
      rl = S.off[sc + 3u] - rs;
    }
    const uint32_t rincl = wave_incl_scan_u32(rl, (int)lane);
    const uint32_t rdv = rs + rl - rincl;
    uint32_t rend[9], rdelta[9];
#pragma unroll
    for (int r = 0; r < 9; r++) {
      rend[r] = readlane_u32(rincl, r);
      rdelta[r] = readlane_u32(rdv, r);
    }
    const uint32_t T = rend[8];
    uint32_t found = 0u;  // list length of home atom k0 + lane
    for (uint32_t c0 = 0; c0 < T; c0 += 64u * kGroup) {
      // kGroup chunks of 64 candidates, one candidate a lane, in registers
      uint32_t q[kGroup];
      double cx[kGroup], cy[kGroup], cz[kGroup];
#pragma unroll
      for (int c = 0; c < kGroup; c++) {
        const uint32_t a = c0 + 64u * (uint32_t)c, s = a + lane;
        q[c] = L::kListPad;
        cx[c] = cy[c] = cz[c] = kFar;
        if (a >= T) continue;  // chunks past the candidates stay unread (a dilute item fills one)
        // the run s falls in: the last one whose predecessor ends at or
        // before s (a chain of selects, no branch)
        const bool valid = s < T;
        uint32_t d = rdelta[0];
#pragma unroll
        for (int r = 1; r < 9; r++) d = s >= rend[r - 1] ? rdelta[r] : d;
        if (valid) q[c] = s + d;
        if (staged) {
          const uint32_t qs = valid ? q[c] : L::kListPad;
          cx[c] = S.x[qs];
          cy[c] = S.y[qs];
          cz[c] = S.z[qs];
        } else if (valid) {
          staged_pos(R, S, false, q[c], X, Y, Z, cx[c], cy[c], cz[c]);
        }
      }
      const uint32_t nc = (T - c0 + 63u) / 64u;  // chunks of this group that hold candidates
      uint32_t qe[kGroup];
#pragma unroll
      for (int c = 0; c < kGroup; c++) qe[c] = list_entry_of(q[c], staged);
      for (uint32_t k = k0; k < k1; k++) {
        const uint32_t e = ebase + k;
        double xi, yi, zi;
        staged_pos(R, S, staged, e, X, Y, Z, xi, yi, zi);
        uint32_t f = readlane_u32(found, k - k0);
        char* lst = (char*)(B.nbr + list_base(R, lbase + k));
#pragma unroll
        for (int c = 0; c < kGroup; c++) {
          if ((uint32_t)c < nc) {
            // a ballot marks the candidates within rc + skin other than the
            // atom itself (the one candidate whose staged index is e); each
            // hit goes at the list length plus its rank
            const double d2 = dist2(cx[c] - xi, cy[c] - yi, cz[c] - zi);
            const bool near = d2 < R.list_r2 && q[c] != e;
            // the same predicate as a lane mask, straight from the compares
            // (ordered less-than, unsigned not-equal)
            const unsigned long long mask =
                __builtin_amdgcn_fcmp(d2, R.list_r2, 4) & __builtin_amdgcn_uicmp(q[c], e, 33);
            const uint32_t pos = f + lane_rank(mask);
            const uint32_t hits = (uint32_t)__builtin_popcountll(mask);
            // past the capacity only while the list fills its last entries
            // (a scalar test, so the usual chunk tests no lane against it)
            bool store = near;
            if (__builtin_expect(f + hits > cap, 0)) store = near && pos < cap;
            // byte offset (pos / 8) KB + 2 (pos % 8) = 2 pos + 1008 (pos / 8)
            if (store) *(uint16_t*)(lst + (size_t)(__umul24(pos >> 3, 1008u) + 2u * pos)) = (uint16_t)qe[c];
            f += hits;
          }
        }
        found = lane == k - k0 ? f : found;
      }
    }
    // Past the capacity the list keeps its first `cap` entries and the
    // overflow is reported (lennard_jones.wgsl:494-497). The last octet is
    // filled up with pads.
    if (lane < k1 - k0) {
      const uint32_t li = lbase + k0 + lane;
      const uint32_t len = found < cap ? found : cap;
      uint16_t* lst = B.nbr + list_base(R, li);
      for (uint32_t p = len; (p & 7u) != 0u; p++) lst[list_entry(p)] = (uint16_t)entry_of(L::kListPad);
      B.nbr_count[R.atom_off + li] = len;
      if (found > cap) atomicMax(&B.ctl[g.rep].overflow, found);
    }
  }
}

// One workgroup per brick (PARTS false), or a grid that walks the bricks
// marked to be cut and builds them part by part (PARTS true; launched after
// the first, on the same step). A brick counts as cut when its
// neighbourhood does not fit in LDS and it has more than one home cell.
// `relist`: every brick of every replica again, from the positions of the
// last build, after the brick plan changed; nothing else changes.
template <class L, bool PARTS>
__global__ void __launch_bounds__(L::kThreads) k_build(Buffers B, int relist) {
  __shared__ Stage<L> S;
  __shared__ uint32_t next_item;
  if (!relist && *B.rebuilding == 0u) return;  // no replica rebuilds this step
  if constexpr (!PARTS) {
    const uint32_t gb = layout_first<L>(B) + dealt_index(blockIdx.x);
    if (gb >= layout_end<L>(B)) return;
    const Brick g = B.bricks[gb];
    const Control& ctl = B.ctl[g.rep];
    if (!relist && !ctl.rebuild) return;
    const bool moved = !relist && ctl.reorder != 0u;
    const Replica R = B.rep[g.rep];
    const double* X = relist ? B.xb : (moved ? B.x2 : B.x);
    const double* Y = relist ? B.yb : (moved ? B.y2 : B.y);
    const double* Z = relist ? B.zb : (moved ? B.z2 : B.z);
    if (threadIdx.x == 0) next_item = 0u;
    const bool staged = stage_brick(B, R, g, X, Y, Z, S);  // ends with a barrier
    // A neighbourhood that does not fit is cut into parts when the handle
    // runs the PARTS launches (parts_on), else the brick is built whole from
    // global memory; either way the overflow is counted.
    const bool over = !staged && home_cells(g) > 1u;
    const bool cut = over && B.parts_on;
    if (threadIdx.x == 0) {  // mark the brick and record its parts for the PARTS launches
      const uint32_t was = B.cut[gb];
      const uint32_t np = cut ? cut_brick(g, S, B.parts + (size_t)gb * kHomeCells) : 0u;
      B.cut[gb] = np;
      if ((was != 0u) != (np != 0u)) {
        atomicAdd(&B.cut_groups[gb / kCutGroup], np != 0u ? 1u : 0xffffffffu);
        atomicAdd(&B.flags[1], np != 0u ? 1u : 0xffffffffu);
      }
      if (over && !relist) atomicAdd(&B.ctl[g.rep].stage_overflow, 1u);
    }
    if (cut) return;
    build_part(B, R, S, g, staged, moved, relist != 0, X, Y, Z, next_item);
  } else {
    each_cut_part<L>(B, [&](uint32_t gb, uint32_t code) {
      const Brick brick = B.bricks[gb];
      const Control& ctl = B.ctl[brick.rep];
      if (!relist && !ctl.rebuild) return;
      const bool moved = !relist && ctl.reorder != 0u;
      const Replica R = B.rep[brick.rep];
      const double* X = relist ? B.xb : (moved ? B.x2 : B.x);
      const double* Y = relist ? B.yb : (moved ? B.y2 : B.y);
      const double* Z = relist ? B.zb : (moved ? B.z2 : B.z);
      const Brick g = part_brick(brick, code, R);
      __syncthreads();  // the previous part is done with S and next_item
      if (threadIdx.x == 0) next_item = 0u;
      const bool staged = stage_brick(B, R, g, X, Y, Z, S);
      build_part(B, R, S, g, staged, moved, relist != 0, X, Y, Z, next_item);
    });
  }
}

// ---------------------------------------------------------------------------
// Pair forces (lennard_jones.wgsl:501-574)
// ---------------------------------------------------------------------------

struct PairAcc {
  double ax, ay, az;
  double pe, w;
  uint32_t near, within;
};

// 1 / x correctly rounded, for x whose reciprocal is a normal number
// (here r^2, from about 1e-4 to the sentinel's 3e36): the operations of the
// IEEE division sequence (reciprocal estimate, two Newton-Raphson steps, one
// residual correction) without its range scaling and special-value fix-up,
// which change nothing for such x.
__device__ __forceinline__ double rcp_rn(double x) {
  double y = __builtin_amdgcn_rcp(x);
  double e = __builtin_fma(-x, y, 1.0);
  y = __builtin_fma(y, e, y);
  e = __builtin_fma(-x, y, 1.0);
  y = __builtin_fma(y, e, y);
  e = __builtin_fma(-x, y, 1.0);
  return __builtin_fma(e, y, y);
}

// 1 / q correctly rounded in FP32, for a positive normal q whose reciprocal
// is normal (R5f's 1.0f / q): the hardware estimate (within one ulp), one
// Newton-Raphson correction y + y (1 - q y) with both steps fused (1 - q y
// is exact), which rounds to the nearest float for every faithful estimate
// except one case: a significand of all ones, whose reciprocal 2^k (1 +
// 2^-24 + 2^-48 + ...) lies just above a midpoint, so that the estimate
// 2^k gives the tie 2^k (1 + 2^-24) and rounds back down to 2^k. There the
// answer is 2^k (1 + 2^-23) from either estimate, i.e. the last
// significand bit set.
__device__ __forceinline__ float rcp_rn_f32(float q) {
  const float y0 = __builtin_amdgcn_rcpf(q);
  const float e = __builtin_fmaf(-q, y0, 1.0f);
  const float y1 = __builtin_fmaf(y0, e, y0);
  const uint32_t m = __float_as_uint(q) & 0x7fffffu;
  return __uint_as_float(__float_as_uint(y1) | ((m + 1u) >> 23));
}

// R5 (FP64) and R5f (FP32 pair arithmetic on FP64 displacements), without
// branches: the terms of a pair outside the cutoff are computed and then
// replaced by zero, which leaves every sum bit for bit as if the pair had
// been skipped (no sum here is ever -0, so adding +0 changes nothing).
// Forces: G += d fscal with fscal = g / r^2, g = r.F / 24, and a = -24 G
// once per atom. ENERGIES: the measurement's pe, w, near and within (R9)
// instead of the force.
template <int PREC, bool SF, bool ENERGIES>
__device__ __forceinline__ void pair_accumulate(const Replica& R, double dx, double dy, double dz,
                                                double r2, PairAcc& a) {
  const bool hit = r2 < R.rc2;
  if (PREC == 0 && !SF && !ENERGIES) {
    // The plain force takes the cutoff through r^2: a pair outside it gets
    // 0x567 (2^360) as r^2's sign and exponent, so 1/r^2 is a normal number
    // below 2^-360, s^6 rounds to exactly 0 (under 2^-1080), g is -0 and so is
    // fscal, and the three fma() return their sums unchanged. One select of a
    // 32-bit word instead of two after the arithmetic.
    const int hi = hit ? __double2hiint(r2) : 0x56700000;
    const double inv2 = rcp_rn(__hiloint2double(hi, __double2loint(r2)));
    const double s6 = inv2 * inv2 * inv2;
    const double fscal = s6 * __builtin_fma(s6, 2.0, -1.0) * inv2;
    a.ax = __builtin_fma(dx, fscal, a.ax);
    a.ay = __builtin_fma(dy, fscal, a.ay);
    a.az = __builtin_fma(dz, fscal, a.az);
  } else if (PREC == 0) {
    const double inv2 = rcp_rn(r2);
    const double s6 = inv2 * inv2 * inv2;
    if (!ENERGIES) {
      double g = s6 * __builtin_fma(s6, 2.0, -1.0);
      if (SF) g = __builtin_fma(-__builtin_sqrt(r2), R.force_shift24, g);
      const double fscal = hit ? g * inv2 : 0.0;
      a.ax = __builtin_fma(dx, fscal, a.ax);
      a.ay = __builtin_fma(dy, fscal, a.ay);
      a.az = __builtin_fma(dz, fscal, a.az);
    } else {
      double rf = s6 * __builtin_fma(48.0, s6, -24.0);
      const double t = 4.0 * s6;
      double u = __builtin_fma(t, s6, -t) - R.shift;
      if (SF) {
        const double r = __builtin_sqrt(r2);
        rf = __builtin_fma(-r, R.force_shift, rf);
        u = __builtin_fma(r - R.rc, R.force_shift, u);
      }
      a.pe += hit ? 0.5 * u : 0.0;
      a.w += hit ? 0.5 * rf : 0.0;
    }
  } else {
    static_assert(PREC == 0 || ENERGIES, "FP32 forces sum through pair_force_f32");
    const float q = (float)r2;
    const float inv2 = rcp_rn_f32(q);
    const float s6 = inv2 * inv2 * inv2;
    {
      float rf = s6 * __builtin_fmaf(48.0f, s6, -24.0f);
      const float t = 4.0f * s6;
      float u = __builtin_fmaf(t, s6, -t) - (float)R.shift;
      if (SF) {
        const float r = __builtin_sqrtf(q);
        const float fs = (float)R.force_shift;
        rf = __builtin_fmaf(-r, fs, rf);
        u = __builtin_fmaf(r - (float)R.rc, fs, u);
      }
      a.pe += hit ? 0.5 * (double)u : 0.0;
      a.w += hit ? 0.5 * (double)rf : 0.0;
    }
  }
  if (ENERGIES) {
    a.near += hit && r2 < kNearR2 ? 1u : 0u;
    a.within += hit ? 1u : 0u;
  }
}

// R5f's force: the pair factor in FP32 (on the FP64 cutoff test) and
// f += (float)d fscal in FP32 over the eight entries of one list octet; the
// caller adds each octet's f to the FP64 sum (pads add exactly 0).
template <bool SF>
__device__ __forceinline__ void pair_force_f32(const Replica& R, double dx, double dy, double dz, double r2,
                                               float& fx, float& fy, float& fz) {
  const bool hit = r2 < R.rc2;
  const float q = (float)r2;
  const float inv2 = rcp_rn_f32(q);
  const float s6 = inv2 * inv2 * inv2;
  float g = s6 * __builtin_fmaf(s6, 2.0f, -1.0f);
  if (SF) g = __builtin_fmaf(-__builtin_sqrtf(q), (float)R.force_shift24, g);
  const float fscal = hit ? g * inv2 : 0.0f;
  fx = __builtin_fmaf((float)dx, fscal, fx);
  fy = __builtin_fmaf((float)dy, fscal, fy);
  fz = __builtin_fmaf((float)dz, fscal, fz);
}

// List octets (8 entries each) in flight per batch of the force loop. FCC-1M
// (two-layer bricks, drift-ahead force launch; same box): 2 octets 129.5 us
// per FP64 launch, 1 130.6, 3 132.8, 4 135.1, 8 170 (spills); the FP32 form
// is the same at 2, 3 and 4.
#ifndef VLJ_OCTET_BATCH
#define VLJ_OCTET_BATCH 2
#endif
constexpr int kOctetBatch = VLJ_OCTET_BATCH;

// Octet o of a list, o clamped to `last`, the list's own last octet (its
// first when it is empty), so the load needs no branch and a clamped load
// hits the cache line just read; octets past an atom's count are loaded but
// never summed.
__device__ __forceinline__ u32x4 load_octet(const uint16_t* base, uint32_t o, uint32_t last) {
  return __builtin_nontemporal_load((const u32x4*)(base + (uint64_t)min(o, last) * (8u * kListStride)));
}

// f(k, entry) for entries k = 0 .. cnt - 1 of a list in order, read an
// octet (16 bytes) at a time with the next octet's load in flight, so that
// a loop over the entries does not wait on one 2-byte load per entry.
template <class F>
__device__ __forceinline__ void each_entry(const uint16_t* lst, uint32_t cnt, F&& f) {
  const uint32_t no = (cnt + 7u) / 8u;
  const uint32_t last = no > 0u ? no - 1u : 0u;
  u32x4 cur = load_octet(lst, 0u, last);
  for (uint32_t o = 0; o < no; o++) {
    const u32x4 nxt = load_octet(lst, o + 1u, last);
    const uint32_t w[4] = {cur.x, cur.y, cur.z, cur.w};
#pragma unroll
    for (uint32_t c = 0; c < 8u; c++) {
      const uint32_t k = 8u * o + c;
      if (k < cnt) f(k, (w[c >> 1] >> ((c & 1u) * 16u)) & 0xffffu);
    }
    cur = nxt;
  }
}

// Atom i's list in order, kOctetBatch octets at a time, double-buffered:
// the next batch's loads are in flight while this one is summed. Positions
// come from the brick's LDS; whole octets are taken (pads meet the far
// sentinel), so the eight pairs of an octet carry no branch and their LDS
// reads can all be in flight together.
template <int PREC, bool SF, bool ENERGIES, class L>
__device__ __forceinline__ PairAcc interactions_staged(const Buffers& B, const Replica& R, const Stage<L>& S,
                                                       uint32_t i, double xi, double yi, double zi) {
  const uint32_t cnt = B.nbr_count[i];
  const uint32_t no = (cnt + 7u) / 8u;
  const uint16_t* lst = B.nbr + list_base(R, i - R.atom_off);
  const uint32_t last = no > 0u ? no - 1u : 0u;
  PairAcc a = {0.0, 0.0, 0.0, 0.0, 0.0, 0u, 0u};
  u32x4 cur[kOctetBatch], nxt[kOctetBatch];
#pragma unroll
  for (int b = 0; b < kOctetBatch; b++) cur[b] = load_octet(lst, (uint32_t)b, last);
  for (uint32_t o0 = 0; o0 < no; o0 += kOctetBatch) {
#pragma unroll
    for (int b = 0; b < kOctetBatch; b++) nxt[b] = load_octet(lst, o0 + kOctetBatch + (uint32_t)b, last);
#pragma unroll
    for (int b = 0; b < kOctetBatch; b++) {
      if (o0 + (uint32_t)b < no) {
        const uint32_t w[4] = {cur[b].x, cur[b].y, cur[b].z, cur[b].w};
        double dx[8], dy[8], dz[8];
        // the FP64 form issues an octet's LDS gathers at raised priority
        // (150 against 154 us per force launch at N = 1M; the FP32 form,
        // with less arithmetic per gather, lost 1% and does not)
        if constexpr (PREC == 0) __builtin_amdgcn_s_setprio(VLJ_READ_PRIO);
#pragma unroll
        for (int c = 0; c < 8; c++) {
          const uint32_t off = (w[c >> 1] >> ((c & 1) * 16)) & 0xffffu;  // 8 e
          dx[c] = *(const double*)((const char*)S.x + off) - xi;
          dy[c] = *(const double*)((const char*)S.y + off) - yi;
          dz[c] = *(const double*)((const char*)S.z + off) - zi;
        }
        if constexpr (PREC == 0) __builtin_amdgcn_s_setprio(0);
        if constexpr (PREC == 1 && !ENERGIES) {
          float fx = 0.0f, fy = 0.0f, fz = 0.0f;
#pragma unroll
          for (int c = 0; c < 8; c++)
            pair_force_f32<SF>(R, dx[c], dy[c], dz[c], dist2(dx[c], dy[c], dz[c]), fx, fy, fz);
          a.ax += (double)fx;
          a.ay += (double)fy;
          a.az += (double)fz;
        } else {
#pragma unroll
          for (int c = 0; c < 8; c++)
            pair_accumulate<PREC, SF, ENERGIES>(R, dx[c], dy[c], dz[c], dist2(dx[c], dy[c], dz[c]), a);
        }
      }
    }
#pragma unroll
    for (int b = 0; b < kOctetBatch; b++) cur[b] = nxt[b];
  }
  return a;
}

// The same sum for a brick whose neighbourhood did not fit in LDS: entry by
// entry up to the count, positions through the cell tables.
template <int PREC, bool SF, bool ENERGIES, class L>
__device__ __forceinline__ PairAcc interactions_global(const Buffers& B, const Replica& R, const Stage<L>& S,
                                                    uint32_t i, double xi, double yi, double zi) {
  const uint32_t cnt = B.nbr_count[i];
  const uint16_t* lst = B.nbr + list_base(R, i - R.atom_off);
  PairAcc a = {0.0, 0.0, 0.0, 0.0, 0.0, 0u, 0u};
  float fx = 0.0f, fy = 0.0f, fz = 0.0f;  // FP32 forces: the current octet's sum
  each_entry(lst, cnt, [&](uint32_t k, uint32_t entry) {
    double xj, yj, zj;
    staged_pos(R, S, false, list_index(entry, false), B.x, B.y, B.z, xj, yj, zj);
    const double dx = xj - xi, dy = yj - yi, dz = zj - zi;
    if constexpr (PREC == 1 && !ENERGIES) {
      pair_force_f32<SF>(R, dx, dy, dz, dist2(dx, dy, dz), fx, fy, fz);
      if ((k & 7u) == 7u || k + 1u == cnt) {
        a.ax += (double)fx;
        a.ay += (double)fy;
        a.az += (double)fz;
        fx = fy = fz = 0.0f;
      }
    } else {
      pair_accumulate<PREC, SF, ENERGIES>(R, dx, dy, dz, dist2(dx, dy, dz), a);
    }
  });
  return a;
}

template <int PREC, bool SF, bool ENERGIES, class L>
__device__ __forceinline__ PairAcc interactions(const Buffers& B, const Replica& R, const Stage<L>& S, bool staged,
                                                uint32_t i, double xi, double yi, double zi) {
  if (staged) return interactions_staged<PREC, SF, ENERGIES>(B, R, S, i, xi, yi, zi);
  return interactions_global<PREC, SF, ENERGIES>(B, R, S, i, xi, yi, zi);
}

// New accelerations, one workgroup per brick (gb: its table entry), and
// with kick 1 the second half kick v += a dt/2 (v holds the half step).
// Kick 2, the drift-ahead form (NVE): v holds v(t) and a a(t); the half
// step v(t) + a(t) dt/2 is formed again exactly as the drift forms it, the
// kick gives v(t + dt), and the next step's drift follows at once: its half
// step and x(t + 2 dt) = x(t + dt) + v' dt go to the second position set
// (x2), which no kernel of this step reads, and the largest moved^2 since
// the list build into one of the replica's kAheadSlots maxima of the next
// step (amax_next), so that step needs no drift launch; the set this step's
// bin read (amax) is zeroed for the step after.
template <int PREC, bool SF, class L>
__device__ __forceinline__ void force_part(const Buffers& B, const Replica& R, const Stage<L>& S, const Brick& g,
                                           uint32_t gb, bool staged, int kick) {
  const uint32_t H = S.hoff[home_cells(g)];
  unsigned long long moved = 0ull;
  for (uint32_t h = threadIdx.x; h < H; h += L::kThreads) {
    const Home at = home_atom(g, S, h);
    const uint32_t i = at.slot;
    double xi, yi, zi;
    staged_pos(R, S, staged, at.e, B.x, B.y, B.z, xi, yi, zi);
    double* const A[3] = {B.ax, B.ay, B.az};
    double* const V[3] = {B.vx, B.vy, B.vz};
    const double* const XB[3] = {B.xb, B.yb, B.zb};
    // The atom's own v and a, loaded before its pairs so that the loads are
    // in flight while they are summed (the 512-replica batch +3.3%, FCC-1M
    // +0.9%) -- in the plain (truncated and shifted) forms; the
    // shifted-force pair loops leave no room for them within 128 VGPRs.
    constexpr bool kEarly = !SF;
    double vo[3], ao[3] = {0.0, 0.0, 0.0};
    auto own_state = [&] {
#pragma unroll
      for (int k = 0; k < 3; k++) vo[k] = V[k][i];
      if (kick == 2) {
#pragma unroll
        for (int k = 0; k < 3; k++) ao[k] = A[k][i];
      }
    };
    if constexpr (kEarly) own_state();
    const PairAcc G = interactions<PREC, SF, false>(B, R, S, staged, i, xi, yi, zi);
    if constexpr (!kEarly) own_state();
    const double a[3] = {-24.0 * G.ax, -24.0 * G.ay, -24.0 * G.az};
    if (kick == 2) {
      const double* const X[3] = {B.x, B.y, B.z};
      double* const X2[3] = {B.x2, B.y2, B.z2};
      double d[3];
#pragma unroll
      for (int k = 0; k < 3; k++) {
        const double half = __builtin_fma(ao[k], R.half_dt, vo[k]);  // the drift's half step
        const double v = __builtin_fma(a[k], R.half_dt, half);
        A[k][i] = a[k];
        V[k][i] = v;
        const double x = __builtin_fma(__builtin_fma(a[k], R.half_dt, v), R.dt, X[k][i]);
        X2[k][i] = x;
        d[k] = x - XB[k][i];
      }
      const unsigned long long m = (unsigned long long)__double_as_longlong(dist2(d[0], d[1], d[2]));
      moved = m > moved ? m : moved;
    } else {
#pragma unroll
      for (int k = 0; k < 3; k++) {
        A[k][i] = a[k];
        if (kick) V[k][i] = __builtin_fma(a[k], R.half_dt, vo[k]);
      }
    }
  }
  if (kick == 2) {
    moved = wave_max_u64(moved);
    if ((threadIdx.x & 63u) == 0 && moved != 0ull)
      atomicMax(B.amax_next + (size_t)g.rep * kAheadSlots + gb % kAheadSlots, moved);
    if (threadIdx.x == 0)
      for (uint32_t k = gb; k < B.n_replicas * kAheadSlots; k += B.n_bricks) B.amax[k] = 0ull;
  }
}

// One workgroup per brick, leaving bricks that are cut (see k_build) to the
// PARTS launch that follows.
template <class L, int PREC, bool SF, bool PARTS>
__global__ void __launch_bounds__(L::kThreads, !PARTS && !L::kSparse ? 4 : 1) k_force(Buffers B, int kick) {
  __shared__ Stage<L> S;
  if constexpr (!PARTS) {
    const uint32_t gb = layout_first<L>(B) + dealt_index(blockIdx.x);
    if (gb >= layout_end<L>(B)) return;
    const Brick g = B.bricks[gb];
    const Replica R = B.rep[g.rep];
    const bool staged = stage_brick(B, R, g, B.x, B.y, B.z, S);
    if (!staged && home_cells(g) > 1u && B.parts_on) return;  // cut into parts
    force_part<PREC, SF>(B, R, S, g, gb, staged, kick);
  } else {
    each_cut_part<L>(B, [&](uint32_t gb, uint32_t code) {
      const Brick brick = B.bricks[gb];
      const Replica R = B.rep[brick.rep];
      const Brick g = part_brick(brick, code, R);
      __syncthreads();  // the previous part is done with S
      const bool staged = stage_brick(B, R, g, B.x, B.y, B.z, S);
      force_part<PREC, SF>(B, R, S, g, gb, staged, kick);
    });
  }
}

// ---------------------------------------------------------------------------
// Measurement (lennard_jones.wgsl:576-626): per atom KE, PE, W, momentum,
// condensed, pairs within rc, each to 64-bit fixed point; a workgroup sums
// its atoms by wave shuffles and LDS and adds once per slot with a 64-bit
// atomic (fixed.wgsl:71-105).
// ---------------------------------------------------------------------------

// Adds slot k of every thread (k < n <= kSlots) into dst[k]: wave sums,
// then one 64-bit atomic per slot for the workgroup.
template <int n = kSlots, int threads = kBlock>
__device__ __forceinline__ void block_add_slots(long long v[kSlots], long long* dst) {
  __shared__ long long part[threads / 64][kSlots];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < n; k++) {
    long long s = wave_sum_i64(v[k]);
    if (lane == 0) part[wave][k] = s;
  }
  __syncthreads();
  if (threadIdx.x < (unsigned)n) {
    long long s = 0;
#pragma unroll
    for (int w = 0; w < threads / 64; w++) s += part[w][threadIdx.x];
    if (s != 0) atomicAdd((unsigned long long*)(dst + threadIdx.x), (unsigned long long)s);
  }
}

// Each pass over up to L::kThreads home atoms ends in the workgroup's sum
// (a pass's sums are not kept across the pair loop).
template <int PREC, bool SF, class L>
__device__ __forceinline__ void measure_part(const Buffers& B, const Replica& R, const Stage<L>& S, const Brick& g,
                                             bool staged, long long* dst) {
  const uint32_t H = S.hoff[home_cells(g)];
  for (uint32_t h0 = 0; h0 < H; h0 += L::kThreads) {
    long long v[kSlots] = {0, 0, 0, 0, 0, 0, 0, 0};
    const uint32_t h = h0 + threadIdx.x;
    if (h < H) {
      const Home at = home_atom(g, S, h);
      const uint32_t i = at.slot;
      double xi, yi, zi;
      staged_pos(R, S, staged, at.e, B.x, B.y, B.z, xi, yi, zi);
      const PairAcc a = interactions<PREC, SF, true>(B, R, S, staged, i, xi, yi, zi);
      const double vx = B.vx[i], vy = B.vy[i], vz = B.vz[i];
      const double ke = 0.5 * dist2(vx, vy, vz);
      v[0] = to_fixed(ke, R.energy_fixed);
      v[1] = to_fixed(a.pe, R.energy_fixed);
      v[2] = to_fixed(a.w, R.energy_fixed);
      v[3] = to_fixed(vx, R.momentum_fixed);
      v[4] = to_fixed(vy, R.momentum_fixed);
      v[5] = to_fixed(vz, R.momentum_fixed);
      v[6] = a.near >= R.near_count ? 1 : 0;
      v[7] = (long long)a.within;
    }
    block_add_slots<kSlots, L::kThreads>(v, dst);
    __syncthreads();  // the sum's LDS is free for the next pass
  }
}

template <class L, int PREC, bool SF, bool PARTS>
__global__ void __launch_bounds__(L::kThreads) k_measure(Buffers B) {
  __shared__ Stage<L> S;
  if constexpr (!PARTS) {
    const uint32_t gb = layout_first<L>(B) + dealt_index(blockIdx.x);
    if (gb >= layout_end<L>(B)) return;
    const Brick g = B.bricks[gb];
    const Replica R = B.rep[g.rep];
    const bool staged = stage_brick(B, R, g, B.x, B.y, B.z, S);
    if (!staged && home_cells(g) > 1u && B.parts_on) return;  // cut into parts
    measure_part<PREC, SF>(B, R, S, g, staged, B.slots + (size_t)g.rep * kSlots);
  } else {
    each_cut_part<L>(B, [&](uint32_t gb, uint32_t code) {
      const Brick brick = B.bricks[gb];
      const Replica R = B.rep[brick.rep];
      const Brick g = part_brick(brick, code, R);
      __syncthreads();  // the previous part is done with S
      const bool staged = stage_brick(B, R, g, B.x, B.y, B.z, S);
      measure_part<PREC, SF>(B, R, S, g, staged, B.slots + (size_t)g.rep * kSlots);
    });
  }
}

// ---------------------------------------------------------------------------
// Profiles along z (lennard_jones.wgsl:628-730). Slab b holds kProfileSums
// sums from b * kProfileSums: particle count, v_a^2 (a = x, y, z), the
// Irving-Kirkwood W_aa, potential energy; after the slabs, the whole-box
// W_aa. Counts, kinetic terms and potential go to the slab of the atom's
// wrapped z. A pair's half of the virial, w_a = (0.5 d_a) d_a q with
// q = rf (1 / r2), is spread over the slabs that the segment from z_i to
// z_i + d_z crosses, in proportion to the length inside each, the last slab
// taking what is left; the slab of z is floor(z (1 / width)). An atom's
// shares to its own slab and the slabs beside it are summed in FP64 in list
// order (SlabWindow) and go to fixed point once per atom, slab and axis; a
// share to a slab farther away (slabs narrower than the cutoff) and every
// per-atom value go to fixed point on their own, as rint(x 2^32), and the
// integers are added, so the sums do not depend on the order of the atoms.
// Pair terms are FP64 in both precision forms. Up to kProfileLdsBins (256)
// slabs a workgroup adds into LDS and then once per touched sum into the
// replica's sums; with more slabs it adds straight into them.
// ---------------------------------------------------------------------------

constexpr uint32_t kProfileLdsBins = 256;
constexpr double kProfileOne = 4294967296.0;  // fixed point: 2^32 per unit

struct ProfileSink {
  long long* lds;
  long long* global;
  bool use_lds;
  __device__ __forceinline__ void add(uint32_t k, long long v) const {
    if (v == 0) return;
    if (use_lds)
      atomicAdd((unsigned long long*)(lds + k), (unsigned long long)v);
    else
      atomicAdd((unsigned long long*)(global + k), (unsigned long long)v);
  }
  __device__ __forceinline__ void add(uint32_t k, double x) const { add(k, to_fixed(x, kProfileOne)); }
};

// The virial shares of one atom's pairs to the slab below its own slab,
// its own and the one above, summed in FP64 in list order and sent to the
// profile in fixed point once per atom. With slabs at least a cutoff wide
// every share lands in these three; a share elsewhere (narrow slabs) goes
// to the profile at once.
struct SlabWindow {
  double w[3][3];  // [slab below, own, above][axis]
  __device__ __forceinline__ void clear() {
#pragma unroll
    for (int k = 0; k < 3; k++)
#pragma unroll
      for (int a = 0; a < 3; a++) w[k][a] = 0.0;
  }
  // w_a share to slab b (unwrapped), k = b - hb from the atom's (-1, 0, 1
  // fit). Each sum takes fma(f, 1, sum) = sum + f or fma(f, 0, sum) = sum:
  // one instruction per sum and no select (a non-finite share, from atoms
  // on top of each other, spoils all three of the atom's sums instead of
  // one).
  __device__ __forceinline__ void add(int k, int b, int bins, const double wv[3], double share,
                                      const ProfileSink& sink) {
    if (k >= -1 && k <= 1) {
#pragma unroll
      for (int kk = 0; kk < 3; kk++) {
        const double sel = kk == k + 1 ? 1.0 : 0.0;
#pragma unroll
        for (int a = 0; a < 3; a++) w[kk][a] = __builtin_fma(wv[a] * share, sel, w[kk][a]);
      }
    } else {
      const uint32_t slab = (uint32_t)(((b % bins) + bins) % bins);
#pragma unroll
      for (int a = 0; a < 3; a++) sink.add(slab * kProfileSums + 4u + (uint32_t)a, wv[a] * share);
    }
  }
};

template <bool SF, class L>
__device__ __forceinline__ void profile_part(const Buffers& B, const Replica& R, const Stage<L>& S, const Brick& g,
                                             bool staged, const ProfileSink& sink, long long whole[3]) {
  const uint32_t bins = R.prof_bins;
  const double width = R.bin_width;
  const double iw = 1.0 / width;  // slab of z: floor(z iw)
  const int ibins = (int)bins;
  const uint32_t H = S.hoff[home_cells(g)];
  for (uint32_t h = threadIdx.x; h < H; h += L::kThreads) {
    const Home at = home_atom(g, S, h);
    const uint32_t i = at.slot;
    double xi, yi, zi;
    staged_pos(R, S, staged, at.e, B.x, B.y, B.z, xi, yi, zi);
    const double zw = wrap_one(zi, R.box[2]);
    const int hb = (int)floor(zw * iw);  // the slab the shares are counted from
    const uint32_t home = (uint32_t)(hb < 0 ? 0 : (hb > ibins - 1 ? ibins - 1 : hb));
    sink.add(home * kProfileSums, 1.0);
    const double v[3] = {B.vx[i], B.vy[i], B.vz[i]};
#pragma unroll
    for (int a = 0; a < 3; a++) sink.add(home * kProfileSums + 1u + (uint32_t)a, v[a] * v[a]);
    double pe = 0.0, total[3] = {0.0, 0.0, 0.0};
    SlabWindow win;
    win.clear();
    const uint32_t cnt = B.nbr_count[i];
    const uint16_t* lst = B.nbr + list_base(R, i - R.atom_off);
    each_entry(lst, cnt, [&](uint32_t k, uint32_t entry) {
      double xj, yj, zj;
      staged_pos(R, S, staged, list_index(entry, staged), B.x, B.y, B.z, xj, yj, zj);
      const double d[3] = {xj - xi, yj - yi, zj - zi};
      const double r2 = dist2(d[0], d[1], d[2]);
      if (!(r2 < R.rc2)) return;
      const double inv2 = rcp_rn(r2);  // = 1.0 / r2, correctly rounded
      const double s6 = inv2 * inv2 * inv2;
      double rf = s6 * __builtin_fma(48.0, s6, -24.0);
      const double t = 4.0 * s6;
      double u = __builtin_fma(t, s6, -t) - R.shift;
      if (SF) {
        const double r = __builtin_sqrt(r2);
        rf = __builtin_fma(-r, R.force_shift, rf);
        u = __builtin_fma(r - R.rc, R.force_shift, u);
      }
      pe += 0.5 * u;
      const double q = rf * inv2;
      double w[3];
#pragma unroll
      for (int a = 0; a < 3; a++) {
        w[a] = (0.5 * d[a]) * d[a] * q;
        total[a] += w[a];
      }
      const double lo = zw + (d[2] < 0.0 ? d[2] : 0.0);
      const double hi = zw + (d[2] > 0.0 ? d[2] : 0.0);
      const double length = hi - lo;
      if (length < 1e-6 * width) {
        win.add((int)home - hb, (int)home, ibins, w, 1.0, sink);
        return;
      }
      int b = (int)floor(lo * iw);
      double left = 1.0;
      // A segment shorter than the box crosses at most bins + 1 slabs.
      for (uint32_t walk = 0; walk < bins + 2u; walk++) {
        const bool last = (double)(b + 1) * width >= hi;
        const double start = lo > (double)b * width ? lo : (double)b * width;
        const double end = hi < (double)(b + 1) * width ? hi : (double)(b + 1) * width;
        double share = last ? left : (end - start) / length;
        share = share > 0.0 ? share : 0.0;
        left -= share;
        win.add(b - hb, b, ibins, w, share, sink);
        if (last) break;
        b++;
      }
    });
    // the window's sums, to the slabs below, at and above the atom's
#pragma unroll
    for (int kk = 0; kk < 3; kk++) {
      const uint32_t slab = (uint32_t)((((hb - 1 + kk) % ibins) + ibins) % ibins);
#pragma unroll
      for (int a = 0; a < 3; a++) sink.add(slab * kProfileSums + 4u + (uint32_t)a, win.w[kk][a]);
    }
    sink.add(home * kProfileSums + 7u, pe);
#pragma unroll
    for (int a = 0; a < 3; a++) whole[a] += to_fixed(total[a], kProfileOne);
  }
}

// A pair's potential u and q = rF / r^2 (R10, FP64 in both forms).
template <bool SF>
__device__ __forceinline__ void profile_pair_terms(const Replica& R, double r2, double& u, double& q) {
  const double inv2 = rcp_rn(r2);  // = 1.0 / r2, correctly rounded
  const double s6 = inv2 * inv2 * inv2;
  double rf = s6 * __builtin_fma(48.0, s6, -24.0);
  const double t = 4.0 * s6;
  u = __builtin_fma(t, s6, -t) - R.shift;
  if (SF) {
    const double r = __builtin_sqrt(r2);
    rf = __builtin_fma(-r, R.force_shift, rf);
    u = __builtin_fma(r - R.rc, R.force_shift, u);
  }
  q = rf * inv2;
}

// A pair's virial half w spread over the slabs its segment from z (wrapped,
// in slab hb) to z + dz crosses, in proportion to the length inside each,
// the last slab taking what is left (R10): into the atom's window sums, or
// to the profile at once for a slab outside them.
__device__ __forceinline__ void walk_shares(double zw, double dz, int hb, int home, double width, double iw,
                                            uint32_t bins, const double w[3], SlabWindow& win,
                                            const ProfileSink& sink) {
  const int ibins = (int)bins;
  const double lo = zw + (dz < 0.0 ? dz : 0.0);
  const double hi = zw + (dz > 0.0 ? dz : 0.0);
  const double length = hi - lo;
  if (length < 1e-6 * width) {
    win.add(home - hb, home, ibins, w, 1.0, sink);
    return;
  }
  int b = (int)floor(lo * iw);
  double left = 1.0;
  for (uint32_t walk = 0; walk < bins + 2u; walk++) {
    const bool last = (double)(b + 1) * width >= hi;
    const double start = lo > (double)b * width ? lo : (double)b * width;
    const double end = hi < (double)(b + 1) * width ? hi : (double)(b + 1) * width;
    double share = last ? left : (end - start) / length;
    share = share > 0.0 ? share : 0.0;
    left -= share;
    win.add(b - hb, b, ibins, w, share, sink);
    if (last) break;
    b++;
  }
}

// One atom's pairs for slabs at least a cutoff wide, from the brick's LDS:
// the list in whole octets (pads meet the far sentinel, outside the cutoff),
// every pair through the same operations without a branch. A segment no
// longer than a slab lies in the atom's slab or crosses one slab edge, into
// the slab below or above; the walk above gives its shares as
//   s1 = (end - start) / length (clamped at 0) to the lower slab b and
//   s2 = 1 - s1 (clamped) to b + 1, or 1 to b when the segment ends in b,
// so each pair adds w s to the atom's own slab and, when it crosses, to the
// slab below or above; adds of pairs outside the cutoff are w 0 (w is
// finite for them, the sentinel's included). A pair the two-slab form does
// not cover (a segment ending exactly on an edge below the atom, a tiny
// segment of an atom whose slab was clamped into the box) takes the walk.
template <bool SF, class L>
__device__ __forceinline__ void wide_pairs_staged(const Replica& R, const Stage<L>& S, const uint16_t* lst,
                                                  uint32_t cnt, double xi, double yi, double zi, double zw, int hb,
                                                  int home, double width, double iw, uint32_t bins,
                                                  const ProfileSink& sink, double& pe, double total[3],
                                                  SlabWindow& win) {
  const double tiny_length = 1e-6 * width;
  const uint32_t no = (cnt + 7u) / 8u;
  const uint32_t last_octet = no > 0u ? no - 1u : 0u;
  u32x4 cur = load_octet(lst, 0u, last_octet);
  for (uint32_t o = 0; o < no; o++) {
    const u32x4 nxt = load_octet(lst, o + 1u, last_octet);
    uint32_t w0 = cur.x, w1 = cur.y, w2 = cur.z, w3 = cur.w;
    // one pair at a time (the pair body is long): the octet's entries shift
    // down 16 bits a pair
#pragma unroll 1
    for (int c = 0; c < 8; c++) {
      const uint32_t off = w0 & 0xffffu;  // 8 e
      w0 = (w0 >> 16) | (w1 << 16);
      w1 = (w1 >> 16) | (w2 << 16);
      w2 = (w2 >> 16) | (w3 << 16);
      w3 >>= 16;
      const double d[3] = {*(const double*)((const char*)S.x + off) - xi,
                           *(const double*)((const char*)S.y + off) - yi,
                           *(const double*)((const char*)S.z + off) - zi};
      const double r2 = dist2(d[0], d[1], d[2]);
      const bool inside = r2 < R.rc2;
      double u, q;
      profile_pair_terms<SF>(R, r2, u, q);
      const double sel = inside ? 1.0 : 0.0;
      pe = __builtin_fma(0.5 * u, sel, pe);
      double w[3];
#pragma unroll
      for (int a = 0; a < 3; a++) {
        w[a] = (0.5 * d[a]) * d[a] * q;
        total[a] = __builtin_fma(w[a], sel, total[a]);
      }
      const double lo = zw + (d[2] < 0.0 ? d[2] : 0.0);
      const double hi = zw + (d[2] > 0.0 ? d[2] : 0.0);
      const double length = hi - lo;
      const bool tiny = length < tiny_length;
      const int b = (int)floor(lo * iw);
      const double bw = (double)b * width, e1 = (double)(b + 1) * width;
      const bool last1 = e1 >= hi;
      const bool last2 = (double)(b + 2) * width >= hi;
      const double start = lo > bw ? lo : bw;
      const double end = hi < e1 ? hi : e1;
      double s1 = last1 ? 1.0 : (end - start) / length;
      s1 = s1 > 0.0 ? s1 : 0.0;
      double s2 = 1.0 - s1;
      s2 = s2 > 0.0 ? s2 : 0.0;
      const int rel = b - hb;
      const bool up = !tiny && !last1 && rel == 0, dn = !tiny && !last1 && rel == -1;
      const bool simple = tiny ? home == hb : (last1 ? rel == 0 : last2 && (up || dn));
      const bool take = inside && simple;
      const double so = take ? (tiny ? 1.0 : (dn ? s2 : s1)) : 0.0;
      const double sx = take ? (up ? s2 : s1) : 0.0;
      const double su = up ? sx : 0.0, sd = dn ? sx : 0.0;
#pragma unroll
      for (int a = 0; a < 3; a++) {
        win.w[0][a] += w[a] * sd;
        win.w[1][a] += w[a] * so;
        win.w[2][a] += w[a] * su;
      }
      if (inside && !simple) walk_shares(zw, d[2], hb, home, width, iw, bins, w, win, sink);
    }
    cur = nxt;
  }
}

// Slabs at least a cutoff wide: a wave's atoms put all their items in the
// slabs hmin - 1 .. hmin + kWaveSlots - 2, hmin the lowest home slab in the
// wave (a wave's home atoms lie in a few consecutive cells of one or two
// cell layers, so its home slabs mostly span one or two). Each (slab, sum) of these is added over the wave's lanes first and
// sent to the profile by one atomic, with no LDS bins (two workgroups per
// CU instead of one). A wave whose home slabs spread wider sends each
// lane's items on their own.
constexpr int kWaveSlots = 5;

__device__ __forceinline__ int wave_min_i32(int v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const int o = __shfl_xor(v, off, 64);
    v = o < v ? o : v;
  }
  return v;
}

template <bool SF, class L>
__device__ __forceinline__ void profile_part_wide(const Buffers& B, const Replica& R, const Stage<L>& S,
                                                  const Brick& g, bool staged, long long* prof, long long whole[3]) {
  const uint32_t bins = R.prof_bins;
  const double width = R.bin_width;
  const double iw = 1.0 / width;  // slab of z: floor(z iw)
  const int ibins = (int)bins;
  const ProfileSink direct{nullptr, prof, false};
  const uint32_t H = S.hoff[home_cells(g)];
  const uint32_t lane = threadIdx.x & 63u;
  // one wave-wide sum per slot of sum q, the lanes' values from `value`
  auto send = [&](int hmin, int q, auto value) {
#pragma unroll
    for (int rs = 0; rs < kWaveSlots; rs++) {
      const long long x = wave_sum_i64(value(rs));
      if (lane == 0 && x != 0) {
        const uint32_t slab = (uint32_t)((((hmin - 1 + rs) % ibins) + ibins) % ibins);
        atomicAdd((unsigned long long*)(prof + slab * kProfileSums + (uint32_t)q), (unsigned long long)x);
      }
    }
  };
  for (uint32_t h0 = 0; h0 < H; h0 += L::kThreads) {
    const uint32_t h = h0 + threadIdx.x;
    const bool active = h < H;
    Home at{0, 0, 0, 0};
    double xi = 0.0, yi = 0.0, zi = 0.0, zw = 0.0;
    int hb = 0x3fffffff;
    uint32_t home = 0;
    if (active) {
      at = home_atom(g, S, h);
      staged_pos(R, S, staged, at.e, B.x, B.y, B.z, xi, yi, zi);
      zw = wrap_one(zi, R.box[2]);
      hb = (int)floor(zw * iw);
      home = (uint32_t)(hb < 0 ? 0 : (hb > ibins - 1 ? ibins - 1 : hb));
    }
    const int hmin = wave_min_i32(hb);
    if (hmin == 0x3fffffff) continue;  // no atom in this wave
    const int hmax = -wave_min_i32(active ? -hb : -hmin);
    const bool together = hmax - hmin <= kWaveSlots - 3;  // the same for the whole wave
    const int hslot = (int)home - hmin + 1;  // this lane's home slab among the slots
    const int wslot = hb - hmin;             // ... and its window's first slab
    // count and v_a^2 to the home slab
    long long item[4] = {0, 0, 0, 0};
    if (active) {
      const uint32_t i = at.slot;
      const double v[3] = {B.vx[i], B.vy[i], B.vz[i]};
      item[0] = to_fixed(1.0, kProfileOne);
#pragma unroll
      for (int a = 0; a < 3; a++) item[1 + a] = to_fixed(v[a] * v[a], kProfileOne);
    }
    if (together) {
#pragma unroll
      for (int q = 0; q < 4; q++) send(hmin, q, [&](int rs) { return active && hslot == rs ? item[q] : 0ll; });
    } else if (active) {
#pragma unroll
      for (int q = 0; q < 4; q++) direct.add(home * kProfileSums + (uint32_t)q, item[q]);
    }
    SlabWindow win;
    win.clear();
    long long pe_fixed = 0;
    if (active) {
      const uint32_t i = at.slot;
      double pe = 0.0, total[3] = {0.0, 0.0, 0.0};
      const uint32_t cnt = B.nbr_count[i];
      const uint16_t* lst = B.nbr + list_base(R, i - R.atom_off);
      if (staged) {
        wide_pairs_staged<SF>(R, S, lst, cnt, xi, yi, zi, zw, hb, (int)home, width, iw, bins, direct, pe, total,
                              win);
      } else {
        each_entry(lst, cnt, [&](uint32_t k, uint32_t entry) {
          double xj, yj, zj;
          staged_pos(R, S, false, list_index(entry, false), B.x, B.y, B.z, xj, yj, zj);
          const double d[3] = {xj - xi, yj - yi, zj - zi};
          const double r2 = dist2(d[0], d[1], d[2]);
          if (!(r2 < R.rc2)) return;
          double u, q;
          profile_pair_terms<SF>(R, r2, u, q);
          pe += 0.5 * u;
          double w[3];
#pragma unroll
          for (int a = 0; a < 3; a++) {
            w[a] = (0.5 * d[a]) * d[a] * q;
            total[a] += w[a];
          }
          walk_shares(zw, d[2], hb, (int)home, width, iw, bins, w, win, direct);
        });
      }
      pe_fixed = to_fixed(pe, kProfileOne);
#pragma unroll
      for (int a = 0; a < 3; a++) whole[a] += to_fixed(total[a], kProfileOne);
    }
    long long wf[3][3];  // the window's sums in fixed point
#pragma unroll
    for (int kk = 0; kk < 3; kk++)
#pragma unroll
      for (int a = 0; a < 3; a++) wf[kk][a] = active ? to_fixed(win.w[kk][a], kProfileOne) : 0ll;
    if (together) {
      send(hmin, 7, [&](int rs) { return active && hslot == rs ? pe_fixed : 0ll; });
#pragma unroll
      for (int a = 0; a < 3; a++)
        send(hmin, 4 + a, [&](int rs) {
          long long x = 0;
#pragma unroll
          for (int kk = 0; kk < 3; kk++) x += active && wslot + kk == rs ? wf[kk][a] : 0ll;
          return x;
        });
    } else if (active) {
      direct.add(home * kProfileSums + 7u, pe_fixed);
#pragma unroll
      for (int kk = 0; kk < 3; kk++) {
        const uint32_t slab = (uint32_t)((((hb - 1 + kk) % ibins) + ibins) % ibins);
#pragma unroll
        for (int a = 0; a < 3; a++) direct.add(slab * kProfileSums + 4u + (uint32_t)a, wf[kk][a]);
      }
    }
  }
}

// Adds the workgroup's sums of one brick to the replica's profile.
template <class L>
__device__ __forceinline__ void profile_flush(const ProfileSink& sink, uint32_t bins, const long long whole[3]) {
  const uint32_t nsums = bins * kProfileSums + 3u;
#pragma unroll
  for (int a = 0; a < 3; a++) sink.add(bins * kProfileSums + (uint32_t)a, whole[a]);
  if (sink.use_lds) {
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < nsums; k += L::kThreads)
      if (sink.lds[k] != 0) atomicAdd((unsigned long long*)(sink.global + k), (unsigned long long)sink.lds[k]);
  }
}

// WIDE: every replica's slabs are at least a cutoff wide (profile_part_wide,
// no LDS bins); its main launch over dense bricks is held to 128 VGPRs, so
// that two 512-thread workgroups share a CU.
template <class L, bool SF, bool PARTS, bool WIDE>
__global__ void __launch_bounds__(L::kThreads, WIDE && !PARTS && !L::kSparse ? 4 : 1) k_profile(Buffers B) {
  __shared__ Stage<L> S;
  // brick gb whole (code 0) or one part of it
  auto profile_wide = [&](uint32_t gb, uint32_t code) {
    const Brick brick = B.bricks[gb];
    const Replica R = B.rep[brick.rep];
    if (R.prof_bins == 0) return;  // the whole workgroup together
    const Brick g = PARTS ? part_brick(brick, code, R) : brick;
    if (PARTS) __syncthreads();  // the previous part is done with S
    const bool staged = stage_brick(B, R, g, B.x, B.y, B.z, S);  // ends with a barrier
    if (!PARTS && !staged && home_cells(g) > 1u && B.parts_on) return;  // cut into parts
    long long whole[3] = {0, 0, 0};
    long long* prof = B.profile + R.prof_off;
    profile_part_wide<SF>(B, R, S, g, staged, prof, whole);
#pragma unroll
    for (int a = 0; a < 3; a++) {
      const long long x = wave_sum_i64(whole[a]);
      if ((threadIdx.x & 63u) == 0 && x != 0)
        atomicAdd((unsigned long long*)(prof + R.prof_bins * kProfileSums + (uint32_t)a), (unsigned long long)x);
    }
  };
  if constexpr (WIDE) {
    if constexpr (!PARTS) {
      const uint32_t gb = layout_first<L>(B) + dealt_index(blockIdx.x);
      if (gb < layout_end<L>(B)) profile_wide(gb, 0u);
    } else {
      each_cut_part<L>(B, profile_wide);
    }
    return;
  }
  __shared__ long long acc[WIDE ? 1 : kProfileLdsBins * kProfileSums + 3];
  auto profile = [&](uint32_t gb, uint32_t code) {
    const Brick brick = B.bricks[gb];
    const Replica R = B.rep[brick.rep];
    const uint32_t bins = R.prof_bins;
    if (bins == 0) return;  // the whole workgroup together
    const uint32_t nsums = bins * kProfileSums + 3u;
    const ProfileSink sink{acc, B.profile + R.prof_off, bins <= kProfileLdsBins};
    const Brick g = PARTS ? part_brick(brick, code, R) : brick;
    if (PARTS) __syncthreads();  // the previous part is done with S and acc
    if (sink.use_lds)
      for (uint32_t k = threadIdx.x; k < nsums; k += L::kThreads) acc[k] = 0;
    const bool staged = stage_brick(B, R, g, B.x, B.y, B.z, S);  // ends with a barrier
    if (!PARTS && !staged && home_cells(g) > 1u && B.parts_on) return;  // cut into parts
    long long whole[3] = {0, 0, 0};
    profile_part<SF>(B, R, S, g, staged, sink, whole);
    profile_flush<L>(sink, bins, whole);
  };
  if constexpr (!WIDE) {
    if constexpr (!PARTS) {
      const uint32_t gb = layout_first<L>(B) + dealt_index(blockIdx.x);
      if (gb < layout_end<L>(B)) profile(gb, 0u);
    } else {
      each_cut_part<L>(B, profile);
    }
  }
}

// ---------------------------------------------------------------------------
// Heat current (lennard_jones.wgsl:872-947, R14): J = sum_i e_i v_i +
// 1/2 sum_i sum_j (F_ij . v_i) r_ij, e_i the kinetic energy plus half of
// each pair energy; per atom over its list in order, each component to
// fixed point at 2^32, one 64-bit atomic per workgroup and component into
// sample `sample` of the series. Pair terms in FP64.
// ---------------------------------------------------------------------------

template <bool SF, class L>
__device__ __forceinline__ void heat_part(const Buffers& B, const Replica& R, const Stage<L>& S, const Brick& g,
                                          bool staged, long long jsum[kSlots]) {
  const uint32_t H = S.hoff[home_cells(g)];
  for (uint32_t h = threadIdx.x; h < H; h += L::kThreads) {
    const Home at = home_atom(g, S, h);
    const uint32_t i = at.slot;
    double xi, yi, zi;
    staged_pos(R, S, staged, at.e, B.x, B.y, B.z, xi, yi, zi);
    const double v[3] = {B.vx[i], B.vy[i], B.vz[i]};
    double e = 0.5 * dist2(v[0], v[1], v[2]);
    double pr[3] = {0.0, 0.0, 0.0};
    const uint32_t cnt = B.nbr_count[i];
    const uint16_t* lst = B.nbr + list_base(R, i - R.atom_off);
    each_entry(lst, cnt, [&](uint32_t k, uint32_t entry) {
      double xj, yj, zj;
      staged_pos(R, S, staged, list_index(entry, staged), B.x, B.y, B.z, xj, yj, zj);
      const double d[3] = {xj - xi, yj - yi, zj - zi};
      const double r2 = dist2(d[0], d[1], d[2]);
      if (!(r2 < R.rc2)) return;
      const double inv2 = 1.0 / r2;
      const double s6 = inv2 * inv2 * inv2;
      double rf = s6 * __builtin_fma(48.0, s6, -24.0);
      const double t = 4.0 * s6;
      double u = __builtin_fma(t, s6, -t) - R.shift;
      if (SF) {
        const double r = __builtin_sqrt(r2);
        rf = __builtin_fma(-r, R.force_shift, rf);
        u = __builtin_fma(r - R.rc, R.force_shift, u);
      }
      e += 0.5 * u;
      const double dv = __builtin_fma(d[2], v[2], __builtin_fma(d[1], v[1], d[0] * v[0]));
      const double q = dv * rf / r2;
#pragma unroll
      for (int a = 0; a < 3; a++) pr[a] = __builtin_fma(d[a], q, pr[a]);
    });
#pragma unroll
    for (int a = 0; a < 3; a++) jsum[a] += to_fixed(__builtin_fma(e, v[a], 0.5 * pr[a]), 4294967296.0);
  }
}

template <class L, bool SF, bool PARTS>
__global__ void __launch_bounds__(L::kThreads) k_heat_current(Buffers B, uint32_t sample) {
  __shared__ Stage<L> S;
  if constexpr (!PARTS) {
    const uint32_t gb = layout_first<L>(B) + dealt_index(blockIdx.x);
    if (gb >= layout_end<L>(B)) return;
    const Brick g = B.bricks[gb];
    const Replica R = B.rep[g.rep];
    if (R.hc_cap == 0u || sample >= R.hc_cap) return;
    const bool staged = stage_brick(B, R, g, B.x, B.y, B.z, S);
    if (!staged && home_cells(g) > 1u && B.parts_on) return;  // cut into parts
    long long jsum[kSlots] = {0, 0, 0, 0, 0, 0, 0, 0};
    heat_part<SF>(B, R, S, g, staged, jsum);
    block_add_slots<3, L::kThreads>(jsum, B.hc + 3ull * sample);
  } else {
    each_cut_part<L>(B, [&](uint32_t gb, uint32_t code) {
      const Brick brick = B.bricks[gb];
      const Replica R = B.rep[brick.rep];
      if (R.hc_cap == 0u || sample >= R.hc_cap) return;
      const Brick g = part_brick(brick, code, R);
      __syncthreads();  // the previous part is done with S and the sums
      const bool staged = stage_brick(B, R, g, B.x, B.y, B.z, S);
      long long jsum[kSlots] = {0, 0, 0, 0, 0, 0, 0, 0};
      heat_part<SF>(B, R, S, g, staged, jsum);
      block_add_slots<3, L::kThreads>(jsum, B.hc + 3ull * sample);
    });
  }
}

// ---------------------------------------------------------------------------
// Render arrays (lennard_jones.rs:1041-1051, lennard_jones.wgsl:596-603):
// float4 per slot, the position wrapped into the box (w 0) and the velocity
// with, in w, the neighbours within 1.5 sigma or (drawn 1) m v^2 / 3
// averaged over the molecule and its neighbours within the cutoff.
// ---------------------------------------------------------------------------

template <class L>
__device__ __forceinline__ void render_part(const Buffers& B, const Replica& R, const Stage<L>& S, const Brick& g,
                                            bool staged) {
  const uint32_t H = S.hoff[home_cells(g)];
  for (uint32_t h = threadIdx.x; h < H; h += L::kThreads) {
    const Home at = home_atom(g, S, h);
    const uint32_t i = at.slot;
    double xi, yi, zi;
    staged_pos(R, S, staged, at.e, B.x, B.y, B.z, xi, yi, zi);
    const double v[3] = {B.vx[i], B.vy[i], B.vz[i]};
    uint32_t near = 0, within = 0;
    double v2 = 0.0;
    const uint32_t cnt = B.nbr_count[i];
    const uint16_t* lst = B.nbr + list_base(R, i - R.atom_off);
    each_entry(lst, cnt, [&](uint32_t k, uint32_t entry) {
      const uint32_t e = list_index(entry, staged);
      double xj, yj, zj;
      staged_pos(R, S, staged, e, B.x, B.y, B.z, xj, yj, zj);
      const double r2 = dist2(xj - xi, yj - yi, zj - zi);
      if (!(r2 < R.rc2)) return;
      within++;
      near += r2 < kNearR2 ? 1u : 0u;
      if (R.drawn == 1u) {
        const uint32_t c = stage_cell_of(S, e);
        const uint32_t j = S.base[c] + e;
        v2 += dist2(B.vx[j], B.vy[j], B.vz[j]);
      }
    });
    const float w = R.drawn == 1u ? (float)((dist2(v[0], v[1], v[2]) + v2) / (3.0 * (double)(within + 1u)))
                                  : (float)near;
    float4* P = reinterpret_cast<float4*>(B.render_pos);
    float4* V = reinterpret_cast<float4*>(B.render_vel);
    P[i] = make_float4((float)wrap_one(xi, R.box[0]), (float)wrap_one(yi, R.box[1]), (float)wrap_one(zi, R.box[2]),
                       0.0f);
    V[i] = make_float4((float)v[0], (float)v[1], (float)v[2], w);
  }
}

template <class L, bool PARTS>
__global__ void __launch_bounds__(L::kThreads) k_render(Buffers B) {
  __shared__ Stage<L> S;
  if constexpr (!PARTS) {
    const uint32_t gb = layout_first<L>(B) + dealt_index(blockIdx.x);
    if (gb >= layout_end<L>(B)) return;
    const Brick g = B.bricks[gb];
    const Replica R = B.rep[g.rep];
    const bool staged = stage_brick(B, R, g, B.x, B.y, B.z, S);
    if (!staged && home_cells(g) > 1u && B.parts_on) return;  // cut into parts
    render_part(B, R, S, g, staged);
  } else {
    each_cut_part<L>(B, [&](uint32_t gb, uint32_t code) {
      const Brick brick = B.bricks[gb];
      const Replica R = B.rep[brick.rep];
      const Brick g = part_brick(brick, code, R);
      __syncthreads();  // the previous part is done with S
      const bool staged = stage_brick(B, R, g, B.x, B.y, B.z, S);
      render_part(B, R, S, g, staged);
    });
  }
}

__global__ void __launch_bounds__(kBlock) k_zero_slots(Buffers B) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < B.n_replicas * kSlots) B.slots[t] = 0;
}

// Generic 64-bit fixed-point sum (the device half of vlj_fixed_sum).
__global__ void __launch_bounds__(kBlock) k_fixed_sum(const double* values, size_t n, double scale,
                                                        long long* out) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  long long v[kSlots] = {0, 0, 0, 0, 0, 0, 0, 0};
  if (t < n) v[0] = to_fixed(values[t], scale);
  block_add_slots(v, out);
}

// rcp_rn over an array (a check of its rounding against IEEE division).
__global__ void __launch_bounds__(kBlock) k_rcp(const double* x, double* out, size_t n) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) out[t] = rcp_rn(x[t]);
}

// rcp_rn_f32 and the FP32 square root of the shifted-force pair over an
// array (checks of their rounding against IEEE division and square root).
__global__ void __launch_bounds__(kBlock) k_rcp32(const float* x, float* rcp, float* root, size_t n) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) {
    rcp[t] = rcp_rn_f32(x[t]);
    root[t] = __builtin_sqrtf(x[t]);
  }
}

__global__ void k_zero_i64(long long* p, uint32_t n) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) p[t] = 0;
}

// Plain streaming copy of n 16-byte words: reads and writes exactly 16 n
// bytes. It calibrates the HBM byte counters and measures the bandwidth a
// streaming kernel reaches on the device.
__global__ void __launch_bounds__(kBlock) k_copy(const u32x4* __restrict__ src, u32x4* __restrict__ dst,
                                                   size_t n) {
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x)
    __builtin_nontemporal_store(__builtin_nontemporal_load(src + t), dst + t);
}

// FP64 FMA chains: eight independent chains per thread, `iters` dependent
// fused multiply-adds each (2 flops), the eight results summed and stored
// so that none is dropped. Measures the FP64 vector rate the device reaches.
__global__ void __launch_bounds__(kBlock) k_fma(double* out, uint32_t iters, double m, double c) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  double a[8];
#pragma unroll
  for (int j = 0; j < 8; j++) a[j] = 1.0 + 1e-9 * (double)t + 0.125 * (double)j;
  for (uint32_t k = 0; k < iters; k++) {
#pragma unroll
    for (int j = 0; j < 8; j++) a[j] = __builtin_fma(a[j], m, c);
  }
  double s = 0.0;
#pragma unroll
  for (int j = 0; j < 8; j++) s += a[j];
  out[t] = s;
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------

static inline dim3 grid1(uint32_t blocks) { return dim3(blocks, 1, 1); }

static inline dim3 decide_grid(const Buffers& B) { return dim3(B.n_replicas, 1, 1); }

hipError_t launch_decide(const Buffers& B, hipStream_t s) {
  hipLaunchKernelGGL(k_decide, decide_grid(B), dim3(kBlock), 0, s, B, 0);
  return hipGetLastError();
}

hipError_t launch_drift(const Buffers& B, bool langevin, bool keep_v, hipStream_t s) {
  if (!B.n_atom_blocks) return hipSuccess;
  if (langevin)
    hipLaunchKernelGGL(k_drift<true>, grid1(dealt_grid(B.n_atom_blocks)), dim3(kBlock), 0, s, B, keep_v ? 1 : 0);
  else
    hipLaunchKernelGGL(k_drift<false>, grid1(dealt_grid(B.n_atom_blocks)), dim3(kBlock), 0, s, B, keep_v ? 1 : 0);
  hipLaunchKernelGGL(k_decide, decide_grid(B), dim3(kBlock), 0, s, B, 1);
  return hipGetLastError();
}


hipError_t launch_nh_half(const Buffers& B, bool scale, hipStream_t s) {
  if (!B.n_atom_blocks) return hipSuccess;
  hipLaunchKernelGGL(k_ke_sum, grid1(dealt_grid(B.n_atom_blocks)), dim3(kBlock), 0, s, B);
  hipLaunchKernelGGL(k_nh_update, decide_grid(B), dim3(kBlock), 0, s, B);
  if (scale) hipLaunchKernelGGL(k_nh_scale, grid1(dealt_grid(B.n_atom_blocks)), dim3(kBlock), 0, s, B);
  return hipGetLastError();
}

hipError_t launch_heat_exchange(const Buffers& B, hipStream_t s) {
  if (!B.n_atom_blocks) return hipSuccess;
  hipLaunchKernelGGL(k_hex_sum, grid1(dealt_grid(B.n_atom_blocks)), dim3(kBlock), 0, s, B);
  hipLaunchKernelGGL(k_hex_update, decide_grid(B), dim3(kBlock), 0, s, B);
  hipLaunchKernelGGL(k_hex_apply, grid1(dealt_grid(B.n_atom_blocks)), dim3(kBlock), 0, s, B);
  return hipGetLastError();
}

// Grid of a launch over the bricks that are cut into parts.
inline dim3 parts_grid(const Buffers& B) { return grid1(std::min<uint32_t>(B.n_bricks, kPartsGrid)); }

// Brick kernels run as up to four launches: the dense bricks, the sparse
// bricks, then the parts of cut bricks of each layout (which return after
// one load when no brick is cut). Layouts without bricks are not launched.
template <class L>
inline uint32_t layout_count(const Buffers& B) {
  return L::kSparse ? B.n_bricks - B.n_dense : B.n_dense;
}
template <class L>
inline dim3 brick_grid(const Buffers& B) {
  return grid1(dealt_grid(layout_count<L>(B)));
}

#define VLJ_BRICK_LAUNCHES(KERNEL, ...)                                                                   \
  do {                                                                                                    \
    if (layout_count<Dense>(B))                                                                           \
      hipLaunchKernelGGL((KERNEL<Dense, false>), brick_grid<Dense>(B), dim3(Dense::kThreads), 0, s, __VA_ARGS__); \
    if (layout_count<Sparse>(B))                                                                          \
      hipLaunchKernelGGL((KERNEL<Sparse, false>), brick_grid<Sparse>(B), dim3(Sparse::kThreads), 0, s, __VA_ARGS__); \
    if (B.parts_on && layout_count<Dense>(B))                                                             \
      hipLaunchKernelGGL((KERNEL<Dense, true>), parts_grid(B), dim3(Dense::kThreads), 0, s, __VA_ARGS__);  \
    if (B.parts_on && layout_count<Sparse>(B))                                                            \
      hipLaunchKernelGGL((KERNEL<Sparse, true>), parts_grid(B), dim3(Sparse::kThreads), 0, s, __VA_ARGS__); \
  } while (0)

hipError_t launch_render(const Buffers& B, hipStream_t s) {
  if (!B.n_bricks || !B.render_pos) return hipSuccess;
  VLJ_BRICK_LAUNCHES(k_render, B);
  return hipGetLastError();
}

template <bool SF>
void heat_launches(const Buffers& B, uint32_t sample, hipStream_t s) {
  const dim3 gp = parts_grid(B);
  if (layout_count<Dense>(B)) {
    hipLaunchKernelGGL((k_heat_current<Dense, SF, false>), brick_grid<Dense>(B), dim3(Dense::kThreads), 0, s, B, sample);
  }
  if (layout_count<Sparse>(B)) {
    hipLaunchKernelGGL((k_heat_current<Sparse, SF, false>), brick_grid<Sparse>(B), dim3(Sparse::kThreads), 0, s, B,
                       sample);
  }
  if (B.parts_on && layout_count<Dense>(B)) hipLaunchKernelGGL((k_heat_current<Dense, SF, true>), gp, dim3(Dense::kThreads), 0, s, B, sample);
  if (B.parts_on && layout_count<Sparse>(B))
    hipLaunchKernelGGL((k_heat_current<Sparse, SF, true>), gp, dim3(Sparse::kThreads), 0, s, B, sample);
}

hipError_t launch_heat_current(const Buffers& B, bool shifted, uint32_t sample, hipStream_t s) {
  if (!B.n_bricks) return hipSuccess;
  if (shifted)
    heat_launches<true>(B, sample, s);
  else
    heat_launches<false>(B, sample, s);
  return hipGetLastError();
}

// The stream the sparse-layout main launch goes to: the side stream, forked
// from s, when both layouts have bricks; s otherwise. join_side() makes s
// wait for it again.
static hipStream_t fork_side(const Buffers& B, hipStream_t s, const Side* side, hipError_t& e) {
  if (!side || !side->stream || !layout_count<Dense>(B) || !layout_count<Sparse>(B)) return s;
  if (e == hipSuccess) e = hipEventRecord(side->fork, s);
  if (e == hipSuccess) e = hipStreamWaitEvent(side->stream, side->fork, 0);
  return side->stream;
}

static void join_side(hipStream_t s, hipStream_t ss, const Side* side, hipError_t& e) {
  if (ss == s) return;
  if (e == hipSuccess) e = hipEventRecord(side->join, ss);
  if (e == hipSuccess) e = hipStreamWaitEvent(s, side->join, 0);
}

hipError_t launch_rebuild(const Buffers& B, hipStream_t s, const Side* side) {
  if (!B.n_atom_blocks) return hipSuccess;
  hipLaunchKernelGGL(k_bin, grid1(dealt_grid(B.n_atom_blocks)), dim3(kBlock), 0, s, B);
  hipLaunchKernelGGL(k_scan, grid1(B.n_scan_blocks), dim3(kBlock), 0, s, B);
  hipLaunchKernelGGL(k_scatter, grid1(dealt_grid(B.n_atom_blocks)), dim3(kBlock), 0, s, B);
  hipLaunchKernelGGL(k_sort, grid1(dealt_grid(B.n_sort_blocks)), dim3(kBlock), 0, s, B);
  hipLaunchKernelGGL(k_reorder, grid1(dealt_grid(B.n_atom_blocks)), dim3(kBlock), 0, s, B);
  if (!B.n_bricks) return hipGetLastError();
  hipError_t e = hipSuccess;
  const hipStream_t ss = fork_side(B, s, side, e);
  if (layout_count<Dense>(B))
    hipLaunchKernelGGL((k_build<Dense, false>), brick_grid<Dense>(B), dim3(Dense::kThreads), 0, s, B, 0);
  if (layout_count<Sparse>(B))
    hipLaunchKernelGGL((k_build<Sparse, false>), brick_grid<Sparse>(B), dim3(Sparse::kThreads), 0, ss, B, 0);
  join_side(s, ss, side, e);
  if (B.parts_on && layout_count<Dense>(B)) hipLaunchKernelGGL((k_build<Dense, true>), parts_grid(B), dim3(Dense::kThreads), 0, s, B, 0);
  if (B.parts_on && layout_count<Sparse>(B)) hipLaunchKernelGGL((k_build<Sparse, true>), parts_grid(B), dim3(Sparse::kThreads), 0, s, B, 0);
  return e != hipSuccess ? e : hipGetLastError();
}

hipError_t launch_relist(const Buffers& B, hipStream_t s) {
  if (!B.n_bricks) return hipSuccess;
  if (layout_count<Dense>(B))
    hipLaunchKernelGGL((k_build<Dense, false>), brick_grid<Dense>(B), dim3(Dense::kThreads), 0, s, B, 1);
  if (layout_count<Sparse>(B))
    hipLaunchKernelGGL((k_build<Sparse, false>), brick_grid<Sparse>(B), dim3(Sparse::kThreads), 0, s, B, 1);
  if (B.parts_on && layout_count<Dense>(B)) hipLaunchKernelGGL((k_build<Dense, true>), parts_grid(B), dim3(Dense::kThreads), 0, s, B, 1);
  if (B.parts_on && layout_count<Sparse>(B)) hipLaunchKernelGGL((k_build<Sparse, true>), parts_grid(B), dim3(Sparse::kThreads), 0, s, B, 1);
  return hipGetLastError();
}

template <int PREC, bool SF>
hipError_t force_launches(const Buffers& B, int k, hipStream_t s, const Side* side) {
  const dim3 gp = parts_grid(B);
  hipError_t e = hipSuccess;
  const hipStream_t ss = fork_side(B, s, side, e);
  if (layout_count<Dense>(B))
    hipLaunchKernelGGL((k_force<Dense, PREC, SF, false>), brick_grid<Dense>(B), dim3(Dense::kThreads), 0, s, B, k);
  if (layout_count<Sparse>(B))
    hipLaunchKernelGGL((k_force<Sparse, PREC, SF, false>), brick_grid<Sparse>(B), dim3(Sparse::kThreads), 0, ss, B, k);
  join_side(s, ss, side, e);
  if (B.parts_on && layout_count<Dense>(B)) hipLaunchKernelGGL((k_force<Dense, PREC, SF, true>), gp, dim3(Dense::kThreads), 0, s, B, k);
  if (B.parts_on && layout_count<Sparse>(B)) hipLaunchKernelGGL((k_force<Sparse, PREC, SF, true>), gp, dim3(Sparse::kThreads), 0, s, B, k);
  return e;
}

hipError_t launch_force(const Buffers& B, int precision, bool shifted, int kick, hipStream_t s, const Side* side) {
  if (!B.n_atom_blocks) return hipSuccess;
  const int k = kick;
  hipError_t e;
  if (precision == 0)
    e = shifted ? force_launches<0, true>(B, k, s, side) : force_launches<0, false>(B, k, s, side);
  else
    e = shifted ? force_launches<1, true>(B, k, s, side) : force_launches<1, false>(B, k, s, side);
  return e != hipSuccess ? e : hipGetLastError();
}

template <int PREC, bool SF>
void measure_launches(const Buffers& B, hipStream_t s) {
  const dim3 gp = parts_grid(B);
  if (layout_count<Dense>(B))
    hipLaunchKernelGGL((k_measure<Dense, PREC, SF, false>), brick_grid<Dense>(B), dim3(Dense::kThreads), 0, s, B);
  if (layout_count<Sparse>(B))
    hipLaunchKernelGGL((k_measure<Sparse, PREC, SF, false>), brick_grid<Sparse>(B), dim3(Sparse::kThreads), 0, s, B);
  if (B.parts_on && layout_count<Dense>(B)) hipLaunchKernelGGL((k_measure<Dense, PREC, SF, true>), gp, dim3(Dense::kThreads), 0, s, B);
  if (B.parts_on && layout_count<Sparse>(B)) hipLaunchKernelGGL((k_measure<Sparse, PREC, SF, true>), gp, dim3(Sparse::kThreads), 0, s, B);
}

hipError_t launch_measure(const Buffers& B, int precision, bool shifted, hipStream_t s) {
  const uint32_t zb = (B.n_replicas * kSlots + kBlock - 1) / kBlock;
  hipLaunchKernelGGL(k_zero_slots, grid1(zb), dim3(kBlock), 0, s, B);
  if (!B.n_atom_blocks) return hipGetLastError();
  if (precision == 0) {
    if (shifted)
      measure_launches<0, true>(B, s);
    else
      measure_launches<0, false>(B, s);
  } else {
    if (shifted)
      measure_launches<1, true>(B, s);
    else
      measure_launches<1, false>(B, s);
  }
  return hipGetLastError();
}

template <bool SF, bool WIDE>
void profile_launches(const Buffers& B, hipStream_t s) {
  const dim3 gp = parts_grid(B);
  if (layout_count<Dense>(B))
    hipLaunchKernelGGL((k_profile<Dense, SF, false, WIDE>), brick_grid<Dense>(B), dim3(Dense::kThreads), 0, s, B);
  if (layout_count<Sparse>(B))
    hipLaunchKernelGGL((k_profile<Sparse, SF, false, WIDE>), brick_grid<Sparse>(B), dim3(Sparse::kThreads), 0, s, B);
  if (B.parts_on && layout_count<Dense>(B))
    hipLaunchKernelGGL((k_profile<Dense, SF, true, WIDE>), gp, dim3(Dense::kThreads), 0, s, B);
  if (B.parts_on && layout_count<Sparse>(B))
    hipLaunchKernelGGL((k_profile<Sparse, SF, true, WIDE>), gp, dim3(Sparse::kThreads), 0, s, B);
}

hipError_t launch_profile(const Buffers& B, bool shifted, bool wide, hipStream_t s) {
  if (!B.n_bricks) return hipSuccess;
  if (shifted)
    wide ? profile_launches<true, true>(B, s) : profile_launches<true, false>(B, s);
  else
    wide ? profile_launches<false, true>(B, s) : profile_launches<false, false>(B, s);
  return hipGetLastError();
}

hipError_t launch_rcp(const double* x, double* out, size_t n, hipStream_t s) {
  if (n) hipLaunchKernelGGL(k_rcp, grid1((uint32_t)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, s, x, out, n);
  return hipGetLastError();
}

hipError_t launch_rcp32(const float* x, float* rcp, float* root, size_t n, hipStream_t s) {
  if (n) hipLaunchKernelGGL(k_rcp32, grid1((uint32_t)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, s, x, rcp, root, n);
  return hipGetLastError();
}

hipError_t launch_zero(long long* p, uint32_t n, hipStream_t s) {
  if (n) hipLaunchKernelGGL(k_zero_i64, grid1((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, p, n);
  return hipGetLastError();
}

hipError_t launch_fixed_sum(const double* values, size_t n, double scale, long long* out,
                            hipStream_t s) {
  hipLaunchKernelGGL(k_zero_i64, dim3(1), dim3(kBlock), 0, s, out, (uint32_t)kSlots);
  if (n) {
    const uint32_t blocks = (uint32_t)((n + kBlock - 1) / kBlock);
    hipLaunchKernelGGL(k_fixed_sum, grid1(blocks), dim3(kBlock), 0, s, values, n, scale, out);
  }
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// The fixed-point sum without allocation (vlj_fixed_sum_device): its
// per-workgroup partial sums and last-workgroup ticket live in a file-scope
// device array, one copy per GPU, whose address is looked up once per device.
// Every call on a GPU shares that GPU's copy: calls that overlap on one GPU
// race. The ticket is zero when the module loads and the last workgroup of
// every call sets it back to zero, so no call needs to clear anything.
// ---------------------------------------------------------------------------

struct FixedSumScratch {
  long long partial[kFixedSumBlocks];  // one per workgroup of the launch
  uint32_t ticket;                     // workgroups done
  uint32_t pad[63];
};
static __device__ __align__(256) FixedSumScratch scratch;

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
FixedSumScratch* g_scratch[kMaxDevices] = {};

// The device address of `scratch` on the current device (looked up on the
// first call there). In host code the name `scratch` is only a handle for
// the lookup, never a device address.
static hipError_t scratch_of_device(FixedSumScratch** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (!p) return hipErrorInvalidSymbol;
    g_scratch[dev] = (FixedSumScratch*)p;
  }
  *out = g_scratch[dev];
  return hipSuccess;
}

// Grid-stride sum of round-ties-even(values[i] * scale) per workgroup into
// sc->partial; the last workgroup adds the partials and writes *out.
__global__ void __launch_bounds__(kBlock) k_fixed_sum_partials(const double* values, size_t n, double scale,
                                                                 FixedSumScratch* sc, long long* out) {
  __shared__ long long part[kBlock / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  long long acc = 0;
  for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock)
    acc += to_fixed(values[i], scale);
  acc = wave_sum_i64(acc);
  if (lane == 0) part[wave] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    long long t = 0;
#pragma unroll
    for (int w = 0; w < kBlock / 64; w++) t += part[w];
    sc->partial[blockIdx.x] = t;
  }
  if (!last_block(&sc->ticket)) return;
  long long t = 0;
  for (uint32_t b = threadIdx.x; b < gridDim.x; b += kBlock)
    t += __hip_atomic_load(&sc->partial[b], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  t = wave_sum_i64(t);
  __syncthreads();  // part[] was read above by thread 0
  if (lane == 0) part[wave] = t;
  __syncthreads();
  if (threadIdx.x == 0) {
    long long total = 0;
#pragma unroll
    for (int w = 0; w < kBlock / 64; w++) total += part[w];
    *out = total;
    atomicExch(&sc->ticket, 0u);
  }
}

hipError_t launch_fixed_sum_noalloc(const double* values, size_t n, double scale, long long* out, hipStream_t s) {
  FixedSumScratch* sc = nullptr;
  hipError_t e = scratch_of_device(&sc);
  if (e != hipSuccess) return e;
  const size_t want = (n + kBlock - 1) / kBlock;
  const uint32_t blocks = (uint32_t)std::max<size_t>(1, std::min<size_t>(want, kFixedSumBlocks));
  hipLaunchKernelGGL(k_fixed_sum_partials, grid1(blocks), dim3(kBlock), 0, s, values, n, scale, sc, out);
  return hipGetLastError();
}

hipError_t launch_fma(double* out, uint32_t blocks, uint32_t iters, hipStream_t s) {
  hipLaunchKernelGGL(k_fma, grid1(blocks), dim3(kBlock), 0, s, out, iters, 0.999999999, 1e-12);
  return hipGetLastError();
}

hipError_t launch_copy(const void* src, void* dst, size_t words, hipStream_t s) {
  const uint32_t blocks = (uint32_t)std::min<size_t>((words + kBlock - 1) / kBlock, 256u * 64u);
  hipLaunchKernelGGL(k_copy, grid1(blocks), dim3(kBlock), 0, s, (const u32x4*)src, (u32x4*)dst, words);
  return hipGetLastError();
}

}  // namespace vlj
