// HIP kernels of the Lennard-Jones path, written for CDNA4 (gfx950, wave64).
//
// One step is (lennard_jones.rs:1092-1138):
//   drift       half kick + drift (or BAOAB), wrap, max displacement; the last
//               workgroup to finish takes the rebuild decision for every
//               replica (lennard_jones.wgsl:204-307)
//   bin         cell of each atom, count per cell          } rebuild chain,
//   scan        exclusive prefix sum of counts             } each kernel
//   scatter     atoms into their cell's range              } returns at once
//   sort        each cell's slots ascending (wave per cell)} when its replica
//   reorder     atoms into cell order (every 4th rebuild)  } does not rebuild
//   build       Verlet lists, transposed                   } (grid.wgsl, wgsl:419-499)
//   force       pair forces from the lists + second half kick (wgsl:501-574)
// The sequence is static, so it is captured once into a hipGraph.
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

__device__ void decide(const Buffers& B, uint32_t r, bool advance_step) {
  Control* c = B.ctl + r;
  const Replica& R = B.rep[r];
  unsigned long long bits = atomicExch(&c->max_moved2, 0ull);
  double moved2 = __longlong_as_double((long long)bits);
  uint32_t built = c->built;
  uint32_t rebuilds = c->rebuilds;
  bool rebuild = !built || moved2 > R.half_skin2;
  c->rebuild = rebuild ? 1u : 0u;
  c->reorder = (rebuild && (rebuilds % 4u) == 0u) ? 1u : 0u;
  if (rebuild) {
    c->rebuilds = rebuilds + 1u;
    c->built = 1u;
  }
  if (advance_step) c->step = c->step + 1ull;
}

__global__ void __launch_bounds__(kBlock) k_decide(Buffers B) {
  for (uint32_t r = threadIdx.x; r < B.n_replicas; r += blockDim.x) decide(B, r, false);
}

// ---------------------------------------------------------------------------
// Drift: v += a dt/2; x += v dt (or BAOAB); wrap; max moved^2 since build.
// ---------------------------------------------------------------------------

// LANGEVIN: some replica of the launch uses BAOAB (others stay NVE).
template <bool LANGEVIN>
__global__ void __launch_bounds__(kBlock) k_drift(Buffers B) {
  const Range blk = B.atom_blocks[blockIdx.x];
  const Replica& R = B.rep[blk.rep];
  const uint32_t i = blk.first + threadIdx.x;
  unsigned long long moved = 0ull;
  if (i < blk.end) {
    const bool baoab = LANGEVIN && R.thermostat == 1u;
    double n[3] = {0.0, 0.0, 0.0};
    if (baoab) langevin_noise(R.seed, R.index, B.orig[i], B.ctl[blk.rep].step, n);
    double* X[3] = {B.x, B.y, B.z};
    double* V[3] = {B.vx, B.vy, B.vz};
    const double* A[3] = {B.ax, B.ay, B.az};
    const double* XB[3] = {B.xb, B.yb, B.zb};
    double d[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
      double v = __builtin_fma(A[k][i], R.half_dt, V[k][i]);
      double x = X[k][i];
      if (baoab) {
        x = __builtin_fma(v, R.half_dt, x);
        v = __builtin_fma(R.c2, n[k], R.c1 * v);
        x = __builtin_fma(v, R.half_dt, x);
      } else {
        x = __builtin_fma(v, R.dt, x);
      }
      // wrap by one box length (lennard_jones.wgsl:188-196)
      x = x >= R.box[k] ? x - R.box[k] : (x < 0.0 ? x + R.box[k] : x);
      V[k][i] = v;
      X[k][i] = x;
      d[k] = nearest(x - XB[k][i], R.box[k], R.inv_box[k]);
    }
    moved = (unsigned long long)__double_as_longlong(dist2(d[0], d[1], d[2]));
  }
  // one global atomic per workgroup (lennard_jones.wgsl:215-220)
  __shared__ unsigned long long wmax[kBlock / 64];
  moved = wave_max_u64(moved);
  if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = moved;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long m = wmax[0];
#pragma unroll
    for (int w = 1; w < kBlock / 64; w++) m = wmax[w] > m ? wmax[w] : m;
    atomicMax(&B.ctl[blk.rep].max_moved2, m);
  }
  if (last_block(B.tickets + 0)) {
    for (uint32_t r = threadIdx.x; r < B.n_replicas; r += blockDim.x) decide(B, r, true);
    if (threadIdx.x == 0) atomicExch(B.tickets + 0, 0u);
  }
}

// ---------------------------------------------------------------------------
// Cell grid (grid.wgsl:22-148) with a deterministic order: counts, scan,
// scatter in any order, then a rank sort inside each cell by slot index.
// ---------------------------------------------------------------------------

__global__ void __launch_bounds__(kBlock) k_bin(Buffers B) {
  const Range blk = B.atom_blocks[blockIdx.x];
  if (!B.ctl[blk.rep].rebuild) return;
  const Replica& R = B.rep[blk.rep];
  const uint32_t i = blk.first + threadIdx.x;
  if (i >= blk.end) return;
  uint32_t c = R.cell_off + cell_local(R, B.x[i], B.y[i], B.z[i]);
  B.cell_of[i] = c;
  atomicAdd(B.cell_count + c, 1u);
}

// Exclusive scan of counts inside each block of 1024 cells (4 per thread);
// the last workgroup scans the block totals of each rebuilding replica.
__global__ void __launch_bounds__(kBlock) k_scan(Buffers B) {
  const Range blk = B.scan_blocks[blockIdx.x];
  const bool active = B.ctl[blk.rep].rebuild != 0u;
  __shared__ uint32_t wsum[kBlock / 64];
  if (active) {
    const uint32_t base = blk.first + threadIdx.x * 4u;
    uint32_t v[4];
    uint32_t local = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
      uint32_t c = base + (uint32_t)k;
      v[k] = c < blk.end ? B.cell_count[c] : 0u;
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
    for (int k = 0; k < 4; k++) {
      uint32_t c = base + (uint32_t)k;
      if (c < blk.end) B.cell_scan[c] = run;
      run += v[k];
    }
    if (threadIdx.x == 0) B.bsum[blockIdx.x] = total;
  }
  if (last_block(B.tickets + 1)) {
    for (uint32_t r = threadIdx.x; r < B.n_replicas; r += blockDim.x) {
      if (!B.ctl[r].rebuild) continue;
      const Replica& R = B.rep[r];
      uint32_t acc = 0;
      for (uint32_t b = R.scan_block_off; b < R.scan_block_off + R.scan_blocks; b++) {
        B.bprefix[b] = acc;
        acc += load_agent(B.bsum + b);
      }
    }
    if (threadIdx.x == 0) atomicExch(B.tickets + 1, 0u);
  }
}

__device__ __forceinline__ uint32_t cell_first_slot(const Buffers& B, const Replica& R, uint32_t c) {
  uint32_t sb = R.scan_block_off + (c - R.cell_off) / (uint32_t)kScanCells;
  return R.atom_off + B.cell_scan[c] + B.bprefix[sb];
}

__global__ void __launch_bounds__(kBlock) k_scatter(Buffers B) {
  const Range blk = B.atom_blocks[blockIdx.x];
  if (!B.ctl[blk.rep].rebuild) return;
  const Replica& R = B.rep[blk.rep];
  const uint32_t i = blk.first + threadIdx.x;
  if (i >= blk.end) return;
  uint32_t c = B.cell_of[i];
  uint32_t slot = cell_first_slot(B, R, c) + atomicAdd(B.cell_cursor + c, 1u);
  B.unsorted[slot] = i;
}

// One wave64 per cell: every entry's rank is the number of smaller entries
// in the cell, which puts the cell in ascending slot order whatever order
// the scatter produced (the result of grid.wgsl:129-148's insertion sort).
__global__ void __launch_bounds__(kBlock) k_sort(Buffers B) {
  const Range blk = B.sort_blocks[blockIdx.x];
  if (!B.ctl[blk.rep].rebuild) return;
  const Replica& R = B.rep[blk.rep];
  const uint32_t c = blk.first + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (c >= blk.end) return;
  const uint32_t start = cell_first_slot(B, R, c);
  const uint32_t n = B.cell_count[c];
  if (n <= 64u) {
    uint32_t e = lane < (int)n ? B.unsorted[start + lane] : 0xffffffffu;
    uint32_t rank = 0;
    for (uint32_t m = 0; m < n; m++) rank += (__shfl(e, (int)m, 64) < e) ? 1u : 0u;
    if (lane < (int)n) B.sorted[start + rank] = e;
  } else {
    for (uint32_t k = (uint32_t)lane; k < n; k += 64u) {
      uint32_t e = B.unsorted[start + k];
      uint32_t rank = 0;
      for (uint32_t m = 0; m < n; m++) rank += (B.unsorted[start + m] < e) ? 1u : 0u;
      B.sorted[start + rank] = e;
    }
  }
  if (lane == 0) {
    B.cell_start[c] = start;
    B.cell_n[c] = n;
    B.cell_count[c] = 0u;
    B.cell_cursor[c] = 0u;
  }
}

// Slot t takes the atom at sorted[t] (lennard_jones.wgsl:419-452); the copy
// goes to the second state set and `build` copies it back.
__global__ void __launch_bounds__(kBlock) k_reorder(Buffers B) {
  const Range blk = B.atom_blocks[blockIdx.x];
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

// Verlet list of each atom: the 27 cells around its own, dz outermost,
// slots in cell order, every j != i within rc + skin (lennard_jones.wgsl:454-499).
// Entry k of atom i sits at list_off + k * count + (i - atom_off).
__global__ void __launch_bounds__(kBlock) k_build(Buffers B) {
  const Range blk = B.atom_blocks[blockIdx.x];
  const Control& ctl = B.ctl[blk.rep];
  if (!ctl.rebuild) return;
  const bool moved = ctl.reorder != 0u;
  const Replica& R = B.rep[blk.rep];
  const uint32_t i = blk.first + threadIdx.x;
  if (i >= blk.end) return;
  const double* X = moved ? B.x2 : B.x;
  const double* Y = moved ? B.y2 : B.y;
  const double* Z = moved ? B.z2 : B.z;
  const double xi = X[i], yi = Y[i], zi = Z[i];
  uint32_t cg = moved ? B.cell_of2[i] : B.cell_of[i];
  if (moved) {
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
    B.cell_of[i] = cg;
  }
  const int nx = (int)R.cells[0], ny = (int)R.cells[1], nz = (int)R.cells[2];
  const int c = (int)(cg - R.cell_off);
  const int cx = c % nx, cy = (c / nx) % ny, cz = c / (nx * ny);
  const uint32_t cap = R.capacity;
  uint32_t* out = B.nbr + R.list_off + (i - R.atom_off);
  const uint32_t stride = R.count;
  uint32_t found = 0;
  for (int dz = -1; dz <= 1; dz++) {
    const int ncz = (cz + dz + nz) % nz;
    for (int dy = -1; dy <= 1; dy++) {
      const int ncy = (cy + dy + ny) % ny;
      for (int dx = -1; dx <= 1; dx++) {
        const int ncx = (cx + dx + nx) % nx;
        const uint32_t cell = R.cell_off + (uint32_t)((ncz * ny + ncy) * nx + ncx);
        const uint32_t s0 = B.cell_start[cell];
        const uint32_t s1 = s0 + B.cell_n[cell];
        for (uint32_t s = s0; s < s1; s++) {
          const uint32_t j = B.sorted[s];
          if (j == i) continue;
          double ddx = nearest(X[j] - xi, R.box[0], R.inv_box[0]);
          double ddy = nearest(Y[j] - yi, R.box[1], R.inv_box[1]);
          double ddz = nearest(Z[j] - zi, R.box[2], R.inv_box[2]);
          if (dist2(ddx, ddy, ddz) < R.list_r2) {
            if (found < cap) out[(size_t)found * stride] = j;
            found++;
          }
        }
      }
    }
  }
  B.nbr_count[i] = found < cap ? found : cap;
  if (found > cap) atomicMax(&B.ctl[blk.rep].overflow, found);
  B.xb[i] = xi;
  B.yb[i] = yi;
  B.zb[i] = zi;
}

// ---------------------------------------------------------------------------
// Pair forces (lennard_jones.wgsl:501-574)
// ---------------------------------------------------------------------------

struct PairAcc {
  double ax, ay, az;
  double pe, w;
  uint32_t near, within;
};

// R5 (FP64) and R5f (FP32 pair arithmetic on FP64 displacements).
template <int PREC, bool SF, bool ENERGIES>
__device__ __forceinline__ void pair_accumulate(const Replica& R, double dx, double dy, double dz,
                                                double r2, PairAcc& a) {
  if (PREC == 0) {
    const double inv2 = 1.0 / r2;
    const double s6 = inv2 * inv2 * inv2;
    double rf = s6 * __builtin_fma(48.0, s6, -24.0);
    double u = 0.0;
    if (ENERGIES) {
      const double t = 4.0 * s6;
      u = __builtin_fma(t, s6, -t) - R.shift;
    }
    if (SF) {
      const double r = __builtin_sqrt(r2);
      rf = __builtin_fma(-r, R.force_shift, rf);
      if (ENERGIES) u = __builtin_fma(r - R.rc, R.force_shift, u);
    }
    const double fscal = rf * inv2;
    a.ax = __builtin_fma(-dx, fscal, a.ax);
    a.ay = __builtin_fma(-dy, fscal, a.ay);
    a.az = __builtin_fma(-dz, fscal, a.az);
    if (ENERGIES) {
      a.pe += 0.5 * u;
      a.w += 0.5 * rf;
    }
  } else {
    const float q = (float)r2;
    const float inv2 = __builtin_amdgcn_rcpf(q);
    const float s6 = inv2 * inv2 * inv2;
    float rf = s6 * __builtin_fmaf(48.0f, s6, -24.0f);
    float u = 0.0f;
    if (ENERGIES) {
      const float t = 4.0f * s6;
      u = __builtin_fmaf(t, s6, -t) - (float)R.shift;
    }
    if (SF) {
      const float r = __builtin_sqrtf(q);
      const float fs = (float)R.force_shift;
      rf = __builtin_fmaf(-r, fs, rf);
      if (ENERGIES) u = __builtin_fmaf(r - (float)R.rc, fs, u);
    }
    const float fscal = rf * inv2;
    a.ax -= (double)((float)dx * fscal);
    a.ay -= (double)((float)dy * fscal);
    a.az -= (double)((float)dz * fscal);
    if (ENERGIES) {
      a.pe += 0.5 * (double)u;
      a.w += 0.5 * (double)rf;
    }
  }
  if (ENERGIES) {
    a.near += r2 < kNearR2 ? 1u : 0u;
    a.within += 1u;
  }
}

template <int PREC, bool SF, bool ENERGIES>
__device__ __forceinline__ PairAcc interactions(const Buffers& B, const Replica& R, uint32_t i) {
  const double xi = B.x[i], yi = B.y[i], zi = B.z[i];
  const uint32_t cnt = B.nbr_count[i];
  const uint32_t* lst = B.nbr + R.list_off + (i - R.atom_off);
  const uint32_t stride = R.count;
  PairAcc a = {0.0, 0.0, 0.0, 0.0, 0.0, 0u, 0u};
  for (uint32_t k = 0; k < cnt; k++) {
    const uint32_t j = lst[(size_t)k * stride];
    const double dx = nearest(B.x[j] - xi, R.box[0], R.inv_box[0]);
    const double dy = nearest(B.y[j] - yi, R.box[1], R.inv_box[1]);
    const double dz = nearest(B.z[j] - zi, R.box[2], R.inv_box[2]);
    const double r2 = dist2(dx, dy, dz);
    if (r2 < R.rc2) pair_accumulate<PREC, SF, ENERGIES>(R, dx, dy, dz, r2, a);
  }
  return a;
}

// New accelerations and (kick) the second half kick v += a dt/2.
template <int PREC, bool SF>
__global__ void __launch_bounds__(kBlock) k_force(Buffers B, int kick) {
  const Range blk = B.atom_blocks[blockIdx.x];
  const Replica& R = B.rep[blk.rep];
  const uint32_t i = blk.first + threadIdx.x;
  if (i >= blk.end) return;
  const PairAcc a = interactions<PREC, SF, false>(B, R, i);
  B.ax[i] = a.ax;
  B.ay[i] = a.ay;
  B.az[i] = a.az;
  if (kick) {
    B.vx[i] = __builtin_fma(a.ax, R.half_dt, B.vx[i]);
    B.vy[i] = __builtin_fma(a.ay, R.half_dt, B.vy[i]);
    B.vz[i] = __builtin_fma(a.az, R.half_dt, B.vz[i]);
  }
}

// ---------------------------------------------------------------------------
// Measurement (lennard_jones.wgsl:576-626): per atom KE, PE, W, momentum,
// condensed, pairs within rc, each to 64-bit fixed point; a workgroup sums
// its atoms by wave shuffles and LDS and adds once per slot with a 64-bit
// atomic (fixed.wgsl:71-105).
// ---------------------------------------------------------------------------

__device__ __forceinline__ void block_add_slots(long long v[kSlots], long long* dst) {
  __shared__ long long part[kBlock / 64][kSlots];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < kSlots; k++) {
    long long s = wave_sum_i64(v[k]);
    if (lane == 0) part[wave][k] = s;
  }
  __syncthreads();
  if (threadIdx.x < kSlots) {
    long long s = 0;
#pragma unroll
    for (int w = 0; w < kBlock / 64; w++) s += part[w][threadIdx.x];
    if (s != 0) atomicAdd((unsigned long long*)(dst + threadIdx.x), (unsigned long long)s);
  }
}

template <int PREC, bool SF>
__global__ void __launch_bounds__(kBlock) k_measure(Buffers B) {
  const Range blk = B.atom_blocks[blockIdx.x];
  const Replica& R = B.rep[blk.rep];
  const uint32_t i = blk.first + threadIdx.x;
  long long v[kSlots] = {0, 0, 0, 0, 0, 0, 0, 0};
  if (i < blk.end) {
    const PairAcc a = interactions<PREC, SF, true>(B, R, i);
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
  block_add_slots(v, B.slots + (size_t)blk.rep * kSlots);
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

__global__ void k_zero_i64(long long* p, uint32_t n) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) p[t] = 0;
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------

static inline dim3 grid1(uint32_t blocks) { return dim3(blocks, 1, 1); }

hipError_t launch_decide(const Buffers& B, hipStream_t s) {
  hipLaunchKernelGGL(k_decide, dim3(1), dim3(kBlock), 0, s, B);
  return hipGetLastError();
}

hipError_t launch_drift(const Buffers& B, bool langevin, hipStream_t s) {
  if (!B.n_atom_blocks) return hipSuccess;
  if (langevin)
    hipLaunchKernelGGL(k_drift<true>, grid1(B.n_atom_blocks), dim3(kBlock), 0, s, B);
  else
    hipLaunchKernelGGL(k_drift<false>, grid1(B.n_atom_blocks), dim3(kBlock), 0, s, B);
  return hipGetLastError();
}

hipError_t launch_rebuild(const Buffers& B, hipStream_t s) {
  if (!B.n_atom_blocks) return hipSuccess;
  hipLaunchKernelGGL(k_bin, grid1(B.n_atom_blocks), dim3(kBlock), 0, s, B);
  hipLaunchKernelGGL(k_scan, grid1(B.n_scan_blocks), dim3(kBlock), 0, s, B);
  hipLaunchKernelGGL(k_scatter, grid1(B.n_atom_blocks), dim3(kBlock), 0, s, B);
  hipLaunchKernelGGL(k_sort, grid1(B.n_sort_blocks), dim3(kBlock), 0, s, B);
  hipLaunchKernelGGL(k_reorder, grid1(B.n_atom_blocks), dim3(kBlock), 0, s, B);
  hipLaunchKernelGGL(k_build, grid1(B.n_atom_blocks), dim3(kBlock), 0, s, B);
  return hipGetLastError();
}

hipError_t launch_force(const Buffers& B, int precision, bool shifted, bool kick, hipStream_t s) {
  if (!B.n_atom_blocks) return hipSuccess;
  const int k = kick ? 1 : 0;
  if (precision == 0) {
    if (shifted)
      hipLaunchKernelGGL((k_force<0, true>), grid1(B.n_atom_blocks), dim3(kBlock), 0, s, B, k);
    else
      hipLaunchKernelGGL((k_force<0, false>), grid1(B.n_atom_blocks), dim3(kBlock), 0, s, B, k);
  } else {
    if (shifted)
      hipLaunchKernelGGL((k_force<1, true>), grid1(B.n_atom_blocks), dim3(kBlock), 0, s, B, k);
    else
      hipLaunchKernelGGL((k_force<1, false>), grid1(B.n_atom_blocks), dim3(kBlock), 0, s, B, k);
  }
  return hipGetLastError();
}

hipError_t launch_measure(const Buffers& B, int precision, bool shifted, hipStream_t s) {
  const uint32_t zb = (B.n_replicas * kSlots + kBlock - 1) / kBlock;
  hipLaunchKernelGGL(k_zero_slots, grid1(zb), dim3(kBlock), 0, s, B);
  if (!B.n_atom_blocks) return hipGetLastError();
  if (precision == 0) {
    if (shifted)
      hipLaunchKernelGGL((k_measure<0, true>), grid1(B.n_atom_blocks), dim3(kBlock), 0, s, B);
    else
      hipLaunchKernelGGL((k_measure<0, false>), grid1(B.n_atom_blocks), dim3(kBlock), 0, s, B);
  } else {
    if (shifted)
      hipLaunchKernelGGL((k_measure<1, true>), grid1(B.n_atom_blocks), dim3(kBlock), 0, s, B);
    else
      hipLaunchKernelGGL((k_measure<1, false>), grid1(B.n_atom_blocks), dim3(kBlock), 0, s, B);
  }
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

}  // namespace vlj
