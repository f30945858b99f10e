// Device-side layout and arithmetic of the Lennard-Jones path.
//
// One handle holds one or more independent periodic systems ("replicas").
// Every per-atom array is one flat FP64/u32 array over all replicas, replica r
// owning atoms [atom_off, atom_off + count) in slot order (slots are kept in
// cell order, lennard_jones.wgsl:31-33). Launches cover all replicas at once:
// each workgroup reads a Range (replica, first, end) from a table, so a block
// never straddles two replicas and the replica's parameters stay uniform
// (scalar registers) inside it.
//
// All kernels are built with -ffp-contract=off; every fused multiply-add is
// written as fma() so the operation sequence is fixed and reproducible.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace vlj {

constexpr int kBlock = 256;       // threads per atom block (4 wave64)
#ifndef VLJ_BRICK_THREADS
#define VLJ_BRICK_THREADS 512
#endif
constexpr int kBrickThreads = VLJ_BRICK_THREADS;  // threads of a brick workgroup
static_assert(kBrickThreads % 64 == 0 && kBrickThreads <= 1024, "whole waves");
constexpr int kScanCells = 1024;  // cells per scan block (256 threads x 4) of a dense replica
constexpr int kScanCellsSparse = 4096;  // ... of a dilute one (256 threads x 16): fewer block sums
// Cells per sort block of a dense replica (a wave64 per cell, each wave
// taking the block's cells in turn). The sort launch returns at once on the
// steps without a rebuild, and fewer workgroups cost less there: against 4
// cells, the 512-replica batch +1.7%, FCC-1M (13,718 blocks at 4) +1.2%.
constexpr int kSortCells = 16;
constexpr int kSortCellsSparse = 256;  // ... of a replica with fewer than kSortDense atoms per cell
constexpr double kSortDense = 6.0;
constexpr uint32_t kSmallCell = 8;  // a cell this small is ranked by one thread
constexpr uint32_t kProfileSums = 8;       // sums per profile slab (lennard_jones.wgsl:644)
constexpr uint32_t kMaxProfileBins = 1u << 20;  // profile slabs per replica, at most
constexpr uint32_t kHexSums = 18;   // per slab (hot, cold): count, momentum x3, v^2, f.v, force x3
constexpr uint32_t kHexState = 18;  // per slab: com x3, mean force x3, s - 1, c_velocity, c_force
constexpr int kBrickMax = 8;      // home cells of a brick along x and along y, at most
constexpr int kStageCells = 3 * (kBrickMax + 2) * (kBrickMax + 2);  // staged cells, at most
constexpr int kHomeCells = kBrickMax * kBrickMax;                    // home cells, at most
static_assert(kHomeCells <= 64, "one lane of a wave per home cell");
constexpr uint32_t kPartsGrid = 128;  // workgroups of a launch over the parts of cut bricks
constexpr uint32_t kCutGroup = 128;   // bricks per group with a count of cut bricks
#ifndef VLJ_STAGE_MAX
#define VLJ_STAGE_MAX 3200
#endif
#ifndef VLJ_SPARSE_STAGE_MAX
#define VLJ_SPARSE_STAGE_MAX 1024
#endif
// Atoms a brick workgroup stages in LDS. With the sentinel the staged arrays
// hold kStageMax + 64 doubles each (a multiple of 64, for ds_read2st64); at
// 3200 the stage takes ~79 KB, two 512-thread workgroups per CU (16 waves).
// Measured on the 1M-atom liquid: 256 threads / 2048 atoms (3 workgroups
// per CU, bricks of ~170 atoms leaving a wave in four idle) 0.359 ms per
// step; 512 / 3008: 0.338; 512 / 3200 with bricks planned to 90%: 0.332.
constexpr int kStageMax = VLJ_STAGE_MAX;
static_assert(kStageMax % 64 == 0, "staged arrays are whole rows of 64");

// A brick layout: threads of its workgroup and atoms it stages. Dense bricks
// (liquid layers) use 512 threads and 3200 atoms, two workgroups per CU.
// Sparse bricks, whose neighbourhood holds at most ~900 atoms (a gas: an
// 8 x 8 brick of a replica at rho 0.1 stages ~650 and owns ~140), use 128
// threads and 1024 atoms: five workgroups per CU instead of two, each with
// most of its lanes busy. The brick table holds the dense bricks first, then
// the sparse ones; each layout has its own launches.
template <int THREADS, int STAGE_MAX, bool SPARSE>
struct Layout {
  static constexpr int kThreads = THREADS;
  static constexpr int kStageMax = STAGE_MAX;
  static constexpr uint32_t kListPad = (uint32_t)STAGE_MAX;  // a list pad: the sentinel's staged index
  static constexpr bool kSparse = SPARSE;
  static constexpr int kCellsPerThread = (kStageCells + THREADS - 1) / THREADS;  // staged cells per thread
  static_assert(THREADS % 64 == 0 && THREADS <= 1024 && THREADS >= (int)kCutGroup, "whole waves");
  static_assert(STAGE_MAX % 64 == 0 && STAGE_MAX < 65535, "whole rows of 64; 16-bit staged indices");
};
using Dense = Layout<kBrickThreads, kStageMax, false>;
using Sparse = Layout<128, VLJ_SPARSE_STAGE_MAX, true>;
constexpr int kListStride = 64;   // lists are transposed per block of 64 atoms (one wave)
constexpr int kSlots = 8;         // fixed-point slots per replica
constexpr uint32_t kFixedSumBlocks = 1024;  // workgroups of the no-allocation fixed-point sum, at most
constexpr uint64_t kFixedSumMaxCount = 1ull << 36;  // values one no-allocation sum takes, at most
constexpr double kNearR2 = 2.25;  // condensed: neighbours within 1.5 sigma (lennard_jones.wgsl:417)

// Parameters of one replica (lennard_jones.rs:282-313, :443-560).
struct Replica {
  double box[3];
  double half_box[3];
  double inv_box[3];
  float cell_size[3];
  uint32_t cells[3];
  uint32_t atom_off, count;
  uint32_t cell_off, total_cells;
  uint32_t scan_block_off, scan_blocks, scan_cells;  // scan blocks of scan_cells cells
  uint32_t atom_block_off, atom_blocks;  // this replica's entries of the atom-block table
  uint32_t capacity;
  uint32_t thermostat;  // 0 none, 1 Langevin (BAOAB), 2 Nose-Hoover
  uint64_t list_off;    // first neighbour entry of the replica
  double rc, rc2, list_r2, half_skin2;
  double shift, force_shift, force_shift24;  // force_shift24 = force_shift / 24
  double dt, half_dt, c1, c2;
  double energy_fixed, momentum_fixed;  // scale * 2^32
  uint64_t seed;
  uint32_t near_count;
  uint32_t index;
  double nh_gkt, nh_q, nh_period;  // Nose-Hoover g k T, mass Q and period (thermostat 2)
  double hex_slabs[4];  // heat exchange: hot [0, 1), cold [2, 3) in z
  double hex_energy;    // energy moved per step
  uint32_t hex_on;      // heat exchange on
  uint32_t hc_cap;      // heat-current samples the series holds (0: none)
  double bin_width;    // profile slab height, L_z / prof_bins
  uint32_t prof_bins;  // profile slabs along z (0: no profiles)
  uint32_t prof_off;   // first of the replica's prof_bins * 8 + 3 profile sums
  uint32_t drawn;      // render w of velocities: 0 neighbours within 1.5, 1 local temperature
};

// Control words of one replica (lennard_jones.wgsl:106-110: MAX_DISPLACEMENT,
// REBUILDS, OVERFLOW, NEVER_BUILT), plus the step counter of the noise.
struct Control {
  unsigned long long max_moved2;  // FP64 bits; non-negative doubles order as integers
  unsigned long long step;        // steps taken
  uint32_t rebuild;               // this step rebuilds the lists
  uint32_t reorder;               // ... and moves atoms into cell order
  uint32_t rebuilds;
  uint32_t overflow;
  uint32_t built;
  uint32_t stage_overflow;        // brick list builds whose neighbourhood outgrew LDS
  uint32_t stage_global;          // ... and parts of one home cell that read global memory
  uint32_t pad_;
  double xi, eta, sm1;            // Nose-Hoover friction, its integral, velocity scale - 1
};

struct Range {
  uint32_t rep, first, end, pad;
};

// A brick: gx x gy x gz home cells from cell0 (its lowest x, y and z, an
// index within the replica's grid) of replica rep; gz is 1 or 2 cell layers.
// Its workgroup stages the (gz + 2) x (gy + 2) x (gx + 2) cells around it.
struct Brick {
  uint32_t rep, cell0, gx, gy, gz;
};
__host__ __device__ __forceinline__ uint32_t home_cells(const Brick& g) { return g.gx * g.gy * g.gz; }

// Everything a kernel touches, passed by value.
struct Buffers {
  const Replica* rep;
  Control* ctl;
  const Range* atom_blocks;
  const Range* scan_blocks;
  const Range* sort_blocks;
  const Brick* bricks;        // list build, forces and measurements: n_dense dense, then sparse
  uint32_t* cut;              // per brick: parts it was cut into at its replica's last list build
  uint16_t* parts;            // per brick: kHomeCells packed parts
  uint32_t* cut_groups;       // per group of kCutGroup bricks: bricks cut
  uint32_t n_atom_blocks, n_scan_blocks, n_sort_blocks, n_bricks, n_replicas, n_dense;
  // 1: a brick whose neighbourhood outgrows LDS is cut into parts, which the
  // PARTS launches after each brick launch take; 0: no PARTS launches, and
  // such a brick runs whole, its positions read through the cell tables
  uint32_t parts_on;
  unsigned long long* bmax;   // per atom block: max moved^2 (FP64 bits)
  long long* kepart;          // per atom block: kinetic energy in fixed point (Nose-Hoover)
  long long* hexpart;         // per atom block: kHexSums heat-exchange sums
  double* hexstate;           // per replica: kHexState scales and corrections
  long long* hc;              // heat-current series of replica 0: 3 sums per sample
  double *x, *y, *z, *vx, *vy, *vz, *ax, *ay, *az;           // state, slot order
  double *x2, *y2, *z2, *vx2, *vy2, *vz2, *ax2, *ay2, *az2;  // reorder target
  double *xb, *yb, *zb;                                      // position at last list build
  uint32_t *orig, *orig2;                                    // original index of each slot
  uint32_t *cell_of, *cell_of2;                              // global cell of each slot
  uint32_t *unsorted, *sorted;                               // slots in cell order
  uint32_t *cell_count, *cell_cursor, *cell_scan, *cell_start, *cell_n;
  uint32_t *bsum, *bprefix;                                  // per scan block
  uint32_t* nbr_count;                                       // list length of each slot
  uint16_t* nbr;                                             // lists of staged indices
  long long* slots;                                          // kSlots per replica
  long long* profile;                                        // profile sums of all replicas
  float *render_pos, *render_vel;                            // float4 per slot, for drawing
  uint32_t* tickets;                                         // last-block counters
  // flags[1]: bricks cut into parts; flags[0] and flags[2]: replicas
  // rebuilding on a step, one word for each of two consecutive steps of the
  // drift-ahead form. Kernels that have nothing to do without them return
  // after one load.
  uint32_t* flags;
  // this step's word (zeroed by the drift, or by the decision of the step
  // before, and counted by the decision) and the next step's
  uint32_t *rebuilding, *rebuilding_next;
  // Drift-ahead form: kAheadSlots per replica of largest moved^2 since the
  // list build, taken by the force launch into amax_next (atomicMax) for
  // the next step, whose k_bin decides from them (bin_decides) in place of
  // a k_decide launch; each force launch zeroes the set the bin just read.
  unsigned long long *amax, *amax_next;
  uint32_t bin_decides;
};
constexpr uint32_t kAheadSlots = 64;  // one wave's worth

// Workgroups are dispatched to the 8 XCDs round-robin (workgroup b runs on
// XCD b % 8). Each table (atom blocks, sort blocks, bricks) is dealt to the
// XCDs in chunks of kChunk consecutive entries, chunk c to XCD c % 8:
// neighbouring entries (adjacent bricks, consecutive atoms) share an XCD and
// its L2, and any stretch longer than 8 kChunk entries -- one replica's
// rebuild, the dense quarter of a slab -- is spread over all eight XCDs. The
// grid is rounded up to whole rounds of 8 kChunk workgroups; workgroups past
// the table's end return at once. Chunks of 4 (measured against 2, 8, 16
// and 32): the 1M-atom force kernel 141 us at 16, 136 us at 4 and 8; on the
// present step 2, 4 and 8 alike (FP32 form 5.51e9-5.53e9 atom-steps/s).
#ifndef VLJ_XCD_CHUNK
#define VLJ_XCD_CHUNK 4
#endif
constexpr uint32_t kChunk = VLJ_XCD_CHUNK;

__host__ __device__ __forceinline__ uint32_t dealt_grid(uint32_t n) {
  return (n + 8u * kChunk - 1u) / (8u * kChunk) * (8u * kChunk);
}

__host__ __device__ __forceinline__ uint32_t dealt_index(uint32_t b) {
  const uint32_t xcd = b % 8u, j = b / 8u;
  return ((j / kChunk) * 8u + xcd) * kChunk + j % kChunk;
}

// A list entry is 8 e, where e is the neighbour's index in its brick's
// staging order (the staged cells in (z, y, x) order, each cell's slots in
// cell order), 16 bits: the byte offset of its position in each of the
// staged coordinate arrays, so the force loop addresses LDS with the entry
// itself (kStageMax * 8 < 2^16).
// Lists are stored in octets of 8 entries (16 bytes), transposed within
// blocks of 64 consecutive atoms (one wave): octet o of the replica's atom li
// is at (u16 units)
//   list_off + (li / 64) * octets * 512 + o * 512 + (li % 64) * 8,
// octets = ceil(capacity / 8). A wave reads octet o of its 64 atoms as 1 KB
// of contiguous memory; all of a wave's lists lie in octets KB. The last
// octet of a list is filled up with the entry of the layout's kListPad, the staged index of a
// sentinel position far outside every box, so the force loop can take whole
// octets without a per-entry bound.
constexpr uint32_t kEntryShift = 3;
static_assert((kStageMax + 1u) << kEntryShift <= 0x10000u, "list entries are 16 bits");
__host__ __device__ __forceinline__ uint32_t entry_of(uint32_t e) { return e << kEntryShift; }
__host__ __device__ __forceinline__ uint32_t staged_of(uint32_t entry) { return entry >> kEntryShift; }
// A brick (or part) whose neighbourhood does not fit in LDS reads positions
// through the cell tables instead, and its list entries are the staged
// indices e themselves, so that neighbourhoods of up to 65535 atoms keep
// exact lists. Every kernel that reads a list knows which case it is in
// (the same staged total, from the same cell counts).
__host__ __device__ __forceinline__ uint32_t list_index(uint32_t entry, bool staged) {
  return staged ? staged_of(entry) : entry;
}
__host__ __device__ __forceinline__ uint32_t list_entry_of(uint32_t e, bool staged) {
  return staged ? entry_of(e) : e;
}
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
constexpr double kFar = 1.0e18;  // sentinel coordinate: r^2 ~ 3e36, beyond any cutoff

__host__ __device__ __forceinline__ uint32_t list_octets(uint32_t capacity) { return (capacity + 7u) / 8u; }

__host__ __device__ __forceinline__ uint64_t list_base(const Replica& R, uint32_t li) {
  return R.list_off + (uint64_t)(li / kListStride) * list_octets(R.capacity) * (8u * kListStride) +
         (uint64_t)(li % kListStride) * 8u;
}

// u16 offset of entry k from list_base.
__host__ __device__ __forceinline__ uint64_t list_entry(uint32_t k) {
  return (uint64_t)(k / 8u) * (8u * kListStride) + (k % 8u);
}

// R1: back into the box by one length (positions are at most about a skin
// outside it when the lists are rebuilt; lennard_jones.wgsl:188-196).
__host__ __device__ __forceinline__ double wrap_one(double x, double side) {
  return x >= side ? x - side : (x < 0.0 ? x + side : x);
}

__device__ __forceinline__ double dist2(double dx, double dy, double dz) {
  return __builtin_fma(dz, dz, __builtin_fma(dy, dy, dx * dx));
}

// ---- R2: cell of a position, f32 rule of grid.wgsl:11-20.
__device__ __forceinline__ uint32_t cell_axis(double x, float cs, uint32_t n) {
  float q = (float)x / cs;  // IEEE f32 division (correctly rounded by default)
  int c = (int)floorf(q);
  int m = (int)n - 1;
  c = c < 0 ? 0 : (c > m ? m : c);
  return (uint32_t)c;
}

__device__ __forceinline__ uint32_t cell_local(const Replica& R, double x, double y, double z) {
  uint32_t cx = cell_axis(x, R.cell_size[0], R.cells[0]);
  uint32_t cy = cell_axis(y, R.cell_size[1], R.cells[1]);
  uint32_t cz = cell_axis(z, R.cell_size[2], R.cells[2]);
  return (cz * R.cells[1] + cy) * R.cells[0] + cx;
}

// ---- Counter-based noise: Philox4x32-10 (Salmon et al., SC'11).
__device__ __forceinline__ void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int round = 0; round < 10; round++) {
    uint32_t lo0 = 0xD2511F53u * c[0];
    uint32_t hi0 = __umulhi(0xD2511F53u, c[0]);
    uint32_t lo1 = 0xCD9E8D57u * c[2];
    uint32_t hi1 = __umulhi(0xCD9E8D57u, c[2]);
    uint32_t n0 = hi1 ^ c[1] ^ k0;
    uint32_t n2 = hi0 ^ c[3] ^ k1;
    c[0] = n0;
    c[1] = lo1;
    c[2] = n2;
    c[3] = lo0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
}

// Natural log from +, *, fma and / only (range reduction to [sqrt(1/2),
// sqrt(2)), 2 atanh((m-1)/(m+1)) by its odd series, ln 2 split in two after
// Cody and Waite). Positive normal arguments.
__device__ __forceinline__ double log_basic(double x) {
  unsigned long long b = (unsigned long long)__double_as_longlong(x);
  int e = (int)((b >> 52) & 0x7ff) - 1023;
  b = (b & 0x000fffffffffffffull) | 0x3ff0000000000000ull;
  double m = __longlong_as_double((long long)b);
  if (m > 1.4142135623730951) {
    m = m * 0.5;
    e += 1;
  }
  double f = m - 1.0;
  double s = f / (2.0 + f);
  double z = s * s;
  double p = 1.0 / 21.0;
  p = __builtin_fma(p, z, 1.0 / 19.0);
  p = __builtin_fma(p, z, 1.0 / 17.0);
  p = __builtin_fma(p, z, 1.0 / 15.0);
  p = __builtin_fma(p, z, 1.0 / 13.0);
  p = __builtin_fma(p, z, 1.0 / 11.0);
  p = __builtin_fma(p, z, 1.0 / 9.0);
  p = __builtin_fma(p, z, 1.0 / 7.0);
  p = __builtin_fma(p, z, 1.0 / 5.0);
  p = __builtin_fma(p, z, 1.0 / 3.0);
  double tail = 2.0 * (s * z * p);
  double de = (double)e;
  const double ln2_hi = 6.93147180369123816490e-01;
  const double ln2_lo = 1.90821492927058770002e-10;
  return __builtin_fma(de, ln2_hi, 2.0 * s + __builtin_fma(de, ln2_lo, tail));
}

// sin and cos of 2 pi u, u in [0, 1]: exact removal of the nearest quarter
// turn, Taylor series on [-pi/4, pi/4], quadrant rotation.
__device__ __forceinline__ void sincos_turns(double u, double* s_out, double* c_out) {
  double q = __builtin_rint(4.0 * u);
  double t = u - 0.25 * q;
  double x = t * 6.283185307179586;
  double z = x * x;
  double sp = -1.0 / 355687428096000.0;
  sp = __builtin_fma(sp, z, 1.0 / 1307674368000.0);
  sp = __builtin_fma(sp, z, -1.0 / 6227020800.0);
  sp = __builtin_fma(sp, z, 1.0 / 39916800.0);
  sp = __builtin_fma(sp, z, -1.0 / 362880.0);
  sp = __builtin_fma(sp, z, 1.0 / 5040.0);
  sp = __builtin_fma(sp, z, -1.0 / 120.0);
  sp = __builtin_fma(sp, z, 1.0 / 6.0);
  double sn = __builtin_fma(-x * z, sp, x);
  double cp = 1.0 / 6402373705728000.0;
  cp = __builtin_fma(cp, z, -1.0 / 20922789888000.0);
  cp = __builtin_fma(cp, z, 1.0 / 87178291200.0);
  cp = __builtin_fma(cp, z, -1.0 / 479001600.0);
  cp = __builtin_fma(cp, z, 1.0 / 3628800.0);
  cp = __builtin_fma(cp, z, -1.0 / 40320.0);
  cp = __builtin_fma(cp, z, 1.0 / 720.0);
  cp = __builtin_fma(cp, z, -1.0 / 24.0);
  cp = __builtin_fma(cp, z, 0.5);
  double cs = __builtin_fma(-z, cp, 1.0);
  int quadrant = ((int)q) & 3;
  double so = quadrant == 0 ? sn : (quadrant == 1 ? cs : (quadrant == 2 ? -sn : -cs));
  double co = quadrant == 0 ? cs : (quadrant == 1 ? -sn : (quadrant == 2 ? -cs : sn));
  *s_out = so;
  *c_out = co;
}

// Three standard normals for (seed, stream, original atom, step): counter
// (step lo, step hi, atom, stream), key (seed lo, seed hi); uniforms
// (w + 0.5) 2^-32; Box-Muller on (u0, u1) and the cosine branch on (u2, u3).
// The thermostat uses stream 0: replicas are told apart by their seeds, so a
// replica's noise does not depend on its place in a batch.
__device__ __forceinline__ void langevin_noise(uint64_t seed, uint32_t stream, uint32_t atom,
                                               unsigned long long step, double n[3]) {
  uint32_t c[4] = {(uint32_t)step, (uint32_t)(step >> 32), atom, stream};
  philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
  double u0 = ((double)c[0] + 0.5) * 2.3283064365386963e-10;
  double u1 = ((double)c[1] + 0.5) * 2.3283064365386963e-10;
  double u2 = ((double)c[2] + 0.5) * 2.3283064365386963e-10;
  double u3 = ((double)c[3] + 0.5) * 2.3283064365386963e-10;
  double r0 = __builtin_sqrt(-2.0 * log_basic(u0));
  double s0, c0;
  sincos_turns(u1, &s0, &c0);
  double r1 = __builtin_sqrt(-2.0 * log_basic(u2));
  double s1, c1;
  sincos_turns(u3, &s1, &c1);
  n[0] = r0 * c0;
  n[1] = r0 * s0;
  n[2] = r1 * c1;
}

// Round-ties-even of x * fixed_scale as a 64-bit integer (fixed.wgsl:16-55
// in FP64: the scale is a power of two, so the product is exact).
__device__ __forceinline__ long long to_fixed(double x, double fixed_scale) {
  return (long long)__builtin_rint(x * fixed_scale);
}

}  // namespace vlj
