/*
 * CPU f64 oracle for the Lennard-Jones hot path. TEST INFRASTRUCTURE ONLY:
 * it produces expected answers and is never the thing measured or shipped.
 * See oracle_lj.h.
 *
 * Built with -ffp-contract=off: every fused multiply-add below is written as
 * fma() on purpose, and nothing else is fused, so the HIP kernels (built the
 * same way) can follow the identical sequence of IEEE operations.
 *
 * Arithmetic rules shared with the HIP kernels:
 *  R1 positions are wrapped into the box (x >= L -> x - L, x < 0 -> x + L)
 *     when the lists are rebuilt, not every step: between rebuilds they move
 *     continuously, so a pair's periodic image fixed at the rebuild stays
 *     right until the next one (the reference wraps every step,
 *     lennard_jones.wgsl:188-196; positions read back are wrapped).
 *     r2 = fma(dz, dz, fma(dy, dy, dx * dx)).
 *  R2 cell of a position: c = clamp(floor((float)x / cs_f32), 0, n - 1) in
 *     f32 per axis, index (cz ny + cy) nx + cx (grid.wgsl:11-20; the
 *     reference bins the f32 hi part, which is (float)x).
 *  R3 cell order: slots ascending within each cell (grid.wgsl:129-148).
 *  R4 list: for dz, dy, dx in -1..1 (dz outermost) the cell (c + d + n) % n,
 *     slots in cell order, j != i with r2 < (rc + skin)^2 in FP64
 *     (lennard_jones.wgsl:454-499). The displacement is taken to the image
 *     of j's cell seen from i's cell, d = (xj + s L) - xi with s = -1 where
 *     c + d wraps below 0, +1 where it wraps past n - 1, else 0; for n >= 3
 *     (n >= 2 is enough: see vo_derive_params) this is the nearest image of
 *     every pair within one cell width
 *     (the reference's nearest(), lennard_jones.wgsl:198-202). Each entry
 *     keeps its image s.
 *  R5 pair, FP64, over the list in order with d = (xj + s L) - xi and the
 *     entry's image s: r2 < rc^2; inv2 = 1 / r2; s6 = inv2 inv2 inv2;
 *     g = s6 fma(s6, 2, -1) (that is r.F / 24); shifted force:
 *     g = fma(-sqrt(r2), fs / 24, g); fscal = g inv2; G = fma(d, fscal, G)
 *     over the list, and a = -24 G at the end (lennard_jones.wgsl:501-545,
 *     with the factor 24 of r.F = 24 s6 (2 s6 - 1) taken out of the sum).
 *  R5f pair, FP32 forces: cutoff test and d as R5 in FP64, then in f32:
 *     q = (float)r2, inv2 = 1/q, s6, g, fscal as R5; f = fmaf((float)d,
 *     fscal, f) in f32 over each octet of the list (entries 8m .. 8m + 7,
 *     pairs past the cutoff adding nothing), each octet's f added to G in
 *     FP64, a = -24 G.
 *  R9 energies: rf = s6 fma(48, s6, -24); u = fma(4 s6, s6, -4 s6) - shift;
 *     shifted force: r = sqrt(r2), rf = fma(-r, fs, rf), u = fma(r - rc, fs, u).
 *  R6 drift (velocity Verlet): v = fma(a, dt/2, v); x = fma(v, dt, x);
 *     moved^2 = r2 of x - x_build (lennard_jones.wgsl:204-276).
 *  R7 kick: v = fma(a, dt/2, v) (lennard_jones.wgsl:555-574).
 *  R8 rebuild iff never built or max moved^2 > (skin/2)^2, and then move the
 *     atoms into cell order (lennard_jones.wgsl:278-307 reorders on every 4th
 *     rebuild only; here every rebuild does, so each cell owns one contiguous
 *     range of slots).
 *  R9 measurement: per atom ke = 0.5 fma(vz, vz, fma(vy, vy, vx vx)), pe and
 *     w summed in list order as pe += 0.5 u, w += 0.5 rf; each per-atom value
 *     goes to 64-bit fixed point as llrint(value * scale * 2^32) and the
 *     integers are summed (lennard_jones.wgsl:576-626, fixed.wgsl:1-15).
 *  R12 Langevin BAOAB: v = fma(a, dt/2, v); x = fma(v, dt/2, x);
 *     v = fma(c2, n, c1 v); x = fma(v, dt/2, x) (lennard_jones.wgsl:249-266),
 *     with normals from Philox4x32-10 and Box-Muller (vo_langevin_noise).
 */
#include "oracle_lj.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define TWO32 4294967296.0
#define VO_PI 3.14159265358979323846 /* the f64 nearest pi (std::f64::consts::PI) */

struct vo_sim {
    vo_params p;
    double *x[3], *v[3], *a[3], *xb[3];
    uint32_t *orig, *cell_of, *sorted, *cell_start, *cell_count;
    uint32_t *nbr_count, *nbr;
    uint8_t *nbr_image; /* (sx + 1) + 3 (sy + 1) + 9 (sz + 1) per entry */
    double *pe, *w;
    uint32_t *near, *within;
    uint32_t rebuilds, overflow;
    int built, primed;
    uint32_t prof_bins;
    uint64_t prof_samples;
    int64_t *prof;
    double max_moved2;
    uint64_t steps;
    double xi, eta, nh_sm1; /* Nose-Hoover friction, its integral, last scale - 1 */
    int hex_on;              /* heat exchange (R15) */
    double hex_slabs[4], hex_energy;
    uint32_t hc_cap, hc_samples; /* heat-current series (R14) */
    int64_t *hc;
};

/* ------------------------------------------------------------------ */
/* Fixed point                                                         */
/* ------------------------------------------------------------------ */

/* fixed.rs:9-41 / fixed.wgsl:20-55: (low, high) words of round-ties-even(x 2^32),
 * by integer operations on the float's bits. Valid for |x| < 2^31. */
void vo_fixed_split(float x, uint32_t *low_out, uint32_t *high_out)
{
    uint32_t bits;
    memcpy(&bits, &x, 4);
    int32_t exponent = (int32_t)((bits >> 23) & 0xffu) - 127;
    if (exponent < -34) {
        *low_out = 0;
        *high_out = 0;
        return;
    }
    uint32_t mantissa = (bits & 0x7fffffu) | 0x800000u;
    int32_t shift = exponent + 9;
    uint32_t low = 0, high = 0;
    if (shift >= 32) {
        high = mantissa << (shift - 32);
    } else if (shift > 0) {
        low = mantissa << shift;
        high = mantissa >> (32 - shift);
    } else if (shift == 0) {
        low = mantissa;
    } else {
        uint32_t r = (uint32_t)(-shift);
        low = mantissa >> r;
        uint32_t rest = mantissa & ((1u << r) - 1u);
        uint32_t half = 1u << (r - 1u);
        if (rest > half || (rest == half && (low & 1u) == 1u))
            low += 1u;
    }
    if ((bits & 0x80000000u) == 0) {
        *low_out = low;
        *high_out = high;
    } else {
        *low_out = ~low + 1u;
        *high_out = ~high + (low == 0 ? 1u : 0u);
    }
}

/* fixed.rs:50-52 */
int64_t vo_fixed_from_words(uint32_t low, uint32_t high)
{
    return (int64_t)(((uint64_t)high << 32) | (uint64_t)low);
}

/* fixed.rs:44-47 */
int64_t vo_fixed_to_raw(float x)
{
    uint32_t low, high;
    vo_fixed_split(x, &low, &high);
    return vo_fixed_from_words(low, high);
}

/* fixed.rs:55-57 */
double vo_fixed_to_f64(int64_t raw) { return (double)raw / TWO32; }

/* The FP64 counterpart of fixed_split that the HIP library uses: round-ties-even
 * of x * fixed_scale, where fixed_scale = scale * 2^32 is a power of two, so
 * the product is exact and the only rounding is llrint's. */
int64_t vo_fixed_f64(double x, double fixed_scale) { return llrint(x * fixed_scale); }

/* lennard_jones.rs:435-440 */
double vo_scale_for(double magnitude)
{
    if (magnitude <= 0.0 || !isfinite(magnitude))
        return 1.0;
    return ldexp(1.0, (int)floor(20.0 - log2(magnitude)));
}

/* ------------------------------------------------------------------ */
/* Initial states                                                      */
/* ------------------------------------------------------------------ */

/* xorshift64 step used throughout the reference (lennard_jones.rs:1411-1413). */
uint64_t vo_xorshift(uint64_t *s)
{
    *s ^= *s << 13;
    *s ^= *s >> 7;
    *s ^= *s << 17;
    return *s;
}

/* lennard_jones.rs:1409-1415: ((s >> 11) + 0.5) / 2^53 */
static double lattice_random(uint64_t *s)
{
    vo_xorshift(s);
    return ((double)(*s >> 11) + 0.5) / 9007199254740992.0;
}

/* lennard_jones.rs:1416-1419: Box-Muller cosine branch */
static double lattice_gaussian(uint64_t *s)
{
    double r = sqrt(-2.0 * log(lattice_random(s)));
    double u = lattice_random(s);
    return r * cos(6.283185307179586 * u);
}

static void remove_mean_velocity(size_t count, double *vel)
{
    /* lennard_jones.rs:1433-1439: mean summed in particle order */
    for (int k = 0; k < 3; k++) {
        double sum = 0.0;
        for (size_t i = 0; i < count; i++)
            sum += vel[3 * i + k];
        double mean = sum / (double)count;
        for (size_t i = 0; i < count; i++)
            vel[3 * i + k] -= mean;
    }
}

/* lennard_jones.rs:1401-1441 */
static void lattice_sites(size_t count, const double size[3], const size_t sites[3],
                          double temperature, uint64_t seed, double *pos, double *vel)
{
    double gap[3];
    for (int k = 0; k < 3; k++)
        gap[k] = size[k] / (double)sites[k];
    uint64_t s = seed | 1u;
    double st = sqrt(temperature);
    for (size_t i = 0; i < count; i++) {
        size_t c[3] = {i % sites[0], (i / sites[0]) % sites[1], i / (sites[0] * sites[1])};
        for (int k = 0; k < 3; k++)
            pos[3 * i + k] = ((double)c[k] + 0.5) * gap[k];
        for (int k = 0; k < 3; k++)
            vel[3 * i + k] = lattice_gaussian(&s) * st;
    }
    remove_mean_velocity(count, vel);
}

/* lennard_jones.rs:1379-1382 */
void vo_lattice(size_t count, double box_size, double temperature, uint64_t seed,
                double *pos, double *vel)
{
    size_t per_side = (size_t)ceil(cbrt((double)count));
    double size[3] = {box_size, box_size, box_size};
    size_t sites[3] = {per_side, per_side, per_side};
    lattice_sites(count, size, sites, temperature, seed, pos, vel);
}

/* lennard_jones.rs:1388-1399. Rust's max_by keeps the last of equal maxima. */
void vo_lattice_box(size_t count, const double size[3], double temperature, uint64_t seed,
                    double *pos, double *vel)
{
    double spacing = cbrt(size[0] * size[1] * size[2] / (double)count);
    size_t sites[3];
    for (int k = 0; k < 3; k++) {
        double r = round(size[k] / spacing);
        sites[k] = r < 1.0 ? 1 : (size_t)r;
    }
    while (sites[0] * sites[1] * sites[2] < count) {
        int widest = 0;
        for (int k = 1; k < 3; k++)
            if (size[k] / (double)sites[k] >= size[widest] / (double)sites[widest])
                widest = k;
        sites[widest] += 1;
    }
    lattice_sites(count, size, sites, temperature, seed, pos, vel);
}

/* tests/thermodynamics.rs:61-69, tests/coexistence.rs:81-89:
 * pos += ((s >> 11) / 2^53 - 0.5) * amplitude, three draws per particle. */
void vo_jitter(size_t count, uint64_t seed, double amplitude, double *pos)
{
    uint64_t s = seed;
    for (size_t i = 0; i < count; i++)
        for (int k = 0; k < 3; k++) {
            vo_xorshift(&s);
            pos[3 * i + k] += ((double)(s >> 11) / 9007199254740992.0 - 0.5) * amplitude;
        }
}

/* FCC lattice, 4 n^3 sites (new: the reference has only simple cubic).
 * The side is rounded to f32 first, as the simulation will round it
 * (lennard_jones.rs:450), and the lattice constant is side / n so the lattice
 * is exactly periodic in that box. Sites are offset by a/4 so none lies on a
 * cell boundary. Velocities come from the same stream as lattice_sites.
 * Returns the side. */
double vo_fcc(uint32_t n, double density, double temperature, uint64_t seed, double *pos,
              double *vel)
{
    double a0 = cbrt(4.0 / density);
    double side = (double)(float)(a0 * (double)n);
    double a = side / (double)n;
    static const double basis[4][3] = {{0, 0, 0}, {0.5, 0.5, 0}, {0.5, 0, 0.5}, {0, 0.5, 0.5}};
    uint64_t s = seed | 1u;
    double st = sqrt(temperature);
    size_t i = 0;
    for (uint32_t cz = 0; cz < n; cz++)
        for (uint32_t cy = 0; cy < n; cy++)
            for (uint32_t cx = 0; cx < n; cx++)
                for (int b = 0; b < 4; b++, i++) {
                    double c[3] = {(double)cx, (double)cy, (double)cz};
                    for (int k = 0; k < 3; k++)
                        pos[3 * i + k] = (c[k] + basis[b][k] + 0.25) * a;
                    for (int k = 0; k < 3; k++)
                        vel[3 * i + k] = lattice_gaussian(&s) * st;
                }
    remove_mean_velocity(i, vel);
    return side;
}

/* Random non-overlapping insertion (new): uniform trial positions from the
 * xorshift stream, rejected when closer than min_distance (nearest image) to
 * an earlier particle. Returns 0 on success, -1 if 1000 trials per particle
 * did not suffice. O(n^2); for fixtures and small replicas. */
int vo_random_insertion(size_t count, const double box[3], double min_distance,
                        double temperature, uint64_t seed, double *pos, double *vel)
{
    uint64_t s = seed | 1u;
    double md2 = min_distance * min_distance;
    for (size_t i = 0; i < count; i++) {
        int placed = 0;
        for (int trial = 0; trial < 1000 && !placed; trial++) {
            double p[3];
            for (int k = 0; k < 3; k++)
                p[k] = lattice_random(&s) * box[k];
            placed = 1;
            for (size_t j = 0; j < i && placed; j++) {
                double r2 = 0;
                for (int k = 0; k < 3; k++) {
                    double d = pos[3 * j + k] - p[k];
                    d -= box[k] * round(d / box[k]);
                    r2 += d * d;
                }
                if (r2 < md2)
                    placed = 0;
            }
            if (placed)
                for (int k = 0; k < 3; k++)
                    pos[3 * i + k] = p[k];
        }
        if (!placed)
            return -1;
    }
    double st = sqrt(temperature);
    for (size_t i = 0; i < count; i++)
        for (int k = 0; k < 3; k++)
            vel[3 * i + k] = lattice_gaussian(&s) * st;
    remove_mean_velocity(count, vel);
    return 0;
}

/* ------------------------------------------------------------------ */
/* Counter-based random numbers (Langevin noise)                        */
/* ------------------------------------------------------------------ */

/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
 * easy as 1, 2, 3", SC'11): ten rounds of two 32x32->64 multiplies with the
 * round keys bumped by the Weyl constants. */
void vo_philox4x32(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4])
{
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int round = 0; round < 10; round++) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

/* Natural log from +, *, fma and / only, so the GPU reproduces it bit for bit:
 * x = 2^e m with m in [sqrt(1/2), sqrt(2)); log m = 2 atanh(s), s = (m-1)/(m+1),
 * by the odd series to s^21 (|s| <= 0.172); e ln 2 with ln 2 split in two
 * (Cody and Waite). Positive normal x only. */
double vo_log(double x)
{
    uint64_t b;
    memcpy(&b, &x, 8);
    int e = (int)((b >> 52) & 0x7ff) - 1023;
    b = (b & 0x000fffffffffffffull) | 0x3ff0000000000000ull;
    double m;
    memcpy(&m, &b, 8);
    if (m > 1.4142135623730951) {
        m = m * 0.5;
        e += 1;
    }
    double f = m - 1.0;
    double s = f / (2.0 + f);
    double z = s * s;
    double p = 1.0 / 21.0;
    p = fma(p, z, 1.0 / 19.0);
    p = fma(p, z, 1.0 / 17.0);
    p = fma(p, z, 1.0 / 15.0);
    p = fma(p, z, 1.0 / 13.0);
    p = fma(p, z, 1.0 / 11.0);
    p = fma(p, z, 1.0 / 9.0);
    p = fma(p, z, 1.0 / 7.0);
    p = fma(p, z, 1.0 / 5.0);
    p = fma(p, z, 1.0 / 3.0);
    double tail = 2.0 * (s * z * p);
    double de = (double)e;
    const double ln2_hi = 6.93147180369123816490e-01; /* 0x3fe62e42fee00000 */
    const double ln2_lo = 1.90821492927058770002e-10; /* 0x3dea39ef35793c76 */
    return fma(de, ln2_hi, 2.0 * s + fma(de, ln2_lo, tail));
}

/* sin(2 pi u) and cos(2 pi u) for u in [0, 1] from +, * and fma only:
 * the nearest quarter turn q = rint(4u) is taken off exactly, the remainder
 * x = 2 pi (u - q/4) lies in [-pi/4, pi/4] and goes through the Taylor series
 * (sin to x^17, cos to x^18), then the quadrant rotates the pair. */
void vo_sincos_turns(double u, double *s_out, double *c_out)
{
    double q = rint(4.0 * u);
    double t = u - 0.25 * q;
    double x = t * 6.283185307179586;
    double z = x * x;
    double sp = -1.0 / 355687428096000.0;
    sp = fma(sp, z, 1.0 / 1307674368000.0);
    sp = fma(sp, z, -1.0 / 6227020800.0);
    sp = fma(sp, z, 1.0 / 39916800.0);
    sp = fma(sp, z, -1.0 / 362880.0);
    sp = fma(sp, z, 1.0 / 5040.0);
    sp = fma(sp, z, -1.0 / 120.0);
    sp = fma(sp, z, 1.0 / 6.0);
    double sn = fma(-x * z, sp, x);
    double cp = 1.0 / 6402373705728000.0;
    cp = fma(cp, z, -1.0 / 20922789888000.0);
    cp = fma(cp, z, 1.0 / 87178291200.0);
    cp = fma(cp, z, -1.0 / 479001600.0);
    cp = fma(cp, z, 1.0 / 3628800.0);
    cp = fma(cp, z, -1.0 / 40320.0);
    cp = fma(cp, z, 1.0 / 720.0);
    cp = fma(cp, z, -1.0 / 24.0);
    cp = fma(cp, z, 0.5);
    double cs = fma(-z, cp, 1.0);
    int quadrant = ((int)q) & 3;
    switch (quadrant) {
    case 0: *s_out = sn; *c_out = cs; break;
    case 1: *s_out = cs; *c_out = -sn; break;
    case 2: *s_out = -sn; *c_out = -cs; break;
    default: *s_out = -cs; *c_out = sn; break;
    }
}

/* Three standard normals for one atom and step: Philox4x32-10 with counter
 * (step low, step high, original atom index, stream) and key (seed low,
 * seed high); uniforms (w + 0.5) 2^-32; Box-Muller on (u0, u1) gives two,
 * the cosine branch on (u2, u3) the third. The reference draws four and uses
 * three per step as well (lennard_jones.wgsl:242-247), from a stateful PCG.
 * The thermostat uses stream 0: replicas differ by seed only. */
void vo_langevin_noise(uint64_t seed, uint32_t stream, uint32_t atom, uint64_t step,
                       double out[3])
{
    uint32_t ctr[4] = {(uint32_t)step, (uint32_t)(step >> 32), atom, stream};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t w[4];
    vo_philox4x32(ctr, key, w);
    double u[4];
    for (int k = 0; k < 4; k++)
        u[k] = ((double)w[k] + 0.5) * 2.3283064365386963e-10;
    double r0 = sqrt(-2.0 * vo_log(u[0]));
    double s0, c0;
    vo_sincos_turns(u[1], &s0, &c0);
    double r1 = sqrt(-2.0 * vo_log(u[2]));
    double s1, c1;
    vo_sincos_turns(u[3], &s1, &c1);
    (void)s1;
    out[0] = r0 * c0;
    out[1] = r0 * s0;
    out[2] = r1 * c1;
}

/* ------------------------------------------------------------------ */
/* Parameters                                                          */
/* ------------------------------------------------------------------ */

static double wrap_initial(double x, double side)
{
    /* Rust rem_euclid (lennard_jones.rs:477) */
    double r = fmod(x, side);
    return r < 0.0 ? r + side : r;
}

/* R2 */
static uint32_t capacity_override; /* 0: the capacity rule */

void vo_set_capacity_override(uint32_t capacity) { capacity_override = capacity; }

static inline uint32_t cell_of_position(const vo_params *p, double x, double y, double z)
{
    float q[3] = {(float)x, (float)y, (float)z};
    int32_t c[3];
    for (int k = 0; k < 3; k++) {
        float t = q[k] / p->cell_size[k];
        int32_t ci = (int32_t)floorf(t);
        int32_t m = (int32_t)p->cells[k] - 1;
        c[k] = ci < 0 ? 0 : (ci > m ? m : ci);
    }
    return ((uint32_t)c[2] * p->cells[1] + (uint32_t)c[1]) * p->cells[0] + (uint32_t)c[0];
}

/* lennard_jones.rs:443-560, with two deliberate changes:
 *  - capacity uses the larger of the box-average density and the densest cell
 *    of the initial state (the reference's box-average rule overflows for a
 *    dense slab in a mostly empty box); the reference's value is kept too;
 *  - dt stays FP64 (the reference rounds it to f32 for its shaders). */
int vo_derive_params(const double *pos, size_t n, const double box_in[3], double cutoff,
                     double skin, double dt, int shifted_force, int precision,
                     int thermostat, double temperature, double friction, double period,
                     uint64_t seed, const double *vel, vo_params *p)
{
    memset(p, 0, sizeof(*p));
    double list_radius = cutoff + skin;
    for (int k = 0; k < 3; k++) {
        p->box[k] = (double)(float)box_in[k];
        p->half_box[k] = 0.5 * p->box[k];
        double c = floor(p->box[k] / list_radius);
        p->cells[k] = c < 0 ? 0 : (uint32_t)c;
    }
    /* The reference asks for 3 cells per axis (lennard_jones.rs:458-461);
     * here 2 suffice: with L >= 2 (rc + skin) at most one image of a pair
     * is within rc + skin, so a stencil that meets a cell twice (once per
     * image) still lists each pair once. */
    for (int k = 0; k < 3; k++)
        if (p->cells[k] < 2)
            return -1;
    while ((uint64_t)p->cells[0] * p->cells[1] * p->cells[2] > (1u << 20)) {
        int longest = 0;
        for (int k = 1; k < 3; k++)
            if (p->cells[k] >= p->cells[longest])
                longest = k;
        p->cells[longest] -= 1;
    }
    p->total_cells = p->cells[0] * p->cells[1] * p->cells[2];
    double cell_size[3];
    for (int k = 0; k < 3; k++) {
        cell_size[k] = p->box[k] / (double)p->cells[k];
        p->cell_size[k] = (float)cell_size[k];
    }
    double volume = p->box[0] * p->box[1] * p->box[2];
    p->count = (uint32_t)n;
    double density = (double)n / volume;
    double sphere = 4.0 / 3.0 * 3.14159265358979323846 * list_radius * list_radius * list_radius;
    p->capacity_reference = (uint32_t)ceil(2.0 * sphere * density + 32.0);
    /* densest cell of the initial state */
    uint32_t *occ = (uint32_t *)calloc(p->total_cells, sizeof(uint32_t));
    uint32_t occ_max = 0;
    for (size_t i = 0; i < n; i++) {
        double w[3];
        for (int k = 0; k < 3; k++)
            w[k] = wrap_initial(pos[3 * i + k], p->box[k]);
        uint32_t c = cell_of_position(p, w[0], w[1], w[2]);
        occ[c]++;
        if (occ[c] > occ_max)
            occ_max = occ[c];
    }
    free(occ);
    double cell_volume = cell_size[0] * cell_size[1] * cell_size[2];
    double local = (double)occ_max / cell_volume;
    double use = local > density ? local : density;
    p->capacity = (uint32_t)ceil(2.0 * sphere * use + 32.0);
    /* vo_set_capacity_override may replace the rule (at least 8): the tests'
     * way to reach list overflow, which the rule avoids for the initial state */
    if (capacity_override)
        p->capacity = capacity_override > 8 ? capacity_override : 8;

    double kinetic = 0.0, speed = 0.0;
    for (size_t i = 0; i < n; i++) {
        double v2 = 0.0;
        for (int k = 0; k < 3; k++)
            v2 += vel[3 * i + k] * vel[3 * i + k];
        kinetic += 0.5 * v2;
        speed += sqrt(v2);
    }
    double thermal = thermostat != 0 ? 1.5 * temperature * (double)n : 0.0;
    double m = kinetic > thermal ? kinetic : thermal;
    if ((double)n > m)
        m = (double)n;
    p->energy_scale = vo_scale_for(m);
    p->momentum_scale = vo_scale_for(speed > (double)n ? speed : (double)n);
    p->energy_fixed = p->energy_scale * TWO32;
    p->momentum_fixed = p->momentum_scale * TWO32;

    p->rc = cutoff;
    p->rc2 = cutoff * cutoff;
    p->list_r2 = list_radius * list_radius;
    p->half_skin2 = 0.25 * skin * skin;
    double inv6 = 1.0 / (p->rc2 * p->rc2 * p->rc2);
    p->shift = 4.0 * (inv6 * inv6 - inv6);
    p->force_shift = shifted_force ? 24.0 * (2.0 * inv6 * inv6 - inv6) / cutoff : 0.0;
    p->shifted_force = shifted_force;
    p->dt = dt;
    p->half_dt = 0.5 * dt;
    p->thermostat = thermostat;
    p->c1 = 1.0;
    p->c2 = 0.0;
    p->nh_q = 1.0;
    p->nh_period = period;
    if (thermostat == 2) {
        /* lennard_jones.rs:516-528: zero total momentum, g = 3N - 3 */
        double dof = 3.0 * (double)n - 3.0;
        p->nh_gkt = dof * temperature;
        p->nh_q = dof * temperature * period * period / (4.0 * (VO_PI * VO_PI));
    }
    if (thermostat == 1) {
        p->c1 = exp(-friction * dt);
        p->c2 = sqrt((1.0 - p->c1 * p->c1) * temperature);
    }
    p->seed = seed;
    p->precision = precision;
    p->near_count = 4;
    return 0;
}

/* ------------------------------------------------------------------ */
/* Simulation                                                          */
/* ------------------------------------------------------------------ */

static void *xcalloc(size_t n, size_t size)
{
    void *p = calloc(n ? n : 1, size);
    if (!p)
        abort();
    return p;
}

vo_sim *vo_create(const double *pos, const double *vel, size_t n, const double box[3],
                  double cutoff, double skin, double dt, int shifted_force, int precision,
                  int thermostat, double temperature, double friction, double period,
                  uint64_t seed)
{
    vo_sim *s = (vo_sim *)xcalloc(1, sizeof(vo_sim));
    if (vo_derive_params(pos, n, box, cutoff, skin, dt, shifted_force, precision, thermostat,
                         temperature, friction, period, seed, vel, &s->p) != 0) {
        free(s);
        return NULL;
    }
    for (int k = 0; k < 3; k++) {
        s->x[k] = (double *)xcalloc(n, sizeof(double));
        s->v[k] = (double *)xcalloc(n, sizeof(double));
        s->a[k] = (double *)xcalloc(n, sizeof(double));
        s->xb[k] = (double *)xcalloc(n, sizeof(double));
    }
    s->orig = (uint32_t *)xcalloc(n, 4);
    s->cell_of = (uint32_t *)xcalloc(n, 4);
    s->sorted = (uint32_t *)xcalloc(n, 4);
    s->cell_start = (uint32_t *)xcalloc(s->p.total_cells + 1, 4);
    s->cell_count = (uint32_t *)xcalloc(s->p.total_cells, 4);
    s->nbr_count = (uint32_t *)xcalloc(n, 4);
    s->nbr = (uint32_t *)xcalloc(n * (size_t)s->p.capacity, 4);
    s->nbr_image = (uint8_t *)xcalloc(n * (size_t)s->p.capacity, 1);
    s->pe = (double *)xcalloc(n, sizeof(double));
    s->w = (double *)xcalloc(n, sizeof(double));
    s->near = (uint32_t *)xcalloc(n, 4);
    s->within = (uint32_t *)xcalloc(n, 4);
    for (size_t i = 0; i < n; i++) {
        for (int k = 0; k < 3; k++) {
            s->x[k][i] = wrap_initial(pos[3 * i + k], s->p.box[k]);
            s->v[k][i] = vel[3 * i + k];
        }
        s->orig[i] = (uint32_t)i;
    }
    return s;
}

void vo_destroy(vo_sim *s)
{
    if (!s)
        return;
    for (int k = 0; k < 3; k++) {
        free(s->x[k]);
        free(s->v[k]);
        free(s->a[k]);
        free(s->xb[k]);
    }
    free(s->orig);
    free(s->cell_of);
    free(s->sorted);
    free(s->cell_start);
    free(s->cell_count);
    free(s->nbr_count);
    free(s->nbr);
    free(s->nbr_image);
    free(s->pe);
    free(s->w);
    free(s->near);
    free(s->within);
    free(s->prof);
    free(s->hc);
    free(s);
}

void vo_get_params(const vo_sim *s, vo_params *out) { *out = s->p; }
void vo_set_near_count(vo_sim *s, uint32_t near_count) { s->p.near_count = near_count; }
uint32_t vo_rebuilds(const vo_sim *s) { return s->rebuilds; }

/* R1: one box length, for positions at most a skin beyond the box */
static inline double wrap_step(double x, double side)
{
    if (x >= side)
        return x - side;
    if (x < 0.0)
        return x + side;
    return x;
}

/* grid.wgsl:22-148: bin, prefix sum, then cell order ascending by slot
 * (a stable counting sort gives exactly the reference's per-cell insertion
 * sort result). */
static void bin_and_sort(vo_sim *s)
{
    const vo_params *p = &s->p;
    uint32_t n = p->count;
    memset(s->cell_count, 0, (size_t)p->total_cells * 4);
    for (uint32_t i = 0; i < n; i++) {
        uint32_t c = cell_of_position(p, s->x[0][i], s->x[1][i], s->x[2][i]);
        s->cell_of[i] = c;
        s->cell_count[c]++;
    }
    uint32_t run = 0;
    for (uint32_t c = 0; c < p->total_cells; c++) {
        s->cell_start[c] = run;
        run += s->cell_count[c];
    }
    s->cell_start[p->total_cells] = n;
    memset(s->cell_count, 0, (size_t)p->total_cells * 4);
    for (uint32_t i = 0; i < n; i++) {
        uint32_t c = s->cell_of[i];
        s->sorted[s->cell_start[c] + s->cell_count[c]++] = i;
    }
}

/* lennard_jones.wgsl:419-452: slot s takes the particle at sorted[s]. */
static void reorder(vo_sim *s)
{
    uint32_t n = s->p.count;
    double *tmp = (double *)xcalloc(n, sizeof(double));
    uint32_t *itmp = (uint32_t *)xcalloc(n, 4);
    double **arrays[3] = {s->x, s->v, s->a};
    for (int g = 0; g < 3; g++)
        for (int k = 0; k < 3; k++) {
            double *arr = arrays[g][k];
            for (uint32_t t = 0; t < n; t++)
                tmp[t] = arr[s->sorted[t]];
            memcpy(arr, tmp, (size_t)n * sizeof(double));
        }
    for (uint32_t t = 0; t < n; t++)
        itmp[t] = s->orig[s->sorted[t]];
    memcpy(s->orig, itmp, (size_t)n * 4);
    for (uint32_t t = 0; t < n; t++)
        itmp[t] = s->cell_of[s->sorted[t]];
    memcpy(s->cell_of, itmp, (size_t)n * 4);
    for (uint32_t t = 0; t < n; t++)
        s->sorted[t] = t;
    free(tmp);
    free(itmp);
}

/* R4, lennard_jones.wgsl:454-499 */
static void build_lists(vo_sim *s)
{
    const vo_params *p = &s->p;
    int32_t n = (int32_t)p->count;
    uint32_t cap = p->capacity;
    int32_t nx = (int32_t)p->cells[0], ny = (int32_t)p->cells[1], nz = (int32_t)p->cells[2];
    uint32_t overflow = 0;
#pragma omp parallel for schedule(static) reduction(max : overflow)
    for (int32_t i = 0; i < n; i++) {
        double xi = s->x[0][i], yi = s->x[1][i], zi = s->x[2][i];
        int32_t c = (int32_t)s->cell_of[i];
        int32_t cx = c % nx, cy = (c / nx) % ny, cz = c / (nx * ny);
        uint32_t found = 0;
        uint32_t *row = s->nbr + (size_t)i * cap;
        for (int32_t dz = -1; dz <= 1; dz++)
            for (int32_t dy = -1; dy <= 1; dy++)
                for (int32_t dx = -1; dx <= 1; dx++) {
                    int32_t ncx = (cx + dx + nx) % nx, ncy = (cy + dy + ny) % ny,
                            ncz = (cz + dz + nz) % nz;
                    uint32_t cell = (uint32_t)((ncz * ny + ncy) * nx + ncx);
                    /* periodic image of the neighbour cell: -L, 0 or +L per axis */
                    int ix = cx + dx < 0 ? -1 : (cx + dx >= nx ? 1 : 0);
                    int iy = cy + dy < 0 ? -1 : (cy + dy >= ny ? 1 : 0);
                    int iz = cz + dz < 0 ? -1 : (cz + dz >= nz ? 1 : 0);
                    double sx = ix * p->box[0], sy = iy * p->box[1], sz = iz * p->box[2];
                    uint8_t code = (uint8_t)((ix + 1) + 3 * (iy + 1) + 9 * (iz + 1));
                    for (uint32_t t = s->cell_start[cell]; t < s->cell_start[cell + 1]; t++) {
                        uint32_t j = s->sorted[t];
                        if ((int32_t)j == i)
                            continue;
                        double ddx = (s->x[0][j] + sx) - xi;
                        double ddy = (s->x[1][j] + sy) - yi;
                        double ddz = (s->x[2][j] + sz) - zi;
                        double r2 = fma(ddz, ddz, fma(ddy, ddy, ddx * ddx));
                        if (r2 < p->list_r2) {
                            if (found < cap) {
                                row[found] = j;
                                s->nbr_image[(size_t)i * cap + found] = code;
                            }
                            found++;
                        }
                    }
                }
        s->nbr_count[i] = found < cap ? found : cap;
        if (found > cap && found > overflow)
            overflow = found;
        s->xb[0][i] = xi;
        s->xb[1][i] = yi;
        s->xb[2][i] = zi;
    }
    if (overflow > s->overflow)
        s->overflow = overflow;
}

/* R8, lennard_jones.wgsl:278-307 and the rebuild chain of lennard_jones.rs:1055-1078 */
static void update_lists(vo_sim *s)
{
    int rebuild = !s->built || s->max_moved2 > s->p.half_skin2;
    s->max_moved2 = 0.0;
    if (!rebuild)
        return;
    s->rebuilds++;
    s->built = 1;
    for (int k = 0; k < 3; k++)
        for (uint32_t i = 0; i < s->p.count; i++)
            s->x[k][i] = wrap_step(s->x[k][i], s->p.box[k]);
    bin_and_sort(s);
    reorder(s);
    build_lists(s);
}

/* R5 / R5f, lennard_jones.wgsl:501-545: the acceleration of atom i; with
 * `energies` also its pe, w, near and within for the measurement (R9). */
static void interactions(vo_sim *s, int32_t i, double acc[3], int energies)
{
    const vo_params *p = &s->p;
    const double fs24 = p->force_shift / 24.0;
    double xi = s->x[0][i], yi = s->x[1][i], zi = s->x[2][i];
    double gx = 0.0, gy = 0.0, gz = 0.0, pe = 0.0, w = 0.0;
    float fx = 0.0f, fy = 0.0f, fz = 0.0f; /* R5f: the current octet's sum */
    uint32_t near = 0, within = 0;
    uint32_t cnt = s->nbr_count[i];
    const uint32_t *row = s->nbr + (size_t)i * p->capacity;
    const uint8_t *img = s->nbr_image + (size_t)i * p->capacity;
    for (uint32_t k = 0; k < cnt; k++) {
        uint32_t j = row[k];
        int code = img[k];
        double sx = (code % 3 - 1) * p->box[0];
        double sy = ((code / 3) % 3 - 1) * p->box[1];
        double sz = (code / 9 - 1) * p->box[2];
        double dx = (s->x[0][j] + sx) - xi;
        double dy = (s->x[1][j] + sy) - yi;
        double dz = (s->x[2][j] + sz) - zi;
        double r2 = fma(dz, dz, fma(dy, dy, dx * dx));
        if (!(r2 < p->rc2))
            goto next; /* past the cutoff: no term (an R5f octet still ends) */
        double rf = 0.0, u = 0.0;
        if (p->precision == 0) {
            double inv2 = 1.0 / r2;
            double s6 = inv2 * inv2 * inv2;
            double g = s6 * fma(s6, 2.0, -1.0); /* r.F / 24 */
            if (p->shifted_force)
                g = fma(-sqrt(r2), fs24, g);
            double fscal = g * inv2;
            gx = fma(dx, fscal, gx);
            gy = fma(dy, fscal, gy);
            gz = fma(dz, fscal, gz);
            if (energies) {
                rf = s6 * fma(48.0, s6, -24.0);
                double t = 4.0 * s6;
                u = fma(t, s6, -t) - p->shift;
                if (p->shifted_force) {
                    double r = sqrt(r2);
                    rf = fma(-r, p->force_shift, rf);
                    u = fma(r - p->rc, p->force_shift, u);
                }
            }
        } else {
            float q = (float)r2;
            float inv2 = 1.0f / q;
            float s6 = inv2 * inv2 * inv2;
            float g = s6 * fmaf(s6, 2.0f, -1.0f);
            if (p->shifted_force)
                g = fmaf(-sqrtf(q), (float)fs24, g);
            float fscal = g * inv2;
            fx = fmaf((float)dx, fscal, fx);
            fy = fmaf((float)dy, fscal, fy);
            fz = fmaf((float)dz, fscal, fz);
            if (energies) {
                float rff = s6 * fmaf(48.0f, s6, -24.0f);
                float t = 4.0f * s6;
                float uf = fmaf(t, s6, -t) - (float)p->shift;
                if (p->shifted_force) {
                    float r = sqrtf(q);
                    float fs = (float)p->force_shift;
                    rff = fmaf(-r, fs, rff);
                    uf = fmaf(r - (float)p->rc, fs, uf);
                }
                rf = (double)rff;
                u = (double)uf;
            }
        }
        if (energies) {
            pe += 0.5 * u;
            w += 0.5 * rf;
            near += r2 < 2.25 ? 1u : 0u;
            within += 1u;
        }
    next:
        if (p->precision != 0 && ((k & 7u) == 7u || k + 1u == cnt)) {
            gx += (double)fx;
            gy += (double)fy;
            gz += (double)fz;
            fx = fy = fz = 0.0f;
        }
    }
    acc[0] = -24.0 * gx;
    acc[1] = -24.0 * gy;
    acc[2] = -24.0 * gz;
    if (energies) {
        s->pe[i] = pe;
        s->w[i] = w;
        s->near[i] = near;
        s->within[i] = within;
    }
}

static void compute_forces(vo_sim *s, int kick)
{
    int32_t n = (int32_t)s->p.count;
    double hdt = s->p.half_dt;
#pragma omp parallel for schedule(static)
    for (int32_t i = 0; i < n; i++) {
        double acc[3];
        interactions(s, i, acc, 0);
        for (int k = 0; k < 3; k++) {
            s->a[k][i] = acc[k];
            if (kick)
                s->v[k][i] = fma(acc[k], hdt, s->v[k][i]);
        }
    }
}

/* lennard_jones.rs:1080-1089 */
void vo_prime(vo_sim *s)
{
    if (s->primed)
        return;
    update_lists(s);
    compute_forces(s, 0);
    s->primed = 1;
}

/* R6 / R12, lennard_jones.wgsl:204-276 */
static void drift(vo_sim *s)
{
    const vo_params *p = &s->p;
    int32_t n = (int32_t)p->count;
    double maxd = s->max_moved2;
    uint64_t step = s->steps;
#pragma omp parallel for schedule(static) reduction(max : maxd)
    for (int32_t i = 0; i < n; i++) {
        double noise[3] = {0.0, 0.0, 0.0};
        if (p->thermostat == 1)
            vo_langevin_noise(p->seed, 0u, s->orig[i], step, noise);
        double x[3];
        for (int k = 0; k < 3; k++) {
            double v0 = s->v[k][i];
            if (p->thermostat == 2) /* R13: the first half step's scaling */
                v0 = fma(v0, s->nh_sm1, v0);
            double v = fma(s->a[k][i], p->half_dt, v0);
            double xk = s->x[k][i];
            if (p->thermostat == 1) {
                xk = fma(v, p->half_dt, xk);
                v = fma(p->c2, noise[k], p->c1 * v);
                xk = fma(v, p->half_dt, xk);
            } else {
                xk = fma(v, p->dt, xk);
            }
            s->v[k][i] = v;
            s->x[k][i] = xk;
            x[k] = xk;
        }
        double dx = x[0] - s->xb[0][i];
        double dy = x[1] - s->xb[1][i];
        double dz = x[2] - s->xb[2][i];
        double d2 = fma(dz, dz, fma(dy, dy, dx * dx));
        if (d2 > maxd)
            maxd = d2;
    }
    s->max_moved2 = maxd;
}

/* R13 Nose-Hoover half step (lennard_jones.wgsl:309-374): KE from the
 * fixed-point sum, then xi, eta and the velocity scale s - 1 = exp(-xi dt/2) - 1
 * by its series, in FP64. */
static void nh_half(vo_sim *s)
{
    const vo_params *p = &s->p;
    int32_t n = (int32_t)p->count;
    int64_t raw = 0;
#pragma omp parallel for schedule(static) reduction(+ : raw)
    for (int32_t i = 0; i < n; i++) {
        double vx = s->v[0][i], vy = s->v[1][i], vz = s->v[2][i];
        raw += vo_fixed_f64(0.5 * fma(vz, vz, fma(vy, vy, vx * vx)), p->energy_fixed);
    }
    double ke = (double)raw / p->energy_fixed;
    double quarter = 0.25 * p->dt;
    s->xi = fma((2.0 * ke - p->nh_gkt) / p->nh_q, quarter, s->xi);
    double x = s->xi * p->half_dt;
    double sm1 = -x * (1.0 - x * (0.5 - x * (1.0 / 6.0 - x / 24.0)));
    s->eta = fma(s->xi, p->half_dt, s->eta);
    double ke2 = fma(ke, sm1 * (2.0 + sm1), ke);
    s->xi = fma((2.0 * ke2 - p->nh_gkt) / p->nh_q, quarter, s->xi);
    s->nh_sm1 = sm1;
}

static void nh_scale(vo_sim *s)
{
    int32_t n = (int32_t)s->p.count;
    double sm1 = s->nh_sm1;
#pragma omp parallel for schedule(static)
    for (int32_t i = 0; i < n; i++)
        for (int k = 0; k < 3; k++)
            s->v[k][i] = fma(s->v[k][i], sm1, s->v[k][i]);
}

/* R15 heat exchange, eHEX/a (Wirnsberger, Frenkel and Dellago 2015, eq. 20;
 * lennard_jones.wgsl:732-870): after the step the hot slab gains `energy`
 * and the cold slab loses it, by scaling velocities about each slab's
 * centre-of-mass velocity, with the O(dt^3) position correction. Slab sums
 * go through 64-bit fixed point (count exact; momentum and force at the
 * momentum scale; v^2 and f.v at the energy scale). Membership uses the
 * wrapped z. */
static void pair_terms(const vo_params *p, double r2, double *rf_out, double *u_out);

static int hex_slab(const vo_sim *s, double z)
{
    const double *b = s->hex_slabs;
    if (z >= b[0] && z < b[1])
        return 0;
    if (z >= b[2] && z < b[3])
        return 1;
    return 2;
}

static void heat_exchange(vo_sim *s)
{
    const vo_params *p = &s->p;
    int32_t n = (int32_t)p->count;
    int64_t cnt0 = 0, cnt1 = 0, q0[8] = {0}, q1[8] = {0};
#pragma omp parallel for schedule(static) reduction(+ : cnt0, cnt1) reduction(+ : q0[:8], q1[:8])
    for (int32_t i = 0; i < n; i++) {
        int slab = hex_slab(s, wrap_step(s->x[2][i], p->box[2]));
        if (slab == 2)
            continue;
        double v[3] = {s->v[0][i], s->v[1][i], s->v[2][i]};
        double f[3] = {s->a[0][i], s->a[1][i], s->a[2][i]};
        int64_t *q = slab == 0 ? q0 : q1;
        if (slab == 0)
            cnt0++;
        else
            cnt1++;
        for (int k = 0; k < 3; k++) {
            q[k] += vo_fixed_f64(v[k], p->momentum_fixed);
            q[5 + k] += vo_fixed_f64(f[k], p->momentum_fixed);
        }
        q[3] += vo_fixed_f64(fma(v[2], v[2], fma(v[1], v[1], v[0] * v[0])), p->energy_fixed);
        q[4] += vo_fixed_f64(fma(f[2], v[2], fma(f[1], v[1], f[0] * v[0])), p->energy_fixed);
    }
    double com[2][3], mf[2][3], sm1[2], cvel[2], cforce[2];
    const double dt3 = p->dt * p->dt * p->dt;
    for (int slab = 0; slab < 2; slab++) {
        const int64_t *q = slab == 0 ? q0 : q1;
        double m = (double)(slab == 0 ? cnt0 : cnt1);
        sm1[slab] = cvel[slab] = cforce[slab] = 0.0;
        for (int k = 0; k < 3; k++)
            com[slab][k] = mf[slab][k] = 0.0;
        if (!(m > 0.0))
            continue;
        double pm[3], fm[3];
        for (int k = 0; k < 3; k++) {
            pm[k] = (double)q[k] / p->momentum_fixed;
            fm[k] = (double)q[5 + k] / p->momentum_fixed;
            com[slab][k] = pm[k] / m;
            mf[slab][k] = fm[k] / m;
        }
        double v2 = (double)q[3] / p->energy_fixed;
        double fv = (double)q[4] / p->energy_fixed;
        double pp = fma(pm[2], pm[2], fma(pm[1], pm[1], pm[0] * pm[0]));
        double kinetic = 0.5 * v2 - 0.5 * pp / m;
        double energy = slab == 0 ? s->hex_energy : -s->hex_energy;
        double flux = energy / p->dt;
        double x = energy / kinetic;
        sm1[slab] = x / (1.0 + sqrt(1.0 + x));
        double power = fv - fma(fm[2], com[slab][2], fma(fm[1], com[slab][1], fm[0] * com[slab][0]));
        cvel[slab] = dt3 * flux / (2.0 * kinetic * kinetic) * (flux / 48.0 + power / 6.0);
        cforce[slab] = dt3 * flux / (12.0 * kinetic);
    }
#pragma omp parallel for schedule(static)
    for (int32_t i = 0; i < n; i++) {
        int slab = hex_slab(s, wrap_step(s->x[2][i], p->box[2]));
        if (slab == 2)
            continue;
        for (int k = 0; k < 3; k++) {
            double rel = s->v[k][i] - com[slab][k];
            double shift = cvel[slab] * rel - cforce[slab] * (s->a[k][i] - mf[slab][k]);
            s->x[k][i] = s->x[k][i] - shift;
            s->v[k][i] = fma(rel, sm1[slab], s->v[k][i]);
        }
    }
}

int vo_set_heat_exchange(vo_sim *s, int on, const double hot[2], const double cold[2], double energy)
{
    if (s->p.thermostat != 0)
        return -1;
    s->hex_on = on;
    s->hex_slabs[0] = on ? hot[0] : 0.0;
    s->hex_slabs[1] = on ? hot[1] : 0.0;
    s->hex_slabs[2] = on ? cold[0] : 0.0;
    s->hex_slabs[3] = on ? cold[1] : 0.0;
    s->hex_energy = on ? energy : 0.0;
    return 0;
}

/* R14 heat current J = sum_i e_i v_i + 1/2 sum_i sum_j (F_ij . v_i) r_ij
 * (lennard_jones.wgsl:872-947), e_i = KE_i + half of each pair energy, over
 * atom i's list in order: q = ((d . v_i) rf) / r2, P = fma(d, q, P);
 * j_i = fma(e_i, v_i, 0.5 P); each component to fixed point at 2^32 and
 * summed as integers, one sample of three sums. */
void vo_record_heat_current(vo_sim *s, uint32_t capacity)
{
    free(s->hc);
    s->hc = capacity ? (int64_t *)xcalloc((size_t)capacity * 3, sizeof(int64_t)) : NULL;
    s->hc_cap = capacity;
    s->hc_samples = 0;
}

int vo_sample_heat_current(vo_sim *s)
{
    if (!s->hc || s->hc_samples >= s->hc_cap)
        return -1;
    vo_prime(s);
    const vo_params *p = &s->p;
    int32_t n = (int32_t)p->count;
    int64_t jx = 0, jy = 0, jz = 0;
#pragma omp parallel for schedule(static) reduction(+ : jx, jy, jz)
    for (int32_t i = 0; i < n; i++) {
        const double xi = s->x[0][i], yi = s->x[1][i], zi = s->x[2][i];
        const double v[3] = {s->v[0][i], s->v[1][i], s->v[2][i]};
        double e = 0.5 * fma(v[2], v[2], fma(v[1], v[1], v[0] * v[0]));
        double pr[3] = {0.0, 0.0, 0.0};
        const uint32_t cnt = s->nbr_count[i];
        const uint32_t *row = s->nbr + (size_t)i * p->capacity;
        const uint8_t *img = s->nbr_image + (size_t)i * p->capacity;
        for (uint32_t k = 0; k < cnt; k++) {
            const uint32_t j = row[k];
            const int code = img[k];
            const double d[3] = {(s->x[0][j] + (code % 3 - 1) * p->box[0]) - xi,
                                 (s->x[1][j] + ((code / 3) % 3 - 1) * p->box[1]) - yi,
                                 (s->x[2][j] + (code / 9 - 1) * p->box[2]) - zi};
            const double r2 = fma(d[2], d[2], fma(d[1], d[1], d[0] * d[0]));
            if (!(r2 < p->rc2))
                continue;
            double rf, u;
            pair_terms(p, r2, &rf, &u);
            e += 0.5 * u;
            const double dv = fma(d[2], v[2], fma(d[1], v[1], d[0] * v[0]));
            const double q = dv * rf / r2;
            for (int a = 0; a < 3; a++)
                pr[a] = fma(d[a], q, pr[a]);
        }
        jx += llrint(fma(e, v[0], 0.5 * pr[0]) * TWO32);
        jy += llrint(fma(e, v[1], 0.5 * pr[1]) * TWO32);
        jz += llrint(fma(e, v[2], 0.5 * pr[2]) * TWO32);
    }
    int64_t *out = s->hc + 3 * (size_t)s->hc_samples;
    out[0] = jx;
    out[1] = jy;
    out[2] = jz;
    s->hc_samples++;
    return 0;
}

void vo_read_heat_current(const vo_sim *s, int64_t *raw, uint32_t *samples)
{
    if (raw && s->hc)
        memcpy(raw, s->hc, (size_t)s->hc_samples * 3 * sizeof(int64_t));
    if (samples)
        *samples = s->hc_samples;
}

/* lennard_jones.rs:1092-1138 (NVE, Langevin, Nose-Hoover) */
void vo_run(vo_sim *s, uint64_t steps)
{
    vo_prime(s);
    int nh = s->p.thermostat == 2;
    for (uint64_t t = 0; t < steps; t++) {
        if (nh)
            nh_half(s);
        drift(s);
        update_lists(s);
        compute_forces(s, 1);
        if (nh) {
            nh_half(s);
            nh_scale(s);
        }
        if (s->hex_on)
            heat_exchange(s);
        s->steps++;
    }
}

/* lennard_jones.rs:905-925 */
void vo_set_temperature(vo_sim *s, double temperature)
{
    vo_params *p = &s->p;
    if (p->thermostat == 1) {
        p->c2 = sqrt((1.0 - p->c1 * p->c1) * temperature);
    } else if (p->thermostat == 2) {
        double dof = 3.0 * (double)p->count - 3.0;
        p->nh_gkt = dof * temperature;
        p->nh_q = p->nh_gkt * p->nh_period * p->nh_period / (4.0 * (VO_PI * VO_PI));
    }
}

/* Pair terms r.F and U at r2 < rc2, in FP64 whatever the force precision
 * (the R5 sequence). */
static void pair_terms(const vo_params *p, double r2, double *rf_out, double *u_out)
{
    double inv2 = 1.0 / r2;
    double s6 = inv2 * inv2 * inv2;
    double rf = s6 * fma(48.0, s6, -24.0);
    double t = 4.0 * s6;
    double u = fma(t, s6, -t) - p->shift;
    if (p->shifted_force) {
        double r = sqrt(r2);
        rf = fma(-r, p->force_shift, rf);
        u = fma(r - p->rc, p->force_shift, u);
    }
    *rf_out = rf;
    *u_out = u;
}

/* R10 profiles along z (lennard_jones.wgsl:628-730): per slab the count,
 * m v_a^2, the Irving-Kirkwood virial diagonal W_aa and the potential, then
 * the whole-box W_aa. A pair's half of the virial, w_a = (0.5 d_a) d_a q with
 * q = rf (1 / r2), is spread over the slabs the segment from z_i to z_i + d_z
 * crosses, in proportion to the length inside each, the last slab taking
 * what is left, starting from slab floor(lo iw), lo the segment's lower end
 * and iw = 1 / width. Slabs use the wrapped z; hb = floor(z_i iw) is atom
 * i's slab. Atom i's shares to the slabs hb - 1, hb and hb + 1 (every share,
 * when slabs are at least a cutoff wide) are summed per slab and axis in
 * FP64, in list order, and each of the nine sums goes to fixed point once,
 * llrint(x 2^32); a share to any other slab goes to fixed point on its own.
 * Each per-atom value (count, v_a^2, potential, whole-box W_aa) goes to fixed
 * point on its own, and the integers are summed, so the totals do not depend
 * on the order of atoms (the reference adds each contribution with
 * fixed_add). The pair terms are FP64 in both precision forms. The walk over
 * slabs stops after bins + 2 slabs, more than a segment shorter than the box
 * can cross. */
static inline void prof_add(int64_t *prof, uint32_t slab, int q, double x)
{
    prof[slab * 8u + (uint32_t)q] += llrint(x * TWO32);
}

/* Share `share` of w (one value per axis) to the slab k - hb = rel from atom
 * i's: into the atom's FP64 window sums when rel is -1, 0 or 1, else to
 * fixed point on its own. */
static inline void prof_share(int64_t *prof, double win[3][3], int rel, uint32_t slab, const double w[3],
                              double share)
{
    for (int a = 0; a < 3; a++) {
        const double f = w[a] * share;
        if (rel >= -1 && rel <= 1)
            win[rel + 1][a] += f;
        else
            prof_add(prof, slab, 4 + a, f);
    }
}

void vo_set_profile_bins(vo_sim *s, uint32_t bins)
{
    free(s->prof);
    s->prof_bins = bins;
    s->prof = bins ? (int64_t *)xcalloc((size_t)bins * 8 + 3, sizeof(int64_t)) : NULL;
    s->prof_samples = 0;
}

void vo_clear_profile(vo_sim *s)
{
    if (s->prof)
        memset(s->prof, 0, ((size_t)s->prof_bins * 8 + 3) * sizeof(int64_t));
    s->prof_samples = 0;
}

void vo_sample_profile(vo_sim *s)
{
    vo_prime(s);
    const vo_params *p = &s->p;
    const uint32_t bins = s->prof_bins;
    if (!bins)
        return;
    int64_t *prof = s->prof;
    const double width = p->box[2] / (double)bins;
    const double iw = 1.0 / width;
    for (uint32_t i = 0; i < p->count; i++) {
        const double xi = s->x[0][i], yi = s->x[1][i], zraw = s->x[2][i];
        const double zi = wrap_step(zraw, p->box[2]);
        int hb = (int)floor(zi * iw);
        const uint32_t home = (uint32_t)(hb < 0 ? 0 : (hb > (int)bins - 1 ? (int)bins - 1 : hb));
        prof_add(prof, home, 0, 1.0);
        for (int k = 0; k < 3; k++)
            prof_add(prof, home, 1 + k, s->v[k][i] * s->v[k][i]);
        double pe = 0.0, total[3] = {0.0, 0.0, 0.0};
        double win[3][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}}; /* slabs hb - 1, hb, hb + 1 */
        const uint32_t cnt = s->nbr_count[i];
        const uint32_t *row = s->nbr + (size_t)i * p->capacity;
        const uint8_t *img = s->nbr_image + (size_t)i * p->capacity;
        for (uint32_t k = 0; k < cnt; k++) {
            const uint32_t j = row[k];
            const int code = img[k];
            const double d[3] = {(s->x[0][j] + (code % 3 - 1) * p->box[0]) - xi,
                                 (s->x[1][j] + ((code / 3) % 3 - 1) * p->box[1]) - yi,
                                 (s->x[2][j] + (code / 9 - 1) * p->box[2]) - zraw};
            const double r2 = fma(d[2], d[2], fma(d[1], d[1], d[0] * d[0]));
            if (!(r2 < p->rc2))
                continue;
            double rf, u;
            pair_terms(p, r2, &rf, &u);
            pe += 0.5 * u;
            const double q = rf * (1.0 / r2);
            double w[3];
            for (int a = 0; a < 3; a++) {
                w[a] = (0.5 * d[a]) * d[a] * q;
                total[a] += w[a];
            }
            const double lo = zi + (d[2] < 0.0 ? d[2] : 0.0);
            const double hi = zi + (d[2] > 0.0 ? d[2] : 0.0);
            const double length = hi - lo;
            if (length < 1e-6 * width) {
                prof_share(prof, win, (int)home - hb, home, w, 1.0);
                continue;
            }
            int b = (int)floor(lo * iw);
            double left = 1.0;
            for (uint32_t walk = 0; walk < bins + 2u; walk++) {
                const int last = (double)(b + 1) * width >= hi;
                const double start = lo > (double)b * width ? lo : (double)b * width;
                const double end = hi < (double)(b + 1) * width ? hi : (double)(b + 1) * width;
                double share = last ? left : (end - start) / length;
                share = share > 0.0 ? share : 0.0;
                left -= share;
                const uint32_t slab = (uint32_t)(((b % (int)bins) + (int)bins) % (int)bins);
                prof_share(prof, win, b - hb, slab, w, share);
                if (last)
                    break;
                b++;
            }
        }
        for (int kk = 0; kk < 3; kk++) {
            const uint32_t slab = (uint32_t)((((hb - 1 + kk) % (int)bins) + (int)bins) % (int)bins);
            for (int a = 0; a < 3; a++)
                prof_add(prof, slab, 4 + a, win[kk][a]);
        }
        prof_add(prof, home, 7, pe);
        for (int a = 0; a < 3; a++)
            prof[(size_t)bins * 8u + (uint32_t)a] += llrint(total[a] * TWO32);
    }
    s->prof_samples++;
}

void vo_read_profile(const vo_sim *s, int64_t *raw, uint64_t *samples)
{
    if (raw && s->prof)
        memcpy(raw, s->prof, ((size_t)s->prof_bins * 8 + 3) * sizeof(int64_t));
    if (samples)
        *samples = s->prof_samples;
}

/* R9, lennard_jones.wgsl:576-626 and lennard_jones.rs:1209-1246 */
void vo_measure(vo_sim *s, vo_measurement *m)
{
    vo_prime(s);
    const vo_params *p = &s->p;
    int32_t n = (int32_t)p->count;
    int64_t ke = 0, pe = 0, w = 0, mx = 0, my = 0, mz = 0, cond = 0, within = 0;
#pragma omp parallel for schedule(static) reduction(+ : ke, pe, w, mx, my, mz, cond, within)
    for (int32_t i = 0; i < n; i++) {
        double acc[3];
        interactions(s, i, acc, 1);
        double vx = s->v[0][i], vy = s->v[1][i], vz = s->v[2][i];
        double kei = 0.5 * fma(vz, vz, fma(vy, vy, vx * vx));
        ke += vo_fixed_f64(kei, p->energy_fixed);
        pe += vo_fixed_f64(s->pe[i], p->energy_fixed);
        w += vo_fixed_f64(s->w[i], p->energy_fixed);
        mx += vo_fixed_f64(vx, p->momentum_fixed);
        my += vo_fixed_f64(vy, p->momentum_fixed);
        mz += vo_fixed_f64(vz, p->momentum_fixed);
        cond += s->near[i] >= p->near_count ? 1 : 0;
        within += s->within[i];
    }
    memset(m, 0, sizeof(*m));
    m->raw[0] = ke;
    m->raw[1] = pe;
    m->raw[2] = w;
    m->raw[3] = mx;
    m->raw[4] = my;
    m->raw[5] = mz;
    m->raw[6] = cond;
    m->raw[7] = within;
    m->step = s->steps;
    m->kinetic = (double)ke / p->energy_fixed;
    m->potential = (double)pe / p->energy_fixed;
    m->virial = (double)w / p->energy_fixed;
    m->momentum[0] = (double)mx / p->momentum_fixed;
    m->momentum[1] = (double)my / p->momentum_fixed;
    m->momentum[2] = (double)mz / p->momentum_fixed;
    m->count = p->count;
    m->volume = p->box[0] * p->box[1] * p->box[2];
    m->list_rebuilds = s->rebuilds;
    m->thermostat_energy = p->thermostat == 2 ? 0.5 * p->nh_q * s->xi * s->xi + p->nh_gkt * s->eta : 0.0;
    m->list_overflow = s->overflow;
    m->condensed = (uint32_t)cond;
    m->pairs_within_cutoff = (uint64_t)within / 2u;
}

void vo_read_state(const vo_sim *s, double *pos, double *vel, double *acc, uint32_t *orig)
{
    uint32_t n = s->p.count;
    for (uint32_t i = 0; i < n; i++)
        for (int k = 0; k < 3; k++) {
            if (pos)
                pos[3 * i + k] = s->x[k][i];
            if (vel)
                vel[3 * i + k] = s->v[k][i];
            if (acc)
                acc[3 * i + k] = s->a[k][i];
        }
    if (orig)
        memcpy(orig, s->orig, (size_t)n * 4);
}

/* lennard_jones.rs:1354-1374: original order, positions wrapped into the box */
void vo_read_particles(const vo_sim *s, double *pos, double *vel)
{
    uint32_t n = s->p.count;
    for (uint32_t t = 0; t < n; t++) {
        uint32_t i = s->orig[t];
        for (int k = 0; k < 3; k++) {
            pos[3 * i + k] = wrap_step(s->x[k][t], s->p.box[k]);
            vel[3 * i + k] = s->v[k][t];
        }
    }
}

void vo_read_cells(const vo_sim *s, uint32_t *cell_of, uint32_t *cell_start, uint32_t *sorted)
{
    uint32_t n = s->p.count;
    if (cell_of)
        memcpy(cell_of, s->cell_of, (size_t)n * 4);
    if (cell_start)
        memcpy(cell_start, s->cell_start, ((size_t)s->p.total_cells + 1) * 4);
    if (sorted)
        memcpy(sorted, s->sorted, (size_t)n * 4);
}

void vo_read_lists(const vo_sim *s, uint32_t *counts, uint32_t *entries)
{
    uint32_t n = s->p.count;
    if (counts)
        memcpy(counts, s->nbr_count, (size_t)n * 4);
    if (entries)
        memcpy(entries, s->nbr, (size_t)n * s->p.capacity * 4);
}

void vo_read_energies(const vo_sim *s, double *pe, double *w)
{
    uint32_t n = s->p.count;
    if (pe)
        memcpy(pe, s->pe, (size_t)n * sizeof(double));
    if (w)
        memcpy(w, s->w, (size_t)n * sizeof(double));
}

/* tests/coexistence.rs:39-53 (both cutoff styles) and tests/thermodynamics.rs:33-53,
 * all pairs i < j, nearest image by x - L round(x / L) (Rust round: half away
 * from zero). Forces on each atom: F_i = -sum_j d_ij rf / r2, d_ij = x_j - x_i. */
void vo_all_pairs(const double *pos, size_t n, const double box[3], double cutoff,
                  int shifted_force, double *pe_out, double *w_out, double *forces)
{
    double rc2 = cutoff * cutoff;
    double uc, rfc;
    {
        double s6 = 1.0 / (rc2 * rc2 * rc2);
        uc = 4.0 * (s6 * s6 - s6);
        rfc = 24.0 * (2.0 * s6 * s6 - s6);
    }
    double fc = rfc / cutoff;
    double pe = 0.0, w[3] = {0.0, 0.0, 0.0};
    if (forces)
        memset(forces, 0, n * 3 * sizeof(double));
    for (size_t i = 0; i < n; i++)
        for (size_t j = i + 1; j < n; j++) {
            double d[3];
            for (int k = 0; k < 3; k++) {
                double x = pos[3 * j + k] - pos[3 * i + k];
                d[k] = x - box[k] * round(x / box[k]);
            }
            double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
            if (r2 < rc2) {
                double s6 = 1.0 / (r2 * r2 * r2);
                double u = 4.0 * (s6 * s6 - s6) - uc;
                double rf = 24.0 * (2.0 * s6 * s6 - s6);
                if (shifted_force) {
                    double r = sqrt(r2);
                    u += (r - cutoff) * fc;
                    rf -= r * fc;
                }
                pe += u;
                for (int k = 0; k < 3; k++) {
                    w[k] += d[k] * d[k] * rf / r2;
                    if (forces) {
                        double f = d[k] * rf / r2;
                        forces[3 * i + k] -= f;
                        forces[3 * j + k] += f;
                    }
                }
            }
        }
    if (pe_out)
        *pe_out = pe;
    if (w_out)
        for (int k = 0; k < 3; k++)
            w_out[k] = w[k];
}

void vo_set_threads(int threads)
{
#ifdef _OPENMP
    omp_set_num_threads(threads);
#else
    (void)threads;
#endif
}

int vo_max_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
