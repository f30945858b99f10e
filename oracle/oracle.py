"""ctypes binding of the CPU oracle (liboracle_lj.so).

TEST INFRASTRUCTURE ONLY: imported by the test suite, the smoke check and the
CPU-baseline leg of the benchmark, to check the product's results. The
product package never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle_lj.so")


class Params(C.Structure):
    _fields_ = [
        ("box", C.c_double * 3),
        ("half_box", C.c_double * 3),
        ("cells", C.c_uint32 * 3),
        ("total_cells", C.c_uint32),
        ("cell_size", C.c_float * 3),
        ("count", C.c_uint32),
        ("capacity", C.c_uint32),
        ("capacity_reference", C.c_uint32),
        ("rc", C.c_double),
        ("rc2", C.c_double),
        ("list_r2", C.c_double),
        ("half_skin2", C.c_double),
        ("shift", C.c_double),
        ("force_shift", C.c_double),
        ("dt", C.c_double),
        ("half_dt", C.c_double),
        ("energy_scale", C.c_double),
        ("momentum_scale", C.c_double),
        ("energy_fixed", C.c_double),
        ("momentum_fixed", C.c_double),
        ("thermostat", C.c_int32),
        ("c1", C.c_double),
        ("c2", C.c_double),
        ("seed", C.c_uint64),
        ("precision", C.c_int32),
        ("shifted_force", C.c_int32),
        ("near_count", C.c_uint32),
        ("nh_gkt", C.c_double),
        ("nh_q", C.c_double),
        ("nh_period", C.c_double),
    ]


class Measurement(C.Structure):
    _fields_ = [
        ("step", C.c_uint64),
        ("kinetic", C.c_double),
        ("potential", C.c_double),
        ("virial", C.c_double),
        ("momentum", C.c_double * 3),
        ("count", C.c_uint32),
        ("volume", C.c_double),
        ("list_rebuilds", C.c_uint32),
        ("list_overflow", C.c_uint32),
        ("condensed", C.c_uint32),
        ("pairs_within_cutoff", C.c_uint64),
        ("raw", C.c_int64 * 8),
        ("thermostat_energy", C.c_double),
    ]


_lib = None


def build() -> None:
    """Compile liboracle_lj.so with its Makefile."""
    subprocess.run(["make", "-s", "-B", "liboracle_lj.so"], cwd=HERE, check=True)


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        d, u32, u64, i32, i64, f = C.c_double, C.c_uint32, C.c_uint64, C.c_int, C.c_int64, C.c_float
        P = C.c_void_p
        dp = C.POINTER(C.c_double)
        up = C.POINTER(C.c_uint32)
        sz = C.c_size_t
        L.vo_fixed_split.argtypes = [f, up, up]
        L.vo_fixed_to_raw.argtypes = [f]
        L.vo_fixed_to_raw.restype = i64
        L.vo_fixed_f64.argtypes = [d, d]
        L.vo_fixed_f64.restype = i64
        L.vo_scale_for.argtypes = [d]
        L.vo_scale_for.restype = d
        L.vo_lattice.argtypes = [sz, d, d, u64, dp, dp]
        L.vo_lattice_box.argtypes = [sz, dp, d, u64, dp, dp]
        L.vo_jitter.argtypes = [sz, u64, d, dp]
        L.vo_fcc.argtypes = [u32, d, d, u64, dp, dp]
        L.vo_fcc.restype = d
        L.vo_random_insertion.argtypes = [sz, dp, d, d, u64, dp, dp]
        L.vo_random_insertion.restype = i32
        L.vo_derive_params.argtypes = [dp, sz, dp, d, d, d, i32, i32, i32, d, d, d, u64, dp, C.POINTER(Params)]
        L.vo_derive_params.restype = i32
        L.vo_create.argtypes = [dp, dp, sz, dp, d, d, d, i32, i32, i32, d, d, d, u64]
        L.vo_set_temperature.argtypes = [P, d]
        L.vo_create.restype = P
        L.vo_destroy.argtypes = [P]
        L.vo_get_params.argtypes = [P, C.POINTER(Params)]
        L.vo_set_near_count.argtypes = [P, u32]
        L.vo_prime.argtypes = [P]
        L.vo_run.argtypes = [P, u64]
        L.vo_measure.argtypes = [P, C.POINTER(Measurement)]
        L.vo_read_state.argtypes = [P, dp, dp, dp, up]
        L.vo_read_particles.argtypes = [P, dp, dp]
        L.vo_read_cells.argtypes = [P, up, up, up]
        L.vo_read_lists.argtypes = [P, up, up]
        L.vo_read_energies.argtypes = [P, dp, dp]
        L.vo_rebuilds.argtypes = [P]
        L.vo_rebuilds.restype = u32
        L.vo_set_profile_bins.argtypes = [P, u32]
        L.vo_set_heat_exchange.argtypes = [P, i32, dp, dp, d]
        L.vo_set_heat_exchange.restype = i32
        L.vo_record_heat_current.argtypes = [P, u32]
        L.vo_sample_heat_current.argtypes = [P]
        L.vo_sample_heat_current.restype = i32
        L.vo_read_heat_current.argtypes = [P, C.POINTER(i64), C.POINTER(u32)]
        L.vo_clear_profile.argtypes = [P]
        L.vo_sample_profile.argtypes = [P]
        L.vo_read_profile.argtypes = [P, C.POINTER(i64), C.POINTER(u64)]
        L.vo_all_pairs.argtypes = [dp, sz, dp, d, i32, dp, dp, dp]
        L.vo_philox4x32.argtypes = [up, up, up]
        L.vo_langevin_noise.argtypes = [u64, u32, u32, u64, dp]
        L.vo_log.argtypes = [d]
        L.vo_log.restype = d
        L.vo_sincos_turns.argtypes = [d, dp, dp]
        L.vo_set_threads.argtypes = [i32]
        L.vo_set_capacity_override.argtypes = [u32]
        L.vo_max_threads.restype = i32
        _lib = L
    return _lib


def _dp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _up(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_uint32))


def fixed_split(x: float) -> tuple[int, int]:
    lo, hi = C.c_uint32(), C.c_uint32()
    lib().vo_fixed_split(C.c_float(x), C.byref(lo), C.byref(hi))
    return lo.value, hi.value


def fixed_to_raw(x: float) -> int:
    return lib().vo_fixed_to_raw(C.c_float(x))


def fixed_f64(x: float, fixed_scale: float) -> int:
    return lib().vo_fixed_f64(x, fixed_scale)


def scale_for(m: float) -> float:
    return lib().vo_scale_for(m)


def lattice(count: int, box_size: float, temperature: float, seed: int):
    pos = np.zeros((count, 3))
    vel = np.zeros((count, 3))
    lib().vo_lattice(count, box_size, temperature, seed, _dp(pos), _dp(vel))
    return pos, vel


def lattice_box(count: int, size, temperature: float, seed: int):
    pos = np.zeros((count, 3))
    vel = np.zeros((count, 3))
    s = np.ascontiguousarray(size, dtype=np.float64)
    lib().vo_lattice_box(count, _dp(s), temperature, seed, _dp(pos), _dp(vel))
    return pos, vel


def jitter(pos: np.ndarray, seed: int = 99, amplitude: float = 0.3) -> np.ndarray:
    out = np.ascontiguousarray(pos, dtype=np.float64).copy()
    lib().vo_jitter(out.shape[0], seed, amplitude, _dp(out))
    return out


def fcc(cells_per_side: int, density: float, temperature: float, seed: int):
    n = 4 * cells_per_side ** 3
    pos = np.zeros((n, 3))
    vel = np.zeros((n, 3))
    side = lib().vo_fcc(cells_per_side, density, temperature, seed, _dp(pos), _dp(vel))
    return pos, vel, side


def random_insertion(count: int, box, min_distance: float, temperature: float, seed: int):
    pos = np.zeros((count, 3))
    vel = np.zeros((count, 3))
    b = np.ascontiguousarray(box, dtype=np.float64)
    rc = lib().vo_random_insertion(count, _dp(b), min_distance, temperature, seed, _dp(pos), _dp(vel))
    if rc != 0:
        raise RuntimeError("random insertion failed")
    return pos, vel


def all_pairs(pos: np.ndarray, box, cutoff: float, shifted_force: bool, forces: bool = False):
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    b = np.ascontiguousarray(box, dtype=np.float64)
    pe = C.c_double()
    w = np.zeros(3)
    f = np.zeros((pos.shape[0], 3)) if forces else None
    lib().vo_all_pairs(_dp(pos), pos.shape[0], _dp(b), cutoff, int(shifted_force), C.byref(pe), _dp(w),
                       _dp(f) if forces else None)
    return (pe.value, w, f) if forces else (pe.value, w)


def derive_params(pos, vel, box, cutoff=2.5, skin=0.3, dt=0.005, shifted_force=False, precision=0,
                  thermostat=0, temperature=0.0, friction=0.0, seed=1, period=0.0) -> Params:
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    vel = np.ascontiguousarray(vel, dtype=np.float64)
    b = np.ascontiguousarray(box, dtype=np.float64)
    p = Params()
    rc = lib().vo_derive_params(_dp(pos), pos.shape[0], _dp(b), cutoff, skin, dt, int(shifted_force), precision,
                                thermostat, temperature, friction, period, seed, _dp(vel), C.byref(p))
    if rc != 0:
        raise ValueError("the box must be at least two (cutoff + skin) wide on every axis")
    return p


def philox4x32(ctr, key):
    c = np.ascontiguousarray(ctr, dtype=np.uint32)
    k = np.ascontiguousarray(key, dtype=np.uint32)
    out = np.zeros(4, dtype=np.uint32)
    lib().vo_philox4x32(_up(c), _up(k), _up(out))
    return out


def langevin_noise(seed: int, replica: int, atom: int, step: int) -> np.ndarray:
    out = np.zeros(3)
    lib().vo_langevin_noise(seed, replica, atom, step, _dp(out))
    return out


def log(x: float) -> float:
    return lib().vo_log(x)


def sincos_turns(u: float):
    s, c = C.c_double(), C.c_double()
    lib().vo_sincos_turns(u, C.byref(s), C.byref(c))
    return s.value, c.value


class capacity_override:
    """List capacity in place of the rule for sims created inside the block."""

    def __init__(self, capacity: int):
        self.capacity = capacity

    def __enter__(self):
        lib().vo_set_capacity_override(self.capacity)
        return self

    def __exit__(self, *exc):
        lib().vo_set_capacity_override(0)


def set_threads(n: int) -> None:
    lib().vo_set_threads(n)


def max_threads() -> int:
    return lib().vo_max_threads()


class Sim:
    """The oracle's simulation, mirroring the product's step/measure API."""

    def __init__(self, pos, vel, box, cutoff=2.5, skin=0.3, dt=0.005, shifted_force=False, precision=0,
                 thermostat=0, temperature=0.0, friction=0.0, seed=1, period=0.0):
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        vel = np.ascontiguousarray(vel, dtype=np.float64)
        b = np.ascontiguousarray(box, dtype=np.float64)
        self.n = pos.shape[0]
        self._h = lib().vo_create(_dp(pos), _dp(vel), self.n, _dp(b), cutoff, skin, dt, int(shifted_force),
                                  precision, thermostat, temperature, friction, period, seed)
        if not self._h:
            raise ValueError("the box must be at least two (cutoff + skin) wide on every axis")
        self.params = Params()
        lib().vo_get_params(self._h, C.byref(self.params))

    def close(self):
        if self._h:
            lib().vo_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def set_temperature(self, t: float):
        lib().vo_set_temperature(self._h, t)
        lib().vo_get_params(self._h, C.byref(self.params))

    def set_near_count(self, k: int):
        lib().vo_set_near_count(self._h, k)

    def prime(self):
        lib().vo_prime(self._h)

    def run(self, steps: int):
        lib().vo_run(self._h, steps)

    def measure(self) -> Measurement:
        m = Measurement()
        lib().vo_measure(self._h, C.byref(m))
        return m

    def state(self):
        n = self.n
        pos, vel, acc = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros((n, 3))
        orig = np.zeros(n, dtype=np.uint32)
        lib().vo_read_state(self._h, _dp(pos), _dp(vel), _dp(acc), _up(orig))
        return pos, vel, acc, orig

    def read_particles(self):
        n = self.n
        pos, vel = np.zeros((n, 3)), np.zeros((n, 3))
        lib().vo_read_particles(self._h, _dp(pos), _dp(vel))
        return pos, vel

    def cells(self):
        n = self.n
        cell_of = np.zeros(n, dtype=np.uint32)
        start = np.zeros(self.params.total_cells + 1, dtype=np.uint32)
        sorted_ = np.zeros(n, dtype=np.uint32)
        lib().vo_read_cells(self._h, _up(cell_of), _up(start), _up(sorted_))
        return cell_of, start, sorted_

    def lists(self):
        n, cap = self.n, self.params.capacity
        counts = np.zeros(n, dtype=np.uint32)
        entries = np.zeros((n, cap), dtype=np.uint32)
        lib().vo_read_lists(self._h, _up(counts), _up(entries))
        return counts, entries

    def energies(self):
        pe, w = np.zeros(self.n), np.zeros(self.n)
        lib().vo_read_energies(self._h, _dp(pe), _dp(w))
        return pe, w

    def rebuilds(self) -> int:
        return lib().vo_rebuilds(self._h)

    # Heat exchange and heat current (lennard_jones.wgsl:732-947)
    def set_heat_exchange(self, hot, cold, energy):
        on = hot is not None
        h = np.ascontiguousarray(hot if on else [0.0, 0.0], dtype=np.float64)
        c = np.ascontiguousarray(cold if on else [0.0, 0.0], dtype=np.float64)
        if lib().vo_set_heat_exchange(self._h, int(on), _dp(h), _dp(c), float(energy)) != 0:
            raise ValueError("heat exchange runs at constant energy, without a thermostat")

    def record_heat_current(self, capacity: int):
        self.hc_capacity = int(capacity)
        lib().vo_record_heat_current(self._h, self.hc_capacity)

    def sample_heat_current(self):
        if lib().vo_sample_heat_current(self._h) != 0:
            raise ValueError("heat current series full or not recorded")

    def read_heat_current(self):
        """Raw (samples, 3) int64 sums, value * 2^32."""
        raw = np.zeros((max(self.hc_capacity, 1), 3), dtype=np.int64)
        n = C.c_uint32(0)
        lib().vo_read_heat_current(self._h, raw.ctypes.data_as(C.POINTER(C.c_int64)), C.byref(n))
        return raw[: n.value]

    # Profiles along z (lennard_jones.wgsl:628-730): raw fixed-point sums,
    # bins * 8 + 3 of them, and the sample count.
    def set_profile_bins(self, bins: int):
        self.profile_bins = int(bins)
        lib().vo_set_profile_bins(self._h, self.profile_bins)

    def clear_profile(self):
        lib().vo_clear_profile(self._h)

    def sample_profile(self):
        lib().vo_sample_profile(self._h)

    def read_profile(self):
        raw = np.zeros(self.profile_bins * 8 + 3, dtype=np.int64)
        samples = C.c_uint64(0)
        lib().vo_read_profile(self._h, raw.ctypes.data_as(C.POINTER(C.c_int64)), C.byref(samples))
        return raw, samples.value
