"""Lennard-Jones particles in a periodic rectangular box, on an MI355X.

Python mirror of `virial_particles::lennard_jones`
(crates/virial-particles/src/lennard_jones.rs): the same types, builder,
step and measurement API and the same panic messages (raised here as
exceptions), over the C ABI of libvirial_lj.so (include/virial_lj.h). All
computation happens in the HIP library; this module only marshals.

Reduced units (sigma = epsilon = mass = k_B = 1). Pressure from the virial
theorem, P = (2 KE + W) / (3 V) (lennard_jones.rs:110-131).
"""
from __future__ import annotations

import ctypes as C
import os
import enum
import math
from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Union

import numpy as np

from . import _lib as L

PARTICLE_DTYPE = np.dtype([("pos", "<f8", (3,)), ("vel", "<f8", (3,))])
FP64 = 0
FP32 = 1


class LennardJonesError(RuntimeError):
    """A failed library call. `code` is the VLJ_E_* status."""

    def __init__(self, code: int, message: str):
        super().__init__(message)
        self.code = code


def _check(code: int, handle=None) -> None:
    if code != L.OK:
        raise LennardJonesError(code, L.lib().vlj_last_error(handle).decode())


@dataclass
class Particle:
    """A particle (mass 1) (lennard_jones.rs:25-30)."""

    pos: Sequence[float]
    vel: Sequence[float]


@dataclass
class Langevin:
    """Langevin thermostat, BAOAB splitting (lennard_jones.rs:32-40)."""

    temperature: float
    friction: float
    seed: int


@dataclass
class NoseHoover:
    """Nose-Hoover thermostat (lennard_jones.rs:42-52)."""

    temperature: float
    period: float


Thermostat = Optional[Union[Langevin, NoseHoover]]


@dataclass
class HeatExchange:
    """Heat exchange between a hot and a cold slab along z
    (lennard_jones.rs:376-387): `energy` moved per step."""

    hot: Sequence[float]
    cold: Sequence[float]
    energy: float


@dataclass
class Settings:
    """Box, cutoff, time step and thermostat (lennard_jones.rs:63-82), plus
    the pair-force precision (FP64, or FP32 on FP64 displacements)."""

    box_size: Sequence[float]
    cutoff: float
    dt: float
    thermostat: Thermostat = None
    shifted_force: bool = False
    skin: float = 0.3
    profile_bins: int = 0
    precision: int = FP64

    def to_c(self) -> L.Settings:
        s = L.Settings()
        for k in range(3):
            s.box_size[k] = float(self.box_size[k])
        s.cutoff = self.cutoff
        s.dt = self.dt
        s.skin = self.skin
        t = self.thermostat
        if t is None:
            s.thermostat.kind = 0
        elif isinstance(t, Langevin):
            s.thermostat.kind = 1
            s.thermostat.temperature = t.temperature
            s.thermostat.friction = t.friction
            s.thermostat.seed = int(t.seed)
        elif isinstance(t, NoseHoover):
            s.thermostat.kind = 2
            s.thermostat.temperature = t.temperature
            s.thermostat.period = t.period
        else:
            raise TypeError("thermostat must be None, Langevin or NoseHoover")
        s.shifted_force = 1 if self.shifted_force else 0
        s.profile_bins = int(self.profile_bins)
        s.precision = int(self.precision)
        return s

    @staticmethod
    def from_c(s: L.Settings) -> "Settings":
        t = s.thermostat
        if t.kind == 1:
            thermostat = Langevin(t.temperature, t.friction, int(t.seed))
        elif t.kind == 2:
            thermostat = NoseHoover(t.temperature, t.period)
        else:
            thermostat = None
        return Settings(list(s.box_size), s.cutoff, s.dt, thermostat, bool(s.shifted_force), s.skin,
                        int(s.profile_bins), int(s.precision))


def _load_handle(path) -> C.c_void_p:
    h = C.c_void_p()
    _check(L.lib().vlj_load(os.fsencode(str(path)), C.byref(h)))
    return h


def _replica_info(h, replica: int):
    st, n, reps = L.Settings(), C.c_uint64(), C.c_size_t()
    _check(L.lib().vlj_replica_info(h, replica, C.byref(st), C.byref(n), C.byref(reps)), h)
    return Settings.from_c(st), int(n.value), int(reps.value)


@dataclass
class Measurement:
    """Totals at one step (lennard_jones.rs:84-131)."""

    step: int
    time: float
    kinetic: float
    potential: float
    virial: float
    momentum: List[float]
    count: int
    volume: float
    list_rebuilds: int
    list_overflow: int
    thermostat_energy: float
    condensed: int
    pairs_within_cutoff: int = 0
    raw: List[int] = field(default_factory=list)

    @classmethod
    def from_c(cls, m: L.Measurement) -> "Measurement":
        return cls(m.step, m.time, m.kinetic, m.potential, m.virial, list(m.momentum), m.count, m.volume,
                   m.list_rebuilds, m.list_overflow, m.thermostat_energy, m.condensed, m.pairs_within_cutoff,
                   list(m.raw))

    def energy(self) -> float:
        return self.kinetic + self.potential

    def temperature(self, dof: float) -> float:
        return 2.0 * self.kinetic / dof

    def pressure(self) -> float:
        return (2.0 * self.kinetic + self.virial) / (3.0 * self.volume)

    def extended_energy(self) -> float:
        return self.energy() + self.thermostat_energy


PROFILE_SUMS = 8  # sums per slab in the profile buffer (lennard_jones.rs:321)


@dataclass
class Slab:
    """One slab of a profile, averaged over samples (lennard_jones.rs:133-147)."""

    z: float
    particles: float
    kinetic: List[float]   # sum of m v_a^2 over the slab's particles
    virial: List[float]    # diagonal of the virial tensor, spread along each pair's line (Irving-Kirkwood)
    potential: float       # half of each pair's energy to each particle's slab


class Drawn(enum.IntEnum):
    """What measurements and render() leave in the w of each velocity
    (lennard_jones.rs:365-373)."""

    NEIGHBOURS = 0  # neighbours within 1.5 sigma, a local density (the default)
    TEMPERATURE = 1  # m v^2 / 3 over the molecule and its neighbours within the cutoff


def _render_methods(cls):
    """positions()/velocities() for drawing (lennard_jones.rs:1041-1051, 885-889)."""

    def set_drawn(self, drawn: Drawn) -> None:
        _check(L.lib().vlj_set_drawn(self._h, int(drawn)), self._h)

    def render(self) -> None:
        """Fills the float4 render arrays on the device (asynchronous)."""
        _check(L.lib().vlj_render(self._h), self._h)

    def render_buffers(self):
        """Device addresses of the float4 render arrays (positions,
        velocities), one float4 per particle in slot order."""
        p, v = C.c_void_p(), C.c_void_p()
        _check(L.lib().vlj_render_buffers(self._h, C.byref(p), C.byref(v)), self._h)
        return p.value, v.value

    def read_render(self):
        """Host copies of the render arrays, (n, 4) float32 each, slot order."""
        n = self.count() if hasattr(self, "count") else sum(self.sizes)
        pos = np.zeros((n, 4), dtype=np.float32)
        vel = np.zeros((n, 4), dtype=np.float32)
        fp = C.POINTER(C.c_float)
        _check(L.lib().vlj_read_render(self._h, pos.ctypes.data_as(fp), vel.ctypes.data_as(fp), n), self._h)
        return pos, vel

    for f in (set_drawn, render, render_buffers, read_render):
        setattr(cls, f.__name__, f)
    return cls


class Phase(enum.Enum):
    """Where a slab sits in a liquid-vapour profile (lennard_jones.rs:149-156)."""

    LIQUID = "liquid"
    VAPOUR = "vapour"
    INTERFACE = "interface"
    BETWEEN = "between"


def _round_half_away(x: float) -> float:
    # f64::round: halves away from zero
    return math.copysign(math.floor(abs(x) + 0.5), x)


@dataclass
class Profile:
    """Profiles along z, averaged over samples (lennard_jones.rs:158-271);
    from LennardJones.read_profile."""

    samples: int
    box_size: List[float]
    slabs: List[Slab]
    virial: List[float]    # whole-box virial diagonal, summed pair by pair

    @classmethod
    def from_sums(cls, sums: Sequence[float], samples: int, box_size: Sequence[float]) -> "Profile":
        """From the profile_bins * 8 + 3 averaged sums of vlj_read_profile
        (lennard_jones.rs:1297-1325)."""
        bins = (len(sums) - 3) // PROFILE_SUMS
        width = box_size[2] / bins
        slabs = []
        for b in range(bins):
            q = sums[b * PROFILE_SUMS:(b + 1) * PROFILE_SUMS]
            slabs.append(Slab((b + 0.5) * width, float(q[0]), [float(q[1]), float(q[2]), float(q[3])],
                              [float(q[4]), float(q[5]), float(q[6])], float(q[7])))
        t = sums[bins * PROFILE_SUMS:]
        return cls(int(samples), [float(x) for x in box_size], slabs, [float(t[0]), float(t[1]), float(t[2])])

    def slab_volume(self) -> float:
        lx, ly, lz = self.box_size
        return lx * ly * lz / len(self.slabs)

    def density(self) -> List[float]:
        """Number density in each slab."""
        v = self.slab_volume()
        return [s.particles / v for s in self.slabs]

    def temperature(self) -> List[float]:
        """Kinetic temperature in each slab (3 degrees of freedom per particle)."""
        return [sum(s.kinetic) / (3.0 * s.particles) if s.particles else math.nan for s in self.slabs]

    def pressure_tensor(self) -> List[List[float]]:
        """Diagonal (P_xx, P_yy, P_zz) of the pressure tensor in each slab."""
        v = self.slab_volume()
        return [[(s.kinetic[a] + s.virial[a]) / v for a in range(3)] for s in self.slabs]

    def total_pressure_tensor(self) -> List[float]:
        """Diagonal of the whole box's pressure tensor."""
        volume = self.box_size[0] * self.box_size[1] * self.box_size[2]
        return [(sum(s.kinetic[a] for s in self.slabs) + self.virial[a]) / volume for a in range(3)]

    def interfaces(self, midpoint: float) -> List[float]:
        """Heights where the density crosses `midpoint`, interpolated between
        slab centres around the periodic z axis."""
        density = self.density()
        n = len(density)
        width = self.box_size[2] / n
        out = []
        for b in range(n):
            a, c = density[b] - midpoint, density[(b + 1) % n] - midpoint
            if (a < 0.0) != (c < 0.0):
                z = self.slabs[b].z + width * a / (a - c)
                out.append(z % self.box_size[2])
        return out

    def phases(self, midpoint: float, margin: float) -> List[Phase]:
        """Liquid or vapour at least `margin` from every interface, interface
        within margin / 2 of one, between otherwise; all between without an
        interface."""
        lz = self.box_size[2]
        interfaces = self.interfaces(midpoint)
        out = []
        for s, rho in zip(self.slabs, self.density()):
            distance = math.inf
            for z in interfaces:
                d = s.z - z
                distance = min(distance, abs(d - lz * _round_half_away(d / lz)))
            if math.isinf(distance):
                out.append(Phase.BETWEEN)
            elif distance >= margin:
                out.append(Phase.LIQUID if rho > midpoint else Phase.VAPOUR)
            elif distance <= margin / 2.0:
                out.append(Phase.INTERFACE)
            else:
                out.append(Phase.BETWEEN)
        return out

    def mean_in(self, values: Sequence[float], phases: Sequence[Phase], phase: Phase) -> Optional[float]:
        """Mean of a per-slab quantity over the slabs in `phase`; None if none."""
        picked = [v for v, p in zip(values, phases) if p == phase]
        return sum(picked) / len(picked) if picked else None

    def surface_tension(self) -> float:
        """gamma = L_z / 2 (P_zz - (P_xx + P_yy) / 2) for a slab spanning x
        and y, the 2 for its two interfaces."""
        pxx, pyy, pzz = self.total_pressure_tensor()
        return 0.5 * self.box_size[2] * (pzz - 0.5 * (pxx + pyy))


def _as_particle_array(particles) -> np.ndarray:
    """Particles as a C-contiguous array of vlj_particle (48 bytes each)."""
    if isinstance(particles, np.ndarray) and particles.dtype == PARTICLE_DTYPE:
        return np.ascontiguousarray(particles)
    if isinstance(particles, tuple) and len(particles) == 2:
        pos, vel = (np.asarray(a, dtype=np.float64).reshape(-1, 3) for a in particles)
        out = np.zeros(pos.shape[0], dtype=PARTICLE_DTYPE)
        out["pos"] = pos
        out["vel"] = vel
        return out
    out = np.zeros(len(particles), dtype=PARTICLE_DTYPE)
    for i, p in enumerate(particles):
        out["pos"][i] = p.pos
        out["vel"][i] = p.vel
    return out


def particles_from_arrays(pos, vel) -> np.ndarray:
    return _as_particle_array((pos, vel))


def _pptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(L.Particle))


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _uptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_uint32))


def to_particles(a: np.ndarray) -> List[Particle]:
    return [Particle(list(map(float, r["pos"])), list(map(float, r["vel"]))) for r in a]


@_render_methods
class LennardJones:
    """Lennard-Jones particles on the GPU (lennard_jones.rs:332-1375).

    `particles` is a list of Particle, a (pos, vel) pair of (n, 3) arrays or
    a PARTICLE_DTYPE array. The reference's constructor also takes the wgpu
    device and queue; here the current HIP device is used.
    """

    def __init__(self, particles, settings: Settings):
        arr = _as_particle_array(particles)
        self._settings = settings
        h = C.c_void_p()
        cs = settings.to_c()
        _check(L.lib().vlj_create(_pptr(arr), arr.shape[0], C.byref(cs), C.byref(h)))
        self._h = h

    def save(self, path) -> None:
        """Writes the complete state, after the steps taken so far, for an
        exact resume with LennardJones.load (no counterpart in the
        reference)."""
        _check(L.lib().vlj_save(self._h, os.fsencode(str(path))), self._h)

    @classmethod
    def load(cls, path) -> "LennardJones":
        """A handle from a file written by save(); it steps on to the same
        bits as the handle that saved it."""
        self = cls.__new__(cls)
        self._h = _load_handle(path)
        self._settings, _, reps = _replica_info(self._h, 0)
        if reps != 1:
            self.close()
            raise LennardJonesError(L.E_INVALID, f"{path}: a batch state; load it with Batch.load")
        return self

    def close(self) -> None:
        if getattr(self, "_h", None):
            L.lib().vlj_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _c(self, code: int) -> None:
        _check(code, self._h)

    # ---- stepping (lennard_jones.rs:1092-1138, 1274-1284)
    def step(self, steps: int) -> None:
        self._c(L.lib().vlj_step(self._h, steps))

    def run(self, steps: int) -> None:
        self._c(L.lib().vlj_run(self._h, steps))

    # ---- measurements (lennard_jones.rs:1141-1260)
    def measure(self) -> bool:
        ok = C.c_int32()
        self._c(L.lib().vlj_measure(self._h, C.byref(ok)))
        return bool(ok.value)

    def after_submit(self) -> None:
        """Kept for call-site shape: the library's read-back needs no mapping step."""

    def poll(self) -> List[Measurement]:
        out = (L.Measurement * 3)()
        n = C.c_size_t()
        self._c(L.lib().vlj_poll(self._h, out, 3, C.byref(n)))
        return [Measurement.from_c(out[k]) for k in range(n.value)]

    def measure_now(self) -> Measurement:
        m = L.Measurement()
        self._c(L.lib().vlj_measure_now(self._h, C.byref(m)))
        return Measurement.from_c(m)

    # ---- state (lennard_jones.rs:1354-1374)
    def read_particle_array(self) -> np.ndarray:
        out = np.zeros(self.count(), dtype=PARTICLE_DTYPE)
        self._c(L.lib().vlj_read_particles(self._h, _pptr(out), out.shape[0]))
        return out

    def read_particles(self) -> List[Particle]:
        return to_particles(self.read_particle_array())

    # ---- accessors (lennard_jones.rs:870-1030)
    def count(self) -> int:
        return L.lib().vlj_count(self._h)

    def box_size(self) -> List[float]:
        out = (C.c_double * 3)()
        self._c(L.lib().vlj_box_size(self._h, out))
        return list(out)

    def volume(self) -> float:
        return L.lib().vlj_volume(self._h)

    def steps(self) -> int:
        return L.lib().vlj_steps(self._h)

    def time(self) -> float:
        return L.lib().vlj_time(self._h)

    def target_temperature(self) -> Optional[float]:
        t = C.c_double()
        has = C.c_int32()
        self._c(L.lib().vlj_target_temperature(self._h, C.byref(t), C.byref(has)))
        return t.value if has.value else None

    def set_temperature(self, temperature: float) -> None:
        self._c(L.lib().vlj_set_temperature(self._h, temperature))

    def set_condensed_neighbours(self, neighbours: int) -> None:
        self._c(L.lib().vlj_set_condensed_neighbours(self._h, neighbours))

    # ---- profiles (lennard_jones.rs:1167-1191, 1297-1325)
    def sample_profile(self) -> None:
        self._c(L.lib().vlj_sample_profile(self._h))

    def sample_profile_now(self) -> None:
        self.sample_profile()

    def clear_profile(self) -> None:
        self._c(L.lib().vlj_clear_profile(self._h))

    # ---- heat exchange and heat current (lennard_jones.rs:927-1012)
    def set_heat_exchange(self, exchange: Optional[HeatExchange]) -> None:
        on = exchange is not None
        hot = np.ascontiguousarray(exchange.hot if on else [0.0, 0.0], dtype=np.float64)
        cold = np.ascontiguousarray(exchange.cold if on else [0.0, 0.0], dtype=np.float64)
        self._c(L.lib().vlj_set_heat_exchange(self._h, int(on), _dptr(hot), _dptr(cold),
                                              float(exchange.energy) if on else 0.0))

    def heat_flux(self) -> Optional[float]:
        flux, has = C.c_double(), C.c_int32()
        self._c(L.lib().vlj_heat_flux(self._h, C.byref(flux), C.byref(has)))
        return flux.value if has.value else None

    def record_heat_current(self, capacity: int) -> None:
        self._c(L.lib().vlj_record_heat_current(self._h, int(capacity)))

    def sample_heat_current(self) -> None:
        self._c(L.lib().vlj_sample_heat_current(self._h))

    def _hc_samples(self) -> int:
        n = C.c_uint32()
        self._c(L.lib().vlj_read_heat_current_raw(self._h, None, 0, C.byref(n)))
        return n.value

    def read_heat_current(self) -> List[List[float]]:
        cap = max(self._hc_samples(), 1)
        out = np.zeros(3 * cap)
        n = C.c_uint32()
        self._c(L.lib().vlj_read_heat_current(self._h, _dptr(out), cap, C.byref(n)))
        return out[: 3 * n.value].reshape(-1, 3).tolist()

    def read_heat_current_raw(self) -> np.ndarray:
        cap = max(self._hc_samples(), 1)
        out = np.zeros((cap, 3), dtype=np.int64)
        n = C.c_uint32()
        self._c(L.lib().vlj_read_heat_current_raw(self._h, out.ctypes.data_as(C.POINTER(C.c_int64)), cap,
                                                  C.byref(n)))
        return out[: n.value]

    def _profile_size(self) -> int:
        return int(self._settings.profile_bins) * PROFILE_SUMS + 3

    def read_profile(self) -> Profile:
        """The profiles averaged over every sample since the last clear."""
        out = np.zeros(self._profile_size())
        samples = C.c_uint64(0)
        self._c(L.lib().vlj_read_profile(self._h, _dptr(out), out.shape[0], C.byref(samples)))
        return Profile.from_sums(out, samples.value, self.box_size())

    def read_profile_raw(self):
        """The 64-bit fixed-point sums (value * 2^32) and the sample count."""
        out = np.zeros(self._profile_size(), dtype=np.int64)
        samples = C.c_uint64(0)
        self._c(L.lib().vlj_read_profile_raw(self._h, out.ctypes.data_as(C.POINTER(C.c_int64)), out.shape[0],
                                             C.byref(samples)))
        return out, samples.value

    # ---- inspection (parity tests; no counterpart in the reference)
    def debug_state(self):
        n = self.count()
        pos, vel, acc = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros((n, 3))
        orig = np.zeros(n, dtype=np.uint32)
        self._c(L.lib().vlj_debug_state(self._h, _dptr(pos), _dptr(vel), _dptr(acc), _uptr(orig)))
        return pos, vel, acc, orig

    def debug_grid(self) -> dict:
        out = np.zeros(6, dtype=np.uint32)
        self._c(L.lib().vlj_debug_grid(self._h, _uptr(out)))
        return {"cells": [int(x) for x in out[:3]], "total_cells": int(out[3]), "capacity": int(out[4]),
                "rebuilds": int(out[5])}

    def debug_counters(self) -> dict:
        return _debug_counters(self._h)

    def debug_params(self) -> dict:
        out = np.zeros(6)
        self._c(L.lib().vlj_debug_params(self._h, _dptr(out)))
        keys = ["energy_scale", "momentum_scale", "shift", "force_shift", "c1", "c2"]
        return dict(zip(keys, map(float, out)))

    def debug_cells(self):
        g = self.debug_grid()
        n = self.count()
        cell_of = np.zeros(n, dtype=np.uint32)
        start = np.zeros(g["total_cells"] + 1, dtype=np.uint32)
        sorted_ = np.zeros(n, dtype=np.uint32)
        self._c(L.lib().vlj_debug_cells(self._h, _uptr(cell_of), _uptr(start), _uptr(sorted_)))
        return cell_of, start, sorted_

    def debug_lists(self):
        g = self.debug_grid()
        n = self.count()
        counts = np.zeros(n, dtype=np.uint32)
        entries = np.zeros((n, g["capacity"]), dtype=np.uint32)
        self._c(L.lib().vlj_debug_lists(self._h, _uptr(counts), _uptr(entries)))
        return counts, entries

    def tune(self, steps_per_candidate: int = 6) -> int:
        """Times each candidate brick plan over real steps and keeps the
        fastest (vlj_tune); returns the steps it ran."""
        return _tune(self._h, steps_per_candidate)

    def plan_info(self) -> dict:
        """The fill of the brick plan in use, re-plans so far and the last
        tuning's trials."""
        return _plan_info(self._h)

    def timed_run(self, steps: int):
        """(force-kernel ms summed over the steps, total ms) from HIP events."""
        f, t = C.c_double(), C.c_double()
        self._c(L.lib().vlj_timed_run(self._h, steps, C.byref(f), C.byref(t)))
        return f.value, t.value


def _tune(h, steps_per_candidate: int) -> int:
    taken = C.c_uint64(0)
    _check(L.lib().vlj_tune(h, steps_per_candidate, C.byref(taken)), h)
    return taken.value


def _plan_info(h) -> dict:
    fill = C.c_double()
    replans = C.c_uint32()
    n = C.c_size_t()
    trials = (L.TuneTrial * 16)()
    _check(L.lib().vlj_plan_info(h, C.byref(fill), C.byref(replans), trials, 16, C.byref(n)), h)
    return {"fill": fill.value, "replans": replans.value,
            "trials": [{"fill": t.fill, "force_us": t.force_us, "build_us": t.build_us, "bricks": t.bricks}
                       for t in trials[: min(n.value, 16)]]}


def _debug_counters(h) -> dict:
    out = np.zeros(6, dtype=np.uint64)
    _check(L.lib().vlj_debug_counters(h, out.ctypes.data_as(C.POINTER(C.c_uint64))), h)
    return {"rebuilds": int(out[0]), "stage_overflow_builds": int(out[1]), "list_overflow": int(out[2]),
            "bricks": int(out[3]), "global_cell_builds": int(out[4]), "sparse_bricks": int(out[5])}


@_render_methods
class Batch:
    """Many independent replicas in the same launches (no counterpart in the
    reference). `replicas` is a list of (particles, Settings)."""

    def __init__(self, replicas):
        arrays = [_as_particle_array(p) for p, _ in replicas]
        self.sizes = [a.shape[0] for a in arrays]
        allp = np.concatenate(arrays) if arrays else np.zeros(0, dtype=PARTICLE_DTYPE)
        offsets = np.zeros(len(arrays) + 1, dtype=np.uint64)
        offsets[1:] = np.cumsum(self.sizes)
        settings = (L.Settings * len(replicas))(*[s.to_c() for _, s in replicas])
        h = C.c_void_p()
        _check(L.lib().vlj_batch_create(_pptr(allp), offsets.ctypes.data_as(C.POINTER(C.c_uint64)), settings,
                                        len(replicas), C.byref(h)))
        self._h = h
        self.replicas = len(replicas)
        self.profile_bins = [int(st.profile_bins) for _, st in replicas]

    def save(self, path) -> None:
        """Writes the complete state of every replica (see LennardJones.save)."""
        _check(L.lib().vlj_save(self._h, os.fsencode(str(path))), self._h)

    @classmethod
    def load(cls, path) -> "Batch":
        """A batch from a file written by save()."""
        self = cls.__new__(cls)
        self._h = _load_handle(path)
        _, _, reps = _replica_info(self._h, 0)
        info = [_replica_info(self._h, r) for r in range(reps)]
        self.sizes = [n for _, n, _ in info]
        self.replicas = reps
        self.profile_bins = [int(st.profile_bins) for st, _, _ in info]
        return self

    def close(self) -> None:
        if getattr(self, "_h", None):
            L.lib().vlj_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self, steps: int) -> None:
        _check(L.lib().vlj_batch_run(self._h, steps), self._h)

    def step(self, steps: int) -> None:
        """Enqueues `steps` steps of every replica and returns at once."""
        _check(L.lib().vlj_step(self._h, steps), self._h)

    def clear_profile(self) -> None:
        """Starts every replica's profile sums afresh."""
        _check(L.lib().vlj_clear_profile(self._h), self._h)

    def tune(self, steps_per_candidate: int = 6) -> int:
        """Times each candidate brick plan over real steps and keeps the
        fastest (vlj_tune); returns the steps it ran."""
        return _tune(self._h, steps_per_candidate)

    def plan_info(self) -> dict:
        return _plan_info(self._h)

    def set_temperature(self, replica: int, temperature: float) -> None:
        """set_temperature (lennard_jones.rs:905-925) of one replica."""
        _check(L.lib().vlj_batch_set_temperature(self._h, replica, temperature), self._h)

    def measure_now(self) -> List[Measurement]:
        out = (L.Measurement * self.replicas)()
        _check(L.lib().vlj_batch_measure_now(self._h, out, self.replicas), self._h)
        return [Measurement.from_c(m) for m in out]

    def read_particle_array(self) -> np.ndarray:
        out = np.zeros(sum(self.sizes), dtype=PARTICLE_DTYPE)
        _check(L.lib().vlj_batch_read_particles(self._h, _pptr(out), out.shape[0]), self._h)
        return out

    def sample_profile(self) -> None:
        _check(L.lib().vlj_sample_profile(self._h), self._h)

    def debug_counters(self) -> dict:
        return _debug_counters(self._h)

    def debug_replica_counters(self, replica: int) -> dict:
        """One replica's list rebuilds, cut brick builds, list overflow and
        global-memory single-cell builds."""
        out = (C.c_uint64 * 4)()
        _check(L.lib().vlj_debug_replica_counters(self._h, replica, out), self._h)
        return {"rebuilds": out[0], "stage_overflow_builds": out[1], "list_overflow": out[2],
                "global_cell_builds": out[3]}

    def read_profile_raw(self, replica: int):
        out = np.zeros(self.profile_bins[replica] * PROFILE_SUMS + 3, dtype=np.int64)
        samples = C.c_uint64(0)
        _check(L.lib().vlj_batch_read_profile_raw(self._h, replica, out.ctypes.data_as(C.POINTER(C.c_int64)),
                                                  out.shape[0], C.byref(samples)), self._h)
        return out, samples.value


def lattice(count: int, box_size: float, temperature: float, seed: int) -> List[Particle]:
    """Simple cubic lattice filling the cube (lennard_jones.rs:1379-1382)."""
    return to_particles(lattice_array(count, box_size, temperature, seed))


def lattice_array(count: int, box_size: float, temperature: float, seed: int) -> np.ndarray:
    out = np.zeros(count, dtype=PARTICLE_DTYPE)
    _check(L.lib().vlj_lattice(count, box_size, temperature, seed, _pptr(out)))
    return out


def lattice_box(count: int, size, temperature: float, seed: int) -> List[Particle]:
    """Lattice filling a rectangular box (lennard_jones.rs:1388-1399)."""
    return to_particles(lattice_box_array(count, size, temperature, seed))


def lattice_box_array(count: int, size, temperature: float, seed: int) -> np.ndarray:
    out = np.zeros(count, dtype=PARTICLE_DTYPE)
    s = (C.c_double * 3)(*map(float, size))
    _check(L.lib().vlj_lattice_box(count, s, temperature, seed, _pptr(out)))
    return out


def fcc_array(cells_per_side: int, density: float, temperature: float, seed: int):
    """FCC lattice of 4 n^3 particles; returns (particles, side)."""
    out = np.zeros(4 * cells_per_side ** 3, dtype=PARTICLE_DTYPE)
    side = C.c_double()
    _check(L.lib().vlj_fcc(cells_per_side, density, temperature, seed, _pptr(out), C.byref(side)))
    return out, side.value


def fcc_count_array(count: int, density: float, temperature: float, seed: int):
    """FCC sites for any count (the first `count` of the smallest 4 n^3
    lattice); returns (particles, side)."""
    out = np.zeros(count, dtype=PARTICLE_DTYPE)
    side = C.c_double()
    _check(L.lib().vlj_fcc_count(count, density, temperature, seed, _pptr(out), C.byref(side)))
    return out, side.value


def random_insertion_array(count: int, box, min_distance: float, temperature: float, seed: int) -> np.ndarray:
    """Random non-overlapping insertion (uniform trials, min_distance apart)."""
    out = np.zeros(count, dtype=PARTICLE_DTYPE)
    b = np.ascontiguousarray(box, dtype=np.float64)
    _check(L.lib().vlj_random_insertion(count, _dptr(b), min_distance, temperature, seed, _pptr(out)))
    return out


class debug_plan:
    """Overrides the planning rules for handles created inside the block
    (tests only): list_capacity in place of the capacity rule, stage_fill and
    sparse_fill in place of the brick planner's 0.9 (see vlj_debug_plan);
    parts True or False to run the PARTS launches from the start or not (see
    vlj_debug_parts)."""

    def __init__(self, list_capacity: int = 0, stage_fill: float = 0.0, sparse_fill: float = 0.0,
                 parts: Optional[bool] = None):
        self.values = (list_capacity, stage_fill, sparse_fill)
        self.parts = -1 if parts is None else int(bool(parts))

    def __enter__(self):
        _check(L.lib().vlj_debug_plan(*self.values))
        _check(L.lib().vlj_debug_parts(self.parts))
        return self

    def __exit__(self, *exc):
        _check(L.lib().vlj_debug_plan(0, 0.0, 0.0))
        _check(L.lib().vlj_debug_parts(-1))


def debug_rcp(x) -> np.ndarray:
    """The force loop's reciprocal of each x, computed on the device."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.zeros_like(x)
    _check(L.lib().vlj_debug_rcp(_dptr(x), _dptr(out), x.shape[0]))
    return out


def debug_rcp32(x):
    """The FP32 force loop's reciprocal and square root of each x (float32),
    computed on the device."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    rcp = np.zeros_like(x)
    root = np.zeros_like(x)
    fp = C.POINTER(C.c_float)
    _check(L.lib().vlj_debug_rcp32(x.ctypes.data_as(fp), rcp.ctypes.data_as(fp), root.ctypes.data_as(fp),
                                   x.shape[0]))
    return rcp, root


def copy_bandwidth(nbytes: int, repeats: int = 10) -> float:
    """Average ms of a plain device copy of nbytes (read once, written once)."""
    ms = C.c_double()
    _check(L.lib().vlj_copy_bandwidth(nbytes, repeats, C.byref(ms)))
    return ms.value


def fp64_fma_rate(repeats: int = 10) -> float:
    """TFLOP/s of FP64 fused multiply-add chains over the whole device."""
    tf = C.c_double()
    _check(L.lib().vlj_fp64_fma_rate(repeats, C.byref(tf)))
    return tf.value


def fixed_sum(values, fixed_scale: float) -> int:
    """Device 64-bit fixed-point sum of round-ties-even(values * fixed_scale)."""
    v = np.ascontiguousarray(values, dtype=np.float64)
    out = C.c_int64()
    _check(L.lib().vlj_fixed_sum(_dptr(v), v.shape[0], fixed_scale, C.byref(out)))
    return out.value
