"""Loads libvirial_lj.so (the HIP library) and declares its C ABI for ctypes.

There is no CPU fallback: without the built library every entry point
raises, and without a gfx950 device the library's own calls return
VLJ_E_NO_DEVICE.
"""
from __future__ import annotations

import ctypes as C
import os

PACKAGE_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PACKAGE_ROOT, "libvirial_lj.so")

OK = 0
E_INVALID = 1
E_HIP = 2
E_STATE = 3
E_UNSUPPORTED = 4
E_NO_DEVICE = 5


class Particle(C.Structure):
    _fields_ = [("pos", C.c_double * 3), ("vel", C.c_double * 3)]


class Thermostat(C.Structure):
    _fields_ = [
        ("kind", C.c_uint32),
        ("temperature", C.c_double),
        ("friction", C.c_double),
        ("period", C.c_double),
        ("seed", C.c_uint64),
    ]


class Settings(C.Structure):
    _fields_ = [
        ("box_size", C.c_double * 3),
        ("cutoff", C.c_double),
        ("dt", C.c_double),
        ("skin", C.c_double),
        ("thermostat", Thermostat),
        ("shifted_force", C.c_int32),
        ("profile_bins", C.c_uint32),
        ("precision", C.c_uint32),
    ]


class Measurement(C.Structure):
    _fields_ = [
        ("step", C.c_uint64),
        ("time", C.c_double),
        ("kinetic", C.c_double),
        ("potential", C.c_double),
        ("virial", C.c_double),
        ("momentum", C.c_double * 3),
        ("count", C.c_uint32),
        ("volume", C.c_double),
        ("list_rebuilds", C.c_uint32),
        ("list_overflow", C.c_uint32),
        ("thermostat_energy", C.c_double),
        ("condensed", C.c_uint32),
        ("pairs_within_cutoff", C.c_uint64),
        ("raw", C.c_int64 * 8),
    ]


class TuneTrial(C.Structure):
    _fields_ = [("fill", C.c_double), ("force_us", C.c_double), ("build_us", C.c_double), ("bricks", C.c_uint32)]


# Every symbol include/virial_lj.h declares, with its ctypes signature.
_P = C.c_void_p
_d = C.c_double
_u32 = C.c_uint32
_u64 = C.c_uint64
_i32 = C.c_int32
_sz = C.c_size_t
_dp = C.POINTER(C.c_double)
_up = C.POINTER(C.c_uint32)
_u64p = C.POINTER(C.c_uint64)
_i64p = C.POINTER(C.c_int64)
_pp = C.POINTER(Particle)

SIGNATURES = {
    "vlj_create": (C.c_int, [_pp, _sz, C.POINTER(Settings), C.POINTER(_P)]),
    "vlj_destroy": (None, [_P]),
    "vlj_step": (C.c_int, [_P, _u32]),
    "vlj_run": (C.c_int, [_P, _u64]),
    "vlj_measure": (C.c_int, [_P, C.POINTER(_i32)]),
    "vlj_poll": (C.c_int, [_P, C.POINTER(Measurement), _sz, C.POINTER(_sz)]),
    "vlj_measure_now": (C.c_int, [_P, C.POINTER(Measurement)]),
    "vlj_read_particles": (C.c_int, [_P, _pp, _sz]),
    "vlj_count": (_u32, [_P]),
    "vlj_box_size": (C.c_int, [_P, _dp]),
    "vlj_volume": (_d, [_P]),
    "vlj_steps": (_u64, [_P]),
    "vlj_time": (_d, [_P]),
    "vlj_target_temperature": (C.c_int, [_P, _dp, C.POINTER(_i32)]),
    "vlj_set_temperature": (C.c_int, [_P, _d]),
    "vlj_set_condensed_neighbours": (C.c_int, [_P, _u32]),
    "vlj_sample_profile": (C.c_int, [_P]),
    "vlj_clear_profile": (C.c_int, [_P]),
    "vlj_read_profile": (C.c_int, [_P, _dp, _sz, _u64p]),
    "vlj_read_profile_raw": (C.c_int, [_P, _i64p, _sz, _u64p]),
    "vlj_batch_read_profile_raw": (C.c_int, [_P, _sz, _i64p, _sz, _u64p]),
    "vlj_set_heat_exchange": (C.c_int, [_P, _i32, _dp, _dp, _d]),
    "vlj_heat_flux": (C.c_int, [_P, _dp, C.POINTER(_i32)]),
    "vlj_record_heat_current": (C.c_int, [_P, _u32]),
    "vlj_sample_heat_current": (C.c_int, [_P]),
    "vlj_read_heat_current": (C.c_int, [_P, _dp, _sz, _up]),
    "vlj_read_heat_current_raw": (C.c_int, [_P, _i64p, _sz, _up]),
    "vlj_last_error": (C.c_char_p, [_P]),
    "vlj_batch_create": (C.c_int, [_pp, _u64p, C.POINTER(Settings), _sz, C.POINTER(_P)]),
    "vlj_save": (C.c_int, [_P, C.c_char_p]),
    "vlj_render_buffers": (C.c_int, [_P, C.POINTER(_P), C.POINTER(_P)]),
    "vlj_render": (C.c_int, [_P]),
    "vlj_read_render": (C.c_int, [_P, C.POINTER(C.c_float), C.POINTER(C.c_float), _sz]),
    "vlj_set_drawn": (C.c_int, [_P, C.c_uint32]),
    "vlj_load": (C.c_int, [C.c_char_p, C.POINTER(_P)]),
    "vlj_replica_info": (C.c_int, [_P, _sz, C.POINTER(Settings), _u64p, C.POINTER(_sz)]),
    "vlj_batch_run": (C.c_int, [_P, _u64]),
    "vlj_batch_set_temperature": (C.c_int, [_P, _sz, _d]),
    "vlj_batch_measure_now": (C.c_int, [_P, C.POINTER(Measurement), _sz]),
    "vlj_batch_replicas": (_sz, [_P]),
    "vlj_batch_read_particles": (C.c_int, [_P, _pp, _sz]),
    "vlj_tune": (C.c_int, [_P, _u32, _u64p]),
    "vlj_plan_info": (C.c_int, [_P, _dp, _up, C.POINTER(TuneTrial), _sz, C.POINTER(_sz)]),
    "vlj_lattice": (C.c_int, [_sz, _d, _d, _u64, _pp]),
    "vlj_lattice_box": (C.c_int, [_sz, _dp, _d, _u64, _pp]),
    "vlj_fcc": (C.c_int, [_u32, _d, _d, _u64, _pp, _dp]),
    "vlj_fcc_count": (C.c_int, [_sz, _d, _d, _u64, _pp, _dp]),
    "vlj_random_insertion": (C.c_int, [_sz, _dp, _d, _d, _u64, _pp]),
    "vlj_fixed_sum": (C.c_int, [_dp, _sz, _d, _i64p]),
    "vlj_fixed_sum_device": (C.c_int, [_P, _sz, _d, _P, _P]),
    "vlj_debug_state": (C.c_int, [_P, _dp, _dp, _dp, _up]),
    "vlj_debug_cells": (C.c_int, [_P, _up, _up, _up]),
    "vlj_debug_lists": (C.c_int, [_P, _up, _up]),
    "vlj_debug_grid": (C.c_int, [_P, _up]),
    "vlj_debug_params": (C.c_int, [_P, _dp]),
    "vlj_debug_counters": (C.c_int, [_P, _u64p]),
    "vlj_debug_replica_counters": (C.c_int, [_P, _sz, _u64p]),
    "vlj_debug_plan": (C.c_int, [_u32, _d, _d]),
    "vlj_debug_parts": (C.c_int, [C.c_int32]),
    "vlj_debug_swizzle": (_u32, [_u32, _u32]),
    "vlj_copy_bandwidth": (C.c_int, [_sz, _u32, _dp]),
    "vlj_fp64_fma_rate": (C.c_int, [_u32, _dp]),
    "vlj_debug_rcp": (C.c_int, [_dp, _dp, _sz]),
    "vlj_debug_rcp32": (C.c_int, [C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_float), _sz]),
    "vlj_timed_run": (C.c_int, [_P, _u64, _dp, _dp]),
}

_lib = None


class LibraryMissing(ImportError):
    pass


def lib() -> C.CDLL:
    """The loaded library; raises LibraryMissing when it is not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise LibraryMissing(
                f"{LIB_PATH} is not built: run `make -C virial-engine_amd` or __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib
