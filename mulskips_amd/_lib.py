"""Loader of the HIP engine library (libmskg.so, the C ABI of include/mskg.h).

There is no CPU fallback: if the library is missing or cannot be loaded this
module raises, and every engine entry point fails loudly.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmskg.so")

MSKG_OK, MSKG_EINVAL, MSKG_ESTATE, MSKG_ECAPACITY, MSKG_EDEVICE, MSKG_ENOMEM = range(6)
MSKG_STOP_LIMIT, MSKG_STOP_FEW_MC, MSKG_STOP_NO_EVENTS = 0, 1, 3  # stop_reason of mskg_run
MSKG_SCHEDULE_COLOURS, MSKG_SCHEDULE_TILES = 0, 1  # mskg_set_schedule

EXPORTS = [
    "mskg_create", "mskg_destroy", "mskg_last_error", "mskg_load_state", "mskg_set_temperature",
    "mskg_refresh_rates", "mskg_run", "mskg_replay", "mskg_store_state", "mskg_site_rates",
    "mskg_counters", "mskg_last_timing", "mskg_geometry_find_nn", "mskg_geometry_find_sn",
    "mskg_build_slab", "mskg_set_event_log", "mskg_get_event_log", "mskg_run_serial", "mskg_adjoint", "mskg_build_alloy_slab",
    "mskg_track_classes", "mskg_copy_state", "mskg_frame", "mskg_set_temperature_field", "mskg_phases",
    "mskg_set_la_variant", "mskg_la_counts", "mskg_set_st_tables", "mskg_species_counts",
    "mskg_run_diagnostics", "mskg_frame_defects", "mskg_frame_device", "mskg_set_schedule",
    "mskg_refresh_active", "mskg_adjoint_checkpointed", "mskg_run_ensemble",
    "mskg_build_shape", "mskg_build_cad",
]
MSKG_FIELD_SOCKET, MSKG_FIELD_PLANES = 0, 1  # layouts of the laser-annealing maps

# rate-class keys of the trajectory likelihood (include/mskg.h)
KEY_ADS, KEY_PH, KEY_ISO, KEY_N, ADJ_NS = 2496, 2499, 2500, 2501, 2510


class Config(C.Structure):
    """mskg_config of include/mskg.h."""
    _fields_ = [
        ("len_x", C.c_int32), ("len_y", C.c_int32), ("len_z", C.c_int32),
        ("n_crystal", C.c_int32), ("n_cov", C.c_int32),
        ("list_crystal", C.c_int32 * 3), ("list_cov", C.c_int32 * 3),
        ("ptrans_e", C.c_void_p), ("ptrans_d", C.c_void_p), ("ptrans_abs", C.c_void_p),
        ("ptrans_des", C.c_void_p), ("ptrans_zig", C.c_double),
        ("la_mode", C.c_int32), ("tm", C.c_double), ("p0", C.c_double), ("sigma", C.c_double),
        ("tflex", C.c_double), ("seed", C.c_uint64), ("tile", C.c_int32), ("tile_capacity", C.c_int32),
        ("events_per_pass", C.c_double), ("tau_policy", C.c_int32),
        ("frozen_below_z", C.c_int32), ("second_factor", C.c_void_p), ("stored_len_z", C.c_int32),
        ("reserved", C.c_int32 * 3),
    ]


EVENT_DTYPE = np.dtype([("iter", "<i8"), ("site", "<i4", (3,)), ("itrans", "<i4"), ("arm", "<i4"),
                        ("coor", "<i4"), ("rtot", "<f8"), ("time", "<f8")])

_lib = None


class EngineUnavailable(RuntimeError):
    pass


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise EngineUnavailable(f"{LIB_PATH} is not built (run __graft_entry__.build())")
    L = C.CDLL(LIB_PATH)
    P, I32, I64, D = C.c_void_p, C.c_int32, C.c_int64, C.c_double
    L.mskg_create.argtypes = [P, P]
    L.mskg_destroy.argtypes = [P]
    L.mskg_last_error.restype = C.c_char_p
    L.mskg_last_error.argtypes = [P]
    L.mskg_load_state.argtypes = [P, P, I32, P, P, P, P, I32, P, P, I32, P, P, D]
    L.mskg_set_temperature.argtypes = [P, P]
    L.mskg_refresh_rates.argtypes = [P]
    L.mskg_refresh_active.argtypes = [P]
    L.mskg_run.argtypes = [P, I64, D, P, P, P]
    L.mskg_replay.argtypes = [P, P, I64, I32]
    L.mskg_store_state.argtypes = [P, P, I32, P, P, P, P, P, I32, P, P, P, I32, P, P, P]
    L.mskg_site_rates.argtypes = [P, I64, P, P, P]
    L.mskg_counters.argtypes = [P, P, P, P, P, P]
    L.mskg_last_timing.argtypes = [P, P, P, P, P]
    L.mskg_run_diagnostics.argtypes = [P, P, P, P, P]
    L.mskg_set_schedule.argtypes = [P, I32]
    L.mskg_build_slab.argtypes = [P, I32, I32]
    L.mskg_set_event_log.argtypes = [P, I64]
    L.mskg_get_event_log.argtypes = [P, P, I64, P]
    L.mskg_run_serial.argtypes = [P, I32, P, P, P, P, P]
    L.mskg_run_ensemble.argtypes = [P, I32, P, P, P, P, P]
    L.mskg_build_alloy_slab.argtypes = [P, I32, D]
    L.mskg_build_shape.argtypes = [P, C.c_char_p, P]
    L.mskg_build_cad.argtypes = [P, P, I32, D]
    L.mskg_track_classes.argtypes = [P, I32]
    L.mskg_copy_state.argtypes = [P, P]
    L.mskg_frame.argtypes = [P, P, I64, P, I64, P, I64, P]
    L.mskg_frame_defects.argtypes = [P, P, I64, P, I64, P]
    L.mskg_frame_device.argtypes = [P, P, I64, P, I64, P, I64, P]
    L.mskg_set_temperature_field.argtypes = [P, P, I32, I32]
    L.mskg_phases.argtypes = [P, P, P, I32, I32]
    L.mskg_set_la_variant.argtypes = [P, P, P, P, P, P, P]
    L.mskg_la_counts.argtypes = [P, P, P]
    L.mskg_set_st_tables.argtypes = [P, I32, P, P, P, P, P]
    L.mskg_species_counts.argtypes = [P, P, P]
    L.mskg_adjoint.argtypes = [P, I32, P, P, P, P, P, P]
    L.mskg_adjoint_checkpointed.argtypes = [P, I32, P, P, P, P, P, P, I64, P]
    L.mskg_geometry_find_nn.argtypes = [P, P, P, I32, I32, P]
    L.mskg_geometry_find_sn.argtypes = [P, P, P, I32, I32, P, P]
    _lib = L
    return L


def ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)
