"""ctypes binding of libcns.so (include/cns.h).

The library is the product: if it is missing or fails to load, every entry
point raises; there is no fallback implementation.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcns.so")

CNS_OK = 0
CNS_RK3_LOWSTORAGE, CNS_RK3_KUTTA, CNS_RK4 = 0, 1, 2
CNS_FP64, CNS_FP32 = 0, 1


class CnsConfig(C.Structure):
    """Mirror of struct cns_config."""
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("stencil_size", C.c_int), ("stencil_visc", C.c_int),
        ("periodic_x", C.c_int), ("nonuniform_x", C.c_int), ("forcing_z", C.c_int),
        ("sponge", C.c_int), ("scheme", C.c_int), ("check_cfl", C.c_int),
        ("check_bulk", C.c_int), ("fix_y_work", C.c_int), ("precision", C.c_int), ("compensated", C.c_int),
        ("device", C.c_int),
        ("len_x", C.c_double), ("len_y", C.c_double), ("len_z", C.c_double),
        ("reynolds", C.c_double), ("prandtl", C.c_double), ("mach", C.c_double), ("gamma", C.c_double),
        ("visc_exp", C.c_double), ("cfl", C.c_double), ("stretch", C.c_double),
        ("twall_bot", C.c_double), ("twall_top", C.c_double), ("dpdz0", C.c_double),
        ("boundary_layer", C.c_int), ("wall_blowing", C.c_int),
    ]


class CnsControl(C.Structure):
    """Mirror of struct cns_control."""
    _fields_ = [
        ("time", C.c_double), ("dt", C.c_double), ("dpdz", C.c_double), ("visc_max", C.c_double),
        ("steps", C.c_int64), ("visc_from_state", C.c_int), ("reserved", C.c_int),
    ]


# Python Config names (the reference's macro names) -> C struct fields
C_NAMES = {"mx": "nx", "my": "ny", "mz": "nz", "forcing": "forcing_z", "Lx": "len_x", "Ly": "len_y",
           "Lz": "len_z", "Re": "reynolds", "Pr": "prandtl", "Ma": "mach", "gam": "gamma",
           "viscexp": "visc_exp", "perturbed": "wall_blowing"}


EXPORTS = [
    "cns_config_init", "cns_create", "cns_destroy", "cns_last_error", "cns_get_grid",
    "cns_fill_taylor_green", "cns_fill_channel", "cns_init_taylor_green", "cns_init_channel", "cns_set_state", "cns_get_state", "cns_rhs", "cns_rhs_noalloc",
    "cns_step", "cns_set_sponge", "cns_get_scalars", "cns_sync", "cns_profile_enable",
    "cns_profile_read", "cns_time_steps", "cns_device_bytes", "cns_mem_info", "cns_release_cache", "cns_sponge_from_profile",
    "cns_write_fields", "cns_read_fields", "cns_write_grid", "cns_get_control", "cns_set_control",
    "cns_channel_profiles", "cns_friction_reynolds", "cns_flow_stats", "cns_set_pencil", "cns_tune",
    "cns_step_stats",
    "cns_batch_create", "cns_batch_tune", "cns_batch_step", "cns_batch_launch", "cns_batch_wait", "cns_batch_time_steps", "cns_batch_group_info", "cns_batch_set_shape",
    "cns_batch_groups", "cns_batch_destroy", "cns_batch_last_error",
]

_lib = None


class CnsError(RuntimeError):
    pass


def load(path=None):
    """Load libcns.so and declare the prototypes; raises if it is absent."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise CnsError(f"{p} is not built; run cudanavierstokes_amd/build.py (hipcc, gfx950)")
    L = C.CDLL(p)
    dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
    vp = C.c_void_p
    L.cns_config_init.argtypes = [C.POINTER(CnsConfig)]
    L.cns_create.argtypes = [C.POINTER(CnsConfig), C.POINTER(vp)]
    L.cns_destroy.argtypes = [vp]
    L.cns_last_error.restype = C.c_char_p
    L.cns_last_error.argtypes = [vp]
    L.cns_get_grid.argtypes = [vp, vp, vp, vp, vp, vp, vp]
    L.cns_fill_taylor_green.argtypes = [vp, dp, dp, dp, dp, dp]
    L.cns_fill_channel.argtypes = [vp, C.c_uint64, C.c_double, C.c_double, dp, dp, dp, dp, dp]
    L.cns_init_taylor_green.argtypes = [vp]
    L.cns_init_channel.argtypes = [vp, C.c_uint64, C.c_double, C.c_double]
    L.cns_set_state.argtypes = [vp, dp, dp, dp, dp, dp]
    L.cns_get_state.argtypes = [vp, dp, dp, dp, dp, dp]
    L.cns_rhs.argtypes = [vp, dp, dp, dp, dp, dp]
    L.cns_rhs_noalloc.argtypes = [vp, vp, vp]
    L.cns_step.argtypes = [vp, C.c_int, vp, vp, vp]
    L.cns_set_sponge.argtypes = [vp, dp, dp, dp, dp, dp, dp, dp]
    L.cns_get_scalars.argtypes = [vp, vp, vp, vp]
    L.cns_sync.argtypes = [vp]
    L.cns_profile_enable.argtypes = [vp, C.c_int]
    L.cns_profile_read.argtypes = [vp, C.POINTER(C.c_double), C.POINTER(C.c_int)]
    L.cns_time_steps.argtypes = [vp, C.c_int, C.POINTER(C.c_double)]
    L.cns_device_bytes.argtypes = [vp, C.POINTER(C.c_uint64)]
    L.cns_mem_info.argtypes = [C.c_int, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
    L.cns_release_cache.argtypes = [C.c_int]
    L.cns_sponge_from_profile.argtypes = [C.c_int, dp, dp, dp, dp, C.c_int, dp, C.c_int, dp, C.c_double,
                                          C.c_double, C.c_double, C.c_double, C.c_double, vp, dp, dp, dp, dp,
                                          dp, dp, dp]
    L.cns_write_fields.argtypes = [vp, C.c_char_p, C.c_int]
    L.cns_read_fields.argtypes = [vp, C.c_char_p, C.c_int]
    L.cns_write_grid.argtypes = [vp, C.c_char_p]
    L.cns_get_control.argtypes = [vp, C.POINTER(CnsControl)]
    L.cns_set_control.argtypes = [vp, C.POINTER(CnsControl)]
    L.cns_channel_profiles.argtypes = [vp, dp]
    L.cns_friction_reynolds.argtypes = [vp, C.POINTER(C.c_double)]
    L.cns_flow_stats.argtypes = [vp, dp]
    L.cns_set_pencil.argtypes = [vp, C.c_int]
    L.cns_tune.argtypes = [vp, C.POINTER(C.c_int)]
    L.cns_step_stats.argtypes = [vp, C.c_int, C.c_int, dp]
    L.cns_batch_create.argtypes = [C.POINTER(vp), C.c_int, C.POINTER(vp)]
    L.cns_batch_tune.argtypes = [vp]
    L.cns_batch_step.argtypes = [vp, C.c_int]
    L.cns_batch_launch.argtypes = [vp, C.c_int]
    L.cns_batch_wait.argtypes = [vp]
    L.cns_batch_time_steps.argtypes = [vp, C.c_int, C.POINTER(C.c_double)]
    ip = C.POINTER(C.c_int)
    L.cns_batch_group_info.argtypes = [vp, C.c_int, ip, ip, ip, ip, ip]
    L.cns_batch_groups.argtypes = [vp]
    L.cns_batch_set_shape.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int]
    L.cns_batch_destroy.argtypes = [vp]
    L.cns_batch_last_error.restype = C.c_char_p
    L.cns_batch_last_error.argtypes = [vp]
    for name in EXPORTS:
        if name not in ("cns_last_error", "cns_batch_last_error"):
            getattr(L, name).restype = C.c_int
    if path is None:
        _lib = L
    return L


def check(rc, handle=None, what=""):
    if rc != CNS_OK:
        msg = load().cns_last_error(handle)
        raise CnsError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")


def check_batch(rc, handle=None, what=""):
    if rc != CNS_OK:
        msg = load().cns_batch_last_error(handle)
        raise CnsError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")
