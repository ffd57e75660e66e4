"""ctypes binding of libspeedy_hip.so (C ABI: include/speedy_abi.h).

The library is built in-tree by ``__graft_entry__.build()``.  There is no
fallback: if the shared library is missing or a GPU is absent the calls raise.
"""
from __future__ import annotations

import ctypes
import os

KX_MAX = 16
ABI_VERSION = 1
LIB_NAME = "libspeedy_hip.so"
LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), LIB_NAME)

FLAG_LFLUXLAND = 0x1
FLAG_VERIFY_STATE = 0x2
FLAG_VERIFY_TENDENCIES = 0x4
FLAG_PLAIN_LAUNCH = 0x8
TERMS = ("set_physics_flags", "set_forcing", "spec_hum_to_rel_hum", "get_convection_tendencies",
         "get_large_scale_condensation_tendencies", "get_clouds", "get_shortwave_rad_fluxes",
         "get_downward_longwave_rad_fluxes", "get_surface_fluxes", "get_upward_longwave_rad_fluxes",
         "get_vertical_diffusion_tend")
TERMS_ALL = (1 << len(TERMS)) - 1

ERRORS = {1: "argument", 2: "kx", 3: "workspace", 4: "hip", 5: "capacity"}

FP = ctypes.POINTER(ctypes.c_float)
CFP = ctypes.c_void_p


class SpeedyParams(ctypes.Structure):
    _fields_ = [(n, ctypes.c_float) for n in (
        "psmin", "trcnv", "rhil", "rhbl", "entmax", "smf",
        "trlsc", "rhlsc", "drhlsc", "rhblsc",
        "albcl", "albcls", "absdry", "absaer", "abswv1", "abswv2", "abscl1", "abscl2",
        "ablwin", "ablwv1", "ablwv2", "ablcl1", "ablcl2", "rhcl1", "rhcl2", "qacl", "wpcl",
        "pmaxcl", "clsmax", "clsminl", "gse_s0", "gse_s1",
        "albsea", "albice", "albsn", "epslw", "emisfc",
        "fwind0", "ftemp0", "fhum0", "cdl", "cds", "chl", "chs", "vgust", "ctday", "dtheta",
        "fstab", "clambda", "clambsn", "hdrag",
        "trshc", "trvdi", "trvds", "redshc", "rhgrad", "segrad",
        "co2_scale", "drag_scale")] + [(n, ctypes.c_int32) for n in (
            "lscasym", "lskineb", "increase_co2", "co2_year_ref")]


class SpeedyVertical(ctypes.Structure):
    _fields_ = [("kx", ctypes.c_int32), ("hsg", ctypes.c_float * (KX_MAX + 1))] + [
        (n, ctypes.c_float * KX_MAX) for n in ("fsg", "dhs", "sigl", "grdsig", "grdscp")] + [
        ("wvi", (ctypes.c_float * 2) * KX_MAX)]


class SpeedyDate(ctypes.Structure):
    _fields_ = [("tyear", ctypes.c_float), ("model_year", ctypes.c_int32),
                ("model_step", ctypes.c_int32), ("dt_seconds", ctypes.c_float)]


FIELD_NAMES = ("u", "v", "t", "q", "phi", "ps", "alb0", "sice_am", "snowc_am", "soilw_am", "stl_am", "sst",
               "fmask", "orog", "phis0", "sia", "coa", "du", "dv", "dt", "dq", "diag", "budget")
NBUDGET = 6
BUDGET_NAMES = ("column_energy_tendency_W_m2", "column_water_tendency_g_m2_s", "evaporation_g_m2_s",
                "precipitation_g_m2_s", "water_residual_g_m2_s", "max_abs_column_water_residual_g_m2_s")


class SpeedyFields(ctypes.Structure):
    _fields_ = [(n, CFP) for n in FIELD_NAMES]


ADJ_NAMES = ("ct_du", "ct_dv", "ct_dt", "ct_dq", "g_u", "g_v", "g_t", "g_q", "g_ps", "g_params")


class SpeedyAdjoint(ctypes.Structure):
    _fields_ = [(n, CFP) for n in ADJ_NAMES]


class SpeedySegment(ctypes.Structure):
    _fields_ = [("fields", SpeedyFields), ("adj", SpeedyAdjoint), ("date", SpeedyDate),
                ("nlon", ctypes.c_int32), ("nlat", ctypes.c_int32), ("nmember", ctypes.c_int32),
                ("flags", ctypes.c_uint32), ("level_set", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("workspace", CFP), ("workspace_bytes", ctypes.c_size_t)]


BATCH_MAX_SEGMENTS = 4096


class SpeedyError(RuntimeError):
    pass


_lib = None


def load(path: str | None = None):
    """Load the library once; raise SpeedyError if it is missing (no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise SpeedyError(f"{p} not found: build it with __graft_entry__.build() (there is no CPU fallback)")
    lib = ctypes.CDLL(p)
    lib.speedy_abi_version.restype = ctypes.c_int
    lib.speedy_workspace_bytes.restype = ctypes.c_size_t
    lib.speedy_workspace_bytes.argtypes = [ctypes.c_int32, ctypes.c_int32, ctypes.c_int32]
    lib.speedy_workspace_init.restype = ctypes.c_int
    lib.speedy_workspace_init.argtypes = [CFP, ctypes.c_size_t, CFP]
    lib.speedy_workspace_status.restype = ctypes.c_int
    lib.speedy_workspace_status.argtypes = [CFP, ctypes.c_size_t, ctypes.POINTER(ctypes.c_uint32), CFP]
    lib.speedy_diag_rows.restype = ctypes.c_int
    lib.speedy_diag_rows.argtypes = [ctypes.c_int32]
    lib.speedy_diag_names.restype = ctypes.c_int
    lib.speedy_diag_names.argtypes = [ctypes.c_int32, ctypes.c_char_p, ctypes.c_size_t]
    lib.speedy_fwd.restype = ctypes.c_int
    lib.speedy_fwd.argtypes = [ctypes.POINTER(SpeedyFields), ctypes.POINTER(SpeedyVertical),
                               ctypes.POINTER(SpeedyParams), ctypes.POINTER(SpeedyDate), ctypes.c_int32,
                               ctypes.c_int32, ctypes.c_int32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int32,
                               CFP, ctypes.c_size_t, CFP]
    lib.speedy_noalloc_max_bytes.restype = ctypes.c_size_t
    lib.speedy_noalloc_max_bytes.argtypes = []
    lib.speedy_fwd_noalloc.restype = ctypes.c_int
    lib.speedy_fwd_noalloc.argtypes = [ctypes.POINTER(SpeedyFields), ctypes.POINTER(SpeedyVertical),
                                       ctypes.POINTER(SpeedyParams), ctypes.POINTER(SpeedyDate), ctypes.c_int32,
                                       ctypes.c_int32, ctypes.c_int32, ctypes.c_uint32, ctypes.c_uint32,
                                       ctypes.c_int32, CFP]
    lib.speedy_fwd_vjp.restype = ctypes.c_int
    lib.speedy_fwd_vjp.argtypes = [ctypes.POINTER(SpeedyFields), ctypes.POINTER(SpeedyAdjoint),
                                   ctypes.POINTER(SpeedyVertical), ctypes.POINTER(SpeedyParams),
                                   ctypes.POINTER(SpeedyDate), ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                   ctypes.c_uint32, ctypes.c_int32, CFP, ctypes.c_size_t, CFP]
    lib.speedy_batch_desc_bytes.restype = ctypes.c_size_t
    lib.speedy_batch_desc_bytes.argtypes = [ctypes.POINTER(SpeedySegment), ctypes.c_int32]
    lib.speedy_fwd_batch.restype = ctypes.c_int
    lib.speedy_fwd_batch.argtypes = [ctypes.POINTER(SpeedySegment), ctypes.c_int32, ctypes.POINTER(SpeedyVertical),
                                     ctypes.POINTER(SpeedyParams), CFP, ctypes.c_size_t, CFP, CFP]
    lib.speedy_read_level_sets.restype = ctypes.c_int
    lib.speedy_read_level_sets.argtypes = [CFP, ctypes.c_size_t, ctypes.c_int32,
                                           ctypes.POINTER(ctypes.c_uint32), CFP]
    lib.speedy_diag_accumulate.restype = ctypes.c_int
    lib.speedy_diag_accumulate.argtypes = [CFP, CFP, ctypes.c_int64, ctypes.c_float, CFP]
    lib.speedy_timer_enable.restype = ctypes.c_int
    lib.speedy_timer_enable.argtypes = [ctypes.c_int32]
    lib.speedy_timer_disable.restype = ctypes.c_int
    lib.speedy_timer_read.restype = ctypes.c_int
    lib.speedy_timer_read.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_int32), ctypes.c_int32]
    lib.speedy_last_error.restype = ctypes.c_char_p
    if lib.speedy_abi_version() != ABI_VERSION:
        raise SpeedyError("libspeedy_hip ABI version mismatch")
    _lib = lib
    return lib


def check(rc: int, what: str):
    if rc != 0:
        msg = _lib.speedy_last_error().decode() if _lib is not None else ""
        raise SpeedyError(f"{what} failed ({ERRORS.get(abs(rc), rc)}): {msg}")


def diag_names(kx: int) -> list[str]:
    """Row names of the diagnostic output; 3-D fields expand to kx rows 'name[k]'."""
    lib = load()
    buf = ctypes.create_string_buffer(8192)
    check(lib.speedy_diag_names(kx, buf, len(buf)), "speedy_diag_names")
    out = []
    for n in buf.value.decode().split(","):
        if n.endswith("[]"):
            out.extend(f"{n[:-2]}[{k}]" for k in range(kx))
        else:
            out.append(n)
    return out


KERNEL_KINDS = ("levelset", "fwd", "fwd_vjp", "param_reduce")


def timer_read():
    """{kind: (total_ms, launches)} of the kernels timed since the last read."""
    lib = load()
    n = len(KERNEL_KINDS)
    ms = (ctypes.c_float * n)()
    cnt = (ctypes.c_int32 * n)()
    check(lib.speedy_timer_read(ms, cnt, n), "speedy_timer_read")
    return {k: (float(ms[i]), int(cnt[i])) for i, k in enumerate(KERNEL_KINDS)}
