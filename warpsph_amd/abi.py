"""ctypes mirror of include/crksph.h and the loader of the HIP library.

The library is built in-tree by ``__graft_entry__.build()`` into
``warpsph_amd/lib/libcrksph_hip.so``.  There is no fallback: if the library is
missing every entry point raises, so a step can never silently run anywhere
but on the GPU kernels.
"""
from __future__ import annotations

import ctypes as C
import os
import threading

LIB_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib")
LIB_PATH = os.path.join(LIB_DIR, "libcrksph_hip.so")

# status codes
CRK_OK = 0
CRK_EINVAL = -1
CRK_EHIP = -2
CRK_ECAPACITY = -3
CRK_ESTATE = -4

# enums (values as in include/crksph.h)
CRK_KERNEL_WENDLAND2 = 0
CRK_KERNEL_WENDLAND4 = 1
CRK_KERNEL_WENDLAND6 = 2
CRK_SUPPORT_NONE = 0
CRK_SUPPORT_MONAGHAN = 1
CRK_SUPPORT_OWEN = 2
CRK_SWITCH_CULLEN_DEHNEN = 2
CRK_SWITCH_NONE = 6
CRK_SWITCH_LEARNED = 100
PASSES = ("owen", "volume", "moments", "velgrad", "switch", "accel", "balance", "compat", "support")


class CrkDomain(C.Structure):
    _pack_ = 4
    _fields_ = [("dim", C.c_int32), ("periodic", C.c_int32 * 3),
                ("min", C.c_float * 3), ("max", C.c_float * 3)]


class CrkDiffusionParams(C.Structure):
    _pack_ = 4
    _fields_ = [("c_s", C.c_float), ("C_l", C.c_float), ("C_q", C.c_float),
                ("Cu_l", C.c_float), ("Cu_q", C.c_float), ("K", C.c_float),
                ("thermalConductivity", C.c_float),
                ("viscosityTerm", C.c_int32), ("thermalConductivityTerm", C.c_int32),
                ("scaleBeta", C.c_int32), ("monaghanSwitch", C.c_int32), ("correctXi", C.c_int32)]


class CrkViscosityParams(C.Structure):
    _pack_ = 4
    _fields_ = [("eta_fold", C.c_float), ("eta_crit", C.c_float),
                ("enableCRKLimiter", C.c_int32), ("enableVanLeerLimiter", C.c_int32),
                ("forceVanLeerOff", C.c_int32), ("forceVanLeerOn", C.c_int32)]


class CrkSwitchParams(C.Structure):
    _pack_ = 4
    _fields_ = [("scheme", C.c_int32), ("limitXi", C.c_int32), ("correctVelocityGradient", C.c_int32),
                ("alpha_min", C.c_float), ("alpha_max", C.c_float), ("beta_c", C.c_float),
                ("beta_d", C.c_float), ("beta_xi", C.c_float)]


class CrkConfig(C.Structure):
    _pack_ = 4
    _fields_ = [("domain", CrkDomain), ("kernel", C.c_int32), ("verletScale", C.c_float),
                ("targetNeighbors", C.c_float), ("cflFactor", C.c_float), ("minDt", C.c_float),
                ("maxDt", C.c_float), ("dtGrowthFactor", C.c_float), ("gamma", C.c_float),
                ("energyScheme", C.c_int32), ("adaptiveSupportScheme", C.c_int32),
                ("adaptiveSupportIterations", C.c_int32), ("adaptiveSupportThreshold", C.c_float),
                ("compatibleEnergy", C.c_int32), ("diffusion", CrkDiffusionParams),
                ("crkViscosity", CrkViscosityParams), ("switchParams", CrkSwitchParams),
                ("alphaWeightedQ", C.c_int32), ("evolveAlpha0", C.c_int32)]


class CrkGrid(C.Structure):
    _pack_ = 4
    _fields_ = [("dim", C.c_int32), ("bits", C.c_int32), ("ncell", C.c_int32 * 3),
                ("origin", C.c_float * 3), ("inv_cell", C.c_float * 3), ("table_size", C.c_int64)]


class CrkAdjacency(C.Structure):
    _fields_ = [("n", C.c_int32), ("n_edges", C.c_int64), ("perm", C.c_void_p), ("rank", C.c_void_p),
                ("row_offsets", C.c_void_p), ("nbr", C.c_void_p), ("tiles", C.c_void_p), ("nbr_loc", C.c_void_p)]


# LDS tiles (include/crksph.h)
CRK_TILE_ROWS = 256
CRK_TILE_CAP = 1152
CRK_TILE_MAXRUNS = 40
CRK_TILE_IDS = 20480
CRK_TILE_WORDS = 4 + 2 * (CRK_TILE_MAXRUNS + 1)


class CrkState(C.Structure):
    _fields_ = [("n", C.c_int32)] + [(k, C.c_void_p) for k in (
        "positions", "velocities", "supports", "masses", "densities", "internalEnergies",
        "entropies", "pressures", "soundspeeds", "divergence", "alpha0s", "alphas")]


class CrkUpdate(C.Structure):
    _fields_ = [(k, C.c_void_p) for k in ("dvdt", "dudt", "drhodt", "dEdt")]


class CrkEdges(C.Structure):
    _fields_ = [(k, C.c_void_p) for k in ("count", "nbr", "pf", "rec", "f", "loc", "th")]


MLP_IN = 8
MLP_HIDDEN = 64
MLP_NPARAM = 64 * 8 + 64 + 64 * 64 + 64 + 64 + 1
MLP_NBF16 = 64 * 16 + 64 * 64 + 64


class CrkMlp(C.Structure):
    _fields_ = [("params", C.c_void_p), ("wbf16", C.c_void_p), ("mu", C.c_float * 8), ("sigma", C.c_float * 8),
                ("lo", C.c_float * 8), ("hi", C.c_float * 8), ("scale", C.c_float), ("multiply", C.c_int32)]


_P = C.c_void_p
_PROTOS = {
    "crk_ctx_create": [C.POINTER(C.c_void_p)],
    "crk_ctx_destroy": [_P],
    "crk_last_error": [C.c_char_p, C.c_size_t],
    "crk_kernel_constants": [C.c_int32, C.c_int32, C.POINTER(C.c_double), C.POINTER(C.c_double)],
    "crk_kernel_xi": [C.c_int32, C.c_int32, C.POINTER(C.c_double)],
    "crk_owen_lut": [C.c_int32, C.c_int32, C.c_int32, C.c_double, C.c_double, _P, _P],
    "crk_set_owen_lut": [_P, C.c_int32, _P, _P, C.c_double, _P],
    "crk_owen_lut_device": [_P, C.c_int32, C.c_int32, C.c_int32, C.c_double, C.c_double, C.c_double, _P],
    "crk_owen_lut_read": [_P, _P, _P, _P],
    "crk_grid_plan": [C.POINTER(CrkConfig), C.c_float, _P, _P, C.POINTER(CrkGrid)],
    "crk_adjacency_count": [_P, C.POINTER(CrkConfig), C.POINTER(CrkGrid), C.c_int32, _P, _P,
                            _P, _P, _P, _P, _P, _P, _P, _P, C.POINTER(C.c_int64), _P],
    "crk_adjacency_fill": [_P, C.POINTER(CrkConfig), C.POINTER(CrkGrid), C.c_int32, _P, _P, _P,
                           _P, _P, _P, C.c_int64, _P],
    "crk_adjacency_count_keep": [_P, C.POINTER(CrkConfig), C.POINTER(CrkGrid), C.c_int32, _P, _P, _P, _P,
                                 C.POINTER(CrkAdjacency), _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P,
                                 C.POINTER(C.c_int64), _P],
    "crk_runs_failed": [_P, C.POINTER(CrkConfig), C.c_int32, _P, _P],
    "crk_adjacency_tiles": [_P, C.POINTER(CrkAdjacency), _P, _P, _P, C.c_int32, _P, _P, _P],
    "crk_adjacency_valid": [_P, C.POINTER(CrkConfig), C.c_int32, _P, _P, _P, _P, _P,
                            C.POINTER(C.c_int32), _P],
    "crk_adjacency_export": [_P, C.c_int32, _P, _P, _P, C.c_int64, _P, _P, _P, _P, _P],
    "crk_stage_begin": [_P, C.POINTER(CrkConfig), C.POINTER(CrkAdjacency), C.POINTER(CrkState), _P],
    "crk_stage_rebind": [_P, C.POINTER(CrkConfig), C.POINTER(CrkAdjacency), C.POINTER(CrkState), _P],
    "crk_state_refresh": [_P, C.POINTER(CrkConfig), C.POINTER(CrkState), _P],
    "crk_update_refresh": [_P, C.POINTER(CrkUpdate), _P],
    "crk_stage_begin_checked": [_P, C.POINTER(CrkConfig), C.POINTER(CrkAdjacency), C.POINTER(CrkState), _P, _P, _P,
                                C.POINTER(C.c_int32), _P],
    "crk_owen_valid": [_P, C.POINTER(CrkConfig), C.POINTER(C.c_int32), _P],
    "crk_owen_iteration": [_P, C.POINTER(CrkConfig), C.POINTER(CrkState), _P, _P, _P],
    "crk_minmax": [_P, C.c_int32, _P, _P, _P],
    "crk_owen_support": [_P, C.POINTER(CrkConfig), C.POINTER(CrkState), _P, _P, _P],
    "crk_monaghan_iteration": [_P, C.POINTER(CrkConfig), C.POINTER(CrkState), _P, _P, _P],
    "crk_pairs_factors_eos": [_P, C.POINTER(CrkConfig), C.POINTER(CrkState), C.POINTER(CrkEdges), _P, _P],
    "crk_support": [_P, C.POINTER(CrkConfig), C.POINTER(CrkEdges), _P],
    "crk_density_gather": [_P, C.POINTER(CrkConfig), _P, _P],
    "crk_factors_eos": [_P, C.POINTER(CrkConfig), C.POINTER(CrkState), _P, _P],
    "crk_velocity_gradient": [_P, C.POINTER(CrkConfig), C.POINTER(CrkUpdate), _P],
    "crk_switch": [_P, C.POINTER(CrkConfig), C.POINTER(CrkState), C.c_float, _P],
    "crk_accel_dudt": [_P, C.POINTER(CrkConfig), C.POINTER(CrkState), C.POINTER(CrkUpdate),
                       C.POINTER(CrkEdges), _P],
    "crk_switch_accel_dudt": [_P, C.POINTER(CrkConfig), C.POINTER(CrkState), C.POINTER(CrkUpdate),
                              C.POINTER(CrkEdges), C.c_float, _P],
    "crk_balance": [_P, C.POINTER(CrkConfig), C.c_float, C.POINTER(CrkEdges), _P],
    "crk_export_field": [_P, C.c_int32, _P, _P],
    "crk_edges_export": [_P, C.c_int32, C.POINTER(CrkAdjacency), C.POINTER(CrkEdges), _P, C.c_float, _P, _P, _P,
                         _P],
    "crk_axpy": [C.c_int64, _P, _P, C.c_float, _P, _P],
    "crk_rk_combine": [C.c_int64, _P, _P, _P, C.c_float, C.c_float, C.c_float, _P, _P],
    "crk_wrap_positions": [C.POINTER(CrkConfig), C.c_int32, _P, _P],
    "crk_compatible_energy": [_P, C.POINTER(CrkConfig), C.c_int32, C.c_float, _P, _P, _P, _P, _P,
                              C.c_float, C.c_float, C.POINTER(CrkAdjacency), C.POINTER(CrkEdges),
                              C.POINTER(CrkAdjacency), C.POINTER(CrkEdges), _P, _P, _P],
    "crk_set_group_size": [_P, C.c_int32, C.c_int32],
    "crk_set_block_size": [_P, C.c_int32, C.c_int32],
    "crk_struct_sizes": [C.POINTER(C.c_int64), C.c_int32],
    "crk_cfl_dt": [_P, C.POINTER(CrkConfig), C.c_int32, _P, _P, C.c_double,
                   C.POINTER(C.c_double), _P],
    "crk_conservation": [_P, C.c_int32, C.c_int32, _P, _P, _P, _P, _P],
    "crk_conservation_noalloc": [_P, C.c_int32, C.c_int32, _P, _P, _P, _P, _P],
    "crk_bind_runs": [_P, C.c_int32, _P, _P, _P],
    "crk_run_ids": [_P, C.c_int32, _P, _P],
    "crk_run_extents": [_P, C.POINTER(CrkConfig), C.c_int32, _P, _P, _P, _P],
    "crk_grid_plan_runs": [C.POINTER(CrkConfig), C.c_int32, _P, _P, C.POINTER(C.c_int64)],
    "crk_runs_dt": [_P, C.c_int32, _P, _P, _P],
    "crk_cfl_dt_runs": [_P, C.POINTER(CrkConfig), C.c_int32, _P, _P, _P, _P, _P, _P],
    "crk_axpy_p": [C.c_int64, C.c_int32, _P, _P, _P, _P, _P],
    "crk_rk_combine_p": [C.c_int64, C.c_int32, _P, _P, _P, C.c_float, C.c_float, _P, _P, _P],
    "crk_flag_clear": [_P, _P],
    "crk_flag_read": [_P, C.POINTER(C.c_int32), _P],
    "crk_flag_read_async": [_P, C.c_int32, _P],
    "crk_flag_wait": [_P, C.c_int32, C.POINTER(C.c_int32)],
    "crk_copy_unless_flag": [_P, _P, _P, C.c_int64, _P],
    "crk_runs_advance_time": [_P, _P, _P, _P],
    "crk_cfl_dt_device": [_P, C.POINTER(CrkConfig), C.c_int32, _P, _P, _P, _P, _P, _P],
    "crk_mlp_pack": [_P, _P, _P],
    "crk_mlp_forward": [C.POINTER(CrkMlp), C.c_int32, _P, _P, _P],
    "crk_mlp_grad": [_P, C.POINTER(CrkMlp), C.c_int32, _P, _P, _P, _P],
    "crk_mlp_adam": [C.c_int32, _P, _P, _P, _P, C.c_float, C.c_float, C.c_float, C.c_float, C.c_int32, _P],
    "crk_learned_features": [_P, C.POINTER(CrkConfig), C.c_float, _P, _P],
    "crk_learned_switch": [_P, C.POINTER(CrkConfig), C.POINTER(CrkState), C.c_float, C.POINTER(CrkMlp), _P],
}
# every symbol include/crksph.h declares (checked by tests/test_abi.py)
EXPORTED = sorted(list(_PROTOS.keys()) + ["crk_version"])

_lib = None
_lock = threading.Lock()


class CrkError(RuntimeError):
    pass


def lib():
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise CrkError(f"HIP library not built: {LIB_PATH} is missing "
                           "(run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        for name, args in _PROTOS.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = C.c_int
        L.crk_version.argtypes = []
        L.crk_version.restype = C.c_char_p
        _lib = L
    return _lib


def last_error() -> str:
    buf = C.create_string_buffer(1024)
    lib().crk_last_error(buf, 1024)
    return buf.value.decode(errors="replace")


def check(status: int, what: str):
    """Translate a status code into the exception the reference would raise."""
    if status == CRK_OK:
        return
    msg = f"{what}: {last_error()}"
    if status in (CRK_EINVAL, CRK_ECAPACITY):
        raise ValueError(msg)
    raise CrkError(msg)


def call(name: str, *args):
    check(getattr(lib(), name)(*args), name)


def ptr(t) -> C.c_void_p:
    """Device (or host) pointer of a tensor / None."""
    if t is None:
        return C.c_void_p(0)
    return C.c_void_p(t.data_ptr())
