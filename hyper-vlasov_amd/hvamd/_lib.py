"""ctypes binding of libhvamd.so (C-ABI declared in include/hv_abi.h and
include/hv_init.h).

Only plain pointers and sizes cross the boundary.  Loading fails loudly when
the library has not been built: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os
import sys
import warnings
from pathlib import Path

import numpy as np

PKG_ROOT = Path(__file__).resolve().parent.parent
LIB_PATH = Path(os.environ.get("HVAMD_LIB", PKG_ROOT / "libhvamd.so"))

HV_OK = 0
HV_ERR_ARG = -1
HV_ERR_HIP = -2
HV_ERR_ALLOC = -3
HV_ERR_STATE = -4
HV_ERR_IO = -5
HV_WARN_SOR_NOT_CONVERGED = 1
HV_WARN_PARTICLE_OUTSIDE = 2
HV_WARN_SMEAR_MARGIN = 4
HV_WARN_INIT_NOT_CONVERGED = 8
HV_NOALLOC_MAX_PARTICLES = 1 << 24

HV_PREC_F64 = 0
HV_PREC_MIXED = 1

HV_PROTONS = 0
HV_NEUTRONS = 1

HV_BOOST_REF_WEIGHTED = 0
HV_BOOST_SYMMETRIC = 1

HV_FIELD_DENSITY = 0
HV_FIELD_POTENTIALS = 1
HV_FIELD_COULOMB = 2
HV_FIELD_FORCES_X = 3
HV_FIELD_FORCES_Y = 4
HV_FIELD_FORCES_Z = 5

HV_OPT_SORT_EVERY = 0
HV_OPT_SOR_RESIDENT = 1
HV_OPT_SOR_FUSED = 2
HV_OPT_SOR_RESIDENT_F32 = 3
HV_OPT_SOR_F32_LAG = 4
HV_OPT_SMEAR_TILE = 5
HV_OPT_BATCH_GROUP_KIB = 6
HV_OPT_SMEAR_FUSED = 7
HV_OPT_BATCH_GRAPH = 8
HV_OPT_SOR_F32_TILES_PER_CU = 9
HV_OPT_SOR_F32_OWNED = 10
HV_OPT_SOR_EXTRAPOLATE = 11
HV_OPT_SOR_F32_TOL_SHIFT = 12
HV_OPT_BATCH_DEVICE_STEER = 13
HV_OPT_SOR_F32_OMEGA = 14

HV_KCLASS_SOR = 0
HV_KCLASS_PARTICLE = 1
HV_KCLASS_SMEAR = 2
HV_KCLASS_POTENTIAL = 3
HV_KCLASS_SOR_F32 = 4
HV_KCLASS_SOR_RES_F32 = 5
HV_KCLASS_SOR_STAGE = 6
KCLASS_NAMES = ("sor_sweep", "particle_pass", "smear", "potential", "sor_sweep_f32", "sor_resident_f32",
                "sor_stage")


class HvWorld(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int * 3), ("d_max", ctypes.c_double * 3)]


class HvSkyrme(ctypes.Structure):
    _fields_ = [("A", ctypes.c_double), ("B", ctypes.c_double), ("C", ctypes.c_double),
                ("gamma", ctypes.c_double)]


class HvParams(ctypes.Structure):
    _fields_ = [
        ("t_f", ctypes.c_double),
        ("steps", ctypes.c_int),
        ("substeps", ctypes.c_int),
        ("t_exc", ctypes.c_double),
        ("eta_exc", ctypes.c_double),
        ("boost_mode", ctypes.c_int),
        ("z", ctypes.c_int),
        ("n", ctypes.c_int),
        ("sigma_r", ctypes.c_double),
        ("part_per_nucleon", ctypes.c_int),
        ("sor_omega", ctypes.c_double),
        ("sor_tol", ctypes.c_double),
        ("sor_max_it", ctypes.c_int),
    ]


class HvWoodsSaxon(ctypes.Structure):
    _fields_ = [("V0", ctypes.c_double), ("R12", ctypes.c_double), ("a", ctypes.c_double)]


class HvError(RuntimeError):
    """An entry point returned a negative status."""

    def __init__(self, fn: str, code: int, message: str):
        super().__init__(f"{fn} failed with status {code}: {message}")
        self.code = code


_P = ctypes.c_void_p
_D = ctypes.c_double
_I = ctypes.c_int
_DP = ctypes.POINTER(ctypes.c_double)
_IP = ctypes.POINTER(ctypes.c_int)

# name -> (restype, argtypes)
SIGNATURES = {
    "hv_ctx_create": (_I, [ctypes.POINTER(HvWorld), _I, _I, _I, ctypes.POINTER(_P)]),
    "hv_ctx_destroy": (_I, [_P]),
    "hv_last_error": (ctypes.c_char_p, [_P]),
    "hv_sync": (_I, [_P]),
    "hv_stream": (_P, [_P]),
    "hv_last_outside": (_I, [_P, _IP]),
    "hv_upload_particles": (_I, [_P] + [_P] * 9),
    "hv_download_particles": (_I, [_P] + [_P] * 9),
    "hv_upload_field": (_I, [_P, _I, _P]),
    "hv_download_field": (_I, [_P, _I, _P]),
    "hv_cic_deposit": (_I, [_P]),
    "hv_smear_density": (_I, [_P, _D, _I]),
    "hv_skyrme_potentials": (_I, [_P, ctypes.POINTER(HvSkyrme)]),
    "hv_set_coulomb_bc": (_I, [_P, _I]),
    "hv_coulomb_sor": (_I, [_P, _D, _D, _I, _IP, _DP]),
    "hv_merge_potentials": (_I, [_P]),
    "hv_forces_fdm": (_I, [_P]),
    "hv_cic_gather": (_I, [_P]),
    "hv_kick_half": (_I, [_P, _D]),
    "hv_drift": (_I, [_P, _D]),
    "hv_excite": (_I, [_P, _D, _I, _I, _I]),
    "hv_center_of_mass": (_I, [_P, _I, _DP]),
    "hv_mean_squared_radius": (_I, [_P, _I, _DP]),
    "hv_run_init": (_I, [_P, ctypes.POINTER(HvSkyrme), ctypes.POINTER(HvParams), _IP]),
    "hv_run_steps": (_I, [_P, ctypes.POINTER(HvSkyrme), ctypes.POINTER(HvParams), _I, _I, _P, _P]),
    "hv_simulate": (_I, [_P, ctypes.POINTER(HvSkyrme), ctypes.POINTER(HvParams), _P, _P]),
    "hv_timer_start": (_I, [_P]),
    "hv_timer_stop": (_I, [_P, _DP]),
    "hv_build_info": (ctypes.c_char_p, []),
    "hv_profile_enable": (_I, [_P, _I]),
    "hv_set_option": (_I, [_P, _I, _I]),
    "hv_get_option": (_I, [_P, _I, _IP]),
    "hv_profile_read": (_I, [_P, _I, _DP, ctypes.POINTER(ctypes.c_longlong)]),
    "hv_profile_issued": (_I, [_P, _I, ctypes.POINTER(ctypes.c_longlong)]),
    "hv_copy_bandwidth": (_I, [ctypes.c_longlong, _I, _DP, _DP]),
    "hv_stats_noalloc": (_I, [ctypes.POINTER(HvWorld), _I, _P, _P, _P, _I, _I, _P, _P]),
    "hv_total_energy": (_I, [_P, ctypes.POINTER(HvSkyrme), _I, _DP]),
    "hv_sor_counts": (_I, [_P, ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(ctypes.c_longlong),
                           ctypes.POINTER(ctypes.c_longlong)]),
    "hv_batch_create": (_I, [_I, ctypes.POINTER(HvWorld), _IP, _IP, _I, ctypes.POINTER(_P)]),
    "hv_batch_destroy": (_I, [_P]),
    "hv_batch_last_error": (ctypes.c_char_p, [_P]),
    "hv_batch_size": (_I, [_P]),
    "hv_batch_instance": (_P, [_P, _I]),
    "hv_batch_set_option": (_I, [_P, _I, _I]),
    "hv_batch_get_option": (_I, [_P, _I, _IP]),
    "hv_batch_run_init": (_I, [_P, ctypes.POINTER(HvSkyrme), ctypes.POINTER(HvParams), _IP]),
    "hv_batch_run_steps": (_I, [_P, ctypes.POINTER(HvSkyrme), ctypes.POINTER(HvParams), _I, _I, _P, _P]),
    "hv_batch_profile_enable": (_I, [_P, _I]),
    "hv_batch_profile_read": (_I, [_P, _I, _DP, ctypes.POINTER(ctypes.c_longlong)]),
    "hv_batch_sor_counts": (_I, [_P] + [ctypes.POINTER(ctypes.c_longlong)] * 4),
    "hv_batch_graph_info": (_I, [_P] + [ctypes.POINTER(ctypes.c_longlong)] * 2),
    "hv_write_vtk": (_I, [_P, ctypes.c_char_p]),
    "hv_vtk_write_fields": (_I, [ctypes.c_char_p, ctypes.POINTER(HvWorld)] + [_P] * 5 + [_I]),
    "hv_dipole_loss_grad": (_I, [_P, ctypes.POINTER(HvSkyrme), ctypes.POINTER(HvParams), _I, _DP, _I, _DP, _DP,
                                 _DP]),
    "hv_fermi_momentum": (_D, [_D]),
    "hv_sample_uniform_spheres": (_I, [ctypes.c_ulonglong, _I, _I, _D, _D] + [_P] * 6),
    "hv_mix64": (ctypes.c_ulonglong, [ctypes.c_ulonglong]),
    "hv_ws_nuclear_radius": (_D, [_I]),
    "hv_ws_max_particles": (_I, [_D, _D, _I]),
    "hv_ws_sample_round": (_I, [_I, _I, _I, _D, ctypes.POINTER(HvWoodsSaxon), _DP, _DP, _IP] + [_P] * 6),
    "hv_ws_particle_densities": (_I, [_P, _P, _P, _I, _I, _D, _I, _P, _P]),
    "hv_ws_fit": (_I, [_P] * 5 + [_I, _I, ctypes.POINTER(HvSkyrme), ctypes.POINTER(HvWoodsSaxon), _IP]),
    "hv_ws_chi_squared": (_I, [_P] * 5 + [_I, _I, ctypes.POINTER(HvSkyrme), ctypes.POINTER(HvWoodsSaxon), _I, _DP]),
    "hv_init_woods_saxon": (_I, [_I, _I, _I, ctypes.POINTER(HvSkyrme), _D, _D, _D, _D, _DP, ctypes.c_uint] + [_P] * 6
                            + [ctypes.POINTER(HvWoodsSaxon), _IP, _DP]),
}

_lib = None

# The ROCm runtime libhvamd.so is built against, loaded under the bare names
# that torch's own HIP libraries ask for (their DT_NEEDED entries carry no
# version suffix), so that a torch imported later in the process binds the
# same HIP and HSA runtime instead of loading the copies bundled in its wheel.
RUNTIME_PRELOAD = ("librocprofiler-register.so", "libamd_comgr.so", "libhsa-runtime64.so", "libamdhip64.so")


def _preload_runtime() -> None:
    if "torch" in sys.modules:
        warnings.warn("torch was imported before hvamd: libhvamd.so binds the HIP runtime torch loaded, not "
                      "the ROCm runtime it was built against; import hvamd first", RuntimeWarning, stacklevel=3)
        return
    for name in RUNTIME_PRELOAD:
        try:
            ctypes.CDLL(name, mode=ctypes.RTLD_GLOBAL)
        except OSError:
            pass  # not on the search path: the library's own RUNPATH finds the runtime


def load() -> ctypes.CDLL:
    """Load libhvamd.so once; raise if it is missing (no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(
            f"{LIB_PATH} is not built: run `make -C hyper-vlasov_amd` or __graft_entry__.build(); "
            "the step loop has no CPU fallback")
    _preload_runtime()
    lib = ctypes.CDLL(str(LIB_PATH))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def ptr(a: np.ndarray | None):
    if a is None:
        return None
    return ctypes.c_void_p(a.ctypes.data)
