"""hvamd: Python layer of the MI355X (gfx950) hyper-vlasov test-particle loop.

The compute lives in libhvamd.so (HIP kernels behind the C-ABI of
include/hv_abi.h); this package binds it with ctypes and adds the workload
builders and post-processing.
"""
from . import _lib as _lib_module
from ._lib import (HV_BOOST_REF_WEIGHTED, HV_BOOST_SYMMETRIC, HV_FIELD_COULOMB, HV_FIELD_DENSITY,
                   HV_FIELD_FORCES_X, HV_FIELD_FORCES_Y, HV_FIELD_FORCES_Z, HV_FIELD_POTENTIALS,
                   HV_NEUTRONS, HV_OPT_BATCH_DEVICE_STEER, HV_OPT_BATCH_GRAPH, HV_OPT_BATCH_GROUP_KIB, HV_OPT_SOR_F32_LAG, HV_OPT_SOR_F32_OMEGA, HV_OPT_SOR_F32_TOL_SHIFT, HV_OPT_SOR_EXTRAPOLATE, HV_OPT_SOR_F32_OWNED, HV_OPT_SOR_F32_TILES_PER_CU, HV_OPT_SOR_FUSED, HV_OPT_SOR_RESIDENT,
                   HV_OPT_SOR_RESIDENT_F32, HV_OPT_SORT_EVERY, HV_OPT_SMEAR_FUSED, HV_OPT_SMEAR_TILE, HV_PREC_F64, HV_PREC_MIXED, HV_PROTONS, HvError,
                   load)
from .sim import (HBARC, MC2, RHO0, Params, Simulation, Skyrme, World, dipole,
                  sample_uniform_spheres, sigma_from_fwhm)

__all__ = [
    "HV_BOOST_REF_WEIGHTED", "HV_BOOST_SYMMETRIC", "HV_FIELD_COULOMB", "HV_FIELD_DENSITY",
    "HV_FIELD_FORCES_X", "HV_FIELD_FORCES_Y", "HV_FIELD_FORCES_Z", "HV_FIELD_POTENTIALS",
    "HV_NEUTRONS", "HV_OPT_BATCH_DEVICE_STEER", "HV_OPT_BATCH_GRAPH", "HV_OPT_BATCH_GROUP_KIB", "HV_OPT_SOR_F32_LAG", "HV_OPT_SOR_F32_OMEGA", "HV_OPT_SOR_F32_TOL_SHIFT", "HV_OPT_SOR_EXTRAPOLATE", "HV_OPT_SOR_F32_OWNED", "HV_OPT_SOR_F32_TILES_PER_CU", "HV_OPT_SOR_FUSED", "HV_OPT_SOR_RESIDENT", "HV_OPT_SOR_RESIDENT_F32",
    "HV_OPT_SORT_EVERY", "HV_OPT_SMEAR_FUSED", "HV_OPT_SMEAR_TILE", "HV_PREC_F64", "HV_PREC_MIXED", "HV_PROTONS", "HvError", "load",
    "HBARC", "MC2", "RHO0", "Params", "Simulation", "Skyrme", "World", "dipole",
    "sample_uniform_spheres", "sigma_from_fwhm",
]

# the ROCm runtime libhvamd.so is built against, in place before anything
# imports torch (which would otherwise bring the HIP runtime of its wheel)
_lib_module._preload_runtime()
