"""ctypes loader for libhipotf.so (the HIP library built by binsim_amd/build.py).

There is no CPU fallback: if the library is missing or a call fails, an exception is raised.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
# HIPOTF_LIBRARY names another build of the same library (the host-sanitizer build of
# binsim_amd/build.py --asan); by default the in-tree libhipotf.so
LIB_PATH = os.environ.get("HIPOTF_LIBRARY") or os.path.join(HERE, "libhipotf.so")

_dp = C.POINTER(C.c_double)
_i64p = C.POINTER(C.c_int64)
_i32p = C.POINTER(C.c_int32)
_u32p = C.POINTER(C.c_uint32)
_dpp = C.POINTER(C.c_void_p)
_vp = C.c_void_p


class C64(C.Structure):
    """complex float64 by value (real, imaginary), the reference's complexf64"""
    _fields_ = [("re", C.c_double), ("im", C.c_double)]

# name -> (restype, argtypes); every symbol include/hipotf.h declares
SIGNATURES = {
    "hipotf_last_error": (C.c_char_p, []),
    "hipotf_clear_error": (None, []),
    "create_basis_manager": (_vp, [C.c_int64, C.c_int64, C.c_int64, C.c_int64, _i64p, C.c_int64]),
    "get_subspace_dim": (C.c_int64, [_vp]),
    "get_num_symmetry_blocks": (C.c_int64, [_vp]),
    "destroy_basis_manager": (None, [_vp]),
    "set_det_coeff_f64": (None, [_vp, C.c_uint32, C.c_uint32, C.c_double, _vp]),
    "set_det_coeff_c64": (None, [_vp, C.c_uint32, C.c_uint32, C64, _vp]),
    "build_network_otf_f64": (_vp, [_i64p, C.c_int64, C.c_int64, _u32p, _u32p, _i64p, _i64p, _i64p,
                                    _u32p, _u32p, _dp, _dp]),
    "destroy_network_otf_f64": (None, [_vp]),
    "sync_device_cuda": (None, []),
    "build_basisdev_f64": (_vp, [_vp]),
    "destroy_basisdev_f64": (None, [_vp]),
    "build_networkdev_f64": (_vp, [_vp]),
    "destroy_networkdev_f64": (None, [_vp]),
    "hvec_cuda": (None, [_vp, _vp, _vp, _vp]),
    "get_diags_elements_cuda": (None, [_vp, _vp, _vp]),
    "expm_cuda": (None, [_vp, _vp, C.c_int64, C.c_double, _vp]),
    "grad_cuda": (C.c_double, [_vp, _vp, C.c_int64, C.c_double, _vp, _vp]),
    "backgrad_cuda": (C.c_double, [_vp, _vp, C.c_int64, C.c_double, _vp, _vp]),
    "batchgrad_cuda": (None, [_vp, _vp, _vp, _vp, _vp, _vp]),
    "tvec_cuda": (None, [_vp, _vp, C.c_int64, _vp, _vp]),
    "backtran_cuda": (None, [_vp, _vp, C.c_int64, C.c_double, _vp, _vp, _vp]),
    "batchexpm_cuda": (None, [_vp, _vp, C.c_int64, C.c_double, _vp, C.c_int64, C.c_int64]),
    "batchtran_cuda": (None, [_vp, _vp, _vp, _vp, _vp]),
    "expm_cuda_2d": (None, [_vp, _vp, C.c_int64, C.c_double, _vp]),
    "grad_cuda_2d": (C.c_double, [_vp, _vp, C.c_int64, C.c_double, _vp, _vp]),
    "backgrad_cuda_2d": (C.c_double, [_vp, _vp, C.c_int64, C.c_double, _vp, _vp]),
    "hipotf_build_ham_f64": (_vp, [C.c_int64, _dp, _dp, C.c_double]),
    "hipotf_build_pool_f64": (_vp, [C.c_int64, C.c_int64, _i32p, _i32p]),
    "hipotf_num_generators": (C.c_int64, [_vp]),
    "hipotf_prepare": (C.c_int, [_vp, _vp]),
    "hipotf_set_tuning": (C.c_int, [C.c_int, _i32p, C.c_int]),
    "hipotf_get_tuning": (C.c_int, [C.c_int, _i32p, C.c_int]),
    "hipotf_hvec_block": (C.c_int, [_vp, _vp, C.c_int, _dpp, _dpp, C.c_int]),
    "hipotf_hvec_noalloc": (C.c_int, [_vp, _vp, _vp, _vp]),
    "hipotf_vqe_energy_grad": (C.c_int, [_vp, _vp, _vp, C.c_int64, _i64p, _dp, _vp, _vp, _dp, _dp]),
    "hipotf_vqe_energy": (C.c_int, [_vp, _vp, _vp, C.c_int64, _i64p, _dp, _vp, _vp, _dp]),
    "hipotf_network_kind": (C.c_int, [_vp]),
    "hipotf_ab_timing": (C.c_int, [C.c_int, _dp, _i64p, _i64p]),
    "hipotf_davidson_f64": (C.c_int, [_vp, _vp, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_int,
                                      C.c_int, C.c_int, _vp, _dp, _vp, _dp]),
    "build_ritz_and_residual_f64": (C.c_double, [C.c_int64, C.c_int, _dp, C.c_double, _dpp, _dpp, _vp]),
    "build_ritz_vector_f64": (None, [C.c_int64, C.c_int, _dp, _dpp, _vp]),
    "apply_preconditioner_inplace_f64": (None, [C.c_int64, _vp, _vp, C.c_double, C.c_double]),
    "mgs_orthogonalize_f64": (C.c_double, [C.c_int64, C.c_int, _dpp, _vp]),
    "inplace_normalize_f64": (C.c_double, [C.c_int64, _vp]),
    "fast_dot_f64": (C.c_double, [C.c_int64, _vp, _vp]),
    "compute_heff_col_f64": (None, [C.c_int64, C.c_int, _dpp, _vp, _dp]),
    "rebuild_heff_f64": (None, [C.c_int64, C.c_int, _dpp, _dpp, _dp, C.c_int]),
}

_lib = None


class HipOTFError(RuntimeError):
    pass


def load() -> C.CDLL:
    """Load libhipotf.so once; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise HipOTFError(f"{LIB_PATH} is missing: build it with `python -m binsim_amd.build`")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def last_error() -> str:
    msg = load().hipotf_last_error()
    return msg.decode() if msg else ""


def check(ok: bool, what: str):
    """Raise with the library's last error message if ``ok`` is false."""
    if not ok:
        raise HipOTFError(f"{what}: {last_error()}")


def call_void(name: str, *args):
    """Call a void entry point and raise if it recorded an error."""
    lib = load()
    lib.hipotf_clear_error()
    getattr(lib, name)(*args)
    err = last_error()
    if err:
        raise HipOTFError(f"{name}: {err}")
