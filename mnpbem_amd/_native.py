"""ctypes binding of ``libmnpbem_amd.so`` (the C ABI declared in
include/mnpbem_amd.h).

The library is required: there is no CPU fallback for the BEM path.  Loading
fails with ImportError when the shared object has not been built
(``python -m mnpbem_amd.build``), and every entry point raises RuntimeError
with the library's message when it returns a non-zero status.
"""

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmnpbem_amd.so")
# GEMM tile table written by mnpbem_amd.autotune (applied at load when present;
# both tiles accumulate in the same order, so the table changes speed only)
TUNING_PATH = os.path.join(_HERE, "data", "gemm_tuning_gfx950.json")

DTYPE_C64 = 1
DTYPE_C128 = 2
DTYPE_C64X = 3  # complex64 data, GEMMs on bf16 matrix cores (3-term split products)

_i, _l, _d, _p = ctypes.c_int, ctypes.c_long, ctypes.c_double, ctypes.c_void_p

# name -> argument types (all return int status)
SIGNATURES = {
    "mnpbem_abi_version": [],
    "mnpbem_device_sync": [],
    "mnpbem_cgemm_sub_batched": [_i, _i, _i, _i, _p, _l, _l, _p, _l, _l, _p, _l, _l, _i, _p],
    "mnpbem_cgemm_sub_variant": [_i, _i, _i, _i, _p, _l, _l, _p, _l, _l, _p, _l, _l, _i, _p],
    "mnpbem_gemm_tuning_set": [_p, _i],
    "mnpbem_gemm_tuning_get": [_p, _i],
    "mnpbem_gemm_tuning_seen": [_p, _i, _i],
    "mnpbem_getrf_batched": [_i, _i, _p, _l, _l, _p, _l, _p, _i, _p],
    "mnpbem_getrf_nopiv_batched": [_i, _i, _p, _l, _l, _p, _l, _p, _p, _i, _p],
    "mnpbem_getrs_batched": [_i, _i, _i, _p, _l, _l, _p, _l, _p, _l, _l, _i, _p],
    "mnpbem_refine_coeffs": [_i, _i, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _i, _p, _i, _p, _i, _p, _p, _p,
                             _p],
    "mnpbem_closed_correction": [_i, _p, _p, _p, _p, _p, _p, _p, _p, _p, _d, _i, _p, _p, _p],
    "mnpbem_green_assemble": [_i, _i, _i, _p, _p, _p, _p, _p, _i, _i, _i, _p, _p, _p, _p, _d, _p, _p, _l, _l,
                              _p, _p, _i, _p, _i, _p],
    "mnpbem_gemv_t_batched": [_i, _i, _i, _p, _l, _l, _p, _l, _l, _p, _l, _l, _d, _d, _d, _i, _p],
    "mnpbem_lu_solve_t_batched": [_i, _i, _i, _p, _l, _l, _p, _l, _p, _l, _l, _p, _l, _l, _p, _l, _i, _p],
    "mnpbem_lu_blockinv_batched": [_i, _i, _p, _l, _l, _p, _l, _i, _p],
    "mnpbem_identity_batched": [_i, _i, _p, _l, _l, _i, _p],
    "mnpbem_sigma_batched": [_i, _i, _i, _p, _p, _p, _p, _l, _l, _p, _p, _i, _i, _p],
    "mnpbem_bem_solve_batched": [_i, _i, _i, _p, _p, _p, _p, _p, _p, _p, _p, _l, _l, _l, _p, _p, _p, _l, _p,
                                 _p, _p, _l, _l, _p, _p, _i, _i, _p],
    "mnpbem_bem_solve_inv_batched": [_i, _i, _i, _p, _p, _p, _p, _p, _l, _l, _p, _p, _p, _l, _l, _p, _p, _i, _i,
                                     _p],
    "mnpbem_bem_solve_batched_noalloc": [_i, _i, _i, _p, _p, _p, _p, _p, _p, _p, _p, _l, _l, _l, _p, _p, _p, _l,
                                         _p, _p, _l, _l, _p, _p, _i, _i, _p],
    "mnpbem_bem_residual_batched": [_i, _i, _p, _p, _p, _p, _l, _l, _p, _p, _p, _l, _l, _p, _l, _l, _p, _p,
                                    _p, _i, _i, _p],
    "mnpbem_bem_solve_adjoint_batched": [_i, _i, _i, _p, _p, _p, _p, _p, _p, _p, _p, _l, _l, _l, _p, _p, _p, _l,
                                         _p, _p, _p, _l, _l, _p, _p, _i, _p],
    "mnpbem_bem_solve_adjoint_inv_batched": [_i, _i, _i, _p, _p, _p, _p, _p, _l, _l, _p, _p, _p, _l, _l, _p, _p,
                                             _i, _p],
    "mnpbem_bem_residual_adjoint_batched": [_i, _i, _p, _p, _p, _p, _l, _l, _p, _p, _p, _l, _l, _p, _p, _l, _l,
                                            _p, _p, _p, _i, _p],
    "mnpbem_gemv_n_batched": [_i, _i, _i, _p, _l, _l, _p, _l, _l, _p, _l, _l, _d, _i, _p],
    "mnpbem_lu_solve_n_batched": [_i, _i, _i, _p, _l, _l, _p, _l, _p, _l, _l, _p, _l, _l, _p, _l, _i, _p],
    "mnpbem_ext_functional": [_i, _i, _p, _p, _p, _p, _p, _p, _p, _p, _l, _l, _i, _p],
    "mnpbem_grad_prep": [_i, _i, _p, _p, _l, _l, _p, _p, _p, _p, _p, _i, _p],
    "mnpbem_grad_far": [_i, _i, _p, _p, _p, _p, _p, _p, _p, _i, _p, _p, _p, _p, _p, _p, _p, _p, _i, _p],
    "mnpbem_grad_refined": [_i, _i, _i, _i] + [_p] * 13 + [_i, _p, _i, _p, _i, _p, _p, _p, _p, _p, _p, _p, _l,
                                                         _p, _i, _p],
    "mnpbem_grad_local": [_i, _i, _p, _p, _p, _p, _l, _l] + [_p] * 11 + [_i, _p],
    "mnpbem_getrf_tall_rows": [_i],
    "mnpbem_trsm_rows_max": [_i],
    "mnpbem_trsm_left_rhs": [_i],
    "mnpbem_gemm_thin_max": [_i],
    "mnpbem_gemv_mfma": [_i],
    "mnpbem_trsv_mfma": [_i],
    "mnpbem_trsm_leaf_fused": [_i],
    "mnpbem_getri_batched": [_i, _i, _p, _l, _l, _p, _l, _p, _l, _l, _p, _l, _i, _p],
    "mnpbem_getrs_blockinv_batched": [_i, _i, _i, _p, _l, _l, _p, _l, _p, _l, _i, _p, _l, _l, _i, _p],
    "mnpbem_getri_blockinv_batched": [_i, _i, _p, _l, _l, _p, _l, _p, _l, _i, _p, _l, _l, _p, _l, _i, _p],
    "mnpbem_getri_nopiv_blockinv_batched": [_i, _i, _p, _l, _l, _p, _l, _i, _p, _l, _l, _i, _p],
    "mnpbem_lu_blockinv128_batched": [_i, _i, _p, _l, _l, _p, _l, _p, _l, _i, _p],
    "mnpbem_accumulate": [_i, _l, _p, _p, _p],
    "mnpbem_convert": [_i, _l, _p, _p, _p],
    "mnpbem_planewave_batched": [_i, _i, _p, _p, _p, _p, _p, _p, _p, _p, _l, _l, _p, _l, _i, _i, _p],
    "mnpbem_extinction_batched": [_i, _i, _p, _l, _l, _p, _p, _p, _p, _p, _p, _p, _p, _i, _i, _p],
    "mnpbem_bem_combine_batched": [_i, _i, _p, _l, _l, _p, _p, _l, _l, _p, _p, _p, _i, _i, _p],
    "mnpbem_widen": [_i, _l, _p, _p, _p],
    "mnpbem_mv_dot_batched": [_l, _i, _i, _p, _l, _l, _p, _l, _p, _p, _i, _p],
    "mnpbem_mv_update_batched": [_l, _i, _i, _p, _l, _l, _p, _d, _p, _l, _i, _p],
    "mnpbem_mv_scale_batched": [_l, _i, _i, _p, _p, _l, _p, _l, _i, _p],
    "mnpbem_hmat_aca": [_i, _p, _p, _p, _p, _l, _l, _p, _p, _p, _i, _d, _p, _p],
    "mnpbem_hmat_dense": [_i, _i, _p, _p, _p, _p, _l, _l, _p, _p],
    "mnpbem_hmat_aca_green": [_i, _p, _p, _p, _i, _d, _p, _p, _p, _p, _p, _i, _p, _p, _p, _p, _p, _p, _p, _i, _d,
                              _p, _p],
    "mnpbem_hmat_dense_green": [_i, _i, _p, _p, _p, _i, _d, _p, _p, _p, _p, _p, _i, _p, _p, _p, _p, _p, _p],
    "mnpbem_hmat_compact": [_i, _i, _p, _p, _p, _p, _p, _p, _p, _p],
    "mnpbem_hmat_matvec": [_i, _i, _i, _i, _i, _p, _p, _p, _p, _p, _p, _p, _i, _p, _l, _p, _l, _i, _p, _p, _p],
    "mnpbem_hmat_pack": [_i, _i, _p, _i, _p, _p, _p, _p, _p, _p, _p, _p, _p],
    "mnpbem_hmat_matvec_packed": [_i, _i, _i, _p, _p, _p, _i, _p, _p, _p, _p, _i, _p, _l, _p, _l, _i, _p, _p],
    "mnpbem_hmat_full": [_i, _i, _p, _p, _p, _p, _p, _p, _l, _p],
    "mnpbem_hmat_truncate": [_i, _i, _p, _p, _p, _p, _p, _i, _d, _p, _p, _p, _p, _p],
    "mnpbem_hmat_trunc_form": [_i, _i, _p, _p, _p, _p, _p, _p, _i, _i, _p, _p, _p],
    "mnpbem_hmat_axpby": [_i, _l, _d, _d, _p, _d, _d, _p, _p, _p],
    "mnpbem_hmat_concat": [_i, _i, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p],
    "mnpbem_hmat_diag": [_i, _i, _p, _p, _p, _p, _p, _i, _p],
    "mnpbem_scattering_batched": [_i, _i, _p, _l, _l, _p, _p, _p, _p, _i, _p, _p, _p, _p, _p, _p, _p, _i, _p],
    "mnpbem_refine_classify": [_i, _p, _p, _d, _d, _i, _p, _p, _p, _p, _p, _p],
    "mnpbem_surface_field": [_i, _i, _i, _d, _p, _p, _p, _p, _p, _i, _p, _i, _p, _p, _p, _p, _p, _l, _l, _p, _p,
                             _l, _l, _i, _p],
    "mnpbem_green_apply": [_i, _i, _d, _p, _p, _p, _p, _p, _i, _i, _p, _p, _p, _p, _p, _l, _l, _i, _p, _l, _l, _i,
                           _i, _i, _p],
    "mnpbem_green_deriv": [_i, _d, _p, _p, _p, _p, _p, _i, _i, _p, _p, _p, _p, _p, _i, _p],
    "mnpbem_green_select": [_i, _i, _d, _p, _p, _p, _p, _p, _i, _i, _p, _p, _p, _p, _l, _p, _p, _p, _p],
    "mnpbem_excitation": [_i, _i, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _d, _p, _l, _p],
    "mnpbem_green_dk": [_i, _p, _p, _p, _p, _p, _i, _i, _i, _p, _p, _p, _p, _p, _p, _l, _l, _i, _p],
    "mnpbem_deps_rhs": [_i, _i, _p, _l, _l, _p, _l, _l, _p, _p, _p, _i, _p],
    "mnpbem_prof_begin": [_i],
    "mnpbem_prof_end": [_p, _p, _p],
    "mnpbem_prof_bytes": [_p],
    "mnpbem_prof_cbytes": [_p],
    "mnpbem_prof_launches": [_i, _p, _p],
}

_lib = None


def load(path=None):
    """Load the shared library and declare its entry points."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    path = path or LIB_PATH
    if not os.path.exists(path):
        raise ImportError("mnpbem_amd: HIP library %s is not built; run `python -m mnpbem_amd.build`" % path)
    lib = ctypes.CDLL(path)
    for name, argtypes in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = argtypes
        fn.restype = ctypes.c_int
    lib.mnpbem_last_error.argtypes = []
    lib.mnpbem_last_error.restype = ctypes.c_char_p
    if path == LIB_PATH:
        _apply_tuning(lib)
        _lib = lib
    return lib


def _apply_tuning(lib):
    if not os.path.exists(TUNING_PATH):
        return
    import json
    with open(TUNING_PATH) as fh:
        table = json.load(fh)["table"]
    buf = (ctypes.c_ubyte * len(table))(*table)
    if lib.mnpbem_gemm_tuning_set(buf, len(table)) != 0:
        raise RuntimeError("mnpbem_amd: bad GEMM tuning table %s: %s"
                           % (TUNING_PATH, lib.mnpbem_last_error().decode(errors="replace")))


def call(name, *args):
    """Call an entry point and raise RuntimeError on a non-zero status."""
    lib = load()
    rc = getattr(lib, name)(*args)
    if rc != 0:
        msg = lib.mnpbem_last_error().decode(errors="replace")
        raise RuntimeError("%s failed (status %d): %s" % (name, rc, msg))
    return rc


def ptr(t):
    """Raw device (or host) address of a torch tensor, or None."""
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def ptr_off(t, elems):
    """Address of element ``elems`` (counted in t's element type) of a
    contiguous torch tensor."""
    return ctypes.c_void_p(t.data_ptr() + int(elems) * t.element_size())


def stream_handle(device=None):
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)
