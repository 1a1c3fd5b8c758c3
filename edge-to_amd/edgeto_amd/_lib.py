"""ctypes binding of libedgeto_amd.so (the C ABI in include/edgeto_amd.h).

The library is loaded from this package directory.  There is no CPU
fallback: if the library is missing or no HIP device is visible, every entry
point raises.  Tensors are torch device tensors used purely as memory; all
arithmetic runs in the library's HIP kernels on torch's current stream.
"""
import ctypes
import os

import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libedgeto_amd.so")

_lib = None

ET_OK, ET_EINVAL, ET_EHIP, ET_ENONFINITE, ET_EBREAKDOWN, ET_ENOCONV, ET_ENOMEM = range(7)


class EdgetoError(RuntimeError):
    """A C-ABI call returned a non-zero status."""

    def __init__(self, name, status, msg):
        super().__init__(f"{name} failed (status {status}): {msg}")
        self.status = status


class MGParams(ctypes.Structure):
    _fields_ = [("dim", ctypes.c_int), ("nel", ctypes.c_int * 3), ("n_level", ctypes.c_int),
                ("n_smooth", ctypes.c_int), ("omega", ctypes.c_double),
                ("min_omega", ctypes.c_double), ("cycle_w", ctypes.c_int),
                ("w_level_mask", ctypes.c_int), ("precision", ctypes.c_int)]


class UMesh(ctypes.Structure):
    """et_umesh: a general mesh's device arrays (include/edgeto_amd.h)."""
    _fields_ = [("n_nodes", ctypes.c_int64), ("n_el", ctypes.c_int64), ("dof", ctypes.c_int),
                ("elem_ptr", ctypes.c_void_p), ("elem_nodes", ctypes.c_void_p), ("K", ctypes.c_void_p),
                ("K_ptr", ctypes.c_void_p), ("star_ptr", ctypes.c_void_p), ("star", ctypes.c_void_p),
                ("el_of", ctypes.c_void_p)]


_P = ctypes.c_void_p
_I = ctypes.c_int
_I64 = ctypes.c_int64
_D = ctypes.c_double
_IP = ctypes.POINTER(ctypes.c_int)

_SIGS = {
    "et_tune_set": [_I, _I],
    "et_mg_apply_level": [_P, _I, _P, _P],
    "et_mg_coarse_solve": [_P, _P, _P],
    "et_tune_get": [_I],
    "et_generate_elements": [_I, _IP, _P, _P],
    "et_k0_structure": [_I, ctypes.POINTER(ctypes.c_double), _IP],
    "et_matvec_f64": [_I, _IP, _P, _P, _P, _P, _P, _P],
    "et_matvec_f32": [_I, _IP, _P, _P, _P, _P, _P, _P],
    "et_matvec_f32_mfma": [_I, _IP, ctypes.POINTER(ctypes.c_double), _P, _P, _P, _P, _P, _P],
    "et_residual_f64": [_I, _IP, _P, _P, _P, _P, _P, _P, _P],
    "et_residual_f32": [_I, _IP, _P, _P, _P, _P, _P, _P, _P],
    "et_jacobi_f64": [_I, _IP, _P, _P, _P, _P, _P, _P, ctypes.c_double, _P, _P],
    "et_jacobi_f32": [_I, _IP, _P, _P, _P, _P, _P, _P, ctypes.c_double, _P, _P],
    "et_diagonal_f64": [_I, _IP, _P, _P, _P, _P, _P],
    "et_elem_energy_f64": [_I, _IP, _P, _P, _P, _P],
    "et_sensitivity_f64": [_I, _IP, _P, _P, _P, _D, _D, _P, _P],
    "et_simp_forward_f64": [_I64, _P, _D, _D, _I, _P, _P],
    "et_fea_integrals_f64": [_I, _IP, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P],
    "et_filter_normalizer_f64": [_I, _IP, _I, _P, _P, _P, _P],
    "et_filter_fwd_f64": [_I, _IP, _I, _P, _P, _P, _P, _P],
    "et_filter_adj_f64": [_I, _IP, _I, _P, _P, _P, _P, _P, _P],
    "et_simp_adjoint_f64": [_I, _IP, _P, _P, _P, _D, _D, _I, _P, _P, _P, _I, _P, _P],
    "et_restrict_f64": [_I, _IP, _I, _P, _P, _P],
    "et_prolong_f64": [_I, _IP, _I, _P, _P, _P],
    "et_stencil_from_operator_f64": [_I, _IP, _P, _P, _P, _P, _P],
    "et_galerkin_f64": [_I, _IP, _I, _P, _P, _P],
    "et_stencil_apply_f64": [_I, _IP, _I, _P, _P, _P, _P],
    "et_galerkin_elements_f64": [_I, _IP, ctypes.POINTER(ctypes.c_double), _P, _P, _P, _P],
    "et_dot_f64": [_I64, _P, _P, _P, _P],
    "et_sqdist_f64": [_I64, _P, _P, _P, _P],
    "et_oc_work_size": [_I64],
    "et_multimat_f64": [_I64, _I, _P, _P, _D, _D, _I, _I, _P, _P],
    "et_oc_multi_work_size": [_I64, _I],
    "et_oc_update_multi_f64": [_I64, _I, _P, _P, _D, _D, _D, ctypes.POINTER(ctypes.c_double),
                               ctypes.POINTER(ctypes.c_double), _P, _P, ctypes.POINTER(ctypes.c_int), _P],
    "et_krylov_work_size": [_I64, _I],
    "et_cg_f64": [_I, _IP, _P, _P, _P, _P, _P, _D, _I, _P, _I64, ctypes.POINTER(ctypes.c_int),
                  ctypes.POINTER(ctypes.c_double), _P],
    "et_gmres_f64": [_I, _IP, _P, _P, _P, _P, _P, _D, _I, _I, _P, _I64, ctypes.POINTER(ctypes.c_int),
                     ctypes.POINTER(ctypes.c_double), _P],
    "et_oc_update_f64": [_I64, _P, _P, _D, _D, _D, _D, _P, _P, ctypes.POINTER(ctypes.c_int), _P],
    "et_oc_update_f64_noalloc": [_I64, _P, _P, _D, _D, _D, _D, _P, ctypes.POINTER(ctypes.c_int), _P],
    "et_oc_noalloc_max_n": [],
    "et_batch_create": [ctypes.POINTER(_P), _I, _P, _I],
    "et_batch_run": [_P, _I, _I, ctypes.POINTER(ctypes.c_double)],
    "et_batch_result": [_P, _I, _P, _P, _P, _P],
    "et_batch_status": [_P, _I, _P, _P],
    "et_batch_destroy": [_P],
    "et_oc_update_elem_f64": [_I64, _P, _P, _P, _D, _D, _D, _P, _P, ctypes.POINTER(ctypes.c_int), _P],
    "et_umesh_star": [_I64, _I64, _P, _P, _P, _P, _P, _P],
    "et_umatvec_f64": [ctypes.POINTER(UMesh), _P, _P, _P, _P, _P],
    "et_udiagonal_f64": [ctypes.POINTER(UMesh), _P, _P, _P, _P],
    "et_uelem_energy_f64": [ctypes.POINTER(UMesh), _P, _P, _P],
    "et_usensitivity_f64": [ctypes.POINTER(UMesh), _P, _P, _D, _D, _P, _P],
    "et_ufea_integrals_f64": [ctypes.POINTER(UMesh), _I, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P],
    "et_csr_spmv_f64": [_I64, _P, _P, _P, _P, _P, _P],
    "et_umesh_pattern": [ctypes.POINTER(UMesh), _P, _P, _I64, ctypes.POINTER(ctypes.c_int64), _P],
    "et_uassemble_f64": [ctypes.POINTER(UMesh), _P, _P, _P, _P, _P, _P],
    "et_csr_cg_f64": [_I64, _P, _P, _P, _P, _P, _D, _I, _P, _I64, ctypes.POINTER(ctypes.c_int),
                      ctypes.POINTER(ctypes.c_double), _P],
    "et_csr_gmres_f64": [_I64, _P, _P, _P, _P, _P, _D, _I, _I, _P, _I64, ctypes.POINTER(ctypes.c_int),
                         ctypes.POINTER(ctypes.c_double), _P],
    "et_ucg_f64": [ctypes.POINTER(UMesh), _P, _P, _P, _P, _D, _I, _P, _I64, ctypes.POINTER(ctypes.c_int),
                   ctypes.POINTER(ctypes.c_double), _P],
    "et_ugmres_f64": [ctypes.POINTER(UMesh), _P, _P, _P, _P, _D, _I, _I, _P, _I64, ctypes.POINTER(ctypes.c_int),
                      ctypes.POINTER(ctypes.c_double), _P],
    "et_mg_create": [ctypes.POINTER(_P), ctypes.POINTER(MGParams), ctypes.POINTER(ctypes.c_double), _P],
    "et_mg_destroy": [_P],
    "et_mg_solve": [_P, _P, _P, _P, _P, _D, _I, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double)],
    "et_mg_last_timings": [_P, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)],
    "et_mg_profile": [_P, _I],
    "et_mg_profile_read": [_P, _P, _P],
}

_I64_RET = ("et_oc_work_size", "et_oc_noalloc_max_n", "et_krylov_work_size", "et_oc_multi_work_size")

PROF_KINDS = ["l0_jacobi", "l0_residual", "l0_jacobi_add", "l0_scale", "matvec64", "coarse_ops",
              "transfer", "coarse_solve", "pcg_vector", "setup", "l0_jacobi0"]

EXPORTED = sorted(list(_SIGS) + ["et_last_error", "et_version"])


def load(require_gpu=True):
    """Load the shared library (once).  Raises if missing or if no GPU is visible."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"edgeto_amd native library not built: {LIB_PATH} "
                           "(run `python edge-to_amd/build.py`)")
    lib = ctypes.CDLL(LIB_PATH)
    for name, args in _SIGS.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = ctypes.c_int64 if name in _I64_RET else ctypes.c_int
    lib.et_last_error.restype = ctypes.c_char_p
    lib.et_last_error.argtypes = []
    lib.et_version.restype = ctypes.c_char_p
    lib.et_version.argtypes = []
    if require_gpu and not torch.cuda.is_available():
        raise RuntimeError("edgeto_amd requires a HIP device (torch.cuda.is_available() is False)")
    _lib = lib
    return lib


def lib():
    return _lib if _lib is not None else load()


def check(name, status):
    if status != ET_OK:
        msg = lib().et_last_error().decode(errors="replace")
        raise EdgetoError(name, status, msg)


def call(name, *args):
    check(name, getattr(lib(), name)(*args))


def stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def ptr(t, dtype=None, numel=None):
    """Device pointer of a contiguous CUDA tensor (or None -> NULL).

    dtype / numel, when given, must match: the kernels read and write exactly
    numel values of that type, so a mismatch would be an out-of-bounds access."""
    if t is None:
        return None
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise ValueError("expected a device tensor")
    if not t.is_contiguous():
        raise ValueError("expected a contiguous tensor")
    if dtype is not None and t.dtype != dtype:
        raise ValueError(f"expected a {dtype} tensor, got {t.dtype}")
    if numel is not None and t.numel() != numel:
        raise ValueError(f"expected {numel} values, got {t.numel()}")
    return ctypes.c_void_p(t.data_ptr())


def f64(t, numel=None):
    return ptr(t, torch.float64, numel)


def f32(t, numel=None):
    return ptr(t, torch.float32, numel)


def i32(t, numel=None):
    return ptr(t, torch.int32, numel)


def i64(t, numel=None):
    return ptr(t, torch.int64, numel)


def u8(t, numel=None):
    """bool or uint8 tensor (one byte per entry)."""
    if t is not None and isinstance(t, torch.Tensor) and t.dtype not in (torch.bool, torch.uint8):
        raise ValueError(f"expected a bool/uint8 tensor, got {t.dtype}")
    return ptr(t, None, numel)


def nel_arg(nel):
    a = (ctypes.c_int * 3)(0, 0, 0)
    for i, v in enumerate(nel):
        a[i] = int(v)
    return a
