"""GMG-preconditioned CG on the device.

Mirrors EDGETO/solvers/_multigrid_cuda.py (CuMultiGrid / MultiGrid, :15-242)
and the Solver base (solvers/_base.py:1-11): the same constructor arguments,
solve(rhs, rho=None, use_last=True) -> (x, residual) with a warm start from
the previous solution, and reset().  The whole solve runs inside the native
library (et_mg_solve): hierarchy construction on the device, V-cycles, the
PCG loop and its convergence test.

Extension beyond the reference: precision='mixed' runs the V-cycle in FP32
inside the FP64 outer loop.  The reference's 'W' default behaves as a
V-cycle (it tests for a lowercase 'w', :165); the same rule applies here.
"""
import ctypes
import time
import logging

import numpy as np
import torch

from . import _lib

logger = logging.getLogger(__name__)


class Solver:
    def __init__(self):
        pass

    def __call__(self, *args, **kwargs):
        return self.solve(*args, **kwargs)

    def solve(self, *args, **kwargs):
        pass

    def reset(self):
        pass


class MultiGrid(Solver):
    def __init__(self, mesh, kernel, maxiter=1000, tol=1e-5, n_smooth=3, omega=0.5, n_level=3,
                 cycle="W", w_level=1, coarse_solver="splu", matrix_free=False,
                 low_level_tol=1e-8, low_level_maxiter=5000, min_omega=0.4, omega_boost=1.06,
                 precision="fp64", strict=False):
        super().__init__()
        # the reference's checks, in its order (solvers/_multigrid_cuda.py:44-54)
        if coarse_solver in ("splu", "spsolve") and matrix_free:
            raise ValueError("Matrix free is not supported with splu solver use cg instead.")
        if coarse_solver not in ("cg", "gmres", "splu", "spsolve", "cholesky"):
            raise ValueError("Coarse solver not recognized.")
        if matrix_free:
            raise NotImplementedError("Matrix free is not implemented yet.")
        if precision not in ("fp64", "mixed"):
            raise ValueError("precision must be 'fp64' or 'mixed'")
        self.kernel = kernel
        self.mesh = mesh
        self.last_x0 = None
        self.tol = tol
        self.maxiter = maxiter
        self.n_smooth = n_smooth
        self.omega = omega
        self.max_omega = omega
        self.min_omega = max(omega / 2, min_omega)
        self.d_omega = self.max_omega - self.min_omega
        self.omega_boost = omega_boost
        self.n_level = n_level
        self.cycle = cycle
        self.w_level = [w_level] if isinstance(w_level, int) else list(w_level)
        self.coarse_solver = coarse_solver  # the coarsest grid is always solved directly (FP64 Cholesky)
        self.dof = mesh.dof
        self.precision = precision
        self.strict = strict
        self.last_iterations = 0
        self.last_timings = (0.0, 0.0)
        self._handle = None
        self._create()

    def _create(self):
        p = _lib.MGParams()
        p.dim = len(self.mesh.nel)
        for i, v in enumerate(self.mesh.nel):
            p.nel[i] = int(v)
        p.n_level = int(self.n_level)
        p.n_smooth = int(self.n_smooth)
        p.omega = float(self.max_omega)
        p.min_omega = float(self.min_omega)  # min = max(omega/2, min_omega) already applied
        p.cycle_w = 1 if self.cycle == "w" else 0
        mask = 0
        for l in self.w_level:
            if 0 <= int(l) < 31:
                mask |= 1 << int(l)
        p.w_level_mask = mask
        p.precision = 1 if self.precision == "mixed" else 0
        K0 = np.ascontiguousarray(self.mesh.K_single_host, dtype=np.float64)
        h = ctypes.c_void_p()
        _lib.call("et_mg_create", ctypes.byref(h), ctypes.byref(p),
                  K0.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), _lib.stream())
        self._handle = h

    def __del__(self):
        h = getattr(self, "_handle", None)
        if h is not None and h.value and _lib._lib is not None:
            _lib._lib.et_mg_destroy(h)
            self._handle = None

    def reset(self):
        self.last_x0 = None

    def solve(self, rhs, rho=None, use_last=True):
        if not (self.kernel.has_rho or rho is not None):
            raise ValueError("Solver requires a density vector to be passed or set on the kernel.")
        if rho is not None:
            self.kernel.set_rho(rho)
        if not (isinstance(rhs, torch.Tensor) and rhs.is_cuda and rhs.dtype == torch.float64):
            raise ValueError("rhs must be a float64 device tensor")
        if use_last and self.last_x0 is not None:
            x = self.last_x0
        else:
            x = torch.zeros_like(rhs)
        rhs = rhs.contiguous()
        it = ctypes.c_int(0)
        res = ctypes.c_double(0.0)
        n = self.kernel.shape[0]
        status = _lib.lib().et_mg_solve(self._handle, _lib.f64(self.kernel.rho, self.mesh.n_el),
                                        _lib.u8(self.kernel.constraints, n), _lib.f64(rhs, n),
                                        _lib.f64(x, n), float(self.tol), int(self.maxiter),
                                        ctypes.byref(it), ctypes.byref(res))
        if status != _lib.ET_OK and (self.strict or status != _lib.ET_ENOCONV):
            _lib.check("et_mg_solve", status)
        self.last_iterations = it.value
        a, b = ctypes.c_double(0), ctypes.c_double(0)
        _lib.lib().et_mg_last_timings(self._handle, ctypes.byref(a), ctypes.byref(b))
        self.last_timings = (a.value, b.value)
        self.last_x0 = x
        return x, res.value


CuMultiGrid = MultiGrid


class _Krylov(Solver):
    """Shared part of CG / GMRES (solvers/_gpu.py:11-107): the reference hands the stiffness
    operator to cupyx.scipy.sparse.linalg.cg / gmres; here the whole iteration runs in the
    library (et_cg_f64 / et_gmres_f64) on the matrix-free structured operator."""

    _restart = 0

    def __init__(self, kernel, maxiter=1000, tol=1e-5, matrix_free=None):
        super().__init__()
        general = hasattr(kernel, "_um")
        if general:
            # general meshes: assembled CSR by default, as the reference (solvers/_gpu.py:18-27)
            matrix_free = False if matrix_free is None else bool(matrix_free)
        else:
            if matrix_free is False:
                logger.warning("matrix_free=False: structured kernels are always applied matrix-free "
                               "(no assembled matrix is built); the result is the same operator.")
            matrix_free = True
        self.kernel = kernel
        self.last_x0 = None
        self.tol = tol
        self.maxiter = maxiter
        self.matrix_free = matrix_free
        self.last_iterations = 0
        self.last_timings = (0.0, 0.0)
        self._work = None

    def _call(self, name, rhs, x):
        k = self.kernel
        n = k.shape[0]
        size = int(_lib.lib().et_krylov_work_size(n, self._restart))
        if self._work is None or self._work.numel() * 8 < size:
            self._work = torch.empty(size // 8 + 1, dtype=torch.float64, device=rhs.device)
        it, res = ctypes.c_int(0), ctypes.c_double(0.0)
        extra = (int(self._restart),) if self._restart else ()
        t0 = time.perf_counter()
        if hasattr(k, "_um") and not self.matrix_free:
            # general meshes, assembled: the operator of this density as CSR, then the iterations
            A = k.construct(k.rho)
            rp, cols = k._pattern()
            status = getattr(_lib.lib(), self._cname)(
                n, _lib.i32(rp), _lib.i32(cols), _lib.f64(k._csr_vals), _lib.f64(rhs, n), _lib.f64(x, n),
                float(self.tol), *extra, int(self.maxiter), _lib.f64(self._work), self._work.numel() * 8,
                ctypes.byref(it), ctypes.byref(res), _lib.stream())
            name = self._cname
            del A
        elif hasattr(k, "_um"):
            # general meshes: the same iterations on the unstructured operator
            status = getattr(_lib.lib(), self._uname)(
                ctypes.byref(k._um), _lib.f64(k.rho, k.n_el), k._cons_ptr(), _lib.f64(rhs, n), _lib.f64(x, n),
                float(self.tol), *extra, int(self.maxiter), _lib.f64(self._work), self._work.numel() * 8,
                ctypes.byref(it), ctypes.byref(res), _lib.stream())
            name = self._uname
        else:
            status = getattr(_lib.lib(), name)(
                len(k._nel), _lib.nel_arg(k._nel), _lib.f64(k.K_single), _lib.f64(k.rho, k.elements.shape[0]),
                k._cons_ptr(), _lib.f64(rhs, n), _lib.f64(x, n), float(self.tol), *extra, int(self.maxiter),
                _lib.f64(self._work), self._work.numel() * 8, ctypes.byref(it), ctypes.byref(res), _lib.stream())
        self.last_timings = (0.0, (time.perf_counter() - t0) * 1e3)
        # maxiter reached above the tolerance is not an error (cupyx reports it in `info` only)
        if status != _lib.ET_OK and status != _lib.ET_ENOCONV:
            _lib.check(name, status)
        self.last_iterations = it.value
        return res.value

    def solve(self, rhs, rho=None, use_last=True):
        if not (self.kernel.has_rho or rho is not None):
            raise ValueError("Solver requires a density vector to be passed or set on the kernel.")
        if rho is not None:
            self.kernel.set_rho(rho)
        if not (isinstance(rhs, torch.Tensor) and rhs.is_cuda and rhs.dtype == torch.float64):
            raise ValueError("rhs must be a float64 device tensor")
        rhs = rhs.contiguous()
        if use_last and self.last_x0 is not None:
            x = self.last_x0.clone()
        else:
            x = torch.zeros_like(rhs)
        residual = self._call(self._name, rhs, x)
        if use_last:
            self.last_x0 = x
        return x, residual

    def reset(self):
        self.last_x0 = None


class CG(_Krylov):
    """Conjugate gradients without preconditioning (solvers/_gpu.py:11-56)."""
    _name = "et_cg_f64"
    _uname = "et_ucg_f64"
    _cname = "et_csr_cg_f64"


class GMRES(_Krylov):
    """Restarted GMRES(20) without preconditioning (solvers/_gpu.py:58-103), cupyx's
    default restart length."""
    _name = "et_gmres_f64"
    _uname = "et_ugmres_f64"
    _cname = "et_csr_gmres_f64"
    _restart = 20


CuCG = CG
CuGMRES = GMRES
