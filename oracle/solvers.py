"""GMG-preconditioned conjugate gradients (oracle restatement).

GPU semantics: solvers/_multigrid_cuda.py:15-242 (CuMultiGrid)
CPU semantics: solvers/_multigrid.py:11-230 (MultiGrid): omega is not reset
at the start of a solve (:200 vs _cuda :210) and the 2-D restriction is R/2
(core/_mgm.py:158-160).

  omega schedule   min = max(omega/2, 0.4), d = omega - min        (_cuda :27-30)
                   omega <- min at solve start (GPU only)           (_cuda :210)
                   omega <- min + d/2 exp((-clip(R, tol, .1) + tol) 500) after
                   each PCG iteration                               (_cuda :227)
                   level l smooths with omega * 1.06^l               (_cuda :139-141)
  V-cycle          damped Jacobi pre/post smoothing, restriction, recursion,
                   prolongation; the 'w' branch only for lowercase cycle == 'w'
                   and level in w_level                              (_cuda :133-186)
  coarse solve     direct (SciPy splu, as the reference)            (_cuda :120-131)
  PCG              Fletcher-Reeves beta, recursive residual, R = |r|/|b| < tol
                                                                     (_cuda :188-242)
Pinned in tests: the converged solution satisfies |b - A x| / |b| < tol and
matches a direct SciPy solve to the tolerance-implied accuracy.
"""
import numpy as np
import scipy.sparse.linalg as spla

from . import mgm


class MultiGrid:
    def __init__(self, nel, dof, maxiter=1000, tol=1e-5, n_smooth=3, omega=0.5, n_level=3,
                 cycle="W", w_level=1, min_omega=0.4, mode="gpu"):
        self.nel = np.asarray(nel, dtype=np.int64)
        self.dof = dof
        self.maxiter = maxiter
        self.tol = tol
        self.n_smooth = n_smooth
        self.omega = omega
        self.max_omega = omega
        self.min_omega = max(omega / 2, min_omega)
        self.d_omega = self.max_omega - self.min_omega
        self.n_level = n_level
        self.cycle = cycle
        self.w_level = [w_level] if isinstance(w_level, int) else list(w_level)
        self.mode = mode
        self.last_x0 = None
        self.last_iters = 0

    def setup(self, A0, A0_matvec=None):
        """Level operators: A0 (CSR of A_hat) and Galerkin products below it."""
        self.levels = []
        A = A0
        self.A0_mv = A0_matvec if A0_matvec is not None else (lambda v: A0 @ v)
        self.levels.append((self.A0_mv, 1.0 / A0.diagonal()))
        mats = [A0]
        for l in range(self.n_level):
            A = mgm.galerkin(A, self.nel // (2 ** l), self.dof)
            mats.append(A)
            self.levels.append(((lambda M: (lambda v: M @ v))(A), 1.0 / A.diagonal()))
        self.mats = mats
        self.coarse = spla.splu(mats[-1].tocsc())
        if not hasattr(self, "_transfer"):
            self._transfer = [(mgm.restriction_matrix(self.nel // (2 ** l), self.dof).tocsr(),
                               mgm.prolongation_matrix(self.nel // (2 ** l), self.dof).tocsr())
                              for l in range(self.n_level)]

    def _jacobi(self, x, b, level, n):
        A, Dinv = self.levels[level]
        for _ in range(n):
            x = x + self.omega * Dinv * (b - A(x))
        return x

    def _vcycle(self, x, b, level):
        if level == self.n_level:
            return self.coarse.solve(b)
        A, _ = self.levels[level]
        quirk = self.mode == "cpu" and len(self.nel) == 2

        def correction(x):
            r = b - A(x)
            R, P = self._transfer[level]
            rc = R @ r
            if quirk:
                rc = 0.5 * rc
            e = self._vcycle(rc * 0.0, rc, level + 1)
            e = P @ e
            self.omega = self.omega * 1.06 ** level
            e = self._jacobi(e, r, level, self.n_smooth)
            self.omega = self.omega / 1.06 ** level
            return x + e

        self.omega = self.omega * 1.06 ** level
        x = self._jacobi(x, b, level, self.n_smooth)
        self.omega = self.omega / 1.06 ** level
        x = correction(x)
        if self.cycle == "w" and level in self.w_level:
            x = correction(x)
        return x

    def solve(self, A0, rhs, A0_matvec=None, use_last=True):
        self.setup(A0, A0_matvec)
        A = self.A0_mv
        x = self.last_x0.copy() if (use_last and self.last_x0 is not None) else np.zeros_like(rhs)
        if self.mode == "gpu":
            self.omega = self.min_omega
        r = rhs - A(x)
        z = self._vcycle(np.zeros_like(r), r, 0)
        p = z.copy()
        rho_old = r @ z
        norm_b = np.linalg.norm(rhs)
        norm_r = np.linalg.norm(r)
        it = 0
        for it in range(1, self.maxiter + 1):
            q = A(p)
            alpha = rho_old / (p @ q)
            x = x + alpha * p
            r = r - alpha * q
            norm_r = np.linalg.norm(r)
            R = norm_r / norm_b
            if R < self.tol:
                break
            self.omega = self.min_omega + self.d_omega / 2 * np.exp((-np.clip(R, self.tol, 1e-1) + self.tol) * 500)
            z = self._vcycle(np.zeros_like(r), r, 0)
            rho_new = r @ z
            beta = rho_new / rho_old
            p = z + beta * p
            rho_old = rho_new
        self.last_x0 = x
        self.last_iters = it
        return x, norm_r / norm_b
