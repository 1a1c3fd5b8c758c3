"""Unpreconditioned Krylov solvers (oracle restatement; test infrastructure only).

CuCG / CuGMRES (solvers/_gpu.py:11-107) hand the matrix-free stiffness operator to
cupyx.scipy.sparse.linalg.cg and gmres (third-party, cupy-cuda12x, unpinned in
requirements.txt:1).  Their published algorithms, restated:

  cg       atol = tol |b|; r = b - A x; rho = r.r; p = r (first) or r + (rho/rho_old) p;
           q = A p; alpha = rho / p.q; x += alpha p; r -= alpha q; iterations += 1;
           stop when |r| <= atol or iterations == maxiter.
  gmres    restart m = min(20, n); per cycle: r = b - A x, stop if |r| <= atol or the
           inner-step count reached maxiter; V[:, 0] = r / |r|; m Arnoldi steps with one
           classical Gram-Schmidt pass (h = V^T u, u -= V h, H[j+1, j] = |u|,
           V[:, j+1] = u / H[j+1, j]); y = argmin |(|r| e1) - H y|; x += V y; count += m.
The reported residual is |b - A x| / |b| of the returned x (solvers/_gpu.py:43, :92).
Pinned in tests: CG iterates equal SciPy's CG iterates (the same recurrence); the first
GMRES cycle equals the dense minimum-residual solution over the Krylov space
span{r, A r, ..., A^(m-1) r}; both converge to the direct solution.
"""
import numpy as np


def cg(matvec, b, x0=None, tol=1e-5, maxiter=1000, callback=None):
    x = np.zeros_like(b) if x0 is None else np.array(x0, dtype=np.float64, copy=True)
    atol = tol * np.linalg.norm(b)
    r = b - matvec(x)
    rho = r @ r
    rho_old = None
    p = None
    it = 0
    resid = np.sqrt(rho)
    if np.linalg.norm(b) == 0 or resid <= atol:
        return x, it
    while it < maxiter:
        p = r.copy() if p is None else r + (rho / rho_old) * p
        q = matvec(p)
        alpha = rho / (p @ q)
        x = x + alpha * p
        r = r - alpha * q
        it += 1
        if callback is not None:
            callback(x)
        rho_old, rho = rho, r @ r
        resid = np.sqrt(rho)
        if resid <= atol:
            break
    return x, it


def gmres(matvec, b, x0=None, tol=1e-5, restart=20, maxiter=1000):
    n = b.shape[0]
    m = min(restart, n)
    x = np.zeros_like(b) if x0 is None else np.array(x0, dtype=np.float64, copy=True)
    atol = tol * np.linalg.norm(b)
    it = 0
    while True:
        r = b - matvec(x)
        rn = np.linalg.norm(r)
        if rn <= atol or it >= maxiter:
            break
        V = np.zeros((n, m))
        H = np.zeros((m + 1, m))
        V[:, 0] = r / rn
        for j in range(m):
            u = matvec(V[:, j])
            h = V[:, :j + 1].T @ u
            u = u - V[:, :j + 1] @ h
            H[:j + 1, j] = h
            H[j + 1, j] = np.linalg.norm(u)
            if j + 1 < m:
                V[:, j + 1] = u / H[j + 1, j]
        e = np.zeros(m + 1)
        e[0] = rn
        y = np.linalg.lstsq(H, e, rcond=None)[0]
        x = x + V @ y
        it += m
    return x, it
