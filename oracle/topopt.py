"""SIMP material model, OC update and the optimisation loop (oracle restatement).

  SingleMaterial      MaterialModels.py:31-99
    forward           E = clip(x~^p, void, 1)                       :47-60
    grad              p * clip(x~, void, 1)^(p-1)                     :62-73
    constraint        sum(x * A) / vol <= vf  (unfiltered x)          :75-81
    ocP               x * nan_to_num(sqrt(max(-dc/A, 0))), 1e-3 if all zero   :83-90
  update_design       TopOpt/_cuda.py:242-280 (CPU TopOpt/_solver.py:271-311)
  solve_and_process   TopOpt/_cuda.py:282-302 (CPU :313-334)
  optimize            TopOpt/_cuda.py:304-382 (CPU :336-417); rho_init is taken as
                      clip(x0, void, 1), the evident intent of :321-322 (which, as
                      written, reads rho before assignment).
Pinned in tests: OC output bounds and move limits, the bisection's volume
bracket, and compliance / sensitivity against central finite differences.
"""
import numpy as np

from . import filters, mesh, ops, physics
from .solvers import MultiGrid


class SingleMaterial:
    def __init__(self, E=1.0, nu=0.33, void=1e-6, penalty=3.0, volume_fraction=0.25):
        self.E, self.nu, self.void = E, nu, void
        self.penalty, self.volume_fraction = penalty, volume_fraction

    def forward(self, x, plain=False):
        if plain:
            return np.clip(x, self.void, 1.0)
        return np.clip(x ** self.penalty, self.void, 1.0)

    def grad(self, x):
        return self.penalty * np.clip(x, self.void, 1.0) ** (self.penalty - 1)

    def constraint(self, x, A, vol):
        return (x.squeeze() * A).sum() / vol <= self.volume_fraction

    def ocP(self, dc, A, x):
        with np.errstate(invalid="ignore", divide="ignore"):
            o = x * np.nan_to_num(np.sqrt(np.maximum(-dc / A, 0)), nan=0)
        if np.abs(o).sum() == 0:
            o = np.ones_like(o) * 1e-3
        return o


def oc_update(x, dc, A, vol, vf, move, material):
    """Bisection OC step; returns (x_new, n_bisection_steps)."""
    xU = np.clip(x + move, 0.0, 1.0)
    xL = np.clip(x - move, 0.0, 1.0)
    ocP = np.maximum(1e-12, material.ocP(dc, A, x))
    l1, l2 = 1e-9, 1e9
    lmid = 0.5 * (l1 + l2)
    xn = np.maximum(0, np.maximum(xL, np.minimum(1.0, np.minimum(xU, ocP / lmid))))
    steps = 0
    while (l2 - l1) / (l2 + l1) > 1e-4:
        lmid = 0.5 * (l1 + l2)
        xn = np.maximum(0, np.maximum(xL, np.minimum(1.0, np.minimum(xU, ocP / lmid))))
        if (xn * A).sum() / vol <= vf:
            l2 = lmid
        else:
            l1 = lmid
        steps += 1
    return xn, steps


class Problem:
    """A structured SIMP problem: mesh, K0, BCs, loads and filter (all host arrays)."""

    def __init__(self, nel, E=1.0, nu=0.3, r_min=1.5, size=None):
        self.nel = tuple(int(n) for n in nel)
        self.dim = len(self.nel)
        size = tuple(float(n) for n in (size if size is not None else self.nel))
        if self.dim == 2:
            self.elements = mesh.elements_2d(*self.nel)
            self.nodes = mesh.nodes_2d(*self.nel, *size)
        else:
            self.elements = mesh.elements_3d(*self.nel)
            self.nodes = mesh.nodes_3d(*self.nel, *size)
        K, _, _, A = physics.auto_stiffness(self.nodes[self.elements[0]], E, nu)
        self.K0 = K
        self.A = A
        self.dof = K.shape[0] // self.elements.shape[1]
        self.n_dof = self.nodes.shape[0] * self.dof
        self.n_el = self.elements.shape[0]
        self.volume = A * self.n_el
        self.cons = np.zeros(self.n_dof, dtype=bool)
        self.f = np.zeros(self.n_dof)
        spacing = [s / n for s, n in zip(size, self.nel)]
        self.offsets, self.weights = filters.offsets_weights(r_min, spacing)
        self.norm = filters.normalizer(self.nel, self.offsets, self.weights)

    def filt(self, x):
        return filters.forward(x, self.nel, self.offsets, self.weights)

    def filt_T(self, v):
        return filters.transpose(v, self.nel, self.offsets, self.weights, self.norm)


def solve_and_process(prob, x, material, solver, use_last=True):
    xt = prob.filt(x)
    E = material.forward(xt)
    A0 = ops.assemble(prob.K0, prob.elements, E, prob.cons, prob.n_dof)
    U, res = solver.solve(A0, prob.f, A0_matvec=lambda v: ops.matvec(prob.K0, prob.elements, E, v, prob.cons),
                          use_last=use_last)
    c = prob.f @ U
    dk = ops.process_dk(prob.K0, prob.elements, U)
    dr = prob.filt_T(material.grad(xt) * dk)
    return c, dr, res, U


def optimize(prob, material, solver, x0, max_iter, move=0.2, ch_tol=1e-4, fun_tol=1e-6,
             callback=None, schedule=None):
    """Returns (x, flag, history dict with compliance, change, residual, pcg iterations).
    schedule(p, i): penalty of iteration i (the initial solve uses i = 0) with the stopping
    test waiting for schedule(p, i) == p (MaterialModels.py:47-60, :95-98)."""
    p_final = material.penalty

    def set_penalty(i):
        if schedule is not None:
            material.penalty = schedule(p_final, i)

    x = material.forward(np.asarray(x0, dtype=np.float64), plain=True)
    hist = {"comp": [], "change": [], "res": [], "pcg": []}
    set_penalty(0)
    c, dr, res, _ = solve_and_process(prob, x, material, solver)
    hist["comp"].append(c)
    hist["res"].append(res)
    hist["pcg"].append(solver.last_iters)
    flag = False
    for i in range(max_iter):
        x_old, c_old = x.copy(), c
        x, _ = oc_update(x, dr, prob.A, prob.volume, material.volume_fraction, move, material)
        set_penalty(i)
        c, dr, res, _ = solve_and_process(prob, x, material, solver)
        change = np.sqrt(np.square(x - x_old).mean())
        change_f = abs(c - c_old) / c
        hist["comp"].append(c)
        hist["change"].append(change)
        hist["res"].append(res)
        hist["pcg"].append(solver.last_iters)
        if callback is not None:
            callback(i, x, c)
        if change < ch_tol and change_f < fun_tol and material.penalty == p_final:
            flag = True
            break
    material.penalty = p_final
    return x, flag, hist


class PenalizedMultiMaterial:
    """Discrete material optimisation with m candidate materials (MaterialModels.py:100-235;
    rp=True: PenalizedMultiMaterialRP, :237-369).  Densities rho (n_el, m), per-material moduli
    E_mat (m,):
      forward   E = clip(sum_i E_i r_i^p prod_{j != i} (1 - r_j^p), void, inf)
                rp: prod_{j != i} (1 - r_j)^p                          (:145-176)
      plain     clip(rho @ E_mat, void, 1)                              (:146-149)
      grad      dE/dr_k = E_k p r_k^(p-1) prod_{j != k} w_j
                          - sum_{i != k} E_i r_i^p dw_k prod_{j != i, k} w_j,
                w_j = 1 - r_j^p (rp: (1 - r_j)^p), dw_k = p r_k^(p-1) (rp: p (1 - r_k)^(p-1))
                (the unclipped expression, as :178-206 evaluates it; r = clip(rho, void, 1))
      constraint per material sum_e r_ek A / vol <= mass_k, or with rho_mat the total
                sum_e sum_k r_ek A rho_mat_k / vol <= mass for every material  (:208-221)
      ocP       rho * nan_to_num(sqrt(-df / A)), 1e-3 everywhere if all zero (:223-229)."""

    def __init__(self, E_mat, mass, penalty=3.0, void=1e-6, rho_mat=None, rp=False):
        self.E_mat = np.asarray(E_mat, dtype=np.float64).reshape(-1)
        self.m = self.E_mat.size
        self.mass = np.asarray(mass, dtype=np.float64).reshape(-1)
        self.rho_mat = None if rho_mat is None else np.asarray(rho_mat, dtype=np.float64).reshape(-1)
        self.penalty, self.void, self.rp = penalty, void, rp

    def _w(self, r):
        p = self.penalty
        return (1.0 - r) ** p if self.rp else 1.0 - r ** p

    def forward(self, rho, plain=False):
        if plain:
            return np.clip(rho @ self.E_mat, self.void, 1.0)
        p = self.penalty
        r = np.clip(rho, self.void, 1.0)
        w = self._w(r)
        E = np.zeros(r.shape[0])
        for i in range(self.m):
            t = r[:, i] ** p
            for j in range(self.m):
                if j != i:
                    t = t * w[:, j]
            E += self.E_mat[i] * t
        return np.clip(E, self.void, np.inf)

    def grad(self, rho):
        p = self.penalty
        r = np.clip(rho, self.void, 1.0)
        w = self._w(r)
        dw = p * (1.0 - r) ** (p - 1) if self.rp else p * r ** (p - 1)
        g = np.zeros_like(r)
        for k in range(self.m):
            t = self.E_mat[k] * p * r[:, k] ** (p - 1)
            for j in range(self.m):
                if j != k:
                    t = t * w[:, j]
            for i in range(self.m):
                if i == k:
                    continue
                s = self.E_mat[i] * r[:, i] ** p * dw[:, k]
                for j in range(self.m):
                    if j != i and j != k:
                        s = s * w[:, j]
                t = t - s
            g[:, k] = t
        return g

    def constraint(self, x, A, vol):
        if self.rho_mat is None:
            return (x * A).sum(axis=0) / vol - self.mass <= 0
        return np.full(self.m, ((x * A) @ self.rho_mat).sum() / vol <= self.mass[0])

    def ocP(self, dc, A, x):
        with np.errstate(invalid="ignore", divide="ignore"):
            o = x * np.nan_to_num(np.sqrt(-dc / A), nan=0)
        if np.abs(o).sum() == 0:
            o = np.ones_like(o) * 1e-3
        return o


def oc_update_multi(x, dc, A, vol, move, material):
    """OC with one multiplier per material, bisected in lockstep until every bracket has closed
    (TopOpt/_cuda.py:242-280 with n_sets = m: lmid has one entry per material and the loop runs
    while any (l2 - l1) / (l2 + l1) > 1e-4).  x, dc: (n_el, m)."""
    m = x.shape[1]
    xU = np.clip(x + move, 0.0, 1.0)
    xL = np.clip(x - move, 0.0, 1.0)
    ocP = np.maximum(1e-12, material.ocP(dc, A, x))
    l1, l2 = np.full(m, 1e-9), np.full(m, 1e9)
    lmid = 0.5 * (l1 + l2)
    xn = np.maximum(0, np.maximum(xL, np.minimum(1.0, np.minimum(xU, ocP / lmid))))
    steps = 0
    while np.any((l2 - l1) / (l2 + l1) > 1e-4):
        lmid = 0.5 * (l1 + l2)
        xn = np.maximum(0, np.maximum(xL, np.minimum(1.0, np.minimum(xU, ocP / lmid))))
        valid = material.constraint(xn, A, vol)
        l2 = np.where(valid, lmid, l2)
        l1 = np.where(valid, l1, lmid)
        steps += 1
    return xn, steps
