"""Pins of the CPU oracle: closed forms, invariants and brute force.

Each test is built so that a plausible slip in the oracle (a dropped term, a
wrong sign or index, a transposed operand) fails it; none retypes the
oracle's own formula.
"""
import itertools

import numpy as np
import pytest
import scipy.sparse.linalg as spla

from oracle import filters, mesh, mgm, ops, physics, topopt
from oracle.solvers import MultiGrid

# ----------------------------------------------------------------- physics


def _top88_KE(nu):
    """Published Q4 element matrix of Andreassen et al. (2011), unit square, E = 1."""
    A11 = np.array([[12, 3, -6, -3], [3, 12, 3, 0], [-6, 3, 12, -3], [-3, 0, -3, 12]])
    A12 = np.array([[-6, -3, 0, 3], [-3, -6, -3, -6], [0, -3, -6, 3], [3, -6, 3, -6]])
    B11 = np.array([[-4, 3, -2, 9], [3, -4, -9, 4], [-2, -9, -4, -3], [9, 4, -3, -4]])
    B12 = np.array([[2, -3, 4, -9], [-3, 2, 9, -2], [4, 9, 2, 3], [-9, -2, 3, 2]])
    return 1 / (1 - nu ** 2) / 24 * (np.block([[A11, A12], [A12.T, A11]])
                                     + nu * np.block([[B11, B12], [B12.T, B11]]))


UNIT_Q4 = np.array([[0, 0], [1, 0], [1, 1], [0, 1]], dtype=float)
UNIT_H8 = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                    [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], dtype=float)


@pytest.mark.parametrize("nu", [0.3, 0.33, 0.0, 0.45])
def test_q4_matches_top88(nu):
    K, _, _ = physics.q4_stiffness(UNIT_Q4, 1.0, nu)
    assert np.allclose(K, _top88_KE(nu), atol=1e-15, rtol=0)


def _affine_energy_check(K, C, corners, dim, V, rng):
    """u = G x (affine field) is interpolated exactly: u^T K u = V * eps : C : eps (Voigt)."""
    for _ in range(5):
        G = rng.standard_normal((dim, dim))
        u = (corners @ G.T).reshape(-1)
        eps = 0.5 * (G + G.T)
        if dim == 2:
            v = np.array([eps[0, 0], eps[1, 1], 2 * eps[0, 1]])
        else:
            v = np.array([eps[0, 0], eps[1, 1], eps[2, 2], 2 * eps[0, 1], 2 * eps[1, 2], 2 * eps[0, 2]])
        assert np.isclose(u @ K @ u, V * v @ C @ v, rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize("scale", [(1.0, 1.0, 1.0), (2.0, 0.5, 1.5)])
def test_h8_invariants(scale):
    rng = np.random.default_rng(0)
    X = UNIT_H8 * np.array(scale)
    K, C, _ = physics.h8_stiffness(X, 1.0, 0.3)
    assert np.allclose(K, K.T, atol=1e-14)
    ev = np.linalg.eigvalsh(K)
    assert np.sum(np.abs(ev) < 1e-10 * ev.max()) == 6 and ev.min() > -1e-12
    _affine_energy_check(K, C, X, 3, float(np.prod(scale)), rng)
    K3, _, _ = physics.h8_stiffness(X, 1.0, 0.3, n_gauss=3)
    assert np.allclose(K, K3, atol=1e-13)
    assert np.isclose(physics.hex_volume(X), np.prod(scale))


def test_q4_invariants():
    rng = np.random.default_rng(1)
    X = UNIT_Q4 * np.array([2.0, 0.7])
    K, D, _ = physics.q4_stiffness(X, 2.5, 0.25)
    ev = np.linalg.eigvalsh(K)
    assert np.sum(np.abs(ev) < 1e-10 * ev.max()) == 3
    _affine_energy_check(K, D, X, 2, 1.4, rng)
    assert np.allclose(K, physics.q4_stiffness(X, 2.5, 0.25, n_gauss=3)[0], atol=1e-13)


def test_product_physics_matches_oracle():
    from edgeto_amd import physics as pphys
    for X in (UNIT_Q4, UNIT_Q4 * [1.5, 0.5]):
        a, b = pphys.auto_stiffness(X, 1.0, 0.3), physics.auto_stiffness(X, 1.0, 0.3)
        for u, v in zip(a, b):
            assert np.allclose(u, v, atol=1e-14)
    for X in (UNIT_H8, UNIT_H8 * [1.0, 2.0, 0.5]):
        a, b = pphys.auto_stiffness(X, 1.0, 0.33), physics.auto_stiffness(X, 1.0, 0.33)
        for u, v in zip(a, b):
            assert np.allclose(u, v, atol=1e-14)

# -------------------------------------------------------------------- mesh


@pytest.mark.parametrize("nel", [(3, 2), (5, 4), (3, 2, 4), (4, 3, 2)])
def test_connectivity_literal_loops(nel):
    """Nested-loop enumeration in the reference's traversal order (i slowest)."""
    dim = len(nel)
    if dim == 2:
        nx = nel[0] + 1
        ref = [[j * nx + i, j * nx + i + 1, (j + 1) * nx + i + 1, (j + 1) * nx + i]
               for i in range(nel[0]) for j in range(nel[1])]
        got = mesh.elements_2d(*nel)
        nodes = mesh.nodes_2d(*nel, *nel)
    else:
        nx, nz = nel[0] + 1, nel[2] + 1
        ref = []
        for i in range(nel[0]):
            for j in range(nel[1]):
                for k in range(nel[2]):
                    b = nz * i + nz * nx * j + k
                    ref.append([b, nz * (i + 1) + nz * nx * j + k, nz * (i + 1) + nz * nx * (j + 1) + k,
                                nz * i + nz * nx * (j + 1) + k, b + 1, nz * (i + 1) + nz * nx * j + k + 1,
                                nz * (i + 1) + nz * nx * (j + 1) + k + 1, nz * i + nz * nx * (j + 1) + k + 1])
        got = mesh.elements_3d(*nel)
        nodes = mesh.nodes_3d(*nel, *nel)
    assert got.dtype == np.int32 and np.array_equal(got, np.array(ref, dtype=np.int32))
    # element e's first node sits at the element's integer origin, nodes are CCW-consistent
    origins = np.array(list(itertools.product(*[range(n) for n in nel])), dtype=float)
    assert np.array_equal(nodes[got[:, 0]], origins)
    corner = UNIT_Q4 if dim == 2 else UNIT_H8
    assert np.array_equal(nodes[got] - nodes[got[:, :1]], np.broadcast_to(corner, (got.shape[0],) + corner.shape))


def test_node_star():
    el = mesh.elements_3d(3, 2, 2)
    n_nodes = 4 * 3 * 3
    flat, el_ids, sorter, node_ids = mesh.node_star(el, n_nodes)
    assert np.all(np.diff(flat[sorter]) >= 0)
    counts = np.diff(np.append(node_ids, flat.size))
    assert np.array_equal(counts, np.bincount(flat, minlength=n_nodes))

# ---------------------------------------------------------------- operator


def _rand_problem(nel, seed):
    rng = np.random.default_rng(seed)
    dim = len(nel)
    el = mesh.elements_3d(*nel) if dim == 3 else mesh.elements_2d(*nel)
    X = UNIT_H8 if dim == 3 else UNIT_Q4
    K0 = physics.auto_stiffness(X, 1.0, 0.3)[0]
    n_dof = int(np.prod(np.array(nel) + 1)) * dim
    return el, K0, rng.uniform(0.01, 1, el.shape[0]), rng.standard_normal(n_dof), rng.random(n_dof) < 0.2, n_dof


@pytest.mark.parametrize("nel", [(4, 3), (3, 2, 2)])
def test_matvec_vs_literal_loop_and_assembly(nel):
    el, K0, E, u, cons, n_dof = _rand_problem(nel, 3)
    star = mesh.node_star(el, n_dof // len(nel))
    y_loop = ops.matvec_node_loop(K0, el, E, u, cons, star)
    y_vec = ops.matvec(K0, el, E, u, cons)
    A = ops.assemble(K0, el, E, cons, n_dof)
    assert np.allclose(y_loop, y_vec, rtol=0, atol=1e-13)
    assert np.allclose(A @ u, y_vec, rtol=0, atol=1e-13)
    # constrained rows copy the input, constrained columns are ignored
    assert np.array_equal(y_vec[cons], u[cons])
    u2 = u.copy()
    u2[cons] += 7.0
    assert np.allclose(ops.matvec(K0, el, E, u2, cons)[~cons], y_vec[~cons], atol=1e-13)


@pytest.mark.parametrize("nel", [(6, 4), (4, 3, 3)])
def test_diagonal_and_energy(nel):
    el, K0, E, u, cons, n_dof = _rand_problem(nel, 4)
    A = ops.assemble(K0, el, E, cons, n_dof)
    assert np.allclose(ops.diagonal(K0, el, E, cons, n_dof), A.diagonal(), atol=1e-14)
    dk = ops.process_dk(K0, el, u)
    Kfull = ops.assemble(K0, el, E, None, n_dof)
    assert np.isclose(-(E * dk).sum(), u @ (Kfull @ u), rtol=1e-12)
    # single element: energy of its own DOFs
    e = 5
    edof = mesh.dof_map(el[e:e + 1], len(nel))[0]
    assert np.isclose(dk[e], -u[edof] @ K0 @ u[edof], rtol=1e-13)

@pytest.mark.parametrize("dim", [2, 3])
def test_fea_integrals_affine_field(dim):
    """A displacement u(x) = G x (G arbitrary, brick elements of non-unit size) has the uniform
    small strain eps = sym(G) in every element: strain = eps (engineering shears), stress =
    C eps, element energy = 1/2 E_e V eps.C.eps, and von Mises of that stress; checked against
    closed forms and the textbook von Mises cases (uniaxial: |s|, pure shear: sqrt(3)|t|)."""
    rng = np.random.default_rng(21)
    nel = (3, 2) if dim == 2 else (3, 2, 2)
    size = (1.5, 0.7) if dim == 2 else (1.5, 0.7, 0.4)
    if dim == 2:
        el, nodes = mesh.elements_2d(*nel), mesh.nodes_2d(*nel, *[n * s for n, s in zip(nel, size)])
        K0, C, B = physics.q4_stiffness(nodes[el[0]], 1.0, 0.3)
        V = size[0] * size[1]
    else:
        el, nodes = mesh.elements_3d(*nel), mesh.nodes_3d(*nel, *[n * s for n, s in zip(nel, size)])
        K0, C, B = physics.h8_stiffness(nodes[el[0]], 1.0, 0.3)
        V = size[0] * size[1] * size[2]
    G = rng.standard_normal((dim, dim))
    U = (nodes @ G.T).ravel()
    E = rng.uniform(0.2, 1.0, el.shape[0])
    eps = 0.5 * (G + G.T)
    if dim == 2:
        ev = np.array([eps[0, 0], eps[1, 1], 2 * eps[0, 1]])
    else:
        ev = np.array([eps[0, 0], eps[1, 1], eps[2, 2], 2 * eps[0, 1], 2 * eps[1, 2], 2 * eps[2, 0]])
    strain, stress, energy, vm = ops.fea_integrals(K0, C, B, el, E, U)
    assert np.allclose(strain, ev[None], atol=1e-13)
    assert np.allclose(stress, (C @ ev)[None], atol=1e-13)
    assert np.allclose(energy, 0.5 * E * V * (ev @ C @ ev), rtol=1e-12)
    # von Mises closed forms: affine fields whose stress is uniaxial or pure shear
    if dim == 2:
        uni, shear = np.array([2.5, 0.0, 0.0]), np.array([0.0, 0.0, 1.7])
    else:
        uni, shear = np.array([2.5, 0, 0, 0, 0, 0.0]), np.array([0, 0, 0, 0, 1.7, 0.0])
    Dinv = np.linalg.inv(C)
    for target, vm_expect in ((uni, 2.5), (shear, np.sqrt(3) * 1.7)):
        e_t = Dinv @ target  # the strain with this stress
        if dim == 2:
            Gt = np.array([[e_t[0], e_t[2] / 2], [e_t[2] / 2, e_t[1]]])
        else:
            Gt = np.array([[e_t[0], e_t[3] / 2, e_t[5] / 2], [e_t[3] / 2, e_t[1], e_t[4] / 2],
                           [e_t[5] / 2, e_t[4] / 2, e_t[2]]])
        _, st, _, vmt = ops.fea_integrals(K0, C, B, el, E, (nodes @ Gt.T).ravel())
        assert np.allclose(st, target[None], atol=1e-12)
        assert np.allclose(vmt, vm_expect, rtol=1e-12)


# ------------------------------------------------------------------ filter


@pytest.mark.parametrize("shape,rmin", [((12, 7), 1.5), ((9, 8), 2.6), ((7, 6, 5), 2.0), ((6, 5, 6), 2.9)])
def test_filter_against_kdtree(shape, rmin):
    offs, w = filters.offsets_weights(rmin, [1.0] * len(shape))
    H = filters.explicit_matrix(shape, rmin)
    x = np.random.default_rng(5).uniform(0, 1, int(np.prod(shape)))
    assert np.allclose(filters.forward(x, shape, offs, w), H @ x, atol=1e-14)
    norm = filters.normalizer(shape, offs, w)
    assert np.allclose(filters.transpose(x, shape, offs, w, norm), H.T @ x, atol=1e-14)
    assert np.allclose(np.asarray(H.sum(axis=1)).ravel(), 1.0)
    y = np.random.default_rng(6).standard_normal(x.size)
    assert np.isclose(filters.forward(x, shape, offs, w) @ y,
                      x @ filters.transpose(y, shape, offs, w, norm), rtol=1e-12)


def test_filter_offsets_counts():
    # counts of integer points strictly inside the radius (known lattice counts)
    assert len(filters.offsets_weights(1.5, [1, 1])[0]) == 9
    assert len(filters.offsets_weights(2.0, [1, 1, 1])[0]) == 19 + 8   # |o| < 2: 1+6+12+8
    assert len(filters.offsets_weights(1.2, [1, 1, 1])[0]) == 7
    assert np.isclose(filters.offsets_weights(2.0, [1, 1, 1])[1].sum(), 1.0)

# ------------------------------------------------------------ multigrid


@pytest.mark.parametrize("nel", [(8, 4), (6, 10), (4, 4, 2), (6, 4, 8)])
def test_transfer_closed_forms(nel):
    dof = len(nel)
    R = mgm.restriction_matrix(nel, dof)
    P = mgm.prolongation_matrix(nel, dof)
    assert np.allclose(np.asarray(R.sum(axis=1)).ravel(), 1.0)        # R 1 = 1
    assert abs(P - R.T).max() < 1e-15                                  # P = R^T on even grids
    v = np.random.default_rng(7).standard_normal(R.shape[1])
    assert np.allclose(mgm.restrict(v, nel, dof), R @ v, atol=1e-14)
    # interior weights: the full-weighting stencil [1/4, 1/2, 1/4] per axis
    if len(nel) == 2:
        c = (nel[1] // 4) * (nel[0] // 2 + 1) + nel[0] // 4
        fine = mesh.node_index(nel, (2 * (nel[0] // 4), 2 * (nel[1] // 4)))
        assert np.isclose(R[c * dof, fine * dof], 1.0 / 4.0)


def _literal_restrict_cpu2d(v, nel, dof):
    """Loop restatement of core/_mgm.py:94-166 including the total_weight accumulation."""
    nxf, nyf = nel[0] + 1, nel[1] + 1
    nxc, nyc = nel[0] // 2 + 1, nel[1] // 2 + 1
    out = np.zeros(nxc * nyc * dof)
    for c in range(nxc * nyc):
        ic, jc = c % nxc, c // nxc
        b = (ic in (0, nxc - 1)) + (jc in (0, nyc - 1))
        tw = 2.25 if b == 2 else (3.0 if b == 1 else 4.0)
        for d in range(dof):
            val = 0.0
            for di in (-1, 0, 1):
                for dj in (-1, 0, 1):
                    i, j = 2 * ic + di, 2 * jc + dj
                    if 0 <= i < nxf and 0 <= j < nyf:
                        wgt = 1.0 if abs(di) + abs(dj) == 0 else (0.5 if abs(di) + abs(dj) == 1 else 0.25)
                        val += wgt * v[(i + nxf * j) * dof + d]
                        if d == 0:
                            tw += wgt
            out[c * dof + d] = val / tw
    return out


def test_cpu2d_restriction_quirk_is_half():
    nel = (6, 4)
    v = np.random.default_rng(8).standard_normal(7 * 5 * 2)
    lit = _literal_restrict_cpu2d(v, nel, 2)
    assert np.allclose(lit, 0.5 * (mgm.restriction_matrix(nel, 2) @ v), atol=1e-14)
    assert np.allclose(mgm.restrict(v, nel, 2, cpu2d_quirk=True), lit, atol=1e-14)


def _literal_galerkin(A, nel, dof):
    """Per-(c, t) restatement of the nnz-based Galerkin kernels: sum over fine
    pairs (f, g) around (2c, 2t) of w(f-2c) w(g-2t) A[f, g] / (tw(c) tw(t))."""
    dim = len(nel)
    nf = np.array(nel) + 1
    ncn = np.array(nel) // 2 + 1
    A = A.toarray()

    def coords(n, shape):
        if dim == 2:
            return (n % shape[0], n // shape[0])
        return ((n // shape[2]) % shape[0], n // (shape[2] * shape[0]), n % shape[2])

    def index(ijk, shape):
        if dim == 2:
            return ijk[1] * shape[0] + ijk[0]
        return (ijk[1] * shape[0] + ijk[0]) * shape[2] + ijk[2]

    def tw(ijk):
        return np.prod([1.5 if (c == 0 or c == n - 1) else 2.0 for c, n in zip(ijk, ncn)])

    Nc = int(np.prod(ncn))
    Ac = np.zeros((Nc * dof, Nc * dof))
    for c in range(Nc):
        cc = coords(c, ncn)
        for t in range(Nc):
            tt = coords(t, ncn)
            if max(abs(a - b) for a, b in zip(cc, tt)) > 1:
                continue
            blk = np.zeros((dof, dof))
            for df in itertools.product((-1, 0, 1), repeat=dim):
                f = [2 * a + d for a, d in zip(cc, df)]
                if any(x < 0 or x >= n for x, n in zip(f, nf)):
                    continue
                for dg in itertools.product((-1, 0, 1), repeat=dim):
                    g = [2 * a + d for a, d in zip(tt, dg)]
                    if any(x < 0 or x >= n for x, n in zip(g, nf)):
                        continue
                    w = 0.5 ** (sum(map(abs, df)) + sum(map(abs, dg)))
                    fi, gi = index(f, nf), index(g, nf)
                    blk += w * A[fi * dof:(fi + 1) * dof, gi * dof:(gi + 1) * dof]
            Ac[c * dof:(c + 1) * dof, t * dof:(t + 1) * dof] = blk / (tw(cc) * tw(tt))
    return Ac


@pytest.mark.parametrize("nel", [(4, 4), (6, 4), (4, 2, 2)])
def test_galerkin_literal(nel):
    el, K0, E, u, cons, n_dof = _rand_problem(nel, 9)
    A = ops.assemble(K0, el, E, cons, n_dof)
    Ac = mgm.galerkin(A, nel, len(nel)).toarray()
    assert np.allclose(Ac, _literal_galerkin(A, nel, len(nel)), atol=1e-13)
    assert np.allclose(Ac, Ac.T, atol=1e-13)
    assert np.linalg.eigvalsh(Ac).min() > 0


@pytest.mark.parametrize("mode", ["gpu", "cpu"])
def test_mg_pcg_matches_direct(mode):
    nel = (16, 8)
    el, K0, E, _, _, n_dof = _rand_problem(nel, 10)
    cons = np.zeros(n_dof, dtype=bool)
    cons[: (nel[1] + 1) * 2] = False
    left = np.arange(nel[1] + 1) * (nel[0] + 1)
    cons[left * 2] = cons[left * 2 + 1] = True
    f = np.zeros(n_dof)
    f[(nel[0]) * 2 + 1] = -1.0
    A = ops.assemble(K0, el, E, cons, n_dof)
    mg = MultiGrid(nel, 2, tol=1e-10, n_smooth=2, n_level=2, mode=mode)
    x, res = mg.solve(A, f, use_last=False)
    assert res < 1e-10
    xd = spla.spsolve(A.tocsc(), f)
    assert np.linalg.norm(x - xd) / np.linalg.norm(xd) < 1e-7

# -------------------------------------------------------------------- SIMP


def test_oc_update_invariants():
    rng = np.random.default_rng(11)
    n = 500
    x = rng.uniform(0, 1, n)
    dc = -rng.uniform(0, 1, n)
    mat = topopt.SingleMaterial(void=1e-9, penalty=3.0, volume_fraction=0.35)
    xn, steps = topopt.oc_update(x, dc, 1.0, float(n), 0.35, 0.2, mat)
    assert np.all(xn >= np.clip(x - 0.2, 0, 1) - 1e-15) and np.all(xn <= np.clip(x + 0.2, 0, 1) + 1e-15)
    assert 20 < steps < 120
    # the final midpoint leaves the volume within the bisection bracket of the target
    assert abs(xn.mean() - 0.35) < 5e-3
    # brute force: the monotone volume function crosses vf near the returned lambda
    ocp = np.maximum(1e-12, x * np.sqrt(-dc))
    lams = np.geomspace(1e-9, 1e9, 20001)
    vols = [np.maximum(np.clip(x - .2, 0, 1), np.minimum(np.clip(x + .2, 0, 1), ocp / lam)).mean() for lam in lams]
    assert np.all(np.diff(vols) <= 1e-15)


def test_sensitivity_central_differences():
    """dc/dx of the full chain (filter, SIMP, solve, adjoint) against central FD."""
    nel = (8, 5)
    prob = topopt.Problem(nel, nu=0.3, r_min=1.5)
    left = np.arange(nel[1] + 1) * (nel[0] + 1)
    prob.cons[left * 2] = prob.cons[left * 2 + 1] = True
    prob.f[(nel[0] + 1) * nel[1] * 2 + nel[0] * 2 + 1] = -1.0
    mat = topopt.SingleMaterial(void=1e-9, penalty=3.0, volume_fraction=0.5)
    x = np.random.default_rng(12).uniform(0.2, 1.0, prob.n_el)

    def comp(xx):
        E = mat.forward(prob.filt(xx))
        A = ops.assemble(prob.K0, prob.elements, E, prob.cons, prob.n_dof)
        U = spla.spsolve(A.tocsc(), prob.f)
        return prob.f @ U, U

    c, U = comp(x)
    dr = prob.filt_T(mat.grad(prob.filt(x)) * ops.process_dk(prob.K0, prob.elements, U))
    h = 1e-6
    for e in (0, 7, 19, 33):
        xp, xm = x.copy(), x.copy()
        xp[e] += h
        xm[e] -= h
        fd = (comp(xp)[0] - comp(xm)[0]) / (2 * h)
        assert np.isclose(dr[e], fd, rtol=1e-6, atol=1e-9 * abs(c))


# ------------------------------------------------------------------ Krylov

def _spd_problem(nel, seed):
    from oracle import krylov  # noqa: F401
    el, K0, E, u, cons, n_dof = _rand_problem(nel, seed)
    A = ops.assemble(K0, el, E, cons, n_dof)
    b = np.random.default_rng(seed + 1).standard_normal(n_dof)
    return A, b


def test_cg_iterates_equal_scipy_cg():
    """Unpreconditioned CG is one recurrence: the oracle's iterates are SciPy's (to rounding),
    and the converged x solves the system."""
    from oracle import krylov
    import scipy.sparse.linalg as sla
    A, b = _spd_problem((6, 4, 3), 3)
    mine, ref = [], []
    krylov.cg(lambda v: A @ v, b, tol=1e-10, maxiter=25, callback=lambda x: mine.append(x.copy()))
    sla.cg(A, b, rtol=1e-10, maxiter=25, callback=lambda x: ref.append(x.copy()))
    k = min(len(mine), len(ref))
    assert k >= 20
    for a, r in zip(mine[:k], ref[:k]):
        assert np.linalg.norm(a - r) <= 1e-9 * np.linalg.norm(r)
    x, it = krylov.cg(lambda v: A @ v, b, tol=1e-10, maxiter=5000)
    assert np.linalg.norm(b - A @ x) <= 1e-10 * np.linalg.norm(b)
    assert np.allclose(x, spla.spsolve(A.tocsc(), b), rtol=0, atol=1e-7 * np.abs(x).max())


def test_gmres_first_cycle_is_minimum_residual():
    """One GMRES(m) cycle returns the minimiser of |b - A x| over x0 + K_m(A, r0), computed here
    by dense least squares on an explicit (orthonormalised) Krylov basis; restarts converge."""
    from oracle import krylov
    A, b = _spd_problem((5, 3, 3), 5)
    m = 6
    x1, it = krylov.gmres(lambda v: A @ v, b, tol=1e-30, restart=m, maxiter=m)
    assert it == m
    K = np.empty((b.size, m))
    v = b.copy()
    for j in range(m):
        K[:, j] = v / np.linalg.norm(v)
        v = A @ K[:, j]
    Q, _ = np.linalg.qr(K)
    y = np.linalg.lstsq(A @ Q, b, rcond=None)[0]
    xr = Q @ y
    assert np.linalg.norm(x1 - xr) <= 1e-8 * np.linalg.norm(xr)
    x, it = krylov.gmres(lambda v: A @ v, b, tol=1e-9, restart=20, maxiter=20000)
    assert np.linalg.norm(b - A @ x) <= 1e-9 * np.linalg.norm(b)
    assert it % 20 == 0


# ---------------------------------------------------------- multi-material

@pytest.mark.parametrize("rp", [False, True])
def test_multimaterial_model_pins(rp):
    """The restated PenalizedMultiMaterial(RP): the gradient is the derivative of the forward map
    (central differences), one material reduces to E_1 r^p, two materials at a vertex give the
    pure material, plain is the clipped linear mix."""
    rng = np.random.default_rng(2)
    mm = topopt.PenalizedMultiMaterial([1.0, 0.5, 0.2], [0.2, 0.1, 0.1], rp=rp)
    r = rng.uniform(0.05, 0.95, (40, 3))
    g = mm.grad(r)
    h = 1e-6
    for k in range(3):
        a, b = r.copy(), r.copy()
        a[:, k] += h
        b[:, k] -= h
        fd = (mm.forward(a) - mm.forward(b)) / (2 * h)
        assert np.max(np.abs(fd - g[:, k])) <= 1e-8 * np.max(np.abs(g))
    one = topopt.PenalizedMultiMaterial([0.7], [0.3], rp=rp)
    x = rng.uniform(0.05, 1.0, (30, 1))
    assert np.allclose(one.forward(x), 0.7 * x[:, 0] ** 3, rtol=1e-15)
    two = topopt.PenalizedMultiMaterial([1.0, 0.5], [0.2, 0.2], rp=rp, void=1e-9)
    assert np.isclose(two.forward(np.array([[1.0, 0.0]]))[0], 1.0)
    assert np.isclose(two.forward(np.array([[0.0, 1.0]]))[0], 0.5)
    assert np.allclose(two.forward(np.array([[0.3, 0.4]]), plain=True), 0.3 + 0.2)


def test_multimaterial_oc_invariants():
    """Lockstep multi-multiplier OC: bounds, move limits, each material's volume within its
    mass target to the bisection tolerance, and the total-mass variant."""
    rng = np.random.default_rng(4)
    n = 2000
    x = rng.uniform(0.05, 0.6, (n, 2))
    dc = -rng.uniform(0.0, 2.0, (n, 2))
    mm = topopt.PenalizedMultiMaterial([1.0, 0.5], [0.25, 0.15])
    xn, steps = topopt.oc_update_multi(x, dc, 1.0, float(n), 0.2, mm)
    assert steps > 10 and np.all(xn >= 0) and np.all(xn <= 1)
    assert np.all(np.abs(xn - x) <= 0.2 + 1e-15)
    vol = xn.mean(axis=0)
    assert np.all(np.abs(vol - np.array([0.25, 0.15])) < 2e-3)
    mt = topopt.PenalizedMultiMaterial([1.0, 0.5], [[0.3]], rho_mat=[1.0, 2.0])
    xt, _ = topopt.oc_update_multi(x, dc, 1.0, float(n), 1.0, mt)  # move 1: the target is reachable
    assert abs((xt.mean(axis=0) @ np.array([1.0, 2.0])) - 0.3) < 2e-3
