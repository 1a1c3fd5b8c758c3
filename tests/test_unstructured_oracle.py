"""General-mesh oracle and host logic (CPU): T3 / T4 element matrices, the assembled
general operator, the KD-tree filter and the node -> element star, each pinned by an
independent check (patch test, rigid-body modes, the structured oracle, brute force)."""
import sys
import os

import numpy as np
import pytest

from oracle import filters as ofilt
from oracle import mesh as omesh
from oracle import ops as oops
from oracle import physics as ophys
from oracle import unstructured as ou

sys.path.insert(0, os.path.dirname(__file__))
import umesh_cases as cases  # noqa: E402


def _linear_field_energy(x, K, D, measure, rng, dim):
    """u = G p for a random affine field: u^T K u must be measure * eps^T D eps (constant strain)."""
    G = rng.standard_normal((dim, dim))
    u = (x @ G.T).reshape(-1)
    if dim == 2:
        eps = np.array([G[0, 0], G[1, 1], G[0, 1] + G[1, 0]])
    else:
        eps = np.array([G[0, 0], G[1, 1], G[2, 2], G[0, 1] + G[1, 0], G[1, 2] + G[2, 1], G[0, 2] + G[2, 0]])
    return u @ K @ u, measure * eps @ D @ eps


@pytest.mark.parametrize("seed", range(4))
def test_simplex_stiffness_pins(seed):
    """Oracle T3 / T4 (the reference's cofactor formulas): symmetric, rigid-body null space of
    dimension 3 / 6, and the constant-strain patch identity u^T K u = A eps:D:eps."""
    rng = np.random.default_rng(seed)
    tri = rng.uniform(0, 1, (3, 2))
    if ophys.tri_area(tri) < 0:
        tri = tri[::-1]
    tet = rng.uniform(0, 1, (4, 3))
    if ophys.tet_volume(tet) < 0:
        tet = tet[[0, 1, 3, 2]]
    for x, dim, nullity in ((tri, 2, 3), (tet, 3, 6)):
        K, D, B, A = ophys.auto_stiffness(x, 1.0, 0.3)
        assert np.allclose(K, K.T, atol=1e-12 * np.abs(K).max())
        ev = np.linalg.eigvalsh(K)
        assert np.sum(np.abs(ev) < 1e-10 * ev.max()) == nullity
        assert ev.min() > -1e-10 * ev.max()
        e1, e2 = _linear_field_energy(x, K, D, A, rng, dim)
        assert abs(e1 - e2) <= 1e-12 * abs(e2)


@pytest.mark.parametrize("seed", range(3))
def test_product_element_matrices_match_oracle(seed):
    """physics.element_matrices / auto_stiffness (gradients of barycentric coordinates, batched
    Gauss quadrature) against the oracle, both orientations of the simplices."""
    import edgeto_amd.physics as P
    rng = np.random.default_rng(10 + seed)
    tris = rng.uniform(0, 1, (6, 3, 2))
    tets = rng.uniform(0, 1, (6, 4, 3))
    q = np.array([[0, 0], [1, 0], [1, 1], [0, 1]], float) + rng.uniform(-0.2, 0.2, (4, 2))
    h = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], float)
    h = h + rng.uniform(-0.15, 0.15, (8, 3))
    for batch in (tris, tets, q[None].repeat(3, 0), h[None].repeat(3, 0)):
        Kb, Db, Bb, Ab = P.element_matrices(batch, 1.0, 0.3)
        for i, x in enumerate(batch):
            K, D, B, A = ophys.auto_stiffness(x, 1.0, 0.3)
            s = np.abs(K).max()
            assert np.abs(Kb[i] - K).max() <= 1e-13 * s
            assert np.abs(Bb[i] - B).max() <= 1e-12 * np.abs(B).max()
            assert abs(Ab[i] - A) <= 1e-14 * abs(A)
            Ks = P.auto_stiffness(x, 1.0, 0.3)[0]
            assert np.abs(Ks - K).max() <= 1e-13 * s


def test_inverted_quad_and_hex_raise():
    """Clockwise Q4 / inverted H8 raise, as the reference's final re-evaluation does."""
    import edgeto_amd.physics as P
    q = np.array([[0, 0], [0, 1], [1, 1], [1, 0]], float)
    with pytest.raises(ValueError):
        P.auto_stiffness(q, 1.0, 0.3)
    h = np.array([[0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1], [0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]], float)
    with pytest.raises(ValueError):
        P.auto_stiffness(h, 1.0, 0.3)


@pytest.mark.parametrize("nel", [(5, 4), (3, 4, 2)])
def test_general_assembly_equals_structured(nel):
    """A structured grid written as a general mesh (every element matrix K0) assembles to the
    structured oracle's operator, constraints included."""
    dim = len(nel)
    prob_el = omesh.elements_2d(*nel) if dim == 2 else omesh.elements_3d(*nel)
    n_nodes = int(np.prod(np.array(nel) + 1))
    K0 = ophys.auto_stiffness(omesh.nodes_2d(*nel, *nel)[prob_el[0]] if dim == 2 else
                              omesh.nodes_3d(*nel, *nel)[prob_el[0]], 1.0, 0.3)[0]
    rng = np.random.default_rng(1)
    E = rng.uniform(0.1, 1.0, len(prob_el))
    cons = np.zeros(n_nodes * dim, dtype=bool)
    cons[rng.choice(cons.size, cons.size // 7, replace=False)] = True
    A1 = ou.assemble([K0] * len(prob_el), prob_el, E, cons, n_nodes * dim, dim)
    A2 = oops.assemble(K0, prob_el, E, cons, n_nodes * dim)
    assert abs(A1 - A2).max() <= 1e-14 * abs(A2).max()


@pytest.mark.parametrize("case", ["tri", "tet"])
def test_general_operator_invariants(case):
    """The assembled operator of a T3 / T4 mesh: symmetric, rigid-body modes in its null space
    without constraints, positive definite once a face is clamped."""
    if case == "tri":
        nodes, els = cases.tri_mesh((4, 3), flip_every=3)
        dim = 2
    else:
        nodes, els = cases.tet_mesh((2, 2, 2), swap_every=4)
        dim = 3
    mats = [ophys.auto_stiffness(nodes[e], 1.0, 0.3) for e in els]
    n = nodes.shape[0] * dim
    E = np.ones(len(els))
    A = ou.assemble([m[0] for m in mats], els, E, np.zeros(n, bool), n, dim).toarray()
    assert np.allclose(A, A.T, atol=1e-12 * np.abs(A).max())
    ev = np.linalg.eigvalsh(A)
    assert np.sum(np.abs(ev) < 1e-10 * ev.max()) == (3 if dim == 2 else 6)
    cons = np.zeros(n, bool)
    face = np.nonzero(nodes[:, 0] == 0.0)[0]
    for d in range(dim):
        cons[face * dim + d] = True
    Ac = ou.assemble([m[0] for m in mats], els, E, cons, n, dim).toarray()
    assert np.linalg.eigvalsh(Ac).min() > 0


@pytest.mark.parametrize("nel,r", [((7, 5), 1.5), ((5, 4, 3), 2.0), ((6, 6), 2.6)])
def test_general_filter_pins(nel, r):
    """The KD-tree filter: equal to brute-force pair enumeration, and on a structured grid's
    element centroids equal to the structured oracle's normalised filter."""
    dim = len(nel)
    els = omesh.elements_2d(*nel) if dim == 2 else omesh.elements_3d(*nel)
    nodes = omesh.nodes_2d(*nel, *nel) if dim == 2 else omesh.nodes_3d(*nel, *nel)
    c = ou.centroids(nodes, els)
    H = ou.general_filter(c, r)
    assert np.abs(H.toarray() - ou.general_filter_bruteforce(c, r)).max() < 1e-15
    offs, w = ofilt.offsets_weights(r, (1.0,) * dim)
    x = np.random.default_rng(2).uniform(0, 1, len(els))
    assert np.abs(H @ x - ofilt.forward(x, nel, offs, w)).max() < 1e-14


def test_star_is_stable_argsort():
    nodes, els = cases.mixed_3d((3, 2, 2), seed=4)
    flat, ptr = ou.flat_layout(els)
    sorter, node_ptr = ou.star(flat, nodes.shape[0])
    assert np.all(np.diff(flat[sorter]) >= 0)
    for i in range(nodes.shape[0]):
        seg = sorter[node_ptr[i]:node_ptr[i + 1]]
        assert np.all(flat[seg] == i) and np.all(np.diff(seg) > 0)
