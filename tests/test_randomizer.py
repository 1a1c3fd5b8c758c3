"""Random configuration generators (edgeto_amd.randomizer) against the reference's outputs for
the same seeds (tests/golden/randomizer_vectors.npz, written by make_randomizer_vectors.py from
randomizer/BoundaryConditions2D.py and randomizer/BoundaryConditions3D.py), plus the
invariants every configuration satisfies (CPU only)."""
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import make_randomizer_vectors as golden  # noqa: E402

from edgeto_amd.randomizer import BoundaryConditions2D as B2  # noqa: E402
from edgeto_amd.randomizer import BoundaryConditions3D as B3  # noqa: E402

MODS = {"BoundaryConditions2D": B2, "BoundaryConditions3D": B3}
VEC = np.load(os.path.join(HERE, "golden", "randomizer_vectors.npz"))


def _check(key, result):
    got = golden.flatten(key, result)
    want = {k: VEC[k] for k in VEC.files if k == key or k.startswith(key + ".")}
    want.pop(key + ".tail")
    assert sorted(got) == sorted(want), key
    for k in want:
        assert got[k].shape == want[k].shape, (k, got[k].shape, want[k].shape)
        assert np.array_equal(got[k], want[k]), k
    assert int(np.random.randint(2 ** 31)) == int(VEC[key + ".tail"]), key + ": draw count differs"


@pytest.mark.parametrize("case", golden.cases(), ids=lambda c: c[0])
def test_matches_reference_draws(case):
    key, mod, fn, args, kwargs, seed = case
    np.random.seed(seed)
    _check(key, getattr(MODS[mod], fn)(*args, **kwargs))


def test_configuration_from_list_matches_reference():
    domains = np.array([[20, 10, 10], [16, 16, 8]])
    BCs = [[(np.array([[0, 0, 0], [0, 10, 10]]), np.array([[1, 1, 1], [1, 0, 1]]))],
           [(np.array([[0, 0, 0]]), np.array([[1, 1, 1]])), (np.array([[16, 16, 8]]), np.array([[0, 1, 1]]))]]
    for s in range(4):
        np.random.seed(1000 + s)
        _check(f"list_{s}", B3.create_random_configuration_from_list(domains, BCs))


@pytest.mark.parametrize("seed", range(20))
def test_configuration_invariants(seed):
    """Nodes inside the closed domain, every axis constrained somewhere, unit-length loads
    scaled by magnitudes whose largest is 1, E_mat largest 1, volume fractions within bounds."""
    np.random.seed(seed)
    dim = 2 + seed % 2
    dims, loads, load_v, bcs, bc_c, E_mat, vf = B3.generate_random_configuration(
        500, 4000, dim=dim, min_nel_per_axis=5, surface_loads=0.2, surface_bc=0.2, n_material=1 + seed % 4)
    assert dims.shape == (dim,) and np.all(dims >= 1)
    for pts in (loads, bcs):
        assert pts.shape[1] == dim
        assert np.all(pts >= 0) and np.all(pts <= dims[None])
    assert np.all(np.asarray(bc_c).sum(axis=0) > 0)
    assert load_v.shape == loads.shape and np.all(np.isfinite(load_v))
    assert E_mat.max() == 1.0 and np.all(E_mat > 0)
    assert np.all(vf >= 0.02) and 0.1 - 1e-12 <= vf.sum() <= 0.75 + 1e-12


def test_thin_domain_raises_like_the_reference():
    """Seed 9 draws a domain with an axis too short for a distributed group's half length
    (choice over range(1, n // 2) is empty): the reference raises the same ValueError."""
    np.random.seed(9)
    with pytest.raises(ValueError):
        B2.generate_random_condition(min_element_num=400, max_element_num=2500)


@pytest.mark.parametrize("seed", [0, 1, 2, 3, 4, 5, 6, 7, 8, 10])
def test_condition_2d_invariants(seed):
    np.random.seed(seed)
    dims, lp, lv, fx, fd, E_mat, vf = B2.generate_random_condition(min_element_num=400, max_element_num=2500)
    assert len(lp) == len({tuple(p) for p in lp.tolist()})  # distinct load positions
    assert np.isclose(np.linalg.norm(lv, axis=1).max(), 1.0)
    assert np.array_equal(fx, np.unique(fx, axis=0)) and fd.shape == fx.shape
    assert fd[:, 0].sum() >= 1 and fd[:, 1].sum() >= 1 and max(fd[:, 0].sum(), fd[:, 1].sum()) >= 2
    assert np.all(fd.sum(axis=1) >= 1)


def test_segment_and_surface_shapes():
    """A line of integer nodes from p1 toward p2 (rows distinct and sorted, consecutive rows one
    step apart; the running sum is truncated, so the far end can round down: here 33 steps of
    6/33 sum to just under 6 and (6, 3) becomes (5, 2), as in the reference) and a face rectangle of
    (|dx|+1)(|dy|+1) nodes."""
    line = B3._segment(np.array([0, 0, 3]), np.array([6, 3, 3]), 3)
    assert np.array_equal(line, np.unique(line, axis=0)) and np.all(line[:, 2] == 3)
    assert np.all(np.abs(np.diff(line, axis=0)) <= 1)
    assert line[0].tolist() == [0, 0, 3] and line[-1].tolist() == [5, 2, 3]
    np.random.seed(3)
    pts = B3._surface(np.array([8, 6, 5]), 3)
    axis = [a for a in range(3) if np.all(pts[:, a] == pts[0, a])]
    assert axis and pts[0, axis[0]] in (0, [8, 6, 5][axis[0]])
    others = [a for a in range(3) if a != axis[0]]
    ext = [int(pts[:, a].max() - pts[:, a].min()) + 1 for a in others]
    assert len(pts) == ext[0] * ext[1]


def test_continuation_values():
    """Penalty continuation of ConfigSolve.py:14-16: 1 + (i // 15)(p - 1)/5 below 75, then 3."""
    from edgeto_amd.randomizer.ConfigSolve import continuation
    got = [continuation(3.0, i) for i in (0, 14, 15, 29, 30, 44, 45, 60, 74, 75, 200)]
    assert got == [1.0, 1.0, 1.4, 1.4, 1.8, 1.8, 2.2, 2.6, 2.6, 3.0, 3.0]


@pytest.mark.parametrize("dims", [(4, 3), (4, 3, 2), (1, 5, 3)])
def test_node_ids_follow_the_mesh_order(dims):
    """Integer grid coordinates -> node ids, against the oracle mesh's node coordinates (the
    reference's node order); out-of-grid points snap to the nearest boundary node."""
    from oracle import mesh as omesh
    from edgeto_amd.randomizer.ConfigSolve import _node_ids
    d = np.array(dims)
    nodes = omesh.nodes_2d(*dims, *dims) if len(dims) == 2 else omesh.nodes_3d(*dims, *dims)
    assert np.array_equal(_node_ids(nodes, d), np.arange(len(nodes)))
    outside = np.array([[-3] + [0] * (len(dims) - 1), list(d + 2)])
    assert np.array_equal(_node_ids(outside, d), [0, len(nodes) - 1])


def test_oracle_schedule_pins():
    """The oracle's penalty schedule: a constant schedule equals no schedule, and the initial
    solve under the continuation equals a plain run with penalty 1 (p(0) = 1)."""
    from oracle import topopt as otop
    from oracle.solvers import MultiGrid as OMG
    from edgeto_amd.randomizer.ConfigSolve import continuation

    def run(penalty, schedule, iters):
        prob = otop.Problem((12, 6), nu=0.3, r_min=1.5)
        prob.cons[:2 * 7] = True
        prob.f[-1] = -1.0
        mat = otop.SingleMaterial(penalty=penalty, volume_fraction=0.4)
        return otop.optimize(prob, mat, OMG(prob.nel, 2, tol=1e-12, n_level=2), np.ones(prob.n_el), iters,
                             ch_tol=0, fun_tol=0, schedule=schedule)

    a = run(3.0, None, 3)[2]["comp"]
    b = run(3.0, lambda p, i: p, 3)[2]["comp"]
    assert a == b
    c = run(3.0, continuation, 0)[2]["comp"]
    d = run(1.0, None, 0)[2]["comp"]
    assert c == d
