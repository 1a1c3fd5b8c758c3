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
