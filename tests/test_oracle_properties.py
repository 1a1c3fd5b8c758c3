"""Property tests of the oracle over randomised sizes and parameters (hypothesis, CPU): the
filter is the transpose of its adjoint and normalised, the transfers satisfy R 1 = 1 and P = R^T,
the assembled operator is symmetric with the rigid-body null space and the Galerkin operator
is R A R^T, and the OC update respects its move limits and volume bracket."""
import numpy as np
import pytest

hypothesis = pytest.importorskip("hypothesis")
from hypothesis import given, settings, strategies as st  # noqa: E402

from oracle import filters as ofilt  # noqa: E402
from oracle import mesh as omesh  # noqa: E402
from oracle import mgm  # noqa: E402
from oracle import ops as oops  # noqa: E402
from oracle import physics as ophys  # noqa: E402
from oracle import topopt as otop  # noqa: E402

SET = settings(max_examples=25, deadline=None, derandomize=True)


def _brick(dim, lx, ly, lz):
    if dim == 2:
        return np.array([[0, 0], [lx, 0], [lx, ly], [0, ly]], float)
    return np.array([[0, 0, 0], [lx, 0, 0], [lx, ly, 0], [0, ly, 0],
                     [0, 0, lz], [lx, 0, lz], [lx, ly, lz], [0, ly, lz]], float)


@SET
@given(dim=st.sampled_from([2, 3]), n=st.lists(st.integers(2, 7), min_size=3, max_size=3),
       rmin=st.floats(1.0, 3.2), sc=st.lists(st.floats(1.0, 2.0), min_size=3, max_size=3),
       seed=st.integers(0, 2 ** 16))
def test_filter_adjoint_and_normalisation(dim, n, rmin, sc, seed):
    """<H x, y> = <x, H^T y> for the normalised filter, H 1 = 1, and every weight positive."""
    nel = tuple(n[:dim])
    offs, w = ofilt.offsets_weights(rmin, tuple(sc[:dim]))
    assert np.all(w > 0) and len(w) == len(offs)
    norm = ofilt.normalizer(nel, offs, w)
    rng = np.random.default_rng(seed)
    x, y = rng.standard_normal(int(np.prod(nel))), rng.standard_normal(int(np.prod(nel)))
    Hx = ofilt.forward(x, nel, offs, w)
    HTy = ofilt.transpose(y, nel, offs, w, norm)
    assert abs(Hx @ y - x @ HTy) <= 1e-12 * (np.abs(Hx).sum() * np.abs(y).max() + 1.0)
    ones = np.ones(int(np.prod(nel)))
    assert np.allclose(ofilt.forward(ones, nel, offs, w), 1.0, atol=1e-14)


@SET
@given(dim=st.sampled_from([2, 3]), half=st.lists(st.integers(1, 4), min_size=3, max_size=3))
def test_transfers_partition_of_unity_and_transpose(dim, half):
    """R 1 = 1 and P = R^T on even grids of any aspect ratio."""
    nel = tuple(2 * h for h in half[:dim])
    R = mgm.restriction_matrix(nel, dim)
    P = mgm.prolongation_matrix(nel, dim)
    assert np.allclose(np.asarray(R.sum(axis=1)).ravel(), 1.0, atol=1e-14)
    assert abs(P - R.T).max() < 1e-15


@SET
@given(dim=st.sampled_from([2, 3]), n=st.lists(st.integers(1, 3), min_size=3, max_size=3),
       l=st.lists(st.floats(0.3, 3.0), min_size=3, max_size=3), nu=st.floats(0.0, 0.45),
       seed=st.integers(0, 2 ** 16))
def test_assembled_operator_symmetric_with_rigid_modes(dim, n, l, nu, seed):
    """Symmetric; without supports it annihilates the 3 (2-D) or 6 (3-D) rigid-body modes of the
    grid's node coordinates, for any brick aspect ratio, Poisson ratio and densities."""
    nel = tuple(2 * v for v in n[:dim])
    K0 = ophys.auto_stiffness(_brick(dim, *l), 1.0, nu)[0]
    els = omesh.elements_2d(*nel) if dim == 2 else omesh.elements_3d(*nel)
    size = tuple(l[a] * nel[a] for a in range(dim))
    nodes = omesh.nodes_2d(*nel, *size) if dim == 2 else omesh.nodes_3d(*nel, *size)
    E = np.random.default_rng(seed).uniform(0.05, 1.0, len(els))
    A = oops.assemble(K0, els, E, None, nodes.shape[0] * dim)
    s = abs(A).max()
    assert abs(A - A.T).max() <= 1e-13 * s
    modes = [np.eye(dim)[d][None].repeat(nodes.shape[0], 0) for d in range(dim)]
    if dim == 2:
        modes.append(np.stack([-nodes[:, 1], nodes[:, 0]], 1))
    else:
        x, y, z = nodes.T
        modes += [np.stack([-y, x, 0 * x], 1), np.stack([0 * x, -z, y], 1), np.stack([z, 0 * x, -x], 1)]
    for m in modes:
        v = m.reshape(-1)
        assert np.abs(A @ v).max() <= 1e-11 * s * np.abs(v).max()


@SET
@given(dim=st.sampled_from([2, 3]), half=st.lists(st.integers(1, 3), min_size=3, max_size=3),
       seed=st.integers(0, 2 ** 16))
def test_galerkin_is_triple_product(dim, half, seed):
    """The oracle's coarse operator equals R A R^T assembled explicitly (constraints included)."""
    nel = tuple(2 * h for h in half[:dim])
    K0 = ophys.auto_stiffness(_brick(dim, 1.0, 1.0, 1.0), 1.0, 0.3)[0]
    els = omesh.elements_2d(*nel) if dim == 2 else omesh.elements_3d(*nel)
    n_dof = int(np.prod(np.array(nel) + 1)) * dim
    rng = np.random.default_rng(seed)
    cons = np.zeros(n_dof, bool)
    cons[rng.choice(n_dof, max(1, n_dof // 9), replace=False)] = True
    A = oops.assemble(K0, els, rng.uniform(0.1, 1.0, len(els)), cons, n_dof)
    R = mgm.restriction_matrix(nel, dim)
    Ac = mgm.galerkin(A, nel, dim)
    ref = R @ A @ R.T
    assert abs(Ac - ref).max() <= 1e-13 * abs(ref).max()


@SET
@given(n=st.integers(8, 400), vf=st.floats(0.1, 0.9), move=st.floats(0.05, 0.4),
       seed=st.integers(0, 2 ** 16))
def test_oc_update_bounds_and_volume(n, vf, move, seed):
    """The new design stays within [max(0, x - move), min(1, x + move)], and the final multiplier
    brackets the volume target: the volume is at most vf unless the move limits forbid it."""
    rng = np.random.default_rng(seed)
    x = rng.uniform(0.0, 1.0, n)
    dc = -rng.exponential(1.0, n)
    mat = otop.SingleMaterial(volume_fraction=vf)
    xn, steps = otop.oc_update(x, dc, 1.0, float(n), vf, move, mat)
    lo, hi = np.clip(x - move, 0, 1), np.clip(x + move, 0, 1)
    assert np.all(xn >= lo - 1e-15) and np.all(xn <= hi + 1e-15)
    assert steps > 0
    if lo.mean() <= vf:  # the target is reachable within the move limits
        assert xn.mean() <= vf * (1 + 1e-3) + 1e-12
