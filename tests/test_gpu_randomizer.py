"""Random configurations optimised on the GPU: ConfigSolve.generate_topology against the oracle
(same grid, supports, loads, penalty continuation and stopping test), the batched
generate_topologies against generate_topology problem by problem, and the multi-material
model on a random configuration."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _configs():
    from edgeto_amd.randomizer import BoundaryConditions2D as B2
    from edgeto_amd.randomizer import BoundaryConditions3D as B3
    np.random.seed(7)
    c3 = B3.generate_random_configuration(300, 700, dim=3, min_nel_per_axis=5, surface_loads=0.3, surface_bc=0.3)
    np.random.seed(11)
    c2 = B3.generate_random_configuration(300, 900, dim=2, min_nel_per_axis=8)
    np.random.seed(3)
    d2 = B2.generate_random_condition(dimensions=(24, 12), n_material=1)
    return [c3, c2, d2]


def _oracle(cfg, max_iter, tol):
    from oracle import topopt as otop
    from oracle.solvers import MultiGrid as OMG
    from edgeto_amd.batch import gmg_levels
    from edgeto_amd.randomizer.ConfigSolve import _node_ids, continuation
    dims, loads, load_v, bcs, bc_c, E_mat, vf = cfg
    dims = np.asarray(dims).astype(int)
    dim = dims.size
    prob = otop.Problem(tuple(dims), nu=0.33, r_min=1.5, size=tuple(dims / dims.max()))
    ids = _node_ids(np.asarray(bcs, float), dims)
    flags = np.asarray(bc_c).astype(bool).reshape(-1, dim)
    for d in range(dim):
        prob.cons[ids[flags[:, d]] * dim + d] = True
    lid = _node_ids(np.asarray(loads, float), dims)
    lv = np.asarray(load_v, float).reshape(-1, dim)
    for n, v in zip(lid, lv):  # a repeated node keeps its last load
        prob.f[n * dim:(n + 1) * dim] = v
    mat = otop.SingleMaterial(penalty=3.0, volume_fraction=float(np.asarray(vf).reshape(-1)[0]))
    solver = OMG(prob.nel, dim, tol=tol, n_smooth=3, n_level=gmg_levels(dims))
    return otop.optimize(prob, mat, solver, np.ones(prob.n_el), max_iter, ch_tol=1e-4, fun_tol=1e-5,
                         schedule=continuation)


@pytest.mark.parametrize("which", [0, 1, 2])
def test_generate_topology_matches_oracle(which):
    """20 OC iterations (penalty 1 then 1.4 from iteration 15) of a random configuration:
    compliance history and design against the oracle."""
    from edgeto_amd.randomizer.ConfigSolve import generate_topology
    cfg = _configs()[which]
    print(f"[randomizer] config {which}: dims {np.asarray(cfg[0]).tolist()}", flush=True)
    rho, flag, hist = generate_topology(cfg, multi=False, max_iter=20, tol=1e-11)
    xo, flag_o, ho = _oracle(cfg, 20, 1e-11)
    comp = np.array(hist["comp_history"])
    assert len(comp) == len(ho["comp"]) - 1 and flag == flag_o
    assert np.max(np.abs(comp - ho["comp"][1:]) / np.abs(ho["comp"][1:])) < 1e-6
    assert np.max(np.abs(rho.reshape(-1).cpu().numpy() - xo)) < 1e-5


def test_generate_topologies_match_single_runs():
    """The batched engine (penalty schedule per problem, per-axis supports, scaled elements,
    stopping test) against generate_topology for each configuration: the same iterations, the
    same stop, designs and compliances to 1e-10, over up to 110 iterations (past the end of
    the continuation at iteration 75)."""
    from edgeto_amd.randomizer.ConfigSolve import generate_topologies, generate_topology
    cfgs = _configs()
    batch = generate_topologies(cfgs, max_iter=110, tol=1e-10)
    for i, cfg in enumerate(cfgs):
        rho, flag, hist = generate_topology(cfg, multi=False, max_iter=110, tol=1e-10)
        rb, fb, hb = batch[i]
        print(f"[randomizer] problem {i}: {len(hist['comp_history'])} iterations, converged {flag}", flush=True)
        assert fb == flag and len(hb["comp_history"]) == len(hist["comp_history"])
        c1, c2 = np.array(hb["comp_history"]), np.array(hist["comp_history"])
        assert np.max(np.abs(c1 - c2) / np.abs(c2)) < 1e-10
        assert np.max(np.abs(rb - rho.reshape(-1).cpu().numpy())) < 1e-10


def test_multi_material_configuration():
    """A random configuration with three materials through the multi-material model: finite,
    positive compliances and, once the move limit lets the designs get there, every material
    within its volume fraction."""
    from edgeto_amd.randomizer import BoundaryConditions3D as B3
    from edgeto_amd.randomizer.ConfigSolve import generate_topology
    np.random.seed(5)
    cfg = B3.generate_random_configuration(300, 700, dim=2, min_nel_per_axis=8, n_material=3)
    E_mat, vf = cfg[5], cfg[6]
    assert E_mat.shape == (3, 1)
    rho, flag, hist = generate_topology(cfg, multi=True, max_iter=12, tol=1e-10)
    comp = np.array(hist["comp_history"])
    assert len(comp) == 12 and np.all(np.isfinite(comp)) and np.all(comp > 0)
    r = rho.reshape(-1, 3).cpu().numpy()
    assert np.all(r >= 0) and np.all(r <= 1)
    assert np.all(r.mean(axis=0) <= vf.reshape(-1) * (1 + 1e-3) + 1e-9)
