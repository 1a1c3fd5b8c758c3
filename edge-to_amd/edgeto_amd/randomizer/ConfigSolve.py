"""Optimise random configurations (reference: randomizer/ConfigSolve.py).

generate_topology(config) runs one configuration (dims, loads, load_v, bcs, bc_c, E_mat, vf)
as the reference does: a structured grid of dims unit elements scaled so that its longest
side is 1, supports and loads placed at the nodes nearest to the given coordinates (scaled
the same way), SIMP with penalty 3 reached by the continuation 1 + (i // 15)(p - 1)/5 for
OC iterations i < 75, move 0.2, stopping when the design change is below 1e-4 and the
compliance change below 1e-5 at the final penalty (or after max_iter iterations).  multi
selects the penalised multi-material model with the configuration's E_mat and volume
fractions (one OC multiplier per material), otherwise the single-material model with vf[0].

The reference builds its solver through a `Solver` class and a `generate_structured_mesh`
helper that its package no longer contains; here the same problem is set up on this
package's classes: CuStructuredMesh2D/3D (E = 1, nu = 0.33), the structured stiffness
kernel, GMG-PCG (depth from batch.gmg_levels) and a density filter of radius rmin elements
(the reference file gives no radius; 1.5 by default).  Supports are per-axis flags taken as
booleans (BCs[:, i] selects nodes).

generate_topologies(configs) runs many single-material configurations concurrently on the
batched engine (BatchTopOpt) with the same continuation and stopping test, and returns what
generate_topology returns for each.
"""
import numpy as np

from .. import geom
from ..MaterialModels import PenalizedMultiMaterial, SingleMaterial
from ..TopOpt import TopOpt
from ..batch import BatchTopOpt, gmg_levels
from ..kernels import CuStructuredStiffnessKernel
from ..solvers import CuMultiGrid

__all__ = ["continuation", "generate_topology", "generate_topologies"]

PENALTY = 3.0


def continuation(p, i):
    """Penalty of OC iteration i (ConfigSolve.py:14-16): 1, 1 + (p-1)/5, ... in steps of 15
    iterations, then 3 from iteration 75 on."""
    return 3.0 if i >= 75 else 1 + (i // 15) * (p - 1) / 5


def _unpack(config):
    dims, loads, load_v, bcs, bc_c, E_mat, vf = config
    dims = np.asarray(dims).astype(int).reshape(-1)
    if dims.size not in (2, 3) or np.any(dims < 1):
        raise ValueError(f"configuration dims must be 2 or 3 positive element counts, got {dims}")
    return (dims, np.asarray(loads, dtype=np.float64), np.asarray(load_v, dtype=np.float64),
            np.asarray(bcs, dtype=np.float64), np.asarray(bc_c), np.asarray(E_mat, dtype=np.float64),
            np.asarray(vf, dtype=np.float64))


def _material(E_mat, vf, multi):
    if multi:
        return PenalizedMultiMaterial(n_material=E_mat.shape[0], E_mat=E_mat, mass=vf, penalty=PENALTY,
                                      penalty_schedule=continuation)
    return SingleMaterial(penalty=PENALTY, volume_fraction=float(vf.reshape(-1)[0]), penalty_schedule=continuation)


def _structured(dims, material, rmin, max_iter, tol, n_smooth, precision):
    """TopOpt on a grid of dims unit elements scaled to a longest side of 1."""
    L = dims / dims.max()
    if dims.size == 2:
        mesh = geom.CuStructuredMesh2D(int(dims[0]), int(dims[1]), float(L[0]), float(L[1]))
        filt = geom.CuStructuredFilter2D(mesh, rmin)
    else:
        mesh = geom.CuStructuredMesh3D(int(dims[0]), int(dims[1]), int(dims[2]), float(L[0]), float(L[1]),
                                       float(L[2]))
        filt = geom.CuStructuredFilter3D(mesh, rmin)
    kernel = CuStructuredStiffnessKernel(mesh)
    solver = CuMultiGrid(mesh, kernel, maxiter=1000, tol=tol, n_smooth=n_smooth, n_level=gmg_levels(dims),
                         precision=precision)
    return TopOpt(mesh, material, kernel, solver, filt, max_iter=max_iter, move=0.2, ch_tol=1e-4, fun_tol=1e-5,
                  verbose=False)


def generate_topology(config, multi=True, solver_instance=None, gpu=False, rmin=1.5, max_iter=1000, tol=1e-5,
                      n_smooth=3, precision="fp64"):
    """Optimise one configuration; returns (rho, converged, {"comp_history": [...]}) as
    TopOpt.optimize does.  solver_instance: a TopOpt of the same grid to reuse (its material
    model is replaced).  gpu is accepted for the reference's signature: this package always
    runs on the GPU."""
    dims, loads, load_v, bcs, bc_c, E_mat, vf = _unpack(config)
    material = _material(E_mat, vf, multi)
    if solver_instance is None:
        solver_instance = _structured(dims, material, rmin, max_iter, tol, n_smooth, precision)
    else:
        solver_instance.material_model = material
    scale = float(dims.max())
    solver_instance.reset_BC()
    solver_instance.reset_F()
    solver_instance.add_BCs(bcs / scale, bc_c.astype(bool))
    solver_instance.add_Forces(loads / scale, load_v)
    return solver_instance.optimize(save_comp_history=True, save_change_history=False, save_rho_history=False)


def _node_ids(points, dims):
    """Ids of the grid nodes nearest to integer-grid coordinates (clipped into the grid), in
    the structured node order (y slowest, z fastest; 2-D: y slowest)."""
    p = np.clip(np.rint(points).astype(np.int64), 0, dims[None])
    nn = dims + 1
    if dims.size == 2:
        return p[:, 1] * nn[0] + p[:, 0]
    return (p[:, 1] * nn[0] + p[:, 0]) * nn[2] + p[:, 2]


def generate_topologies(configs, max_iter=1000, rmin=1.5, tol=1e-5, n_smooth=3, precision="fp64", n_workers=None):
    """Optimise many single-material configurations at once on the batched engine; returns
    [(rho, converged, {"comp_history": [...]}), ...] in the order of configs."""
    problems = []
    for cfg in configs:
        dims, loads, load_v, bcs, bc_c, E_mat, vf = _unpack(cfg)
        if E_mat.shape[0] != 1:
            raise ValueError("generate_topologies runs single-material configurations; "
                             "use generate_topology for multi-material ones")
        # element edges as the scaled grid has them: (dims / max) / dims per axis
        h = (dims / dims.max()) / dims
        problems.append(dict(dim=int(dims.size), nel=tuple(int(v) for v in dims), h=h,
                             nu=0.33, loads=_node_ids(loads, dims), dirs=load_v.reshape(-1, dims.size),
                             fix_nodes=_node_ids(bcs, dims), fix_flags=bc_c.astype(bool).reshape(-1, dims.size),
                             vf=float(vf.reshape(-1)[0]), rmin=rmin, penal=PENALTY,
                             penal_schedule=[continuation(PENALTY, k) for k in range(max(min(max_iter, 76), 1))],
                             ch_tol=1e-4, fun_tol=1e-5, move=0.2, n_level=gmg_levels(dims), x0=1.0))
    bt = BatchTopOpt(problems, precision=precision, tol=tol, maxiter=1000, n_smooth=n_smooth, void=1e-6)
    res = bt.run(max_iter, n_workers=n_workers)
    return [(bt.design(i), pr["converged"], {"comp_history": list(pr["compliance"][1:])})
            for i, pr in enumerate(res["problems"])]
