"""Each configuration of the benchmarked workload set run at its own size on the GPU.

c1  2-D MBB half-beam 120 x 40 Q4, 50 OC iterations: compliance history against the oracle
    (GPU semantics of CuMultiGrid, TopOpt/_cuda.py:304-382) to 1e-6, final design to 1e-4.
c2  3-D cantilever 96 x 48 x 48 H8, FP64, 4-grid GMG V(2,2)-PCG: the first OC iterations
    against the oracle, then 50 OC iterations with per-iteration properties.
c3  3-D cantilever 256 x 128 x 128 H8: 30 OC iterations mixed (FP32 V-cycle) against FP64,
    relative compliance difference <= 1e-6; the full-size fine operator three ways.
c4  3-D 128 x 64 x 64 H8, seeded design and loads: the adjoint gradient of compliance()
    against central differences (step 1e-6, PCG tol 1e-12) on 64 seeded elements.
Also a 3-D 50-iteration FP64 trajectory against the oracle at a size the oracle runs in seconds.
"""
import numpy as np
import pytest

from oracle import topopt as otop
from oracle.solvers import MultiGrid as OMultiGrid

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _cantilever(ET, nel, precision="fp64", n_level=3, tol=1e-8, vf=0.3, rmin=2.0, max_iter=0):
    nx, ny, nz = nel
    mesh = ET.CuStructuredMesh3D(nx, ny, nz, nx, ny, nz,
                                 stiffness=lambda x0: ET.auto_stiffness(x0, 1.0, 0.3))
    kernel = ET.CuStructuredStiffnessKernel(mesh)
    filt = ET.CuStructuredFilter3D(mesh, rmin)
    mat = ET.SingleMaterial(void=1e-9, penalty=3.0, volume_fraction=vf)
    solver = ET.CuMultiGrid(mesh, kernel, tol=tol, n_smooth=2, n_level=n_level, precision=precision)
    tp = ET.CuTopOpt(mesh, mat, kernel, solver, filt, max_iter=max_iter, ch_tol=0, fun_tol=0, verbose=False)
    nx1, ny1, nz1 = nx + 1, ny + 1, nz + 1
    j, k = np.meshgrid(np.arange(ny1), np.arange(nz1), indexing="ij")
    clamp = (j * nx1 * nz1 + k).ravel()
    load = np.array([nx * nz1 + kk for kk in range(nz1)])  # x = L, y = 0, every z
    tp.add_BC_nodal(clamp, np.array([[1, 1, 1]]))
    tp.add_F_nodal(load, np.array([[0.0, -1.0 / len(load), 0.0]]))
    return mesh, tp, clamp, load


def _oracle_cantilever(nel, clamp, load, vf=0.3, rmin=2.0):
    prob = otop.Problem(nel, nu=0.3, r_min=rmin)
    for d in range(3):
        prob.cons[clamp * 3 + d] = True
    prob.f[load * 3 + 1] = -1.0 / len(load)
    return prob, otop.SingleMaterial(void=1e-9, penalty=3.0, volume_fraction=vf)


def _history(tp, hist):
    return np.array([tp.iteration_log[0]["comp"]] + hist["comp_history"])


def test_c1_mbb_120x40_50_iterations():
    import edgeto_amd as ET
    nelx, nely = 120, 40
    mesh = ET.CuStructuredMesh2D(nelx, nely, nelx, nely,
                                 stiffness=lambda x0: ET.auto_stiffness(x0, 1.0, 0.3))
    kernel = ET.CuStructuredStiffnessKernel(mesh)
    filt = ET.CuStructuredFilter2D(mesh, 1.5)
    mat = ET.SingleMaterial(void=1e-9, penalty=3.0, volume_fraction=0.5)
    solver = ET.CuMultiGrid(mesh, kernel, tol=1e-8, n_smooth=3, omega=0.5, n_level=3)
    tp = ET.CuTopOpt(mesh, mat, kernel, solver, filt, max_iter=50, move=0.2, ch_tol=0, fun_tol=0,
                     verbose=False)
    nx1 = nelx + 1
    left = np.arange(nely + 1) * nx1
    tp.add_BC_nodal(left, np.array([[1, 0]]))
    tp.add_BC_nodal(np.array([nelx]), np.array([[0, 1]]))
    tp.add_F_nodal(np.array([nely * nx1]), np.array([[0.0, -1.0]]))
    x, flag, hist = tp.optimize(save_comp_history=True, rho_init=np.full(mesh.n_el, 0.5))
    comp_gpu = _history(tp, hist)
    prob = otop.Problem((nelx, nely), nu=0.3, r_min=1.5)
    prob.cons[left * 2] = True
    prob.cons[nelx * 2 + 1] = True
    prob.f[nely * nx1 * 2 + 1] = -1.0
    omat = otop.SingleMaterial(void=1e-9, penalty=3.0, volume_fraction=0.5)
    osol = OMultiGrid(prob.nel, 2, tol=1e-8, n_smooth=3, omega=0.5, n_level=3, mode="gpu")
    xo, _, oh = otop.optimize(prob, omat, osol, np.full(prob.n_el, 0.5), 50, ch_tol=0, fun_tol=0)
    comp_ref = np.array(oh["comp"])
    assert comp_gpu.shape == comp_ref.shape == (51,)
    assert np.max(np.abs(comp_gpu - comp_ref) / comp_ref) < 1e-6
    assert abs(comp_gpu[-1] - comp_ref[-1]) / comp_ref[-1] < 1e-6
    xg = x.cpu().numpy()
    assert np.max(np.abs(xg - xo)) < 1e-4
    # the design meets the volume constraint to the bisection tolerance
    assert abs(xg.mean() - 0.5) < 1e-3


def test_c2_cantilever_96x48x48_fp64():
    import edgeto_amd as ET
    nel = (96, 48, 48)
    mesh, tp, clamp, load = _cantilever(ET, nel, "fp64", n_level=3, max_iter=2)
    x, _, hist = tp.optimize(save_comp_history=True, rho_init=np.full(mesh.n_el, 0.3))
    comp_gpu = _history(tp, hist)
    prob, omat = _oracle_cantilever(nel, clamp, load)
    osol = OMultiGrid(prob.nel, 3, tol=1e-8, n_smooth=2, n_level=3, mode="gpu")
    xo, _, oh = otop.optimize(prob, omat, osol, np.full(prob.n_el, 0.3), 2, ch_tol=0, fun_tol=0)
    comp_ref = np.array(oh["comp"])
    assert np.max(np.abs(comp_gpu - comp_ref) / comp_ref) < 1e-6, (comp_gpu, comp_ref)
    assert np.max(np.abs(x.cpu().numpy() - xo)) < 1e-6
    # PCG iteration counts of the same algorithm agree (rounding may move one by one)
    pcg_gpu = [e["pcg"] for e in tp.iteration_log]
    assert all(abs(a - b) <= 1 for a, b in zip(pcg_gpu, oh["pcg"])), (pcg_gpu, oh["pcg"])

    # 50 OC iterations on the GPU alone: every solve converged, designs feasible, finite
    mesh, tp, clamp, load = _cantilever(ET, nel, "fp64", n_level=3, max_iter=50)
    x, _, hist = tp.optimize(save_comp_history=True, rho_init=np.full(mesh.n_el, 0.3))
    log = tp.iteration_log
    assert len(log) == 51
    assert all(e["res"] < 1e-8 and np.isfinite(e["comp"]) and e["comp"] > 0 for e in log)
    xs = x.cpu().numpy()
    assert xs.min() >= 0.0 and xs.max() <= 1.0
    assert abs(xs.mean() - 0.3) < 1e-3
    comp = _history(tp, hist)
    assert comp[-1] < 0.5 * comp[0]  # the design improves on the uniform start


def test_c3_mixed_vs_fp64_30_iterations():
    """The benchmark's validity gate as a test: both precisions from the same start."""
    import edgeto_amd as ET
    nel = (256, 128, 128)
    comp, pcg = {}, {}
    for precision in ("mixed", "fp64"):
        mesh, tp, _, _ = _cantilever(ET, nel, precision, n_level=4, max_iter=30)
        x, _, hist = tp.optimize(save_comp_history=True, rho_init=np.full(mesh.n_el, 0.3))
        comp[precision] = _history(tp, hist)
        pcg[precision] = [e["pcg"] for e in tp.iteration_log]
        assert all(e["res"] < 1e-8 for e in tp.iteration_log)
        del tp, mesh
        torch.cuda.empty_cache()
    rel = np.abs(comp["mixed"] - comp["fp64"]) / comp["fp64"]
    assert comp["mixed"].shape == (31,)
    assert rel[-1] <= 1e-6 and rel.max() <= 1e-6, rel.max()
    # the FP32 preconditioner changes no iteration count by more than a few
    assert max(abs(a - b) for a, b in zip(pcg["mixed"], pcg["fp64"])) <= 3


def test_c3_fine_operator_full_size():
    """y = A x at 256 x 128 x 128 (clamped face, random densities) by the transformed-basis
    operator (default), the LDS-tiled node gather and the FP64 matrix-core operator."""
    import edgeto_amd as ET
    from edgeto_amd import _lib as L
    nel = (256, 128, 128)
    nx, ny, nz = nel
    mesh = ET.CuStructuredMesh3D(nx, ny, nz, nx, ny, nz, stiffness=lambda x0: ET.auto_stiffness(x0, 1.0, 0.3))
    kernel = ET.CuStructuredStiffnessKernel(mesh)
    nx1, ny1, nz1 = nx + 1, ny + 1, nz + 1
    j, k = np.meshgrid(np.arange(ny1), np.arange(nz1), indexing="ij")
    clamp = (j * nx1 * nz1 + k).ravel()
    kernel.add_constraints(np.concatenate([clamp * 3 + d for d in range(3)]))
    g = torch.Generator(device="cuda").manual_seed(3)
    E = torch.rand(mesh.n_el, dtype=torch.float64, device="cuda", generator=g) * 0.95 + 0.05
    u = torch.randn(kernel.shape[0], dtype=torch.float64, device="cuda", generator=g)
    kernel.set_rho(E)
    knobs = (7, 3, 4)  # ET_TUNE_WHT, ET_TUNE_MX_F32, ET_TUNE_MX_F64
    old = [L.lib().et_tune_get(k) for k in knobs]
    out = {}
    try:
        for name, vals in (("wht", (1, 0, 0)), ("tiled", (0, 0, 0)), ("matrix", (0, 1, 1))):
            for k, v in zip(knobs, vals):
                L.call("et_tune_set", k, v)
            out[name] = kernel.dot(u).clone()
    finally:
        for k, v in zip(knobs, old):
            L.call("et_tune_set", k, v)
    ref = out["tiled"]
    scale = ref.abs().max().item()
    for name in ("wht", "matrix"):
        assert (out[name] - ref).abs().max().item() / scale < 1e-12, name
    # constrained rows copy the input
    cons = kernel.constraints
    assert torch.equal(out["wht"][cons], u[cons])


def test_c4_gradient_central_differences_128x64x64():
    import edgeto_amd as ET
    nel = (128, 64, 64)
    nx, ny, nz = nel
    rng = np.random.default_rng(0)
    mesh = ET.CuStructuredMesh3D(nx, ny, nz, nx, ny, nz, stiffness=lambda x0: ET.auto_stiffness(x0, 1.0, 0.3))
    kernel = ET.CuStructuredStiffnessKernel(mesh)
    filt = ET.CuStructuredFilter3D(mesh, 2.0)
    mat = ET.SingleMaterial(void=1e-9, penalty=3.0, volume_fraction=0.3)
    solver = ET.CuMultiGrid(mesh, kernel, tol=1e-12, n_smooth=2, n_level=3)
    tp = ET.CuTopOpt(mesh, mat, kernel, solver, filt, verbose=False)
    nx1, ny1, nz1 = nx + 1, ny + 1, nz + 1
    j, k = np.meshgrid(np.arange(ny1), np.arange(nz1), indexing="ij")
    tp.add_BC_nodal((j * nx1 * nz1 + k).ravel(), np.array([[1, 1, 1]]))
    x = rng.uniform(0.1, 1.0, mesh.n_el)
    top = np.array([(ny * nx1 + i) * nz1 + kk for i in range(1, nx1) for kk in range(nz1)])
    nodes = rng.choice(top, 16, replace=False)
    F = rng.standard_normal((16, 3))
    tp.add_F_nodal(nodes, F)
    picks = rng.choice(mesh.n_el, 64, replace=False)
    xt = torch.tensor(x, device="cuda", requires_grad=True)
    c = ET.compliance(xt, tp)
    c.backward()
    g = xt.grad.cpu().numpy()
    h = 1e-6

    def c_of(xv):
        ET.compliance(torch.tensor(xv, device="cuda"), tp)
        return ET.energy_compliance(tp.solver.last_x0, tp)

    fd = np.empty(len(picks))
    for n, e in enumerate(picks):
        xp, xm = x.copy(), x.copy()
        xp[e] += h
        xm[e] -= h
        fd[n] = (c_of(xp) - c_of(xm)) / (2 * h)
    ge = g[picks]
    # rounding of c at step h bounds the difference quotient's accuracy: ~1e-16 |c| / h
    floor = 1e-9 * abs(c.item())
    err = np.abs(fd - ge)
    bad = int(np.sum(err > 1e-4 * np.abs(ge) + floor))
    assert bad == 0, (bad, float(err.max()), float(np.abs(ge).min()), float(np.median(err / np.abs(ge))))
    assert np.all(ge <= 0)  # adding material never raises compliance


def test_fp64_trajectory_3d_50_iterations():
    """FP64 parity over a whole run: compliance after 50 OC iterations within 1e-6 of the
    oracle (3-D cantilever small enough for the oracle's 50 solves)."""
    import edgeto_amd as ET
    nel = (24, 12, 12)
    mesh, tp, clamp, load = _cantilever(ET, nel, "fp64", n_level=2, max_iter=50)
    x, _, hist = tp.optimize(save_comp_history=True, rho_init=np.full(mesh.n_el, 0.3))
    comp_gpu = _history(tp, hist)
    prob, omat = _oracle_cantilever(nel, clamp, load)
    osol = OMultiGrid(prob.nel, 3, tol=1e-8, n_smooth=2, n_level=2, mode="gpu")
    xo, _, oh = otop.optimize(prob, omat, osol, np.full(prob.n_el, 0.3), 50, ch_tol=0, fun_tol=0)
    comp_ref = np.array(oh["comp"])
    assert abs(comp_gpu[-1] - comp_ref[-1]) / comp_ref[-1] < 1e-6
    assert np.max(np.abs(comp_gpu - comp_ref) / comp_ref) < 1e-6
    assert np.max(np.abs(x.cpu().numpy() - xo)) < 1e-4
