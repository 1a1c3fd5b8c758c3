"""GPU GMG-PCG solver and SIMP loop against the CPU oracle (small meshes)."""
import numpy as np
import pytest

from oracle import ops as oops
from oracle import topopt as otop
from oracle.solvers import MultiGrid as OMultiGrid

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _edgeto():
    import edgeto_amd
    return edgeto_amd


def _cantilever_3d(ET, nel, nu=0.3, rmin=2.0):
    nx, ny, nz = nel
    mesh = ET.CuStructuredMesh3D(nx, ny, nz, nx, ny, nz,
                                 stiffness=lambda x0: ET.auto_stiffness(x0, 1.0, nu))
    kernel = ET.CuStructuredStiffnessKernel(mesh)
    filt = ET.CuStructuredFilter3D(mesh, rmin)
    # clamp face x = 0: node (i=0, j, k) -> j*nx1*nz1 + k
    nx1, ny1, nz1 = nx + 1, ny + 1, nz + 1
    j, k = np.meshgrid(np.arange(ny1), np.arange(nz1), indexing="ij")
    clamp = (j * nx1 * nz1 + k).ravel()
    load = np.array([0 * nx1 * nz1 + nx * nz1 + kk for kk in range(nz1)])  # x = L, y = 0 edge
    return mesh, kernel, filt, clamp, load


def _cantilever_nodes(nel):
    """Clamped face x = 0 and the loaded edge x = L, y = 0 (node ids)."""
    nx, ny, nz = nel
    nx1, ny1, nz1 = nx + 1, ny + 1, nz + 1
    j, k = np.meshgrid(np.arange(ny1), np.arange(nz1), indexing="ij")
    clamp = (j * nx1 * nz1 + k).ravel()
    load = np.array([nx * nz1 + kk for kk in range(nz1)])
    return clamp, load


def _oracle_problem(nel, nu, rmin, clamp, load, fy):
    prob = otop.Problem(nel, nu=nu, r_min=rmin)
    dof = prob.dof
    for d in range(dof):
        prob.cons[clamp * dof + d] = True
    prob.f[load * dof + 1] = fy
    return prob


# fine-grid operator variants: (ET_TUNE_WHT, ET_TUNE_MX_F32, ET_TUNE_MX_F64); None = library default
_OPS = {"default": None, "tiled": (0, 0, 0, 0, 0, 0), "matrix": (0, 1, 1, 1, 1 << 30, 1),
        "wht": (1, 0, 0, 1, 300000, 1)}
# ET_TUNE_WHT, ET_TUNE_MX_F32, ET_TUNE_MX_F64, ET_TUNE_COARSE_LIB, ET_TUNE_ROW_STENCIL_NODES,
# ET_TUNE_LEVEL1_MF
_KNOBS = (7, 3, 4, 9, 12, 13)


@pytest.fixture
def fine_ops(request):
    from edgeto_amd import _lib as L
    L.load()
    knobs = _OPS[request.param]
    old = [L.lib().et_tune_get(k) for k in _KNOBS]
    if knobs is not None:
        for k, v in zip(_KNOBS, knobs):
            L.call("et_tune_set", k, v)
    yield request.param
    for k, v in zip(_KNOBS, old):
        L.call("et_tune_set", k, v)


@pytest.mark.parametrize("fine_ops", ["default", "tiled", "matrix", "wht"], indirect=True)
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("nel,n_level", [((16, 8, 8), 2), ((24, 12, 12), 2)])
def test_solve_matches_direct(precision, nel, n_level, fine_ops):
    ET = _edgeto()
    mesh, kernel, filt, clamp, load = _cantilever_3d(ET, nel)
    rng = np.random.default_rng(0)
    E = rng.uniform(0.05, 1.0, mesh.n_el)
    kernel.set_rho(torch.tensor(E, device="cuda"))
    tp = ET.CuTopOpt(mesh, ET.SingleMaterial(void=1e-9, penalty=3.0, volume_fraction=0.3), kernel,
                     None, filt, verbose=False)
    tp.add_BC_nodal(clamp, np.array([[1, 1, 1]]))
    tp.add_F_nodal(load, np.array([[0.0, -1.0 / len(load), 0.0]]))
    solver = ET.CuMultiGrid(mesh, kernel, tol=1e-10, n_smooth=2, n_level=n_level, precision=precision)
    x, res = solver.solve(tp.f, use_last=False)
    assert res < 1e-10
    prob = _oracle_problem(nel, 0.3, 2.0, clamp, load, -1.0 / len(load))
    A = oops.assemble(prob.K0, prob.elements, E, prob.cons, prob.n_dof)
    import scipy.sparse.linalg as spla
    xd = spla.spsolve(A.tocsc(), prob.f)
    xg = x.cpu().numpy()
    assert np.linalg.norm(prob.f - A @ xg) / np.linalg.norm(prob.f) < 2e-10
    assert np.linalg.norm(xg - xd) / np.linalg.norm(xd) < 1e-7
    if precision == "fp64":
        # same algorithm as the oracle's GPU-semantics MultiGrid: PCG iteration counts agree
        om = OMultiGrid(prob.nel, prob.dof, tol=1e-10, n_smooth=2, n_level=n_level, mode="gpu")
        om.solve(A, prob.f, use_last=False)
        assert abs(om.last_iters - solver.last_iterations) <= 1


def test_simp_trajectory_2d():
    """Short 2-D MBB run (config-1 setup at reduced size) vs the oracle, GPU omega semantics."""
    ET = _edgeto()
    nelx, nely = 40, 20
    mesh = ET.CuStructuredMesh2D(nelx, nely, nelx, nely,
                                 stiffness=lambda x0: ET.auto_stiffness(x0, 1.0, 0.3))
    kernel = ET.CuStructuredStiffnessKernel(mesh)
    filt = ET.CuStructuredFilter2D(mesh, 1.5)
    mat = ET.SingleMaterial(void=1e-9, penalty=3.0, volume_fraction=0.5)
    solver = ET.CuMultiGrid(mesh, kernel, tol=1e-8, n_smooth=3, n_level=2)
    tp = ET.CuTopOpt(mesh, mat, kernel, solver, filt, max_iter=10, ch_tol=0, fun_tol=0, verbose=False)
    nx1 = nelx + 1
    left = np.arange(nely + 1) * nx1
    tp.add_BC_nodal(left, np.array([[1, 0]]))
    tp.add_BC_nodal(np.array([nelx]), np.array([[0, 1]]))
    tp.add_F_nodal(np.array([nely * nx1]), np.array([[0.0, -1.0]]))
    x, flag, hist = tp.optimize(save_comp_history=True, rho_init=np.full(mesh.n_el, 0.5))
    prob = otop.Problem((nelx, nely), nu=0.3, r_min=1.5)
    prob.cons[left * 2] = True
    prob.cons[nelx * 2 + 1] = True
    prob.f[nely * nx1 * 2 + 1] = -1.0
    omat = otop.SingleMaterial(void=1e-9, penalty=3.0, volume_fraction=0.5)
    osol = OMultiGrid(prob.nel, 2, tol=1e-8, n_smooth=3, n_level=2, mode="gpu")
    xo, _, ohist = otop.optimize(prob, omat, osol, np.full(prob.n_el, 0.5), 10, ch_tol=0, fun_tol=0)
    comp_gpu = np.array([tp.iteration_log[0]["comp"]] + hist["comp_history"])
    comp_ref = np.array(ohist["comp"])
    assert np.max(np.abs(comp_gpu - comp_ref) / comp_ref) < 1e-6
    assert np.max(np.abs(x.cpu().numpy() - xo)) < 1e-4


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("level1_mf", [1, 0])
@pytest.mark.parametrize("nel", [(16, 8, 8), (20, 12, 10)])
def test_level_operators_match_galerkin(precision, level1_mf, nel):
    """Every smoothing level's operator (fine grid, matrix-free or stencil first coarse level,
    stencil levels) against the oracle's Galerkin products R A R^T on a clamped cantilever with
    random densities, so the constrained-DOF correction of the first coarse level is exercised."""
    ET = _edgeto()
    from edgeto_amd import _lib as L
    old = L.lib().et_tune_get(13)
    L.call("et_tune_set", 13, level1_mf)
    try:
        mesh, kernel, filt, clamp, load = _cantilever_3d(ET, nel)
        rng = np.random.default_rng(3)
        E = rng.uniform(0.01, 1.0, mesh.n_el)
        kernel.set_rho(torch.tensor(E, device="cuda"))
        tp = ET.CuTopOpt(mesh, ET.SingleMaterial(void=1e-9, penalty=3.0, volume_fraction=0.3), kernel,
                         None, filt, verbose=False)
        tp.add_BC_nodal(clamp, np.array([[1, 1, 1]]))
        tp.add_F_nodal(load, np.array([[0.0, -1.0 / len(load), 0.0]]))
        n_level = 3
        solver = ET.CuMultiGrid(mesh, kernel, tol=1e-6, n_smooth=2, n_level=n_level, precision=precision)
        solver.solve(tp.f, use_last=False)
    finally:
        L.call("et_tune_set", 13, old)
    prob = _oracle_problem(nel, 0.3, 2.0, clamp, load, -1.0 / len(load))
    A = oops.assemble(prob.K0, prob.elements, E, prob.cons, prob.n_dof).tocsr()
    from oracle import mgm as omgm
    dt = torch.float32 if precision == "mixed" else torch.float64
    tol = 2e-5 if precision == "mixed" else 1e-11
    lnel = tuple(nel)
    for level in range(n_level):
        if level > 0:
            R = omgm.restriction_matrix(lnel, 3)
            A = (R @ A @ R.T).tocsr()
            lnel = tuple(n // 2 for n in lnel)
        x = rng.standard_normal(A.shape[0])
        xt = torch.tensor(x, dtype=dt, device="cuda")
        yt = torch.empty_like(xt)
        L.call("et_mg_apply_level", solver._handle, level, L.ptr(xt), L.ptr(yt))
        ref = A @ x
        err = np.abs(yt.cpu().numpy() - ref).max() / np.abs(ref).max()
        assert err < tol, (level, err)


@pytest.mark.parametrize("coarse_lib", [2, 1, 0])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("nel", [(16, 8, 8), (24, 12, 12), (32, 16, 16)])
def test_coarse_solve_matches_dense(coarse_lib, precision, nel):
    """The coarsest-grid direct solve (band Cholesky with explicit L^-1, rocSOLVER, in-house
    dense) against a dense solve of the oracle's coarsest Galerkin operator R A R^T."""
    ET = _edgeto()
    from edgeto_amd import _lib as L
    old = L.lib().et_tune_get(9)
    L.call("et_tune_set", 9, coarse_lib)
    n_level = 2
    try:
        mesh, kernel, filt, clamp, load = _cantilever_3d(ET, nel)
        rng = np.random.default_rng(5)
        E = rng.uniform(0.01, 1.0, mesh.n_el)
        kernel.set_rho(torch.tensor(E, device="cuda"))
        tp = ET.CuTopOpt(mesh, ET.SingleMaterial(void=1e-9, penalty=3.0, volume_fraction=0.3), kernel,
                         None, filt, verbose=False)
        tp.add_BC_nodal(clamp, np.array([[1, 1, 1]]))
        tp.add_F_nodal(load, np.array([[0.0, -1.0 / len(load), 0.0]]))
        solver = ET.CuMultiGrid(mesh, kernel, tol=1e-6, n_smooth=2, n_level=n_level, precision=precision)
        solver.solve(tp.f, use_last=False)
    finally:
        L.call("et_tune_set", 9, old)
    prob = _oracle_problem(nel, 0.3, 2.0, clamp, load, -1.0 / len(load))
    A = oops.assemble(prob.K0, prob.elements, E, prob.cons, prob.n_dof).tocsr()
    from oracle import mgm as omgm
    lnel = tuple(nel)
    for _ in range(n_level):
        R = omgm.restriction_matrix(lnel, 3)
        A = (R @ A @ R.T).tocsr()
        lnel = tuple(n // 2 for n in lnel)
    b = rng.standard_normal(A.shape[0])
    xd = np.linalg.solve(A.toarray(), b)
    dt = torch.float32 if precision == "mixed" else torch.float64
    bt = torch.tensor(b, dtype=dt, device="cuda")
    xt = torch.empty_like(bt)
    L.call("et_mg_coarse_solve", solver._handle, L.ptr(bt), L.ptr(xt))
    err = np.abs(xt.cpu().numpy() - xd).max() / np.abs(xd).max()
    assert err < (1e-5 if precision == "mixed" else 1e-9), err


def _pcg_run(precision, maxiter, fused_rz, strict=False):
    """One solve of a clamped 24x12x12 cantilever with random densities under ET_TUNE_FUSED_RZ."""
    ET = _edgeto()
    from edgeto_amd import _lib as L
    nel = (24, 12, 12)
    old = L.lib().et_tune_get(14)
    L.call("et_tune_set", 14, fused_rz)
    try:
        mesh, kernel, filt, clamp, load = _cantilever_3d(ET, nel)
        rng = np.random.default_rng(11)
        E = rng.uniform(0.05, 1.0, mesh.n_el)
        kernel.set_rho(torch.tensor(E, device="cuda"))
        tp = ET.CuTopOpt(mesh, ET.SingleMaterial(void=1e-9, penalty=3.0, volume_fraction=0.3), kernel,
                         None, filt, verbose=False)
        tp.add_BC_nodal(clamp, np.array([[1, 1, 1]]))
        tp.add_F_nodal(load, np.array([[0.0, -1.0 / len(load), 0.0]]))
        solver = ET.CuMultiGrid(mesh, kernel, tol=1e-10, maxiter=maxiter, n_smooth=2, n_level=2,
                                precision=precision, strict=strict)
        x, res = solver.solve(tp.f, use_last=False)
        iters = solver.last_iterations
    finally:
        L.call("et_tune_set", 14, old)
    prob = _oracle_problem(nel, 0.3, 2.0, clamp, load, -1.0 / len(load))
    A = oops.assemble(prob.K0, prob.elements, E, prob.cons, prob.n_dof)
    xg = x.cpu().numpy()
    true_res = np.linalg.norm(prob.f - A @ xg) / np.linalg.norm(prob.f)
    return iters, res, true_res, xg


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("maxiter", [1000, 7])
def test_pcg_vector_updates(precision, maxiter):
    """PCG with r.z summed inside the last fine-grid sweep (ET_TUNE_FUSED_RZ = 1) and by its own
    pass (0), and with x += alpha p deferred to the p update: the returned x must satisfy
    |f - A x| / |f| = the reported residual, whether the loop ends on the tolerance or on
    maxiter, and both settings give the same iterates."""
    runs = []
    for fused_rz in (1, 0):
        iters, res, true_res, xg = _pcg_run(precision, maxiter, fused_rz)
        if maxiter == 7:
            assert iters == 7
            assert res > 1e-10
        else:
            assert iters < maxiter and res < 1e-10
        # the recursively updated residual tracks the true one
        assert abs(true_res - res) <= 1e-6 * res + 2e-11, (fused_rz, true_res, res)
        runs.append((iters, xg))
    (i0, x0), (i1, x1) = runs
    assert abs(i0 - i1) <= 1
    if i0 == i1:
        tol = 1e-6 if precision == "mixed" else 1e-9
        assert np.linalg.norm(x0 - x1) / np.linalg.norm(x0) < tol


def test_pcg_maxiter_status():
    """maxiter reached above the tolerance: et_mg_solve returns ET_ENOCONV; MultiGrid(strict=True)
    raises it, strict=False returns the last iterate and its residual."""
    from edgeto_amd import _lib as L
    with pytest.raises(L.EdgetoError) as ei:
        _pcg_run("fp64", 7, 1, strict=True)
    assert ei.value.status == L.ET_ENOCONV
    iters, res, true_res, _ = _pcg_run("fp64", 7, 1, strict=False)
    assert iters == 7 and res > 1e-10


# (ET_TUNE_WHT_THREADS_F32, ET_TUNE_WHT_THREADS_F64, ET_TUNE_WHT_SLICES_F32, ET_TUNE_WHT_SLICES_F64)
_TILES = {"default": (1024, 512, 16, 8), "f32_half": (512, 512, 8, 8), "f64_full": (1024, 512, 16, 16),
          "f64_256": (1024, 256, 16, 8)}


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_solve_tile_variants(precision):
    """The V-cycle's fused fine-grid modes (two zero-guess sweeps in one pass, the last sweep with
    the r.z partials) on every transformed-basis tile shape: the solution of a clamped cantilever
    with random densities against a direct solve, and each iteration count against the default
    tile's."""
    ET = _edgeto()
    from edgeto_amd import _lib as L
    import scipy.sparse.linalg as spla
    knobs = (10, 11, 15, 16)
    nel = (24, 12, 12)
    prob = _oracle_problem(nel, 0.3, 2.0, *_cantilever_nodes(nel), -1.0 / (nel[2] + 1))
    rng = np.random.default_rng(17)
    E = rng.uniform(0.05, 1.0, int(np.prod(nel)))
    A = oops.assemble(prob.K0, prob.elements, E, prob.cons, prob.n_dof)
    xd = spla.spsolve(A.tocsc(), prob.f)
    iters = {}
    for tiles, values in _TILES.items():
        old = [L.lib().et_tune_get(k) for k in knobs]
        for k, v in zip(knobs, values):
            L.call("et_tune_set", k, v)
        try:
            mesh, kernel, filt, clamp, load = _cantilever_3d(ET, nel)
            kernel.set_rho(torch.tensor(E, device="cuda"))
            tp = ET.CuTopOpt(mesh, ET.SingleMaterial(void=1e-9, penalty=3.0, volume_fraction=0.3), kernel,
                             None, filt, verbose=False)
            tp.add_BC_nodal(clamp, np.array([[1, 1, 1]]))
            tp.add_F_nodal(load, np.array([[0.0, -1.0 / len(load), 0.0]]))
            solver = ET.CuMultiGrid(mesh, kernel, tol=1e-10, n_smooth=2, n_level=2, precision=precision)
            x, res = solver.solve(tp.f, use_last=False)
            iters[tiles] = solver.last_iterations
        finally:
            for k, v in zip(knobs, old):
                L.call("et_tune_set", k, v)
        assert res < 1e-10, tiles
        xg = x.cpu().numpy()
        assert np.linalg.norm(xg - xd) / np.linalg.norm(xd) < 1e-7, tiles
    for tiles, it in iters.items():
        assert abs(iters["default"] - it) <= 1, (tiles, iters)


@pytest.mark.parametrize("nel", [(16, 8), (8, 4, 4), (12, 6, 6)])
@pytest.mark.parametrize("method", ["cg", "gmres"])
def test_krylov_solvers(nel, method):
    """CuCG / CuGMRES (unpreconditioned, matrix-free) against the oracle's restatement of the
    cupyx iterations on the same clamped problem: iteration counts, the reported true residual
    and the solution."""
    ET = _edgeto()
    from oracle import krylov
    from oracle import mesh as omesh
    dim = len(nel)
    if dim == 3:
        mesh, kernel, filt, clamp, load = _cantilever_3d(ET, nel)
        fvec = np.array([[0.0, -1.0 / len(load), 0.0]])
    else:
        mesh = ET.CuStructuredMesh2D(*nel, *nel, stiffness=lambda x0: ET.auto_stiffness(x0, 1.0, 0.3))
        kernel = ET.CuStructuredStiffnessKernel(mesh)
        filt = ET.CuStructuredFilter2D(mesh, 1.5)
        clamp = np.arange(nel[1] + 1) * (nel[0] + 1)
        load = np.array([nel[0]])
        fvec = np.array([[0.0, -1.0]])
    rng = np.random.default_rng(21)
    E = rng.uniform(0.2, 1.0, mesh.n_el)
    kernel.set_rho(torch.tensor(E, device="cuda"))
    tp = ET.CuTopOpt(mesh, ET.SingleMaterial(void=1e-9, penalty=3.0, volume_fraction=0.3), kernel,
                     None, filt, verbose=False)
    tp.add_BC_nodal(clamp, np.ones((1, dim), dtype=np.int64))
    tp.add_F_nodal(load, fvec)
    tol = 1e-8
    solver = (ET.CuCG if method == "cg" else ET.CuGMRES)(kernel, maxiter=20000, tol=tol)
    x, res = solver.solve(tp.f, use_last=False)
    cons = kernel.constraints.cpu().numpy()
    f = tp.f.cpu().numpy()
    el = omesh.elements_3d(*nel) if dim == 3 else omesh.elements_2d(*nel)
    A = oops.assemble(mesh.K_single_host, el, E, cons, kernel.shape[0])
    xg = x.cpu().numpy()
    true_res = np.linalg.norm(f - A @ xg) / np.linalg.norm(f)
    assert res <= tol and abs(true_res - res) <= 1e-3 * res
    if method == "cg":
        xo, it = krylov.cg(lambda v: A @ v, f, tol=tol, maxiter=20000)
        # rounding moves the stopping iteration of a long CG run by a few at most
        assert abs(solver.last_iterations - it) <= max(2, it // 100), (solver.last_iterations, it)
    else:
        xo, it = krylov.gmres(lambda v: A @ v, f, tol=tol, restart=20, maxiter=20000)
        assert abs(solver.last_iterations - it) <= 20, (solver.last_iterations, it)
    import scipy.sparse.linalg as spla
    xd = spla.spsolve(A.tocsc(), f)
    assert np.linalg.norm(xg - xd) / np.linalg.norm(xd) < 1e-5
    # warm start from the previous solution: converged at once
    x2, res2 = solver.solve(tp.f, use_last=True)
    assert res2 <= tol
