"""Differentiation (config-4 setup at reduced size): the adjoint backward of the
compliance function against central finite differences and against the
oracle's sensitivity array."""
import numpy as np
import pytest

from oracle import ops as oops
from oracle import topopt as otop

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _setup(nel, seed, tol):
    import edgeto_amd as ET
    nx, ny, nz = nel
    rng = np.random.default_rng(seed)
    mesh = ET.CuStructuredMesh3D(nx, ny, nz, nx, ny, nz, stiffness=lambda x0: ET.auto_stiffness(x0, 1.0, 0.3))
    kernel = ET.CuStructuredStiffnessKernel(mesh)
    filt = ET.CuStructuredFilter3D(mesh, 2.0)
    mat = ET.SingleMaterial(void=1e-9, penalty=3.0, volume_fraction=0.3)
    solver = ET.CuMultiGrid(mesh, kernel, tol=tol, n_smooth=2, n_level=2)
    tp = ET.CuTopOpt(mesh, mat, kernel, solver, filt, verbose=False)
    nx1, ny1, nz1 = nx + 1, ny + 1, nz + 1
    j, k = np.meshgrid(np.arange(ny1), np.arange(nz1), indexing="ij")
    clamp = (j * nx1 * nz1 + k).ravel()
    tp.add_BC_nodal(clamp, np.array([[1, 1, 1]]))
    # 16 distinct top-surface nodes (y = ny), x > 0, loads ~ N(0, 1)
    top = np.array([(ny * nx1 + i) * nz1 + kk for i in range(1, nx1) for kk in range(nz1)])
    nodes = rng.choice(top, 16, replace=False)
    F = rng.standard_normal((16, 3))
    tp.add_F_nodal(nodes, F)
    x = rng.uniform(0.1, 1.0, mesh.n_el)
    return ET, tp, x, clamp, nodes, F


def test_backward_matches_oracle_sensitivity():
    nel = (16, 8, 8)
    ET, tp, x, clamp, nodes, F = _setup(nel, 0, 1e-12)
    xt = torch.tensor(x, device="cuda", requires_grad=True)
    c = ET.compliance(xt, tp)
    c.backward()
    g = xt.grad.cpu().numpy()
    prob = otop.Problem(nel, nu=0.3, r_min=2.0)
    for d in range(3):
        prob.cons[clamp * 3 + d] = True
    for n, fv in zip(nodes, F):
        prob.f[n * 3:(n + 1) * 3] = fv
    mat = otop.SingleMaterial(void=1e-9, penalty=3.0)
    import scipy.sparse.linalg as spla
    E = mat.forward(prob.filt(x))
    A = oops.assemble(prob.K0, prob.elements, E, prob.cons, prob.n_dof)
    U = spla.spsolve(A.tocsc(), prob.f)
    dr = prob.filt_T(mat.grad(prob.filt(x)) * oops.process_dk(prob.K0, prob.elements, U))
    assert abs(c.item() - prob.f @ U) / abs(prob.f @ U) < 1e-9
    assert np.max(np.abs(g - dr)) / np.max(np.abs(dr)) < 1e-7


def test_backward_central_differences():
    nel = (32, 16, 16)
    ET, tp, x, *_ = _setup(nel, 1, 1e-11)
    xt = torch.tensor(x, device="cuda", requires_grad=True)
    c = ET.compliance(xt, tp)
    c.backward()
    g = xt.grad.cpu().numpy()
    rng = np.random.default_rng(2)
    picks = rng.choice(x.size, 8, replace=False)
    h = 1e-6

    def c_of(xv):
        xtt = torch.tensor(xv, device="cuda")
        ET.compliance(xtt, tp)
        return ET.energy_compliance(tp.solver.last_x0, tp)

    for e in picks:
        xp, xm = x.copy(), x.copy()
        xp[e] += h
        xm[e] -= h
        fd = (c_of(xp) - c_of(xm)) / (2 * h)
        assert abs(fd - g[e]) <= 2e-3 * abs(g[e]) + 1e-9 * abs(c.item()), (e, fd, g[e])
