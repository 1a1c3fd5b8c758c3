"""Failure handling, forced from the host through the inputs (every kernel still exits
normally; every loop is bounded by maxiter): a zero load, a structure without supports,
non-finite densities.  None may end as a silent success."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _setup(ET, nel=(16, 8, 8), precision="mixed", clamp=True, maxiter=1000, tol=1e-8):
    nx, ny, nz = nel
    mesh = ET.CuStructuredMesh3D(nx, ny, nz, nx, ny, nz, stiffness=lambda x0: ET.auto_stiffness(x0, 1.0, 0.3))
    kernel = ET.CuStructuredStiffnessKernel(mesh)
    nx1, ny1, nz1 = nx + 1, ny + 1, nz + 1
    if clamp:
        j, k = np.meshgrid(np.arange(ny1), np.arange(nz1), indexing="ij")
        kernel.set_constraints(np.concatenate([(j * nx1 * nz1 + k).ravel() * 3 + d for d in range(3)]))
    kernel.set_rho(torch.full((mesh.n_el,), 0.5, dtype=torch.float64, device="cuda"))
    f = torch.zeros(kernel.shape[0], dtype=torch.float64, device="cuda")
    f[(nx * nz1 + np.arange(nz1)) * 3 + 1] = -1.0 / nz1
    solver = ET.CuMultiGrid(mesh, kernel, tol=tol, maxiter=maxiter, n_smooth=2, n_level=2, precision=precision)
    return mesh, kernel, f, solver


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_zero_load_returns_zero(precision):
    """b = 0: x = 0 with residual 0 and no iterations, also from a non-zero warm start (the
    reference's relative test divides by |b| = 0 and returns NaN)."""
    import edgeto_amd as ET
    mesh, kernel, f, solver = _setup(ET, precision=precision)
    x, res = solver.solve(f)  # a warm start for the next solve
    assert torch.count_nonzero(x) > 0
    x0, res0 = solver.solve(torch.zeros_like(f))
    assert res0 == 0.0 and solver.last_iterations == 0
    assert torch.count_nonzero(x0) == 0


@pytest.mark.parametrize("method", ["cg", "gmres"])
def test_krylov_zero_load_returns_zero(method):
    """CuCG / CuGMRES with b = 0 return x = b = 0 at once, as cupyx does, whatever the warm start."""
    import edgeto_amd as ET
    mesh, kernel, f, _ = _setup(ET)
    S = ET.CuCG if method == "cg" else ET.CuGMRES
    solver = S(kernel, maxiter=50, tol=1e-8)
    x, _ = solver.solve(f)
    assert torch.count_nonzero(x) > 0
    x0, res0 = solver.solve(torch.zeros_like(f))
    assert res0 == 0.0 and torch.count_nonzero(x0) == 0


def test_no_supports_is_not_a_silent_success():
    """Without supports the stiffness is singular (rigid-body modes) and the load is not in its
    range: the solve raises, or returns with a residual above the tolerance that matches the
    true relative residual -- never a converged-looking answer."""
    import edgeto_amd as ET
    mesh, kernel, f, solver = _setup(ET, precision="fp64", clamp=False, maxiter=60)
    try:
        x, res = solver.solve(f)
    except RuntimeError as e:
        assert any(w in str(e).lower() for w in ("non-finite", "breakdown", "factor", "pivot", "coarse")), str(e)
        return
    assert not (np.isfinite(res) and res < 1e-8)
    if np.isfinite(res) and torch.isfinite(x).all():
        true = float(torch.linalg.norm(f - kernel.dot(x)) / torch.linalg.norm(f))
        assert true > 1e-8


def test_nonfinite_density_raises():
    """A NaN density reaches the operator: the solve reports it (ET_ENONFINITE) instead of
    returning."""
    import edgeto_amd as ET
    mesh, kernel, f, solver = _setup(ET, precision="fp64", maxiter=50)
    rho = torch.full((mesh.n_el,), 0.5, dtype=torch.float64, device="cuda")
    rho[mesh.n_el // 2] = float("nan")
    kernel.set_rho(rho)
    with pytest.raises(RuntimeError):
        solver.solve(f, use_last=False)
