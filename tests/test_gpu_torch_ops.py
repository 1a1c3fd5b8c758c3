"""torch.ops.edgeto_amd.* against the class methods that call the same entry points."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def test_torch_ops_match_classes():
    import edgeto_amd as ET
    nel = (10, 6, 5)
    mesh = ET.CuStructuredMesh3D(*nel, *nel, stiffness=lambda x0: ET.auto_stiffness(x0, 1.0, 0.3))
    kernel = ET.CuStructuredStiffnessKernel(mesh)
    filt = ET.CuStructuredFilter3D(mesh, 2.0)
    rng = np.random.default_rng(0)
    E = torch.tensor(rng.uniform(0.1, 1.0, mesh.n_el), device="cuda")
    kernel.set_rho(E)
    kernel.add_constraints(np.arange(30))
    u = torch.tensor(rng.standard_normal(kernel.shape[0]), device="cuda")
    y1 = kernel.dot(u).clone()
    y2 = torch.ops.edgeto_amd.matvec(mesh.K_single, E, kernel.constraints, u, list(nel))
    assert torch.equal(y1, y2)
    x = torch.tensor(rng.uniform(0, 1, mesh.n_el), device="cuda")
    assert torch.equal(filt.dot(x), torch.ops.edgeto_amd.filter_fwd(filt.offsets, filt.weights, x, list(nel)))
    assert torch.equal(filt._rmatvec(x),
                       torch.ops.edgeto_amd.filter_adj(filt.offsets, filt.weights, filt.normalizer, x, list(nel)))
    a1 = filt.simp_adjoint(kernel, u, x, 3.0, 1e-9)
    a2 = torch.ops.edgeto_amd.simp_adjoint(mesh.K_single, u, x, 3.0, 1e-9, filt.offsets, filt.weights,
                                           filt.normalizer, list(nel), int(kernel.k0_flags))
    assert torch.equal(a1, a2)
    dc = -torch.tensor(rng.uniform(0, 1, mesh.n_el), device="cuda")
    tp = ET.CuTopOpt(mesh, ET.SingleMaterial(void=1e-9, penalty=3.0, volume_fraction=0.3), kernel, None, filt,
                     verbose=False)
    o1 = tp.update_design(x, dc)
    o2 = torch.ops.edgeto_amd.oc_update(x, dc, mesh.A_host, mesh.volume, 0.3, 0.2)
    assert torch.equal(o1, o2)
