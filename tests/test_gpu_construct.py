"""construct() of the structured stiffness kernel and the matrix products of the kernels
(kernels/_cuda.py:158-170, :213-302): the device CSR against the oracle's assembled operator
(2-D and 3-D grids of several aspect ratios, constraints included), its product against the
matrix-free operator, dot() of a sparse identity, of a random sparse block and of a dense block."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

from oracle import mesh as omesh  # noqa: E402
from oracle import ops as oops  # noqa: E402


def _kernel(dim, nel, size, seed):
    import edgeto_amd as ET
    if dim == 2:
        mesh = ET.CuStructuredMesh2D(*nel, *size)
        els = omesh.elements_2d(*nel)
    else:
        mesh = ET.CuStructuredMesh3D(*nel, *size)
        els = omesh.elements_3d(*nel)
    kernel = ET.CuStructuredStiffnessKernel(mesh)
    assert np.array_equal(kernel.elements.cpu().numpy(), els)
    n = kernel.shape[0]
    rng = np.random.default_rng(seed)
    E = rng.uniform(0.05, 1.0, els.shape[0])
    cons = np.zeros(n, bool)
    cons[rng.choice(n, max(1, n // 7), replace=False)] = True
    kernel.set_rho(torch.tensor(E, device="cuda"))
    kernel.add_constraints(np.nonzero(cons)[0])
    return kernel, mesh, els, E, cons, rng


CASES = [(2, (5, 4), (5.0, 4.0)), (2, (6, 3), (1.0, 0.25)), (3, (4, 3, 3), (4.0, 3.0, 3.0)),
         (3, (3, 2, 4), (1.0, 0.5, 2.0))]


@pytest.mark.parametrize("dim,nel,size", CASES)
def test_structured_construct_matches_oracle(dim, nel, size):
    """The assembled A_hat = M K(E) M + I_c equals the oracle's assembly entry by entry, and
    A_hat u equals the matrix-free product; a second construct() reuses the pattern."""
    print(f"[construct] dim {dim} nel {nel}", flush=True)
    kernel, mesh, els, E, cons, rng = _kernel(dim, nel, size, 3)
    n = kernel.shape[0]
    A = kernel.construct(kernel.rho)
    assert A.layout == torch.sparse_csr and tuple(A.shape) == (n, n)
    Ad = A.to_dense().cpu().numpy()
    Ao = oops.assemble(mesh.K_single_host, els, E, cons, n).toarray()
    assert np.abs(Ad - Ao).max() <= 1e-13 * np.abs(Ao).max()
    u = torch.tensor(rng.standard_normal(n), device="cuda")
    mf = kernel.dot(u).clone()
    y = (A @ u.reshape(-1, 1)).reshape(-1)
    assert float((y - mf).abs().max()) <= 1e-13 * float(mf.abs().max())
    rp = kernel._pattern()[0]
    E2 = torch.tensor(rng.uniform(0.05, 1.0, els.shape[0]), device="cuda")
    A2 = kernel.construct(E2)
    assert kernel._pattern()[0] is rp
    Ao2 = oops.assemble(mesh.K_single_host, els, E2.cpu().numpy(), cons, n).toarray()
    assert np.abs(A2.to_dense().cpu().numpy() - Ao2).max() <= 1e-13 * np.abs(Ao2).max()


def test_structured_matmat_sparse_and_dense():
    """dot() of the sparse identity gives the assembled operator (the reference's construct
    route); dot() of a dense block gives one operator application per column."""
    kernel, mesh, els, E, cons, rng = _kernel(3, (3, 3, 2), (3.0, 3.0, 2.0), 8)
    n = kernel.shape[0]
    A = kernel.construct(kernel.rho).to_dense()
    eye = torch.eye(n, dtype=torch.float64, device="cuda").to_sparse_csr()
    KI = kernel.dot(eye)
    assert KI.layout == torch.sparse_csr
    assert float((KI.to_dense() - A).abs().max()) <= 1e-14 * float(A.abs().max())
    B = torch.tensor(rng.standard_normal((n, 5)), device="cuda")
    KB = kernel @ B
    for j in range(5):
        col = kernel.dot(B[:, j].contiguous())
        assert torch.equal(KB[:, j], col)
    assert kernel._rmatmat(kernel.rho, B).shape == (n, 5)
    with pytest.raises(ValueError):
        kernel.dot(torch.zeros((n + 3, 2), dtype=torch.float64, device="cuda"))


def test_general_kernel_matmat_dense():
    """The unstructured kernels share the same products: K @ B for a dense block against the
    column-wise matrix-free products and the assembled operator."""
    import edgeto_amd as ET
    nx, ny = 4, 3
    nodes = np.array([[i, j] for j in range(ny + 1) for i in range(nx + 1)], float)
    tris = []
    for j in range(ny):
        for i in range(nx):
            a, b = j * (nx + 1) + i, j * (nx + 1) + i + 1
            c, d = b + nx + 1, a + nx + 1
            tris += [[a, b, c], [a, c, d]]
    mesh = ET.CuGeneralMesh(nodes, np.array(tris))
    kernel = ET.CuGeneralStiffnessKernel(mesh)
    rng = np.random.default_rng(2)
    kernel.set_rho(torch.tensor(rng.uniform(0.1, 1.0, len(tris)), device="cuda"))
    kernel.add_constraints(np.arange(4))
    n = kernel.shape[0]
    B = torch.tensor(rng.standard_normal((n, 3)), device="cuda")
    KB = kernel.dot(B)
    A = kernel.construct(kernel.rho).to_dense()
    assert float((KB - A @ B).abs().max()) <= 1e-13 * float(KB.abs().max())
    for j in range(3):
        assert torch.equal(KB[:, j], kernel.dot(B[:, j].contiguous()))


def test_structured_matmat_random_sparse():
    """dot() of a random sparse CSR block (8 columns, ~5 % dense) with the structured kernel: the
    sparse product equals the assembled operator times the dense block and the oracle's product."""
    kernel, mesh, els, E, cons, rng = _kernel(3, (4, 2, 3), (4.0, 2.0, 3.0), 12)
    n = kernel.shape[0]
    Bd = rng.standard_normal((n, 8)) * (rng.random((n, 8)) < 0.05)
    Bd[0, 0] = 1.0  # at least one entry
    B = torch.tensor(Bd, device="cuda").to_sparse_csr()
    KB = kernel.dot(B)
    assert KB.layout == torch.sparse_csr and tuple(KB.shape) == (n, 8)
    ref = kernel.construct(kernel.rho).to_dense() @ torch.tensor(Bd, device="cuda")
    assert float((KB.to_dense() - ref).abs().max()) <= 1e-13 * float(ref.abs().max())
    Ao = oops.assemble(mesh.K_single_host, els, E, cons, n).toarray()
    assert np.abs(KB.to_dense().cpu().numpy() - Ao @ Bd).max() <= 1e-12 * np.abs(Ao @ Bd).max()
