"""General (unstructured) meshes on the GPU against the oracle, through the C ABI and the
reference-named classes: the device-built node -> element star (bit-exact), the operator,
diagonal, element energy / sensitivity and FEA integrals on mixed T3 / Q4 and T4 / H8 meshes
(1e-12), the CSR filter, CG / GMRES solves against a direct solve, the per-element-measure OC
update, and a short SIMP run on a tetrahedral mesh against an oracle loop."""
import ctypes
import os
import sys

import numpy as np
import pytest
import scipy.sparse.linalg as spla

from oracle import physics as ophys
from oracle import topopt as otop
from oracle import unstructured as ou

sys.path.insert(0, os.path.dirname(__file__))
import umesh_cases as cases  # noqa: E402

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

_KEEP = []


@pytest.fixture(autouse=True)
def _release():
    yield
    torch.cuda.synchronize()
    _KEEP.clear()


def _et():
    import edgeto_amd as ET
    from edgeto_amd import _lib as L
    L.load()
    return ET, L


def rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def _oracle_mats(nodes, els):
    mats = [ophys.auto_stiffness(nodes[e], 1.0, 0.33) for e in els]
    return [m[0] for m in mats], [m[1] for m in mats], [m[2] for m in mats], np.array([m[3] for m in mats])


MESHES = {
    "mixed2d": lambda: cases.mixed_2d((6, 5), seed=1),
    "mixed3d": lambda: cases.mixed_3d((4, 3, 3), seed=2),
    "tri": lambda: cases.tri_mesh((5, 4), flip_every=3),
    "tet": lambda: cases.tet_mesh((3, 2, 2), swap_every=5),
}


@pytest.mark.parametrize("name", sorted(MESHES))
def test_star_bit_exact(name):
    ET, L = _et()
    nodes, els = MESHES[name]()
    mesh = ET.CuGeneralMesh(nodes, els)
    sp, st, eo = (t.cpu().numpy() for t in mesh.star())
    flat, ptr = ou.flat_layout(els)
    sorter, node_ptr = ou.star(flat, nodes.shape[0])
    assert np.array_equal(sp, node_ptr)
    assert np.array_equal(st, sorter)
    assert np.array_equal(eo, np.repeat(np.arange(len(els)), np.diff(ptr)))
    k = (ET.CuUniformStiffnessKernel if mesh.is_uniform else ET.CuGeneralStiffnessKernel)(mesh)
    assert np.array_equal(k.sorter.cpu().numpy(), sorter)
    assert np.array_equal(k.node_ids.cpu().numpy(), node_ptr[:-1])


def test_star_rejects_bad_node_ids():
    ET, L = _et()
    ptr = torch.tensor([0, 3], dtype=torch.int32, device="cuda")
    nodes = torch.tensor([0, 1, 7], dtype=torch.int32, device="cuda")
    out = [torch.empty(n, dtype=torch.int32, device="cuda") for n in (4, 3, 3)]
    st = L.lib().et_umesh_star(3, 1, L.i32(ptr), L.i32(nodes), L.i32(out[0]), L.i32(out[1]), L.i32(out[2]),
                               L.stream())
    assert st == L.ET_EINVAL


@pytest.mark.parametrize("name", sorted(MESHES))
def test_operators_match_oracle(name):
    ET, L = _et()
    nodes, els = MESHES[name]()
    mesh = ET.CuGeneralMesh(nodes, els)
    kernel = ET.CuGeneralStiffnessKernel(mesh)
    dim = nodes.shape[1]
    n = nodes.shape[0] * dim
    rng = np.random.default_rng(7)
    E = rng.uniform(0.05, 1.0, len(els))
    cons = np.zeros(n, bool)
    cons[rng.choice(n, n // 8, replace=False)] = True
    u = rng.standard_normal(n)
    Ks, Ds, Bs, As = _oracle_mats(nodes, els)
    kernel.set_rho(torch.tensor(E, device="cuda"))
    kernel.add_constraints(np.nonzero(cons)[0])
    A = ou.assemble(Ks, els, E, cons, n, dim)
    y = kernel.dot(torch.tensor(u, device="cuda"))
    assert rel(y.cpu().numpy(), A @ u) < 1e-12
    assert rel(kernel.diagonal().cpu().numpy(), A.diagonal()) < 1e-12
    dk = kernel.process_grad(torch.tensor(u, device="cuda"))
    assert rel(dk.cpu().numpy(), ou.process_dk(Ks, els, u, dim)) < 1e-12
    xt = rng.uniform(0.0, 1.0, len(els))
    mat = otop.SingleMaterial(void=1e-9, penalty=3.0)
    s = kernel.sensitivity(torch.tensor(u, device="cuda"), torch.tensor(xt, device="cuda"), 3.0, 1e-9)
    assert rel(s.cpu().numpy(), mat.grad(xt) * ou.process_dk(Ks, els, u, dim)) < 1e-12
    # element matrices and measures of the mesh against the oracle's
    assert rel(mesh.As.cpu().numpy().reshape(-1), As) < 1e-14
    assert abs(mesh.volume - As.sum()) <= 1e-12 * As.sum()


@pytest.mark.parametrize("nel", [(5, 4), (3, 3, 2)])
def test_structured_grid_as_general_mesh(nel):
    """A structured grid given as a general mesh: the general operator equals the structured
    library operator (et_matvec_f64) on the same densities, constraints and vector."""
    ET, L = _et()
    dim = len(nel)
    from oracle import mesh as omesh
    els = omesh.elements_2d(*nel) if dim == 2 else omesh.elements_3d(*nel)
    nodes = omesh.nodes_2d(*nel, *nel) if dim == 2 else omesh.nodes_3d(*nel, *nel)
    mesh = ET.CuGeneralMesh(nodes, els, stiffness=ET.AutoStiffness(1.0, 0.3))
    kernel = ET.CuUniformStiffnessKernel(mesh)
    n = nodes.shape[0] * dim
    rng = np.random.default_rng(3)
    E = torch.tensor(rng.uniform(0.1, 1.0, len(els)), device="cuda")
    cons = torch.zeros(n, dtype=torch.bool, device="cuda")
    cons[torch.tensor(rng.choice(n, n // 6, replace=False), device="cuda")] = True
    u = torch.tensor(rng.standard_normal(n), device="cuda")
    kernel.set_rho(E)
    kernel.add_constraints(torch.nonzero(cons).reshape(-1))
    y = kernel.dot(u).clone()
    K0 = torch.tensor(np.ascontiguousarray(mesh.Ks[0].cpu().numpy()), device="cuda")
    ys = torch.empty_like(u)
    _KEEP.extend([K0, E, cons, u])
    L.call("et_matvec_f64", dim, L.nel_arg(nel), L.f64(K0), L.f64(E), L.u8(cons), L.f64(u), L.f64(ys), L.stream())
    assert rel(y.cpu().numpy(), ys.cpu().numpy()) < 1e-12


@pytest.mark.parametrize("name", ["mixed2d", "mixed3d"])
def test_fea_integrals_match_oracle(name):
    ET, L = _et()
    nodes, els = MESHES[name]()
    mesh = ET.CuGeneralMesh(nodes, els)
    dim = nodes.shape[1]
    n = nodes.shape[0] * dim
    rng = np.random.default_rng(9)
    E = rng.uniform(0.1, 1.0, len(els))
    U = rng.standard_normal(n)
    Ks, Ds, Bs, As = _oracle_mats(nodes, els)
    ref = ou.fea(Ks, Bs, Ds, els, E, U, dim)
    kernel = ET.CuGeneralStiffnessKernel(mesh)
    ne, nb = len(els), mesh.B_size
    outs = [torch.empty((ne, nb), dtype=torch.float64, device="cuda") for _ in range(2)] + \
           [torch.empty(ne, dtype=torch.float64, device="cuda") for _ in range(2)]
    Et, Ut = torch.tensor(E, device="cuda"), torch.tensor(U, device="cuda")
    L.call("et_ufea_integrals_f64", ctypes.byref(kernel._um), nb, L.f64(mesh.B_flat), L.i64(mesh.B_ptr),
           L.f64(mesh.D_flat), L.i64(mesh.D_ptr), L.f64(Et), L.f64(Ut), *(L.f64(o) for o in outs), L.stream())
    for got, want in zip(outs, ref):
        assert rel(got.cpu().numpy(), want) < 1e-12


@pytest.mark.parametrize("name,r", [("mixed2d", 1.6), ("mixed3d", 1.8), ("tet", 1.1)])
def test_general_filter(name, r):
    ET, L = _et()
    nodes, els = MESHES[name]()
    mesh = ET.CuGeneralMesh(nodes, els)
    filt = ET.CuGeneralFilter(mesh, r)
    H = ou.general_filter(ou.centroids(nodes, els), r)
    x = np.random.default_rng(4).uniform(0, 1, len(els))
    xt = torch.tensor(x, device="cuda")
    assert rel(filt.dot(xt).cpu().numpy(), H @ x) < 1e-13
    assert rel(filt.T.dot(xt).cpu().numpy(), H.T @ x) < 1e-13
    assert rel((filt @ xt).cpu().numpy(), H @ x) < 1e-13


def _clamped(nodes, dim):
    face = np.nonzero(nodes[:, 0] <= 1e-12)[0]
    cons = np.zeros(nodes.shape[0] * dim, bool)
    for d in range(dim):
        cons[face * dim + d] = True
    return face, cons


@pytest.mark.parametrize("name", sorted(MESHES))
def test_construct_assembled_operator(name):
    """construct(): the device CSR equals the oracle's assembled A_hat (constraints included),
    its columns are sorted per row, and its product equals the matrix-free operator."""
    ET, L = _et()
    nodes, els = MESHES[name]()
    mesh = ET.CuGeneralMesh(nodes, els)
    kernel = ET.CuGeneralStiffnessKernel(mesh)
    dim = nodes.shape[1]
    n = nodes.shape[0] * dim
    rng = np.random.default_rng(11)
    E = rng.uniform(0.05, 1.0, len(els))
    cons = np.zeros(n, bool)
    cons[rng.choice(n, n // 9, replace=False)] = True
    kernel.set_rho(torch.tensor(E, device="cuda"))
    kernel.add_constraints(np.nonzero(cons)[0])
    A = kernel.construct(kernel.rho)
    rp, cols = (t.cpu().numpy() for t in kernel._pattern())
    vals = kernel._csr_vals.cpu().numpy()
    import scipy.sparse as sps
    Ad = sps.csr_matrix((vals, cols, rp), shape=(n, n)).toarray()
    Ks = _oracle_mats(nodes, els)[0]
    Ao = ou.assemble(Ks, els, E, cons, n, dim).toarray()
    assert rel(Ad, Ao) < 1e-12
    assert all(np.all(np.diff(cols[rp[r]:rp[r + 1]]) > 0) for r in range(n))
    u = torch.tensor(rng.standard_normal(n), device="cuda")
    y = torch.empty_like(u)
    L.call("et_csr_spmv_f64", n, L.i32(kernel._pattern()[0]), L.i32(kernel._pattern()[1]), L.f64(kernel._csr_vals),
           L.f64(u), L.f64(y), L.stream())
    assert rel(y.cpu().numpy(), kernel.dot(u).cpu().numpy()) < 1e-12
    assert A.layout == torch.sparse_csr


@pytest.mark.parametrize("name,solver,mf", [("tri", "CuCG", None), ("tet", "CuCG", None), ("tet", "CuCG", True),
                                            ("tet", "CuGMRES", None), ("tri", "CuGMRES", True)])
def test_krylov_on_general_mesh(name, solver, mf):
    """CuCG / CuGMRES with a general stiffness kernel (assembled CSR by default, or matrix-free)
    against a direct solve of the assembled oracle operator."""
    ET, L = _et()
    nodes, els = MESHES[name]()
    dim = nodes.shape[1]
    mesh = ET.CuGeneralMesh(nodes, els)
    kernel = ET.CuGeneralStiffnessKernel(mesh)
    face, cons = _clamped(nodes, dim)
    rng = np.random.default_rng(5)
    E = rng.uniform(0.2, 1.0, len(els))
    kernel.set_rho(torch.tensor(E, device="cuda"))
    kernel.add_constraints(np.nonzero(cons)[0])
    b = rng.standard_normal(nodes.shape[0] * dim)
    Ks = _oracle_mats(nodes, els)[0]
    A = ou.assemble(Ks, els, E, cons, b.size, dim).tocsc()
    xd = spla.spsolve(A, b)
    s = getattr(ET, solver)(kernel, maxiter=5000, tol=1e-11, matrix_free=mf)
    assert s.matrix_free == bool(mf)
    x, res = s.solve(torch.tensor(b, device="cuda"), use_last=False)
    assert res < 1e-10
    assert rel(x.cpu().numpy(), xd) < 1e-8


def test_oc_update_elem():
    """et_oc_update_elem_f64: equal to the scalar update when every measure is the same, and to
    the oracle bisection with per-element measures."""
    ET, L = _et()
    rng = np.random.default_rng(6)
    n = 20000
    x = rng.uniform(0.05, 1.0, n)
    dc = -rng.uniform(0.0, 2.0, n)
    A = rng.uniform(0.5, 1.5, n)
    xt, dct, At = (torch.tensor(v, device="cuda") for v in (x, dc, A))
    work = torch.empty(int(L.lib().et_oc_work_size(n)) // 8 + 1, dtype=torch.float64, device="cuda")
    out = torch.empty_like(xt)
    steps = ctypes.c_int(0)
    vol = float(A.sum())
    L.call("et_oc_update_elem_f64", n, L.f64(xt), L.f64(dct), L.f64(At), vol, 0.4, 0.2, L.f64(out), L.f64(work),
           ctypes.byref(steps), L.stream())
    mat = otop.SingleMaterial(void=1e-9, penalty=3.0)
    ref, ref_steps = otop.oc_update(x, dc, A, vol, 0.4, 0.2, mat)
    assert steps.value == ref_steps
    assert rel(out.cpu().numpy(), ref) < 1e-14
    ones = torch.full_like(xt, 0.75)
    out2 = torch.empty_like(xt)
    L.call("et_oc_update_elem_f64", n, L.f64(xt), L.f64(dct), L.f64(ones), 0.75 * n, 0.4, 0.2, L.f64(out),
           L.f64(work), None, L.stream())
    L.call("et_oc_update_f64", n, L.f64(xt), L.f64(dct), 0.75, 0.75 * n, 0.4, 0.2, L.f64(out2), L.f64(work),
           None, L.stream())
    assert torch.equal(out, out2)


def test_simp_on_tet_mesh_matches_oracle_loop():
    """CuTopOpt on a tetrahedral cantilever with CuCG: compliance history of 6 OC iterations
    against an oracle loop (assembled operator, direct solves, KD-tree filter, per-element
    measures)."""
    ET, L = _et()
    nodes, els = cases.tet_mesh((6, 3, 3))
    dim = 3
    mesh = ET.CuGeneralMesh(nodes, els, stiffness=ET.AutoStiffness(1.0, 0.3))
    kernel = ET.CuGeneralStiffnessKernel(mesh)
    filt = ET.CuGeneralFilter(mesh, 1.2)
    solver = ET.CuCG(kernel, maxiter=20000, tol=1e-12)
    mat = ET.SingleMaterial(void=1e-9, penalty=3.0, volume_fraction=0.4)
    tp = ET.CuTopOpt(mesh, mat, kernel, solver, filt, max_iter=6, ch_tol=0.0, fun_tol=0.0, verbose=False)
    face, cons = _clamped(nodes, dim)
    tp.add_BC_nodal(face, np.array([[1, 1, 1]]))
    load = np.nonzero((nodes[:, 0] == 6.0) & (nodes[:, 1] == 0.0))[0]
    tp.add_F_nodal(load, np.array([[0.0, -1.0 / len(load), 0.0]]))
    rho, flag, hist = tp.optimize(save_comp_history=True, rho_init=np.full(len(els), 0.4))
    # oracle loop
    mats = [ophys.auto_stiffness(nodes[e], 1.0, 0.3) for e in els]
    Ks, As = [m[0] for m in mats], np.array([m[3] for m in mats])
    H = ou.general_filter(ou.centroids(nodes, els), 1.2)
    f = np.zeros(nodes.shape[0] * dim)
    f[load * dim + 1] = -1.0 / len(load)
    omat = otop.SingleMaterial(void=1e-9, penalty=3.0, volume_fraction=0.4)
    x = np.clip(np.full(len(els), 0.4), 1e-9, 1.0)
    comps = []

    def solve(x):
        xt = H @ x
        E = omat.forward(xt)
        A = ou.assemble(Ks, els, E, cons, f.size, dim).tocsc()
        U = spla.spsolve(A, f)
        dc = H.T @ (omat.grad(xt) * ou.process_dk(Ks, els, U, dim))
        return f @ U, dc

    c, dc = solve(x)
    for _ in range(6):
        x, _s = otop.oc_update(x, dc, As, As.sum(), 0.4, 0.2, omat)
        c, dc = solve(x)
        comps.append(c)
    got = np.asarray(hist["comp_history"], dtype=float)
    assert rel(got, np.array(comps)) < 1e-7, (got, comps)
