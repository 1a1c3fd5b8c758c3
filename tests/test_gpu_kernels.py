"""GPU parity of every native kernel against the CPU oracle, through the C ABI.

Tolerances: integer and index work bit-exact; FP64 operators 1e-12 relative
(max-norm, relative to the max-norm of the reference output).
"""
import ctypes

import numpy as np
import pytest

from oracle import filters as ofilt
from oracle import mesh as omesh
from oracle import mgm as omgm
from oracle import ops as oops
from oracle import physics as ophys
from oracle import topopt as otop

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

MESHES = [(7, 5), (12, 8), (5, 4, 3), (8, 6, 4), (6, 4, 4)]


def _lib():
    from edgeto_amd import _lib as L
    L.load()
    return L


_KEEP = []


@pytest.fixture(autouse=True)
def _release_inputs():
    """Device inputs stay referenced until the test ends: a pointer handed to the
    C ABI must outlive the call (torch's allocator reuses freed blocks at once)."""
    yield
    torch.cuda.synchronize()
    _KEEP.clear()


def dev(a, dtype=None):
    t = torch.from_numpy(np.ascontiguousarray(a))
    if dtype is not None:
        t = t.to(dtype)
    t = t.cuda()
    _KEEP.append(t)
    return t


def rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _problem(nel, seed=0, nu=0.3, frac_cons=0.15):
    rng = np.random.default_rng(seed)
    dim = len(nel)
    el = omesh.elements_3d(*nel) if dim == 3 else omesh.elements_2d(*nel)
    nodes = omesh.nodes_3d(*nel, *nel) if dim == 3 else omesh.nodes_2d(*nel, *nel)
    K0, _, _, _ = ophys.auto_stiffness(nodes[el[0]], 1.0, nu)
    n_dof = nodes.shape[0] * dim
    E = rng.uniform(1e-3, 1.0, el.shape[0])
    u = rng.standard_normal(n_dof)
    cons = rng.random(n_dof) < frac_cons
    return dict(nel=nel, dim=dim, el=el, nodes=nodes, K0=K0, E=E, u=u, cons=cons, n_dof=n_dof)


@pytest.mark.parametrize("nel", MESHES)
def test_connectivity_bit_exact(nel):
    L = _lib()
    dim = len(nel)
    ref = omesh.elements_3d(*nel) if dim == 3 else omesh.elements_2d(*nel)
    out = torch.empty(ref.shape, dtype=torch.int32, device="cuda")
    L.call("et_generate_elements", dim, L.nel_arg(nel), L.ptr(out), L.stream())
    assert np.array_equal(out.cpu().numpy(), ref)


@pytest.mark.parametrize("nel", MESHES)
@pytest.mark.parametrize("with_cons", [False, True])
def test_matvec_f64(nel, with_cons):
    L = _lib()
    p = _problem(nel, seed=1)
    cons = p["cons"] if with_cons else None
    ref = oops.matvec(p["K0"], p["el"], p["E"], p["u"], cons)
    y = torch.empty(p["n_dof"], dtype=torch.float64, device="cuda")
    ct = dev(cons.astype(np.uint8)) if with_cons else None
    L.call("et_matvec_f64", p["dim"], L.nel_arg(nel), L.ptr(dev(p["K0"])), L.ptr(dev(p["E"])),
           L.ptr(ct), L.ptr(dev(p["u"])), L.ptr(y), L.stream())
    assert rel(y.cpu().numpy(), ref) < 1e-12


@pytest.mark.parametrize("nel", [(12, 8), (6, 4, 4)])
def test_matvec_f32(nel):
    L = _lib()
    p = _problem(nel, seed=2)
    ref = oops.matvec(p["K0"], p["el"], p["E"], p["u"], p["cons"])
    y = torch.empty(p["n_dof"], dtype=torch.float32, device="cuda")
    L.call("et_matvec_f32", p["dim"], L.nel_arg(nel), L.ptr(dev(p["K0"], torch.float32)),
           L.ptr(dev(p["E"], torch.float32)), L.ptr(dev(p["cons"].astype(np.uint8))),
           L.ptr(dev(p["u"], torch.float32)), L.ptr(y), L.stream())
    assert rel(y.cpu().numpy(), ref) < 1e-5


@pytest.mark.parametrize("nel", MESHES)
def test_diagonal_energy_sensitivity(nel):
    L = _lib()
    p = _problem(nel, seed=3)
    ct = dev(p["cons"].astype(np.uint8))
    d = torch.empty(p["n_dof"], dtype=torch.float64, device="cuda")
    L.call("et_diagonal_f64", p["dim"], L.nel_arg(nel), L.ptr(dev(p["K0"])), L.ptr(dev(p["E"])),
           L.ptr(ct), L.ptr(d), L.stream())
    assert rel(d.cpu().numpy(), oops.diagonal(p["K0"], p["el"], p["E"], p["cons"], p["n_dof"])) < 1e-12
    dk = torch.empty(p["el"].shape[0], dtype=torch.float64, device="cuda")
    L.call("et_elem_energy_f64", p["dim"], L.nel_arg(nel), L.ptr(dev(p["K0"])), L.ptr(dev(p["u"])),
           L.ptr(dk), L.stream())
    ref_dk = oops.process_dk(p["K0"], p["el"], p["u"])
    assert rel(dk.cpu().numpy(), ref_dk) < 1e-12
    xt = np.random.default_rng(4).uniform(0.0, 1.0, p["el"].shape[0])
    xt[:3] = [0.0, 1e-12, 1.0]
    s = torch.empty_like(dk)
    L.call("et_sensitivity_f64", p["dim"], L.nel_arg(nel), L.ptr(dev(p["K0"])), L.ptr(dev(p["u"])),
           L.ptr(dev(xt)), 3.0, 1e-9, L.ptr(s), L.stream())
    mat = otop.SingleMaterial(void=1e-9, penalty=3.0)
    assert rel(s.cpu().numpy(), mat.grad(xt) * ref_dk) < 1e-12
    E = torch.empty_like(dk)
    L.call("et_simp_forward_f64", xt.shape[0], L.ptr(dev(xt)), 3.0, 1e-9, 0, L.ptr(E), L.stream())
    assert rel(E.cpu().numpy(), mat.forward(xt)) < 1e-15


FILTER_CASES = [((12, 8), 1.5, (1, 1)), ((20, 7), 2.6, (1, 1)), ((300, 270), 3.0, (1, 1)),
                ((9, 7, 5), 2.0, (1, 1, 1)), ((8, 6, 7), 2.9, (1, 1, 1)),
                ((10, 6, 4), 2.2, (1.0, 0.5, 2.0)), ((40, 9, 70), 2.0, (1, 1, 1)),
                ((12, 10, 9), 4.5, (1, 1, 1)), ((33, 20), 5.2, (1, 1))]


@pytest.mark.parametrize("nel,rmin,spacing", FILTER_CASES)
def test_filter(nel, rmin, spacing):
    """Normaliser, forward and transpose against the oracle; the last two cases have taps
    beyond the LDS ring (radius above 3 elements), which take the direct-gather path."""
    L = _lib()
    dim = len(nel)
    offs, w = ofilt.offsets_weights(rmin, spacing)
    norm = ofilt.normalizer(nel, offs, w)
    n = int(np.prod(nel))
    x = np.random.default_rng(5).uniform(0, 1, n)
    o_t, w_t = dev(offs.astype(np.int32)), dev(w)
    nt = torch.empty(n, dtype=torch.float64, device="cuda")
    L.call("et_filter_normalizer_f64", dim, L.nel_arg(nel), len(w), L.ptr(o_t), L.ptr(w_t),
           L.ptr(nt), L.stream())
    assert rel(nt.cpu().numpy(), norm) < 1e-14
    y = torch.empty(n, dtype=torch.float64, device="cuda")
    L.call("et_filter_fwd_f64", dim, L.nel_arg(nel), len(w), L.ptr(o_t), L.ptr(w_t), L.ptr(dev(x)),
           L.ptr(y), L.stream())
    assert rel(y.cpu().numpy(), ofilt.forward(x, nel, offs, w)) < 1e-12
    L.call("et_filter_adj_f64", dim, L.nel_arg(nel), len(w), L.ptr(o_t), L.ptr(w_t), L.ptr(nt),
           L.ptr(dev(x)), L.ptr(y), L.stream())
    assert rel(y.cpu().numpy(), ofilt.transpose(x, nel, offs, w, norm)) < 1e-12


@pytest.mark.parametrize("nel,rmin,spacing", [((70, 9, 66), 2.0, (1, 1, 1)), ((17, 11, 130), 2.6, (1, 1, 1)),
                                               ((10, 6, 4), 2.2, (1.0, 0.5, 2.0)), ((300, 41), 1.5, (1, 1))])
def test_filter_tap_orders_and_tiles(nel, rmin, spacing):
    """The stencils sum an offset table in ascending (dx, dy, dz) order over the unrolled
    cube of radius <= 2, and any other order with the per-tap loop; both agree with the
    oracle, and 4- and 8-row tiles give the same bits."""
    L = _lib()
    dim = len(nel)
    offs, w = ofilt.offsets_weights(rmin, spacing)
    norm = ofilt.normalizer(nel, offs, w)
    n = int(np.prod(nel))
    x = dev(np.random.default_rng(12).uniform(-1, 1, n))
    nt = dev(norm)
    perm = np.random.default_rng(13).permutation(len(w))
    ref_f = ofilt.forward(x.cpu().numpy(), nel, offs, w)
    ref_t = ofilt.transpose(x.cpu().numpy(), nel, offs, w, norm)
    outs = {}
    try:
        for rows in (4, 8):
            L.call("et_tune_set", 17, rows)
            for order, (o, ww) in (("asc", (offs, w)), ("perm", (offs[perm], w[perm]))):
                o_t, w_t = dev(np.ascontiguousarray(o).astype(np.int32)), dev(np.ascontiguousarray(ww))
                yf = torch.empty(n, dtype=torch.float64, device="cuda")
                yt = torch.empty(n, dtype=torch.float64, device="cuda")
                L.call("et_filter_fwd_f64", dim, L.nel_arg(nel), len(ww), L.ptr(o_t), L.ptr(w_t), L.ptr(x),
                       L.ptr(yf), L.stream())
                L.call("et_filter_adj_f64", dim, L.nel_arg(nel), len(ww), L.ptr(o_t), L.ptr(w_t), L.ptr(nt),
                       L.ptr(x), L.ptr(yt), L.stream())
                assert rel(yf.cpu().numpy(), ref_f) < 1e-12, (rows, order)
                assert rel(yt.cpu().numpy(), ref_t) < 1e-12, (rows, order)
                outs[(rows, order)] = (yf, yt)
    finally:
        L.call("et_tune_set", 17, 8)
    for order in ("asc", "perm"):
        assert torch.equal(outs[(4, order)][0], outs[(8, order)][0])
        assert torch.equal(outs[(4, order)][1], outs[(8, order)][1])


@pytest.mark.parametrize("nel", [(8, 4), (12, 6), (8, 4, 4), (6, 4, 8), (7, 5, 3)])
def test_transfers(nel):
    L = _lib()
    dim = len(nel)
    dof = dim
    nf = int(np.prod(np.array(nel) + 1)) * dof
    ncn = int(np.prod(np.array(nel) // 2 + 1)) * dof
    rng = np.random.default_rng(6)
    vf, vc = rng.standard_normal(nf), rng.standard_normal(ncn)
    out_c = torch.empty(ncn, dtype=torch.float64, device="cuda")
    L.call("et_restrict_f64", dim, L.nel_arg(nel), dof, L.ptr(dev(vf)), L.ptr(out_c), L.stream())
    assert rel(out_c.cpu().numpy(), omgm.restrict(vf, nel, dof)) < 1e-14
    out_f = torch.empty(nf, dtype=torch.float64, device="cuda")
    L.call("et_prolong_f64", dim, L.nel_arg(nel), dof, L.ptr(dev(vc)), L.ptr(out_f), L.stream())
    assert rel(out_f.cpu().numpy(), omgm.prolong(vc, nel, dof)) < 1e-14


def _stencil_to_dense(S, nel, dof):
    """Expand the library's block-stencil layout into a dense matrix (test helper)."""
    dim = len(nel)
    nn = np.array(nel) + 1
    N = int(np.prod(nn))
    NO = 27 if dim == 3 else 9
    S = S.reshape(NO, dof, dof, N)
    A = np.zeros((N * dof, N * dof))
    for n in range(N):
        if dim == 3:
            nz, nx = nn[2], nn[0]
            k, t = n % nz, n // nz
            i, j = t % nx, t // nx
        else:
            nx = nn[0]
            i, j, k = n % nx, n // nx, 0
        for o in range(NO):
            if dim == 3:
                oi, oj, ok = o // 9 - 1, (o // 3) % 3 - 1, o % 3 - 1
            else:
                oi, oj, ok = o // 3 - 1, o % 3 - 1, 0
            mi, mj, mk = i + oi, j + oj, k + ok
            if not (0 <= mi < nn[0] and 0 <= mj < nn[1]) or (dim == 3 and not 0 <= mk < nn[2]):
                assert np.all(S[o, :, :, n] == 0)
                continue
            m = (mj * nn[0] + mi) * nn[2] + mk if dim == 3 else mj * nn[0] + mi
            A[n * dof:(n + 1) * dof, m * dof:(m + 1) * dof] = S[o, :, :, n]
    return A


@pytest.mark.parametrize("nel", [(8, 4), (6, 4, 4), (4, 4, 2)])
def test_stencil_and_galerkin(nel):
    L = _lib()
    p = _problem(nel, seed=7)
    dim, dof = p["dim"], p["dim"]
    NO = 27 if dim == 3 else 9
    N = int(np.prod(np.array(nel) + 1))
    S = torch.empty(NO * dof * dof * N, dtype=torch.float64, device="cuda")
    L.call("et_stencil_from_operator_f64", dim, L.nel_arg(nel), L.ptr(dev(p["K0"])), L.ptr(dev(p["E"])),
           L.ptr(dev(p["cons"].astype(np.uint8))), L.ptr(S), L.stream())
    A = oops.assemble(p["K0"], p["el"], p["E"], p["cons"], p["n_dof"]).toarray()
    A_gpu = _stencil_to_dense(S.cpu().numpy(), nel, dof)
    assert rel(A_gpu, A) < 1e-13
    # stencil application equals the matrix-free operator
    y = torch.empty(p["n_dof"], dtype=torch.float64, device="cuda")
    L.call("et_stencil_apply_f64", dim, L.nel_arg(nel), dof, L.ptr(S), L.ptr(dev(p["u"])), L.ptr(y),
           L.stream())
    assert rel(y.cpu().numpy(), A @ p["u"]) < 1e-12
    # Galerkin
    nc = tuple(v // 2 for v in nel)
    Nc = int(np.prod(np.array(nc) + 1))
    Sc = torch.empty(NO * dof * dof * Nc, dtype=torch.float64, device="cuda")
    L.call("et_galerkin_f64", dim, L.nel_arg(nel), dof, L.ptr(S), L.ptr(Sc), L.stream())
    R = omgm.restriction_matrix(nel, dof).toarray()
    assert rel(_stencil_to_dense(Sc.cpu().numpy(), nc, dof), R @ A @ R.T) < 1e-12


def test_dot_and_sqdist():
    L = _lib()
    rng = np.random.default_rng(8)
    for n in (1, 255, 4097, 1_000_003):
        a, b = rng.standard_normal(n), rng.standard_normal(n)
        out = torch.zeros(1, dtype=torch.float64, device="cuda")
        L.call("et_dot_f64", n, L.ptr(dev(a)), L.ptr(dev(b)), L.ptr(out), L.stream())
        assert abs(out.item() - a @ b) <= 1e-12 * np.abs(a * b).sum()
        L.call("et_sqdist_f64", n, L.ptr(dev(a)), L.ptr(dev(b)), L.ptr(out), L.stream())
        assert abs(out.item() - np.sum((a - b) ** 2)) <= 1e-12 * np.sum((a - b) ** 2)


@pytest.mark.parametrize("n,seed", [(4800, 0), (100_000, 1), (7, 2), (4_194_304, 3)])
@pytest.mark.parametrize("scale", [1.0, 1e-20, 1e20])
def test_oc_update(n, seed, scale):
    """Device OC bisection against the oracle's: the same number of steps and the design at
    the last midpoint to 1e-14.  scale moves the multiplier to the bracket's ends (1e-20: the
    longest bisection the stopping rule allows; 1e20: the shortest)."""
    L = _lib()
    rng = np.random.default_rng(seed)
    x = rng.uniform(0.01, 1.0, n)
    dc = -rng.uniform(0.0, 2.0, n) * rng.uniform(0, 1, n) ** 3 * scale
    mat = otop.SingleMaterial(void=1e-9, penalty=3.0, volume_fraction=0.4)
    ref, steps = otop.oc_update(x, dc, 1.0, float(n), 0.4, 0.2, mat)
    work = torch.empty(int(L.lib().et_oc_work_size(n)) // 8 + 1, dtype=torch.float64, device="cuda")
    out = torch.empty(n, dtype=torch.float64, device="cuda")
    st = ctypes.c_int(0)
    L.call("et_oc_update_f64", n, L.ptr(dev(x)), L.ptr(dev(dc)), 1.0, float(n), 0.4, 0.2, L.ptr(out),
           L.ptr(work), ctypes.byref(st), L.stream())
    assert st.value == steps
    assert rel(out.cpu().numpy(), ref) < 1e-14
    # asynchronous form: no host step count; the state at the start of work holds it
    out2 = torch.empty_like(out)
    L.call("et_oc_update_f64", n, L.ptr(dev(x)), L.ptr(dev(dc)), 1.0, float(n), 0.4, 0.2, L.ptr(out2),
           L.ptr(work), None, L.stream())
    assert torch.equal(out, out2)
    assert int(work[:5].view(torch.int32)[9].item()) == steps
    assert int(work[:5].view(torch.int32)[8].item()) == 1  # done
    # all-zero sensitivities: ocP falls back to 1e-3
    ref0, steps0 = otop.oc_update(x, np.zeros(n), 1.0, float(n), 0.4, 0.2, mat)
    L.call("et_oc_update_f64", n, L.ptr(dev(x)), L.ptr(dev(np.zeros(n))), 1.0, float(n), 0.4, 0.2,
           L.ptr(out), L.ptr(work), ctypes.byref(st), L.stream())
    assert st.value == steps0
    assert rel(out.cpu().numpy(), ref0) < 1e-14


@pytest.mark.parametrize("nel", [(8, 4), (12, 6), (6, 4, 4), (8, 4, 6), (4, 2, 2)])
def test_galerkin_from_elements(nel):
    """Element-based first coarse operator equals R A_hat R^T (with constraints)."""
    L = _lib()
    p = _problem(nel, seed=11, frac_cons=0.1)
    dim, dof = p["dim"], p["dim"]
    # clamp one face as well so whole elements carry constraints
    nn = np.array(nel) + 1
    cons = p["cons"].copy()
    if dim == 3:
        j, k = np.meshgrid(np.arange(nn[1]), np.arange(nn[2]), indexing="ij")
        face = (j * nn[0] * nn[2] + k).ravel()
    else:
        face = np.arange(nn[1]) * nn[0]
    for d in range(dof):
        cons[face * dof + d] = True
    NO = 27 if dim == 3 else 9
    nc = tuple(v // 2 for v in nel)
    Nc = int(np.prod(np.array(nc) + 1))
    Sc = torch.empty(NO * dof * dof * Nc, dtype=torch.float64, device="cuda")
    K0 = np.ascontiguousarray(p["K0"])
    L.call("et_galerkin_elements_f64", dim, L.nel_arg(nel), K0.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
           L.ptr(dev(p["E"])), L.ptr(dev(cons.astype(np.uint8))), L.ptr(Sc), L.stream())
    A = oops.assemble(p["K0"], p["el"], p["E"], cons, p["n_dof"])
    R = omgm.restriction_matrix(nel, dof)
    ref = (R @ A @ R.T).toarray()
    assert rel(_stencil_to_dense(Sc.cpu().numpy(), nc, dof), ref) < 1e-12


@pytest.mark.parametrize("nel,scale", [((6, 4, 4), 1.0), ((8, 6, 20), 1e-7), ((5, 3, 33), 3e4), ((12, 5, 17), 1.0)])
def test_matvec_f32_mfma(nel, scale):
    """Matrix-core fine operator (FP16 hi/lo split, per-element scaling) at FP32-level accuracy,
    including tiny and large vector magnitudes and z extents that are not multiples of 16."""
    L = _lib()
    p = _problem(nel, seed=12)
    u = p["u"] * scale
    ref = oops.matvec(p["K0"], p["el"], p["E"], u, p["cons"])
    y = torch.empty(p["n_dof"], dtype=torch.float32, device="cuda")
    K0 = np.ascontiguousarray(p["K0"])
    L.call("et_matvec_f32_mfma", 3, L.nel_arg(nel), K0.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
           L.ptr(dev(p["K0"], torch.float32)), L.ptr(dev(p["E"], torch.float32)),
           L.ptr(dev(p["cons"].astype(np.uint8))), L.ptr(dev(u, torch.float32)), L.ptr(y), L.stream())
    assert rel(y.cpu().numpy(), ref) < 2e-5


MX_MESHES = [(6, 4, 4), (8, 6, 20), (5, 3, 33), (12, 5, 16), (20, 3, 128), (9, 1, 16), (100, 60, 8)]


WHT_MESHES = MX_MESHES + [(7, 4, 64), (5, 3, 65), (30, 4, 100), (4, 2, 1)]

# fine-operator variants by knob values
# (ET_TUNE_WHT, ET_TUNE_MX_F32, ET_TUNE_MX_F64, ET_TUNE_WHT_THREADS_F32, ET_TUNE_WHT_THREADS_F64,
#  ET_TUNE_WHT_SLICES_F32, ET_TUNE_WHT_SLICES_F64)
_KN = (7, 3, 4, 10, 11, 15, 16)
_FINE = {"tiled": (0, 0, 0, 1024, 512, 16, 8), "matrix": (0, 1, 1, 1024, 512, 16, 8),
         "wht": (1, 0, 0, 1024, 512, 16, 8), "wht_alt": (1, 0, 0, 512, 1024, 16, 16),
         "wht_half": (1, 0, 0, 512, 256, 8, 8)}


def _matvec_variant(L, nel, prec, variant, seed):
    old = [L.lib().et_tune_get(k) for k in _KN]
    p = _problem(nel, seed=seed)
    ref = oops.matvec(p["K0"], p["el"], p["E"], p["u"], p["cons"])
    dt = torch.float32 if prec == "f32" else torch.float64
    y = torch.empty(p["n_dof"], dtype=dt, device="cuda")
    for k, v in zip(_KN, _FINE[variant]):
        L.call("et_tune_set", k, v)
    try:
        L.call("et_matvec_" + prec, 3, L.nel_arg(nel), L.ptr(dev(p["K0"], dt)), L.ptr(dev(p["E"], dt)),
               L.ptr(dev(p["cons"].astype(np.uint8))), L.ptr(dev(p["u"], dt)), L.ptr(y), L.stream())
        torch.cuda.synchronize()
    finally:
        for k, v in zip(_KN, old):
            L.call("et_tune_set", k, v)
    return y.cpu().numpy(), ref


@pytest.mark.parametrize("nel", MX_MESHES)
@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_matvec_matrix_core(nel, prec):
    """Matrix-core fine operator at working precision (FP32 / FP64 MFMA): z extents that are
    and are not multiples of 16, the 128-element z limit, a single element layer, and a mesh
    whose (x tile, plane) units outnumber the persistent workgroups."""
    y, ref = _matvec_variant(_lib(), nel, prec, "matrix", 21)
    assert rel(y, ref) < (1e-5 if prec == "f32" else 1e-12)


@pytest.mark.parametrize("nel", WHT_MESHES)
@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("variant", ["wht", "wht_alt", "wht_half"])
def test_matvec_transformed_basis(nel, prec, variant):
    """Walsh-Hadamard-basis fine operator (every workgroup size): both tile widths (nelz <= 64
    and <= 128), the 64 and 128 seams, a single z layer, and more (x tile, plane) units than
    persistent workgroups."""
    y, ref = _matvec_variant(_lib(), nel, prec, variant, 22)
    assert rel(y, ref) < (1e-5 if prec == "f32" else 1e-12)


SWEEP_MESHES = [(6, 4, 4), (7, 5), (20, 3, 128), (48, 20, 100), (40, 30, 40)]


@pytest.mark.parametrize("nel", SWEEP_MESHES)
@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("variant", ["tiled", "matrix", "wht", "wht_alt", "wht_half"])
def test_residual_and_jacobi(nel, prec, variant):
    """Fused residual and damped-Jacobi modes of every fine-grid operator variant against the
    oracle (r = b - A x, y = x + w D^-1 r), on meshes with many y planes per workgroup; the
    transformed-basis sweeps on 16- and 8-slice tiles with every workgroup size."""
    L = _lib()
    dim = len(nel)
    old = [L.lib().et_tune_get(k) for k in _KN]
    p = _problem(nel, seed=23)
    rng = np.random.default_rng(5)
    b = rng.standard_normal(p["n_dof"])
    A = oops.assemble(p["K0"], p["el"], p["E"], p["cons"], p["n_dof"])
    Dinv = 1.0 / A.diagonal()
    omega = 0.6
    r_ref = b - A @ p["u"]
    y_ref = p["u"] + omega * Dinv * r_ref
    dt = torch.float32 if prec == "f32" else torch.float64
    tol = 2e-5 if prec == "f32" else 1e-12
    args = (L.ptr(dev(p["K0"], dt)), L.ptr(dev(p["E"], dt)), L.ptr(dev(p["cons"].astype(np.uint8))))
    x = dev(p["u"], dt)
    bt = dev(b, dt)
    r = torch.empty(p["n_dof"], dtype=dt, device="cuda")
    y = torch.empty(p["n_dof"], dtype=dt, device="cuda")
    for k, v in zip(_KN, _FINE[variant]):
        L.call("et_tune_set", k, v)
    try:
        L.call("et_residual_" + prec, dim, L.nel_arg(nel), *args, L.ptr(x), L.ptr(bt), L.ptr(r), L.stream())
        L.call("et_jacobi_" + prec, dim, L.nel_arg(nel), *args, L.ptr(x), L.ptr(bt),
               L.ptr(dev(Dinv, dt)), omega, L.ptr(y), L.stream())
        torch.cuda.synchronize()
    finally:
        for k, v in zip(_KN, old):
            L.call("et_tune_set", k, v)
    assert rel(r.cpu().numpy(), r_ref) < tol
    assert rel(y.cpu().numpy(), y_ref) < tol


def test_dtype_checked_before_launch():
    """Every FP64 wrapper rejects a float32 tensor (ValueError before any launch: the kernels
    would read and write 8 bytes per entry of a 4-byte buffer); compliance() casts its input
    and returns the gradient in the input's dtype."""
    import edgeto_amd as ET
    nel = (8, 4, 4)
    mesh = ET.CuStructuredMesh3D(*nel, *nel, stiffness=lambda x0: ET.auto_stiffness(x0, 1.0, 0.3))
    kernel = ET.CuStructuredStiffnessKernel(mesh)
    filt = ET.CuStructuredFilter3D(mesh, 1.5)
    mat = ET.SingleMaterial(void=1e-9, penalty=3.0, volume_fraction=0.3)
    x32 = torch.full((mesh.n_el,), 0.5, dtype=torch.float32, device="cuda")
    with pytest.raises(ValueError):
        filt.dot(x32)
    with pytest.raises(ValueError):
        filt._rmatvec(x32)
    with pytest.raises(ValueError):
        mat(x32, 0)
    with pytest.raises(ValueError):
        kernel.sensitivity(torch.zeros(kernel.shape[0], dtype=torch.float64, device="cuda"), x32, 3.0, 1e-9)
    with pytest.raises(ValueError):  # wrong length
        filt.dot(torch.zeros(mesh.n_el + 1, dtype=torch.float64, device="cuda"))
    solver = ET.CuMultiGrid(mesh, kernel, tol=1e-12, n_smooth=2, n_level=1)
    tp = ET.CuTopOpt(mesh, mat, kernel, solver, filt, verbose=False)
    nx1, ny1, nz1 = nel[0] + 1, nel[1] + 1, nel[2] + 1
    j, k = np.meshgrid(np.arange(ny1), np.arange(nz1), indexing="ij")
    tp.add_BC_nodal((j * nx1 * nz1 + k).ravel(), np.array([[1, 1, 1]]))
    tp.add_F_nodal(np.array([nel[0] * nz1]), np.array([[0.0, -1.0, 0.0]]))
    x = x32.clone().requires_grad_(True)
    c = ET.compliance(x, tp)
    c.backward()
    assert x.grad.dtype == torch.float32 and torch.isfinite(x.grad).all()
    x64 = x32.double().requires_grad_(True)
    c64 = ET.compliance(x64, tp)
    c64.backward()
    # the second solve starts from the first one's solution: equal to the solve tolerance
    assert abs(float(c) - float(c64)) <= 1e-9 * abs(float(c64))
    assert torch.allclose(x.grad.double(), x64.grad, rtol=1e-6, atol=0)


@pytest.mark.parametrize("nel", MESHES + [(64, 32, 32)])
def test_fea_integrals(nel):
    """et_fea_integrals_f64 (strain, stress, energy, von Mises per element) against the oracle
    on a random displacement and density, on unit and non-cubic elements."""
    L = _lib()
    p = _problem(nel, seed=13)
    dim = p["dim"]
    size = [n * s for n, s in zip(nel, (1.0, 0.6, 1.3)[:dim])]
    nodes = omesh.nodes_3d(*nel, *size) if dim == 3 else omesh.nodes_2d(*nel, *size)
    K0, D, B, _ = ophys.auto_stiffness(nodes[p["el"][0]], 1.0, 0.3)
    ne, nb = p["el"].shape[0], B.shape[0]
    out = [torch.empty((ne, nb), dtype=torch.float64, device="cuda") for _ in range(2)]
    out += [torch.empty(ne, dtype=torch.float64, device="cuda") for _ in range(2)]
    L.call("et_fea_integrals_f64", dim, L.nel_arg(nel), L.f64(dev(K0)), L.f64(dev(D)), L.f64(dev(B)),
           L.f64(dev(p["E"])), L.f64(dev(p["u"])), *[L.f64(t) for t in out], L.stream())
    ref = oops.fea_integrals(K0, D, B, p["el"], p["E"], p["u"])
    for name, got, want in zip(("strain", "stress", "energy", "von_mises"), out, ref):
        assert rel(got.cpu().numpy(), want) < 1e-12, name


def test_topopt_fea_integrals():
    """TopOpt.FEA_integrals: solve at E = clip(rho, void, 1), then the element pass, against a
    direct solve and the oracle's element post-processing."""
    import scipy.sparse.linalg as spla
    import edgeto_amd as ET
    nel = (16, 8, 8)
    nx, ny, nz = nel
    mesh = ET.CuStructuredMesh3D(nx, ny, nz, nx, ny, nz, stiffness=lambda x0: ET.auto_stiffness(x0, 1.0, 0.3))
    kernel = ET.CuStructuredStiffnessKernel(mesh)
    mat = ET.SingleMaterial(void=1e-9, penalty=3.0, volume_fraction=0.3)
    solver = ET.CuMultiGrid(mesh, kernel, tol=1e-11, n_smooth=2, n_level=2)
    tp = ET.CuTopOpt(mesh, mat, kernel, solver, ET.CuStructuredFilter3D(mesh, 2.0), verbose=False)
    nx1, ny1, nz1 = nx + 1, ny + 1, nz + 1
    j, k = np.meshgrid(np.arange(ny1), np.arange(nz1), indexing="ij")
    tp.add_BC_nodal((j * nx1 * nz1 + k).ravel(), np.array([[1, 1, 1]]))
    tp.add_F_nodal(np.array([nx * nz1 + kk for kk in range(nz1)]), np.array([[0.0, -1.0 / nz1, 0.0]]))
    rho = np.random.default_rng(4).uniform(0.0, 1.0, mesh.n_el)
    strain, stress, energy, vm = tp.FEA_integrals(rho)
    E = np.clip(rho, 1e-9, 1.0)
    el = omesh.elements_3d(*nel)
    cons = kernel.constraints.cpu().numpy()
    f = tp.f.cpu().numpy()
    A = oops.assemble(mesh.K_single_host, el, E, cons, kernel.shape[0])
    U = spla.spsolve(A.tocsc(), f)
    ref = oops.fea_integrals(mesh.K_single_host, mesh.D_single.cpu().numpy(), mesh.B_single.cpu().numpy(),
                             el, E, U)
    for name, got, want in zip(("strain", "stress", "energy", "von_mises"), (strain, stress, energy, vm), ref):
        assert rel(got.cpu().numpy(), want) < 1e-8, name
    # total strain energy is half the compliance f.u
    assert abs(energy.sum().item() - 0.5 * f @ U) <= 1e-8 * abs(f @ U)


@pytest.mark.parametrize("nel,rmin,spacing", FILTER_CASES[:2] + FILTER_CASES[3:8])
@pytest.mark.parametrize("k0", ["iso", "rotated"])
def test_simp_adjoint_fused(nel, rmin, spacing, k0):
    """et_simp_adjoint_f64 = H^T [p clip(xt)^(p-1) (-u^T K0 u)] in one kernel, against the
    oracle chain filt_T(grad * process_dk); 'rotated' is an element matrix without the brick's
    reflection structure (3-D: the full-product energy instead of the transformed basis)."""
    L = _lib()
    dim = len(nel)
    p = _problem(nel, seed=17)
    K0 = p["K0"]
    if k0 == "rotated":
        if dim == 2:
            pytest.skip("2-D always evaluates the full 8 x 8 product")
        # congruence with a random rotation of each node's displacement frame: symmetric,
        # positive semi-definite, no reflection structure
        rng = np.random.default_rng(3)
        Q, _ = np.linalg.qr(rng.standard_normal((3, 3)))
        T = np.kron(np.eye(8), Q)
        K0 = T.T @ K0 @ T
    offs, w = ofilt.offsets_weights(rmin, spacing)
    norm = ofilt.normalizer(nel, offs, w)
    n = int(np.prod(nel))
    xt = np.random.default_rng(8).uniform(0.0, 1.0, n)
    flags = ctypes.c_int(0)
    K0c = np.ascontiguousarray(K0)
    L.call("et_k0_structure", dim, K0c.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), ctypes.byref(flags))
    mat = otop.SingleMaterial(void=1e-9, penalty=3.0)
    ref = ofilt.transpose(mat.grad(xt) * oops.process_dk(K0, p["el"], p["u"]), nel, offs, w, norm)
    # the library's own structure flags, and 0 (the full-product energy for every matrix)
    for fl in sorted({flags.value, 0}):
        out = torch.empty(n, dtype=torch.float64, device="cuda")
        L.call("et_simp_adjoint_f64", dim, L.nel_arg(nel), L.f64(dev(K0)), L.f64(dev(p["u"])), L.f64(dev(xt)), 3.0,
               1e-9, len(w), L.i32(dev(offs.astype(np.int32))), L.f64(dev(w)), L.f64(dev(norm)), fl, L.f64(out),
               L.stream())
        err = rel(out.cpu().numpy(), ref)
        assert err < 1e-12, (fl, err)
    if k0 == "rotated":
        assert not flags.value & 4  # no transformed-basis structure: the full product ran
