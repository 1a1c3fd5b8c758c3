"""General (unstructured) meshes, their stiffness operators and density filter.

Mirrors EDGETO/geom/_mesh.py:146-253 (GeneralMesh / CuGeneralMesh),
kernels/_cuda.py:305-815 (UniformStiffnessKernel / GeneralStiffnessKernel,
exported as CuUniformStiffnessKernel / CuGeneralStiffnessKernel) and
geom/_filters.py:256-298 (GeneralFilter / CuGeneralFilter).

Element matrices are computed on the host once per mesh (a batch per element
type for the default auto_stiffness, one call per element for any other
stiffness= callable, as the reference does); the node -> element star is built
on the device (et_umesh_star); every operator runs in the library
(et_umatvec_f64, et_udiagonal_f64, et_uelem_energy_f64, et_usensitivity_f64,
et_ufea_integrals_f64), and the filter is a CSR product (et_csr_spmv_f64) with
its transpose stored as a second CSR matrix.

construct() assembles the operator as a CSR matrix on the device (pattern once,
values per call: et_umesh_pattern / et_uassemble_f64); CuCG / CuGMRES use it by
default for these kernels, as the reference does (matrix_free=False).

Deviation: redundant nodes are removed and the element lists renumbered
consistently (the reference renumbers elements_flat only and then evaluates the
element matrices with the old numbering, geom/_mesh.py:172-190).
"""
import ctypes

import numpy as np
import torch

from . import _lib
from .geom import FilterKernel, Mesh, _device
from .kernels import CSRAssembly, StiffnessKernel
from .physics import AutoStiffness


def _as_lists(elements):
    if isinstance(elements, np.ndarray) and elements.ndim == 2:
        return [row for row in elements.astype(np.int64)]
    if isinstance(elements, torch.Tensor):
        return [row for row in elements.cpu().numpy().astype(np.int64)]
    return [np.asarray(e, dtype=np.int64) for e in elements]


class GeneralMesh(Mesh):
    """Mesh of arbitrary T3 / Q4 (2-D) or T4 / H8 (3-D) elements (geom/_mesh.py:146-229).

    nodes: (n_nodes, dim) coordinates; elements: node lists (a 2-D array for a single
    element type, or a list of lists for mixed types).  stiffness: callable x0s -> (K, D, B,
    A) per element, or an AutoStiffness (batched).  Attributes follow the reference:
    nodes, elements, centeroids, is_uniform, elements_flat, elements_ptr, element_sizes,
    Ks / Ds / Bs (uniform) or K_flat / D_flat / B_flat with K_ptr / D_ptr / B_ptr (mixed),
    As (n_el, 1), volume, dof, dtype."""

    def __init__(self, nodes, elements, dtype=np.float64, stiffness=None):
        if dtype not in (np.float64, torch.float64):
            raise ValueError("general meshes run in float64")
        stiffness = AutoStiffness(1.0, 0.33) if stiffness is None else stiffness
        nodes = np.asarray(nodes.cpu().numpy() if isinstance(nodes, torch.Tensor) else nodes, dtype=np.float64)
        els = _as_lists(elements)
        if not els:
            raise ValueError("a mesh needs at least one element")
        flat = np.concatenate(els)
        used = np.unique(flat)
        if used.size != nodes.shape[0] or used[-1] != nodes.shape[0] - 1:
            remap = -np.ones(max(nodes.shape[0], int(used[-1]) + 1), dtype=np.int64)
            remap[used] = np.arange(used.size)
            els = [remap[e] for e in els]
            nodes = nodes[used]
        self.dim = nodes.shape[1]
        self.element_sizes = np.array([e.size for e in els], dtype=np.int32)
        self.is_uniform = bool(np.all(self.element_sizes == self.element_sizes[0]))
        self.elements_flat_host = np.concatenate(els).astype(np.int32)
        self.elements_ptr_host = np.concatenate([[0], np.cumsum(self.element_sizes)]).astype(np.int32)
        self.centeroids = (np.add.reduceat(nodes[self.elements_flat_host], self.elements_ptr_host[:-1], axis=0)
                           / self.element_sizes[:, None])
        self.n_el = len(els)
        self.n_nodes = nodes.shape[0]
        # element matrices: one batch per element size where the stiffness has a batched form
        Ks, Ds, Bs, As = [None] * self.n_el, [None] * self.n_el, [None] * self.n_el, np.empty(self.n_el)
        if hasattr(stiffness, "batch"):
            for size in np.unique(self.element_sizes):
                ids = np.nonzero(self.element_sizes == size)[0]
                x = np.stack([nodes[els[i]] for i in ids])
                K, D, B, A = stiffness.batch(x)
                for j, i in enumerate(ids):
                    Ks[i], Ds[i], Bs[i], As[i] = K[j], D[j], B[j], A[j]
        else:
            for i, e in enumerate(els):
                K, D, B, A = stiffness(nodes[e])
                Ks[i], Ds[i], Bs[i], As[i] = (np.asarray(K, dtype=np.float64), np.asarray(D, dtype=np.float64),
                                              np.asarray(B, dtype=np.float64), float(A))
        self.dof = int(Ks[-1].shape[0] // self.element_sizes[-1])
        if any(K.shape[0] != self.dof * s for K, s in zip(Ks, self.element_sizes)):
            raise ValueError("element matrices do not match the element sizes and one dof per node")
        self._Ks_host, self._Ds_host, self._Bs_host = Ks, Ds, Bs
        self.As_host = As
        self.volume = float(As.sum())
        self.dtype = np.float64
        dev = _device()
        self.nodes = torch.tensor(nodes, device=dev)
        self.elements = (torch.tensor(np.stack(els).astype(np.int32), device=dev) if self.is_uniform
                         else [torch.tensor(e.astype(np.int32), device=dev) for e in els])
        self.elements_flat = torch.tensor(self.elements_flat_host, device=dev)
        self.elements_ptr = torch.tensor(self.elements_ptr_host, device=dev)
        self.As = torch.tensor(As.reshape(-1, 1), device=dev)
        self.A_elem = self.As.reshape(-1).contiguous()
        self.A_host = None

        def cat(mats):
            sizes = np.array([m.size for m in mats], dtype=np.int64)
            ptr = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
            return torch.tensor(np.concatenate([m.reshape(-1) for m in mats]), device=dev), torch.tensor(ptr, device=dev)

        self.K_flat, self.K_ptr = cat(Ks)
        self.D_flat, self.D_ptr = cat(Ds)
        self.B_flat, self.B_ptr = cat(Bs)
        self.B_size = int(Bs[0].shape[0])
        if self.is_uniform:
            self.Ks = self.K_flat.reshape(self.n_el, *Ks[0].shape)
            self.Ds = self.D_flat.reshape(self.n_el, *Ds[0].shape)
            self.Bs = self.B_flat.reshape(self.n_el, *Bs[0].shape)
        self._star = None

    def star(self):
        """(star_ptr, star, el_of) device arrays (built once, et_umesh_star)."""
        if self._star is None:
            dev = self.elements_flat.device
            sp = torch.empty(self.n_nodes + 1, dtype=torch.int32, device=dev)
            st = torch.empty(self.elements_flat.numel(), dtype=torch.int32, device=dev)
            eo = torch.empty_like(st)
            _lib.call("et_umesh_star", self.n_nodes, self.n_el, _lib.i32(self.elements_ptr),
                      _lib.i32(self.elements_flat), _lib.i32(sp), _lib.i32(st), _lib.i32(eo), _lib.stream())
            self._star = (sp, st, eo)
        return self._star

    def umesh(self):
        """The et_umesh descriptor of this mesh (its arrays stay referenced by the mesh)."""
        sp, st, eo = self.star()
        return _lib.UMesh(self.n_nodes, self.n_el, self.dof, self.elements_ptr.data_ptr(),
                          self.elements_flat.data_ptr(), self.K_flat.data_ptr(), self.K_ptr.data_ptr(),
                          sp.data_ptr(), st.data_ptr(), eo.data_ptr())


CuGeneralMesh = GeneralMesh


class _UnstructuredKernel(CSRAssembly, StiffnessKernel):
    """Shared part of the Uniform / General stiffness kernels (kernels/_cuda.py:305-815)."""

    def __init__(self, mesh):
        super().__init__()
        self.mesh = mesh
        self.nodes, self.elements = mesh.nodes, mesh.elements
        self.dof = mesh.dof
        self.dtype = mesh.dtype
        self.elements_flat = mesh.elements_flat
        self.elements_size = int(mesh.element_sizes.max())
        self.n_nodes = mesh.n_nodes
        self.n_el = mesh.n_el
        dev = mesh.elements_flat.device
        self.constraints = torch.zeros(self.n_nodes * self.dof, dtype=torch.bool, device=dev)
        self.shape = (self.n_nodes * self.dof, self.n_nodes * self.dof)
        self.has_con = False
        self._um = mesh.umesh()

    # reference index arrays (kernels/_cuda.py:320-331): sorter = stable argsort of
    # elements_flat, el_ids = element of each flat position, node_ids = start of each node's run
    @property
    def sorter(self):
        return self.mesh.star()[1]

    @property
    def el_ids(self):
        return self.mesh.star()[2]

    @property
    def node_ids(self):
        return self.mesh.star()[0][:-1]

    def set_rho(self, rho):
        if not (isinstance(rho, torch.Tensor) and rho.is_cuda and rho.dtype == torch.float64):
            raise ValueError("rho must be a float64 device tensor")
        if rho.numel() != self.n_el:
            raise ValueError("rho must have one entry per element")
        self.rho = rho.reshape(-1).contiguous()
        self.has_rho = True

    def set_constraints(self, constraints):
        self.constraints[:] = False
        self.add_constraints(constraints)

    def add_constraints(self, constraints):
        idx = torch.as_tensor(np.asarray(constraints, dtype=np.int64) if not isinstance(constraints, torch.Tensor)
                              else constraints, device=self.constraints.device).long().reshape(-1)
        if idx.numel():
            self.constraints[idx] = True
        self.has_con = True

    def reset(self):
        self.has_rho = False
        self.rho = None
        self.constraints[:] = False
        self.has_con = False
        self.mat_vec = self.diag = self.dk = None

    def _cons_ptr(self):
        return _lib.u8(self.constraints, self.shape[0]) if self.has_con else None

    def _matvec(self, rho, vec):
        if self.mat_vec is None:
            self.mat_vec = torch.empty_like(vec)
        vec = vec.contiguous()
        _lib.call("et_umatvec_f64", ctypes.byref(self._um), _lib.f64(rho, self.n_el), self._cons_ptr(),
                  _lib.f64(vec, self.shape[0]), _lib.f64(self.mat_vec, self.shape[0]), _lib.stream())
        return self.mat_vec

    def _rmatvec(self, rho, vec):
        return self._matvec(rho, vec)

    def dot(self, rhs):
        return self._dot_checked(rhs)

    def _n_elements(self):
        return self.n_el

    def _assembly_mesh(self):
        return self._um

    def diagonal(self, rho=None):
        if rho is None and not self.has_rho:
            raise ValueError("Rho has not been set. diagonal works only after setting rho or if rho is provided.")
        r = self.rho if rho is None else rho.contiguous()
        if self.diag is None:
            self.diag = torch.empty(self.shape[0], dtype=torch.float64, device=self.constraints.device)
        _lib.call("et_udiagonal_f64", ctypes.byref(self._um), _lib.f64(r, self.n_el), self._cons_ptr(),
                  _lib.f64(self.diag, self.shape[0]), _lib.stream())
        return self.diag

    def process_grad(self, U):
        if self.dk is None:
            self.dk = torch.empty(self.n_el, dtype=torch.float64, device=U.device)
        U = U.contiguous()
        _lib.call("et_uelem_energy_f64", ctypes.byref(self._um), _lib.f64(U, self.shape[0]),
                  _lib.f64(self.dk, self.n_el), _lib.stream())
        return self.dk

    def sensitivity(self, U, xt, penal, void, out=None):
        """p * clip(xt, void, 1)^(p-1) * (-u_e^T K_e u_e) in one kernel."""
        out = torch.empty(self.n_el, dtype=torch.float64, device=U.device) if out is None else out
        U, xt = U.contiguous(), xt.contiguous()
        _lib.call("et_usensitivity_f64", ctypes.byref(self._um), _lib.f64(U, self.shape[0]),
                  _lib.f64(xt, self.n_el), float(penal), float(void), _lib.f64(out, self.n_el), _lib.stream())
        return out

    def __matmul__(self, rhs):
        return self.dot(rhs)


class UniformStiffnessKernel(_UnstructuredKernel):
    """Elements of one type with their own matrices Ks (kernels/_cuda.py:305-552)."""

    def __init__(self, mesh):
        if not mesh.is_uniform:
            raise ValueError("Mesh is not uniform, you should use GeneralStiffnessKernel instead.")
        super().__init__(mesh)
        self.Ks = mesh.Ks


class GeneralStiffnessKernel(_UnstructuredKernel):
    """Mixed element types, matrices in K_flat at K_ptr (kernels/_cuda.py:554-815)."""

    def __init__(self, mesh):
        super().__init__(mesh)
        self.K_flat, self.K_ptr = mesh.K_flat, mesh.K_ptr
        self.elements_ptr = mesh.elements_ptr


CuUniformStiffnessKernel = UniformStiffnessKernel
CuGeneralStiffnessKernel = GeneralStiffnessKernel


class GeneralFilter(FilterKernel):
    """Density filter of a general mesh (geom/_filters.py:256-298): row-normalised hat
    weights over the element centroids within r_min, from a KD-tree ball query on the host
    (core/_filter.py:126-182), applied as a CSR product on the device; the transpose is
    stored as its own CSR matrix."""

    def __init__(self, mesh, r_min):
        super().__init__()
        from scipy.spatial import cKDTree
        import scipy.sparse as sps
        self.dtype = mesh.dtype
        self.nd = mesh.dim
        c = np.asarray(mesh.centeroids, dtype=np.float64)
        tree = cKDTree(c)
        nb = tree.query_ball_point(c, r_min)
        lens = np.array([len(v) for v in nb], dtype=np.int64)
        cols = np.concatenate([np.asarray(v, dtype=np.int64) for v in nb])
        rows = np.repeat(np.arange(c.shape[0]), lens)
        w = (r_min - np.linalg.norm(c[cols] - c[rows], axis=-1)) / r_min
        sums = np.bincount(rows, weights=w, minlength=c.shape[0])
        w = w / sums[rows]
        H = sps.coo_matrix((w, (rows, cols)), shape=(c.shape[0], c.shape[0])).tocsr()
        self.kernel = H
        self.shape = H.shape
        dev = _device()

        def up(M):
            M = M.tocsr()
            M.sort_indices()
            return (torch.tensor(M.indptr.astype(np.int32), device=dev), torch.tensor(M.indices.astype(np.int32), device=dev),
                    torch.tensor(M.data.astype(np.float64), device=dev))

        self._H = up(H)
        self._HT = up(H.T)
        self.weights = torch.empty(self.shape[0], dtype=torch.float64, device=dev)

    def _apply(self, M, rho, out=None):
        rho = rho.contiguous()
        n = self.shape[0]
        out = torch.empty(n, dtype=torch.float64, device=rho.device) if out is None else out
        _lib.call("et_csr_spmv_f64", n, _lib.i32(M[0]), _lib.i32(M[1]), _lib.f64(M[2]), _lib.f64(rho, n),
                  _lib.f64(out, n), _lib.stream())
        return out

    def _matvec(self, rho, out=None):
        return self._apply(self._H, rho, out)

    def _rmatvec(self, rho, out=None):
        return self._apply(self._HT, rho, out)


CuGeneralFilter = GeneralFilter
