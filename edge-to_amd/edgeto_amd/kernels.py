"""Structured stiffness operator (device).

Mirrors EDGETO/kernels/_cuda.py: StiffnessKernel (:7-59) and
StructuredStiffnessKernel (:62-302) with set_rho, dot / @, diagonal,
process_grad, add_constraints / set_constraints and the reference's
output-buffer reuse (dot, diagonal and process_grad return the same buffer on
every call).  The node->element index arrays the reference keeps for its
CSR-based multigrid (el_ids, sorter, node_ids) are provided lazily for
compatibility; the structured kernels index the grid in closed form.

construct(rho) assembles the constrained operator as a device CSR matrix (the
reference builds it as K(rho) @ I through _matmat, kernels/_cuda.py:158-170):
the grid is handed to the unstructured assembler (et_umesh_pattern /
et_uassemble_f64) as a mesh whose elements all share the one element matrix.
_matmat / _rmatmat (kernels/_cuda.py:213-276) and dot() of a sparse or 2-D
right-hand side give K(rho) @ B.
"""
import ctypes
import warnings

import numpy as np
import torch

from . import _lib


class StiffnessKernel:
    def __init__(self):
        self.has_rho = False
        self.rho = None
        self.shape = None
        self.matvec = self.dot
        self.rmatvec = self.dot
        self.mat_vec = None
        self.diag = None
        self.dk = None

    def __matmul__(self, rhs):
        return self.dot(rhs)


class CSRAssembly:
    """construct / _matmat / _rmatmat for kernels that can describe themselves as an et_umesh
    (_assembly_mesh()); shared by the structured and the unstructured kernels."""

    def _pattern(self):
        """CSR row offsets and columns of the assembled operator (built once, et_umesh_pattern)."""
        if getattr(self, "_csr", None) is None:
            um = self._assembly_mesh()
            dev = self.constraints.device
            rp = torch.empty(self.shape[0] + 1, dtype=torch.int32, device=dev)
            nnz = ctypes.c_int64(0)
            _lib.call("et_umesh_pattern", ctypes.byref(um), _lib.i32(rp), None, 0, ctypes.byref(nnz), _lib.stream())
            cols = torch.empty(int(nnz.value), dtype=torch.int32, device=dev)
            _lib.call("et_umesh_pattern", ctypes.byref(um), _lib.i32(rp), _lib.i32(cols), int(nnz.value),
                      ctypes.byref(nnz), _lib.stream())
            self._csr = (rp, cols)
        return self._csr

    def construct(self, rho):
        """The assembled operator A_hat = M K(rho) M + I_c (M zeroes the constrained rows and
        columns, I_c puts 1 on their diagonal) as a device CSR tensor (kernels/_cuda.py:158-170,
        :407-420); the sparsity pattern is built on the first call."""
        rp, cols = self._pattern()
        n_el = self._n_elements()
        rho = rho.reshape(-1).contiguous()
        if rho.numel() != n_el:
            raise ValueError("rho must have one entry per element")
        vals = torch.empty(cols.numel(), dtype=torch.float64, device=cols.device)
        _lib.call("et_uassemble_f64", ctypes.byref(self._assembly_mesh()), _lib.f64(rho, n_el), self._cons_ptr(),
                  _lib.i32(rp), _lib.i32(cols), _lib.f64(vals), _lib.stream())
        with warnings.catch_warnings():
            warnings.simplefilter("ignore", UserWarning)  # torch marks sparse CSR tensors as beta
            self.CSR = torch.sparse_csr_tensor(rp, cols, vals, size=self.shape)
        self._csr_vals = vals
        self.ptr = rp
        self.has_been_constructed = True
        return self.CSR

    def _matmat(self, rho, mat, Cp=None):
        """A_hat(rho) @ mat for a device CSR matrix (a sparse product with the assembled
        operator, kernels/_cuda.py:213-273) or a dense 2-D tensor (one operator application per
        column).  Cp (the reference's cached result row offsets) is accepted and unused."""
        if mat.shape[0] != self.shape[0]:
            raise ValueError("Shape of the input matrix does not match the number of nodes and dof.")
        if mat.layout == torch.strided:
            if mat.dim() != 2 or mat.dtype != torch.float64:
                raise ValueError("dense right-hand sides must be 2-D float64 tensors")
            out = torch.empty(mat.shape, dtype=torch.float64, device=mat.device)
            for j in range(mat.shape[1]):
                out[:, j] = self._matvec(rho, mat[:, j].contiguous())
            return out
        A = self.construct(rho)
        B = mat if mat.layout == torch.sparse_csr else mat.to_sparse_csr()
        with warnings.catch_warnings():
            warnings.simplefilter("ignore", UserWarning)
            return torch.sparse.mm(A, B.to(torch.float64))

    def _rmatmat(self, rho, mat, Cp=None):
        return self._matmat(rho, mat, Cp)

    def _dot_checked(self, rhs):
        """dot(): a vector through _matvec, a matrix (sparse, or dense 2-D) through _matmat
        (kernels/_cuda.py:278-302)."""
        if not self.has_rho:
            raise ValueError("Rho has not been set. dot works only after setting rho.")
        if not isinstance(rhs, torch.Tensor) or not rhs.is_cuda:
            raise NotImplementedError("Only device tensors are supported as vectors and matrices.")
        if rhs.shape[0] != self.shape[0]:
            raise ValueError("Shape of the input vector does not match the number of nodes and dof.")
        if rhs.layout != torch.strided or rhs.dim() == 2:
            return self._matmat(self.rho, rhs)
        if rhs.dtype != torch.float64:
            raise ValueError("vectors must be float64")
        return self._matvec(self.rho, rhs)


class StructuredStiffnessKernel(CSRAssembly, StiffnessKernel):
    def __init__(self, mesh):
        super().__init__()
        self.mesh = mesh
        self.nodes, self.elements = mesh.nodes, mesh.elements
        self.dof = mesh.dof
        self.K_single = mesh.K_single
        self.dtype = mesh.dtype
        self.elements_flat = self.elements.reshape(-1)
        self.elements_size = self.elements.shape[1]
        self.n_nodes = self.nodes.shape[0]
        self._dim = len(mesh.nel)
        self._nel = tuple(int(v) for v in mesh.nel)
        self.constraints = torch.zeros(self.n_nodes * self.dof, dtype=torch.bool,
                                       device=self.elements.device)
        self.shape = (self.n_nodes * self.dof, self.n_nodes * self.dof)
        self.has_con = False
        self._star = None

    @property
    def k0_flags(self):
        """et_k0_structure of the element matrix (host check, computed once)."""
        if getattr(self, "_k0_flags", None) is None:
            import ctypes
            K = np.ascontiguousarray(self.mesh.K_single_host, dtype=np.float64)
            f = ctypes.c_int(0)
            _lib.call("et_k0_structure", self._dim, K.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                      ctypes.byref(f))
            self._k0_flags = f.value
        return self._k0_flags

    # --- reference index arrays (kernels/_cuda.py:77-96), built on first use ---
    def _build_star(self):
        flat = self.elements_flat.cpu().numpy()
        el_ids = np.repeat(np.arange(self.elements.shape[0], dtype=np.int32), self.elements_size)
        sorter = np.argsort(flat, kind="stable").astype(np.int32)
        node_ids = np.searchsorted(flat, np.arange(self.n_nodes, dtype=np.int32), sorter=sorter,
                                   side="left").astype(np.int32)
        dev = self.elements.device
        self._star = tuple(torch.from_numpy(a).to(dev) for a in (el_ids, sorter, node_ids))

    @property
    def el_ids(self):
        if self._star is None:
            self._build_star()
        return self._star[0]

    @property
    def sorter(self):
        if self._star is None:
            self._build_star()
        return self._star[1]

    @property
    def node_ids(self):
        if self._star is None:
            self._build_star()
        return self._star[2]

    # --- state ---
    def set_rho(self, rho):
        if not (isinstance(rho, torch.Tensor) and rho.is_cuda and rho.dtype == torch.float64):
            raise ValueError("rho must be a float64 device tensor")
        if rho.numel() != self.elements.shape[0]:
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

    # --- operator ---
    def _cons_ptr(self):
        return _lib.u8(self.constraints, self.shape[0]) if self.has_con else None

    def _matvec(self, rho, vec):
        if self.mat_vec is None:
            self.mat_vec = torch.empty_like(vec)
        vec = vec.contiguous()  # keep the (possibly new) tensor referenced across the call
        n = self.shape[0]
        _lib.call("et_matvec_f64", self._dim, _lib.nel_arg(self._nel), _lib.f64(self.K_single),
                  _lib.f64(rho, self.elements.shape[0]), self._cons_ptr(), _lib.f64(vec, n),
                  _lib.f64(self.mat_vec, n),
                  _lib.stream())
        return self.mat_vec

    def _rmatvec(self, rho, vec):
        return self._matvec(rho, vec)

    def dot(self, rhs):
        return self._dot_checked(rhs)

    def _n_elements(self):
        return self.elements.shape[0]

    def _assembly_mesh(self):
        """The grid as an et_umesh: elements of elements_size nodes, all pointing at the one
        element matrix K_single (K_ptr = 0); its device arrays stay referenced by the kernel."""
        if getattr(self, "_asm_um", None) is None:
            ne, npe = self.elements.shape
            if (ne + 1) * npe >= 2 ** 31:
                raise ValueError("construct: the grid has too many element nodes for 32-bit offsets")
            dev = self.elements.device
            ep = torch.arange(0, (ne + 1) * npe, npe, dtype=torch.int32, device=dev)
            en = self.elements_flat.to(torch.int32).contiguous()
            K = self.K_single.to(torch.float64).contiguous()
            Kp = torch.zeros(ne, dtype=torch.int64, device=dev)
            sp = torch.empty(self.n_nodes + 1, dtype=torch.int32, device=dev)
            st = torch.empty(en.numel(), dtype=torch.int32, device=dev)
            eo = torch.empty_like(st)
            _lib.call("et_umesh_star", self.n_nodes, ne, _lib.i32(ep), _lib.i32(en), _lib.i32(sp), _lib.i32(st),
                      _lib.i32(eo), _lib.stream())
            self._um_arrays = (ep, en, K, Kp, sp, st, eo)
            self._asm_um = _lib.UMesh(self.n_nodes, ne, self.dof, ep.data_ptr(), en.data_ptr(), K.data_ptr(),
                                      Kp.data_ptr(), sp.data_ptr(), st.data_ptr(), eo.data_ptr())
        return self._asm_um

    def diagonal(self, rho=None):
        if rho is None and not self.has_rho:
            raise ValueError("Rho has not been set. diagonal works only after setting rho or if rho is provided.")
        r = self.rho if rho is None else rho.contiguous()
        if self.diag is None:
            self.diag = torch.empty(self.shape[0], dtype=torch.float64, device=self.constraints.device)
        _lib.call("et_diagonal_f64", self._dim, _lib.nel_arg(self._nel), _lib.f64(self.K_single),
                  _lib.f64(r, self.elements.shape[0]), self._cons_ptr(), _lib.f64(self.diag, self.shape[0]),
                  _lib.stream())
        return self.diag

    def process_grad(self, U):
        if self.dk is None:
            self.dk = torch.empty(self.elements.shape[0], dtype=torch.float64, device=U.device)
        U = U.contiguous()
        _lib.call("et_elem_energy_f64", self._dim, _lib.nel_arg(self._nel), _lib.f64(self.K_single),
                  _lib.f64(U, self.shape[0]), _lib.f64(self.dk, self.elements.shape[0]), _lib.stream())
        return self.dk

    def sensitivity(self, U, xt, penal, void, out=None):
        """p * clip(xt, void, 1)^(p-1) * (-u_e^T K0 u_e) in one kernel."""
        out = torch.empty(self.elements.shape[0], dtype=torch.float64, device=U.device) if out is None else out
        U, xt = U.contiguous(), xt.contiguous()
        ne = self.elements.shape[0]
        _lib.call("et_sensitivity_f64", self._dim, _lib.nel_arg(self._nel), _lib.f64(self.K_single),
                  _lib.f64(U, self.shape[0]), _lib.f64(xt, ne), float(penal), float(void),
                  _lib.f64(out, ne), _lib.stream())
        return out


CuStructuredStiffnessKernel = StructuredStiffnessKernel
