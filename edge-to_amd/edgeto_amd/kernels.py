"""Structured stiffness operator (device).

Mirrors EDGETO/kernels/_cuda.py: StiffnessKernel (:7-59) and
StructuredStiffnessKernel (:62-302) with set_rho, dot / @, diagonal,
process_grad, add_constraints / set_constraints and the reference's
output-buffer reuse (dot, diagonal and process_grad return the same buffer on
every call).  The node->element index arrays the reference keeps for its
CSR-based multigrid (el_ids, sorter, node_ids) are provided lazily for
compatibility; the structured kernels index the grid in closed form.
"""
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


class StructuredStiffnessKernel(StiffnessKernel):
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
        if not self.has_rho:
            raise ValueError("Rho has not been set. dot works only after setting rho.")
        if not isinstance(rhs, torch.Tensor) or not rhs.is_cuda:
            raise NotImplementedError("Only device tensors are supported as vectors.")
        if rhs.shape[0] != self.n_nodes * self.dof:
            raise ValueError("Shape of the input vector does not match the number of nodes and dof.")
        if rhs.dtype != torch.float64:
            raise ValueError("vectors must be float64")
        return self._matvec(self.rho, rhs)

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
