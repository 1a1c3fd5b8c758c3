"""SIMP minimum-compliance optimiser (device).

Mirrors EDGETO/TopOpt/_cuda.py (TopOpt / CuTopOpt, :17-382): constructor
arguments, BC and load setup (add_BC_nodal / add_F_nodal :200-240, add_BCs /
add_Forces :144-198 via a host KD-tree), update_design (:242-280),
solve_and_process (:282-302) and optimize (:304-382).  Deviations, on
purpose: rho_init works (the reference reads rho before assignment at
:321-322; the intended clip(x0, void, 1) is used), and the OC bisection runs on
the device with one host synchronisation per chunk of steps.
"""
import ctypes
import logging
import time

import numpy as np
import torch

from . import _lib
from .MaterialModels import PenalizedMultiMaterial

logger = logging.getLogger(__name__)


def _n_el(kernel):
    """Element count of a structured or general stiffness kernel."""
    n = getattr(kernel, "n_el", None)
    return int(n) if n is not None else int(kernel.elements.shape[0])


class TopOpt:
    def __init__(self, mesh, material_model, kernel, solver, filter_kernel=None, max_iter=500,
                 move=0.2, ch_tol=1e-4, fun_tol=1e-6, reuse_sol=True, solve_min_tol=np.inf,
                 min_tol_patience=10, abandon_patience=10, verbose=True):
        self.mesh = mesh
        self.material_model = material_model
        self.kernel = kernel
        self.solver = solver
        self.filter_kernel = filter_kernel
        self.max_iter = max_iter
        self.move = move
        self.ch_tol = ch_tol
        self.fun_tol = fun_tol
        self.solve_min_tol = solve_min_tol
        self.min_tol_patience = min_tol_patience
        self.abandon_patience = abandon_patience
        self.reuse_sol = reuse_sol
        self.verbose = verbose
        self.dtype = mesh.dtype
        if verbose:
            logger.setLevel(logging.INFO)
        dev = kernel.constraints.device
        self.f = torch.zeros(kernel.shape[0], dtype=torch.float64, device=dev)
        self.KDTree = None
        n = _n_el(kernel)
        self._oc_work = torch.empty(int(_lib.lib().et_oc_work_size(n)) // 8 + 1, dtype=torch.float64,
                                    device=dev)
        self._scalar = torch.zeros(4, dtype=torch.float64, device=dev)
        self.iteration_log = []

    # ------------------------------------------------------------- BCs / loads
    def reset_BC(self):
        self.kernel.set_constraints([])
        self.kernel.has_con = False

    def reset_F(self):
        self.f[:] = 0

    def _tree(self):
        if self.KDTree is None:
            from scipy.spatial import KDTree
            self.KDTree = KDTree(self.mesh.nodes.cpu().numpy())
        return self.KDTree

    def add_BCs(self, positions, BCs):
        _, idx = self._tree().query(np.asarray(positions))
        BCs = np.asarray(BCs)
        if BCs.shape[0] != idx.shape[0] and BCs.shape[0] > 1:
            raise ValueError("BCs shape does not match positions shape and is not broadcastable.")
        self.add_BC_nodal(np.asarray(idx), BCs)

    def add_Forces(self, positions, Fs):
        _, idx = self._tree().query(np.asarray(positions))
        Fs = np.asarray(Fs, dtype=np.float64)
        if Fs.shape[0] != idx.shape[0] and Fs.shape[0] > 1:
            raise ValueError("Fs shape does not match positions shape and is not broadcastable.")
        self._set_loads(np.asarray(idx).reshape(-1), Fs)

    def _set_loads(self, ids, Fs):
        """f[node * dof + i] = Fs[., i]; a node listed more than once keeps its last row (NumPy
        assignment order, made explicit: a device scatter with repeated indices has none)."""
        dof = self.mesh.dof
        for i in range(dof):
            dofs = ids * dof + i
            vals = np.broadcast_to(Fs[0, i] if Fs.shape[0] == 1 else Fs[:, i], dofs.shape)
            _, first_rev = np.unique(dofs[::-1], return_index=True)
            keep = len(dofs) - 1 - first_rev
            self.f[torch.as_tensor(dofs[keep], device=self.f.device)] = torch.as_tensor(
                np.ascontiguousarray(vals[keep]), dtype=torch.float64, device=self.f.device)

    def add_BC_nodal(self, node_ids, BCs):
        node_ids = np.asarray(node_ids).reshape(-1)
        BCs = np.asarray(BCs).astype(bool)
        if BCs.shape[0] != node_ids.shape[0] and BCs.shape[0] > 1:
            raise ValueError("BCs shape does not match positions shape and is not broadcastable.")
        dof = self.mesh.dof
        for i in range(dof):
            if BCs.shape[0] == 1:
                if BCs[0, i]:
                    self.kernel.add_constraints(node_ids * dof + i)
            else:
                self.kernel.add_constraints(node_ids[BCs[:, i]] * dof + i)

    def add_F_nodal(self, node_ids, Fs):
        node_ids = np.asarray(node_ids)
        Fs = np.asarray(Fs, dtype=np.float64)
        if Fs.shape[0] != node_ids.shape[0] and Fs.shape[0] > 1:
            raise ValueError("Fs shape does not match positions shape and is not broadcastable.")
        self._set_loads(node_ids.reshape(-1), Fs)

    # ------------------------------------------------------- post-processing
    def FEA_integrals(self, rho):
        """Strain, stress, strain energy and von Mises stress of every element for the design
        rho (TopOpt/_cuda.py:78-129): E = clip(rho, void, 1) (the material model's plain map),
        a warm-started solve, then one fused element pass (et_fea_integrals_f64).  Returns
        (strain (n_el, NB), stress (n_el, NB), strain_energy (n_el,), von_mises (n_el,))."""
        rho = torch.as_tensor(rho, dtype=torch.float64, device=self.f.device).reshape(-1).contiguous()
        rho_ = self.material_model(rho, 1e6, plain=True)
        self.kernel.set_rho(rho_)
        U, residual = self.solver.solve(self.f, use_last=True)
        if residual > 1e-5:
            logger.warning(f"Solver residual is above 1e-5 ({residual:.4e}). Consider higher iterations "
                           "(rerun this function and more iteration from prior solve will be applied).")
        mesh = self.mesh
        ne = rho_.numel()
        dev = self.f.device
        if hasattr(self.kernel, "_um"):
            # general meshes: per-element B, D and K (FEA_Integrals_..._{full,flat}, :1722-1800)
            nb = mesh.B_size
            strain = torch.empty((ne, nb), dtype=torch.float64, device=dev)
            stress = torch.empty((ne, nb), dtype=torch.float64, device=dev)
            energy = torch.empty(ne, dtype=torch.float64, device=dev)
            vm = torch.empty(ne, dtype=torch.float64, device=dev)
            _lib.call("et_ufea_integrals_f64", ctypes.byref(self.kernel._um), nb, _lib.f64(mesh.B_flat),
                      _lib.i64(mesh.B_ptr), _lib.f64(mesh.D_flat), _lib.i64(mesh.D_ptr), _lib.f64(rho_, ne),
                      _lib.f64(U, self.kernel.shape[0]), _lib.f64(strain), _lib.f64(stress), _lib.f64(energy),
                      _lib.f64(vm), _lib.stream())
            return strain, stress, energy, vm
        nb = mesh.B_single.shape[0]
        strain = torch.empty((ne, nb), dtype=torch.float64, device=dev)
        stress = torch.empty((ne, nb), dtype=torch.float64, device=dev)
        energy = torch.empty(ne, dtype=torch.float64, device=dev)
        vm = torch.empty(ne, dtype=torch.float64, device=dev)
        B = mesh.B_single.contiguous()
        D = mesh.D_single.contiguous()
        if tuple(D.shape) != (nb, nb) or B.shape[1] != mesh.K_single.shape[0]:
            raise ValueError("D_single / B_single do not match the element matrix")
        _lib.call("et_fea_integrals_f64", len(mesh.nel), _lib.nel_arg(mesh.nel), _lib.f64(mesh.K_single),
                  _lib.f64(D), _lib.f64(B), _lib.f64(rho_, ne), _lib.f64(U, self.kernel.shape[0]),
                  _lib.f64(strain), _lib.f64(stress), _lib.f64(energy), _lib.f64(vm), _lib.stream())
        return strain, stress, energy, vm

    # --------------------------------------------------------------- OC / loop
    def _multi(self):
        """Multi-material updates (one OC multiplier per material): any PenalizedMultiMaterial,
        one material included, as the reference's class methods do for m = 1."""
        mm = self.material_model
        return mm.base_properties()["n_sets"] > 1 or isinstance(mm, PenalizedMultiMaterial)

    def _update_design_multi(self, rho, df):
        """One multiplier per material (n_sets = m), TopOpt/_cuda.py:242-280."""
        mm = self.material_model
        m = mm.n_sets
        if getattr(self.mesh, "A_host", None) is None:
            raise NotImplementedError("multi-material design updates need a mesh with one element measure "
                                      "(structured meshes)")
        rho = rho.reshape(-1, m).contiguous()
        df = df.reshape(-1, m).contiguous()
        n = rho.shape[0]
        if getattr(self, "_oc_multi_work", None) is None:
            size = int(_lib.lib().et_oc_multi_work_size(n, m))
            self._oc_multi_work = torch.empty(size // 8 + 1, dtype=torch.float64, device=rho.device)
        out = torch.empty_like(rho)
        mass = np.ascontiguousarray(np.asarray(mm.mass, dtype=np.float64).reshape(-1))
        rhom = None if mm.rho_mat is None else np.ascontiguousarray(np.asarray(mm.rho_mat, dtype=np.float64).reshape(-1))
        steps = ctypes.c_int(0)
        dp = ctypes.POINTER(ctypes.c_double)
        _lib.call("et_oc_update_multi_f64", n, m, _lib.f64(rho), _lib.f64(df, n * m), float(self.mesh.A_host),
                  float(self.mesh.volume), float(self.move), mass.ctypes.data_as(dp),
                  None if rhom is None else rhom.ctypes.data_as(dp), _lib.f64(out, n * m),
                  _lib.f64(self._oc_multi_work), ctypes.byref(steps), _lib.stream())
        self._multi_steps = steps.value
        return out

    def update_design(self, rho, df):
        if self._multi():
            return self._update_design_multi(rho, df)
        rho = rho.reshape(-1).contiguous()
        df = df.reshape(-1).contiguous()
        out = torch.empty_like(rho)
        n = rho.numel()
        # asynchronous: the step count stays in the work buffer (int32 at byte 36)
        A_elem = getattr(self.mesh, "A_elem", None)
        if A_elem is not None:
            # general meshes: one measure per element (mesh.As)
            _lib.call("et_oc_update_elem_f64", n, _lib.f64(rho), _lib.f64(df, n), _lib.f64(A_elem, n),
                      float(self.mesh.volume), float(self.material_model.volume_fraction), float(self.move),
                      _lib.f64(out, n), _lib.f64(self._oc_work), None, _lib.stream())
            return out
        _lib.call("et_oc_update_f64", n, _lib.f64(rho), _lib.f64(df, n), float(self.mesh.A_host),
                  float(self.mesh.volume), float(self.material_model.volume_fraction), float(self.move),
                  _lib.f64(out, n), _lib.f64(self._oc_work), None, _lib.stream())
        return out

    @property
    def last_oc_steps(self):
        """Bisection steps of the last update_design (reads the device state: synchronises)."""
        if self._multi():
            return getattr(self, "_multi_steps", 0)
        return int(self._oc_work[:5].view(torch.int32)[9].item())

    def _dot(self, a, b):
        _lib.call("et_dot_f64", a.numel(), _lib.f64(a), _lib.f64(b, a.numel()), _lib.f64(self._scalar[0:1]),
                  _lib.stream())
        return float(self._scalar[0].item())

    def _rms_change(self, a, b):
        _lib.call("et_sqdist_f64", a.numel(), _lib.f64(a), _lib.f64(b, a.numel()), _lib.f64(self._scalar[1:2]),
                  _lib.stream())
        return float(np.sqrt(self._scalar[1].item() / a.numel()))

    def _solve_and_process_multi(self, rho, iteration):
        """m materials: each column filtered, E from the multi-material map, dc_k = H^T
        (dE/dr_k * (-u_e^T K0 u_e)) per column (TopOpt/_cuda.py:282-302)."""
        mm = self.material_model
        m = mm.n_sets
        rho = rho.reshape(-1, m)
        rho_ = torch.stack([self.filter_kernel.dot(rho[:, k].contiguous()) for k in range(m)], dim=1).contiguous()
        E = mm(rho_, iteration)
        self.kernel.set_rho(E)
        U, residual = self.solver.solve(self.f, use_last=self.reuse_sol)
        compliance = self._dot(self.f, U)
        dk = self.kernel.process_grad(U)
        dr = mm.grad(rho_, iteration) * dk.reshape(-1, 1)
        dr = torch.stack([self.filter_kernel._rmatvec(dr[:, k].contiguous()) for k in range(m)], dim=1)
        return compliance, dr, residual

    def solve_and_process(self, rho, iteration):
        mm = self.material_model
        if self._multi():
            return self._solve_and_process_multi(rho, iteration)
        rho_ = self.filter_kernel.dot(rho)
        rho__ = mm(rho_, iteration)
        self.kernel.set_rho(rho__)
        U, residual = self.solver.solve(self.f, use_last=self.reuse_sol)
        compliance = self._dot(self.f, U)
        pen = mm.penalty_schedule(mm.penalty, iteration) if mm.penalty_schedule else mm.penalty
        if hasattr(self.filter_kernel, "simp_adjoint"):
            # sensitivity and filter transpose in one kernel
            dr = self.filter_kernel.simp_adjoint(self.kernel, U, rho_, pen, mm.void)
        else:
            dr = self.filter_kernel._rmatvec(self.kernel.sensitivity(U, rho_, pen, mm.void))
        return compliance, dr.reshape(-1, 1), residual

    def optimize(self, save_comp_history=False, save_change_history=False, save_rho_history=False,
                 rho_init=None):
        flag = False
        multi = self._multi()
        if rho_init is None:
            rho = self.material_model.init_desvars(_n_el(self.kernel))
        elif multi:
            # clip(x0, void, 1) per material (the intent of :321-322)
            x0 = torch.as_tensor(rho_init, dtype=torch.float64, device=self.f.device).reshape(-1).contiguous()
            rho = torch.empty_like(x0)
            _lib.call("et_simp_forward_f64", x0.numel(), _lib.f64(x0), 1.0, float(self.material_model.void), 1,
                      _lib.f64(rho, x0.numel()), _lib.stream())
            rho = rho.reshape(-1, self.material_model.n_sets)
        else:
            x0 = torch.as_tensor(rho_init, dtype=torch.float64, device=self.f.device).reshape(-1)
            rho = self.material_model(x0.contiguous(), 0, plain=True)
        if not multi:
            self.kernel.set_rho(rho)
        t0 = time.perf_counter()
        comp, df, res = self.solve_and_process(rho, 0)
        self.iteration_log = [dict(it=-1, comp=comp, res=res, pcg=self.solver.last_iterations,
                                   t=time.perf_counter() - t0, setup_ms=self.solver.last_timings[0],
                                   pcg_ms=self.solver.last_timings[1], oc_steps=0)]
        comp_history, change_history, rho_history = [], [], []
        for i in range(self.max_iter):
            t0 = time.perf_counter()
            rho_old = rho.clone()
            comp_old = comp
            rho = self.update_design(rho, df)
            comp, df, res = self.solve_and_process(rho, i)
            change = self._rms_change(rho, rho_old)
            change_f = abs(comp - comp_old) / comp
            self.iteration_log.append(dict(it=i, comp=comp, res=res, pcg=self.solver.last_iterations,
                                           t=time.perf_counter() - t0,
                                           setup_ms=self.solver.last_timings[0],
                                           pcg_ms=self.solver.last_timings[1],
                                           oc_steps=self.last_oc_steps, change=change))
            if save_change_history:
                change_history.append(change)
            if save_rho_history:
                rho_history.append(rho.cpu().numpy())
            if save_comp_history:
                comp_history.append(comp)
            if self.verbose:
                logger.info(f"it {i}: Compliance: {comp:.4e}, Change: {change:.4e}, "
                            f"Function Change: {change_f:.4e}, Residual: {res:.4e}")
            if change < self.ch_tol and change_f < self.fun_tol and self.material_model.is_terminal(i):
                flag = True
                break
        hist = {}
        if save_comp_history:
            hist["comp_history"] = comp_history
        if save_change_history:
            hist["change_history"] = change_history
        if save_rho_history:
            hist["rho_history"] = rho_history
        return rho, flag, hist


CuTopOpt = TopOpt
