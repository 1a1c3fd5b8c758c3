"""SIMP material model (device).

Mirrors EDGETO/MaterialModels.py: MaterialModel (:3-28) and SingleMaterial
(:31-99).  The elementwise maps run in the library (et_simp_forward_f64);
the OC multiplier and the volume test are evaluated inside the device OC
update (et_oc_update_f64) rather than as separate host-synchronising calls.
evaluate_constraint / ocP are kept for callers that drive their own OC loop:
they take device tensors (or host arrays, moved to the device) and compute
with torch on the device.
"""
import numpy as np
import torch

from . import _lib


def _dev_args(*xs):
    """float64 tensors on the device of the first tensor argument (else cuda:0)."""
    dev = next((x.device for x in xs if isinstance(x, torch.Tensor)), None)
    dev = torch.device("cuda") if dev is None else dev
    return [torch.as_tensor(x, dtype=torch.float64, device=dev) for x in xs]


def _oc_product(df, dg, rho, clamp):
    """rho * nan_to_num(sqrt(-df / dg[:, None])), negative ratios giving 0
    (MaterialModels.py:83-90, :217-225); all zero -> 1e-3 everywhere."""
    df, dg, rho = _dev_args(df, dg, rho)
    q = -df / dg.reshape(-1, 1)
    if clamp:
        q = torch.clamp(q, min=0.0)
    out = rho * torch.nan_to_num(torch.sqrt(q), nan=0.0)
    if float(out.abs().sum()) == 0.0:
        out = torch.full_like(out, 1e-3)
    return out


class MaterialModel:
    def __init__(self, E=1.0, nu=0.33, n_sets=1):
        self.E = E
        self.nu = nu
        self.n_sets = n_sets

    def __call__(self, rho, iteration, **kwargs):
        return self.__forward__(rho, iteration, **kwargs)

    def evaluate_constraint(self, rho, dg, vol, **kwargs):
        return self.__con__(rho, dg, vol, **kwargs)

    def ocP(self, df, dg, rho, **kwargs):
        return self.__ocP__(df, dg, rho, **kwargs)

    def grad(self, rho, iteration, **kwargs):
        return self.__backward__(rho, iteration, **kwargs)

    def base_properties(self):
        return {"E": self.E, "nu": self.nu, "n_sets": self.n_sets}

    def is_terminal(self, iteration):
        return True

    def init_desvars(self, nel, **kwargs):
        return self.__initvars__(nel, **kwargs)


class SingleMaterial(MaterialModel):
    def __init__(self, E=1.0, nu=0.33, void=1e-6, penalty=3.0, volume_fraction=0.25,
                 penalty_schedule=None):
        super().__init__(E=E, nu=nu, n_sets=1)
        self.void = void
        self.penalty = penalty
        self.volume_fraction = volume_fraction
        self.penalty_schedule = penalty_schedule

    def _pen(self, iteration):
        if self.penalty_schedule is not None:
            return self.penalty_schedule(self.penalty, iteration)
        return self.penalty

    def __forward__(self, rho, iteration, plain=False, out=None, **_):
        rho = rho.contiguous()
        out = torch.empty_like(rho) if out is None else out
        _lib.call("et_simp_forward_f64", rho.numel(), _lib.f64(rho), float(self._pen(iteration)),
                  float(self.void), 1 if plain else 0, _lib.f64(out, rho.numel()), _lib.stream())
        return out

    def __backward__(self, rho, iteration, **_):
        """p * clip(rho, void, 1)^(p-1) -- the SIMP factor of the sensitivity."""
        pen = float(self._pen(iteration))
        clipped = self.__forward__(rho, iteration, plain=True)
        E = torch.empty_like(clipped)
        # p * x^(p-1) = p * (x^p) / x with x >= void > 0
        _lib.call("et_simp_forward_f64", clipped.numel(), _lib.f64(clipped), pen - 1.0, 0.0, 0,
                  _lib.f64(E, clipped.numel()), _lib.stream())
        return E * pen

    def __con__(self, rho, dg, vol, **_):
        """[sum(rho * dg) / vol <= volume_fraction] as a host bool array (MaterialModels.py:75-81)."""
        rho, dg = _dev_args(rho, dg)
        V = float((rho.reshape(-1) * dg.reshape(-1)).sum())
        return np.array([V / vol <= self.volume_fraction])

    def __ocP__(self, df, dg, rho, **_):
        return _oc_product(df, dg, rho, clamp=True)

    def __initvars__(self, nel, dtype=torch.float64, **_):
        return torch.ones(nel, dtype=torch.float64, device=torch.device("cuda"))

    def is_terminal(self, iteration):
        if self.penalty_schedule is not None:
            return self.penalty_schedule(self.penalty, iteration) == self.penalty
        return True


class PenalizedMultiMaterial(MaterialModel):
    """Discrete multi-material SIMP (MaterialModels.py:100-235): densities rho (n_el, m),
    E = clip(sum_i E_i r_i^p prod_{j != i} (1 - r_j^p), void, inf), gradient, per-material
    (or, with rho_mat, total-mass) volume constraints and the OC multiplier.  The elementwise
    maps run in the library (et_multimat_f64, m <= 4); the constraints are evaluated inside the
    multi-multiplier OC update (et_oc_update_multi_f64)."""

    _rp = 0

    def __init__(self, E=1.0, nu=0.33, n_material=2, mass=np.array([[0.25], [0.25]]),
                 E_mat=np.array([[1.0], [0.5]]), rho_mat=None, void=1e-6, penalty=3.0, penalty_schedule=None):
        super().__init__(E=E, nu=nu, n_sets=n_material)
        self.n_material = n_material
        self.mass = np.asarray(mass, dtype=np.float64)
        self.E_mat = np.asarray(E_mat, dtype=np.float64)
        self.rho_mat = None if rho_mat is None else np.asarray(rho_mat, dtype=np.float64)
        self.void = void
        self.penalty = penalty
        self.penalty_schedule = penalty_schedule
        if self.E_mat.shape[0] != self.n_material:
            raise ValueError("E_mat must have the same number of materials as n_material")
        if self.rho_mat is not None:
            if self.rho_mat.shape[0] != self.n_material:
                raise ValueError("rho_mat must have the same number of materials as n_material")
            if self.mass.shape[0] != 1:
                raise ValueError("if rho_mat is not None, mass must have only one value")
            self.mass = self.mass.reshape(-1)[:1]
        elif self.mass.shape[0] != self.n_material:
            raise ValueError("Mass and E_mat must have the same number of materials")
        self._E_dev = None

    def _pen(self, iteration):
        if self.penalty_schedule is not None:
            return self.penalty_schedule(self.penalty, iteration)
        return self.penalty

    def _emat(self, device):
        if self._E_dev is None or self._E_dev.device != device:
            self._E_dev = torch.tensor(self.E_mat.reshape(-1), dtype=torch.float64, device=device)
        return self._E_dev

    def _call(self, rho, iteration, mode):
        rho = rho.reshape(-1, self.n_material).contiguous()
        n = rho.shape[0]
        out = torch.empty((n, self.n_material) if mode == 2 else (n,), dtype=torch.float64, device=rho.device)
        _lib.call("et_multimat_f64", n, self.n_material, _lib.f64(rho), _lib.f64(self._emat(rho.device)),
                  float(self._pen(iteration)), float(self.void), self._rp, mode, _lib.f64(out), _lib.stream())
        return out

    def __forward__(self, rho, iteration, plain=False, **_):
        return self._call(rho, iteration, 1 if plain else 0)

    def __backward__(self, rho, iteration, **_):
        return self._call(rho, iteration, 2)

    def __con__(self, rho, dg, vol, **_):
        """Constraint flags per material as host booleans (MaterialModels.py:203-215): each
        material's volume sum(rho_k * dg) / vol against its mass, or with rho_mat the total
        mass against the one budget.  dg is one measure per element or one shared value."""
        rho, dg = _dev_args(rho, dg)
        V = (rho.reshape(-1, self.n_material) * dg.reshape(-1, 1)).sum(dim=0).cpu().numpy()
        if self.rho_mat is None:
            return (V / vol - self.mass.reshape(-1)) <= 0
        V = float(V @ self.rho_mat.reshape(-1))
        return np.ones(self.n_material, dtype=bool) * (V / vol <= self.mass.reshape(-1)[0])

    def __ocP__(self, df, dg, rho, **_):
        return _oc_product(df, dg, rho, clamp=False)

    def __initvars__(self, nel, dtype=torch.float64, **_):
        return torch.ones((nel, self.n_material), dtype=torch.float64, device=torch.device("cuda"))

    def is_terminal(self, iteration):
        if self.penalty_schedule is not None:
            return self.penalty_schedule(self.penalty, iteration) == self.penalty
        return True


class PenalizedMultiMaterialRP(PenalizedMultiMaterial):
    """The RAMP-like variant (MaterialModels.py:237-369): w_j = (1 - r_j)^p."""

    _rp = 1
