"""SIMP material model (device).

Mirrors EDGETO/MaterialModels.py: MaterialModel (:3-28) and SingleMaterial
(:31-99).  The elementwise maps run in the library (et_simp_forward_f64);
the OC multiplier and the volume test are evaluated inside the device OC
update (et_oc_update_f64) rather than as separate host-synchronising calls.
"""
import torch

from . import _lib


class MaterialModel:
    def __init__(self, E=1.0, nu=0.33, n_sets=1):
        self.E = E
        self.nu = nu
        self.n_sets = n_sets

    def __call__(self, rho, iteration, **kwargs):
        return self.__forward__(rho, iteration, **kwargs)

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

    def __initvars__(self, nel, dtype=torch.float64, **_):
        return torch.ones(nel, dtype=torch.float64, device=torch.device("cuda"))

    def is_terminal(self, iteration):
        if self.penalty_schedule is not None:
            return self.penalty_schedule(self.penalty, iteration) == self.penalty
        return True
