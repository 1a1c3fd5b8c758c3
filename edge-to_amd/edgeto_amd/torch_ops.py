"""The hot-path operations as torch custom operators (torch.ops.edgeto_amd.*).

Each operator is a functional wrapper of one C-ABI entry point (no object state),
registered with torch's dispatcher together with a shape function, so the
operations can be traced (torch.compile, FX) and appear in torch's profiler by
name.  All of them take float64 device tensors and run on torch's current
stream, exactly as the classes do:

    matvec(K0, E, cons, u, nel)                     A_hat(E) u          (et_matvec_f64)
    filter_fwd(offsets, w, x, nel)                  H x                 (et_filter_fwd_f64)
    filter_adj(offsets, w, norm, v, nel)            H^T v               (et_filter_adj_f64)
    simp_adjoint(K0, U, xt, penal, void, offsets, w, norm, nel, k0_flags)
                                                    H^T [p clip(xt)^(p-1) (-u^T K0 u)] (et_simp_adjoint_f64)
    oc_update(x, dc, A, vol, vf, move)              the OC design update (et_oc_update_f64)
"""
from typing import List, Optional

import torch

from . import _lib


def _nel(nel):
    return _lib.nel_arg([int(v) for v in nel])


@torch.library.custom_op("edgeto_amd::matvec", mutates_args=())
def matvec(K0: torch.Tensor, E: torch.Tensor, cons: Optional[torch.Tensor], u: torch.Tensor,
           nel: List[int]) -> torch.Tensor:
    u = u.contiguous()
    y = torch.empty_like(u)
    _lib.call("et_matvec_f64", len(nel), _nel(nel), _lib.f64(K0.contiguous()), _lib.f64(E.contiguous()),
              None if cons is None else _lib.u8(cons.contiguous(), u.numel()), _lib.f64(u), _lib.f64(y, u.numel()),
              _lib.stream())
    return y


@matvec.register_fake
def _(K0, E, cons, u, nel):
    return torch.empty_like(u)


@torch.library.custom_op("edgeto_amd::filter_fwd", mutates_args=())
def filter_fwd(offsets: torch.Tensor, w: torch.Tensor, x: torch.Tensor, nel: List[int]) -> torch.Tensor:
    x = x.contiguous()
    y = torch.empty_like(x)
    _lib.call("et_filter_fwd_f64", len(nel), _nel(nel), w.numel(), _lib.i32(offsets.contiguous()),
              _lib.f64(w.contiguous()), _lib.f64(x), _lib.f64(y), _lib.stream())
    return y


@filter_fwd.register_fake
def _(offsets, w, x, nel):
    return torch.empty_like(x)


@torch.library.custom_op("edgeto_amd::filter_adj", mutates_args=())
def filter_adj(offsets: torch.Tensor, w: torch.Tensor, norm: torch.Tensor, v: torch.Tensor,
               nel: List[int]) -> torch.Tensor:
    v = v.contiguous()
    y = torch.empty_like(v)
    _lib.call("et_filter_adj_f64", len(nel), _nel(nel), w.numel(), _lib.i32(offsets.contiguous()),
              _lib.f64(w.contiguous()), _lib.f64(norm.contiguous()), _lib.f64(v), _lib.f64(y), _lib.stream())
    return y


@filter_adj.register_fake
def _(offsets, w, norm, v, nel):
    return torch.empty_like(v)


@torch.library.custom_op("edgeto_amd::simp_adjoint", mutates_args=())
def simp_adjoint(K0: torch.Tensor, U: torch.Tensor, xt: torch.Tensor, penal: float, void: float,
                 offsets: torch.Tensor, w: torch.Tensor, norm: torch.Tensor, nel: List[int],
                 k0_flags: int) -> torch.Tensor:
    xt = xt.contiguous()
    out = torch.empty_like(xt)
    _lib.call("et_simp_adjoint_f64", len(nel), _nel(nel), _lib.f64(K0.contiguous()), _lib.f64(U.contiguous()),
              _lib.f64(xt), float(penal), float(void), w.numel(), _lib.i32(offsets.contiguous()),
              _lib.f64(w.contiguous()), _lib.f64(norm.contiguous()), int(k0_flags), _lib.f64(out), _lib.stream())
    return out


@simp_adjoint.register_fake
def _(K0, U, xt, penal, void, offsets, w, norm, nel, k0_flags):
    return torch.empty_like(xt)


@torch.library.custom_op("edgeto_amd::oc_update", mutates_args=())
def oc_update(x: torch.Tensor, dc: torch.Tensor, A: float, vol: float, vf: float, move: float) -> torch.Tensor:
    x = x.contiguous()
    n = x.numel()
    work = torch.empty(int(_lib.lib().et_oc_work_size(n)) // 8 + 1, dtype=torch.float64, device=x.device)
    out = torch.empty_like(x)
    _lib.call("et_oc_update_f64", n, _lib.f64(x), _lib.f64(dc.contiguous(), n), float(A), float(vol), float(vf),
              float(move), _lib.f64(out), _lib.f64(work), None, _lib.stream())
    return out


@oc_update.register_fake
def _(x, dc, A, vol, vf, move):
    return torch.empty_like(x)
