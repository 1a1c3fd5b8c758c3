"""Timing harness of the fine-grid operator entry points on a synthetic structured grid.

    python edge-to_amd/edgeto_amd/opbench.py [nelx nely nelz] [--reps N]

Runs et_matvec / et_residual / et_jacobi in FP64 and FP32 on random inputs (unit-cube H8,
nu 0.3, densities U(0, 1), the y = 0 node plane constrained), `reps` times each, and prints
the median time per call measured with HIP events around each call.  Each call also does
its host-side work (argument checks, the element-matrix structure check), so the kernel
durations are the ones a kernel trace or counter collection of this script reports
(rocprofv3 --kernel-trace / --pmc), not the printed times.
"""
import argparse
import os
import sys

import numpy as np
import torch

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
from edgeto_amd import _lib  # noqa: E402
from edgeto_amd import physics  # noqa: E402

_CUBE = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], float)


def _median_us(fn, reps):
    fn()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ts = []
    for _ in range(reps):
        e0.record()
        fn()
        e1.record()
        e1.synchronize()
        ts.append(e0.elapsed_time(e1) * 1e3)
    return float(np.median(ts))


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("nel", nargs="*", type=int, default=[256, 128, 128])
    ap.add_argument("--reps", type=int, default=30)
    a = ap.parse_args(argv)
    nel = tuple(a.nel)
    if len(nel) != 3:
        raise SystemExit("three element counts expected")
    _lib.load()
    K0 = physics.auto_stiffness(_CUBE, 1.0, 0.3)[0]
    ne = int(np.prod(nel))
    nn = int(np.prod([v + 1 for v in nel]))
    g = torch.Generator(device="cuda").manual_seed(1)
    out = {}
    for dt, s in ((torch.float64, "f64"), (torch.float32, "f32")):
        K = torch.as_tensor(K0, device="cuda").to(dt)
        E = torch.rand(ne, device="cuda", dtype=dt, generator=g)
        u = torch.randn(3 * nn, device="cuda", dtype=dt, generator=g)
        b = torch.randn(3 * nn, device="cuda", dtype=dt, generator=g)
        D = torch.rand(3 * nn, device="cuda", dtype=dt, generator=g) + 1.0
        y = torch.empty_like(u)
        cons = torch.zeros(3 * nn, device="cuda", dtype=torch.uint8)
        cons[: 3 * (nel[0] + 1) * (nel[2] + 1)] = 1  # node plane y = 0
        P, na = _lib.ptr, _lib.nel_arg(nel)
        calls = {
            "matvec": lambda: _lib.call("et_matvec_" + s, 3, na, P(K), P(E), P(cons), P(u), P(y), _lib.stream()),
            "residual": lambda: _lib.call("et_residual_" + s, 3, na, P(K), P(E), P(cons), P(u), P(b), P(y),
                                          _lib.stream()),
            "jacobi": lambda: _lib.call("et_jacobi_" + s, 3, na, P(K), P(E), P(cons), P(u), P(b), P(D), 0.6, P(y),
                                        _lib.stream()),
        }
        for name, fn in calls.items():
            out[f"{name}_{s}"] = _median_us(fn, a.reps)
    print(f"[opbench] {nel}: " + "  ".join(f"{k} {v:.1f} us" for k, v in out.items()), flush=True)


if __name__ == "__main__":
    main()
