"""Headline benchmark: SIMP OC iterations per second on one MI355X.

Workload (default): 3-D cantilever of 256 x 128 x 128 unit H8 elements
(4 194 304 elements, 12 830 211 DOF), E0 = 1, Emin = 1e-9, nu = 0.3, volume
fraction 0.3, penalty 3, hat filter rmin = 2.0, x0 = 0.3, face x = 0 clamped,
unit total downward load spread over the bottom edge at x = L, 5-grid GMG
(n_level = 4) V(2,2)-PCG to relative residual 1e-8.  One step = one OC
iteration: OC update, filter, SIMP map, GMG-PCG solve (hierarchy rebuilt),
compliance, sensitivity, filter transpose.

Both legs run the same steps from the same start: the mixed leg (FP32 V-cycle
inside FP64 PCG) gives `value`; the plain FP64 leg is reported beside it with
the compliance agreement.  `roofline` is the dominant kernel class of the
mixed leg, timed with HIP events on the solver stream over the timed region.
`cpu_baseline` times the NumPy/SciPy oracle on a bounded sample (see
`cpu_baseline.sample`).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c3|c2|c5]
(c5: 64 seeded random 2-D/3-D problems solved in turn; `steps` OC iterations each.)
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "edge-to_amd"))

METRIC = "SIMP iters/s, 3D 12.8M-DOF H8: mixed vs FP64 GMG-PCG at rel. compliance 1e-6"
HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    "c3": dict(nel=(256, 128, 128), n_level=4),
    "c2": dict(nel=(96, 48, 48), n_level=3),
}


def build_problem(ET, nel, n_level, precision):
    import torch
    nx, ny, nz = nel
    mesh = ET.CuStructuredMesh3D(nx, ny, nz, nx, ny, nz,
                                 stiffness=lambda x0: ET.auto_stiffness(x0, 1.0, 0.3))
    kernel = ET.CuStructuredStiffnessKernel(mesh)
    filt = ET.CuStructuredFilter3D(mesh, 2.0)
    mat = ET.SingleMaterial(void=1e-9, penalty=3.0, volume_fraction=0.3)
    solver = ET.CuMultiGrid(mesh, kernel, tol=1e-8, n_smooth=2, n_level=n_level, precision=precision)
    tp = ET.CuTopOpt(mesh, mat, kernel, solver, filt, max_iter=0, ch_tol=0, fun_tol=0, verbose=False)
    nx1, ny1, nz1 = nx + 1, ny + 1, nz + 1
    j, k = np.meshgrid(np.arange(ny1), np.arange(nz1), indexing="ij")
    tp.add_BC_nodal((j * nx1 * nz1 + k).ravel(), np.array([[1, 1, 1]]))
    load = np.array([nx * nz1 + kk for kk in range(nz1)])  # x = L, y = 0, all z
    tp.add_F_nodal(load, np.array([[0.0, -1.0 / len(load), 0.0]]))
    torch.cuda.synchronize()
    return mesh, tp


def _profile_read(_lib, h):
    ms = (ctypes.c_double * len(_lib.PROF_KINDS))()
    cnt = (ctypes.c_int64 * len(_lib.PROF_KINDS))()
    _lib.call("et_mg_profile_read", h, ctypes.cast(ms, ctypes.c_void_p), ctypes.cast(cnt, ctypes.c_void_p))
    return {k: (ms[i], cnt[i]) for i, k in enumerate(_lib.PROF_KINDS)}


def run_leg(ET, wl, precision, warmup, steps, profile=False):
    """Initial solve + warmup OC iterations, then `steps` timed OC iterations.

    With `profile`, the last warmup iteration (untimed) records every kernel class
    with HIP events; the timed iterations record only the dominant fine-grid class
    (two event records per launch of it), whose average launch time feeds the
    roofline figure."""
    import torch
    from edgeto_amd import _lib
    mesh, tp = build_problem(ET, wl["nel"], wl["n_level"], precision)
    x = tp.material_model(torch.full((mesh.n_el,), 0.3, dtype=torch.float64, device="cuda"), 0, plain=True)
    tp.kernel.set_rho(x)
    comp, df, res = tp.solve_and_process(x, 0)
    h = tp.solver._handle
    classes, dominant = None, None
    for i in range(warmup):
        if profile and i == warmup - 1:
            _lib.call("et_mg_profile", h, -1)
        x = tp.update_design(x, df)
        comp, df, res = tp.solve_and_process(x, i)
    if profile and warmup > 0:
        classes = _profile_read(_lib, h)
        cand = {k: v for k, v in classes.items() if k in ROOFLINE_CLASSES and v[1] > 0}
        dominant = max(cand, key=lambda k: cand[k][0]) if cand else None
    if profile:
        kind = _lib.PROF_KINDS.index(dominant or "l0_jacobi")
        _lib.call("et_mg_profile", h, 1 << kind)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    pcg, per_step = [], []
    for i in range(steps):
        ts = time.perf_counter()
        x = tp.update_design(x, df)
        comp, df, res = tp.solve_and_process(x, warmup + i)   # compliance read back: host sync
        per_step.append(time.perf_counter() - ts)
        pcg.append(tp.solver.last_iterations)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    prof = None
    if profile:
        prof = _profile_read(_lib, h)
        _lib.call("et_mg_profile", h, 0)
    return dict(elapsed=elapsed, per_step=per_step, comp=comp, res=res, pcg=pcg, prof=prof,
                classes=classes, n_dof=tp.kernel.shape[0], n_el=mesh.n_el)


def batch_problem(i):
    """Problem i of the 64-problem batch: seeded by default_rng(1000 + i)."""
    rng = np.random.default_rng(1000 + i)
    dim = 2 if rng.random() < 0.5 else 3
    if dim == 2:
        nel = (int(rng.integers(64, 513)), int(rng.integers(32, 257)))
    else:
        nx = int(rng.integers(32, 97)) // 2 * 2
        nel = (nx, nx // 2, nx // 2)
    side = int(rng.integers(0, 4))  # clamped side: x = 0, x = L, y = 0, y = H
    n_loads = int(rng.integers(1, 5))
    vf = float(rng.uniform(0.2, 0.6))
    rmin = float(rng.uniform(1.2, 3.0))
    nn = [n + 1 for n in nel]
    idx = np.indices(nn[::-1] if dim == 2 else (nn[1], nn[0], nn[2]))
    if dim == 2:
        j, i_ = idx[0].ravel(), idx[1].ravel()
        k = np.zeros_like(i_)
        node = j * nn[0] + i_
    else:
        j, i_, k = idx[0].ravel(), idx[1].ravel(), idx[2].ravel()
        node = (j * nn[0] + i_) * nn[2] + k
    on_side = [i_ == 0, i_ == nel[0], j == 0, j == nel[1]]
    boundary = on_side[0] | on_side[1] | on_side[2] | on_side[3]
    if dim == 3:
        boundary |= (k == 0) | (k == nel[2])
    clamp = node[on_side[side]]
    cand = node[boundary & ~on_side[side]]
    loads = rng.choice(cand, size=n_loads, replace=False)
    dirs = rng.standard_normal((n_loads, dim))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    n_level = 3
    return dict(dim=dim, nel=nel, clamp=clamp, loads=loads, dirs=dirs, vf=vf, rmin=rmin, n_level=n_level)


def run_batch(ET, n_problems, iters):
    """Batched workload: independent random problems solved one after another on
    one GPU (mixed precision); aggregate OC iterations/s and per-problem latency."""
    import torch
    lat, total_it, t_all = [], 0, 0.0
    for i in range(n_problems):
        pb = batch_problem(i)
        nel = pb["nel"]
        stiff = lambda x0: ET.auto_stiffness(x0, 1.0, 0.3)
        if pb["dim"] == 2:
            mesh = ET.CuStructuredMesh2D(*nel, *nel, stiffness=stiff)
            filt = ET.CuStructuredFilter2D(mesh, pb["rmin"])
        else:
            mesh = ET.CuStructuredMesh3D(*nel, *nel, stiffness=stiff)
            filt = ET.CuStructuredFilter3D(mesh, pb["rmin"])
        kernel = ET.CuStructuredStiffnessKernel(mesh)
        mat = ET.SingleMaterial(void=1e-9, penalty=3.0, volume_fraction=pb["vf"])
        solver = ET.CuMultiGrid(mesh, kernel, tol=1e-8, n_smooth=2, n_level=pb["n_level"], precision="mixed")
        tp = ET.CuTopOpt(mesh, mat, kernel, solver, filt, max_iter=0, ch_tol=0, fun_tol=0, verbose=False)
        tp.add_BC_nodal(pb["clamp"], np.ones((1, pb["dim"]), dtype=np.int64))
        tp.add_F_nodal(pb["loads"], pb["dirs"])
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        x = tp.material_model(torch.full((mesh.n_el,), pb["vf"], dtype=torch.float64, device="cuda"), 0,
                              plain=True)
        tp.kernel.set_rho(x)
        comp, df, _ = tp.solve_and_process(x, 0)
        for it in range(iters):
            x = tp.update_design(x, df)
            comp, df, _ = tp.solve_and_process(x, it)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        lat.append(dt / iters)
        total_it += iters
        t_all += dt
        print(f"[bench] batch problem {i}: {pb['dim']}-D {'x'.join(map(str, nel))}, "
              f"{dt / iters * 1e3:.1f} ms/iter, comp {comp:.6e}", file=sys.stderr, flush=True)
    lat = np.array(lat)
    return {"metric": "SIMP iters/s, 64-problem batch (mixed GMG-PCG)", "value": round(total_it / t_all, 3),
            "unit": "SIMP iters/s", "n_gpus": 1, "steps": iters, "warmup": 0,
            "ms_per_step": round(t_all / total_it * 1e3, 3), "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f32 V-cycle + f64 PCG", "data": "synthetic (seeded random problems)",
            "config": {"workload": f"{n_problems} random 2-D/3-D cantilever problems, default_rng(1000 + i), "
                                   f"initial solve + {iters} OC iterations each, 4-grid GMG", "parallelism": "single GPU"},
            "latency_ms_per_iter": {"p50": round(float(np.percentile(lat, 50)) * 1e3, 3),
                                    "p99": round(float(np.percentile(lat, 99)) * 1e3, 3)}}


def measured_traffic(kernel_class, workload):
    """HBM bytes per launch of the dominant kernel from the committed rocprofv3 --pmc
    summary (FETCH_SIZE and WRITE_SIZE, each scaled by its streaming-kernel calibration
    factor), or None when no summary for this kernel and workload is present."""
    path = os.path.join(ROOT, "profiles", f"{workload}_{kernel_class}_traffic.json")
    try:
        with open(path) as fh:
            t = json.load(fh)
    except (OSError, ValueError):
        return None
    return (t["fetch_kib"] * 1024.0 * t["fetch_factor"] + t["write_kib"] * 1024.0 * t["write_factor"],
            os.path.relpath(path, ROOT))


ROOFLINE_CLASSES = ("l0_jacobi", "l0_jacobi0", "l0_residual", "l0_jacobi_add", "matvec64")


def roofline(prof, n_dof, n_el, workload):
    """Dominant kernel class of the mixed leg and its achieved algorithmic bandwidth."""
    s = 4  # FP32 V-cycle vectors / E
    mask = n_dof  # one byte per DOF
    bytes_per = {  # algorithmic bytes per launch
        "l0_jacobi": s * (4 * n_dof + n_el) + mask,       # read x, b, D^-1, E; write y
        "l0_jacobi0": s * (3 * n_dof + n_el) + mask,      # read b, D^-1, E; write y (x0 = w D^-1 b)
        "l0_residual": s * (3 * n_dof + n_el) + mask,     # read x, b, E; write r
        "l0_jacobi_add": s * (6 * n_dof + n_el) + mask,   # + read x_prev and r (FP32 copy, for r.z)
        "matvec64": 8 * (2 * n_dof + n_el) + mask,        # FP64: read x, E; write y
    }
    cand = {k: v for k, v in prof.items() if k in bytes_per and v[1] > 0}
    name = max(cand, key=lambda k: cand[k][0])
    ms, count = cand[name]
    avg_s = ms / count * 1e-3
    achieved = bytes_per[name] / avg_s / 1e9
    tr = measured_traffic(name, workload)
    out = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None if tr is None else round(tr[0]),
           "kernel": name, "avg_us": round(avg_s * 1e6, 2), "launches": count,
           "bytes_per_launch": bytes_per[name]}
    if tr is not None:
        out["traffic_source"] = tr[1]
    return out


def cpu_baseline(nel_metric, sample_nel=(96, 48, 48), n_level=3):
    """Oracle (NumPy/SciPy, FP64, one thread) on a bounded sample, scaled by element count."""
    from threadpoolctl import threadpool_limits
    sys.path.insert(0, ROOT)
    from oracle import topopt
    from oracle.solvers import MultiGrid
    with threadpool_limits(limits=1):
        nx, ny, nz = sample_nel
        p = topopt.Problem(sample_nel, nu=0.3, r_min=2.0)
        nx1, ny1, nz1 = nx + 1, ny + 1, nz + 1
        j, k = np.meshgrid(np.arange(ny1), np.arange(nz1), indexing="ij")
        clamp = (j * nx1 * nz1 + k).ravel()
        for d in range(3):
            p.cons[clamp * 3 + d] = True
        load = np.array([nx * nz1 + kk for kk in range(nz1)])
        p.f[load * 3 + 1] = -1.0 / len(load)
        mat = topopt.SingleMaterial(void=1e-9, penalty=3.0, volume_fraction=0.3)
        solver = MultiGrid(p.nel, 3, tol=1e-8, n_smooth=2, n_level=n_level, mode="gpu")
        x = np.full(p.n_el, 0.3)
        c, dr, _, _ = topopt.solve_and_process(p, x, mat, solver)   # initial solve (untimed)
        t = time.perf_counter()
        x, _ = topopt.oc_update(x, dr, p.A, p.volume, 0.3, 0.2, mat)
        topopt.solve_and_process(p, x, mat, solver)
        dt = time.perf_counter() - t
    scale = np.prod(nel_metric) / np.prod(sample_nel)
    return {"value": round(1.0 / (dt * scale), 6), "unit": "SIMP iters/s", "cores": 1, "kind": "port",
            "sample": f"1 OC iteration of the oracle on a {sample_nel[0]}x{sample_nel[1]}x{sample_nel[2]} "
                      f"H8 cantilever (same BCs/load/filter, {n_level + 1}-grid GMG, tol 1e-8): {dt:.2f} s, "
                      f"scaled by the element-count ratio {scale:.0f} to {nel_metric[0]}x{nel_metric[1]}x{nel_metric[2]}",
            "sample_seconds": round(dt, 3),
            "host_cpus": os.cpu_count(), "affinity": len(os.sched_getaffinity(0))}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="c3", choices=sorted(WORKLOADS) + ["c5"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--legs", default="mixed,fp64")
    ap.add_argument("--validate", type=int, default=30,
                    help="OC iterations of the mixed-vs-FP64 compliance gate (0: skip)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (requested --gpus {args.gpus}); no multi-GPU path "
              "is built (the slab sharding along y is a design only).", file=sys.stderr)
        sys.exit(2)
    import torch
    import edgeto_amd as ET
    if args.workload == "c5":
        print(json.dumps(run_batch(ET, 64, args.steps)))
        return
    wl = WORKLOADS[args.workload]
    legs = {}
    for leg in args.legs.split(","):
        print(f"[bench] leg {leg}: warmup {args.warmup}, steps {args.steps}", file=sys.stderr, flush=True)
        legs[leg] = run_leg(ET, wl, leg, args.warmup, args.steps, profile=(leg == "mixed"))
        print(f"[bench] leg {leg}: {args.steps / legs[leg]['elapsed']:.3f} it/s, "
              f"pcg {legs[leg]['pcg']}, comp {legs[leg]['comp']:.12e}", file=sys.stderr, flush=True)
    main_leg = legs.get("mixed") or legs[list(legs)[0]]
    value = args.steps / main_leg["elapsed"]
    out = {"metric": METRIC, "value": round(value, 4), "unit": "SIMP iters/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": round(main_leg["elapsed"] / args.steps * 1e3, 2), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None,
           "dtype": "f32 V-cycle + f64 PCG" if "mixed" in legs else "f64",
           "data": "synthetic (closed-form cantilever BCs and load, uniform x0 = 0.3)",
           "config": {"workload": f"3-D H8 cantilever {'x'.join(map(str, wl['nel']))}, "
                                  f"{wl['n_level'] + 1}-grid GMG V(2,2)-PCG tol 1e-8, rmin 2.0, vf 0.3",
                      "n_dof": main_leg["n_dof"], "n_elem": main_leg["n_el"],
                      "pcg_iterations": main_leg["pcg"], "parallelism": "single GPU"}}
    if "fp64" in legs and "mixed" in legs:
        f = legs["fp64"]
        out["fp64"] = {"value": round(args.steps / f["elapsed"], 4), "pcg_iterations": f["pcg"],
                       "compliance": f["comp"]}
        out["speedup_vs_fp64"] = round(f["elapsed"] / main_leg["elapsed"], 3)
        out["compliance_rel_diff"] = abs(main_leg["comp"] - f["comp"]) / abs(f["comp"])
    if main_leg["prof"] is not None:
        out["roofline"] = roofline(main_leg["prof"], main_leg["n_dof"], main_leg["n_el"], args.workload)
    if main_leg.get("classes") is not None:
        out["kernel_classes_ms_one_step"] = {k: round(v[0], 2) for k, v in main_leg["classes"].items()}
    if args.validate > 0 and "fp64" in legs and "mixed" in legs:
        # validity gate: both precisions from the same start for `validate` OC iterations
        comp = {}
        for leg in ("mixed", "fp64"):
            print(f"[bench] validity run {leg}: {args.validate} OC iterations", file=sys.stderr, flush=True)
            comp[leg] = run_leg(ET, wl, leg, args.validate, 0)["comp"]
        rd = abs(comp["mixed"] - comp["fp64"]) / abs(comp["fp64"])
        out["validity"] = {"oc_iterations": args.validate, "compliance_mixed": comp["mixed"],
                           "compliance_fp64": comp["fp64"], "rel_diff": rd, "gate": 1e-6,
                           "pass": bool(rd <= 1e-6)}
    if not args.no_cpu_baseline:
        print("[bench] cpu baseline (oracle sample)", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(wl["nel"])
    print(json.dumps(out))


if __name__ == "__main__":
    main()
