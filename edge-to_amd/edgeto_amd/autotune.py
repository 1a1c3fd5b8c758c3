"""Autotuner of the batched engine's launch choices; writes batch_tiles.json.

Two choices are measured on the GPU, on problems of the seeded batch
(batch.random_problem):
  * per 3-D size class, the transformed-basis tile of the FP32 fine-grid sweeps
    (workgroup size and z slices per layer): each candidate runs the class's
    problems for a few OC iterations on one worker, and the fastest median OC
    iteration wins;
  * the number of host worker threads (concurrent problems): the first problems
    of the batch run with each candidate count, and the highest aggregate
    OC-iteration rate wins.
The table records the choice and the measured times of every candidate.

Usage (on a GPU host):  python -m edgeto_amd.autotune [--problems 16] [--iters 2] [--out PATH]
"""
import argparse
import json
import sys

import numpy as np

from .batch import TILE_TABLE, BatchTopOpt, random_problem, size_class

TILE_CANDIDATES = [
    {"threads_f32": 1024, "slices_f32": 16},
    {"threads_f32": 512, "slices_f32": 16},
    {"threads_f32": 512, "slices_f32": 8},
    {"threads_f32": 256, "slices_f32": 8},
]
WORKER_CANDIDATES = [2, 4, 8, 16]


def _median_iter_ms(out):
    return float(np.median(np.concatenate([p["iter_ms"] for p in out["problems"]])))


def tune(n_problems=16, iters=2, out_path=TILE_TABLE, log=sys.stderr):
    problems = [random_problem(i) for i in range(n_problems)]
    table = {"note": "written by edgeto_amd.autotune on a GPU host; times in ms", "candidates": {}}
    classes = sorted({size_class(p["dim"], p["nel"]) for p in problems if p["dim"] == 3})
    for cls in classes:
        members = [p for p in problems if p["dim"] == 3 and size_class(3, p["nel"]) == cls][:3]
        times = []
        for cand in TILE_CANDIDATES:
            b = BatchTopOpt(members, precision="mixed", tiles={cls: cand})
            ms = _median_iter_ms(b.run(iters, n_workers=1))
            del b
            times.append(ms)
            print(f"[autotune] {cls} {cand}: median {ms:.2f} ms per OC iteration", file=log, flush=True)
        best = int(np.argmin(times))
        table[cls] = dict(TILE_CANDIDATES[best])
        table["candidates"][cls] = [dict(c, median_ms=round(t, 3)) for c, t in zip(TILE_CANDIDATES, times)]
    rates = []
    for w in WORKER_CANDIDATES:
        b = BatchTopOpt(problems, precision="mixed", tiles=table)
        out = b.run(iters, n_workers=w)
        del b
        rate = len(problems) * iters / (out["wall_ms"] * 1e-3)
        rates.append(rate)
        print(f"[autotune] workers {w}: {rate:.2f} OC iterations/s", file=log, flush=True)
    best = int(np.argmax(rates))
    table["workers"] = WORKER_CANDIDATES[best]
    table["candidates"]["workers"] = [{"workers": w, "iters_per_s": round(r, 3)}
                                      for w, r in zip(WORKER_CANDIDATES, rates)]
    with open(out_path, "w") as fh:
        json.dump(table, fh, indent=1)
    return table


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--problems", type=int, default=16)
    ap.add_argument("--iters", type=int, default=2)
    ap.add_argument("--out", default=TILE_TABLE)
    a = ap.parse_args()
    print(json.dumps(tune(a.problems, a.iters, a.out)))


if __name__ == "__main__":
    main()
