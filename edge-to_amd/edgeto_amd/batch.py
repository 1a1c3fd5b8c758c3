"""Batched SIMP optimisation: many independent structured problems on one GPU.

The reference optimises one problem per CuTopOpt instance (TopOpt/_cuda.py:304-382);
BatchTopOpt takes a list of problem descriptors (mesh size, clamped nodes, loads,
volume fraction, filter radius, penalty) and runs all of them concurrently in the
library's batched engine (et_batch_*): every problem's state and GMG-PCG solver
stay resident on the GPU, host worker threads drive the problems on their own
streams, and the work between two solves of a problem is one replayed graph.

Per-problem choices made here:
  * GMG depth (gmg_levels): halve while every axis is even and the coarsest grid
    is above `coarse_dof` DOFs; an odd axis stops the even halving, and if the
    coarsest grid is then too large for the dense coarse solve the remaining
    levels truncate (nel // 2, the reference's own coarsening,
    solvers/_multigrid_cuda.py:145);
  * transformed-basis tiles of the 3-D sweeps and the number of worker threads,
    from the autotuned table batch_tiles.json next to this file (autotune.py).
"""
import ctypes
import json
import os

import numpy as np

from . import _lib
from .geom import _offsets_weights
from .physics import auto_stiffness

_HERE = os.path.dirname(os.path.abspath(__file__))
TILE_TABLE = os.path.join(_HERE, "batch_tiles.json")


class BatchProblemC(ctypes.Structure):
    _fields_ = [("dim", ctypes.c_int), ("nel", ctypes.c_int * 3),
                ("K0", ctypes.c_void_p), ("cons", ctypes.c_void_p), ("f", ctypes.c_void_p),
                ("x0", ctypes.c_void_p), ("n_off", ctypes.c_int), ("offsets", ctypes.c_void_p),
                ("weights", ctypes.c_void_p), ("A", ctypes.c_double), ("volfrac", ctypes.c_double),
                ("penal", ctypes.c_double), ("void_", ctypes.c_double), ("move", ctypes.c_double),
                ("tol", ctypes.c_double), ("maxiter", ctypes.c_int), ("n_level", ctypes.c_int),
                ("n_smooth", ctypes.c_int), ("tile_threads_f32", ctypes.c_int),
                ("tile_threads_f64", ctypes.c_int), ("tile_slices_f32", ctypes.c_int),
                ("tile_slices_f64", ctypes.c_int), ("penal_schedule", ctypes.c_void_p),
                ("n_schedule", ctypes.c_int), ("ch_tol", ctypes.c_double), ("fun_tol", ctypes.c_double)]


def gmg_levels(nel, coarse_dof=2000, max_coarse_dof=6000, max_levels=6):
    """Coarsenings for a structured grid of `nel` elements (see the module docstring)."""
    nel = [int(n) for n in nel]
    dim = len(nel)

    def dofs(counts):
        return int(np.prod([c + 1 for c in counts])) * dim

    levels, cur = 0, list(nel)
    while levels < max_levels and dofs(cur) > coarse_dof:
        nxt = [c // 2 for c in cur]
        if min(nxt) < 2:
            break
        even = all(c % 2 == 0 for c in cur)
        if not even and dofs(cur) <= max_coarse_dof:
            break
        cur, levels = nxt, levels + 1
    return max(levels, 1)


def size_class(dim, nel):
    """Key of the tile table: 2-D, or 3-D by the z extent the transformed-basis tile covers."""
    if dim == 2:
        return "2d"
    return "3d_nz<=64" if int(nel[2]) <= 64 else "3d_nz<=128"


def load_tiles(path=TILE_TABLE):
    try:
        with open(path) as fh:
            return json.load(fh)
    except (OSError, ValueError):
        return {}


class BatchTopOpt:
    """problems: list of dicts with keys dim, nel, loads (node ids), dirs (load vectors, one
    row per load; a node listed twice keeps its last vector), vf, rmin, the supports as
    clamp (node ids, every DOF fixed) and / or fix_nodes with fix_flags (node ids and their
    per-axis 0/1 flags, one row per node or one row for all), and optionally penal (3),
    penal_schedule (penalty per OC iteration, the last one repeating; the final value
    should be penal, which the stopping test waits for), ch_tol and fun_tol (stop once the
    RMS design change and the relative compliance change are below both, at the final
    penalty; default 0: run every iteration), move (0.2), x0 (scalar or array, default vf),
    n_level, nu (0.3), h (element edge length, one value or one per axis; default 1: the
    element matrix, the element measure and the filter's spacing follow it)."""

    def __init__(self, problems, precision="mixed", tol=1e-8, maxiter=1000, n_smooth=2, void=1e-9,
                 tiles=None):
        if precision not in ("fp64", "mixed"):
            raise ValueError("precision must be 'fp64' or 'mixed'")
        self.problems = list(problems)
        self.tiles = load_tiles() if tiles is None else tiles
        self._keep = []
        arr = (BatchProblemC * len(self.problems))()
        self.n_el, self.n_dof, self.levels = [], [], []
        for i, pb in enumerate(self.problems):
            dim, nel = int(pb["dim"]), tuple(int(v) for v in pb["nel"])
            corners = np.array([[0, 0], [1, 0], [1, 1], [0, 1]], dtype=np.float64) if dim == 2 else \
                np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]],
                         dtype=np.float64)
            h = np.broadcast_to(np.asarray(pb.get("h", 1.0), dtype=np.float64), (dim,))
            K, _, _, A = auto_stiffness(corners * h[None, :], 1.0, float(pb.get("nu", 0.3)))
            n_el = int(np.prod(nel))
            n_nodes = int(np.prod([n + 1 for n in nel]))
            n_dof = n_nodes * dim
            cons = np.zeros(n_dof, dtype=np.uint8)
            clamp = np.asarray(pb.get("clamp", []), dtype=np.int64).reshape(-1)
            for d in range(dim):
                cons[clamp * dim + d] = 1
            if "fix_nodes" in pb:
                ids = np.asarray(pb["fix_nodes"], dtype=np.int64).reshape(-1)
                flags = np.asarray(pb["fix_flags"]).astype(bool).reshape(-1, dim)
                flags = np.broadcast_to(flags, (len(ids), dim))
                for d in range(dim):
                    cons[ids[flags[:, d]] * dim + d] = 1
            f = np.zeros(n_dof)
            for node, vec in zip(np.asarray(pb["loads"], dtype=np.int64), np.asarray(pb["dirs"], dtype=np.float64)):
                f[node * dim:(node + 1) * dim] = vec
            x0 = np.broadcast_to(np.asarray(pb.get("x0", pb["vf"]), dtype=np.float64), (n_el,)).copy()
            offs, w = _offsets_weights(float(pb["rmin"]), [float(v) for v in h])
            K = np.ascontiguousarray(K, dtype=np.float64)
            offs = np.ascontiguousarray(offs, dtype=np.int32)
            w = np.ascontiguousarray(w, dtype=np.float64)
            sched = np.ascontiguousarray(np.asarray(pb.get("penal_schedule", []), dtype=np.float64).reshape(-1))
            self._keep += [K, cons, f, x0, offs, w, sched]
            n_level = int(pb.get("n_level") or gmg_levels(nel))
            t = self.tiles.get(size_class(dim, nel), {})
            c = arr[i]
            c.dim = dim
            for a in range(3):
                c.nel[a] = nel[a] if a < dim else 0
            c.K0, c.cons, c.f, c.x0 = (K.ctypes.data, cons.ctypes.data, f.ctypes.data, x0.ctypes.data)
            c.n_off, c.offsets, c.weights = len(w), offs.ctypes.data, w.ctypes.data
            c.A, c.volfrac, c.penal = float(A), float(pb["vf"]), float(pb.get("penal", 3.0))
            c.void_, c.move, c.tol = float(void), float(pb.get("move", 0.2)), float(tol)
            c.maxiter, c.n_level, c.n_smooth = int(maxiter), n_level, int(n_smooth)
            c.tile_threads_f32 = int(t.get("threads_f32", 0))
            c.tile_threads_f64 = int(t.get("threads_f64", 0))
            c.tile_slices_f32 = int(t.get("slices_f32", 0))
            c.tile_slices_f64 = int(t.get("slices_f64", 0))
            c.penal_schedule = sched.ctypes.data if sched.size else None
            c.n_schedule = int(sched.size)
            c.ch_tol, c.fun_tol = float(pb.get("ch_tol", 0.0)), float(pb.get("fun_tol", 0.0))
            self.n_el.append(n_el)
            self.n_dof.append(n_dof)
            self.levels.append(n_level)
        h = ctypes.c_void_p()
        _lib.call("et_batch_create", ctypes.byref(h), len(self.problems), ctypes.cast(arr, ctypes.c_void_p),
                  1 if precision == "mixed" else 0)
        self._handle = h
        self.last_iters = 0

    def __del__(self):
        h = getattr(self, "_handle", None)
        if h is not None and h.value and _lib._lib is not None:
            _lib._lib.et_batch_destroy(h)
            self._handle = None

    def run(self, n_iter, n_workers=None):
        """Initial solve plus up to n_iter OC iterations of every problem.  Returns a dict with
        the wall time and, per problem, the OC iterations run, whether the stopping test
        ended them, the compliance history (initial solve first), PCG iteration counts and
        the wall time of each OC iteration."""
        if n_workers is None:
            n_workers = int(self.tiles.get("workers", 8))
        wall = ctypes.c_double(0.0)
        _lib.call("et_batch_run", self._handle, int(n_iter), int(n_workers), ctypes.byref(wall))
        self.last_iters = int(n_iter)
        out = {"wall_ms": wall.value, "n_workers": int(n_workers), "problems": []}
        for i in range(len(self.problems)):
            done, conv = ctypes.c_int(0), ctypes.c_int(0)
            _lib.call("et_batch_status", self._handle, i, ctypes.byref(done), ctypes.byref(conv))
            k = done.value
            comp = np.zeros(n_iter + 1)
            pcg = np.zeros(n_iter + 1, dtype=np.int32)
            ms = np.zeros(max(n_iter, 1))
            _lib.call("et_batch_result", self._handle, i, comp.ctypes.data_as(ctypes.c_void_p),
                      pcg.ctypes.data_as(ctypes.c_void_p), ms.ctypes.data_as(ctypes.c_void_p), None)
            out["problems"].append({"compliance": comp[:k + 1], "pcg": pcg[:k + 1], "iter_ms": ms[:k],
                                    "iterations": k, "converged": bool(conv.value),
                                    "n_level": self.levels[i]})
        return out

    def design(self, i):
        x = np.zeros(self.n_el[i])
        _lib.call("et_batch_result", self._handle, i, None, None, None, x.ctypes.data_as(ctypes.c_void_p))
        return x


def random_problem(i, seed_base=1000):
    """Problem i of the seeded batch: default_rng(seed_base + i).  2-D (Q4) or 3-D (H8) with
    probability 1/2; 2-D nx in [64, 512], ny in [32, 256]; 3-D nx in [32, 96], ny = nz = nx / 2
    (nx even); one of the four sides x = 0, x = L, y = 0, y = H clamped (every DOF of its
    nodes); 1-4 unit point loads in random directions on boundary nodes off the clamp;
    volume fraction ~ U(0.2, 0.6), filter radius ~ U(1.2, 3.0), penalty 3."""
    rng = np.random.default_rng(seed_base + i)
    dim = 2 if rng.random() < 0.5 else 3
    if dim == 2:
        nel = (int(rng.integers(64, 513)), int(rng.integers(32, 257)))
    else:
        nx = int(rng.integers(32, 97)) // 2 * 2
        nel = (nx, nx // 2, nx // 2)
    side = int(rng.integers(0, 4))
    n_loads = int(rng.integers(1, 5))
    vf = float(rng.uniform(0.2, 0.6))
    rmin = float(rng.uniform(1.2, 3.0))
    nn = [n + 1 for n in nel]
    if dim == 2:
        j, i_ = (a.ravel() for a in np.indices((nn[1], nn[0])))
        k = np.zeros_like(i_)
        node = j * nn[0] + i_
    else:
        j, i_, k = (a.ravel() for a in np.indices((nn[1], nn[0], nn[2])))
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
    return dict(dim=dim, nel=nel, clamp=clamp, loads=loads, dirs=dirs, vf=vf, rmin=rmin, penal=3.0)
