"""Configuration c1 (2-D MBB half-beam 120 x 40 Q4, 50 OC iterations) on the oracle in both
reference semantics: CPU MultiGrid (solvers/_multigrid.py: omega kept across solves, 2-D
restriction R/2, core/_mgm.py:158-160) and GPU CuMultiGrid (solvers/_multigrid_cuda.py: omega
reset per solve, R).  The preconditioners differ, the solves do not: both converge to
|r|/|b| < 1e-8, so the compliance histories agree to solve accuracy while the PCG iteration
counts differ."""
import numpy as np

from oracle import topopt as otop
from oracle.solvers import MultiGrid as OMultiGrid


def _c1(mode):
    nelx, nely = 120, 40
    prob = otop.Problem((nelx, nely), nu=0.3, r_min=1.5)
    nx1 = nelx + 1
    prob.cons[np.arange(nely + 1) * nx1 * 2] = True
    prob.cons[nelx * 2 + 1] = True
    prob.f[nely * nx1 * 2 + 1] = -1.0
    mat = otop.SingleMaterial(void=1e-9, penalty=3.0, volume_fraction=0.5)
    sol = OMultiGrid(prob.nel, 2, tol=1e-8, n_smooth=3, omega=0.5, n_level=3, mode=mode)
    x, _, hist = otop.optimize(prob, mat, sol, np.full(prob.n_el, 0.5), 50, ch_tol=0, fun_tol=0)
    return x, hist


def test_c1_cpu_and_gpu_semantics():
    xc, hc = _c1("cpu")
    xg, hg = _c1("gpu")
    cc, cg = np.array(hc["comp"]), np.array(hg["comp"])
    assert cc.shape == cg.shape == (51,)
    assert np.all(np.array(hc["res"]) < 1e-8) and np.all(np.array(hg["res"]) < 1e-8)
    assert np.max(np.abs(cc - cg) / cg) < 1e-6
    assert np.max(np.abs(xc - xg)) < 1e-4
    assert hc["pcg"] != hg["pcg"]  # the two preconditioners are different
    assert abs(xc.mean() - 0.5) < 1e-3
    assert cc[-1] < 0.5 * cc[0]
