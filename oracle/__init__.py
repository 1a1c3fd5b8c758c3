"""CPU oracle: a NumPy/SciPy restatement of EDGE-TO's SIMP minimum-compliance path.

TEST INFRASTRUCTURE ONLY. Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this package, and only as the
checker (or, in ``bench.py``, as the timed CPU baseline). The product path in
``edge-to_amd/`` never imports it and has no CPU fallback.

Every function cites the reference file:line it restates (paths relative to
the reference repository root). Functions whose results no closed form,
invariant, brute-force construction or reference fixture pins are marked
"parity unpinned" in their docstring; at present there are none.

Modules
-------
physics   element stiffness K0, constitutive D, averaged B (Q4 plane stress, H8)
mesh      structured connectivity, node coordinates, node->element star
ops       constrained matrix-free matvec, diagonal, element energy, COO assembly
filters   linear-hat density filter, its normaliser and exact transpose
mgm       restriction / prolongation (GPU and CPU-2D variants), Galerkin operators
solvers   GMG-preconditioned CG with the reference's omega schedule
topopt    SingleMaterial, OC bisection, the SIMP loop
"""
