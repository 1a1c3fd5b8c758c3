"""Host-side logic of the package that needs no GPU: the batched engine's per-problem GMG
depth rule and the seeded problem generator."""
import numpy as np
import pytest


def _batch():
    from edgeto_amd import batch
    return batch


def test_gmg_levels_rule():
    """Depth per problem: even halving while the coarsest grid is above 2000 DOFs; an odd axis
    stops it unless the coarsest grid would stay above 6000 DOFs (then truncating levels,
    the reference's nel // 2)."""
    gmg_levels = _batch().gmg_levels
    assert gmg_levels((256, 128, 128)) == 5       # 8 x 4 x 4: 675 DOFs
    assert gmg_levels((96, 48, 48)) == 3          # 12 x 6 x 6: 1911 DOFs
    assert gmg_levels((512, 256)) == 4            # 32 x 16: 1122 DOFs
    assert gmg_levels((64, 32)) == 1              # 32 x 16: 1122 DOFs
    assert gmg_levels((33, 17)) == 1              # odd at once, small: one level
    assert gmg_levels((325, 89)) == 2             # odd and large: truncating levels
    for nel in ((86, 43, 43), (494, 241), (40, 20, 20)):
        lv = gmg_levels(nel)
        coarse = [n // 2 ** lv for n in nel]
        assert min(coarse) >= 2
        assert np.prod([c + 1 for c in coarse]) * len(nel) <= 6000


def test_random_problem_is_seeded_and_well_formed():
    b = _batch()
    for i in range(16):
        p, q = b.random_problem(i), b.random_problem(i)
        assert p["nel"] == q["nel"] and np.array_equal(p["loads"], q["loads"])
        dim, nel = p["dim"], p["nel"]
        if dim == 2:
            assert 64 <= nel[0] <= 512 and 32 <= nel[1] <= 256
        else:
            assert 32 <= nel[0] <= 96 and nel[1] == nel[2] == nel[0] // 2
        assert 0.2 <= p["vf"] <= 0.6 and 1.2 <= p["rmin"] <= 3.0
        assert 1 <= len(p["loads"]) <= 4
        assert not np.intersect1d(p["loads"], p["clamp"]).size
        assert np.allclose(np.linalg.norm(p["dirs"], axis=1), 1.0)


def test_multigrid_argument_errors_in_reference_order():
    """CuMultiGrid's argument checks run before any device work and in the reference's order
    (solvers/_multigrid_cuda.py:44-54): matrix_free with a direct coarse solver is a
    ValueError, an unknown coarse solver a ValueError, matrix_free otherwise not implemented."""
    from edgeto_amd.solvers import MultiGrid
    with pytest.raises(ValueError, match="splu"):
        MultiGrid(None, None, matrix_free=True)
    with pytest.raises(ValueError, match="not recognized"):
        MultiGrid(None, None, coarse_solver="lu")
    with pytest.raises(NotImplementedError):
        MultiGrid(None, None, coarse_solver="cg", matrix_free=True)


def test_grid_size_limit_is_an_error_before_any_device_access():
    """A grid whose DOF count does not fit 32-bit indices (2000^3 elements) is rejected with
    ET_EINVAL by every grid entry point before it touches memory (host-only: dummy pointers)."""
    import ctypes
    from edgeto_amd import _lib
    L = _lib.lib()
    nel = (ctypes.c_int * 3)(2000, 2000, 2000)
    p = ctypes.cast((ctypes.c_double * 8)(), ctypes.c_void_p)
    q = ctypes.cast((ctypes.c_double * 8)(), ctypes.c_void_p)
    calls = {
        "et_matvec_f64": lambda: L.et_matvec_f64(3, nel, p, p, None, p, q, None),
        "et_residual_f64": lambda: L.et_residual_f64(3, nel, p, p, None, p, p, q, None),
        "et_jacobi_f64": lambda: L.et_jacobi_f64(3, nel, p, p, None, p, p, p, 0.6, q, None),
        "et_diagonal_f64": lambda: L.et_diagonal_f64(3, nel, p, p, None, q, None),
        "et_elem_energy_f64": lambda: L.et_elem_energy_f64(3, nel, p, p, q, None),
        "et_galerkin_f64": lambda: L.et_galerkin_f64(3, nel, 3, p, q, None),
        "et_generate_elements": lambda: L.et_generate_elements(3, nel, p, None),
    }
    for name, call in calls.items():
        assert call() == 1, name
        assert b"32-bit" in L.et_last_error(), (name, L.et_last_error())
