"""Host-side checks of the C ABI (no GPU needed): the shared library loads and
exports every entry point declared in include/*.h, the ctypes table matches,
and argument validation fails cleanly without touching a device."""
import ctypes
import glob
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        src = open(h).read()
        names |= set(re.findall(r"^(?:int|int64_t|const char\*)\s+(et_\w+)\(", src, re.M))
    return names


def test_library_exports_every_declared_symbol():
    from edgeto_amd import _lib
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in sorted(_declared()) if not hasattr(lib, n)]
    assert not missing, f"declared but not exported: {missing}"
    assert set(_lib.EXPORTED) == _declared(), "ctypes table out of sync with the header"


def test_version_and_error_channel():
    from edgeto_amd import _lib
    lib = _lib.load(require_gpu=False)
    assert b"gfx950" in lib.et_version()
    # invalid dimension: rejected before any device work
    st = lib.et_generate_elements(4, _lib.nel_arg((2, 2, 2)), None, None)
    assert st == _lib.ET_EINVAL
    assert b"dim" in lib.et_last_error()
    st = lib.et_matvec_f64(3, _lib.nel_arg((0, 2, 2)), None, None, None, None, None, None)
    assert st == _lib.ET_EINVAL


def test_mg_create_rejects_bad_params():
    from edgeto_amd import _lib
    lib = _lib.load(require_gpu=False)
    p = _lib.MGParams()
    p.dim, p.n_level, p.n_smooth, p.precision = 3, 9, 2, 0
    for i, v in enumerate((8, 8, 8)):
        p.nel[i] = v
    K0 = np.zeros(576)
    h = ctypes.c_void_p()
    st = lib.et_mg_create(ctypes.byref(h), ctypes.byref(p), K0.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), None)
    assert st == _lib.ET_EINVAL and not h.value
    assert b"n_level" in lib.et_last_error()


def test_product_filter_offsets_match_oracle():
    from edgeto_amd.geom import _offsets_weights
    from oracle import filters
    for rmin, sp in ((1.5, (1, 1)), (2.0, (1, 1, 1)), (2.7, (1.0, 0.5, 2.0)), (1.2, (1, 1, 1))):
        a = _offsets_weights(rmin, sp)
        b = filters.offsets_weights(rmin, sp)
        assert np.array_equal(a[0], b[0]) and np.allclose(a[1], b[1], rtol=0, atol=1e-16)



def test_torch_ops_registered_with_shape_functions():
    """The hot-path ops are torch custom operators; their shape functions run under fake
    tensors (no device work), as tracing needs."""
    import torch
    import edgeto_amd  # noqa: F401 (registers the ops)
    from torch._subclasses.fake_tensor import FakeTensorMode
    for name in ("matvec", "filter_fwd", "filter_adj", "simp_adjoint", "oc_update"):
        assert hasattr(torch.ops.edgeto_amd, name)
    with FakeTensorMode():
        u = torch.empty(81, dtype=torch.float64)
        K0 = torch.empty(576, dtype=torch.float64)
        E = torch.empty(8, dtype=torch.float64)
        y = torch.ops.edgeto_amd.matvec(K0, E, None, u, [2, 2, 2])
        assert y.shape == u.shape and y.dtype == torch.float64
        x = torch.empty(8, dtype=torch.float64)
        z = torch.ops.edgeto_amd.oc_update(x, x, 1.0, 8.0, 0.3, 0.2)
        assert z.shape == x.shape
