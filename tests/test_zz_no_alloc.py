"""The allocation-free OC entry point (et_oc_update_f64_noalloc) against et_oc_update_f64 on
the same inputs, through the C ABI: identical designs and step counts on two calls in a row,
the second with a smaller input (the shared static work array keeps the first call's
contents), and a clear error for sizes beyond the array."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _case(n, seed):
    rng = np.random.default_rng(seed)
    x = rng.uniform(0.01, 1.0, n)
    dc = -rng.uniform(0.0, 2.0, n) * rng.uniform(0, 1, n) ** 3
    return (torch.tensor(x, device="cuda"), torch.tensor(dc, device="cuda"))


def test_oc_update_noalloc_matches_allocating_entry_point():
    from edgeto_amd import _lib as L
    L.load()
    nmax = int(L.lib().et_oc_noalloc_max_n())
    assert nmax == 1 << 20
    for n, seed in ((nmax, 0), (5000, 1)):
        x, dc = _case(n, seed)
        vol, vf, move = float(n), 0.35, 0.2
        work = torch.empty(int(L.lib().et_oc_work_size(n)) // 8 + 1, dtype=torch.float64, device="cuda")
        ref = torch.empty(n, dtype=torch.float64, device="cuda")
        out = torch.empty(n, dtype=torch.float64, device="cuda")
        s_ref, s_out = ctypes.c_int(-1), ctypes.c_int(-2)
        L.call("et_oc_update_f64", n, L.f64(x), L.f64(dc), 1.0, vol, vf, move, L.f64(ref), L.f64(work),
               ctypes.byref(s_ref), L.stream())
        L.call("et_oc_update_f64_noalloc", n, L.f64(x), L.f64(dc), 1.0, vol, vf, move, L.f64(out),
               ctypes.byref(s_out), L.stream())
        assert s_out.value == s_ref.value > 0
        assert torch.equal(out, ref)
        # asynchronous form of the same call
        out2 = torch.full_like(out, -1.0)
        L.call("et_oc_update_f64_noalloc", n, L.f64(x), L.f64(dc), 1.0, vol, vf, move, L.f64(out2), None,
               L.stream())
        torch.cuda.synchronize()
        assert torch.equal(out2, ref)
    # beyond the static array: rejected on the host, before any launch
    big = nmax + 1
    st = L.lib().et_oc_update_f64_noalloc(big, ctypes.c_void_p(1), ctypes.c_void_p(1), 1.0, float(big), 0.3,
                                          0.2, ctypes.c_void_p(1), None, L.stream())
    assert st == L.ET_EINVAL
    assert b"1048576" in L.lib().et_last_error()
