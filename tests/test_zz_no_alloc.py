"""tnlhip_multidot_noalloc (no counterpart in the reference; for callers that cannot allocate):
the same column sums as tnlhip_multidot, bit for bit, through the C ABI, on two calls in a row
(the second smaller, so it reuses the fixed device array the first one filled)."""
import ctypes as C

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def multidot_host(tnl, dtype, n, count, V, ld, w):
    out = np.zeros(count, np.float64)
    st = tnl.lib.tnlhip_multidot(dtype, n, count, V.data_ptr(), ld, w.data_ptr(), out.ctypes.data_as(C.c_void_p),
                                 None)
    assert st == 0, tnl.lib.tnlhip_last_error()
    return out


def multidot_noalloc(tnl, dtype, n, count, V, ld, w, dev):
    res = torch.full((count,), np.nan, dtype=torch.float64, device=dev)
    st = tnl.lib.tnlhip_multidot_noalloc(dtype, n, count, V.data_ptr(), ld, w.data_ptr(), res.data_ptr(), None)
    assert st == 0, tnl.lib.tnlhip_last_error()
    torch.cuda.synchronize()
    return res.cpu().numpy()


@pytest.mark.parametrize("dt", ["f64", "f32"])
def test_noalloc_multidot_matches_multidot(tnl, dev, dt):
    torch_dt = torch.float64 if dt == "f64" else torch.float32
    dtype = 0 if dt == "f64" else 1
    g = torch.Generator(device="cpu").manual_seed(11)
    # (n, count): a full 32-column pass over 3 M entries, then a smaller call of 5 columns
    for n, count in ((3_000_001, 32), (70_001, 5)):
        ld = (n + 63) // 64 * 64
        V = (torch.rand(ld * count, generator=g, dtype=torch.float64) * 2 - 1).to(torch_dt).to(dev)
        w = (torch.rand(n, generator=g, dtype=torch.float64) * 2 - 1).to(torch_dt).to(dev)
        want = multidot_host(tnl, dtype, n, count, V, ld, w)
        got = multidot_noalloc(tnl, dtype, n, count, V, ld, w, dev)
        print(dt, n, count, got[:3], want[:3], flush=True)
        assert np.array_equal(got, want), (n, count)
        # against a float64 restatement of the column sums
        Vh = V.cpu().double().numpy().reshape(count, ld)[:, :n]
        ref = Vh @ w.cpu().double().numpy()
        assert np.allclose(got, ref, rtol=1e-9, atol=1e-9 * np.sqrt(n))


def test_noalloc_multidot_refuses_more_columns_than_its_array(tnl, dev):
    V = torch.zeros(64 * 40, dtype=torch.float64, device=dev)
    w = torch.zeros(64, dtype=torch.float64, device=dev)
    res = torch.zeros(40, dtype=torch.float64, device=dev)
    for count in (0, 33, 40):
        st = tnl.lib.tnlhip_multidot_noalloc(0, 64, count, V.data_ptr(), 64, w.data_ptr(), res.data_ptr(), None)
        assert st == 1
        assert b"1 to 32 columns" in tnl.lib.tnlhip_last_error()
