"""The allocation-free band compaction (dcl_band_compact_static) against
dcl_band_compact through the C ABI: two calls in a row on one GPU, the second
on a smaller window (the static scratch keeps the first call's counts), both
bit-exact with the allocating entry point, in FP64 and FP32; a window past
its 2^28-node limit is an error before any GPU work."""
import ctypes as C

import numpy as np
import pytest
import torch

import dcell_amd as D
from dcell_amd import _lib as L
from dcell_amd import configs
from dcell_amd import metrics as M
from dcell_amd.levelset import _p, stream_handle

pytestmark = pytest.mark.gpu


def _compact(fn, phi, win, width, cap):
    out = torch.full((cap, 3), -7, dtype=torch.int32, device="cuda:0")
    count = C.c_int64()
    prec = L.DCL_F64 if phi.dtype == torch.float64 else L.DCL_F32
    L.call(fn, _p(phi), C.byref(win.c()), float(width), prec, _p(out), cap, C.byref(count), stream_handle())
    torch.cuda.synchronize()
    return out.cpu().numpy(), count.value


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_band_compact_static_matches_allocating(dev, dtype):
    cases = [(200, (0.4, 0.45, 0.5), 0.3, 6.0), (37, (0.5, 0.5, 0.5), 0.25, 3.0)]
    for n, c, r, width in cases:
        win = D.GridWindow((3, -2, 5), (n, n, n), (1.0 / (n - 1),) * 3)
        phi = torch.from_numpy(configs.sphere_sdf(n, c, r)).to("cuda:0", dtype)
        ref, nref = _compact("dcl_band_compact", phi, win, width, win.nodes)
        got, ngot = _compact("dcl_band_compact_static", phi, win, width, win.nodes)
        assert nref > 0 and ngot == nref, (n, ngot, nref)
        assert np.array_equal(got, ref), n
        # a capacity below the count: the count is still the total, only the first entries are written
        cap = nref // 3
        got, ngot = _compact("dcl_band_compact_static", phi, win, width, cap)
        assert ngot == nref and np.array_equal(got, ref[:cap]), n
    # through the Python mirror (out= selects the allocation-free entry point)
    out = torch.empty((win.nodes, 3), dtype=torch.int32, device="cuda:0")
    band, cnt = M.band_compact(phi, win, width, out=out)
    torch.cuda.synchronize()
    assert cnt == nref and np.array_equal(band.cpu().numpy(), ref[:nref])


def test_band_compact_static_limit(dev):
    phi = torch.zeros(8, dtype=torch.float64, device="cuda:0")
    win = D.GridWindow((0, 0, 0), (1024, 1024, 257), (1.0, 1.0, 1.0))   # 2^28 + 2^20 nodes
    count = C.c_int64()
    rc = L.lib().dcl_band_compact_static(_p(phi), C.byref(win.c()), 1.0, L.DCL_F64, None, 0, C.byref(count),
                                         stream_handle())
    assert rc == 60 and b"2^28" not in L.lib().dcl_last_error() and b"268435456" in L.lib().dcl_last_error()
