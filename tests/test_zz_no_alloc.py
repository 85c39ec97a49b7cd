"""pseg_spectrum_average_noalloc (the entry point that allocates nothing: its
partial sums live in a static device array) against pseg_spectrum_average
(the same kernels with a caller workspace), through the C ABI: the same bits
on two calls in a row, the second with fewer rows, another window and a
narrower panel, so stale partials of the first call would show."""
import ctypes as C

import numpy as np
import pytest

import processseg_amd as P
from processseg_amd import _lib

pytestmark = pytest.mark.gpu


def _both(sp, src, rl, first):
    import torch
    L = _lib.lib()
    rows, cols = src.shape
    stream = torch.cuda.current_stream().cuda_stream
    need = C.c_size_t()
    assert L.pseg_spectrum_workspace_size(sp._h, rows, C.byref(need)) == _lib.OK
    ws = torch.empty(max(need.value, 1), dtype=torch.uint8, device="cuda")
    a = torch.full((sp.bins,), -1.0, dtype=torch.float64, device="cuda")
    ua = torch.full((1,), -1.0, dtype=torch.float64, device="cuda")
    b = torch.full((sp.bins,), -2.0, dtype=torch.float64, device="cuda")
    ub = torch.full((1,), -2.0, dtype=torch.float64, device="cuda")
    assert L.pseg_spectrum_average(sp._h, src.data_ptr(), rows, cols, rl.data_ptr(), first,
                                   ws.data_ptr(), need.value, a.data_ptr(), ua.data_ptr(),
                                   stream) == _lib.OK
    assert L.pseg_spectrum_average_noalloc(sp._h, src.data_ptr(), rows, cols, rl.data_ptr(),
                                           first, b.data_ptr(), ub.data_ptr(),
                                           stream) == _lib.OK
    torch.cuda.synchronize()
    return a.cpu().numpy(), ua.item(), b.cpu().numpy(), ub.item()


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_noalloc_equals_workspace_form(gpu, dtype):
    import torch
    from processseg_amd.workloads import big_panel
    x, live = big_panel(2000, 3001, dtype)
    sp = P.initSpectrumPlan(dtype, 2.0, 2001, 0.1)
    src = torch.from_numpy(x).cuda()
    rl = torch.from_numpy(live.astype(np.int32)).cuda()
    print(f"noalloc spectrum {np.dtype(dtype).name}: 2000 x 3001 from column 500", flush=True)
    a, ua, b, ub = _both(sp, src, rl, 500)
    assert ua == ub == 2000
    assert np.array_equal(a.view(np.uint64), b.view(np.uint64))
    assert np.all(np.isfinite(a)) and a.max() > 0
    # second call: fewer rows, a narrower panel, another window, a dead row
    small = torch.from_numpy(np.ascontiguousarray(x[:37, :2500])).cuda()
    rl2 = np.minimum(live[:37], 2500).astype(np.int32)
    rl2[5] = -1
    print(f"noalloc spectrum {np.dtype(dtype).name}: 37 x 2500 from column 100", flush=True)
    a2, ua2, b2, ub2 = _both(sp, small, torch.from_numpy(rl2).cuda(), 100)
    assert ua2 == ub2 == 36
    assert np.array_equal(a2.view(np.uint64), b2.view(np.uint64))
    assert not np.array_equal(a2, a)


def test_noalloc_no_rows_and_refusals(gpu):
    import torch
    L = _lib.lib()
    sp = P.initSpectrumPlan(np.float32, 2.0, 500, 0.1)
    amp = torch.full((sp.bins,), 5.0, dtype=torch.float64, device="cuda")
    used = torch.full((1,), 5.0, dtype=torch.float64, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    assert L.pseg_spectrum_average_noalloc(sp._h, None, 0, 0, None, 0, amp.data_ptr(),
                                           used.data_ptr(), stream) == _lib.OK
    torch.cuda.synchronize()
    assert used.item() == 0 and torch.all(amp == 0)
    assert L.pseg_spectrum_average_noalloc(None, None, 1, 1, None, 0, amp.data_ptr(), None,
                                           stream) == _lib.E_INVALID
    assert L.pseg_spectrum_average_noalloc(sp._h, None, 1, 1, None, -1, amp.data_ptr(), None,
                                           stream) == _lib.E_INVALID
    assert L.pseg_spectrum_average_noalloc(sp._h, None, 1, 1, None, 0, None, None,
                                           stream) == _lib.E_INVALID
