"""The allocation-free sector binning entry point, rick_sector_bin_noalloc,
against rick_sector_bin with a caller workspace: bit-exact histograms, list
offsets, row lists, totals and dropped counts, on two calls in a row (the
second smaller, so stale scratch contents from the first would show), plus
the size limit of its fixed workspace.  All calls go through the C ABI."""
import ctypes
import sys
from pathlib import Path

import numpy as np
import pytest
import torch

sys.path.insert(0, str(Path(__file__).resolve().parents[1]))

from rick_amd import _lib  # noqa: E402
from rick_amd._lib import RICK_ERR_INVALID, RICK_OK, lib  # noqa: E402

pytestmark = pytest.mark.gpu


def _vp(t):
    return ctypes.c_void_p(t.data_ptr())


def _bin(vv, S, yaxis, dx, sup, noalloc):
    n = int(vv.numel())
    cap = max(1, 3 * n)
    histo = torch.full((S + 1,), -1, dtype=torch.int32, device="cuda")
    start = torch.full((S + 1,), -1, dtype=torch.int32, device="cuda")
    rows = torch.full((cap,), -1, dtype=torch.int32, device="cuda")
    total, dropped = ctypes.c_ulong(), ctypes.c_uint()
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    if noalloc:
        st = lib().rick_sector_bin_noalloc(n, _vp(vv), S, yaxis, dx, sup, _vp(histo), _vp(start),
                                           _vp(rows), cap, stream, ctypes.byref(total),
                                           ctypes.byref(dropped))
    else:
        ws = lib().rick_sector_workspace_bytes(n, S)
        work = torch.empty(ws, dtype=torch.uint8, device="cuda")
        st = lib().rick_sector_bin(n, _vp(vv), S, yaxis, dx, sup, _vp(histo), _vp(start),
                                   _vp(rows), cap, _vp(work), stream, ctypes.byref(total),
                                   ctypes.byref(dropped))
    torch.cuda.synchronize()
    assert st == RICK_OK, (lib().rick_last_error() or b"").decode()
    return (histo.cpu().numpy(), start.cpu().numpy(), rows[:total.value].cpu().numpy(),
            int(total.value), int(dropped.value))


def _rows(n, seed):
    rng = np.random.default_rng(seed)
    v = rng.uniform(-0.001, 1.001, n)        # rows near and past both borders
    v[::997] = 1.6                           # slab index past nsectors: dropped entries
    v[5::1009] = -0.7                        # negative slab index: dropped entries
    return torch.from_numpy(v).to("cuda")


@pytest.mark.parametrize("calls", [[(1_000_003, 8, 7), (120_001, 3, 9)],
                                   [(400_000, 64, 7), (5, 1, 7)]])
def test_noalloc_sector_bin_matches_workspace_entry(calls):
    for n, S, sup in calls:
        print(f"case sector_bin_noalloc n={n} S={S} support={sup}", flush=True)
        vv = _rows(n, n + S)
        yaxis = max(1, 4096 // S)
        ref = _bin(vv, S, yaxis, 1 / 4096, sup, noalloc=False)
        got = _bin(vv, S, yaxis, 1 / 4096, sup, noalloc=True)
        for a, b in zip(got[:3], ref[:3]):
            np.testing.assert_array_equal(a, b)
        assert got[3:] == ref[3:]
        assert ref[4] > 0 or n < 6


def test_noalloc_sector_bin_rejects_oversized_workspace():
    cap = lib().rick_sector_bin_noalloc_max_workspace()
    assert cap == 4 << 20
    n, S = 8_000_000, 400      # within the scan's capacity, above the fixed workspace
    assert lib().rick_sector_workspace_bytes(n, S) > cap
    vv = torch.zeros(1, dtype=torch.float64, device="cuda")
    out = torch.zeros(S + 1, dtype=torch.int32, device="cuda")
    total, dropped = ctypes.c_ulong(), ctypes.c_uint()
    # Rejected on the host before anything is launched (vv is never read).
    st = lib().rick_sector_bin_noalloc(n, _vp(vv), S, 8, 1 / 4096, 7, _vp(out), _vp(out),
                                       _vp(out), 1, None, ctypes.byref(total),
                                       ctypes.byref(dropped))
    assert st == RICK_ERR_INVALID
    assert b"allocation-free" in lib().rick_last_error()
    assert _lib.SIGNATURES["rick_sector_bin_noalloc"]
