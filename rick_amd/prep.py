"""Device-side preparation of the rows in front of wstack(), over the C entry
points of librick_hip.so (CUDA tensors, the current torch stream).

* ``normalize_uvw`` -- the NORMALIZE_UVW block of gridding.c:26-82, in place.
* ``sector_bin`` -- initialize_array(), gridding.c:139-197: rows into v-slab
  sectors (one per rank) with halo copies.
* ``sector_gather`` -- gridding_cpu.c:90-138: one sector's rows, in the order
  and with the v shift the reference passes to wstack().
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

from . import _lib
from ._lib import check, lib


def _torch():
    import torch
    return torch


def _stream(device):
    torch = _torch()
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def _need(t, dtype, n, name):
    torch = _torch()
    if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == dtype and t.is_contiguous()
            and t.numel() >= n):
        raise TypeError(f"{name} must be a contiguous CUDA {dtype} tensor of >= {n} elements")
    return ctypes.c_void_p(t.data_ptr())


def normalize_uvw(uu, vv, ww, workspace=None):
    """Normalizes device u, v, w in place; returns the extrema tensor
    (min u, v, w, max u, v, w, maxg, 0)."""
    torch = _torch()
    n = int(uu.numel())
    dev = uu.device
    ws_bytes = lib().rick_normalize_uvw_workspace_bytes()
    if workspace is None:
        workspace = torch.empty(ws_bytes, dtype=torch.uint8, device=dev)
    ext = torch.empty(8, dtype=torch.float64, device=dev)
    check(lib().rick_normalize_uvw(
        n, _need(uu, torch.float64, n, "uu"), _need(vv, torch.float64, n, "vv"),
        _need(ww, torch.float64, n, "ww"), ctypes.c_void_p(ext.data_ptr()),
        _need(workspace, torch.uint8, ws_bytes, "workspace"), _stream(dev)))
    return ext


@dataclass
class SectorBins:
    histo: object    # CUDA uint32-as-int32 tensor [nsectors+1]
    start: object    # offsets of each sector's list in rows
    rows: object     # concatenated row lists
    total: int
    dropped: int
    nsectors: int

    def sector(self, s: int):
        """Row list of sector s (a view into rows)."""
        a = int(self.start[s])
        return self.rows[a:a + int(self.histo[s])]


def sector_bin(vv, nsectors: int, yaxis: int, dx: float, w_support: int) -> SectorBins:
    torch = _torch()
    n = int(vv.numel())
    dev = vv.device
    ws = lib().rick_sector_workspace_bytes(n, nsectors)
    if ws == 0:
        raise ValueError(f"sector binning does not handle {n} rows x {nsectors} sectors")
    work = torch.empty(ws, dtype=torch.uint8, device=dev)
    cap = max(1, 3 * n)
    histo = torch.empty(nsectors + 1, dtype=torch.int32, device=dev)
    start = torch.empty(nsectors + 1, dtype=torch.int32, device=dev)
    rows = torch.empty(cap, dtype=torch.int32, device=dev)
    total = ctypes.c_ulong()
    dropped = ctypes.c_uint()
    check(lib().rick_sector_bin(
        n, _need(vv, torch.float64, n, "vv"), nsectors, yaxis, float(dx), w_support,
        ctypes.c_void_p(histo.data_ptr()), ctypes.c_void_p(start.data_ptr()),
        ctypes.c_void_p(rows.data_ptr()), cap, ctypes.c_void_p(work.data_ptr()), _stream(dev),
        ctypes.byref(total), ctypes.byref(dropped)))
    return SectorBins(histo.cpu(), start.cpu(), rows[:total.value], int(total.value),
                      int(dropped.value), nsectors)


def sector_gather(list_rows, isector, yaxis, dx, freq_per_chan, polarizations, uu, vv, ww,
                  weight, vis_real, vis_img) -> dict:
    """Inputs of wstack() for sector ``isector`` from its row list."""
    torch = _torch()
    n = int(uu.numel())
    nsec = int(list_rows.numel())
    F, P = int(freq_per_chan), int(polarizations)
    dev = uu.device
    out = {
        "uu": torch.empty(nsec, dtype=torch.float64, device=dev),
        "vv": torch.empty(nsec, dtype=torch.float64, device=dev),
        "ww": torch.empty(nsec, dtype=torch.float64, device=dev),
        "weight": torch.empty(nsec * P, dtype=torch.float32, device=dev),
        "vis_real": torch.empty(nsec * F * P, dtype=torch.float32, device=dev),
        "vis_img": torch.empty(nsec * F * P, dtype=torch.float32, device=dev),
    }
    if list_rows.dtype != torch.int32 or not list_rows.is_contiguous() or not list_rows.is_cuda:
        raise TypeError("list_rows must be a contiguous CUDA int32 tensor")
    check(lib().rick_sector_gather(
        ctypes.c_void_p(list_rows.data_ptr()), nsec, isector, yaxis, float(dx), n, F, P,
        _need(uu, torch.float64, n, "uu"), _need(vv, torch.float64, n, "vv"),
        _need(ww, torch.float64, n, "ww"), _need(weight, torch.float32, n * P, "weight"),
        _need(vis_real, torch.float32, n * F * P, "vis_real"),
        _need(vis_img, torch.float32, n * F * P, "vis_img"),
        *(ctypes.c_void_p(out[k].data_ptr()) for k in
          ("uu", "vv", "ww", "weight", "vis_real", "vis_img")), _stream(dev)))
    return out


__all__ = ["normalize_uvw", "sector_bin", "sector_gather", "SectorBins"]
