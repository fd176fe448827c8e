"""binMS measurement sets, raw images and WRITE_DATA cubes (the reference's
on-disk formats), over the C entry points of librick_hip.so.

* ``read_meta`` / ``partition`` / ``read`` -- readMetaData, metaData_calculation
  and readData (init.c:300-333, 335-372, 424-490); ``read`` can target host
  numpy arrays or device tensors (streamed through pinned buffers).
* ``write`` -- the layout scripts/create_binMS.py:210-245 produces.
* ``write_image`` -- WRITE_IMAGE's raw files (fourier_transform.c:280-289).
* ``write_cube`` -- WRITE_DATA's grid cubes (gridding.c:210-282,
  fourier_transform.c:134-201), one rank.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass, fields

import numpy as np

from . import _lib
from ._lib import FP32, FP64, check, lib

FILES = ("ucoord.bin", "vcoord.bin", "wcoord.bin", "weights.bin", "visibilities_real.bin",
         "visibilities_img.bin")
META = "meta.txt"
DROP_ROWS = 1000   # metaData_calculation leaves out the last 1000 rows (init.c:337)


@dataclass
class Meta:
    nmeasures: int
    nvis: int
    freq_per_chan: int
    polarisations: int
    ntimes: int = 0
    dt: float = 0.0
    thours: float = 0.0
    baselines: int = 0
    uvmin: float = 0.0
    uvmax: float = 0.0
    wmin: float = 0.0
    wmax: float = 0.0

    def to_c(self) -> _lib.BinmsMeta:
        return _lib.BinmsMeta(**{f.name: getattr(self, f.name) for f in fields(self)})

    @classmethod
    def from_c(cls, m: _lib.BinmsMeta) -> "Meta":
        return cls(**{f.name: getattr(m, f.name) for f in fields(cls)})

    @property
    def resolution(self) -> float:
        """1/max(|uvmin|, |uvmax|), gridding_cpu.c:78."""
        return 1.0 / max(abs(self.uvmin), abs(self.uvmax))


def _path(p) -> bytes:
    return os.fsencode(str(p))


def read_meta(path) -> Meta:
    m = _lib.BinmsMeta()
    check(lib().rick_binms_read_meta(_path(path), ctypes.byref(m)))
    return Meta.from_c(m)


def partition(meta: Meta, rank: int = 0, size: int = 1, drop: int = DROP_ROWS):
    """(startrow, nrows) of ``rank`` out of ``size`` after dropping ``drop`` rows."""
    start = ctypes.c_ulong()
    nrows = ctypes.c_uint()
    c = meta.to_c()
    check(lib().rick_binms_partition(ctypes.byref(c), rank, size, drop, ctypes.byref(start),
                                     ctypes.byref(nrows)))
    return int(start.value), int(nrows.value)


def _alloc(n, F, P, device):
    shapes = (n, n, n, n * P, n * F * P, n * F * P)
    dts = ("f8", "f8", "f8", "f4", "f4", "f4")
    if device is None:
        return [np.empty(s, dtype=d) for s, d in zip(shapes, dts)]
    import torch
    tdt = {"f8": torch.float64, "f4": torch.float32}
    return [torch.empty(s, dtype=tdt[d], device=f"cuda:{device}") for s, d in zip(shapes, dts)]


def _addr(a):
    return ctypes.c_void_p(a.ctypes.data if isinstance(a, np.ndarray) else a.data_ptr())


def read(path, meta: Meta, startrow: int, nrows: int, device: int | None = None) -> dict:
    """Rows [startrow, startrow+nrows) of a binMS directory as a dict of the
    six arrays; numpy on the host, or CUDA tensors on ``device``."""
    F, P = meta.freq_per_chan, meta.polarisations
    arrs = _alloc(nrows, F, P, device)
    if device is not None:
        import torch
        torch.cuda.synchronize(device)
    check(lib().rick_binms_read(_path(path), F, P, startrow, nrows, *map(_addr, arrs),
                                -1 if device is None else int(device)))
    return dict(zip(("uu", "vv", "ww", "weight", "vis_real", "vis_img"), arrs))


def write(path, uu, vv, ww, weight, vis_real, vis_img, freq_per_chan, polarisations, *,
          ntimes=0, dt=0.0, thours=None, baselines=0, uvmin=0.0, uvmax=0.0, wmin=0.0,
          wmax=0.0) -> Meta:
    """Writes a binMS directory from host arrays; returns its Meta."""
    os.makedirs(path, exist_ok=True)
    arrs = [np.ascontiguousarray(a, dtype=d) for a, d in
            zip((uu, vv, ww, weight, vis_real, vis_img), ("f8", "f8", "f8", "f4", "f4", "f4"))]
    n = arrs[0].size
    F, P = int(freq_per_chan), int(polarisations)
    sizes = (n, n, n, n * P, n * F * P, n * F * P)
    for a, s, name in zip(arrs, sizes, FILES):
        if a.size != s:
            raise ValueError(f"{name}: {a.size} values, expected {s}")
    meta = Meta(n, n * F * P, F, P, ntimes, dt, dt * ntimes / 3600 if thours is None else thours,
                baselines, uvmin, uvmax, wmin, wmax)
    c = meta.to_c()
    check(lib().rick_binms_write(_path(path), ctypes.byref(c), *map(_addr, arrs)))
    return meta


def write_image(real_path, imag_path, image_real, image_imag, xaxis, yaxis, rank=0):
    """Raw image files of WRITE_IMAGE; host arrays or CUDA tensors (copied)."""
    re, im = (np.ascontiguousarray(_host(a), dtype=np.float64) for a in (image_real, image_imag))
    if re.size < xaxis * yaxis or im.size < xaxis * yaxis:
        raise ValueError("image arrays smaller than xaxis*yaxis")
    check(lib().rick_write_image(_path(real_path), _path(imag_path), _addr(re), _addr(im),
                                 xaxis, yaxis, rank))


def write_cube(real_path, imag_path, grid, num_w_planes, grid_size_x, grid_size_y):
    """Real and imaginary cubes of a grid (host numpy or CUDA tensor, FP64 or
    FP32), written as doubles plane after plane."""
    n = 2 * num_w_planes * grid_size_x * grid_size_y
    if isinstance(grid, np.ndarray):
        if grid.dtype not in (np.float64, np.float32) or not grid.flags.c_contiguous:
            raise TypeError("grid must be a contiguous float64/float32 array")
        prec, dev = (FP32 if grid.dtype == np.float32 else FP64), -1
    else:
        import torch
        if not (grid.is_cuda and grid.is_contiguous()
                and grid.dtype in (torch.float64, torch.float32)):
            raise TypeError("grid must be a contiguous CUDA float64/float32 tensor")
        prec = FP32 if grid.dtype == torch.float32 else FP64
        dev = grid.device.index
        torch.cuda.synchronize(dev)
    if (grid.size if isinstance(grid, np.ndarray) else grid.numel()) < n:
        raise ValueError(f"grid has fewer than {n} values")
    check(lib().rick_write_cube(_path(real_path), _path(imag_path), _addr(grid), prec,
                                num_w_planes, grid_size_x, grid_size_y, dev))


def _host(a):
    if isinstance(a, np.ndarray):
        return a
    return a.detach().cpu().numpy()


__all__ = ["Meta", "read_meta", "partition", "read", "write", "write_image", "write_cube",
           "FILES", "META", "DROP_ROWS"]
