"""Host-side interface of the imaging chain.

Two layers over librick_hip.so:

* ``wstack``, ``phase_correction``, ``fft_planes`` -- the reference's own entry
  points (w-stacking.h:39-57, 70-83; the FFT stage of fourier_transform.c:13-93)
  with the same argument order and meaning, on host numpy arrays; outputs are
  accumulated into caller-provided arrays exactly as the reference does.
* ``Plan`` -- the device-resident API: torch tensors on the GPU, the current
  torch stream, no allocation per call.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import FP32, FP64, GAUSS, GAUSS_HI_PRECISION, KAISERBESSEL, RickError, check, lib

KERNELS = {"gauss": GAUSS, "kb": KAISERBESSEL, "kaiserbessel": KAISERBESSEL,
           "gauss_hi_precision": GAUSS_HI_PRECISION}


def _dp(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _fp(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _need(a, dtype, n, name):
    if not isinstance(a, np.ndarray) or a.dtype != dtype or not a.flags.c_contiguous:
        raise TypeError(f"{name} must be a C-contiguous numpy {np.dtype(dtype).name} array")
    if a.size < n:
        raise ValueError(f"{name} has {a.size} elements, needs {n}")


# ---------------------------------------------------------------------------
# Reference entry points (host arrays).

def set_kernel(kind: int) -> None:
    check(lib().rick_set_kernel(kind))


def set_phase(on: bool) -> None:
    check(lib().rick_set_phase(int(bool(on))))


def wstack(num_w_planes, num_points, freq_per_chan, polarizations, uu, vv, ww, vis_real,
           vis_img, weight, dx, dw, w_support, grid_size_x, grid_size_y, grid, num_threads=1,
           rank=0):
    """wstack() of w-stacking.cu:211 on host arrays; ``grid`` is accumulated into."""
    n = int(num_points)
    nv = n * freq_per_chan * polarizations
    _need(uu, np.float64, n, "uu")
    _need(vv, np.float64, n, "vv")
    _need(ww, np.float64, n, "ww")
    _need(vis_real, np.float32, nv, "vis_real")
    _need(vis_img, np.float32, nv, "vis_img")
    _need(weight, np.float32, n * polarizations, "weight")
    _need(grid, np.float64, 2 * num_w_planes * grid_size_x * grid_size_y, "grid")
    lib().wstack(num_w_planes, n, freq_per_chan, polarizations, _dp(uu), _dp(vv), _dp(ww),
                 _fp(vis_real), _fp(vis_img), _fp(weight), dx, dw, w_support, grid_size_x,
                 grid_size_y, _dp(grid), num_threads, rank)
    check()
    return grid


def fft_planes(num_w_planes, grid_size_x, grid_size_y, grid, gridss=None, rank=0):
    """The FFT stage (fourier_transform.c:13-93) on host arrays: gridss = FFT(grid)."""
    n = 2 * num_w_planes * grid_size_x * grid_size_y
    _need(grid, np.float64, n, "grid")
    if gridss is None:
        gridss = np.empty(n, dtype=np.float64)
    _need(gridss, np.float64, n, "gridss")
    lib().rick_fft_planes(num_w_planes, grid_size_x, grid_size_y, grid_size_x, grid_size_y,
                          _dp(grid), _dp(gridss), rank)
    check()
    return gridss


def phase_correction(gridss, image_real, image_imag, xaxis, yaxis, num_w_planes, xaxistot,
                     yaxistot, resolution, wmin, wmax, num_threads=1, rank=0):
    """phase_correction() of phase_correction.cu:97; images are accumulated into."""
    _need(gridss, np.float64, 2 * num_w_planes * xaxis * yaxis, "gridss")
    _need(image_real, np.float64, xaxis * yaxis, "image_real")
    _need(image_imag, np.float64, xaxis * yaxis, "image_imag")
    lib().phase_correction(_dp(gridss), _dp(image_real), _dp(image_imag), xaxis, yaxis,
                           num_w_planes, xaxistot, yaxistot, resolution, wmin, wmax, num_threads,
                           rank)
    check()
    return image_real, image_imag


# ---------------------------------------------------------------------------
# Device-resident plans.

def _torch():
    import torch
    return torch


class Plan:
    """A rick_plan: device tables and gridding workspace for one problem shape."""

    def __init__(self, grid_x, grid_y, planes, kernel=KAISERBESSEL, support=7, dx=None,
                 dw=None, precision=FP64, max_points=0, resolution=1.0, wmin=0.0, wmax=1.0,
                 phase_on=True, tile=(0, 0), rows_per_task=0, device=0, xaxistot=0,
                 yaxistot=0, unroll=0):
        d = _lib.PlanDesc()
        d.num_w_planes = planes
        d.grid_size_x = grid_x
        d.grid_size_y = grid_y
        d.kernel = KERNELS.get(kernel, kernel) if isinstance(kernel, str) else kernel
        d.w_support = support
        d.dx = 1.0 / grid_x if dx is None else dx      # init.c:11
        d.dw = 1.0 / planes if dw is None else dw      # init.c:12
        d.precision = precision
        d.max_points = max_points
        d.resolution = resolution
        d.wmin = wmin
        d.wmax = wmax
        d.phase_on = int(bool(phase_on))
        d.tile_u, d.tile_v = tile
        d.rows_per_task = rows_per_task
        d.device = device
        d.xaxistot, d.yaxistot = xaxistot, yaxistot   # a rank's slab of the image
        d.footprint_unroll = unroll                    # 0: smallest of 7, 9, 11, 15 >= support
        self.desc = d
        self._h = ctypes.c_void_p()
        check(lib().rick_plan_create(ctypes.byref(self._h), ctypes.byref(d)))
        self.nx, self.ny, self.planes = grid_x, grid_y, planes
        self.precision = precision
        self.device = device

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and h.value and _lib._lib is not None:
            _lib._lib.rick_plan_destroy(h)
            self._h = ctypes.c_void_p()

    @property
    def device_bytes(self) -> int:
        return lib().rick_plan_device_bytes(self._h)

    @property
    def grid_bytes(self) -> int:
        return lib().rick_plan_grid_bytes(self._h)

    @property
    def grid_dtype(self):
        torch = _torch()
        return torch.float64 if self.precision == FP64 else torch.float32

    def profile(self, enable: bool = True) -> None:
        """Bracket every kernel this plan launches with HIP events."""
        check(lib().rick_plan_profile(self._h, int(bool(enable))))

    def stage_times(self) -> dict:
        """{stage: (summed kernel ms, launches)} since the last query (waits)."""
        n = len(_lib.STAGES)
        ms = (ctypes.c_double * n)()
        cnt = (ctypes.c_int * n)()
        check(lib().rick_plan_stage_times(self._h, ms, cnt, n))
        return {name: (ms[i], cnt[i]) for i, name in enumerate(_lib.STAGES)}

    def grid_stats(self) -> dict:
        s = (ctypes.c_uint * 4)()
        check(lib().rick_plan_grid_stats(self._h, s, 4))
        return {"records": s[0], "tasks": s[1], "split_buckets": s[2], "nonempty_buckets": s[3]}

    def new_grid(self):
        torch = _torch()
        return torch.empty(2 * self.planes * self.ny * self.nx, dtype=self.grid_dtype,
                           device=f"cuda:{self.device}")

    def _stream(self, stream):
        torch = _torch()
        s = stream if stream is not None else torch.cuda.current_stream(self.device)
        return ctypes.c_void_p(s.cuda_stream)

    def _check_grid(self, t, name="grid"):
        torch = _torch()
        n = 2 * self.planes * self.ny * self.nx
        if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == self.grid_dtype
                and t.is_contiguous() and t.numel() >= n):
            raise TypeError(f"{name} must be a contiguous CUDA {self.grid_dtype} tensor of "
                            f">= {n} elements")

    @staticmethod
    def _ptr(t, dtype, n, name):
        torch = _torch()
        if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == dtype
                and t.is_contiguous() and t.numel() >= n):
            raise TypeError(f"{name} must be a contiguous CUDA {dtype} tensor of >= {n} elements")
        return ctypes.c_void_p(t.data_ptr())

    def grid(self, uu, vv, ww, vis_real, vis_img, weight, freq_per_chan=1, polarizations=1,
             out=None, accumulate=False, stream=None):
        torch = _torch()
        n = int(uu.numel())
        nv = n * freq_per_chan * polarizations
        if out is None:
            out = self.new_grid()
            accumulate = False
        self._check_grid(out)
        st = lib().rick_plan_grid(
            self._h, n, freq_per_chan, polarizations,
            self._ptr(uu, torch.float64, n, "uu"), self._ptr(vv, torch.float64, n, "vv"),
            self._ptr(ww, torch.float64, n, "ww"),
            self._ptr(vis_real, torch.float32, nv, "vis_real"),
            self._ptr(vis_img, torch.float32, nv, "vis_img"),
            self._ptr(weight, torch.float32, n * polarizations, "weight"),
            ctypes.c_void_p(out.data_ptr()), int(bool(accumulate)), self._stream(stream))
        check(st)
        return out

    def fft(self, grid, out=None, stream=None):
        self._check_grid(grid)
        if out is None:
            out = grid
        self._check_grid(out, "out")
        check(lib().rick_plan_fft(self._h, ctypes.c_void_p(grid.data_ptr()),
                                  ctypes.c_void_p(out.data_ptr()), self._stream(stream)))
        return out

    def phase(self, gridss, image_real=None, image_imag=None, accumulate=False, stream=None):
        torch = _torch()
        self._check_grid(gridss, "gridss")
        npix = self.nx * self.ny
        if image_real is None or image_imag is None:
            image_real = torch.empty(npix, dtype=torch.float64, device=gridss.device)
            image_imag = torch.empty(npix, dtype=torch.float64, device=gridss.device)
            accumulate = False
        check(lib().rick_plan_phase(
            self._h, ctypes.c_void_p(gridss.data_ptr()),
            self._ptr(image_real, torch.float64, npix, "image_real"),
            self._ptr(image_imag, torch.float64, npix, "image_imag"), int(bool(accumulate)),
            self._stream(stream)))
        return image_real, image_imag

    def image(self, uu, vv, ww, vis_real, vis_img, weight, freq_per_chan=1, polarizations=1,
              grid=None, image_real=None, image_imag=None, stream=None):
        torch = _torch()
        n = int(uu.numel())
        nv = n * freq_per_chan * polarizations
        npix = self.nx * self.ny
        if grid is None:
            grid = self.new_grid()
        self._check_grid(grid)
        if image_real is None or image_imag is None:
            image_real = torch.empty(npix, dtype=torch.float64, device=grid.device)
            image_imag = torch.empty(npix, dtype=torch.float64, device=grid.device)
        check(lib().rick_plan_image(
            self._h, n, freq_per_chan, polarizations,
            self._ptr(uu, torch.float64, n, "uu"), self._ptr(vv, torch.float64, n, "vv"),
            self._ptr(ww, torch.float64, n, "ww"),
            self._ptr(vis_real, torch.float32, nv, "vis_real"),
            self._ptr(vis_img, torch.float32, nv, "vis_img"),
            self._ptr(weight, torch.float32, n * polarizations, "weight"),
            ctypes.c_void_p(grid.data_ptr()),
            self._ptr(image_real, torch.float64, npix, "image_real"),
            self._ptr(image_imag, torch.float64, npix, "image_imag"), self._stream(stream)))
        return image_real, image_imag, grid


def plan_for(ds, precision=FP64, **kw) -> Plan:
    """Plan sized for a synthetic Dataset (rick_amd.synth)."""
    cfg = ds.cfg
    return Plan(cfg.grid, cfg.grid, cfg.planes, kernel=cfg.kernel, support=cfg.support,
                dx=ds.dx, dw=ds.dw, precision=precision, max_points=ds.rows,
                resolution=ds.resolution, wmin=ds.wmin, wmax=ds.wmax, **kw)


__all__ = ["Plan", "plan_for", "wstack", "fft_planes", "phase_correction", "set_kernel",
           "set_phase", "RickError", "FP64", "FP32", "GAUSS", "KAISERBESSEL",
           "GAUSS_HI_PRECISION"]
