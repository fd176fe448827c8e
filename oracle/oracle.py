"""ctypes front end of the CPU oracle.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.

Two back ends:
  * ``liboracle.so`` -- the C restatement in rick_oracle.c (always built);
  * ``_ref/libref_*.so`` -- the reference's own wstack()/phase_correction()
    compiled from its sources (w-stacking.cu, phase_correction.cu) by
    ``make ref``; present only where those sources were available at build time.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
GAUSS, KAISERBESSEL, GAUSS_HI_PRECISION = 0, 1, 2
KERNEL_NAMES = {GAUSS: "gauss", KAISERBESSEL: "kb", GAUSS_HI_PRECISION: "gausshp"}

_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)
_c_int, _c_uint, _c_double = ctypes.c_int, ctypes.c_uint, ctypes.c_double


def _ptr(a: np.ndarray, ctype):
    return a.ctypes.data_as(ctype)


_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        path = HERE / "liboracle.so"
        if not path.exists():
            raise FileNotFoundError(f"{path} is not built (run `make -C oracle`)")
        L = ctypes.CDLL(str(path))
        L.oracle_bessel0.restype = _c_double
        L.oracle_bessel0.argtypes = [_c_double, _c_double]
        L.oracle_kernel_table.restype = _c_int
        L.oracle_kernel_table.argtypes = [_c_int, _c_int, _dp]
        L.oracle_wstack.restype = None
        L.oracle_wstack.argtypes = [_c_int, _c_int, _c_uint, _c_uint, _c_uint, _dp, _dp, _dp,
                                    _fp, _fp, _fp, _c_double, _c_double, _c_int, _c_int,
                                    _c_int, _dp, _c_int]
        L.oracle_fft_planes.restype = None
        L.oracle_fft_planes.argtypes = [_c_int, _c_int, _c_int, _dp, _dp, _c_int]
        L.oracle_phase_correction.restype = None
        L.oracle_phase_correction.argtypes = [_dp, _dp, _dp, _c_int, _c_int, _c_int, _c_int,
                                              _c_int, _c_double, _c_double, _c_double,
                                              _c_int, _c_int]
        L.oracle_normalize_uvw.restype = None
        L.oracle_normalize_uvw.argtypes = [ctypes.c_size_t, _dp, _dp, _dp]
        _up = ctypes.POINTER(_c_uint)
        L.oracle_sector_bin.restype = _c_uint
        L.oracle_sector_bin.argtypes = [ctypes.c_size_t, _dp, _c_int, _c_int, _c_double, _c_int,
                                        _up, _up, _up]
        L.oracle_sector_gather.restype = None
        L.oracle_sector_gather.argtypes = [_up, _c_uint, _c_int, _c_int, _c_double, _c_uint,
                                           _c_uint, _dp, _dp, _dp, _fp, _fp, _fp, _dp, _dp, _dp,
                                           _fp, _fp, _fp]
        _lib = L
    return _lib


def bessel0(x: float, precision: float = 1e-8) -> float:
    return lib().oracle_bessel0(x, precision)


def kernel_table(kind: int, support: int) -> np.ndarray:
    t = np.zeros(5 * support, dtype=np.float64)
    n = lib().oracle_kernel_table(kind, support, _ptr(t, _dp))
    if n < 0:
        raise ValueError(f"no table for kernel kind {kind}")
    return t


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


def wstack(kind, planes, uu, vv, ww, vis_real, vis_img, weight, freq_per_chan,
           polarizations, dx, dw, support, nx, ny, grid=None, threads=0):
    """oracle_wstack; returns the (accumulated) grid as float64[2*planes*ny*nx]."""
    uu, vv, ww = _c(uu, np.float64), _c(vv, np.float64), _c(ww, np.float64)
    vr, vi, wt = _c(vis_real, np.float32), _c(vis_img, np.float32), _c(weight, np.float32)
    n = uu.shape[0]
    assert vr.size == n * freq_per_chan * polarizations and wt.size == n * polarizations
    if grid is None:
        grid = np.zeros(2 * planes * nx * ny, dtype=np.float64)
    lib().oracle_wstack(kind, planes, n, freq_per_chan, polarizations, _ptr(uu, _dp),
                        _ptr(vv, _dp), _ptr(ww, _dp), _ptr(vr, _fp), _ptr(vi, _fp),
                        _ptr(wt, _fp), dx, dw, support, nx, ny, _ptr(grid, _dp), threads)
    return grid


def fft_planes(planes, nx, ny, grid, threads=0):
    grid = _c(grid, np.float64)
    out = np.empty_like(grid)
    lib().oracle_fft_planes(planes, nx, ny, _ptr(grid, _dp), _ptr(out, _dp), threads)
    return out


def phase_correction(gridss, nx, ny, planes, resolution, wmin, wmax, phase_on=True,
                     threads=0, xtot=None, ytot=None):
    gridss = _c(gridss, np.float64)
    re = np.zeros(nx * ny, dtype=np.float64)
    im = np.zeros(nx * ny, dtype=np.float64)
    lib().oracle_phase_correction(_ptr(gridss, _dp), _ptr(re, _dp), _ptr(im, _dp), nx, ny,
                                  planes, xtot or nx, ytot or ny, resolution, wmin, wmax,
                                  int(phase_on), threads)
    return re, im


def normalize_uvw(uu, vv, ww):
    uu, vv, ww = (np.array(a, dtype=np.float64, copy=True) for a in (uu, vv, ww))
    lib().oracle_normalize_uvw(uu.size, _ptr(uu, _dp), _ptr(vv, _dp), _ptr(ww, _dp))
    return uu, vv, ww


def sector_bin(vv, nsectors, yaxis, dx, w_support):
    """initialize_array (gridding.c:139-197): (histo, start, rows, dropped)."""
    vv = _c(vv, np.float64)
    histo = np.zeros(nsectors + 1, dtype=np.uint32)
    start = np.zeros(nsectors + 1, dtype=np.uint32)
    rows = np.zeros(max(1, 3 * vv.size), dtype=np.uint32)
    up = ctypes.POINTER(_c_uint)
    bad = lib().oracle_sector_bin(vv.size, _ptr(vv, _dp), nsectors, yaxis, dx, w_support,
                                  _ptr(histo, up), _ptr(start, up), _ptr(rows, up))
    return histo, start, rows[:int(histo.sum())], int(bad)


def sector_gather(list_rows, isector, yaxis, dx, freq_per_chan, polarizations, uu, vv, ww,
                  weight, vis_real, vis_img):
    """gridding_cpu.c:90-138 for one sector: dict of the arrays passed to wstack."""
    lst = _c(list_rows, np.uint32)
    n = lst.size
    F, P = freq_per_chan, polarizations
    out = {"uu": np.zeros(n), "vv": np.zeros(n), "ww": np.zeros(n),
           "weight": np.zeros(n * P, np.float32), "vis_real": np.zeros(n * F * P, np.float32),
           "vis_img": np.zeros(n * F * P, np.float32)}
    ins = [_c(uu, np.float64), _c(vv, np.float64), _c(ww, np.float64), _c(weight, np.float32),
           _c(vis_real, np.float32), _c(vis_img, np.float32)]
    lib().oracle_sector_gather(
        _ptr(lst, ctypes.POINTER(_c_uint)), n, isector, yaxis, dx, F, P,
        *(_ptr(a, _dp if a.dtype == np.float64 else _fp) for a in ins),
        *(_ptr(out[k], _dp if out[k].dtype == np.float64 else _fp) for k in
          ("uu", "vv", "ww", "weight", "vis_real", "vis_img")))
    return out


def read_binms(path, freq_per_chan, polarizations, startrow, nrows):
    """readData (init.c:424-490): the six arrays of rows [startrow, startrow+nrows)
    read at the reference's byte offsets."""
    F, P = freq_per_chan, polarizations
    spec = [("ucoord.bin", np.float64, 1), ("vcoord.bin", np.float64, 1),
            ("wcoord.bin", np.float64, 1), ("weights.bin", np.float32, P),
            ("visibilities_real.bin", np.float32, F * P),
            ("visibilities_img.bin", np.float32, F * P)]
    out = []
    for name, dt, per in spec:
        item = np.dtype(dt).itemsize
        out.append(np.fromfile(os.path.join(str(path), name), dtype=dt, count=nrows * per,
                               offset=startrow * per * item))
    return dict(zip(("uu", "vv", "ww", "weight", "vis_real", "vis_img"), out))


def partition(nmeasures, rank, size, drop=1000):
    """metaData_calculation (init.c:335-372): (startrow, nrows) of rank."""
    total = nmeasures - drop
    per, rem = divmod(total, size)
    return rank * per, per + (rem if rank == size - 1 else 0)


# ---------------------------------------------------------------------------
# Reference-compiled libraries (oracle/_ref).

_ref_libs: dict = {}


def ref_available() -> bool:
    return all((HERE / "_ref" / f"libref_{k}.so").exists()
               for k in ("gauss", "kb", "gausshp", "phase_on", "phase_off"))


def _ref(name: str) -> ctypes.CDLL:
    if name not in _ref_libs:
        path = HERE / "_ref" / f"libref_{name}.so"
        if not path.exists():
            raise FileNotFoundError(f"{path} is not built (run `make -C oracle ref`)")
        L = ctypes.CDLL(str(path), mode=os.RTLD_LOCAL)
        if name.startswith("phase"):
            L.phase_correction.restype = None
            L.phase_correction.argtypes = [_dp, _dp, _dp, _c_int, _c_int, _c_int, _c_int,
                                           _c_int, _c_double, _c_double, _c_double,
                                           _c_int, _c_int]
        else:
            L.wstack.restype = None
            L.wstack.argtypes = [_c_int, _c_uint, _c_uint, _c_uint, _dp, _dp, _dp, _fp, _fp,
                                 _fp, _c_double, _c_double, _c_int, _c_int, _c_int, _dp,
                                 _c_int, _c_int]
            L.makeGaussKernel.restype = None
            L.makeGaussKernel.argtypes = [_dp, _c_int, _c_int, _c_double]
            L.makeKaiserBesselKernel.restype = None
            L.makeKaiserBesselKernel.argtypes = [_dp, _c_int, _c_int, _c_double, _c_double,
                                                 _c_int]
        _ref_libs[name] = L
    return _ref_libs[name]


def ref_kernel_table(kind: int, support: int) -> np.ndarray:
    """The reference's own makeGaussKernel / makeKaiserBesselKernel output."""
    t = np.zeros(5 * support, dtype=np.float64)
    if kind == GAUSS:
        _ref("gauss").makeGaussKernel(_ptr(t, _dp), support, 5, 0.5)
    elif kind == KAISERBESSEL:
        _ref("kb").makeKaiserBesselKernel(_ptr(t, _dp), support, 5, 8.6, 1.0, 0)
    else:
        raise ValueError("GAUSS_HI_PRECISION has no table")
    return t


def ref_wstack(kind, planes, uu, vv, ww, vis_real, vis_img, weight, freq_per_chan,
               polarizations, dx, dw, support, nx, ny, threads=1):
    """The reference's wstack() on a zeroed grid.

    The grid is allocated with one plane of slack because the reference writes
    rows with w == 1.0 (plane index == planes) past the grid (init.c:391-409
    gives it that slack); the slack is discarded, as the reference's FFT never
    reads it.
    """
    uu, vv, ww = _c(uu, np.float64), _c(vv, np.float64), _c(ww, np.float64)
    vr, vi, wt = _c(vis_real, np.float32), _c(vis_img, np.float32), _c(weight, np.float32)
    n = uu.shape[0]
    plane = 2 * nx * ny
    grid = np.zeros(plane * (planes + 1), dtype=np.float64)
    _ref(KERNEL_NAMES[kind]).wstack(planes, n, freq_per_chan, polarizations, _ptr(uu, _dp),
                                    _ptr(vv, _dp), _ptr(ww, _dp), _ptr(vr, _fp),
                                    _ptr(vi, _fp), _ptr(wt, _fp), dx, dw, support, nx, ny,
                                    _ptr(grid, _dp), threads, 0)
    return grid[: plane * planes].copy()


def ref_wstack_into(kind, planes, uu, vv, ww, vis_real, vis_img, weight, freq_per_chan,
                    polarizations, dx, dw, support, nx, ny, grid, threads):
    """The reference's wstack() accumulating into ``grid``, which must hold
    planes+1 planes (the extra one absorbs rows at w == 1.0, see ref_wstack)."""
    assert grid.dtype == np.float64 and grid.size >= 2 * nx * ny * (planes + 1)
    _ref(KERNEL_NAMES[kind]).wstack(planes, uu.shape[0], freq_per_chan, polarizations,
                                    _ptr(uu, _dp), _ptr(vv, _dp), _ptr(ww, _dp),
                                    _ptr(vis_real, _fp), _ptr(vis_img, _fp), _ptr(weight, _fp),
                                    dx, dw, support, nx, ny, _ptr(grid, _dp), threads, 0)
    return grid


def ref_phase_correction(gridss, nx, ny, planes, resolution, wmin, wmax, phase_on=True,
                         threads=1, xtot=None, ytot=None):
    """The reference's phase_correction(); xtot/ytot are its xaxistot/yaxistot
    (a rank's slab of nx x ny pixels out of xtot x ytot), default nx/ny."""
    gridss = _c(gridss, np.float64)
    re = np.zeros(nx * ny, dtype=np.float64)
    im = np.zeros(nx * ny, dtype=np.float64)
    _ref("phase_on" if phase_on else "phase_off").phase_correction(
        _ptr(gridss, _dp), _ptr(re, _dp), _ptr(im, _dp), nx, ny, planes, xtot or nx, ytot or ny,
        resolution, wmin, wmax, threads, 0)
    return re, im
