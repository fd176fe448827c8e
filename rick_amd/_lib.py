"""ctypes binding of librick_hip.so (declarations in include/rick.h).

The library is loaded from this package directory only; a missing library is an
error (there is no CPU fallback in the product).
"""
from __future__ import annotations

import ctypes
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "librick_hip.so"

RICK_OK, RICK_ERR_INVALID, RICK_ERR_UNSUPPORTED, RICK_ERR_DEVICE, RICK_ERR_NOMEM = range(5)
GAUSS, KAISERBESSEL, GAUSS_HI_PRECISION = 0, 1, 2
FP64, FP32 = 0, 1
STAGES = ("count", "scan", "scatter", "reserved", "tile", "fft_cols", "fft_cols_b",
          "fft_rows", "phase", "zero_split")

_c_int, _c_uint, _c_double, _c_size, _vp = (ctypes.c_int, ctypes.c_uint, ctypes.c_double,
                                            ctypes.c_size_t, ctypes.c_void_p)
_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)


class PlanDesc(ctypes.Structure):
    """rick_plan_desc (include/rick.h)."""
    _fields_ = [
        ("num_w_planes", _c_int), ("grid_size_x", _c_int), ("grid_size_y", _c_int),
        ("kernel", _c_int), ("w_support", _c_int), ("dx", _c_double), ("dw", _c_double),
        ("precision", _c_int), ("max_points", _c_uint), ("resolution", _c_double),
        ("wmin", _c_double), ("wmax", _c_double), ("phase_on", _c_int), ("tile_u", _c_int),
        ("tile_v", _c_int), ("rows_per_task", _c_int), ("device", _c_int),
        ("xaxistot", _c_int), ("yaxistot", _c_int), ("footprint_unroll", _c_int),
    ]


class BinmsMeta(ctypes.Structure):
    """rick_binms_meta (include/rick.h): the twelve parsed values of meta.txt."""
    _fields_ = [
        ("nmeasures", _c_uint), ("nvis", ctypes.c_ulong), ("freq_per_chan", _c_uint),
        ("polarisations", _c_uint), ("ntimes", _c_uint), ("dt", _c_double),
        ("thours", _c_double), ("baselines", _c_uint), ("uvmin", _c_double),
        ("uvmax", _c_double), ("wmin", _c_double), ("wmax", _c_double),
    ]


class BatchJob(ctypes.Structure):
    """rick_batch_job (include/rick.h)."""
    _fields_ = [
        ("num_points", _c_uint), ("uu", _vp), ("vv", _vp), ("ww", _vp), ("vis_real", _vp),
        ("vis_img", _vp), ("weight", _vp), ("num_w_planes", _c_int), ("dw", _c_double),
        ("resolution", _c_double), ("wmin", _c_double), ("wmax", _c_double),
        ("image_real", _vp), ("image_imag", _vp),
    ]


class BatchDesc(ctypes.Structure):
    """rick_batch_desc (include/rick.h)."""
    _fields_ = [
        ("grid_size", _c_int), ("kernel", _c_int), ("w_support", _c_int), ("precision", _c_int),
        ("phase_on", _c_int), ("freq_per_chan", _c_uint), ("polarizations", _c_uint),
        ("tile_u", _c_int), ("tile_v", _c_int), ("rows_per_task", _c_int),
        ("footprint_unroll", _c_int), ("device", _c_int),
    ]


_cp = ctypes.c_char_p
_up = ctypes.POINTER(_c_uint)
_ulp = ctypes.POINTER(ctypes.c_ulong)

# name -> (restype, argtypes); every symbol declared in include/rick.h.
SIGNATURES = {
    "wstack": (None, [_c_int, _c_uint, _c_uint, _c_uint, _dp, _dp, _dp, _fp, _fp, _fp, _c_double,
                      _c_double, _c_int, _c_int, _c_int, _dp, _c_int, _c_int]),
    "phase_correction": (None, [_dp, _dp, _dp, _c_int, _c_int, _c_int, _c_int, _c_int,
                                _c_double, _c_double, _c_double, _c_int, _c_int]),
    "rick_fft_planes": (None, [_c_int, _c_int, _c_int, _c_int, _c_int, _dp, _dp, _c_int]),
    "rick_set_kernel": (_c_int, [_c_int]),
    "rick_get_kernel": (_c_int, []),
    "rick_set_phase": (_c_int, [_c_int]),
    "rick_get_phase": (_c_int, []),
    "rick_last_status": (_c_int, []),
    "rick_last_error": (ctypes.c_char_p, []),
    "rick_device_count": (_c_int, []),
    "rick_version": (ctypes.c_char_p, []),
    "rick_plan_create": (_c_int, [ctypes.POINTER(_vp), ctypes.POINTER(PlanDesc)]),
    "rick_plan_destroy": (_c_int, [_vp]),
    "rick_plan_device_bytes": (_c_size, [_vp]),
    "rick_plan_grid_bytes": (_c_size, [_vp]),
    "rick_plan_grid": (_c_int, [_vp, _c_uint, _c_uint, _c_uint, _vp, _vp, _vp, _vp, _vp, _vp,
                                _vp, _c_int, _vp]),
    "rick_plan_fft": (_c_int, [_vp, _vp, _vp, _vp]),
    "rick_plan_phase": (_c_int, [_vp, _vp, _vp, _vp, _c_int, _vp]),
    "rick_plan_profile": (_c_int, [_vp, _c_int]),
    "rick_plan_stage_times": (_c_int, [_vp, _dp, ctypes.POINTER(_c_int), _c_int]),
    "rick_plan_grid_stats": (_c_int, [_vp, ctypes.POINTER(_c_uint), _c_int]),
    "rick_plan_image": (_c_int, [_vp, _c_uint, _c_uint, _c_uint, _vp, _vp, _vp, _vp, _vp, _vp,
                                 _vp, _vp, _vp, _vp]),
    "rick_batch_create": (_c_int, [ctypes.POINTER(_vp), ctypes.POINTER(BatchDesc),
                                   ctypes.POINTER(BatchJob), _c_int]),
    "rick_batch_run": (_c_int, [_vp, _vp]),
    "rick_batch_destroy": (_c_int, [_vp]),
    "rick_batch_device_bytes": (_c_size, [_vp]),
    "rick_binms_read_meta": (_c_int, [_cp, ctypes.POINTER(BinmsMeta)]),
    "rick_binms_partition": (_c_int, [ctypes.POINTER(BinmsMeta), _c_int, _c_int, _c_uint, _ulp,
                                      _up]),
    "rick_binms_read": (_c_int, [_cp, _c_uint, _c_uint, ctypes.c_ulong, _c_uint, _vp, _vp, _vp,
                                 _vp, _vp, _vp, _c_int]),
    "rick_binms_write": (_c_int, [_cp, ctypes.POINTER(BinmsMeta), _vp, _vp, _vp, _vp, _vp, _vp]),
    "rick_write_image": (_c_int, [_cp, _cp, _vp, _vp, _c_int, _c_int, _c_int]),
    "rick_write_cube": (_c_int, [_cp, _cp, _vp, _c_int, _c_int, _c_int, _c_int, _c_int]),
    "rick_normalize_uvw_workspace_bytes": (_c_size, []),
    "rick_normalize_uvw": (_c_int, [_c_uint, _vp, _vp, _vp, _vp, _vp, _vp]),
    "rick_sector_workspace_bytes": (_c_size, [_c_uint, _c_int]),
    "rick_sector_bin": (_c_int, [_c_uint, _vp, _c_int, _c_int, _c_double, _c_int, _vp, _vp, _vp,
                                 _c_uint, _vp, _vp, _ulp, _up]),
    "rick_sector_bin_noalloc_max_workspace": (_c_size, []),
    "rick_sector_bin_noalloc": (_c_int, [_c_uint, _vp, _c_int, _c_int, _c_double, _c_int, _vp,
                                         _vp, _vp, _c_uint, _vp, _ulp, _up]),
    "rick_sector_gather": (_c_int, [_vp, _c_uint, _c_int, _c_int, _c_double, _c_uint, _c_uint,
                                    _c_uint, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                    _vp, _vp, _vp]),
}

_lib = None


class RickError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"rick status {status}: {message}")
        self.status = status


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(f"{LIB_PATH} is not built; run `python -m rick_amd.build`")
        L = ctypes.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(status: int | None = None):
    """Raise RickError if ``status`` (or the library's last status) is not OK."""
    L = lib()
    st = L.rick_last_status() if status is None else status
    if st != RICK_OK:
        raise RickError(st, (L.rick_last_error() or b"").decode())
