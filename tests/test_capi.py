"""librick_hip.so: it loads, exports every entry point include/rick.h declares,
and rejects bad arguments before touching a device (no GPU needed)."""
import ctypes
import re
import subprocess
from pathlib import Path

import pytest

from conftest import ROOT
from rick_amd import _lib

HEADER = ROOT / "include" / "rick.h"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?(?:void|int|size_t|char)\s*\*?\s*(\w+)\s*\(", text,
                       flags=re.M)
    return sorted(set(names))


def test_library_is_built():
    assert _lib.LIB_PATH.exists(), "run `python -m rick_amd.build`"


def test_every_declared_symbol_is_exported():
    names = declared_functions()
    assert {"wstack", "phase_correction", "rick_fft_planes", "rick_plan_create",
            "rick_plan_grid", "rick_plan_fft", "rick_plan_phase", "rick_plan_image"} <= set(names)
    out = subprocess.run(["nm", "-D", "--defined-only", str(_lib.LIB_PATH)], capture_output=True,
                         text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    # The ctypes binding declares every one of them.
    assert set(names) == set(_lib.SIGNATURES)
    L = _lib.lib()
    for n in names:
        assert hasattr(L, n)


def test_no_oracle_dependency():
    out = subprocess.run(["ldd", str(_lib.LIB_PATH)], capture_output=True, text=True).stdout
    assert "oracle" not in out and "libref" not in out


def test_argument_errors_without_device():
    L = _lib.lib()
    assert L.rick_set_kernel(7) == _lib.RICK_ERR_INVALID
    assert b"unknown kernel" in L.rick_last_error()
    assert L.rick_set_kernel(_lib.KAISERBESSEL) == _lib.RICK_OK
    assert L.rick_get_kernel() == _lib.KAISERBESSEL
    assert L.rick_version().startswith(b"rick")

    def create(**kw):
        d = _lib.PlanDesc(num_w_planes=4, grid_size_x=64, grid_size_y=64, kernel=1,
                          w_support=7, dx=1 / 64, dw=1 / 4, precision=0, max_points=100)
        for k, v in kw.items():
            setattr(d, k, v)
        h = ctypes.c_void_p()
        return L.rick_plan_create(ctypes.byref(h), ctypes.byref(d))

    assert create(w_support=8) == _lib.RICK_ERR_INVALID          # even support
    assert create(w_support=17) == _lib.RICK_ERR_INVALID
    assert create(kernel=5) == _lib.RICK_ERR_INVALID
    assert create(num_w_planes=0) == _lib.RICK_ERR_INVALID
    assert create(precision=3) == _lib.RICK_ERR_INVALID
    assert create(dx=0.0) == _lib.RICK_ERR_INVALID
    assert create(tile_u=48, tile_v=48) == _lib.RICK_ERR_INVALID
    assert create(max_points=2 ** 31) == _lib.RICK_ERR_INVALID
    # Launch limits (fewer than 2^32 threads per launch), checked at plan
    # creation: 8192^2 x 128 planes in 32x32 tiles is 2^23 buckets, so the
    # one-workgroup-per-task tile launch would need 2^32 threads.
    big = dict(num_w_planes=128, grid_size_x=8192, grid_size_y=8192, dx=1 / 8192, dw=1 / 128)
    assert create(tile_u=32, tile_v=32, **big) == _lib.RICK_ERR_INVALID
    assert b"tile launch" in L.rick_last_error()
    assert create(tile_u=32, tile_v=32, **dict(big, num_w_planes=64)) != _lib.RICK_ERR_INVALID
    # One record per task: up to 2 * max_points split buckets.
    assert create(rows_per_task=1, max_points=2 ** 29) == _lib.RICK_ERR_INVALID
    assert b"split buckets" in L.rick_last_error()
    assert L.rick_plan_grid(None, 0, 1, 1, *([None] * 7), 0, None) == _lib.RICK_ERR_INVALID


REFERENCE = Path("/root/reference")
MPI_INCLUDE = Path("/opt/conda/include")


@pytest.mark.skipif(not (REFERENCE / "w-stacking.h").exists() or
                    not (MPI_INCLUDE / "mpi.h").exists(),
                    reason="needs the reference's w-stacking.h (and the mpi.h it includes)")
def test_header_agrees_with_reference_prototypes():
    """include/rick.h and the reference's w-stacking.h in one C translation
    unit: repeated declarations must be compatible, so wstack() and
    phase_correction() keep the reference's C signatures.  A planted wrong
    prototype is rejected, so the check is not vacuous."""
    cmd = ["gcc", "-std=gnu11", "-fsyntax-only", f"-I{REFERENCE}", f"-I{MPI_INCLUDE}",
           f"-I{ROOT / 'include'}", str(ROOT / "tests" / "c" / "header_compat.c")]
    ok = subprocess.run(cmd, capture_output=True, text=True)
    assert ok.returncode == 0, ok.stderr
    bad = subprocess.run(cmd + ["-DRICK_PLANT_MISMATCH"], capture_output=True, text=True)
    assert bad.returncode != 0 and "conflicting types" in bad.stderr


def build_c_caller(out: Path) -> Path:
    """Builds tests/c/dropin_caller.c against include/rick.h, linked with
    -lrick_hip as a C program of the reference would be."""
    lib_dir = _lib.LIB_PATH.parent
    subprocess.run(["gcc", "-O2", "-std=gnu11", f"-I{ROOT / 'include'}",
                    str(ROOT / "tests" / "c" / "dropin_caller.c"), f"-L{lib_dir}", "-lrick_hip",
                    f"-Wl,-rpath,{lib_dir}", "-o", str(out)], check=True)
    return out


def test_c_caller_links_with_library(tmp_path):
    exe = build_c_caller(tmp_path / "dropin_caller")
    r = subprocess.run([str(exe), "--version"], capture_output=True, text=True, check=True)
    assert r.stdout.startswith("rick")


def test_python_wrappers_validate_arrays():
    import numpy as np
    from rick_amd import imaging
    g = np.zeros(2 * 4 * 8 * 8)
    with pytest.raises(TypeError):
        imaging.wstack(4, 1, 1, 1, np.zeros(1, np.float32), np.zeros(1), np.zeros(1),
                       np.zeros(1, np.float32), np.zeros(1, np.float32), np.ones(1, np.float32),
                       1 / 8, 1 / 4, 7, 8, 8, g)
    with pytest.raises(ValueError):
        imaging.wstack(4, 3, 1, 1, np.zeros(1), np.zeros(1), np.zeros(1),
                       np.zeros(1, np.float32), np.zeros(1, np.float32), np.ones(1, np.float32),
                       1 / 8, 1 / 4, 7, 8, 8, g)
