"""One dataset through the whole reference flow on one GPU, device-resident
from the file read to the image write (main.c:120-190 with one rank):

  init        readMetaData, metaData_calculation, readData  (init.c:300-490)
  gridding    NORMALIZE_UVW, initialize_array, per-sector assembly, wstack
              (gridding.c:26-197, gridding_cpu.c:78-204)
  fftw_data   per-plane backward FFT / (Nx*Ny)           (fourier_transform.c:13-93)
  phase       phase_correction                            (fourier_transform.c:213)
  write       WRITE_IMAGE raw files, WRITE_DATA cubes       (fourier_transform.c:280-289)

    python -m rick_amd.pipeline DATADIR --grid 512 --planes 8 --out OUTDIR
"""
from __future__ import annotations

import argparse
import json
import os
import time
from dataclasses import dataclass, field

from . import binms, prep
from ._lib import FP32, FP64, KAISERBESSEL
from .imaging import KERNELS, Plan


@dataclass
class Result:
    image_real: object
    image_imag: object
    grid: object
    meta: binms.Meta
    rows: int
    timings_ms: dict = field(default_factory=dict)


def image_from_binms(path, grid_size: int, planes: int, kernel=KAISERBESSEL, support: int = 7,
                     precision: int = FP64, drop: int = binms.DROP_ROWS, normalize: bool = True,
                     phase_on: bool = True, device: int = 0, out_dir=None,
                     write_data: bool = False) -> Result:
    import torch
    t = {}

    def mark(name, t0):
        torch.cuda.synchronize(device)
        t[name] = (time.perf_counter() - t0) * 1e3
        return time.perf_counter()

    t0 = time.perf_counter()
    meta = binms.read_meta(path)
    startrow, nrows = binms.partition(meta, 0, 1, drop)
    d = binms.read(path, meta, startrow, nrows, device=device)
    t0 = mark("read", t0)
    nx = ny = int(grid_size)
    dx, dw = 1.0 / nx, 1.0 / planes                      # init.c:11-12
    if normalize:
        prep.normalize_uvw(d["uu"], d["vv"], d["ww"])
    # One rank: one sector, the whole grid (yaxis = Ny); the reference still
    # routes the rows through the binning and walks them backwards.
    bins = prep.sector_bin(d["vv"], 1, ny, dx, support)
    rows = prep.sector_gather(bins.sector(0), 0, ny, dx, meta.freq_per_chan, meta.polarisations,
                              d["uu"], d["vv"], d["ww"], d["weight"], d["vis_real"],
                              d["vis_img"])
    del d
    t0 = mark("prepare", t0)
    kind = KERNELS.get(kernel, kernel) if isinstance(kernel, str) else kernel
    n = int(rows["uu"].numel())
    plan = Plan(nx, ny, planes, kernel=kind, support=support, dx=dx, dw=dw, precision=precision,
                max_points=max(n, 1), resolution=meta.resolution, wmin=meta.wmin,
                wmax=meta.wmax, phase_on=phase_on, device=device)
    grid = plan.grid(rows["uu"], rows["vv"], rows["ww"], rows["vis_real"], rows["vis_img"],
                     rows["weight"], meta.freq_per_chan, meta.polarisations)
    t0 = mark("grid", t0)
    if out_dir is not None and write_data:
        os.makedirs(out_dir, exist_ok=True)
        binms.write_cube(os.path.join(out_dir, "gridded_data_real.bin"),
                         os.path.join(out_dir, "gridded_data_img.bin"), grid, planes, nx, ny)
        t0 = time.perf_counter()
    plan.fft(grid)
    t0 = mark("fft", t0)
    if out_dir is not None and write_data:
        binms.write_cube(os.path.join(out_dir, "ffted_real.bin"),
                         os.path.join(out_dir, "ffted_img.bin"), grid, planes, nx, ny)
        t0 = time.perf_counter()
    re, im = plan.phase(grid)
    t0 = mark("phase", t0)
    if out_dir is not None:
        os.makedirs(out_dir, exist_ok=True)
        binms.write_image(os.path.join(out_dir, "fft_real.bin"),
                          os.path.join(out_dir, "fft_img.bin"), re, im, nx, ny, 0)
        t0 = mark("write", t0)
    return Result(re, im, grid, meta, n, t)


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("path")
    ap.add_argument("--grid", type=int, default=4096)
    ap.add_argument("--planes", type=int, default=4)
    ap.add_argument("--kernel", default="kb")
    ap.add_argument("--support", type=int, default=7)
    ap.add_argument("--precision", choices=["fp64", "fp32"], default="fp64")
    ap.add_argument("--drop", type=int, default=binms.DROP_ROWS)
    ap.add_argument("--no-normalize", action="store_true")
    ap.add_argument("--write-data", action="store_true")
    ap.add_argument("--out", default=None)
    a = ap.parse_args()
    r = image_from_binms(a.path, a.grid, a.planes, kernel=a.kernel, support=a.support,
                         precision=FP32 if a.precision == "fp32" else FP64, drop=a.drop,
                         normalize=not a.no_normalize, out_dir=a.out, write_data=a.write_data)
    print(json.dumps({"rows": r.rows, "timings_ms": r.timings_ms}))


if __name__ == "__main__":
    main()
