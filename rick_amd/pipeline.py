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
                     write_data: bool = False, nsectors: int = 1) -> Result:
    """``nsectors`` > 1 grids the rows as that many ranks would: v-slabs of
    grid_size/nsectors rows, each with its halo rows (gridding_cpu.c:90-204),
    gridded one after the other on this GPU into their place in the grid (the
    reduce onto the slab's owner becomes a placement)."""
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
    if nsectors < 1 or ny % nsectors:
        raise ValueError(f"{nsectors} sectors do not divide {ny} grid rows")
    yaxis = ny // nsectors
    # With one sector the whole grid is one slab (yaxis = Ny); the reference
    # still routes the rows through the binning and walks them backwards.
    bins = prep.sector_bin(d["vv"], nsectors, yaxis, dx, support)
    sectors = [prep.sector_gather(bins.sector(s), s, yaxis, dx, meta.freq_per_chan,
                                  meta.polarisations, d["uu"], d["vv"], d["ww"], d["weight"],
                                  d["vis_real"], d["vis_img"]) for s in range(nsectors)]
    del d
    t0 = mark("prepare", t0)
    kind = KERNELS.get(kernel, kernel) if isinstance(kernel, str) else kernel
    n = int(sum(int(r["uu"].numel()) for r in sectors))   # halo rows count once per slab
    common = dict(kernel=kind, support=support, dx=dx, dw=dw, precision=precision,
                  resolution=meta.resolution, wmin=meta.wmin, wmax=meta.wmax,
                  phase_on=phase_on, device=device)
    plan = Plan(nx, ny, planes, max_points=1 if nsectors > 1 else max(n, 1), **common)
    F, P = meta.freq_per_chan, meta.polarisations
    if nsectors == 1:
        r = sectors[0]
        grid = plan.grid(r["uu"], r["vv"], r["ww"], r["vis_real"], r["vis_img"], r["weight"], F, P)
    else:
        grid = plan.new_grid()
        g3 = grid.view(planes, ny, 2 * nx)
        cap = max(1, max(int(r["uu"].numel()) for r in sectors))
        slab_plan = Plan(nx, yaxis, planes, max_points=cap, **common)
        for s, r in enumerate(sectors):
            slab = slab_plan.grid(r["uu"], r["vv"], r["ww"], r["vis_real"], r["vis_img"],
                                  r["weight"], F, P)
            g3[:, s * yaxis:(s + 1) * yaxis, :] = slab.view(planes, yaxis, 2 * nx)
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
    t["entries"] = n
    return Result(re, im, grid, meta, nrows, t)


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
    ap.add_argument("--sectors", type=int, default=1,
                    help="grid as this many v-slabs, as that many ranks would")
    a = ap.parse_args()
    r = image_from_binms(a.path, a.grid, a.planes, kernel=a.kernel, support=a.support,
                         precision=FP32 if a.precision == "fp32" else FP64, drop=a.drop,
                         normalize=not a.no_normalize, out_dir=a.out, write_data=a.write_data,
                         nsectors=a.sectors)
    print(json.dumps({"rows": r.rows, "timings_ms": r.timings_ms}))


if __name__ == "__main__":
    main()
