"""Autotuner for the gridder's launch shape.

The gridder's tile shape (TU x TV cells per LDS tile, i.e. per bucket) and the
records per task (rows_per_task) trade occupancy (LDS per workgroup), record
duplication across tile borders and load balance on hot tiles; the footprint
unroll (the compile-time length of the tile kernel's footprint loops, any
instantiated length >= the support) trades registers against dead
iterations.  This module times the candidates on the actual workload and
caches the fastest per workload signature in a JSON file
(rick_amd/tuning.json by default).  The image chain is tuned on its own
(``--image``): there the grid is left sparse and the row FFT skips empty tile
rows, so the shape that grids fastest is not always the one that images
fastest.

    python -m rick_amd.autotune --config lofar --precision fp32
    python -m rick_amd.autotune --config lofar --precision fp32 --image
    python -m rick_amd.autotune --batch 96 --precision fp32   # the batched job mix
"""
from __future__ import annotations

import argparse
import json
import os
import time
from pathlib import Path

TILES = [(64, 64), (64, 32), (32, 32), (128, 32), (32, 64)]
ROWS_PER_TASK = [1024, 4096, 16384]
UNROLLS = [7, 9, 11, 15]     # instantiated footprint unrolls (rick_plan_desc.footprint_unroll)
CACHE = Path(__file__).resolve().parent / "tuning.json"


def signature(grid: int, planes: int, support: int, precision: int, rows: int,
              kind: str = "grid") -> str:
    """Tuning key: grid, planes, support, precision and log2-bucketed row count;
    image-chain entries carry the suffix _img."""
    key = f"g{grid}_p{planes}_s{support}_{'fp32' if precision else 'fp64'}_r{rows.bit_length()}"
    return key + ("_img" if kind == "image" else "")


def load(path: Path = CACHE) -> dict:
    try:
        return json.loads(Path(path).read_text())
    except (OSError, ValueError):
        return {}


def lookup(grid, planes, support, precision, rows, path: Path = CACHE, kind: str = "grid"):
    """(tile, rows_per_task, unroll) tuned for this shape (kind "grid": the
    gridding step, "image": the image chain), or None; entries tuned before
    the unroll was a candidate give unroll 0 (the smallest)."""
    e = load(path).get(signature(grid, planes, support, precision, rows, kind))
    if not e:
        return None
    return tuple(e["tile"]), int(e["rows_per_task"]), int(e.get("unroll", 0))


def unrolls_for(support: int) -> list:
    return [u for u in UNROLLS if u >= support]


def tune_grid(ds, precision, tiles=None, rpts=None, steps=3, log=print, unrolls=None):
    """Time rick_plan_grid for each candidate on dataset ``ds`` (inputs on the
    GPU): tile shapes x rows per task at the smallest unroll, then the other
    unrolls at the best shape.  Returns {"best": {...}, "results": [...]}."""
    import torch

    from .imaging import plan_for
    cfg = ds.cfg
    ins = [ds.uu, ds.vv, ds.ww, ds.vis_real, ds.vis_img, ds.weight]
    results = []

    def run(tile, rpt, unroll):
        plan = plan_for(ds, precision=precision, tile=tile, rows_per_task=rpt, unroll=unroll)
        grid = plan.new_grid()
        plan.grid(*ins, cfg.n_chan, cfg.n_pol, out=grid)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            plan.grid(*ins, cfg.n_chan, cfg.n_pol, out=grid)
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / steps
        results.append({"tile": list(tile), "rows_per_task": rpt, "unroll": unroll, "ms": ms})
        log(f"[tune] tile {tile[0]}x{tile[1]} rows/task {rpt} unroll {unroll}: {ms:.3f} ms")
        del grid, plan
        torch.cuda.empty_cache()

    cand = unrolls or unrolls_for(cfg.support)
    for tile in tiles or TILES:
        if tile[0] < cfg.support or tile[1] < cfg.support:
            continue
        for rpt in rpts or ROWS_PER_TASK:
            run(tile, rpt, cand[0])
    shape = min(results, key=lambda r: r["ms"])
    for u in cand[1:]:
        run(tuple(shape["tile"]), shape["rows_per_task"], u)
    best = min(results, key=lambda r: r["ms"])
    return {"best": best, "results": results}


def tune_image(ds, precision, tiles=None, rpts=None, steps=3, log=print, unroll=0):
    """Time the image chain (rick_plan_image) for each tile shape x rows per
    task on dataset ``ds``.  Returns {"best": {...}, "results": [...]}."""
    import torch

    from .imaging import plan_for
    cfg = ds.cfg
    ins = [ds.uu, ds.vv, ds.ww, ds.vis_real, ds.vis_img, ds.weight]
    npix = cfg.grid * cfg.grid
    re = torch.empty(npix, dtype=torch.float64, device="cuda:0")
    im = torch.empty_like(re)
    results = []
    for tile in tiles or TILES:
        if tile[0] < cfg.support or tile[1] < cfg.support:
            continue
        for rpt in rpts or ROWS_PER_TASK:
            plan = plan_for(ds, precision=precision, tile=tile, rows_per_task=rpt, unroll=unroll)
            grid = plan.new_grid()
            plan.image(*ins, cfg.n_chan, cfg.n_pol, grid=grid, image_real=re, image_imag=im)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(steps):
                plan.image(*ins, cfg.n_chan, cfg.n_pol, grid=grid, image_real=re, image_imag=im)
            torch.cuda.synchronize()
            ms = (time.perf_counter() - t0) * 1e3 / steps
            results.append({"tile": list(tile), "rows_per_task": rpt, "unroll": unroll, "ms": ms})
            log(f"[tune image] tile {tile[0]}x{tile[1]} rows/task {rpt}: {ms:.3f} ms")
            del grid, plan
            torch.cuda.empty_cache()
    best = min(results, key=lambda r: r["ms"])
    return {"best": best, "results": results}


def store(key: str, best: dict, path: Path = CACHE) -> None:
    data = load(path)
    data[key] = {"tile": best["tile"], "rows_per_task": best["rows_per_task"],
                 "unroll": best.get("unroll", 0), "ms": best["ms"]}
    tmp = Path(str(path) + ".tmp")
    tmp.write_text(json.dumps(data, indent=1, sort_keys=True))
    os.replace(tmp, path)


def tune_batch(n_jobs, precision, out, steps=2, log=print, seed=3):
    """Tune every distinct signature of the batched-imaging job mix
    (rick_amd.batch.job_configs) and store each best shape."""
    import torch

    from . import synth
    from .batch import job_configs
    seen = {}
    for cfg in job_configs(n_jobs, seed):
        key = signature(cfg.grid, cfg.planes, cfg.support, precision, cfg.n_rows)
        if key in seen:
            continue
        ds = synth.make_dataset(cfg, device="cuda:0")
        r = tune_grid(ds, precision, steps=steps, log=lambda m: None)
        store(key, r["best"], Path(out))
        seen[key] = r["best"]
        log(f"[tune] {key}: tile {r['best']['tile']} rows/task {r['best']['rows_per_task']} "
            f"unroll {r['best']['unroll']} {r['best']['ms']:.3f} ms")
        del ds
        torch.cuda.empty_cache()
    return seen


def main():
    from . import FP32, FP64, synth
    ap = argparse.ArgumentParser()
    ap.add_argument("--config", default="lofar")
    ap.add_argument("--precision", choices=["fp64", "fp32"], default="fp32")
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--batch", type=int, default=0,
                    help="tune the signatures of this many batched-imaging jobs instead")
    ap.add_argument("--image", action="store_true",
                    help="tune the image chain (sparse grid + FFT + phase) instead of gridding")
    ap.add_argument("--out", default=str(CACHE))
    args = ap.parse_args()
    prec = FP32 if args.precision == "fp32" else FP64
    if args.batch:
        best = tune_batch(args.batch, prec, args.out, steps=max(1, args.steps - 1))
        print(json.dumps({"batch_jobs": args.batch, "signatures": len(best)}))
        return
    cfg = synth.CONFIGS[args.config]
    ds = synth.make_dataset(cfg, device="cuda:0")
    if args.image:
        g = lookup(cfg.grid, cfg.planes, cfg.support, prec, cfg.n_rows, Path(args.out))
        r = tune_image(ds, prec, steps=args.steps, unroll=g[2] if g else 0)
        key = signature(cfg.grid, cfg.planes, cfg.support, prec, cfg.n_rows, "image")
        store(key, r["best"], Path(args.out))
        print(json.dumps({"key": key, **r}))
        return
    r = tune_grid(ds, prec, steps=args.steps)
    key = signature(cfg.grid, cfg.planes, cfg.support, prec, cfg.n_rows)
    store(key, r["best"], Path(args.out))
    print(json.dumps({"key": key, **r}))


if __name__ == "__main__":
    main()
