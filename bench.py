"""Benchmark of the imaging chain on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config lofar]

Workload (default): the LOFAR-like observation -- 62 stations (dense 2 km core,
log-radial to 50 km), 1,891 baselines x 4,096 times x 16 channels
(120-136 MHz) x 1 polarization = 123,928,576 visibilities, Kaiser-Bessel
support 7, a 4096 x 4096 grid with 32 w-planes -- generated synthetically on
the GPU (seeded), inputs resident in HBM before timing.

A step is one gridding pass (wstack) over all visibilities.  It is timed for
the FP64 path (the reference's arithmetic) and the FP32 path; ``value`` is the
FP32 path's gridded visibilities per second.  The whole image chain
(grid -> per-plane FFT -> phase correction) is timed beside it, and the FP32
image is compared with the FP64 one.  Per-kernel durations come from HIP
events on the launch stream (Plan.profile); the roofline object describes the
dominant gridding kernel.  The CPU baseline is the reference's own wstack()
(compiled from its sources into oracle/_ref) with OpenMP on this host's cores
over a bounded sample of the same rows.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, AMD published


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="lofar")
    ap.add_argument("--cpu-seconds", type=float, default=12.0,
                    help="target CPU time of the bounded CPU-baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--only", choices=["fp64", "fp32"], default=None,
                    help="run one path only (for profiler runs)")
    ap.add_argument("--tile", default=None, help="gridder tile TUxTV (default: tuned/64x64)")
    ap.add_argument("--rows-per-task", type=int, default=0)
    ap.add_argument("--tune", action="store_true",
                    help="autotune tile shape and rows/task first (rick_amd.autotune)")
    ap.add_argument("--no-prep", action="store_true",
                    help="skip the device preparation rates (normalize, binning, gather)")
    ap.add_argument("--binms", default=None, metavar="DIR",
                    help="also time a binMS write and read-back of the workload under DIR")
    ap.add_argument("--no-dropin", action="store_true",
                    help="skip the host-pointer wstack() drop-in timing (PCIe-inclusive)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="first run one 1 GiB device-to-device copy (known HBM bytes) so that "
                         "profiler byte counters can be calibrated against it")
    return ap.parse_args()


def measured_traffic(kernel, workload, precision, mode="dense"):
    """HBM bytes per dispatch of `kernel` from the committed PMC summary
    profiles/<kernel>_<workload>_traffic.txt (FETCH_SIZE/WRITE_SIZE runs scaled
    by the copy-probe factors), or None when there is none."""
    path = ROOT / "profiles" / f"k_{kernel}_{workload}_traffic.txt"
    try:
        for line in path.read_text().splitlines():
            f = line.split()
            if len(f) == 4 and not line.startswith("#") and f[0] == precision and f[1] == mode:
                return int(f[2]) + int(f[3])
    except OSError:
        pass
    return None


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def timed(fn, steps, torch):
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) * 1e3 / steps


def run_path(ds, ins, precision, steps, warmup, torch, tile=(0, 0), rpt=0):
    from rick_amd import imaging
    cfg = ds.cfg
    F, P = cfg.n_chan, cfg.n_pol
    plan = imaging.plan_for(ds, precision=precision, tile=tile, rows_per_task=rpt)
    grid = plan.new_grid()
    npix = cfg.grid * cfg.grid
    re = torch.empty(npix, dtype=torch.float64, device="cuda:0")
    im = torch.empty(npix, dtype=torch.float64, device="cuda:0")

    def grid_step():
        plan.grid(*ins, F, P, out=grid)

    def image_step():
        plan.image(*ins, F, P, grid=grid, image_real=re, image_imag=im)

    for _ in range(warmup):
        grid_step()
        image_step()
    grid_ms = timed(grid_step, steps, torch)
    image_ms = timed(image_step, steps, torch)

    # Per-kernel durations (HIP events on the launch stream), for the
    # gridding step (dense grid) and for the image chain (sparse grid).
    def stage_profile(fn):
        plan.profile(True)
        for _ in range(steps):
            fn()
        st = plan.stage_times()
        plan.profile(False)
        return {k: {"ms_per_launch": (ms / n if n else 0.0), "launches_per_step": n / steps,
                    "ms_per_step": ms / steps}
                for k, (ms, n) in st.items() if n}

    stages = stage_profile(grid_step)
    stats = plan.grid_stats()
    image_stages = stage_profile(image_step)
    img = (re.cpu().numpy() + 1j * im.cpu().numpy())
    return plan, grid_ms, image_ms, stages, image_stages, stats, img


def cpu_baseline(ds_host_fn, cfg, target_s):
    """Reference wstack() (oracle/_ref) on a bounded sample of the same rows."""
    from oracle import oracle as O
    threads = max(1, min(16, len(os.sched_getaffinity(0))))
    kind = "reference" if O.ref_available() else "port"
    nx = cfg.grid
    grid = np.zeros(2 * nx * nx * (cfg.planes + 1), dtype=np.float64)

    def run(n):
        uu, vv, ww, vr, vi, wt = ds_host_fn(n)
        t0 = time.perf_counter()
        if kind == "reference":
            O.ref_wstack_into(cfg.kernel, cfg.planes, uu, vv, ww, vr, vi, wt, cfg.n_chan,
                              cfg.n_pol, 1.0 / nx, 1.0 / cfg.planes, cfg.support, nx, nx, grid,
                              threads)
        else:
            O.wstack(cfg.kernel, cfg.planes, uu, vv, ww, vr, vi, wt, cfg.n_chan, cfg.n_pol,
                     1.0 / nx, 1.0 / cfg.planes, cfg.support, nx, nx, grid=grid, threads=threads)
        return time.perf_counter() - t0

    n = 16384
    t = run(n)
    while t < 1.0 and n < cfg.n_rows:
        n = min(cfg.n_rows, n * 4)
        t = run(n)
    n = int(min(cfg.n_rows, max(n, n * target_s / max(t, 1e-6))))
    t = run(n)
    vis = n * cfg.n_chan * cfg.n_pol
    return {"value": vis / t, "unit": "vis/s", "cores": threads, "kind": kind,
            "sample": f"{n} of {cfg.n_rows} rows ({vis} vis, time-major block from the middle of "
                      f"the observation) gridded by the reference's wstack() "
                      f"(KAISERBESSEL, support {cfg.support}) into the full {nx}^2 x "
                      f"{cfg.planes}-plane FP64 grid with {threads} OpenMP threads; "
                      f"{t:.2f} s", "cpu": _cpu_model()}


def dropin_rate(ds, cfg, torch):
    """The host-pointer drop-in wstack() on the full workload: inputs copied
    to the GPU, gridded, and the device grid added into the caller's host
    grid (FP64).  PCIe-inclusive; reported beside ``value``, never as it."""
    from rick_amd import imaging
    h = ds.to_numpy()
    nx = cfg.grid
    grid = np.zeros(2 * cfg.planes * nx * nx, dtype=np.float64)
    args = (cfg.planes, cfg.n_rows, cfg.n_chan, cfg.n_pol, h.uu, h.vv, h.ww, h.vis_real,
            h.vis_img, h.weight, 1.0 / nx, 1.0 / cfg.planes, cfg.support, nx, nx, grid)
    threads = max(1, min(16, len(os.sched_getaffinity(0))))
    imaging.set_kernel(cfg.kernel)
    imaging.wstack(*args, num_threads=threads)          # warm-up: plan + pinned buffers
    t0 = time.perf_counter()
    imaging.wstack(*args, num_threads=threads)
    t = time.perf_counter() - t0
    del grid, h
    return {"vis_per_s": cfg.n_vis / t, "s_per_call": t, "host_threads": threads,
            "what": "wstack() on host arrays: H2D of uvw/vis/weights, FP64 gridding, D2H of "
                    "the FP64 grid through pinned chunks, += into the host grid"}


def prep_rates(ds, cfg, torch, steps=5):
    """Device preparation in front of the gridder on the workload's rows:
    NORMALIZE_UVW (in place, on copies), sector binning for one GPU and for
    an 8-slab split, and the gather of one GPU's sector.  Timed with events on
    the current stream; algorithmic bytes per call are stated with each."""
    from rick_amd import prep
    rows, F, P = cfg.n_rows, cfg.n_chan, cfg.n_pol
    dx = 1.0 / cfg.grid
    out = {}

    def clock(fn, reset=None):
        ts = []
        for i in range(steps + 1):
            if reset:
                reset()
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record()
            fn()
            b.record()
            b.synchronize()
            if i:
                ts.append(a.elapsed_time(b))
        return sum(ts) / len(ts)

    u, v, w = ds.uu.clone(), ds.vv.clone(), ds.ww.clone()
    ws = torch.empty(1 << 20, dtype=torch.uint8, device="cuda:0")

    def reset():
        u.copy_(ds.uu), v.copy_(ds.vv), w.copy_(ds.ww)

    ms = clock(lambda: prep.normalize_uvw(u, v, w, ws), reset)
    alg = rows * 72            # extrema pass reads u,v,w; apply reads and writes them
    out["normalize_uvw"] = {"ms": ms, "alg_bytes": alg, "GBs": alg / ms / 1e6,
                            "frac_of_peak": alg / ms / 1e6 / HBM_PEAK_GBS}
    for S in (1, 8):
        bins = {}

        def run():
            bins["b"] = prep.sector_bin(ds.vv, S, cfg.grid // S, dx, cfg.support)
        ms = clock(run)
        b = bins["b"]
        alg = rows * 16 + b.total * 4        # v read by the count and fill passes, lists written
        out[f"sector_bin_{S}"] = {"ms": ms, "entries": b.total, "alg_bytes": alg,
                                  "GBs": alg / ms / 1e6,
                                  "note": "includes the stream sync and the total's readback"}
    b = prep.sector_bin(ds.vv, 1, cfg.grid, dx, cfg.support)
    lst = b.sector(0)
    ms = clock(lambda: prep.sector_gather(lst, 0, cfg.grid, dx, F, P, ds.uu, ds.vv, ds.ww,
                                          ds.weight, ds.vis_real, ds.vis_img))
    per_row = 24 + 4 * P + 8 * F * P
    alg = rows * (4 + 2 * per_row)           # list read, row data read and written
    out["sector_gather"] = {"ms": ms, "alg_bytes": alg, "GBs": alg / ms / 1e6,
                            "frac_of_peak": alg / ms / 1e6 / HBM_PEAK_GBS,
                            "note": "includes the output allocations of the Python wrapper"}
    return out


def binms_rates(ds, cfg, torch, path):
    """binMS round trip on the workload's rows (physical values do not matter
    for the rate): write the six files, read them back into host memory and
    into device memory.  Rates count the bytes of the six files; the page
    cache will usually hold them, so this is memory and PCIe, not disk."""
    import shutil

    from rick_amd import binms
    h = ds.to_numpy()
    F, P = cfg.n_chan, cfg.n_pol
    t0 = time.perf_counter()
    meta = binms.write(path, h.uu, h.vv, h.ww, h.weight, h.vis_real, h.vis_img, F, P)
    tw = time.perf_counter() - t0
    nbytes = sum(os.path.getsize(os.path.join(path, f)) for f in binms.FILES)
    t0 = time.perf_counter()
    binms.read(path, meta, 0, meta.nmeasures)
    th = time.perf_counter() - t0
    binms.read(path, meta, 0, meta.nmeasures, device=0)   # warm-up: pinned buffers
    t0 = time.perf_counter()
    d = binms.read(path, meta, 0, meta.nmeasures, device=0)
    torch.cuda.synchronize()
    td = time.perf_counter() - t0
    ok = bool(torch.equal(d["vis_img"], ds.vis_img))
    del d, h
    shutil.rmtree(path, ignore_errors=True)
    return {"bytes": nbytes, "write_GBs": nbytes / tw / 1e9, "read_host_GBs": nbytes / th / 1e9,
            "read_device_GBs": nbytes / td / 1e9, "device_read_matches": ok}


def copy_bandwidth(torch, gib=1, reps=5):
    """Device-to-device copy of ``gib`` GiB (read + write counted), timed with
    events: the practical HBM ceiling beside the published 8 TB/s."""
    n = (gib << 30) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda:0")
    dst = torch.empty_like(src)
    dst.copy_(src)
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(reps):
        dst.copy_(src)
    b.record()
    b.synchronize()
    ms = a.elapsed_time(b) / reps
    del src, dst
    torch.cuda.empty_cache()
    return 2 * (gib << 30) / ms / 1e6


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark measures one GPU; --gpus must be 1 (the "
                          "multi-GPU design shards w-planes and is not built)"}))
        sys.exit(2)
    import torch

    from rick_amd import FP32, FP64, autotune, synth

    cfg = synth.CONFIGS[args.config]
    if args.copy_probe:
        src = torch.ones(1 << 28, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log("copy probe: 1 GiB read + 1 GiB written by one copy kernel")
        del src, dst
    log(f"generating {cfg.name}: {cfg.n_rows} rows x {cfg.n_chan} ch = {cfg.n_vis} vis")
    ds = synth.make_dataset(cfg, device="cuda:0")
    ins = [ds.uu, ds.vv, ds.ww, ds.vis_real, ds.vis_img, ds.weight]
    torch.cuda.synchronize()
    vis = cfg.n_vis
    res = {}
    for name, prec in (("fp64", FP64), ("fp32", FP32)):
        if args.only and args.only != name:
            continue
        tile, rpt = (0, 0), args.rows_per_task
        if args.tile:
            tile = tuple(int(x) for x in args.tile.lower().split("x"))
        elif args.tune:
            best = autotune.tune_grid(ds, prec, log=log)["best"]
            tile, rpt = tuple(best["tile"]), best["rows_per_task"]
        else:
            tuned = autotune.lookup(cfg.grid, cfg.planes, cfg.support, prec, cfg.n_rows)
            if tuned:
                tile, rpt = tuned
        log(f"{name}: tile {tile} rows/task {rpt}; warmup {args.warmup}, steps {args.steps}")
        plan, gms, ims, stages, istages, stats, img = run_path(ds, ins, prec, args.steps,
                                                               args.warmup, torch, tile, rpt)
        res[name] = {"grid_ms": gms, "image_ms": ims, "grid_vis_per_s": vis / (gms * 1e-3),
                     "stages": stages, "image_stages": istages, "grid_stats": stats,
                     "img": img,
                     "grid_bytes": plan.grid_bytes,
                     "tile": [plan.desc.tile_u or 64, plan.desc.tile_v or 64],
                     "rows_per_task": plan.desc.rows_per_task or 4096}
        log(f"{name}: grid {gms:.3f} ms ({vis / gms / 1e6:.3f} Gvis/s), image {ims:.3f} ms")
        del plan
        torch.cuda.empty_cache()

    copy_gbs = copy_bandwidth(torch)
    log(f"device copy: {copy_gbs:.0f} GB/s (read + write)")
    main_path = "fp32" if "fp32" in res else "fp64"
    r = res[main_path]
    rows, F, P = cfg.n_rows, cfg.n_chan, cfg.n_pol
    cb = 8 if main_path == "fp32" else 16
    # Dominant gridding kernel and its algorithmic bytes.
    grid_stages = {k: v for k, v in r["stages"].items()
                   if k in ("count", "scan", "scatter", "zero_split", "tile")}
    dom = max(grid_stages, key=lambda k: grid_stages[k]["ms_per_step"])
    grid_bytes = cfg.planes * cfg.grid * cfg.grid * cb
    rec_bytes = 32 if main_path == "fp64" else 24
    records = r["grid_stats"]["records"]
    alg = {"tile": grid_bytes + records * rec_bytes,        # every cell written once, records read
           "scatter": rows * (24 + 4 * P) + vis * 8 + records * rec_bytes,  # inputs read, records written
           "count": rows * 24}.get(dom)
    t_dom = grid_stages[dom]["ms_per_launch"] * 1e-3
    achieved = alg / t_dom / 1e9 if alg else None
    stage_alg = rows * (24 + 4 * P) + vis * 8 + grid_bytes
    out = {
        "metric": "gridded vis/s + image time, FP32 vs FP64 path at rel-L2<=1e-5; serve p99",
        "value": r["grid_vis_per_s"],
        "unit": "vis/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["grid_ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if main_path == "fp32" else "f64",
        "data": "synthetic (seeded Earth-rotation uvw, 20 point sources + noise), generated on "
                "the GPU",
        "config": {"workload": f"{cfg.name}-like: {cfg.n_ant} stations, {cfg.n_baselines} "
                               f"baselines x {cfg.n_times} times x {cfg.n_chan} channels = "
                               f"{vis} visibilities; {cfg.grid}^2 x {cfg.planes} w-planes, "
                               f"Kaiser-Bessel support {cfg.support}",
                   "step": "one wstack() gridding pass over all visibilities",
                   "path": main_path},
        "roofline": {"bound": "hbm", "kernel": dom,
                     "measured_copy_GBs": copy_gbs,
                     "frac_of_measured_copy": (achieved / copy_gbs) if achieved else None,
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": (achieved / HBM_PEAK_GBS) if achieved else None,
                     "traffic": measured_traffic(dom, cfg.name, main_path),
                     "traffic_source": f"profiles/k_{dom}_{cfg.name}_traffic.txt",
                     "alg_bytes_per_launch": alg,
                     "ms_per_launch": grid_stages[dom]["ms_per_launch"]},
        "gridding_stage": {"alg_bytes": stage_alg,
                           "achieved_GBs": stage_alg / (r["grid_ms"] * 1e-3) / 1e9,
                           "frac_of_peak": stage_alg / (r["grid_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS},
    }
    for name, v in res.items():
        out[name] = {"grid_ms": v["grid_ms"], "image_ms": v["image_ms"],
                     "grid_vis_per_s": v["grid_vis_per_s"], "grid_stats": v["grid_stats"],
                     "tile": v["tile"], "rows_per_task": v["rows_per_task"],
                     "stages_ms_per_step": {k: round(s["ms_per_step"], 4)
                                            for k, s in v["stages"].items()},
                     "image_stages_ms_per_step": {k: round(s["ms_per_step"], 4)
                                                  for k, s in v["image_stages"].items()}}
    if "fp32" in res and "fp64" in res:
        a, b = res["fp32"]["img"], res["fp64"]["img"]
        out["fp32_vs_fp64"] = {
            "grid_speedup": res["fp64"]["grid_ms"] / res["fp32"]["grid_ms"],
            "image_speedup": res["fp64"]["image_ms"] / res["fp32"]["image_ms"],
            "image_rel_l2": float(np.linalg.norm(a - b) / np.linalg.norm(b)),
            "image_peak_rel": float(np.abs(a - b).max() / np.abs(b).max()),
        }
    if not args.no_prep:
        log("device preparation: normalize, sector binning, gather")
        out["prep"] = prep_rates(ds, cfg, torch)
    if args.binms:
        log(f"binMS write and read-back under {args.binms}")
        out["binms"] = binms_rates(ds, cfg, torch, args.binms)
    if not args.no_dropin and not args.only:
        log("drop-in wstack() on host arrays")
        out["dropin_pcie_inclusive"] = dropin_rate(ds, cfg, torch)
    if not args.no_cpu:
        log("cpu baseline")
        mid = cfg.n_rows // 2

        def host_rows(n):
            s = max(0, min(mid, cfg.n_rows - n))
            sub = ds.subset(slice(s, s + n)).to_numpy()
            return (np.ascontiguousarray(sub.uu), np.ascontiguousarray(sub.vv),
                    np.ascontiguousarray(sub.ww), np.ascontiguousarray(sub.vis_real),
                    np.ascontiguousarray(sub.vis_img), np.ascontiguousarray(sub.weight))

        out["cpu_baseline"] = cpu_baseline(host_rows, cfg, args.cpu_seconds)
        out["gpu_vs_cpu"] = out["value"] / out["cpu_baseline"]["value"]
    print(json.dumps(out))


if __name__ == "__main__":
    main()
