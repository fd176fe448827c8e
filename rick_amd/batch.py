"""Batched imaging: many independent jobs of uneven size in one graph launch.

Each job is a full imaging chain (gridding -> per-plane FFT -> phase
correction and plane sum) with its own grid size, w-planes, convolution kernel
and visibility count.  Every job gets a plan sized for it (tile shape and
rows per task from the autotuner cache when present); the chains are issued
round-robin on a few streams so that small jobs overlap, and the whole batch is
captured once into a HIP graph: a replay is one launch of every kernel of every
job, with no host work in between.

    python -m rick_amd.batch --jobs 96 --precision fp32      # prints one JSON line
"""
from __future__ import annotations

import argparse
import ctypes
import json
import math
import sys
import time
from dataclasses import dataclass

import numpy as np

from . import autotune, synth
from ._lib import FP32, FP64, GAUSS, KAISERBESSEL
from .imaging import plan_for

GRIDS = (256, 512, 1024, 2048)
KERNELS = ((GAUSS, 7), (KAISERBESSEL, 9))
BASELINES_64 = 64 * 63 // 2


def job_configs(n_jobs: int = 96, seed: int = 3, vis_lo: float = 5e4, vis_hi: float = 8e6):
    """Job mix: grid in {256..2048}, 4-16 w-planes, visibility counts
    log-uniform in [vis_lo, vis_hi] from Earth-rotation tracks of 64 antennas
    (2,016 baselines) at 150 MHz and declinations 0..80 deg, Gaussian support 7
    or Kaiser-Bessel support 9.  The array is a 1.5 km-radius disk scaled with
    the grid (x grid/512 above 512): the phase stage's pixel angle reaches
    (N-1) * resolution with resolution = 1/max|uv| (phase_correction.cu:204-
    209), so a grid much wider than the longest baseline in wavelengths would
    put most pixels beyond l^2 + m^2 = 1 (NaN, as in the reference)."""
    rng = np.random.default_rng(seed)
    out = []
    for j in range(n_jobs):
        grid = int(rng.choice(GRIDS))
        planes = int(rng.integers(4, 17))
        vis = math.exp(rng.uniform(math.log(vis_lo), math.log(vis_hi)))
        kernel, support = KERNELS[int(rng.integers(0, 2))]
        n_times = max(1, int(round(vis / BASELINES_64)))
        dec = float(rng.uniform(0.0, 80.0))
        radius = 1500.0 * max(1.0, grid / 512)
        out.append(synth.ObsConfig(f"job{j}", "disk", 64, radius, 45.0, n_times, -3.0, 3.0, dec,
                                   150e6, 150e6, 1, 1, 5, grid, planes, kernel, support,
                                   seed=seed * 1000 + j))
    return out


@dataclass
class Job:
    ds: object          # synth.Dataset on the GPU
    plan: object        # imaging.Plan
    grid: object        # device grid (scratch of the chain)
    re: object          # device image, real part
    im: object          # device image, imaginary part

    @property
    def vis(self) -> int:
        return self.ds.cfg.n_vis

    def run(self):
        ds = self.ds
        self.plan.image(ds.uu, ds.vv, ds.ww, ds.vis_real, ds.vis_img, ds.weight,
                        ds.cfg.n_chan, ds.cfg.n_pol, grid=self.grid, image_real=self.re,
                        image_imag=self.im)


def make_jobs(cfgs, precision, device="cuda:0", tuned=True, datasets=None):
    """Plans and buffers for each job (allocations happen here, not per run)."""
    import torch
    jobs = []
    for i, cfg in enumerate(cfgs):
        ds = datasets[i] if datasets is not None else synth.make_dataset(cfg, device=device)
        tile, rpt, unroll = (0, 0), 0, 0
        if tuned:
            t = autotune.lookup(cfg.grid, cfg.planes, cfg.support, precision, cfg.n_rows)
            if t:
                tile, rpt, unroll = t
        plan = plan_for(ds, precision=precision, tile=tile, rows_per_task=rpt, unroll=unroll)
        npix = cfg.grid * cfg.grid
        jobs.append(Job(ds, plan, plan.new_grid(),
                        torch.empty(npix, dtype=torch.float64, device=device),
                        torch.empty(npix, dtype=torch.float64, device=device)))
    return jobs


class Batch:
    """Runs a list of jobs on `n_streams` streams, eagerly or as one graph."""

    def __init__(self, jobs, n_streams=4):
        import torch
        self.jobs = jobs
        self.streams = [torch.cuda.Stream() for _ in range(n_streams)]
        self.graph = None

    def _issue(self):
        import torch
        main = torch.cuda.current_stream()
        for s in self.streams:
            s.wait_stream(main)
        # Largest jobs first so the long chains start early.
        order = sorted(range(len(self.jobs)), key=lambda i: -self.jobs[i].vis)
        for k, i in enumerate(order):
            with torch.cuda.stream(self.streams[k % len(self.streams)]):
                self.jobs[i].run()
        for s in self.streams:
            main.wait_stream(s)

    def run_eager(self):
        self._issue()

    def capture(self):
        import torch
        self.run_eager()                      # warm-up outside the graph
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self._issue()
        self.graph = g
        return g

    def replay(self):
        self.graph.replay()

    def images(self):
        return [(j.re.cpu().numpy() + 1j * j.im.cpu().numpy()) for j in self.jobs]


class FusedBatch:
    """The jobs imaged by one launch per stage for each group of jobs that
    share grid size, kernel and support (rick_batch_*, include/rick.h): one
    binning, scan, tile and FFT pass over the group's stacked w-planes, and
    one fused column/phase pass writing every job's image.  Groups are dealt
    to `n_streams` streams, largest first, so one group's LDS-bound
    accumulation overlaps another's memory-bound transforms (96-job mix,
    FP32: 28.5 ms on one stream, 23.1 ms on four; eight are no better, the
    device exposing four hardware queues per process), and the whole batch
    is one HIP graph replay after capture()."""

    def __init__(self, datasets, precision, device="cuda:0", n_streams=4, tile=(0, 0),
                 rows_per_task=0, unroll=0):
        import torch

        from ._lib import BatchDesc, BatchJob, check, lib
        self.torch = torch
        self.datasets = datasets
        self.precision = precision
        dev = torch.device(device)
        self.re, self.im = [], []
        groups = {}
        for i, ds in enumerate(datasets):
            c = ds.cfg
            groups.setdefault((c.grid, c.kernel, c.support, c.n_chan, c.n_pol), []).append(i)
        self.groups = []
        for (grid, kernel, support, F, P), idx in sorted(groups.items()):
            jobs = (BatchJob * len(idx))()
            for k, i in enumerate(idx):
                ds = datasets[i]
                re = torch.empty(grid * grid, dtype=torch.float64, device=dev)
                im = torch.empty_like(re)
                self.re.append((i, re))
                self.im.append((i, im))
                j = jobs[k]
                j.num_points = ds.rows
                for name in ("uu", "vv", "ww", "vis_real", "vis_img", "weight"):
                    setattr(j, name, getattr(ds, name).data_ptr())
                j.num_w_planes = ds.cfg.planes
                j.dw = ds.dw
                j.resolution, j.wmin, j.wmax = ds.resolution, ds.wmin, ds.wmax
                j.image_real, j.image_imag = re.data_ptr(), im.data_ptr()
            d = BatchDesc()
            d.grid_size, d.kernel, d.w_support, d.precision = grid, kernel, support, precision
            d.phase_on = 1
            d.freq_per_chan, d.polarizations = F, P
            d.tile_u, d.tile_v = tile
            d.rows_per_task, d.footprint_unroll = rows_per_task, unroll
            d.device = dev.index or 0
            h = ctypes.c_void_p()
            check(lib().rick_batch_create(ctypes.byref(h), ctypes.byref(d), jobs, len(idx)))
            self.groups.append((h, jobs, idx))
        self.re = [t for _, t in sorted(self.re, key=lambda x: x[0])]
        self.im = [t for _, t in sorted(self.im, key=lambda x: x[0])]
        self.streams = [torch.cuda.Stream(device=dev) for _ in range(max(1, n_streams))]
        self.graph = None

    def __del__(self):
        from . import _lib
        for h, _, _ in getattr(self, "groups", []):
            if h and h.value and _lib._lib is not None:
                _lib._lib.rick_batch_destroy(h)
        self.groups = []

    @property
    def device_bytes(self) -> int:
        from ._lib import lib
        return sum(lib().rick_batch_device_bytes(h) for h, _, _ in self.groups)

    def _issue(self):
        from ._lib import check, lib
        torch = self.torch
        main = torch.cuda.current_stream()
        for s in self.streams:
            s.wait_stream(main)
        # Largest groups first.
        order = sorted(range(len(self.groups)),
                       key=lambda g: -sum(self.datasets[i].cfg.n_vis for i in self.groups[g][2]))
        for k, gi in enumerate(order):
            s = self.streams[k % len(self.streams)]
            check(lib().rick_batch_run(self.groups[gi][0], ctypes.c_void_p(s.cuda_stream)))
        for s in self.streams:
            main.wait_stream(s)

    def run_eager(self):
        self._issue()

    def capture(self):
        torch = self.torch
        self.run_eager()
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self._issue()
        self.graph = g
        return g

    def replay(self):
        self.graph.replay()

    def images(self):
        return [(r.cpu().numpy() + 1j * i.cpu().numpy()) for r, i in zip(self.re, self.im)]


def _timed(fn, steps):
    import torch
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps


def main():
    import torch
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=96)
    ap.add_argument("--streams", type=int, default=4)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--seed", type=int, default=3)
    args = ap.parse_args()
    cfgs = job_configs(args.jobs, args.seed)
    datasets = [synth.make_dataset(c, device="cuda:0") for c in cfgs]
    total_vis = sum(c.n_vis for c in cfgs)
    print(f"[batch] {len(cfgs)} jobs, {total_vis} visibilities", file=sys.stderr, flush=True)
    # Sequential FP64, one job after another on one stream (the reference's
    # arithmetic, one image at a time).
    seq = make_jobs(cfgs, FP64, datasets=datasets)
    for j in seq:
        j.run()
    t_seq = _timed(lambda: [j.run() for j in seq], args.steps)
    ref_imgs = [(j.re.cpu().numpy() + 1j * j.im.cpu().numpy()) for j in seq]
    del seq
    torch.cuda.empty_cache()
    out = {"jobs": len(cfgs), "visibilities": total_vis, "seq_fp64_s": t_seq}
    for name, prec in (("fp64", FP64), ("fp32", FP32)):
        b = Batch(make_jobs(cfgs, prec, datasets=datasets), args.streams)
        b.capture()
        t_graph = _timed(b.replay, args.steps)
        t_eager = _timed(b.run_eager, args.steps)
        imgs = b.images()
        errs = [float(np.linalg.norm(a - r) / np.linalg.norm(r)) for a, r in zip(imgs, ref_imgs)]
        out[name] = {"graph_s": t_graph, "eager_s": t_eager,
                     "images_per_s": len(cfgs) / t_graph, "vis_per_s": total_vis / t_graph,
                     "max_rel_l2_vs_fp64": max(errs)}
        del b
        torch.cuda.empty_cache()
    for name, prec in (("fused_fp64", FP64), ("fused_fp32", FP32)):
        fb = FusedBatch(datasets, prec)
        fb.capture()
        t_graph = _timed(fb.replay, args.steps)
        imgs = fb.images()
        errs = [float(np.linalg.norm(a - r) / np.linalg.norm(r)) for a, r in zip(imgs, ref_imgs)]
        out[name] = {"graph_s": t_graph, "groups": len(fb.groups),
                     "images_per_s": len(cfgs) / t_graph, "vis_per_s": total_vis / t_graph,
                     "max_rel_l2_vs_fp64": max(errs)}
        del fb
        torch.cuda.empty_cache()
    out["fp32_graph_speedup_vs_seq_fp64"] = t_seq / out["fp32"]["graph_s"]
    out["fused_fp32_speedup_vs_seq_fp64"] = t_seq / out["fused_fp32"]["graph_s"]
    print(json.dumps(out))


if __name__ == "__main__":
    main()
