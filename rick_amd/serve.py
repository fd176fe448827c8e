"""Imaging service: continuous batching of imaging requests on one GPU.

Requests (one imaging job each: visibilities already resident on the GPU)
arrive over time.  The scheduler keeps up to `slots` jobs on the device at
once, each on its own stream, and admits queued requests in arrival order as
soon as a slot and enough HBM are free: a job's footprint (its grid, images and
plan workspace, bytes computed from planes x grid^2 x element size and the
gridder's capacities) is charged against `hbm_budget`.  Plans and buffers are
pooled per job signature, so after warm-up no request allocates; idle pooled
buffers stay charged to the budget too, and when a job would not fit, idle
entries of other signatures are freed, least recently used first.

Faults handled without disturbing other jobs:
  * a request whose footprint exceeds the whole budget is rejected at arrival;
  * a queued request not started within its timeout is cancelled;
  * a job whose launch fails (bad arguments) completes with an error;
  * rows with NaN or out-of-range uvw are dropped by the gridder itself.

    python -m rick_amd.serve --duration 20 --rate 400    # prints one JSON line
"""
from __future__ import annotations

import argparse
import collections
import json
import sys
import time
from dataclasses import dataclass, field

import numpy as np

from . import synth
from ._lib import FP32, FP64


@dataclass
class Request:
    rid: int
    key: int                 # which job (dataset) this request images
    arrival: float           # seconds since the start of the run
    timeout: float = float("inf")
    # filled in by the server
    start: float = None
    finish: float = None
    status: str = "queued"   # queued, running, done, rejected, cancelled, failed
    error: str = ""
    result: object = None


def grid_footprint(cfg, precision) -> int:
    """Device bytes a job holds while it runs: grid, two FP64 images, and the
    plan (gridding workspace sized for its rows, tables, FFT scratch)."""
    cb = 16 if precision == FP64 else 8
    grid = cfg.planes * cfg.grid * cfg.grid * cb
    images = 2 * 8 * cfg.grid * cfg.grid
    rec = (32 if precision == FP64 else 24) * 4 * cfg.n_rows
    tiles = cfg.planes * max(1, cfg.grid // 32) ** 2
    work = rec + 48 * tiles + (cfg.grid * cfg.grid * cb if cfg.grid > 512 else 0)
    return grid + images + work


class PlanPool:
    """Idle per-signature resources with their budgeted bytes, least recently
    used first.  The executor returns a finished job's resources here and
    takes them back for the next job of the same signature."""

    def __init__(self):
        self.entries = collections.OrderedDict()   # token -> (sig, nbytes, resources)
        self._next = 0

    @property
    def nbytes(self) -> int:
        return sum(e[1] for e in self.entries.values())

    def put(self, sig, nbytes, res):
        self.entries[self._next] = (sig, nbytes, res)
        self._next += 1

    def reusable(self, sig) -> int:
        """Bytes of the idle entry a job of `sig` would take (0: none)."""
        for s, n, _ in reversed(self.entries.values()):
            if s == sig:
                return n
        return 0

    def take(self, sig):
        for tok in reversed(self.entries):
            if self.entries[tok][0] == sig:
                return self.entries.pop(tok)[2]
        return None

    def evict(self, nbytes, keep_sig=None) -> int:
        """Drops idle entries, oldest first, until at least `nbytes` are freed,
        sparing the entry take(keep_sig) would return; returns the bytes freed."""
        keep = next((t for t in reversed(self.entries) if self.entries[t][0] == keep_sig), None)
        freed = 0
        for tok in list(self.entries):
            if freed >= nbytes:
                break
            if tok == keep:
                continue
            freed += self.entries.pop(tok)[1]
        return freed


class GpuExecutor:
    """Runs imaging jobs on a GPU: one stream per slot, pooled plans."""

    def __init__(self, datasets, precision=FP32, slots=8, device="cuda:0"):
        import torch
        self.torch = torch
        self.ds = datasets
        self.precision = precision
        self.device = device
        self.streams = [torch.cuda.Stream(device=device) for _ in range(slots)]
        self.free_streams = list(range(slots))
        self.pool = PlanPool()                       # idle (plan, grid, re, im)

    def footprint(self, key) -> int:
        return grid_footprint(self.ds[key].cfg, self.precision)

    def pooled_bytes(self) -> int:
        return self.pool.nbytes

    def reusable_bytes(self, key) -> int:
        return self.pool.reusable(self._sig(key))

    def evict(self, nbytes, keep_key=None) -> int:
        freed = self.pool.evict(nbytes, None if keep_key is None else self._sig(keep_key))
        if freed:
            self.torch.cuda.empty_cache()     # grids and images go back to the device
        return freed

    def _sig(self, key):
        ds = self.ds[key]
        c = ds.cfg
        return (c.grid, c.planes, c.kernel, c.support, ds.rows, ds.resolution, ds.wmin, ds.wmax)

    def _acquire(self, key):
        from .imaging import plan_for
        res = self.pool.take(self._sig(key))
        if res is not None:
            return res
        torch = self.torch
        ds = self.ds[key]
        plan = plan_for(ds, precision=self.precision)
        npix = ds.cfg.grid * ds.cfg.grid
        return (plan, plan.new_grid(),
                torch.empty(npix, dtype=torch.float64, device=self.device),
                torch.empty(npix, dtype=torch.float64, device=self.device))

    def start(self, req):
        """Issue the job; returns a handle.  Raises on launch errors, after
        returning the stream slot (and any resources taken) for later jobs."""
        torch = self.torch
        sidx = self.free_streams.pop()
        res = None
        try:
            res = self._acquire(req.key)
            plan, grid, re, im = res
            ds = self.ds[req.key]
            stream = self.streams[sidx]
            with torch.cuda.stream(stream):
                plan.image(ds.uu, ds.vv, ds.ww, ds.vis_real, ds.vis_img, ds.weight,
                           ds.cfg.n_chan, ds.cfg.n_pol, grid=grid, image_real=re, image_imag=im)
                ev = torch.cuda.Event()
                ev.record(stream)
        except Exception:
            self.free_streams.append(sidx)
            if res is not None:
                self.pool.put(self._sig(req.key), self.footprint(req.key), res)
            raise
        return (sidx, res, ev, req.key)

    def done(self, handle) -> bool:
        return handle[2].query()

    def finish(self, handle, keep_result=False):
        sidx, res, ev, key = handle
        ev.synchronize()
        out = None
        if keep_result:
            out = res[2].cpu().numpy() + 1j * res[3].cpu().numpy()
        self.free_streams.append(sidx)
        self.pool.put(self._sig(key), self.footprint(key), res)
        return out


class Server:
    """Continuous-batching scheduler over an executor (see GpuExecutor)."""

    def __init__(self, executor, hbm_budget: int, slots: int, clock=time.perf_counter):
        self.ex = executor
        self.budget = hbm_budget
        self.slots = slots
        self.clock = clock
        self.used = 0
        self.queue = collections.deque()
        self.running = []       # (req, handle, bytes)
        self.finished = []

    def submit(self, req, now):
        need = self.ex.footprint(req.key)
        if need > self.budget:
            req.status, req.error, req.finish = "rejected", f"needs {need} B > budget", now
            self.finished.append(req)
            return False
        self.queue.append(req)
        return True

    def _admit(self, now):
        while self.queue and len(self.running) < self.slots:
            req = self.queue[0]
            if now - req.arrival > req.timeout:
                self.queue.popleft()
                req.status, req.finish = "cancelled", now
                self.finished.append(req)
                continue
            need = self.ex.footprint(req.key)
            if self.used + need > self.budget:
                break                      # wait for memory (arrival order kept)
            # Idle pooled buffers hold device memory too: the job's own
            # signature's entry is reused, others are freed (LRU) if needed.
            over = self.used + need + self.ex.pooled_bytes() - self.ex.reusable_bytes(req.key) \
                - self.budget
            if over > 0:
                self.ex.evict(over, keep_key=req.key)
            self.queue.popleft()
            try:
                handle = self.ex.start(req)
            except Exception as e:          # launch failure: this job only
                req.status, req.error, req.finish = "failed", str(e), now
                self.finished.append(req)
                continue
            req.status, req.start = "running", now
            self.used += need
            self.running.append((req, handle, need))

    def _reap(self, now, keep_results):
        still = []
        for req, handle, need in self.running:
            if self.ex.done(handle):
                req.result = self.ex.finish(handle, keep_results)
                req.status, req.finish = "done", self.clock() - self.t0
                self.used -= need
                self.finished.append(req)
            else:
                still.append((req, handle, need))
        self.running = still

    def run(self, requests, keep_results=False, poll=0.0002, progress=0.0, log=None):
        """Replay `requests` (sorted by arrival) in real time; returns them.
        With progress > 0, `log` (default: stderr) gets a status line every
        `progress` seconds: elapsed time, completed, queued and running jobs."""
        self.t0 = self.clock()
        pending = collections.deque(sorted(requests, key=lambda r: r.arrival))
        next_report = progress
        while pending or self.queue or self.running:
            now = self.clock() - self.t0
            if progress > 0 and now >= next_report:
                msg = (f"[serve] t={now:.0f} s: {len(self.finished)} finished, "
                       f"{len(self.queue)} queued, {len(self.running)} running, "
                       f"{len(pending)} to arrive")
                (log or (lambda m: print(m, file=sys.stderr, flush=True)))(msg)
                next_report += progress
            while pending and pending[0].arrival <= now:
                self.submit(pending.popleft(), now)
            self._reap(now, keep_results)
            self._admit(self.clock() - self.t0)
            if not self.running and not self.queue and pending:
                wait = pending[0].arrival - (self.clock() - self.t0)
                if wait > 0:
                    time.sleep(min(wait, 0.01))
            elif self.running:
                time.sleep(poll)
        return requests


def latency_stats(reqs):
    done = [r for r in reqs if r.status == "done"]
    lat = np.array([r.finish - r.arrival for r in done]) if done else np.zeros(1)
    span = max((r.finish for r in done), default=0.0) - min((r.arrival for r in reqs), default=0.0)
    return {"completed": len(done), "rejected": sum(r.status == "rejected" for r in reqs),
            "cancelled": sum(r.status == "cancelled" for r in reqs),
            "failed": sum(r.status == "failed" for r in reqs),
            "p50_ms": float(np.percentile(lat, 50) * 1e3),
            "p95_ms": float(np.percentile(lat, 95) * 1e3),
            "p99_ms": float(np.percentile(lat, 99) * 1e3),
            "images_per_s": len(done) / span if span > 0 else 0.0}


def poisson_requests(n_keys, weights, rate, duration, seed=7):
    rng = np.random.default_rng(seed)
    t, out = 0.0, []
    while True:
        t += rng.exponential(1.0 / rate)
        if t > duration:
            return out
        out.append(Request(len(out), int(rng.choice(n_keys, p=weights)), t))


def workload(n_small=24, n_lofar=2, seed=5, lofar_vis=20_000_000):
    """The job mix: batched-imaging jobs plus 10% LOFAR-like 4096^2 x 32-plane
    jobs of ~20M visibilities (16 channels)."""
    from .batch import job_configs
    small = job_configs(n_small, seed)
    base = synth.CONFIGS["lofar"]
    lofar = [base.scaled(name=f"lofar{i}", n_times=max(1, lofar_vis // (base.n_baselines *
                                                                      base.n_chan)),
                         seed=base.seed + 100 + i) for i in range(n_lofar)]
    cfgs = small + lofar
    w = np.array([0.9 / n_small] * n_small + [0.1 / n_lofar] * n_lofar)
    return cfgs, w / w.sum()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--duration", type=float, default=20.0, help="seconds of arrivals")
    ap.add_argument("--rate", type=float, default=400.0, help="requests per second")
    ap.add_argument("--slots", type=int, default=8)
    ap.add_argument("--budget-gb", type=float, default=256.0)
    ap.add_argument("--paths", choices=["both", "fp32", "fp64"], default="both",
                    help="FP32 continuous batching, FP64 sequential, or both")
    ap.add_argument("--progress", type=float, default=30.0,
                    help="seconds between status lines on stderr (0: none)")
    args = ap.parse_args()
    cfgs, weights = workload()
    datasets = [synth.make_dataset(c, device="cuda:0") for c in cfgs]
    reqs = poisson_requests(len(cfgs), weights, args.rate, args.duration)
    print(f"[serve] {len(reqs)} requests over {args.duration} s", file=sys.stderr, flush=True)
    budget = int(args.budget_gb * 1e9)
    out = {"requests": len(reqs), "rate_per_s": args.rate, "duration_s": args.duration,
           "slots": args.slots, "budget_gb": args.budget_gb}
    # Warm the plan pools (plan creation allocates) outside the measured runs.
    precs = {"both": (FP32, FP64), "fp32": (FP32,), "fp64": (FP64,)}[args.paths]
    for prec in precs:
        ex = GpuExecutor(datasets, prec, slots=1 if prec == FP64 else args.slots)
        Server(ex, budget, 1).run([Request(i, i, 0.0) for i in range(len(cfgs))])
        if prec == FP32:
            srv = Server(ex, budget, args.slots)
            out["fp32_continuous"] = latency_stats(srv.run(
                [Request(r.rid, r.key, r.arrival) for r in reqs], progress=args.progress))
        else:
            seq = Server(ex, budget, 1)
            out["fp64_sequential"] = latency_stats(seq.run(
                [Request(r.rid, r.key, r.arrival) for r in reqs], progress=args.progress))
        del ex
    print(json.dumps(out))


if __name__ == "__main__":
    main()
