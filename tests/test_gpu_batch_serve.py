"""Batched (HIP-graph) imaging and the continuous-batching service on the GPU.

Graph replays are compared with eager runs of the same jobs, replay after
replay with eager work and synchronisation in between; the service's results
for a job are compared alone and in a mix of concurrently running jobs.
Tolerance: FP64 relative L2 <= 1e-12 (atomic order only), FP32 <= 1e-6.

At the full workloads: the 96-job batch (grids 256-2048, 4-16 w-planes,
5e4-8e6 visibilities, Gaussian support 7 / Kaiser-Bessel support 9), one job
per grid size and kernel against the CPU oracle (FP64 <= 1e-12, FP32 <=
1e-5) and every job's FP32 image against its FP64 image (<= 1e-5); and the
serving mix with LOFAR-like 20M-visibility jobs, where each FP32 result is
bitwise the same alone and in the mix (FP32 gridding accumulates in
order-independent fixed point) and a subset matches the oracle.
"""
import os

import numpy as np
import pytest
import torch

from conftest import rel_l2
from oracle import oracle as O
from rick_amd import FP32, FP64, batch, plan_for, serve, synth

pytestmark = pytest.mark.gpu


def small_cfgs(n, seed=11):
    cfgs = batch.job_configs(n, seed, vis_lo=2e4, vis_hi=2e5)
    # keep the test quick: grids up to 1024
    return [c.scaled(grid=min(c.grid, 1024)) for c in cfgs]


@pytest.mark.parametrize("precision", [FP64, FP32])
def test_batch_graph_replays_match_eager(precision):
    cfgs = small_cfgs(6)
    jobs = batch.make_jobs(cfgs, precision)
    b = batch.Batch(jobs, n_streams=3)
    b.run_eager()
    torch.cuda.synchronize()
    eager = b.images()
    # each job alone, one at a time
    for j in jobs:
        j.run()
        torch.cuda.synchronize()
    alone = b.images()
    tol = 1e-12 if precision == FP64 else 1e-6
    for a, e in zip(alone, eager):
        assert rel_l2(a, e) <= tol
    b.capture()
    for rep in range(3):
        for j in jobs:                       # clobber the outputs between replays
            j.re.fill_(float("nan"))
            j.im.fill_(float("nan"))
        b.replay()
        torch.cuda.synchronize()
        got = b.images()
        print(f"replay {rep}: max rel-L2 {max(rel_l2(g, e) for g, e in zip(got, eager)):.2e}")
        for g, e in zip(got, eager):
            assert np.isfinite(g).all()
            assert rel_l2(g, e) <= tol
        b.run_eager()                        # eager work between replays
        torch.cuda.synchronize()


def test_serve_results_same_alone_and_in_mix():
    cfgs = small_cfgs(8, seed=21)
    datasets = [synth.make_dataset(c, device="cuda:0") for c in cfgs]
    # A job whose first half of rows has NaN u: those rows are dropped; the
    # rest must image as the same job without them.
    bad = datasets[0].subset(slice(0, datasets[0].rows))
    half = bad.rows // 2
    uu = bad.uu.clone()
    uu[:half] = float("nan")
    bad.uu = uu
    datasets.append(bad)
    clean = datasets[0].subset(slice(half, datasets[0].rows))
    datasets.append(clean)
    ex = serve.GpuExecutor(datasets, FP32, slots=4)
    probes = [1, 3, len(datasets) - 2, len(datasets) - 1]
    alone = {}
    for k in probes:
        r = serve.Request(k, k, 0.0)
        serve.Server(ex, int(1e12), 1).run([r], keep_results=True)
        assert r.status == "done"
        alone[k] = r.result
    mix = [serve.Request(i, k, 0.0) for i, k in enumerate(list(range(len(datasets))) * 2)]
    serve.Server(ex, int(1e12), 4).run(mix, keep_results=True)
    assert all(r.status == "done" for r in mix)
    for r in mix:
        if r.key in alone:
            assert np.array_equal(r.result, alone[r.key], equal_nan=True), r.key
    # NaN rows dropped: the bad job equals the clean half.
    assert rel_l2(alone[len(datasets) - 2], alone[len(datasets) - 1]) <= 1e-6


def host_threads():
    return max(1, min(16, len(os.sched_getaffinity(0))))


def oracle_image(ds):
    """wstack -> FFT -> phase_correction of the CPU oracle on a job's rows."""
    h, cfg, th = ds.to_numpy(), ds.cfg, host_threads()
    g = O.wstack(cfg.kernel, cfg.planes, h.uu, h.vv, h.ww, h.vis_real, h.vis_img, h.weight,
                 cfg.n_chan, cfg.n_pol, ds.dx, ds.dw, cfg.support, cfg.grid, cfg.grid, threads=th)
    gs = O.fft_planes(cfg.planes, cfg.grid, cfg.grid, g, threads=th)
    del g
    re, im = O.phase_correction(gs, cfg.grid, cfg.grid, cfg.planes, ds.resolution, ds.wmin,
                                ds.wmax, threads=th)
    return re + 1j * im


def test_batch_96_jobs_against_oracle():
    cfgs = batch.job_configs(96)
    print(f"case batch 96 jobs: {sum(c.n_vis for c in cfgs)} vis", flush=True)
    datasets = [synth.make_dataset(c, device="cuda:0") for c in cfgs]
    picks = {}
    for i, c in enumerate(cfgs):
        picks.setdefault((c.grid, c.kernel), i)
    imgs = {}
    for prec in (FP64, FP32):
        b = batch.Batch(batch.make_jobs(cfgs, prec, datasets=datasets))
        b.capture()
        for j in b.jobs:
            j.re.fill_(float("nan"))
            j.im.fill_(float("nan"))
        b.replay()
        torch.cuda.synchronize()
        imgs[prec] = b.images()
        del b
        torch.cuda.empty_cache()
    # The same jobs through the fused batch (one launch per stage per group of
    # jobs sharing grid size and kernel), captured and replayed.
    fused = {}
    for prec in (FP64, FP32):
        fb = batch.FusedBatch(datasets, prec)
        fb.capture()
        for rep in range(2):
            for t in fb.re + fb.im:
                t.fill_(float("nan"))
            fb.replay()
            torch.cuda.synchronize()
        fused[prec] = fb.images()
        del fb
        torch.cuda.empty_cache()
    worst64 = worst32 = 0.0
    for (grid, kernel), i in sorted(picks.items()):
        ref = oracle_image(datasets[i])
        e64, e32 = rel_l2(imgs[FP64][i], ref), rel_l2(imgs[FP32][i], ref)
        f64, f32 = rel_l2(fused[FP64][i], ref), rel_l2(fused[FP32][i], ref)
        print(f"job {i} ({grid}^2 x {cfgs[i].planes}, kernel {kernel}, {cfgs[i].n_vis} vis): "
              f"FP64 {e64:.2e}, FP32 {e32:.2e}, fused FP64 {f64:.2e}, fused FP32 {f32:.2e} "
              f"vs oracle", flush=True)
        worst64 = max(worst64, e64, f64)
        worst32 = max(worst32, e32, f32)
    e = max(rel_l2(a, b) for a, b in zip(imgs[FP32], imgs[FP64]))
    ef64 = max(rel_l2(a, b) for a, b in zip(fused[FP64], imgs[FP64]))
    ef32 = max(rel_l2(a, b) for a, b in zip(fused[FP32], imgs[FP64]))
    print(f"oracle picks: FP64 max {worst64:.2e}, FP32 max {worst32:.2e}; all 96 FP32 vs FP64 "
          f"max {e:.2e}; fused FP64 vs per-job FP64 max {ef64:.2e}, fused FP32 vs FP64 max "
          f"{ef32:.2e}", flush=True)
    assert len(picks) == 8
    assert worst64 <= 1e-12 and worst32 <= 1e-5 and e <= 1e-5
    assert ef64 <= 1e-12 and ef32 <= 1e-5


@pytest.mark.parametrize("precision", [FP64, FP32])
def test_fused_batch_edge_jobs(precision):
    # Jobs of one group with no rows, rows not a multiple of the binning
    # block, one plane, NaN rows, and two groups; each image against the job
    # imaged alone by its own plan.
    cfgs = small_cfgs(6, seed=5)
    cfgs = [c.scaled(grid=256) for c in cfgs[:4]] + [c.scaled(grid=512) for c in cfgs[4:]]
    cfgs[1] = cfgs[1].scaled(planes=1)
    datasets = [synth.make_dataset(c, device="cuda:0") for c in cfgs]
    datasets[2] = datasets[2].subset(slice(0, 0))               # no rows
    datasets[3] = datasets[3].subset(slice(0, 16384 + 3))       # past one block
    bad = datasets[0].subset(slice(0, datasets[0].rows))
    uu = bad.uu.clone()
    uu[::7] = float("nan")                                       # dropped rows
    bad.uu = uu
    datasets[0] = bad
    print(f"case fused batch edge jobs prec={precision}: rows {[d.rows for d in datasets]}",
          flush=True)
    fb = batch.FusedBatch(datasets, precision)
    assert len(fb.groups) == len({(c.grid, c.kernel) for c in cfgs})
    fb.run_eager()
    torch.cuda.synchronize()
    got = fb.images()
    for k, ds in enumerate(datasets):
        if ds.rows == 0:
            assert np.count_nonzero(got[k]) == 0
            continue
        plan = plan_for(ds, precision=precision)
        re, im, _ = plan.image(ds.uu, ds.vv, ds.ww, ds.vis_real, ds.vis_img, ds.weight,
                               ds.cfg.n_chan, ds.cfg.n_pol)
        ref = re.cpu().numpy() + 1j * im.cpu().numpy()
        err = rel_l2(got[k], ref)
        print(f"job {k}: fused vs alone {err:.2e}", flush=True)
        assert np.isfinite(got[k]).all()
        assert err <= (1e-12 if precision == FP64 else 1e-6)


def test_serve_lofar_mix_bitwise_and_oracle():
    cfgs, weights = serve.workload(n_small=8, n_lofar=2)
    lofar = [k for k, c in enumerate(cfgs) if c.grid == 4096]
    print(f"case serve mix: {len(cfgs)} jobs, LOFAR-like {[cfgs[k].n_vis for k in lofar]} vis",
          flush=True)
    datasets = [synth.make_dataset(c, device="cuda:0") for c in cfgs]
    budget = int(256e9)
    ex = serve.GpuExecutor(datasets, FP32, slots=4)
    alone = {}
    for k in range(len(cfgs)):
        r = serve.Request(k, k, 0.0)
        serve.Server(ex, budget, 1).run([r], keep_results=True)
        assert r.status == "done"
        alone[k] = r.result
    rng = np.random.default_rng(3)
    keys = list(range(len(cfgs))) * 2
    mix = [serve.Request(i, k, float(t)) for i, (k, t) in
           enumerate(zip(keys, np.sort(rng.uniform(0.0, 0.02, len(keys)))))]
    serve.Server(ex, budget, 4).run(mix, keep_results=True)
    assert all(r.status == "done" for r in mix)
    for r in mix:
        assert np.array_equal(r.result, alone[r.key], equal_nan=True), r.key
    print("FP32 results bitwise equal alone and in the mix", flush=True)
    del ex
    torch.cuda.empty_cache()
    ex64 = serve.GpuExecutor(datasets, FP64, slots=1)
    for k in (lofar[0], 0, 1):
        ref = oracle_image(datasets[k])
        r = serve.Request(k, k, 0.0)
        serve.Server(ex64, budget, 1).run([r], keep_results=True)
        e64, e32 = rel_l2(r.result, ref), rel_l2(alone[k], ref)
        print(f"job {k} ({cfgs[k].grid}^2, {cfgs[k].n_vis} vis): served FP64 {e64:.2e}, "
              f"FP32 {e32:.2e} vs oracle", flush=True)
        assert e64 <= 1e-12 and e32 <= 1e-5
