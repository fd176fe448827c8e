"""Batched (HIP-graph) imaging and the continuous-batching service on the GPU.

Graph replays are compared with eager runs of the same jobs, replay after
replay with eager work and synchronisation in between; the service's results
for a job are compared alone and in a mix of concurrently running jobs.
Tolerance: FP64 relative L2 <= 1e-12 (atomic order only), FP32 <= 1e-6.
"""
import numpy as np
import pytest
import torch

from conftest import rel_l2
from rick_amd import FP32, FP64, batch, serve, synth

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
            assert rel_l2(r.result, alone[r.key]) <= 1e-6, r.key
    # NaN rows dropped: the bad job equals the clean half.
    assert rel_l2(alone[len(datasets) - 2], alone[len(datasets) - 1]) <= 1e-6
