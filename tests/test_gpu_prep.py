"""Device-side data preparation and I/O against the CPU oracle, bit for bit:
NORMALIZE_UVW (gridding.c:26-82), sector binning and gather (gridding.c:
139-197, gridding_cpu.c:90-138), binMS reads into device memory, cube
writes from device grids, and one dataset through the whole flow from the
binMS files to the image files (FP64, rel-L2 <= 1e-12 against the oracle).
"""
import os

import numpy as np
import pytest
import torch

from conftest import rel_l2
from oracle import oracle as O
from rick_amd import FP32, FP64, KAISERBESSEL, binms, pipeline, prep, synth

pytestmark = pytest.mark.gpu


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def host(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


def bits(a):
    return np.ascontiguousarray(a).view(np.uint64)


# --- NORMALIZE_UVW ------------------------------------------------------------

@pytest.mark.parametrize("n", [1, 63, 4097, 1_000_003])
def test_normalize_bit_exact(n):
    print(f"case normalize n={n}", flush=True)
    rng = np.random.default_rng(n)
    u, v, w = rng.normal(0, 2e4, n), rng.normal(0, 3e4, n), rng.uniform(-400, 900, n)
    if n == 1:
        w = np.array([0.5])            # wmax == wmin: 0/0 as in the reference
    want = O.normalize_uvw(u, v, w)
    du, dv, dw = dev(u), dev(v), dev(w)
    ext = prep.normalize_uvw(du, dv, dw)
    for g, e in zip((du, dv, dw), want):
        g = host(g)
        if n == 1 and g is not None and np.isnan(e).all():
            assert np.isnan(g).all()
            continue
        np.testing.assert_array_equal(bits(g), bits(e))
    ex = host(ext)
    np.testing.assert_array_equal(ex[:6], [min(1e20, u.min()), min(1e20, v.min()),
                                           min(1e20, w.min()), max(0.0, u.max()),
                                           max(0.0, v.max()), max(0.0, w.max())])


def test_normalize_zero_extrema_signs():
    # Extremum ties between +0.0 and -0.0: the reference's fold keeps the last
    # one, and the max starts from +0.0 (gridding.c:37-49).
    w = np.array([-0.0, 3.0, 0.0, 5.0, -0.0] * 1000 + [2.0])       # min is a zero, last -0.0
    u = -np.abs(np.random.default_rng(0).normal(size=w.size)) - 1   # all negative: max stays 0
    v = np.linspace(-7, 7, w.size)
    want = O.normalize_uvw(u, v, w)
    du, dv, dw = dev(u), dev(v), dev(w)
    prep.normalize_uvw(du, dv, dw)
    for g, e in zip((du, dv, dw), want):
        np.testing.assert_array_equal(bits(host(g)), bits(e))


def test_normalize_empty_input():
    z = torch.empty(0, dtype=torch.float64, device="cuda")
    ext = host(prep.normalize_uvw(z, z.clone(), z.clone()))
    assert ext[0] == 1e20 and ext[3] == 0.0 and ext[6] == 1e20 + 0.001 * 1e20


# --- sector binning / gather --------------------------------------------------

def check_bins(vv, S, yaxis, dx, sup):
    histo, start, rows, bad = O.sector_bin(vv, S, yaxis, dx, sup)
    b = prep.sector_bin(dev(vv), S, yaxis, dx, sup)
    np.testing.assert_array_equal(b.histo.numpy().view(np.uint32), histo)
    np.testing.assert_array_equal(b.start.numpy().view(np.uint32), start)
    np.testing.assert_array_equal(host(b.rows).view(np.uint32), rows)
    assert b.dropped == bad and b.total == rows.size
    return b


@pytest.mark.parametrize("S", [1, 2, 4, 7, 64, 1000])
@pytest.mark.parametrize("sup", [7, 9])
def test_sector_bin_random(S, sup):
    print(f"case sector_bin S={S} support={sup}", flush=True)
    n = 300_001
    vv = np.random.default_rng(S * 10 + sup).uniform(0.0005, 0.9995, n)
    check_bins(vv, S, max(1, 4096 // S), 1 / 4096, sup)


def test_sector_bin_borders_and_invalid_rows():
    S, yaxis, dx, sup = 4, 16, 1 / 64, 7
    h = (sup - 1) // 2 * dx
    base = np.array([k / S for k in range(S + 1)])
    edges = np.concatenate([base, base + h, base - h, np.nextafter(base + h, 0),
                            np.nextafter(base - h, 1), np.nextafter(base, 2),
                            np.nextafter(base, -1)])
    vv = np.concatenate([edges, [-2.0, 1.5, np.nan, -0.0, 1.0, 0.9999999999]])
    vv = np.tile(vv, 700)          # several blocks, many rows per sector per wave
    b = check_bins(vv, S, yaxis, dx, sup)
    assert b.dropped > 0


def test_sector_bin_empty_and_tiny():
    check_bins(np.zeros(0), 3, 10, 0.1 / 3, 7)
    check_bins(np.array([0.5]), 1, 512, 1 / 512, 7)


@pytest.mark.parametrize("isector,F,P", [(0, 4, 2), (2, 4, 2), (1, 3, 1)])
def test_sector_gather_matches_oracle(isector, F, P):
    n, S = 70_001, 4
    rng = np.random.default_rng(isector)
    a = {"uu": rng.uniform(0, 1, n), "vv": rng.uniform(0, 1, n), "ww": rng.uniform(0, 1, n),
         "weight": rng.uniform(0.5, 2, n * P).astype(np.float32),
         "vis_real": rng.normal(size=n * F * P).astype(np.float32),
         "vis_img": rng.normal(size=n * F * P).astype(np.float32)}
    yaxis, dx = 256, 1 / 1024
    d = {k: dev(v) for k, v in a.items()}
    b = prep.sector_bin(d["vv"], S, yaxis, dx, 7)
    lst = b.sector(isector)
    got = prep.sector_gather(lst, isector, yaxis, dx, F, P, d["uu"], d["vv"], d["ww"],
                             d["weight"], d["vis_real"], d["vis_img"])
    want = O.sector_gather(host(lst).view(np.uint32), isector, yaxis, dx, F, P, a["uu"], a["vv"],
                           a["ww"], a["weight"], a["vis_real"], a["vis_img"])
    for k in want:
        np.testing.assert_array_equal(host(got[k]), want[k])


@pytest.mark.parametrize("precision", [FP64, FP32])
def test_sector_slabs_grid_as_ranks_do(precision):
    # Four v-slabs of a 256^2 grid, as four ranks would grid them: each
    # slab's rows (halo included), walked backwards and shifted into the slab
    # (gridding_cpu.c:118-124), gridded into a 256 x 64 slab whose halo rows
    # reach below v = 0.  FP64 within 1e-12 of the oracle, FP32 within 1e-5.
    from rick_amd import Plan
    cfg = synth.CONFIGS["small"].scaled(n_ant=24, n_times=64, grid=256, planes=4,
                                        kernel=KAISERBESSEL, n_chan=2, n_pol=1)
    ds = synth.make_dataset(cfg)
    S, nx, F, P = 4, 256, cfg.n_chan, cfg.n_pol
    yaxis, dx, dw = nx // S, 1.0 / nx, 1.0 / cfg.planes
    keys = ("uu", "vv", "ww", "weight", "vis_real", "vis_img")
    h = {k: np.ascontiguousarray(getattr(ds, k)) for k in keys}
    d = {k: dev(v) for k, v in h.items()}
    b = prep.sector_bin(d["vv"], S, yaxis, dx, cfg.support)
    histo, start, rows, _ = O.sector_bin(h["vv"], S, yaxis, dx, cfg.support)
    assert histo[:S].sum() > h["uu"].size                  # halo rows present
    for s in range(S):
        print(f"case slab {s} precision={precision}", flush=True)
        g = prep.sector_gather(b.sector(s), s, yaxis, dx, F, P, *(d[k] for k in keys))
        o = O.sector_gather(rows[start[s]:start[s] + histo[s]], s, yaxis, dx, F, P,
                            *(h[k] for k in keys))
        assert host(g["vv"]).min() < 0.0 or s == 0
        plan = Plan(nx, yaxis, cfg.planes, kernel=KAISERBESSEL, support=cfg.support, dx=dx,
                    dw=dw, precision=precision, max_points=int(histo[s]))
        grid = plan.grid(g["uu"], g["vv"], g["ww"], g["vis_real"], g["vis_img"], g["weight"],
                         F, P)
        want = O.wstack(KAISERBESSEL, cfg.planes, o["uu"], o["vv"], o["ww"], o["vis_real"],
                        o["vis_img"], o["weight"], F, P, dx, dw, cfg.support, nx, yaxis)
        err = rel_l2(host(grid), want)
        assert err <= (1e-12 if precision == FP64 else 1e-5), err


# --- binMS reads into device memory, cube writes from device grids ------------

def test_binms_device_read_matches_host_read(tmp_path):
    n, F, P = 150_000, 16, 2            # 38 MB per visibility file: several staging chunks
    rng = np.random.default_rng(7)
    a = [rng.normal(size=n), rng.normal(size=n), rng.normal(size=n),
         rng.uniform(size=n * P).astype(np.float32),
         rng.normal(size=n * F * P).astype(np.float32),
         rng.normal(size=n * F * P).astype(np.float32)]
    m = binms.write(tmp_path, *a, F, P)
    start, rows = binms.partition(m, 1, 3)
    h = binms.read(tmp_path, m, start, rows)
    d = binms.read(tmp_path, m, start, rows, device=0)
    for k in h:
        np.testing.assert_array_equal(host(d[k]), h[k])


@pytest.mark.parametrize("precision", [FP64, FP32])
def test_write_cube_from_device(tmp_path, precision):
    planes, nx, ny = 3, 1024, 1536       # 4.7M cells: the device path works in chunks
    dt = np.float64 if precision == FP64 else np.float32
    g = np.random.default_rng(3).normal(size=2 * planes * nx * ny).astype(dt)
    binms.write_cube(tmp_path / "r.bin", tmp_path / "i.bin", dev(g), planes, nx, ny)
    np.testing.assert_array_equal(np.fromfile(tmp_path / "r.bin"), g[0::2].astype(np.float64))
    np.testing.assert_array_equal(np.fromfile(tmp_path / "i.bin"), g[1::2].astype(np.float64))


# --- whole flow ----------------------------------------------------------------

def write_synthetic_binms(path, cfg):
    """cfg's rows with physical uvw plus 1000 trailing rows (the reader drops
    the last 1000, init.c:337), meta extrema as the converter computes them."""
    ds = synth.make_dataset(cfg)
    u, v, w = synth.physical_uvw(cfg)
    F, P = cfg.n_chan, cfg.n_pol
    pad = 1000
    cat = lambda x, k: np.concatenate([x, x[:pad * k]])
    return binms.write(path, cat(u, 1), cat(v, 1), cat(w, 1), cat(ds.weight, P),
                       cat(ds.vis_real, F * P), cat(ds.vis_img, F * P), F, P,
                       ntimes=cfg.n_times, dt=10.0, baselines=cfg.n_baselines,
                       uvmin=ds.uvmin, uvmax=ds.uvmax, wmin=ds.wmin, wmax=ds.wmax)


def oracle_flow(path, nx, planes, kind, support, nsectors=1):
    m = binms.read_meta(path)
    start, rows = O.partition(m.nmeasures, 0, 1)
    d = O.read_binms(path, m.freq_per_chan, m.polarisations, start, rows)
    uu, vv, ww = O.normalize_uvw(d["uu"], d["vv"], d["ww"])
    dx = 1.0 / nx
    yaxis = nx // nsectors
    histo, st, lst, _ = O.sector_bin(vv, nsectors, yaxis, dx, support)
    grid = np.zeros((planes, nx, 2 * nx))
    for s in range(nsectors):
        g = O.sector_gather(lst[st[s]:st[s] + histo[s]], s, yaxis, dx, m.freq_per_chan,
                            m.polarisations, uu, vv, ww, d["weight"], d["vis_real"],
                            d["vis_img"])
        slab = O.wstack(kind, planes, g["uu"], g["vv"], g["ww"], g["vis_real"], g["vis_img"],
                        g["weight"], m.freq_per_chan, m.polarisations, dx, 1.0 / planes,
                        support, nx, yaxis)
        grid[:, s * yaxis:(s + 1) * yaxis, :] = slab.reshape(planes, yaxis, 2 * nx)
    grid = grid.reshape(-1)
    gridss = O.fft_planes(planes, nx, nx, grid)
    re, im = O.phase_correction(gridss, nx, nx, planes, m.resolution, m.wmin, m.wmax)
    return grid, re, im


def test_pipeline_binms_to_image_files(tmp_path):
    cfg = synth.CONFIGS["small"].scaled(n_ant=24, n_times=64, grid=256, planes=4,
                                        kernel=KAISERBESSEL, n_chan=2, n_pol=2)
    print(f"case pipeline rows={cfg.n_rows}", flush=True)
    write_synthetic_binms(tmp_path / "ms", cfg)
    out = tmp_path / "out"
    r = pipeline.image_from_binms(tmp_path / "ms", 256, 4, kernel=KAISERBESSEL, support=7,
                                  precision=FP64, out_dir=out, write_data=True)
    grid_o, re_o, im_o = oracle_flow(tmp_path / "ms", 256, 4, KAISERBESSEL, 7)
    assert r.rows == cfg.n_rows
    re, im = host(r.image_real), host(r.image_imag)
    err = rel_l2(np.concatenate([re, im]), np.concatenate([re_o, im_o]))
    print(f"pipeline image rel-L2 {err:.3e}", flush=True)
    assert err <= 1e-12
    np.testing.assert_array_equal(np.fromfile(out / "fft_real.bin"), re)
    np.testing.assert_array_equal(np.fromfile(out / "fft_img.bin"), im)
    gridded = np.fromfile(out / "gridded_data_real.bin")
    assert rel_l2(gridded, grid_o[0::2]) <= 1e-12
    assert os.path.getsize(out / "ffted_img.bin") == 4 * 256 * 256 * 8


def test_pipeline_as_four_ranks(tmp_path):
    # The same flow gridding four v-slabs, as four ranks would, against the
    # oracle doing the same.
    cfg = synth.CONFIGS["small"].scaled(n_ant=24, n_times=64, grid=256, planes=4,
                                        kernel=KAISERBESSEL, n_chan=2, n_pol=1)
    write_synthetic_binms(tmp_path / "ms", cfg)
    r = pipeline.image_from_binms(tmp_path / "ms", 256, 4, kernel=KAISERBESSEL, support=7,
                                  precision=FP64, nsectors=4)
    one = pipeline.image_from_binms(tmp_path / "ms", 256, 4, kernel=KAISERBESSEL, support=7,
                                    precision=FP64)
    _, re_o, im_o = oracle_flow(tmp_path / "ms", 256, 4, KAISERBESSEL, 7, nsectors=4)
    re, im = host(r.image_real), host(r.image_imag)
    err = rel_l2(np.concatenate([re, im]), np.concatenate([re_o, im_o]))
    slab_vs_whole = rel_l2(re, host(one.image_real))
    print(f"four-slab image rel-L2 {err:.3e}; four slabs vs one grid {slab_vs_whole:.3e}",
          flush=True)
    assert err <= 1e-12
    assert r.timings_ms["entries"] > r.rows       # halo rows gridded twice


def test_pipeline_fp32_and_polarizations(tmp_path):
    # FP32 path through the whole flow, four polarizations per row, against
    # the FP64 oracle over the same files: rel-L2 <= 1e-5.
    cfg = synth.CONFIGS["small"].scaled(n_ant=20, n_times=48, grid=128, planes=4,
                                        kernel=KAISERBESSEL, n_chan=3, n_pol=4)
    write_synthetic_binms(tmp_path / "ms", cfg)
    r = pipeline.image_from_binms(tmp_path / "ms", 128, 4, kernel=KAISERBESSEL, support=7,
                                  precision=FP32)
    _, re_o, im_o = oracle_flow(tmp_path / "ms", 128, 4, KAISERBESSEL, 7)
    re, im = host(r.image_real), host(r.image_imag)
    err = rel_l2(np.concatenate([re, im]), np.concatenate([re_o, im_o]))
    print(f"fp32 pipeline image rel-L2 {err:.3e}", flush=True)
    assert err <= 1e-5
