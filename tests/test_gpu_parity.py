"""Parity of the HIP path with the reference's outputs and the CPU oracle.

Bars: FP64 grid / image relative L2 <= 1e-12 against the reference-produced
vectors and the oracle (only the atomic/reduction order differs); FP32 path
relative L2 <= 1e-5 against the FP64 result.  FFT: against numpy's ifft2
(pocketfft) since FFTW, the reference's FFT, is not available.
"""
import numpy as np
import pytest
import torch

from conftest import load_golden, rel_l2
from oracle import oracle as O
import rick_amd
from rick_amd import FP32, FP64, Plan

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12
FP32_TOL = 1e-5


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def host(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


def grid_plan(z, precision=FP64, **kw):
    return Plan(int(z["nx"]), int(z["ny"]), int(z["planes"]), kernel=int(z["kind"]),
                support=int(z["support"]), dx=1.0 / int(z["nx"]), dw=1.0 / int(z["planes"]),
                precision=precision, max_points=int(z["uu"].size), **kw)


def grid_inputs(z):
    return [dev(z[k]) for k in ("uu", "vv", "ww", "vis_real", "vis_img", "weight")]


def nan_grid(plan):
    g = plan.new_grid()
    g.fill_(float("nan"))
    return g


# --- gridding -----------------------------------------------------------------

@pytest.mark.parametrize("kernel", ["gauss", "kb", "gausshp"])
@pytest.mark.parametrize("support", [7, 9])
def test_grid_fp64_matches_reference_vectors(kernel, support):
    z = load_golden(f"wstack_{kernel}_s{support}.npz")
    print(f"case wstack_{kernel}_s{support}", flush=True)
    plan = grid_plan(z)
    g = plan.grid(*grid_inputs(z), int(z["F"]), int(z["P"]), out=nan_grid(plan))
    got = host(g)
    assert np.isfinite(got).all()
    assert rel_l2(got, z["grid"]) <= FP64_TOL
    assert np.abs(got - z["grid"]).max() <= 1e-13 * np.abs(z["grid"]).max()
    assert np.array_equal(got == 0, z["grid"] == 0)


@pytest.mark.parametrize("tile", [(64, 64), (64, 32), (32, 32), (128, 32), (32, 64)])
@pytest.mark.parametrize("rows_per_task", [0, 5])
def test_grid_tiles_and_split_buckets(tile, rows_per_task):
    # rows_per_task=5 splits almost every bucket: exercises the atomic flush.
    z = load_golden("wstack_kb_s7.npz")
    print(f"case tile={tile} rows_per_task={rows_per_task}", flush=True)
    plan = grid_plan(z, tile=tile, rows_per_task=rows_per_task)
    got = host(plan.grid(*grid_inputs(z), int(z["F"]), int(z["P"]), out=nan_grid(plan)))
    assert rel_l2(got, z["grid"]) <= FP64_TOL


@pytest.mark.parametrize("precision,tile", [(FP64, (128, 32)), (FP32, (64, 64))])
def test_grid_persistent_tile_launch_matches_oracle(precision, tile):
    """2048^2 x 40 planes in tiles whose LDS allows two workgroups per CU:
    40,960 buckets, enough task slots that the tile kernel runs persistent
    workgroups pulling slots from a counter.  A cluster of 30,000 rows on
    three planes splits buckets over several tasks.  Two calls in a row: the
    slot counter must restart with each launch.  A third call with no rows
    runs only zero tasks, far more than the resident workgroups, and must
    clear the whole grid."""
    rng = np.random.default_rng(77)
    n_uni, n_hot, nx, planes, F, P = 50000, 30000, 2048, 40, 2, 1
    uu = np.concatenate([rng.uniform(0, 1, n_uni), rng.uniform(0.40, 0.41, n_hot)])
    vv = np.concatenate([rng.uniform(0, 1, n_uni), rng.uniform(0.55, 0.56, n_hot)])
    ww = np.concatenate([rng.uniform(0, 1, n_uni), rng.uniform(0.50, 0.575, n_hot)])
    n = uu.size
    vr = rng.normal(size=n * F * P).astype(np.float32)
    vi = rng.normal(size=n * F * P).astype(np.float32)
    wt = rng.uniform(0.5, 1.5, n * P).astype(np.float32)
    ref = O.wstack(O.KAISERBESSEL, planes, uu, vv, ww, vr, vi, wt, F, P, 1.0 / nx, 1.0 / planes,
                   7, nx, nx, threads=8)
    print(f"case persistent tiles precision={precision} tile={tile}", flush=True)
    plan = Plan(nx, nx, planes, kernel=rick_amd.KAISERBESSEL, support=7, precision=precision,
                max_points=n, tile=tile)
    ins = [dev(a) for a in (uu, vv, ww, vr, vi, wt)]
    tol = FP64_TOL if precision == FP64 else 1e-6
    for call in range(2):
        got = host(plan.grid(*ins, F, P, out=nan_grid(plan)))
        assert np.isfinite(got).all()
        err = rel_l2(got, ref)
        print(f"call {call}: rel L2 {err:.3e}", flush=True)
        assert err <= tol
        del got
    assert plan.grid_stats()["split_buckets"] > 0
    empty = [t[:0] for t in ins]
    got = plan.grid(*empty, F, P, out=nan_grid(plan))
    assert not host(got).any()


@pytest.mark.parametrize("precision", [FP64, FP32])
@pytest.mark.parametrize("F,P", [(2, 1), (16, 1), (4, 2), (64, 1)])
def test_grid_scattered_rows_overflow_block_tables(precision, F, P):
    """20,000 rows spread uniformly over 16 planes of 1024^2 in 32x32 tiles:
    each row block touches thousands of distinct buckets, more than its hash
    table holds, so many records take the direct-atomic path; also spans
    several row blocks and a partial last block.  Includes rows at w = 1.0
    (dropped) and NaN visibilities (skipped).  F*P in {2, 16, 8, 64} covers
    the per-row and the cooperative (coalesced) visibility sums."""
    rng = np.random.default_rng(123)
    n, nx, planes = 20000, 1024, 16
    uu = rng.uniform(0.0, 1.0, n)
    vv = rng.uniform(0.0, 1.0, n)
    ww = rng.uniform(0.0, 1.0, n)
    ww[::97] = 1.0
    vr = rng.normal(size=n * F * P).astype(np.float32)
    vi = rng.normal(size=n * F * P).astype(np.float32)
    vr[::53] = np.nan
    wt = rng.uniform(0.5, 1.5, n * P).astype(np.float32)
    ref = O.wstack(O.KAISERBESSEL, planes, uu, vv, ww, vr, vi, wt, F, P, 1.0 / nx, 1.0 / planes,
                   7, nx, nx, threads=8)
    print(f"case scattered precision={precision} F={F} P={P}", flush=True)
    plan = Plan(nx, nx, planes, kernel=rick_amd.KAISERBESSEL, support=7, precision=precision,
                max_points=n, tile=(32, 32))
    ins = [dev(a) for a in (uu, vv, ww, vr, vi, wt)]
    got = host(plan.grid(*ins, F, P, out=nan_grid(plan)))
    assert np.isfinite(got).all()
    assert rel_l2(got, ref) <= (FP64_TOL if precision == FP64 else 1e-6)
    st = plan.grid_stats()
    assert st["records"] >= n - n // 97 - 1


@pytest.mark.parametrize("name,unroll", [("wstack_kb_s7", 9), ("wstack_kb_s7", 11),
                                         ("wstack_gauss_s9", 15), ("wstack_gausshp_s7", 11)])
def test_grid_footprint_unroll_choices(name, unroll):
    # Every footprint unroll the tuner may pick grids the golden vectors.
    z = load_golden(f"{name}.npz")
    plan = grid_plan(z, unroll=unroll)
    got = host(plan.grid(*grid_inputs(z), int(z["F"]), int(z["P"]), out=nan_grid(plan)))
    assert rel_l2(got, z["grid"]) <= FP64_TOL
    with pytest.raises(rick_amd.RickError):
        grid_plan(z, unroll=5)        # below the support / not instantiated


@pytest.mark.parametrize("precision", [FP64, FP32])
def test_grid_nan_imaginary_part_propagates(precision):
    """Only the real part is tested for NaN (w-stacking.cu:517): a NaN
    imaginary part with a finite real part poisons the imaginary grid cells of
    that row's footprint, as in the reference.  (In the FP32 path this takes
    the FP64-atomic fallback of the fixed-point tile accumulation.)"""
    rng = np.random.default_rng(7)
    n, nx, planes = 3000, 256, 4
    uu, vv, ww = rng.uniform(0.1, 0.9, n), rng.uniform(0.1, 0.9, n), rng.uniform(0.0, 0.99, n)
    vr = rng.normal(size=n).astype(np.float32)
    vi = rng.normal(size=n).astype(np.float32)
    vi[17] = np.nan
    wt = np.ones(n, dtype=np.float32)
    ref = O.wstack(O.GAUSS, planes, uu, vv, ww, vr, vi, wt, 1, 1, 1.0 / nx, 1.0 / planes, 7, nx,
                   nx, threads=4)
    print(f"case nan-imag precision={precision}", flush=True)
    plan = Plan(nx, nx, planes, kernel=rick_amd.GAUSS, support=7, precision=precision,
                max_points=n)
    got = host(plan.grid(*[dev(a) for a in (uu, vv, ww, vr, vi, wt)], 1, 1, out=nan_grid(plan)))
    assert np.array_equal(np.isnan(got), np.isnan(ref))
    assert np.isnan(ref).sum() > 0
    ok = ~np.isnan(ref)
    assert rel_l2(got[ok], ref[ok]) <= (FP64_TOL if precision == FP64 else 1e-6)


def test_grid_accumulate_and_empty_input():
    z = load_golden("wstack_gauss_s7.npz")
    plan = grid_plan(z)
    ins = grid_inputs(z)
    g = plan.grid(*ins, int(z["F"]), int(z["P"]))
    plan.grid(*ins, int(z["F"]), int(z["P"]), out=g, accumulate=True)
    assert rel_l2(host(g), 2 * z["grid"]) <= FP64_TOL
    empty = [t[:0] for t in ins]
    g = plan.grid(*empty, int(z["F"]), int(z["P"]), out=nan_grid(plan))
    assert not host(g).any()


@pytest.mark.parametrize("name", ["wstack_kb_s9", "wstack_gauss_s7", "wstack_gausshp_s9"])
def test_grid_fp32_within_tolerance(name):
    """FP32 path (fixed-point tile accumulation) vs the reference's FP64 grid,
    for every convolution kernel (the hi-precision Gaussian bounds its cell
    weights by its norm)."""
    z = load_golden(f"{name}.npz")
    plan = grid_plan(z, precision=FP32)
    got = host(plan.grid(*grid_inputs(z), int(z["F"]), int(z["P"]), out=nan_grid(plan)))
    assert got.dtype == np.float32
    assert rel_l2(got, z["grid"]) <= 1e-6


def test_dropin_wstack_accumulates_into_host_grid():
    z = load_golden("wstack_gausshp_s9.npz")
    rick_amd.set_kernel(int(z["kind"]))
    try:
        pre = np.random.default_rng(0).normal(size=z["grid"].size)
        g = pre.copy()
        rick_amd.wstack(int(z["planes"]), int(z["uu"].size), int(z["F"]), int(z["P"]), z["uu"],
                        z["vv"], z["ww"], z["vis_real"], z["vis_img"], z["weight"],
                        1.0 / int(z["nx"]), 1.0 / int(z["planes"]), int(z["support"]),
                        int(z["nx"]), int(z["ny"]), g, 4, 0)
        assert rel_l2(g - pre, z["grid"]) <= 1e-12
    finally:
        rick_amd.set_kernel(rick_amd.KAISERBESSEL)


# --- FFT ------------------------------------------------------------------------

def cplx(g, planes, ny, nx):
    x = g.reshape(planes, ny, nx, 2).astype(np.float64)
    return x[..., 0] + 1j * x[..., 1]


@pytest.mark.parametrize("shape", [(16, 16), (32, 64), (64, 32), (512, 512), (1024, 256),
                                   (2048, 64), (64, 4096), (8192, 16), (16, 8192)])
@pytest.mark.parametrize("precision", [FP64, FP32])
@pytest.mark.parametrize("inplace", [False, True])
def test_fft_matches_numpy(shape, precision, inplace):
    ny, nx = shape
    planes = 3
    print(f"case fft {shape} prec={precision} inplace={inplace}", flush=True)
    rng = np.random.default_rng(ny + nx)
    c = rng.normal(size=(planes, ny, nx)) + 1j * rng.normal(size=(planes, ny, nx))
    flat = np.stack([c.real, c.imag], -1).reshape(-1)
    plan = Plan(nx, ny, planes, precision=precision)
    src = dev(flat if precision == FP64 else flat.astype(np.float32))
    out = src if inplace else torch.empty_like(src)
    plan.fft(src, out)
    got = cplx(host(out), planes, ny, nx)
    ref = np.fft.ifft2(c, axes=(1, 2))
    assert rel_l2(got, ref) <= (1e-14 if precision == FP64 else 2e-6)


# Axis lengths beside the power-of-two engine: mixed radix (2..16, 3, 5, 7,
# 11, 13), four-step for long rows and columns, Bluestein for other primes
# (FFTW plans any size, fourier_transform.c:34-36).
@pytest.mark.parametrize("shape,planes", [
    ((2000, 3000), 2),        # 2^4*5^3 x 2^3*3*5^3: four-step columns, direct rows of 3000
    ((6000, 6000), 1),        # four-step on both axes
    ((96, 128), 3),           # direct mixed radix
    ((1013, 1009), 1),        # primes: Bluestein on both axes (M = 2048)
    ((300, 4100), 1),         # 4100 = 4 * 5^2 * 41: Bluestein rows (M = 16384, four-step)
    ((64, 16384), 1),         # power of two above 8192: four-step rows
    ((7, 11), 2),             # short odd axes
    ((2, 8), 1),
])
@pytest.mark.parametrize("precision", [FP64, FP32])
@pytest.mark.parametrize("inplace", [False, True])
def test_fft_any_size_matches_numpy(shape, planes, precision, inplace):
    ny, nx = shape
    print(f"case fft any-size {shape} planes={planes} prec={precision} inplace={inplace}",
          flush=True)
    rng = np.random.default_rng(ny * 7 + nx)
    c = rng.normal(size=(planes, ny, nx)) + 1j * rng.normal(size=(planes, ny, nx))
    flat = np.stack([c.real, c.imag], -1).reshape(-1)
    plan = Plan(nx, ny, planes, precision=precision)
    src = dev(flat if precision == FP64 else flat.astype(np.float32))
    out = src if inplace else torch.empty_like(src)
    plan.fft(src, out)
    got = cplx(host(out), planes, ny, nx)
    ref = np.fft.ifft2(c, axes=(1, 2))
    err = rel_l2(got, ref)
    print(f"rel-L2 vs numpy {err:.3e}", flush=True)
    assert err <= (1e-14 if precision == FP64 else 2e-6)


@pytest.mark.parametrize("shape", [(6, 10), (13, 17), (12, 7)])
def test_fft_any_size_matches_brute_force_dft(shape):
    # The defining sum, X[v,u] = 1/(Nx Ny) sum x[k,j] exp(+2 pi i (ju/Nx + kv/Ny)),
    # in exact-index form (no FFT library involved).
    ny, nx = shape
    rng = np.random.default_rng(nx + 100 * ny)
    c = rng.normal(size=(ny, nx)) + 1j * rng.normal(size=(ny, nx))
    ku = np.exp(2j * np.pi * (np.outer(np.arange(nx), np.arange(nx)) % nx) / nx)
    kv = np.exp(2j * np.pi * (np.outer(np.arange(ny), np.arange(ny)) % ny) / ny)
    ref = kv @ c @ ku / (nx * ny)
    plan = Plan(nx, ny, 1)
    out = plan.fft(dev(np.stack([c.real, c.imag], -1).reshape(-1)))
    assert rel_l2(cplx(host(out), 1, ny, nx)[0], ref) <= 1e-14


@pytest.mark.parametrize("precision", [FP64, FP32])
def test_chain_any_size_matches_oracle(precision):
    # Gridding -> FFT -> phase on a 120 x 96 grid (mixed radix on both axes)
    # through rick_plan_image and the separate stages, against the oracle.
    nx, ny, planes, n = 120, 96, 4, 30000
    print(f"case chain any-size {nx}x{ny} prec={precision}", flush=True)
    rng = np.random.default_rng(5)
    uu = rng.uniform(0.05, 0.95, n)
    vv = rng.uniform(0.05, 0.75, n)          # v < Ny/Nx: dx = 1/Nx on both axes
    ww = rng.uniform(0.0, 0.999, n)
    vr = rng.normal(size=n).astype(np.float32)
    vi = rng.normal(size=n).astype(np.float32)
    wt = np.ones(n, np.float32)
    res, wmin, wmax = 1.0 / (3.0 * nx), -80.0, 120.0
    ref = O.wstack(O.KAISERBESSEL, planes, uu, vv, ww, vr, vi, wt, 1, 1, 1.0 / nx,
                   1.0 / planes, 7, nx, ny, threads=8)
    gs = O.fft_planes(planes, nx, ny, ref, threads=8)
    ore, oim = O.phase_correction(gs, nx, ny, planes, res, wmin, wmax, threads=8)
    want = ore + 1j * oim
    plan = Plan(nx, ny, planes, kernel=O.KAISERBESSEL, support=7, precision=precision,
                max_points=n, resolution=res, wmin=wmin, wmax=wmax)
    ins = [dev(a) for a in (uu, vv, ww, vr, vi, wt)]
    grid = nan_grid(plan)
    re, im, _ = plan.image(*ins, grid=grid)
    e_chain = rel_l2(host(re) + 1j * host(im), want)
    g = plan.grid(*ins)
    plan.fft(g)
    re2, im2 = plan.phase(g)
    e_sep = rel_l2(host(re2) + 1j * host(im2), want)
    print(f"image rel-L2 vs oracle: chain {e_chain:.3e}, separate {e_sep:.3e}", flush=True)
    tol = FP64_TOL if precision == FP64 else FP32_TOL
    assert e_chain <= tol and e_sep <= tol


def test_dropin_fft_planes():
    z = load_golden("chain_tiny.npz")
    gs = rick_amd.fft_planes(int(z["planes"]), int(z["nx"]), int(z["ny"]), z["grid"])
    assert rel_l2(gs, z["gridss"]) <= 1e-14


# --- phase correction ---------------------------------------------------------

@pytest.mark.parametrize("tag", ["on_p1", "on_p4", "off_p1", "off_p4"])
def test_phase_matches_reference_vectors(tag):
    z = load_golden(f"phase_{tag}.npz")
    nx, ny, planes = int(z["nx"]), int(z["ny"]), int(z["planes"])
    plan = Plan(nx, ny, planes, resolution=float(z["resolution"]), wmin=float(z["wmin"]),
                wmax=float(z["wmax"]), phase_on=bool(z["phase_on"]))
    re, im = plan.phase(dev(z["gridss"]))
    assert rel_l2(host(re) + 1j * host(im), z["image_real"] + 1j * z["image_imag"]) <= 1e-14
    # FP32 plane stack: phase argument still FP64.
    plan32 = Plan(nx, ny, planes, resolution=float(z["resolution"]), wmin=float(z["wmin"]),
                  wmax=float(z["wmax"]), phase_on=bool(z["phase_on"]), precision=FP32)
    re, im = plan32.phase(dev(z["gridss"].astype(np.float32)))
    assert rel_l2(host(re) + 1j * host(im), z["image_real"] + 1j * z["image_imag"]) <= 1e-6


def test_dropin_phase_correction_accumulates():
    z = load_golden("phase_on_p4.npz")
    nx, ny, planes = int(z["nx"]), int(z["ny"]), int(z["planes"])
    re = np.ones(nx * ny)
    im = np.full(nx * ny, -2.0)
    rick_amd.set_phase(True)
    rick_amd.phase_correction(z["gridss"], re, im, nx, ny, planes, nx, ny,
                              float(z["resolution"]), float(z["wmin"]), float(z["wmax"]), 2, 0)
    assert rel_l2(re - 1, z["image_real"]) <= 1e-14
    assert rel_l2(im + 2, z["image_imag"]) <= 1e-14


def test_phase_slab_matches_reference_vector():
    """One rank's slab, 32 x 12 pixels of a 36 x 48 image (the reference's
    call at fourier_transform.c:215): through the plan API (FP64 and FP32) and
    through the drop-in phase_correction(), which accumulates."""
    z = load_golden("phase_on_slab.npz")
    nx, ny, planes = int(z["nx"]), int(z["ny"]), int(z["planes"])
    xtot, ytot = int(z["xtot"]), int(z["ytot"])
    ref = z["image_real"] + 1j * z["image_imag"]
    for precision, tol in ((FP64, 1e-14), (FP32, 1e-6)):
        plan = Plan(nx, ny, planes, resolution=float(z["resolution"]), wmin=float(z["wmin"]),
                    wmax=float(z["wmax"]), precision=precision, xaxistot=xtot, yaxistot=ytot)
        src = z["gridss"] if precision == FP64 else z["gridss"].astype(np.float32)
        re, im = plan.phase(dev(src))
        assert rel_l2(host(re) + 1j * host(im), ref) <= tol
    re = np.full(nx * ny, 0.5)
    im = np.zeros(nx * ny)
    rick_amd.set_phase(True)
    rick_amd.phase_correction(z["gridss"], re, im, nx, ny, planes, xtot, ytot,
                              float(z["resolution"]), float(z["wmin"]), float(z["wmax"]), 2, 0)
    assert rel_l2((re - 0.5) + 1j * im, ref) <= 1e-14


def test_phase_rejects_degenerate_w_range():
    plan = Plan(16, 16, 4, wmin=5.0, wmax=5.0)
    with pytest.raises(rick_amd.RickError):
        plan.phase(torch.zeros(2 * 4 * 16 * 16, dtype=torch.float64, device="cuda"))


# --- whole chain --------------------------------------------------------------

def chain(z, precision):
    nx, ny, planes = int(z["nx"]), int(z["ny"]), int(z["planes"])
    plan = Plan(nx, ny, planes, kernel=int(z["kind"]), support=int(z["support"]), dx=1.0 / nx,
                dw=1.0 / planes, precision=precision, max_points=int(z["uu"].size),
                resolution=float(z["resolution"]), wmin=float(z["wmin"]), wmax=float(z["wmax"]))
    re, im, _ = plan.image(*grid_inputs(z), int(z["F"]), int(z["P"]))
    return host(re) + 1j * host(im)


def test_chain_tiny_matches_reference_vectors():
    z = load_golden("chain_tiny.npz")
    ref = z["image_real"] + 1j * z["image_imag"]
    assert rel_l2(chain(z, FP64), ref) <= FP64_TOL
    assert rel_l2(chain(z, FP32), ref) <= FP32_TOL


@pytest.fixture(scope="module")
def small_case():
    from rick_amd import synth
    ds = synth.make_dataset(synth.CONFIGS["small"])
    cfg = ds.cfg
    grid = O.wstack(cfg.kernel, cfg.planes, ds.uu, ds.vv, ds.ww, ds.vis_real, ds.vis_img,
                    ds.weight, cfg.n_chan, cfg.n_pol, ds.dx, ds.dw, cfg.support, cfg.grid,
                    cfg.grid, threads=8)
    gridss = O.fft_planes(cfg.planes, cfg.grid, cfg.grid, grid, threads=8)
    re, im = O.phase_correction(gridss, cfg.grid, cfg.grid, cfg.planes, ds.resolution, ds.wmin,
                                ds.wmax, threads=8)
    return ds, grid, re + 1j * im


@pytest.mark.parametrize("precision", [FP64, FP32])
def test_small_config_matches_oracle(small_case, precision):
    """512^2 x 8 planes, 516,096 visibilities, Gaussian support 7."""
    ds, grid_ref, img_ref = small_case
    print(f"case small precision={precision}", flush=True)
    plan = rick_amd.plan_for(ds, precision=precision)
    ins = [dev(a) for a in (ds.uu, ds.vv, ds.ww, ds.vis_real, ds.vis_img, ds.weight)]
    g = plan.grid(*ins, ds.cfg.n_chan, ds.cfg.n_pol)
    e_grid = rel_l2(host(g), grid_ref)
    re, im, _ = plan.image(*ins, ds.cfg.n_chan, ds.cfg.n_pol)
    e_img = rel_l2(host(re) + 1j * host(im), img_ref)
    print(f"small precision={precision}: grid rel-L2 {e_grid:.3e}, image rel-L2 {e_img:.3e}")
    tol = FP64_TOL if precision == FP64 else FP32_TOL
    assert e_grid <= tol and e_img <= tol


@pytest.mark.parametrize("rows_per_task", [1024, 16384])
@pytest.mark.parametrize("tile", [(64, 64), (64, 32), (32, 32), (128, 32), (32, 64)])
@pytest.mark.parametrize("precision", [FP64, FP32])
def test_chain_every_tile_shape_matches_oracle(small_case, precision, tile, rows_per_task):
    """The image chain on each tile shape the gridder instantiates (the image
    chain is tuned separately and may run on a shape the gridding step does
    not use): the sparse grid, the row FFT's tile and band skipping and the
    fused phase pass against the oracle's image, 512^2 x 8 planes."""
    ds, _, img_ref = small_case
    print(f"case chain tile={tile} rows_per_task={rows_per_task} precision={precision}",
          flush=True)
    plan = rick_amd.plan_for(ds, precision=precision, tile=tile, rows_per_task=rows_per_task)
    ins = [dev(a) for a in (ds.uu, ds.vv, ds.ww, ds.vis_real, ds.vis_img, ds.weight)]
    grid = plan.new_grid()
    grid.fill_(float("nan"))          # cells of empty tiles are never read
    re, im, _ = plan.image(*ins, ds.cfg.n_chan, ds.cfg.n_pol, grid=grid)
    err = rel_l2(host(re) + 1j * host(im), img_ref)
    print(f"chain tile={tile} rpt={rows_per_task} precision={precision}: rel-L2 {err:.3e}")
    assert err <= (FP64_TOL if precision == FP64 else FP32_TOL)


@pytest.mark.parametrize("tile", [(128, 32), (32, 64)])
@pytest.mark.parametrize("precision", [FP64, FP32])
def test_chain_fourstep_matches_oracle(precision, tile):
    """1024^2 x 4 planes: the chain's in-place four-step column FFT (rows left
    permuted and read back through the permutation by the phase stage)."""
    from rick_amd import synth
    cfg = synth.CONFIGS["small"].scaled(name="small1024", grid=1024, planes=4, n_times=64,
                                        kernel=synth.KAISERBESSEL, support=9)
    ds = synth.make_dataset(cfg)
    grid = O.wstack(cfg.kernel, cfg.planes, ds.uu, ds.vv, ds.ww, ds.vis_real, ds.vis_img,
                    ds.weight, cfg.n_chan, cfg.n_pol, ds.dx, ds.dw, cfg.support, cfg.grid,
                    cfg.grid, threads=8)
    gridss = O.fft_planes(cfg.planes, cfg.grid, cfg.grid, grid, threads=8)
    re, im = O.phase_correction(gridss, cfg.grid, cfg.grid, cfg.planes, ds.resolution, ds.wmin,
                                ds.wmax, threads=8)
    ref = re + 1j * im
    print(f"case fourstep precision={precision} tile={tile}", flush=True)
    plan = rick_amd.plan_for(ds, precision=precision, tile=tile)
    ins = [dev(a) for a in (ds.uu, ds.vv, ds.ww, ds.vis_real, ds.vis_img, ds.weight)]
    gre, gim, _ = plan.image(*ins, cfg.n_chan, cfg.n_pol)
    err = rel_l2(host(gre) + 1j * host(gim), ref)
    # The natural-order FFT entry point on the same grid agrees too, and so
    # does the unfused grid -> fft -> phase sequence.
    g = plan.grid(*ins, cfg.n_chan, cfg.n_pol)
    out = torch.empty_like(g)
    plan.fft(g, out)
    e_fft = rel_l2(host(out), gridss)
    sre, sim = plan.phase(out)
    e_sep = rel_l2(host(sre) + 1j * host(sim), ref)
    e_cs = rel_l2(host(gre) + 1j * host(gim), host(sre) + 1j * host(sim))
    print(f"fourstep precision={precision}: image rel-L2 {err:.3e}, fft rel-L2 {e_fft:.3e}, "
          f"separate-path image {e_sep:.3e}, chain vs separate {e_cs:.3e}")
    tol = FP64_TOL if precision == FP64 else FP32_TOL
    assert err <= tol and e_fft <= tol and e_sep <= tol


@pytest.mark.parametrize("precision", [FP64, FP32])
def test_phase_large_w_matches_oracle(precision):
    """|w| up to 9e4 wavelengths: phase arguments of ~1e3 rad exercise the
    argument reduction of the device sin/cos."""
    nx, ny, planes = 64, 48, 8
    rng = np.random.default_rng(17)
    gridss = rng.normal(size=2 * planes * nx * ny)
    res, wmin, wmax = 1.0 / 900.0, -3.0e4, 9.0e4
    re, im = O.phase_correction(gridss, nx, ny, planes, res, wmin, wmax, threads=4)
    plan = Plan(nx, ny, planes, resolution=res, wmin=wmin, wmax=wmax, precision=precision)
    src = dev(gridss if precision == FP64 else gridss.astype(np.float32))
    gre, gim = plan.phase(src)
    err = rel_l2(host(gre) + 1j * host(gim), re + 1j * im)
    print(f"large-w phase precision={precision}: rel-L2 {err:.3e}")
    assert err <= (1e-13 if precision == FP64 else 1e-6)


@pytest.mark.parametrize("ny", [256, 512])
@pytest.mark.parametrize("precision", [FP64, FP32])
def test_chain_band_flags_differ_between_planes(ny, precision):
    """Image chain with the single fused column pass (ny <= 512): each plane's
    rows occupy different tile bands and the grid starts as NaN.  The column
    pass takes each plane's band flags from LDS while the FP64 variant loads
    the next plane ahead; flags read for the wrong plane would either drop
    live rows of a plane or let the NaN of never-written rows into the image."""
    rng = np.random.default_rng(31 + ny)
    nx, planes, n_per = 256, 6, 3000
    uu, vv, ww = [], [], []
    for p in range(planes):
        uu.append(rng.uniform(0.05, 0.95, n_per))
        lo = 0.05 + 0.15 * p
        vv.append(rng.uniform(lo, lo + 0.05, n_per))
        ww.append((p + rng.uniform(0.1, 0.9, n_per)) / planes)
    uu, vv, ww = (np.concatenate(a) for a in (uu, vv, ww))
    n = uu.size
    vr = rng.normal(size=n).astype(np.float32)
    vi = rng.normal(size=n).astype(np.float32)
    wt = rng.uniform(0.5, 1.5, n).astype(np.float32)
    res, wmin, wmax = 0.4 / max(nx, ny), -200.0, 200.0
    g = O.wstack(O.KAISERBESSEL, planes, uu, vv, ww, vr, vi, wt, 1, 1, 1.0 / nx, 1.0 / planes, 7,
                 nx, ny, threads=8)
    gs = O.fft_planes(planes, nx, ny, g, threads=8)
    ore, oim = O.phase_correction(gs, nx, ny, planes, res, wmin, wmax, threads=8)
    ref = ore + 1j * oim
    print(f"case band flags ny={ny} precision={precision}", flush=True)
    plan = Plan(nx, ny, planes, kernel=rick_amd.KAISERBESSEL, support=7, dx=1.0 / nx,
                dw=1.0 / planes, precision=precision, max_points=n, resolution=res, wmin=wmin,
                wmax=wmax)
    ins = [dev(a) for a in (uu, vv, ww, vr, vi, wt)]
    for call in range(2):
        re, im, _ = plan.image(*ins, 1, 1, grid=nan_grid(plan))
        got = host(re) + 1j * host(im)
        assert np.isfinite(got).all()
        err = rel_l2(got, ref)
        print(f"call {call}: rel-L2 {err:.3e}", flush=True)
        assert err <= (FP64_TOL if precision == FP64 else FP32_TOL)


@pytest.mark.parametrize("tile", [(128, 32), (32, 64)])
@pytest.mark.parametrize("precision", [FP64, FP32])
def test_chain_row_flags_skip_rows_inside_bands(precision, tile):
    """Image chain on 512 x 1024 x 4 planes (four-step column FFT; the row FFT
    takes 8 rows per workgroup) with the visibilities in thin v-stripes that
    start and end inside tile bands, over part of the u range.  The tile
    kernel flags only the grid rows it writes a non-zero cell to; the row FFT
    skips the others (whole workgroups, or the rows' loads) and the column
    passes read them as zero.  The grid starts as NaN, so a row read that
    should have been skipped, in a tile that was never written, shows as NaN;
    the second call puts the stripes elsewhere, so flags or rows left from
    the first call change the image."""
    rng = np.random.default_rng(47)
    nx, ny, planes, per = 512, 1024, 4, 900
    res, wmin, wmax = 0.4 / ny, -200.0, 200.0
    plan = None
    for call, centres in enumerate(([0.2013, 0.4470, 0.7302], [0.3108, 0.6219, 0.8825])):
        uu = np.concatenate([rng.uniform(0.1, 0.6, per) for _ in centres])
        vv = np.concatenate([c + rng.uniform(0.0, 3.0 / ny, per) for c in centres])
        ww = rng.uniform(0.02, 0.98, uu.size)
        n = uu.size
        vr = rng.normal(size=n).astype(np.float32)
        vi = rng.normal(size=n).astype(np.float32)
        wt = rng.uniform(0.5, 1.5, n).astype(np.float32)
        g = O.wstack(O.KAISERBESSEL, planes, uu, vv, ww, vr, vi, wt, 1, 1, 1.0 / nx,
                     1.0 / planes, 7, nx, ny, threads=8)
        gs = O.fft_planes(planes, nx, ny, g, threads=8)
        ore, oim = O.phase_correction(gs, nx, ny, planes, res, wmin, wmax, threads=8)
        ref = ore + 1j * oim
        print(f"case row flags call={call} tile={tile} precision={precision}", flush=True)
        if plan is None:
            plan = Plan(nx, ny, planes, kernel=rick_amd.KAISERBESSEL, support=7, dx=1.0 / nx,
                        dw=1.0 / planes, precision=precision, max_points=n, resolution=res,
                        wmin=wmin, wmax=wmax, tile=tile)
            grid = nan_grid(plan)
        ins = [dev(a) for a in (uu, vv, ww, vr, vi, wt)]
        re, im, _ = plan.image(*ins, 1, 1, grid=grid)
        got = host(re) + 1j * host(im)
        assert np.isfinite(got).all()
        err = rel_l2(got, ref)
        print(f"call {call}: rel-L2 {err:.3e}", flush=True)
        assert err <= (FP64_TOL if precision == FP64 else FP32_TOL)


_ONE_ROW_CASE = {}


def one_row_case():
    """4096^2 x 2 planes, visibilities in two v-bands over the middle of the
    u range, and the oracle's image of them (computed once)."""
    if not _ONE_ROW_CASE:
        rng = np.random.default_rng(5)
        nx = ny = 4096
        planes, n = 2, 20000
        uu = rng.uniform(0.3, 0.7, n)
        vv = np.concatenate([rng.uniform(0.30, 0.36, n // 2), rng.uniform(0.55, 0.58, n - n // 2)])
        ww = rng.uniform(0.02, 0.98, n)
        vr = rng.normal(size=n).astype(np.float32)
        vi = rng.normal(size=n).astype(np.float32)
        wt = rng.uniform(0.5, 1.5, n).astype(np.float32)
        res, wmin, wmax = 0.4 / ny, -200.0, 200.0
        g = O.wstack(O.KAISERBESSEL, planes, uu, vv, ww, vr, vi, wt, 1, 1, 1.0 / nx,
                     1.0 / planes, 7, nx, ny, threads=8)
        gs = O.fft_planes(planes, nx, ny, g, threads=8)
        ore, oim = O.phase_correction(gs, nx, ny, planes, res, wmin, wmax, threads=8)
        _ONE_ROW_CASE.update(nx=nx, ny=ny, planes=planes, n=n, res=res, wmin=wmin, wmax=wmax,
                             ins=(uu, vv, ww, vr, vi, wt), ref=ore + 1j * oim)
    return _ONE_ROW_CASE


@pytest.mark.parametrize("tile", [(64, 64), (32, 32)])
@pytest.mark.parametrize("precision", [FP64, FP32])
def test_chain_one_row_workgroups_both_occupancy_paths(precision, tile):
    """Image chain at 4096^2, where the row FFT takes one row per workgroup.
    With 64-wide tiles a row has 64 tiles and the row kernel's occupancy-mask
    instance runs; with 32-wide tiles a row has 128 and the instance with the
    LDS occupancy table runs.  The grid starts as NaN, so a load from a tile
    that was never written shows in the image."""
    c = one_row_case()
    print(f"case one-row chain tile={tile} precision={precision}", flush=True)
    plan = Plan(c["nx"], c["ny"], c["planes"], kernel=rick_amd.KAISERBESSEL, support=7,
                dx=1.0 / c["nx"], dw=1.0 / c["planes"], precision=precision, max_points=c["n"],
                resolution=c["res"], wmin=c["wmin"], wmax=c["wmax"], tile=tile)
    re, im, _ = plan.image(*[dev(a) for a in c["ins"]], 1, 1, grid=nan_grid(plan))
    got = host(re) + 1j * host(im)
    assert np.isfinite(got).all()
    err = rel_l2(got, c["ref"])
    print(f"one-row chain tile={tile} precision={precision}: rel-L2 {err:.3e}", flush=True)
    assert err <= (FP64_TOL if precision == FP64 else FP32_TOL)


def test_c_program_calls_dropin_chain(tmp_path):
    """A C program (tests/c/dropin_caller.c) linked with -lrick_hip calls
    wstack() -> rick_fft_planes() -> phase_correction() on host arrays; its
    grid and image match the oracle run on the inputs it wrote."""
    import subprocess
    from test_capi import build_c_caller
    exe = build_c_caller(tmp_path / "dropin_caller")
    print("case C drop-in caller", flush=True)
    r = subprocess.run([str(exe), str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    nx, ny, planes, F, P, n = 128, 96, 4, 2, 1, 20000
    raw = (tmp_path / "inputs.bin").read_bytes()
    f64 = np.frombuffer(raw, np.float64, 3 * n)
    f32 = np.frombuffer(raw, np.float32, 2 * n * F * P + n * P, offset=24 * n)
    uu, vv, ww = f64[:n], f64[n:2 * n], f64[2 * n:]
    vr, vi, wt = f32[:n * F * P], f32[n * F * P:2 * n * F * P], f32[2 * n * F * P:]
    grid = np.fromfile(tmp_path / "grid.bin", np.float64)
    image = np.fromfile(tmp_path / "image.bin", np.float64)
    ref = O.wstack(O.KAISERBESSEL, planes, uu, vv, ww, vr, vi, wt, F, P, 1.0 / nx, 1.0 / planes,
                   7, nx, ny, threads=8)
    gs = O.fft_planes(planes, nx, ny, ref, threads=8)
    ore, oim = O.phase_correction(gs, nx, ny, planes, 1.0 / (3.0 * nx), -150.0, 250.0, threads=8)
    e_grid = rel_l2(grid, ref)
    e_img = rel_l2(image[:nx * ny] + 1j * image[nx * ny:], ore + 1j * oim)
    print(f"C caller: grid rel-L2 {e_grid:.3e}, image rel-L2 {e_img:.3e}")
    assert e_grid <= FP64_TOL and e_img <= FP64_TOL
