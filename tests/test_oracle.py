"""The CPU oracle against the reference's own outputs and against mathematics.

The oracle (oracle/rick_oracle.c) is the checker for the HIP path, so it is
pinned first: bitwise against vectors produced by the reference's compiled
wstack()/phase_correction() (tests/golden), against the reference-compiled
library directly where it is built, and against closed forms (I0 series,
DFT, Parseval, impulses, zero-resolution phase) where no reference vector
exists (the FFT stage: FFTW is absent).
"""
import numpy as np
import pytest
import scipy.special

from conftest import load_golden, rel_l2
from oracle import oracle as O

KINDS = {"gauss": O.GAUSS, "kb": O.KAISERBESSEL, "gausshp": O.GAUSS_HI_PRECISION}


def run_oracle_wstack(z, threads=1):
    return O.wstack(int(z["kind"]), int(z["planes"]), z["uu"], z["vv"], z["ww"], z["vis_real"],
                    z["vis_img"], z["weight"], int(z["F"]), int(z["P"]), 1.0 / int(z["nx"]),
                    1.0 / int(z["planes"]), int(z["support"]), int(z["nx"]), int(z["ny"]),
                    threads=threads)


@pytest.mark.parametrize("kernel", ["gauss", "kb", "gausshp"])
@pytest.mark.parametrize("support", [7, 9])
def test_wstack_matches_reference_vectors(kernel, support):
    z = load_golden(f"wstack_{kernel}_s{support}.npz")
    g = run_oracle_wstack(z)
    # Single-threaded: same operations in the same order -> bitwise equal.
    assert np.array_equal(g, z["grid"])
    assert np.count_nonzero(z["grid"]) > 1000


@pytest.mark.parametrize("kernel", ["gauss", "kb"])
def test_wstack_threads_only_reorder(kernel):
    z = load_golden(f"wstack_{kernel}_s7.npz")
    assert rel_l2(run_oracle_wstack(z, threads=4), z["grid"]) < 1e-14


@pytest.mark.parametrize("tag", ["on_p1", "on_p4", "off_p1", "off_p4"])
def test_phase_matches_reference_vectors(tag):
    z = load_golden(f"phase_{tag}.npz")
    re, im = O.phase_correction(z["gridss"], int(z["nx"]), int(z["ny"]), int(z["planes"]),
                                float(z["resolution"]), float(z["wmin"]), float(z["wmax"]),
                                phase_on=bool(z["phase_on"]), threads=1)
    assert np.array_equal(re, z["image_real"]) and np.array_equal(im, z["image_imag"])


def test_phase_slab_matches_reference_vector():
    """One rank's slab (xaxis != xaxistot, yaxis != yaxistot): pixel
    coordinates use the local iu, iv against the totals (phase_correction.cu:
    204-209), the gridss stride the local xaxis."""
    z = load_golden("phase_on_slab.npz")
    re, im = O.phase_correction(z["gridss"], int(z["nx"]), int(z["ny"]), int(z["planes"]),
                                float(z["resolution"]), float(z["wmin"]), float(z["wmax"]),
                                phase_on=True, threads=1, xtot=int(z["xtot"]),
                                ytot=int(z["ytot"]))
    assert np.array_equal(re, z["image_real"]) and np.array_equal(im, z["image_imag"])
    # The totals matter: the same slab as a whole image differs.
    re2, _ = O.phase_correction(z["gridss"], int(z["nx"]), int(z["ny"]), int(z["planes"]),
                                float(z["resolution"]), float(z["wmin"]), float(z["wmax"]),
                                phase_on=True, threads=1)
    assert not np.allclose(re2, z["image_real"])


def test_chain_tiny_matches_reference_vectors():
    z = load_golden("chain_tiny.npz")
    g = run_oracle_wstack(z)
    assert np.array_equal(g, z["grid"])
    gs = O.fft_planes(int(z["planes"]), int(z["nx"]), int(z["ny"]), g, threads=1)
    assert np.array_equal(gs, z["gridss"])
    re, im = O.phase_correction(gs, int(z["nx"]), int(z["ny"]), int(z["planes"]),
                                float(z["resolution"]), float(z["wmin"]), float(z["wmax"]),
                                phase_on=True, threads=1)
    assert np.array_equal(re, z["image_real"]) and np.array_equal(im, z["image_imag"])


@pytest.mark.skipif(not O.ref_available(), reason="reference sources not built here")
@pytest.mark.parametrize("kind", [O.GAUSS, O.KAISERBESSEL, O.GAUSS_HI_PRECISION])
def test_wstack_matches_reference_library(kind):
    rng = np.random.default_rng(kind)
    n, F, P, nx, ny, planes = 500, 3, 1, 40, 56, 3
    uu, vv, ww = rng.uniform(0, 1, n), rng.uniform(0, 1, n), rng.uniform(0, 1, n)
    vr = rng.normal(size=n * F * P).astype(np.float32)
    vi = rng.normal(size=n * F * P).astype(np.float32)
    wt = rng.uniform(0.5, 1.5, n * P).astype(np.float32)
    a = O.wstack(kind, planes, uu, vv, ww, vr, vi, wt, F, P, 1 / nx, 1 / planes, 9, nx, ny,
                 threads=1)
    b = O.ref_wstack(kind, planes, uu, vv, ww, vr, vi, wt, F, P, 1 / nx, 1 / planes, 9, nx, ny)
    assert np.array_equal(a, b)


@pytest.mark.skipif(not O.ref_available(), reason="reference sources not built here")
@pytest.mark.parametrize("support", [3, 5, 7, 9, 11])
def test_tables_match_reference_library(support):
    for kind in (O.GAUSS, O.KAISERBESSEL):
        assert np.array_equal(O.kernel_table(kind, support), O.ref_kernel_table(kind, support))


# --- tables: closed forms ---------------------------------------------------

@pytest.mark.parametrize("support", [7, 9])
def test_kaiser_bessel_table_closed_form(support):
    t = O.kernel_table(O.KAISERBESSEL, support)
    n, mid = 5 * support, 5 * support // 2
    assert t.size == n
    assert np.array_equal(t, t[::-1])                # mirrored
    assert t[mid] == 1.0                             # I0(8.6)/I0(8.6)
    i = np.arange(mid + 1)
    expect = scipy.special.i0(8.6 * np.sqrt(1 - (i / mid) ** 2)) / scipy.special.i0(8.6)
    # series stopped at ds <= 1e-8*sum (w-stacking.cu:62)
    np.testing.assert_allclose(t[mid:], expect, rtol=2e-8, atol=0)
    assert abs(t[-1] - 1 / scipy.special.i0(8.6)) < 1e-8 * t[-1]


@pytest.mark.parametrize("support", [7, 9])
def test_gauss_table_closed_form(support):
    t = O.kernel_table(O.GAUSS, support)
    mid = 5 * support // 2
    assert np.array_equal(t, t[::-1])
    x = np.arange(mid + 1) / 5.0
    pi_ref = 3.14159265359
    np.testing.assert_allclose(t[mid:], np.sqrt(0.5 / pi_ref) * np.exp(-0.5 * x * x),
                               rtol=1e-15)


def test_bessel0_series():
    for x in (0.0, 0.5, 3.0, 8.6):
        assert abs(O.bessel0(x, 1e-14) - scipy.special.i0(x)) < 1e-13 * scipy.special.i0(x)


# --- wstack: structural properties --------------------------------------------

def one_row(u, v, w, vis=(1.0, 0.0), weight=1.0, kind=O.KAISERBESSEL, support=7, nx=32, ny=32,
            planes=2):
    return O.wstack(kind, planes, np.array([u]), np.array([v]), np.array([w]),
                    np.array([vis[0]], np.float32), np.array([vis[1]], np.float32),
                    np.array([weight], np.float32), 1, 1, 1 / nx, 1 / planes, support, nx, ny,
                    threads=1).reshape(planes, ny, nx, 2)


def test_single_row_footprint_and_separable_weights():
    nx = ny = 32
    t = O.kernel_table(O.KAISERBESSEL, 7)
    pu, pv = 10.3, 17.8                       # cells
    g = one_row(pu / nx, pv / nx, 0.6)
    re = g[..., 0]
    nz = np.argwhere(re != 0)
    assert set(nz[:, 0]) == {1}               # plane int(0.6/0.5) = 1
    assert nz[:, 1].min() == 17 - 3 and nz[:, 1].max() == 17 + 3
    assert nz[:, 2].min() == 10 - 3 and nz[:, 2].max() == 10 + 3
    # weight = K[int(5|j+0.5-pu+KL|)] * K[int(5|k+0.5-pv+KL|)]
    ju = [t[int(5 * abs(j + 0.5 - pu + 3))] for j in range(7, 14)]
    kv = [t[int(5 * abs(k + 0.5 - pv + 3))] for k in range(14, 21)]
    np.testing.assert_allclose(re[1, 14:21, 7:14], np.outer(kv, ju), rtol=1e-15)
    assert np.all(g[..., 1] == 0)


def test_rows_outside_planes_are_dropped_and_edges_clip():
    assert not one_row(0.5, 0.5, 1.0).any()            # w = 1.0 -> plane index == planes
    g = one_row(0.5 / 32, 0.5, 0.1)                     # u at the left edge: 4 columns
    cols = np.unique(np.argwhere(g[..., 0] != 0)[:, 2])
    assert cols.tolist() == [0, 1, 2, 3]
    g = one_row(0.999, 0.5, 0.1)                        # right edge
    cols = np.unique(np.argwhere(g[..., 0] != 0)[:, 2])
    assert cols.max() == 31


def test_nan_real_part_skips_visibility():
    assert not one_row(0.5, 0.5, 0.3, vis=(np.nan, 1.0)).any()


def test_float_product_rounding():
    # weight*vis rounds in float32 before the double weight multiplies it.
    w, v = np.float32(0.1), np.float32(3.3)
    g = one_row(0.5, 0.5, 0.3, vis=(float(v), 0.0), weight=float(w), kind=O.GAUSS)
    ref = one_row(0.5, 0.5, 0.3, vis=(float(np.float32(w * v)), 0.0), weight=1.0, kind=O.GAUSS)
    assert np.array_equal(g, ref)
    assert np.float64(np.float32(w * v)) != np.float64(w) * np.float64(v)


def test_linearity_and_total_weight():
    rng = np.random.default_rng(3)
    n, nx, planes = 200, 48, 3
    uu, vv, ww = rng.uniform(0.1, 0.9, n), rng.uniform(0.1, 0.9, n), rng.uniform(0, 0.99, n)
    vr = rng.normal(size=n).astype(np.float32)
    vi = rng.normal(size=n).astype(np.float32)
    wt = np.ones(n, np.float32)
    args = (planes, uu, vv, ww)
    g1 = O.wstack(O.GAUSS, *args, vr, vi, wt, 1, 1, 1 / nx, 1 / planes, 7, nx, nx, threads=1)
    g2 = O.wstack(O.GAUSS, *args, 2 * vr, 2 * vi, wt, 1, 1, 1 / nx, 1 / planes, 7, nx, nx,
                  threads=1)
    assert np.array_equal(2 * g1, g2)
    # Total real weight = sum_rows v * (sum_j K[jKer]) * (sum_k K[kKer]).
    t = O.kernel_table(O.GAUSS, 7)
    tot = 0.0
    for i in range(n):
        pu, pv = uu[i] * nx, vv[i] * nx
        su = sum(t[int(5 * abs(j + 0.5 - pu + 3))] for j in range(int(pu) - 3, int(pu) + 4))
        sv = sum(t[int(5 * abs(k + 0.5 - pv + 3))] for k in range(int(pv) - 3, int(pv) + 4))
        tot += float(vr[i]) * su * sv
    assert abs(g1[0::2].sum() - tot) < 1e-12 * np.abs(g1).sum()


# --- FFT ----------------------------------------------------------------------

def as_complex(g, planes, ny, nx):
    x = g.reshape(planes, ny, nx, 2)
    return x[..., 0] + 1j * x[..., 1]


def from_complex(c):
    return np.stack([c.real, c.imag], axis=-1).reshape(-1)


def dft2_backward(c):
    ny, nx = c.shape
    ky = np.exp(2j * np.pi * np.outer(np.arange(ny), np.arange(ny)) / ny)
    kx = np.exp(2j * np.pi * np.outer(np.arange(nx), np.arange(nx)) / nx)
    return ky @ c @ kx.T / (nx * ny)


@pytest.mark.parametrize("shape", [(4, 8), (8, 8), (16, 4), (6, 10), (12, 16)])
def test_fft_matches_brute_force_dft(shape):
    ny, nx = shape
    rng = np.random.default_rng(ny * 100 + nx)
    c = rng.normal(size=(2, ny, nx)) + 1j * rng.normal(size=(2, ny, nx))
    out = as_complex(O.fft_planes(2, nx, ny, from_complex(c)), 2, ny, nx)
    for p in range(2):
        assert rel_l2(out[p], dft2_backward(c[p])) < 1e-14


def test_fft_matches_numpy_ifft2_and_invariants():
    ny, nx, planes = 64, 128, 3
    rng = np.random.default_rng(1)
    c = rng.normal(size=(planes, ny, nx)) + 1j * rng.normal(size=(planes, ny, nx))
    out = as_complex(O.fft_planes(planes, nx, ny, from_complex(c)), planes, ny, nx)
    assert rel_l2(out, np.fft.ifft2(c, axes=(1, 2))) < 1e-15
    # Parseval for the 1/N-scaled backward transform.
    assert abs(np.sum(np.abs(out) ** 2) * nx * ny / np.sum(np.abs(c) ** 2) - 1) < 1e-13
    # Impulse at (v0, u0) -> plane wave exp(+2 pi i (u0 x / nx + v0 y / ny)) / N.
    d = np.zeros((1, ny, nx), complex)
    d[0, 5, 9] = 1.0
    out = as_complex(O.fft_planes(1, nx, ny, from_complex(d)), 1, ny, nx)[0]
    y, x = np.mgrid[0:ny, 0:nx]
    assert rel_l2(out, np.exp(2j * np.pi * (9 * x / nx + 5 * y / ny)) / (nx * ny)) < 1e-14
    # Constant -> impulse at the origin.
    out = as_complex(O.fft_planes(1, nx, ny, from_complex(np.ones((1, ny, nx), complex))), 1,
                     ny, nx)[0]
    e = np.zeros((ny, nx), complex)
    e[0, 0] = 1
    assert np.abs(out - e).max() < 1e-15


# --- phase correction -------------------------------------------------------

def test_phase_single_plane_is_plain_copy():
    rng = np.random.default_rng(2)
    g = rng.normal(size=2 * 16 * 8)
    re, im = O.phase_correction(g, 16, 8, 1, 1 / 300.0, -5.0, 9.0, phase_on=True)
    assert np.array_equal(re, g[0::2]) and np.array_equal(im, g[1::2])


def test_phase_zero_resolution_is_weighted_plane_sum():
    # resolution 0 -> l = m = 0 -> n = 1 -> phase 0 -> image = dwnorm * sum_w g_w.
    rng = np.random.default_rng(4)
    planes, nx, ny = 5, 16, 12
    g = rng.normal(size=2 * planes * nx * ny)
    re, im = O.phase_correction(g, nx, ny, planes, 0.0, -40.0, 60.0, phase_on=True)
    c = as_complex(g, planes, ny, nx)
    s = c.sum(axis=0).reshape(-1) / planes            # dwnorm = 1/planes
    assert rel_l2(re + 1j * im, s) < 1e-15


def test_phase_is_a_rotation_per_plane():
    # One non-zero plane value: |image| = |g| * dwnorm * n, for every pixel.
    planes, nx, ny, res = 4, 16, 16, 1 / 40.0
    rng = np.random.default_rng(5)
    c = np.zeros((planes, ny, nx), complex)
    c[2] = rng.normal(size=(ny, nx)) + 1j * rng.normal(size=(ny, nx))
    re, im = O.phase_correction(from_complex(c), nx, ny, planes, res, 0.0, 800.0)
    img = (re + 1j * im).reshape(ny, nx)
    idx = np.arange(nx)
    sh = np.where(idx - nx // 2 < 0, idx + nx // 2, idx - nx // 2)
    lx = np.sin(sh * res)
    n = np.sqrt(1 - lx[None, :] ** 2 - lx[:, None] ** 2)
    np.testing.assert_allclose(np.abs(img), np.abs(c[2]) * 0.25 * n, rtol=1e-13)
    # plane 2 of wmin=0, dw=200: phase 2*pi*400*(n-1)
    ang = np.angle(img / c[2])
    expect = np.angle(np.exp(1j * 2 * np.pi * 400.0 * (n - 1)))
    assert np.abs(np.angle(np.exp(1j * (ang - expect)))).max() < 1e-9


# --- NORMALIZE_UVW -----------------------------------------------------------

def test_normalize_uvw_matches_numpy_restatement():
    from rick_amd import synth
    rng = np.random.default_rng(8)
    u, v, w = rng.normal(0, 900, 1000), rng.normal(0, 1200, 1000), rng.normal(0, 300, 1000)
    a = O.normalize_uvw(u, v, w)
    b = synth.normalize_uvw(u, v, w)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    uu, vv, ww = a
    assert uu.min() > 0 and uu.max() < 1 and vv.min() > 0 and vv.max() < 1
    assert ww.min() == 0.0 and ww.max() == 1.0


def test_normalize_uvw_closed_forms():
    # gridding.c:37-82 fixed by hand on small cases.  The scale is the
    # largest |extremum| of u and v, with the maxima folded from 0.0 and the
    # minima from 1e20, plus 0.1 %; u, v map to (x + maxg) / (2 maxg); w maps
    # its own extremes to exactly 0 and 1.
    # All-negative u and v: the maxima stay 0.0, so the scale comes from the
    # minima: maxg = 4 * 1.001, and u = -4 maps to 0.004 / 8.008 = 1/2002.
    u = np.array([-4.0, -2.0])
    v = np.array([-1.0, -3.0])
    w = np.array([5.0, 9.0])
    uu, vv, ww = O.normalize_uvw(u, v, w)
    assert abs(uu[0] - 1 / 2002) < 1e-15 and abs(uu[1] - 2.004 / 8.008) < 1e-15
    assert abs(vv[1] - 1.004 / 8.008) < 1e-15
    assert ww.tolist() == [0.0, 1.0]
    # A dominant positive maximum: maxg = 10 * 1.001; x = 10 -> 20.01 / 20.02.
    u = np.array([10.0, -1.0, 3.0])
    v = np.array([2.0, 0.5, -6.0])
    w = np.array([-2.0, 0.0, 2.0])
    uu, vv, ww = O.normalize_uvw(u, v, w)
    assert abs(uu[0] - 20.01 / 20.02) < 1e-15 and abs(vv[2] - 4.01 / 20.02) < 1e-15
    assert ww.tolist() == [0.0, 0.5, 1.0]
    # Mirror symmetry: x and -x land symmetric about 1/2.
    u = np.array([7.0, -7.0, 1.5, -1.5])
    uu, _, _ = O.normalize_uvw(u, u[::-1].copy(), np.array([0.0, 1.0, 2.0, 3.0]))
    assert abs((uu[0] - 0.5) + (uu[1] - 0.5)) < 1e-15 and abs((uu[2] - 0.5) + (uu[3] - 0.5)) < 1e-15
    # A minimum above the 1e20 start value is never taken: with every w
    # above 1e20, wmin stays 1e20.
    w = np.array([2e20, 3e20])
    _, _, ww = O.normalize_uvw(np.array([1.0, 2.0]), np.array([1.0, 2.0]), w)
    assert ww[0] == (2e20 - 1e20) / (3e20 - 1e20)
