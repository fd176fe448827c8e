"""binMS reader/writer, row partition, image and cube writers (host paths, no
GPU), and the oracle's sector binning / gather against hand-derived cases.

The reference's own tests hold no binMS fixture; meta.txt is pinned by a file
written the way scripts/create_binMS.py:225-245 writes it (Python str() of
each value, one per line, 15 lines), the binary layout by numpy reads at the
byte offsets of init.c:424-490.
"""
import os

import numpy as np
import pytest

from oracle import oracle as O
from rick_amd import binms
from rick_amd._lib import RICK_ERR_INVALID, RickError

CREATE_BINMS_META = [7745536, 123928576, 16, 1, 4096, 10.013, 10.013 * 4096 / 3600, 1891,
                     -50123.456789, 50234.5, -1234.25, 987.0625, 7745536, 1, "1.0"]


def write_meta_like_converter(path, values):
    os.makedirs(path, exist_ok=True)
    with open(os.path.join(path, "meta.txt"), "w") as f:
        for v in values:
            f.write(str(v) + "\n")


def test_meta_parses_converter_output(tmp_path):
    write_meta_like_converter(tmp_path, CREATE_BINMS_META)
    m = binms.read_meta(tmp_path)
    v = CREATE_BINMS_META
    assert (m.nmeasures, m.nvis, m.freq_per_chan, m.polarisations, m.ntimes) == tuple(v[:5])
    assert m.dt == v[5] and m.thours == v[6] and m.baselines == v[7]
    assert (m.uvmin, m.uvmax, m.wmin, m.wmax) == tuple(v[8:12])
    assert m.resolution == 1.0 / max(abs(v[8]), abs(v[9]))   # gridding_cpu.c:78


def test_meta_errors(tmp_path):
    with pytest.raises(RickError) as e:
        binms.read_meta(tmp_path / "missing")
    assert e.value.status == RICK_ERR_INVALID
    write_meta_like_converter(tmp_path, CREATE_BINMS_META[:7])   # short file
    with pytest.raises(RickError) as e:
        binms.read_meta(tmp_path)
    assert "7 of 12" in str(e.value)


@pytest.mark.parametrize("nmeasures,size", [(1000 + 10_000, 3), (1000 + 7, 1), (1000 + 5, 8),
                                            (1000, 2), (4_000_000_000, 7)])
def test_partition(nmeasures, size):
    meta = binms.Meta(nmeasures, nmeasures, 1, 1)
    total = nmeasures - 1000
    parts = [binms.partition(meta, r, size) for r in range(size)]
    assert parts == [O.partition(nmeasures, r, size) for r in range(size)]
    # Contiguous, covering rows [0, total) with the remainder on the last rank.
    assert parts[0][0] == 0
    for (s0, n0), (s1, _) in zip(parts, parts[1:]):
        assert s0 + n0 == s1
    assert parts[-1][0] + parts[-1][1] == total
    assert all(n == total // size for _, n in parts[:-1])


def test_partition_errors():
    meta = binms.Meta(999, 999, 1, 1)
    with pytest.raises(RickError):
        binms.partition(meta, 0, 1)              # fewer rows than the 1000 dropped
    meta = binms.Meta(5000, 5000, 1, 1)
    with pytest.raises(RickError):
        binms.partition(meta, 2, 2)
    with pytest.raises(RickError):
        binms.partition(meta, 0, 0)
    assert binms.partition(meta, 0, 1, drop=0) == (0, 5000)


def random_ms(n, F, P, seed=0):
    rng = np.random.default_rng(seed)
    return {"uu": rng.normal(0, 1e4, n), "vv": rng.normal(0, 1e4, n), "ww": rng.normal(0, 3e2, n),
            "weight": rng.uniform(0.5, 2, n * P).astype(np.float32),
            "vis_real": rng.normal(0, 1, n * F * P).astype(np.float32),
            "vis_img": rng.normal(0, 1, n * F * P).astype(np.float32)}


def test_binms_roundtrip_and_reference_offsets(tmp_path):
    n, F, P = 3001, 3, 2
    a = random_ms(n, F, P)
    meta = binms.write(tmp_path, a["uu"], a["vv"], a["ww"], a["weight"], a["vis_real"],
                       a["vis_img"], F, P, ntimes=3, dt=8.0, baselines=1000, uvmin=-1.5,
                       uvmax=2.25, wmin=-0.5, wmax=0.75)
    # File sizes of the converter's layout.
    for name, per in zip(binms.FILES, (8, 8, 8, 4 * P, 4 * F * P, 4 * F * P)):
        assert os.path.getsize(tmp_path / name) == n * per
    lines = (tmp_path / "meta.txt").read_text().split()
    assert len(lines) == 15 and lines[12:] == [str(n), str(P), "1.0"]
    m = binms.read_meta(tmp_path)
    assert m == meta and m.nvis == n * F * P and m.thours == 8.0 * 3 / 3600
    start, rows = binms.partition(m, 1, 2)        # second of two ranks
    assert (start, rows) == (1000, 1001)
    got = binms.read(tmp_path, m, start, rows)
    want = O.read_binms(tmp_path, F, P, start, rows)
    for k in want:
        np.testing.assert_array_equal(got[k], want[k])
    np.testing.assert_array_equal(got["vis_img"],
                                  a["vis_img"][start * F * P:(start + rows) * F * P])


def test_binms_short_file_is_an_error(tmp_path):
    n, F, P = 100, 2, 1
    a = random_ms(n, F, P)
    m = binms.write(tmp_path, *(a[k] for k in ("uu", "vv", "ww", "weight", "vis_real",
                                               "vis_img")), F, P)
    with open(tmp_path / "visibilities_img.bin", "r+b") as f:
        f.truncate(n * F * P * 4 - 4)
    with pytest.raises(RickError) as e:
        binms.read(tmp_path, m, 0, n)
    assert "short read" in str(e.value)
    got = binms.read(tmp_path, m, 0, n - 1)          # rows that are there still read
    assert got["uu"].size == n - 1
    assert binms.read(tmp_path, m, n, 0)["vis_img"].size == 0   # no rows: nothing read


def test_write_image_offsets(tmp_path):
    x, y = 12, 5
    re = np.arange(x * y, dtype=np.float64)
    im = -re
    binms.write_image(tmp_path / "r.bin", tmp_path / "i.bin", re, im, x, y, 0)
    np.testing.assert_array_equal(np.fromfile(tmp_path / "r.bin"), re)
    np.testing.assert_array_equal(np.fromfile(tmp_path / "i.bin"), im)
    # rank 2: the file is recreated and the slab lands at byte 2*x*y*8
    # (fourier_transform.c:282-286), with a hole of zeros before it.
    binms.write_image(tmp_path / "r.bin", tmp_path / "i.bin", re, im, x, y, 2)
    r = np.fromfile(tmp_path / "r.bin")
    assert r.size == 3 * x * y
    np.testing.assert_array_equal(r[:2 * x * y], 0.0)
    np.testing.assert_array_equal(r[2 * x * y:], re)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_write_cube_host(tmp_path, dtype):
    planes, nx, ny = 3, 8, 4
    g = np.random.default_rng(1).normal(size=2 * planes * nx * ny).astype(dtype)
    binms.write_cube(tmp_path / "r.bin", tmp_path / "i.bin", g, planes, nx, ny)
    np.testing.assert_array_equal(np.fromfile(tmp_path / "r.bin"), g[0::2].astype(np.float64))
    np.testing.assert_array_equal(np.fromfile(tmp_path / "i.bin"), g[1::2].astype(np.float64))


# --- oracle: sector binning and gather ----------------------------------------

def test_oracle_sector_bin_hand_case():
    # 4 slabs of 16 rows at dx = 1/64 (slab height 0.25), support 7: rows
    # within 3/64 of a border are also listed in the neighbouring slab.
    vv = np.array([0.26, 0.49, 0.1, 0.0, 0.99, 0.25, 0.5])
    histo, start, rows, bad = O.sector_bin(vv, 4, 16, 1 / 64, 7)
    lists = [rows[start[s]:start[s] + histo[s]].tolist() for s in range(5)]
    assert lists == [[0, 2, 3, 5],      # 0.26 and 0.25 reach down into slab 0
                     [0, 1, 5, 6],      # 0.5 (lower border of slab 2) reaches down
                     [1, 6],            # 0.49 reaches up
                     [4],
                     [4]]               # 0.99: the extra slot above the last slab
    assert bad == 0


def test_oracle_sector_bin_properties():
    rng = np.random.default_rng(5)
    n, S, yaxis, dx, sup = 20000, 5, 100, 1 / 500, 9
    vv = rng.uniform(0, 1, n)
    histo, start, rows, bad = O.sector_bin(vv, S, yaxis, dx, sup)
    h = (sup - 1) // 2 * dx
    assert bad == 0 and histo.sum() == rows.size
    seen_main = np.zeros(n, bool)
    for s in range(S + 1):
        lst = rows[start[s]:start[s] + histo[s]]
        assert np.all(np.diff(lst.astype(np.int64)) > 0)     # increasing row order
        main = (vv[lst] * S).astype(int) == s
        seen_main[lst[main]] = True
        # Halo entries lie within h of a border of slab s.
        lo, hi = s * yaxis * dx, (s + 1) * yaxis * dx
        d = np.minimum(np.abs(vv[lst[~main]] - lo), np.abs(vv[lst[~main]] - hi))
        assert np.all(d < h)
    assert seen_main.all()
    # One slab: every row in slab 0 in order, rows near the top also in slot 1.
    histo, start, rows, _ = O.sector_bin(vv, 1, 500, dx, sup)
    np.testing.assert_array_equal(rows[:histo[0]], np.arange(n))
    np.testing.assert_array_equal(rows[histo[0]:], np.nonzero(1.0 - vv < h)[0])


def test_oracle_sector_bin_out_of_range_rows():
    vv = np.array([0.5, -2.0, 1.5, np.nan, 0.999])
    histo, start, rows, bad = O.sector_bin(vv, 2, 10, 1 / 20, 7)
    assert bad == 3      # three rows outside slots 0..2
    assert 1 not in rows and 2 not in rows and 3 not in rows


def test_oracle_sector_gather_reverses_and_shifts():
    rng = np.random.default_rng(2)
    n, F, P = 50, 3, 2
    a = random_ms(n, F, P, seed=2)
    lst = np.sort(rng.choice(n, 17, replace=False)).astype(np.uint32)
    g = O.sector_gather(lst, 3, 16, 1 / 64, F, P, a["uu"], a["vv"], a["ww"], a["weight"],
                        a["vis_real"], a["vis_img"])
    rev = lst[::-1].astype(np.int64)
    np.testing.assert_array_equal(g["uu"], a["uu"][rev])
    np.testing.assert_array_equal(g["vv"], a["vv"][rev] - 3 * (1 / 64 * 16))
    np.testing.assert_array_equal(g["weight"], a["weight"].reshape(n, P)[rev].ravel())
    np.testing.assert_array_equal(g["vis_img"], a["vis_img"].reshape(n, F * P)[rev].ravel())
