"""Tuning cache of the gridder's launch shape, on the CPU: keys, lookups of
gridding and image-chain entries, and the stored file's round trip."""
import json

from rick_amd import FP32, FP64, autotune, synth


def test_signature_separates_precision_rows_and_image_chain():
    a = autotune.signature(4096, 32, 7, FP32, 7_745_536)
    assert a == "g4096_p32_s7_fp32_r23"
    assert autotune.signature(4096, 32, 7, FP64, 7_745_536) == "g4096_p32_s7_fp64_r23"
    # Row counts are bucketed by their bit length.
    assert autotune.signature(4096, 32, 7, FP32, 1 << 22) == a
    assert autotune.signature(4096, 32, 7, FP32, (1 << 23) - 1) == a
    assert autotune.signature(4096, 32, 7, FP32, 1 << 23) != a
    assert autotune.signature(4096, 32, 7, FP32, 7_745_536, "image") == a + "_img"


def test_lookup_keeps_gridding_and_image_entries_apart(tmp_path):
    path = tmp_path / "tuning.json"
    grid_key = autotune.signature(1024, 8, 7, FP32, 100_000)
    autotune.store(grid_key, {"tile": [64, 64], "rows_per_task": 16384, "unroll": 9,
                              "ms": 1.0}, path)
    assert autotune.lookup(1024, 8, 7, FP32, 100_000, path, kind="image") is None
    img_key = autotune.signature(1024, 8, 7, FP32, 100_000, "image")
    autotune.store(img_key, {"tile": [128, 32], "rows_per_task": 4096, "ms": 2.0}, path)
    assert autotune.lookup(1024, 8, 7, FP32, 100_000, path) == ((64, 64), 16384, 9)
    # Entries without an unroll give 0 (the smallest instantiated one).
    assert autotune.lookup(1024, 8, 7, FP32, 100_000, path, kind="image") == ((128, 32), 4096, 0)
    assert set(json.loads(path.read_text())) == {grid_key, img_key}
    assert autotune.lookup(1024, 8, 7, FP64, 100_000, path) is None


def test_shipped_cache_has_image_entries_for_the_lofar_like_workload():
    cfg = synth.CONFIGS["lofar"]
    for prec in (FP32, FP64):
        e = autotune.lookup(cfg.grid, cfg.planes, cfg.support, prec, cfg.n_rows, kind="image")
        assert e is not None
        tile, rpt, unroll = e
        assert tile[0] >= cfg.support and tile[1] >= cfg.support
        assert tuple(tile) in autotune.TILES and rpt in autotune.ROWS_PER_TASK
        assert unroll == 0 or unroll in autotune.unrolls_for(cfg.support)


def test_unrolls_cover_the_support():
    assert autotune.unrolls_for(7) == [7, 9, 11, 15]
    assert autotune.unrolls_for(9) == [9, 11, 15]
    assert autotune.unrolls_for(15) == [15]
