"""GPU: batched passes over many scenes.  A batch runs every pass stage as one
launch over all its scenes; each scene's blocks run the single-scene kernel
bodies on that scene's buffers, so every scene's image must equal the image
of its own session with the same schedule and the same number of pass
pipelines (a batch runs round r on pipeline r mod K, as a session runs pass
r), bit for bit except in pixels whose eye tree overflowed its 2 * max_depth
query slots (their extra queries are added with float atomics, in no fixed
order, in a session as well)."""
import numpy as np
import pytest

from gpu_common import ctx, say

import ppmpa_amd as P

pytestmark = pytest.mark.gpu


def _small_batch(n, seed):
    out = []
    for sc, cfg, passes in P.batch_scenes(n=n, seed=seed, max_side=48, min_side=24):
        cfg = dict(cfg, n_photon=max(2000, cfg["n_photon"] // 40))
        out.append((sc, cfg, max(2, passes // 32)))
    return out


def _same(a, b):
    return np.allclose(a, b, rtol=1e-6, atol=0.0) and (a == b).mean() >= 0.99


def _cfg(c, pipelines=2):
    return P.pass_cfg(seed=c["seed"], n_photon=c["n_photon"], radius0=c["radius0"], alpha=c["alpha"],
                      max_depth=c["max_depth"], pipelines=pipelines)


def _plain_images(c, items, dss, pipelines=2, tile_width=8):
    imgs = []
    for (sc, cfg, passes), ds in zip(items, dss):
        s = P.Session(c, ds, _cfg(cfg, pipelines))
        s.set_tile_width(tile_width)
        s.run(passes)
        m, v, k = s.image()
        s.close()
        assert k == passes
        imgs.append((m, v))
    return imgs


@pytest.mark.parametrize("seed,pipelines", [(3, 2), (11, 1), (7, 3)])
def test_batch_equals_sessions(seed, pipelines):
    say("batch == sessions", seed, pipelines)
    c = ctx()
    items = _small_batch(10, seed)
    dss = [P.DeviceScene(c, sc) for sc, _, _ in items]
    kinds = {ds.stats()[1] > 0 for ds in dss}
    plain = _plain_images(c, items, dss, pipelines)
    b = P.Batch(c, dss, [_cfg(cfg, pipelines) for _, cfg, _ in items], [p for _, _, p in items])
    b.run()
    for i, (sc, cfg, passes) in enumerate(items):
        m, v, k = b.image(i)
        assert k == passes
        assert _same(m, plain[i][0]), i
        assert np.allclose(v, plain[i][1], rtol=1e-5, atol=1e-12), i
    t = b.timings()
    assert t["photons_emitted"] > 0 and t["ms_total"] > 0
    b.close()
    print("scenes with a BVH:", kinds)


def test_batch_mixed_bvh_and_tuning():
    """A batch holding scenes with and without a BVH; the autotuner's choice
    of launch sizes changes the speed, not the images."""
    say("batch tune")
    c = ctx()
    items = _small_batch(12, 5)
    # make sure both layouts are present: one room with many spheres
    big = P.batch_scenes(n=40, seed=5, max_side=40, min_side=32)
    many = [x for x in big if len(x[0].shapes) - 6 >= 80][0]
    items.append((many[0], dict(many[1], n_photon=20000), 3))
    dss = [P.DeviceScene(c, sc) for sc, _, _ in items]
    assert any(ds.stats()[1] > 0 for ds in dss) and any(ds.stats()[1] == 0 for ds in dss)
    b = P.Batch(c, dss, [_cfg(cfg) for _, cfg, _ in items], [p for _, _, p in items])
    best = b.tune(rounds=1)
    assert best["eye_waves_per_cu"] in (4, 8, 16, 32) and best["shade_blocks_per_cu"] in (2, 4, 8, 16)
    assert best["tile_width"] in (4, 8, 16, 32) and b.launch() == best
    # sessions with the tuned tile shape give the batch's images bit for bit
    plain = _plain_images(c, items, dss, tile_width=best["tile_width"])
    b.run()
    for i, (_, _, passes) in enumerate(items):
        m, _, k = b.image(i, with_var=False)
        assert k == passes and _same(m, plain[i][0]), i
    # partial runs: rounds continue where they stopped
    b2 = P.Batch(c, dss, [_cfg(cfg) for _, cfg, _ in items], [p for _, _, p in items])
    b2.set_launch(eye_waves_per_cu=4, shade_blocks_per_cu=2, tile_width=best["tile_width"])
    b2.run(max_rounds=1)
    b2.run()
    for i in range(len(items)):
        m, _, _ = b2.image(i, with_var=False)
        assert _same(m, plain[i][0]), i
    b.close()
    b2.close()


def test_fine_query_keys_match_batch():
    """A session of at least 2^20 pixels sorts its queries on fine (10-bit per
    axis) Morton keys, a batch on coarse (7-bit) ones whatever the size; the
    order is only the gather's, so the images agree bit for bit (outside
    overflowed pixels)."""
    say("fine query keys: session == batch")
    c = ctx()
    sc = P.caustic_scene(1024, 1024)
    cfg = dict(seed=5, n_photon=50000, radius0=0.02, alpha=2 / 3, max_depth=6)
    ds = P.DeviceScene(c, sc)
    s = P.Session(c, ds, _cfg(cfg))
    s.run(2)
    m0, v0, k0 = s.image()
    s.close()
    b = P.Batch(c, [ds], [_cfg(cfg)], [2])
    b.run()
    m1, v1, k1 = b.image(0)
    b.close()
    assert k0 == k1 == 2
    assert m0.mean() > 0
    assert _same(m1, m0)
    assert np.allclose(v1, v0, rtol=1e-5, atol=1e-12)


def test_batch_rooms_match_oracle():
    """Three rooms of the batched workload at reduced size (their 32-38 px
    images, 12k photons per pass, 3 passes), rendered through ppm_batch_*:
    each room's image against the FP64 oracle's passes on the same Philox
    streams, within the session bands.  Rooms 0 (6 spheres, 4 of glass),
    6 (101 spheres, a BVH) and 10 (34 spheres)."""
    import oracle_lib as O
    from gpu_common import assert_crn
    say("batch rooms vs oracle")
    c = ctx()
    all_items = P.batch_scenes(n=64, seed=1, max_side=40, min_side=32)
    items = [(all_items[i][0], dict(all_items[i][1], n_photon=12000), 3) for i in (0, 6, 10)]
    dss = [P.DeviceScene(c, sc) for sc, _, _ in items]
    assert dss[1].stats()[1] > 0 and dss[0].stats()[1] == 0
    b = P.Batch(c, dss, [_cfg(cfg) for _, cfg, _ in items], [p for _, _, p in items])
    b.run()
    for i, (sc, cfg, passes) in enumerate(items):
        m, _, k = b.image(i, with_var=False)
        assert k == passes
        omean, _, _, _, _ = O.OScene(sc).run_passes(_cfg(cfg), passes)
        assert omean.max() > 0
        # the rooms' spheres are a third mirrors and a third glass
        assert_crn(m, omean, "session_specular", f"batch room {(0, 6, 10)[i]}")
    b.close()


def test_batch_full_size_rooms_equal_sessions():
    """All 64 rooms of the batched workload at their full sizes (images of
    128-512 px, 50k-2M photons per pass; passes capped at 24 per room): the
    autotuned batch gives every room its own session's image."""
    say("batch full size == sessions")
    c = ctx()
    items = [(sc, cfg, min(p, 24)) for sc, cfg, p in P.batch_scenes(n=64, seed=1)]
    dss = [P.DeviceScene(c, sc) for sc, _, _ in items]
    b = P.Batch(c, dss, [_cfg(cfg) for _, cfg, _ in items], [p for _, _, p in items])
    tw = b.tune(rounds=1)["tile_width"]
    b.run()
    for i, ((sc, cfg, passes), ds) in enumerate(zip(items, dss)):
        m, _, k = b.image(i, with_var=False)
        assert k == passes
        s = P.Session(c, ds, _cfg(cfg))
        s.set_tile_width(tw)
        s.run(passes)
        m0, _, _ = s.image(with_var=False)
        s.close()
        assert m0.mean() > 0
        assert _same(m, m0), i
    b.close()


def test_tile_widths():
    """The eye's tile shape (4x16, 8x8, 16x4, 32x2 pixels per wave) changes
    the order in which a pixel's eye-tree terms are summed, nothing else:
    images agree with the 8x8 one to float rounding (the per-pixel sums of a
    few terms, 1e-6 relative), a session and a batch at the same width bit
    for bit, and switching the width between passes keeps the run going."""
    say("tile widths")
    c = ctx()
    items = _small_batch(6, 21)
    dss = [P.DeviceScene(c, sc) for sc, _, _ in items]
    ref = _plain_images(c, items, dss, tile_width=8)
    for tw in (4, 16, 32):
        got = _plain_images(c, items, dss, tile_width=tw)
        for i in range(len(items)):
            assert np.allclose(got[i][0], ref[i][0], rtol=1e-6, atol=1e-12), (tw, i)
        b = P.Batch(c, dss, [_cfg(cfg) for _, cfg, _ in items], [p for _, _, p in items])
        b.set_launch(tile_width=tw)
        assert b.launch()["tile_width"] == tw
        b.run()
        for i in range(len(items)):
            m, _, _ = b.image(i, with_var=False)
            assert _same(m, got[i][0]), (tw, i)
        b.close()
    # a width change part-way: the same passes as at 8 throughout, to rounding
    sc, cfg, passes = items[0]
    s = P.Session(c, dss[0], _cfg(cfg))
    s.run(1)
    s.set_tile_width(32)
    s.run(passes - 1)
    m, _, k = s.image(with_var=False)
    s.close()
    assert k == passes and np.allclose(m, ref[0][0], rtol=1e-6, atol=1e-12)
    with pytest.raises(Exception):
        s2 = P.Session(c, dss[0], _cfg(cfg))
        try:
            s2.set_tile_width(12)
        finally:
            s2.close()
