"""GPU parity of the scene BVH (ray-shape intersection, calc_intersection
tracer.rs:306-335).  The BVH only changes which shapes a ray is tested
against: the nearest hit (distance and shape index, ties to the lower index)
must equal testing every shape one by one, bit for bit, and must agree with
the FP64 oracle except at near-ties."""
import numpy as np
import pytest

import oracle_lib as O
from gpu_common import ctx, say

import ppmpa_amd as P
from ppmpa_amd import _abi as A

pytestmark = pytest.mark.gpu


def _rays(n, seed):
    """Rays from inside the box in random directions, axis-aligned directions
    (slab planes parallel to the ray), and rays from outside the box."""
    rng = np.random.default_rng(seed)
    o = rng.uniform(0.02, 0.98, (n, 3))
    d = rng.normal(size=(n, 3))
    k = n // 8
    axes = np.array([(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)], float)
    d[:k] = axes[rng.integers(0, 6, k)]
    o[k:2 * k] = rng.uniform(-2.0, 3.0, (k, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return o, d


def _pair(sc):
    c = ctx()
    sc.bvh = A.SCENE_BVH_ON
    on = P.DeviceScene(c, sc)
    sc.bvh = A.SCENE_BVH_OFF
    off = P.DeviceScene(c, sc)
    sc.bvh = A.SCENE_BVH_AUTO
    return on, off


def test_bvh_layout():
    say("bvh layout")
    sc = P.mesh_scene(64, 48, subdiv=3, n_spheres=40)
    c = ctx()
    ds = P.DeviceScene(c, sc)
    n_lin, n_bvh, depth = ds.stats()
    # the five 1 m walls (wider than a quarter of the scene) and the light's
    # own panel are tested one by one; the spheres and the triangles are in
    # the BVH
    assert n_lin == 6 and n_bvh == sc.n_shapes - 6
    per_leaf = 4 if n_bvh <= 512 else 1  # small trees: 4 shapes per leaf, fitting the LDS copy whole
    assert (1 << depth) >= -(-n_bvh // per_leaf) > (1 << depth) // 2
    small = P.DeviceScene(c, P.caustic_scene(32, 24))
    assert small.stats() == (9, 0, 0)  # fewer than 64 bounded shapes: one by one
    sc2 = P.mesh_scene(32, 24, subdiv=1, n_spheres=8, walls="plane")
    sc2.bvh = A.SCENE_BVH_ON
    # planes stay one by one, and so does the 0.25 m light quad: the bounded
    # shapes of this scene span 0.92 m, and it is wider than a quarter of that
    assert P.DeviceScene(c, sc2).stats() == (6, 8 + 80, 5)
    # a small tree: 320 triangles and 64 spheres in it, the walls and the
    # light's own panel one by one
    sc3 = P.mesh_scene(32, 24, subdiv=2, n_spheres=64)
    assert P.DeviceScene(c, sc3).stats() == (6, 320 + 64, 7)


# subdivision 2: 320 triangles, 64 spheres and the light quad, a small tree
# walked from its LDS copies (triangles and the quad read from global memory
# beside the LDS spheres, leaf slots past the last shape padded)
@pytest.mark.parametrize("subdiv,walls", [(3, "quad"), (4, "plane"), (2, "quad")])
def test_bvh_equals_linear(subdiv, walls):
    say("bvh == linear", subdiv, walls)
    sc = P.mesh_scene(64, 48, subdiv=subdiv, n_spheres=64, walls=walls)
    on, off = _pair(sc)
    assert (on.stats()[1] <= 512) == (subdiv == 2)
    o, d = _rays(40000, subdiv)
    t1, i1 = on.intersect(o, d)
    t0, i0 = off.intersect(o, d)
    assert np.array_equal(i1, i0), np.flatnonzero(i1 != i0)[:10]
    assert np.array_equal(t1, t0)
    assert (i1 >= 0).mean() > 0.5


def test_bvh_against_oracle():
    say("bvh vs oracle")
    sc = P.mesh_scene(64, 48, subdiv=3, n_spheres=64)
    sc.bvh = A.SCENE_BVH_ON
    ds = P.DeviceScene(ctx(), sc)
    osc = O.OScene(sc)
    o, d = _rays(3000, 11)
    t, obj = ds.intersect(o, d)
    same = 0
    for i in range(len(o)):
        oi, pos, _ = osc.intersect(o[i], d[i])
        if oi == obj[i]:
            same += 1
            if oi >= 0:
                to = np.linalg.norm(pos - o[i])
                assert abs(t[i] - to) <= 2e-5 * max(1.0, to), (i, t[i], to)
        else:
            # a near-tie: both shapes are hit at (almost) the same distance
            assert oi >= 0 and obj[i] >= 0, (i, oi, obj[i])
            assert abs(t[i] - np.linalg.norm(pos - o[i])) <= 1e-4
    assert same >= 0.995 * len(o)


def test_bvh_full_mesh():
    """327,680 triangles + 64 spheres + the box: BVH == every shape tested."""
    say("bvh full mesh")
    sc = P.mesh_scene(64, 48)
    assert sc.n_shapes == 327680 + 64 + 6
    on, off = _pair(sc)
    n_lin, n_bvh, depth = on.stats()
    assert n_lin == 6 and n_bvh == sc.n_shapes - 6 and depth == 19  # one shape per leaf: 2^19 >= 327,744
    o, d = _rays(4096, 5)
    t1, i1 = on.intersect(o, d)
    t0, i0 = off.intersect(o, d)
    assert np.array_equal(i1, i0) and np.array_equal(t1, t0)
    # rays towards the icosphere hit a triangle
    dd = np.array([0.5, 0.3, 0.6]) - o[:512]
    dd /= np.linalg.norm(dd, axis=1, keepdims=True)
    t2, i2 = on.intersect(o[:512], dd)
    assert (i2 >= 70).mean() > 0.5


@pytest.mark.parametrize("subdiv", [3, 2])
def test_bvh_pass_identical(subdiv):
    """Whole passes (photon trace, eye rays, shadow rays) give the same image
    through the BVH as through the one-by-one test: a large tree read from
    global memory, and a small tree of triangles, spheres and the light quad
    walked from its LDS copies."""
    say("bvh pass identical", subdiv)
    sc = P.mesh_scene(48, 32, subdiv=subdiv, n_spheres=24)
    on, off = _pair(sc)
    c = ctx()
    cfg = P.pass_cfg(seed=3, n_photon=30000, radius0=0.02, alpha=2.0 / 3.0, max_depth=8)
    img = []
    for ds in (on, off):
        s = P.Session(c, ds, cfg)
        s.run(2)
        m, _, k = s.image()
        s.close()
        assert k == 2
        img.append(m)
    assert np.array_equal(img[0], img[1])
    assert img[0].mean() > 0


def test_bvh_photon_trace_crn():
    """Photon walks through the BVH against the FP64 oracle (which tests every
    shape) on common random numbers."""
    say("bvh photon trace crn")
    sc = P.mesh_scene(32, 24, subdiv=3, n_spheres=32)
    sc.bvh = A.SCENE_BVH_ON
    c = ctx()
    ds = P.DeviceScene(c, sc)
    n = 8000
    ph = P.trace_photons(c, ds, seed=5, pass_=2, n_photon=n, use_classic=True, max_depth=8)
    gpos, gdir, gwl, gmeta = ph.download()
    opos, odir, owl, ometa, _ = O.OScene(sc).trace_photons(5, 2, n, True, 8)
    gk = (gmeta >> 7).astype(np.int64) * 64 + ((gmeta >> 2) & 31)
    ok = (ometa >> 7).astype(np.int64) * 64 + ((ometa >> 2) & 31)
    common, gi, oi = np.intersect1d(gk, ok, return_indices=True)
    assert len(common) >= 0.98 * len(ok), (len(common), len(ok), len(gk))
    dpos = np.abs(gpos[gi] - opos[oi]).max(axis=1)
    assert np.quantile(dpos, 0.98) < 1e-4
