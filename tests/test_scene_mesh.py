"""The mesh workload's scene generator and the oracle's nearest hit over many
triangles and spheres.  CPU only.

The oracle tests every shape (as calc_intersection, tracer.rs:306-335, does);
here it is checked against an independent numpy solution: spheres by the
quadratic formula, triangles by solving o + t d = p0 + u e1 + v e2 as a 3x3
linear system (not the Moller-Trumbore form the oracle uses)."""
import numpy as np
import pytest

import oracle_lib as O

import ppmpa_amd as P
from ppmpa_amd import _abi as A
from ppmpa_amd import scene as S


@pytest.mark.parametrize("k", [0, 1, 2, 4])
def test_icosphere_topology(k):
    v, f = S.icosphere(k)
    assert len(f) == 20 * 4 ** k
    e = np.sort(np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]]), axis=1)
    u, cnt = np.unique(e, axis=0, return_counts=True)
    assert np.all(cnt == 2)                      # closed, manifold
    assert len(v) - len(u) + len(f) == 2         # Euler characteristic of a sphere
    assert np.allclose(np.linalg.norm(v, axis=1), 1.0)
    t = v[f]
    n = np.cross(t[:, 1] - t[:, 0], t[:, 2] - t[:, 0])
    assert np.all(np.einsum("ij,ij->i", n, t.mean(axis=1)) > 0)  # outward winding


def test_icosphere_noise_seeded():
    c = np.array([0.5, 0.3, 0.6])
    v1, f1 = S.icosphere(3, c, 0.22, noise=0.02, seed=1)
    v2, _ = S.icosphere(3, c, 0.22, noise=0.02, seed=1)
    v3, _ = S.icosphere(3, c, 0.22, noise=0.02, seed=2)
    assert np.array_equal(v1, v2) and not np.array_equal(v1, v3)
    r = np.linalg.norm(v1 - c, axis=1) / 0.22
    assert r.min() >= 0.98 and r.max() <= 1.02 and r.std() > 0.005


def test_mesh_scene_layout():
    sc = S.mesh_scene(64, 48, subdiv=3, n_spheres=64)
    assert sc.n_shapes == 6 + 64 + 1280
    sph = [s for s in sc.shapes if s.kind == A.SHAPE_SPHERE]
    assert len(sph) == 64
    c = np.array([s.pos for s in sph])
    r = np.array([s.radius for s in sph])
    assert np.all((r >= 0.02) & (r <= 0.08))
    assert np.all(c - r[:, None] > 0) and np.all(c + r[:, None] < 1)
    d = np.linalg.norm(c[:, None] - c[None], axis=2) + np.eye(64) * 9
    assert np.all(d > r[:, None] + r[None])       # no two spheres overlap
    assert np.all(np.linalg.norm(c - [0.5, 0.3, 0.6], axis=1) > 0.22 * 1.02 + r)
    mats = {sc.materials[s.material].metalness for s in sph} | {sc.materials[s.material].ior[0] for s in sph}
    assert 1.0 in mats and 1.5 in mats            # mirrors and glass both drawn
    arr = sc.shape_array()
    tri = arr[arr["kind"] == A.SHAPE_POLYGON]
    assert len(tri) == 1280
    cr = np.cross(tri["dir1"], tri["dir2"])
    assert np.allclose(cr / np.linalg.norm(cr, axis=1, keepdims=True), tri["nvec"])
    assert np.allclose(np.linalg.norm(tri["nvec"], axis=1), 1.0)


def _numpy_nearest(sc, o, d):
    """Nearest hit with t >= 1e-4 by closed forms (spheres, walls) and a
    linear solve per triangle."""
    arr = sc.shape_array()
    best_t, best_i = np.inf, -1
    for i, s in enumerate(arr):
        ts = []
        if s["kind"] == A.SHAPE_SPHERE:
            oc = o - s["pos"]
            b = oc @ d
            disc = b * b - (oc @ oc - s["radius"] ** 2)
            if disc > 0:
                ts = [-b - np.sqrt(disc), -b + np.sqrt(disc)]
        elif s["kind"] in (A.SHAPE_POLYGON, A.SHAPE_PARALLELOGRAM):
            M = np.column_stack([d, -s["dir1"], -s["dir2"]])
            if abs(np.linalg.det(M)) > 1e-14:
                t, u, v = np.linalg.solve(M, s["pos"] - o)
                lim = 1.0 if s["kind"] == A.SHAPE_POLYGON else 2.0
                if u >= 0 and v >= 0 and u <= 1 and v <= 1 and u + v <= lim:
                    ts = [t]
        for t in ts:
            if 1e-4 <= t < best_t:
                best_t, best_i = t, i
    return best_t, best_i


def test_oracle_nearest_over_mesh():
    sc = S.mesh_scene(16, 16, subdiv=2, n_spheres=12)
    osc = O.OScene(sc)
    rng = np.random.default_rng(4)
    n_hit = 0
    for _ in range(150):
        o = rng.uniform(0.05, 0.95, 3)
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        oi, pos, _ = osc.intersect(o, d)
        t, i = _numpy_nearest(sc, o, d)
        assert oi == i
        if i >= 0:
            n_hit += 1
            assert abs(np.linalg.norm(pos - o) - t) < 1e-9
    assert n_hit > 100  # the rest leave through the open front
