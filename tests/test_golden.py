"""The committed golden fixtures (tests/golden/) against the oracle: the
photon map in the reference's text format parses back exactly, and the
oracle reproduces the committed index sets and radiances.  CPU only."""
import os

import numpy as np
import pytest

import oracle_lib as O
from ppmpa_amd import photonmap_io as M

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def gold():
    m = M.parse_map(os.path.join(GOLD, "map_two_spheres.txt"))
    z = np.load(os.path.join(GOLD, "queries_two_spheres.npz"), allow_pickle=False)
    return m, {k: z[k] for k in z.files}


def test_map_text_round_trip(gold):
    m, _ = gold
    assert m.n_emitted == 4000 and m.power == 0.025
    assert len(m.wl) > 3000
    text = open(os.path.join(GOLD, "map_two_spheres.txt")).read()
    assert M.format_map(m.n_emitted, m.power, m.pos, m.dir, m.wl) == text


def test_oracle_reproduces_index_sets(gold):
    m, z = gold
    for j, r2 in enumerate(z["r2"]):
        sets = O.within_brute(m.pos, z["q"], r2, mode=0)
        off, idx = z[f"off_{j}"], z[f"idx_{j}"]
        for i, s in enumerate(sets):
            assert np.array_equal(s, idx[int(off[i]):int(off[i + 1])])
        # the kd-tree path returns the same sets
        om = O.OMap(m.pos, m.dir, m.wl, m.power)
        for i in range(0, len(z["q"]), 7):
            got, _ = om.within(z["q"][i], r2)
            assert sorted(got.tolist()) == sets[i].tolist()


def test_oracle_reproduces_radiance(gold):
    m, z = gold
    dn = m.dir / np.linalg.norm(m.dir, axis=1, keepdims=True)
    om = O.OMap(m.pos, dn, m.wl, m.power)
    for j, r2 in enumerate(z["r2"]):
        for pf in (0, 1, 2):
            assert np.array_equal(om.gather(z["q"], z["nrm"], r2, pf), z[f"rad_{j}_{pf}"])
