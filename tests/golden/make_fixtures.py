"""Write the golden fixtures of the photon-map parity tests.

The reference publishes no vectors and cannot be built here (Rust, no
cargo), so the photon set is produced by the FP64 oracle's restatement of
trace_photon and written in the reference's own pm text format
(pm.rs:44-45,62-75), which read_map parses (photonmap.rs:31-74).  Expected
outputs: the FP64 ball-query index sets (brute force, the reference's
predicate) and the FP64 radiance estimates (estimate_radiance, all three
filters).

Run from the repository root:  python tests/golden/make_fixtures.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle_lib as O  # noqa: E402
import ppmpa_amd as P  # noqa: E402
from ppmpa_amd import photonmap_io as M  # noqa: E402

N_PHOTON = 4000
R2S = (0.01 ** 2, 0.03 ** 2)


def main():
    sc = P.two_spheres_scene(32, 32)
    osc = O.OScene(sc)
    pos, dir, wl, meta, power = osc.trace_photons(seed=1, pass_=1, n_photon=N_PHOTON, uc=True, maxd=8)
    text = M.format_map(N_PHOTON, power, pos, dir, wl)
    with open(os.path.join(HERE, "map_two_spheres.txt"), "w") as f:
        f.write(text)
    m = M.parse_map(text)  # what read_map sees (exact round trip of the doubles)
    rng = np.random.default_rng(20261202)
    k = 300
    near = m.pos[rng.integers(0, len(m.wl), k)] + rng.normal(0, 0.004, (k, 3))
    q = np.concatenate([near, m.pos[rng.integers(0, len(m.wl), 40)], rng.uniform(0, 1, (60, 3))])
    nrm = rng.normal(size=q.shape)
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    dn = m.dir / np.linalg.norm(m.dir, axis=1, keepdims=True)  # Ray::new_from_elem normalises
    om = O.OMap(m.pos, dn, m.wl, m.power)
    out = {"q": q, "nrm": nrm, "r2": np.array(R2S)}
    for j, r2 in enumerate(R2S):
        sets = O.within_brute(m.pos, q, r2, mode=0)
        out[f"idx_{j}"] = np.concatenate(sets).astype(np.uint32) if sets else np.zeros(0, np.uint32)
        out[f"off_{j}"] = np.cumsum([0] + [len(s) for s in sets]).astype(np.uint64)
        for pf in (0, 1, 2):
            out[f"rad_{j}_{pf}"] = om.gather(q, nrm, r2, pf)
    np.savez(os.path.join(HERE, "queries_two_spheres.npz"), **out)
    print(f"{len(m.wl)} photons, {len(q)} queries")


if __name__ == "__main__":
    main()
