"""Full Monte Carlo images: the device PPM-PA passes and the FP64 oracle's
passes, run on INDEPENDENT random streams, agree within their 3-sigma band.

z = (mean_gpu - mean_oracle) / sqrt(s2_gpu / K + s2_oracle / K), with the
per-pass sample variances s2.  Per pixel the pass values are heavy-tailed at
these photon counts, so the band is checked on 4x4-pixel block means (closer
to normal) and the per-pixel fraction is reported."""
import numpy as np
import pytest

import oracle_lib as O
from gpu_common import ctx, say

import ppmpa_amd as P

pytestmark = pytest.mark.gpu


def _block(img, b):
    h, w, _ = img.shape
    return img[:h - h % b, :w - w % b].reshape(h // b, b, w // b, b, 3).mean(axis=(1, 3))


def test_images_agree_within_3_sigma():
    say("3 sigma")
    sc = P.two_spheres_scene(64, 64)
    K = 24
    c = ctx()
    ds = P.DeviceScene(c, sc)
    cfg_g = P.pass_cfg(seed=101, n_photon=20000, radius0=0.02, alpha=2 / 3, max_depth=8)
    cfg_o = P.pass_cfg(seed=202, n_photon=20000, radius0=0.02, alpha=2 / 3, max_depth=8)
    s = P.Session(c, ds, cfg_g)
    # per-pass images for block statistics
    imgs_g = []
    prev = None
    for k in range(K):
        s.run(1)
        mean, _, n = s.image(with_var=False)
        cur = mean * n
        imgs_g.append(cur - (prev if prev is not None else 0))
        prev = cur
    s.close()
    osc = O.OScene(sc)
    imgs_o = []
    for k in range(K):
        c1 = P.pass_cfg(seed=202, n_photon=20000, radius0=0.02, alpha=2 / 3, max_depth=8, first_pass=k + 1)
        m1, _, _, _, _ = osc.run_passes(c1, 1)
        imgs_o.append(m1)
    G = np.stack([_block(x, 4) for x in imgs_g])
    Ob = np.stack([_block(x, 4) for x in imgs_o])
    mg, mo = G.mean(0), Ob.mean(0)
    se = np.sqrt(G.var(0, ddof=1) / K + Ob.var(0, ddof=1) / K)
    ok = se > 0
    z = np.abs(mg - mo)[ok] / se[ok]
    frac_blocks = (z <= 3).mean()
    # per-pixel fraction (reported)
    pg, po = np.stack(imgs_g), np.stack(imgs_o)
    sep = np.sqrt(pg.var(0, ddof=1) / K + po.var(0, ddof=1) / K)
    okp = sep > 0
    frac_pix = (np.abs(pg.mean(0) - po.mean(0))[okp] / sep[okp] <= 3).mean()
    print(f"3-sigma: blocks {frac_blocks:.4f}, pixels {frac_pix:.4f}, mean gpu {mg.mean():.5g} oracle {mo.mean():.5g}")
    assert frac_blocks >= 0.98
    # the image means agree within 3 sigma of their own spread
    tot_g, tot_o = G.mean(axis=(1, 2, 3)), Ob.mean(axis=(1, 2, 3))
    zt = abs(tot_g.mean() - tot_o.mean()) / np.sqrt(tot_g.var(ddof=1) / K + tot_o.var(ddof=1) / K)
    assert zt <= 3.5, zt
