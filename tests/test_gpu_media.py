"""Gaussian-well media built on the GPU (wv_set_wells_tables) against the
reference's init_ior_2d tables (tests/golden/media_cases.npz) and the oracle.
The sum uses the GPU's exp and tanh (within an ulp or two of glibc's), so tcc
agrees to a relative 1e-14; tgamma and the table rule are exact, and the
installed tables step exactly as the oracle does with the same tables."""
import os

import numpy as np
import pytest

from conftest import say
from gen_inputs import case_inputs

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
BOX = (-2.0, 2.0, -1.125, 1.125)


@pytest.fixture(scope="module")
def wh():
    import wavehip
    wavehip.load()
    return wavehip


def test_wells_tables_match_reference_fixture(wh):
    g = np.load(os.path.join(GOLD, "media_cases.npz"))
    for tag, rand, ts in (("ior_wells_p", 0, 0), ("ior_wells_r", 1, 1)):
        say("wells fixture", tag)
        d = case_inputs(tag, 40, 30)
        W = wh.Wave(dict(nx=40, ny=30, twospeeds=ts))
        tcc, tc, tg = W.set_wells_tables(d["xy_in"], g[tag + "/xc"], g[tag + "/yc"], g[tag + "/h"] if rand else None,
                                         float(g[tag + "/sigma"]), rand, 0.1, 0.05, 1.0e-4, BOX)
        W.close()
        assert np.allclose(tcc, g[tag + "/tcc"], rtol=1e-14, atol=0)
        assert np.allclose(tc, g[tag + "/tc"], rtol=1e-14, atol=0)
        assert np.array_equal(tg, g[tag + "/tgamma"])
        assert np.array_equal(tc == 0, g[tag + "/tc"] == 0)


def test_wells_tables_match_oracle_and_step(wh, oracle):
    nx, ny, nw = 600, 400, 150
    rng = np.random.default_rng(12)
    xc, yc = rng.uniform(-2, 2, nw), rng.uniform(-1.125, 1.125, nw)
    h = np.clip(0.5 + 0.5 * rng.standard_normal(nw), 0, 1)
    x = -2.0 + np.arange(nx) * 4.0 / nx
    y = -1.125 + np.arange(ny) * 2.25 / ny
    X, Y = np.meshgrid(x, y, indexing="ij")
    xy = (X * X + Y * Y > 0.09).astype(np.int16)
    sigma = 0.2 * (4.0 * 2.25) / nw
    p = oracle.make_params(nx, ny, bcond=2)
    ref = oracle.wells_tables(p, xy, xc, yc, h, sigma, 1, 0.1, 0.05, 1e-4, BOX)
    W = wh.Wave(dict(nx=nx, ny=ny, bcond=2))
    got = W.set_wells_tables(xy, xc, yc, h, sigma, 1, 0.1, 0.05, 1e-4, BOX)
    for a, b in zip(got, ref):
        assert np.allclose(a, b, rtol=1e-14, atol=0)
    # the installed tables step like the oracle with the same tables
    phi = np.exp(-((X + 1.0) ** 2 + Y ** 2) / 0.01) * (xy != 0)
    W.set_state(phi, phi)
    W.step(40)
    u = W.get_state()
    W.close()
    r = oracle.run(p, phi, phi, 40, xy, got[0], got[1], got[2])
    assert np.array_equal(u[0].view(np.uint64), r[0].view(np.uint64))


def test_mask_matches_reference_fixture(wh):
    """wv_build_mask against the masks of the reference's own
    xy_in_billiard_single_domain (tests/golden/mask_cases.npz): exact"""
    from gen_inputs import MASK_BOX, mask_cases, mask_circles
    g = np.load(os.path.join(GOLD, "mask_cases.npz"))
    sizes = {"mask_97x61_l03": (97, 61), "mask_97x61_l06_m03": (97, 61), "mask_1280x720_l03": (1280, 720)}
    for tag, lam, mu in mask_cases():
        nx, ny = sizes[tag]
        W = wh.Wave(dict(nx=nx, ny=ny))
        for dom in (wh.D_SINAI, wh.D_EXT_ELLIPSE, wh.D_CIRCLES):
            say("mask fixture", tag, dom)
            got = W.build_mask(dom, MASK_BOX, lam, mu, mask_circles() if dom == wh.D_CIRCLES else None)
            assert np.array_equal(got, g["%s/%d" % (tag, dom)]), (tag, dom)
        W.close()


def test_mask_matches_oracle(wh, oracle):
    """other sizes, boxes and circle lists (one longer than a wave), and a
    mask that steps: the GPU mask fed to wv_set_tables equals the oracle's"""
    rng = np.random.default_rng(11)
    n = 150
    circ = oracle.circle_list(rng.uniform(-1.5, 3.0, n), rng.uniform(-1.2, 1.2, n), rng.uniform(0.01, 0.2, n),
                              (rng.uniform(0, 1, n) > 0.1).astype(float))
    box = (-0.75, 3.25, -1.197916667, 1.197916667)
    for nx, ny in ((300, 203), (64, 4100)):
        W = wh.Wave(dict(nx=nx, ny=ny))
        for dom, lam, mu in ((wh.D_SINAI, 0.7, 0.0), (wh.D_EXT_ELLIPSE, 1.1, 0.45), (wh.D_CIRCLES, 0.0, 0.0)):
            say("mask", nx, ny, dom)
            c = circ if dom == wh.D_CIRCLES else None
            got = W.build_mask(dom, box, lam, mu, c)
            assert np.array_equal(got, oracle.mask(nx, ny, box, dom, lam, mu, c))
        W.close()
    W = wh.Wave(dict(nx=40, ny=30))
    for bad in ((1, BOX), (wh.D_SINAI, (1.0, 1.0, 0.0, 1.0))):
        with pytest.raises(wh.WaveError):
            W.build_mask(bad[0], bad[1], 0.5)
    W.close()
