"""The media oracle (orc_wells_tables) against the reference's init_ior_2d.

tests/golden/media_cases.npz holds the tables init_ior_2d (sub_wave.c:12164,
compiled from its sources) builds for IOR_PERIODIC_WELLS and IOR_RANDOM_WELLS
on a 40x30 grid, with the wells it placed; the oracle reproduces them bit for
bit from the wells."""
import os

import numpy as np
import pytest

from gen_inputs import case_inputs

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
BOX = (-2.0, 2.0, -1.125, 1.125)
CASES = {"ior_wells_p": dict(rand=0, ts=0), "ior_wells_r": dict(rand=1, ts=1)}


def bits_equal(a, b):
    return np.array_equal(np.asarray(a).view(np.uint64), np.asarray(b).view(np.uint64))


def test_wells_oracle_matches_reference_fixture(oracle):
    g = np.load(os.path.join(GOLD, "media_cases.npz"))
    for tag, c in CASES.items():
        d = case_inputs(tag, 40, 30)
        p = oracle.make_params(40, 30, twospeeds=c["ts"])
        h = g[tag + "/h"] if c["rand"] else None
        t = oracle.wells_tables(p, d["xy_in"], g[tag + "/xc"], g[tag + "/yc"], h, float(g[tag + "/sigma"]),
                                c["rand"], 0.1, 0.05, 1.0e-4, BOX)
        for name, a in zip(("tcc", "tc", "tgamma"), t):
            assert bits_equal(a, g[tag + "/" + name]), (tag, name)
        # planted mistakes: without tanh, or without the heights, the tables differ
        if c["rand"]:
            t2 = oracle.wells_tables(p, d["xy_in"], g[tag + "/xc"], g[tag + "/yc"], h, float(g[tag + "/sigma"]), 0,
                                     0.1, 0.05, 1.0e-4, BOX)
            t3 = oracle.wells_tables(p, d["xy_in"], g[tag + "/xc"], g[tag + "/yc"], None,
                                     float(g[tag + "/sigma"]), 1, 0.1, 0.05, 1.0e-4, BOX)
            assert not bits_equal(t2[0], g[tag + "/tcc"]) and not bits_equal(t3[0], g[tag + "/tcc"])


def test_wells_closed_forms(oracle):
    """A single well on a grid point gives s = h there (exp(0) = 1), so
    tcc = c*h + cb*(1-h); far from it tcc = cb; tc = sqrt(tcc) on the domain,
    cb outside it with TWOSPEEDS, 0 otherwise."""
    nx, ny = 32, 20
    x = -2.0 + np.arange(nx) * 4.0 / nx
    y = -1.125 + np.arange(ny) * 2.25 / ny
    xy = np.ones((nx, ny), np.int16)
    xy[0, :] = 0
    p = oracle.make_params(nx, ny, twospeeds=0)
    tcc, tc, tg = oracle.wells_tables(p, xy, [x[10]], [y[7]], [0.75], 1e-3, 0, 0.1, 0.05, 2e-4, BOX)
    assert tcc[10, 7] == 0.1 * 0.75 + 0.05 * (1.0 - 0.75)
    assert tcc[25, 18] == 0.05 and np.all(tg == 2e-4)
    assert tc[10, 7] == np.sqrt(tcc[10, 7]) and np.all(tc[0] == 0)
    p2 = oracle.make_params(nx, ny, twospeeds=1)
    _, tc2, _ = oracle.wells_tables(p2, xy, [x[10]], [y[7]], None, 1e-3, 1, 0.1, 0.05, 2e-4, BOX)
    assert np.all(tc2[0] == 0.05)


# ---- init_ior_2d's per-cell models ------------------------------------------

from gen_inputs import IOR_CODES, IOR_PARAMS, MASK_BOX, ROT_ANGLES, ior_mask  # noqa: E402


def ior_gold():
    return np.load(os.path.join(GOLD, "ior_cases.npz"))


def test_ior_oracle_matches_reference_fixture(oracle):
    """every per-cell model (reference builds ior_m<code>, 97x61, TWOSPEEDS =
    code % 2) and the VARIABLE_IOR == 0 branch, bit for bit; the Mandelbrot
    models are pinned by the one-thread run of the reference (its loop shares
    u, v, k between OpenMP threads)"""
    g = ior_gold()
    xy = ior_mask()
    P = oracle.IorParams(**IOR_PARAMS)
    for m in IOR_CODES:
        p = oracle.make_params(97, 61, twospeeds=m % 2)
        t = oracle.ior_tables(p, xy, m, MASK_BOX, P)
        for name, a in zip(("tcc", "tc", "tgamma"), t):
            assert bits_equal(a, g["m%d/%s" % (m, name)]), (m, name)
    for ts in (0, 1):
        t = oracle.ior_tables(oracle.make_params(97, 61, twospeeds=ts), xy, 0, MASK_BOX, P)
        for name, a in zip(("tcc", "tc", "tgamma"), t):
            assert bits_equal(a, g["const_ts%d/%s" % (ts, name)]), (ts, name)


def test_ior_fixtures_catch_planted_mistakes(oracle):
    """GAMMA and GAMMAB swapped, TWOSPEEDS flipped, the cell centre for the
    corner, LAMBDA or MU off by 5 % (the grid step is 0.04), one Mandelbrot
    level fewer: each fails the fixture of the models it touches"""
    g = ior_gold()
    xy = ior_mask()
    hx, hy = 0.5 * 4.0 / 97, 0.5 * 2.25 / 61
    centre = (MASK_BOX[0] + hx, MASK_BOX[1] + hx, MASK_BOX[2] + hy, MASK_BOX[3] + hy)
    for m in IOR_CODES:
        p = oracle.make_params(97, 61, twospeeds=m % 2)
        want = [g["m%d/%s" % (m, n)] for n in ("tcc", "tc", "tgamma")]

        def differs(box=MASK_BOX, pp=p, **over):
            t = oracle.ior_tables(pp, xy, m, box, oracle.IorParams(**dict(IOR_PARAMS, **over)))
            return any(not bits_equal(a, w) for a, w in zip(t, want))
        swapped = differs(gamma=IOR_PARAMS["gammab"], gammab=IOR_PARAMS["gamma"])
        geometric = m not in (14, 16, 17, 191)  # mask only; 191 reads (x - XMIN) / (XMAX - XMIN)
        assert swapped or m in (1, 100, 2, 19, 191), m
        if geometric:
            assert differs(box=centre), m
        if m in (10, 11, 12, 13, 15, 151, 18, 181, 182, 183):  # the models that read LAMBDA or MU
            assert differs(lam=0.84) or differs(mu=0.315), m
        assert differs(pp=oracle.make_params(97, 61, twospeeds=1 - m % 2)), m
    for m in (1, 100):
        p = oracle.make_params(97, 61, twospeeds=m % 2)
        t = oracle.ior_tables(p, xy, m, MASK_BOX, oracle.IorParams(**dict(IOR_PARAMS, mandellevel=99)))
        assert not bits_equal(t[0], g["m%d/tcc" % m])


def test_ior_closed_forms(oracle):
    """values the formulas fix: IOR_LINEAR_X_B is COURANT^2 at x = XMIN and
    tends to COURANTB^2; the Luneburg lens is COURANT^2/2 at its centre and
    COURANT^2 outside; inside the Mandelbrot set (c = 0, -1) the media is
    COURANT^2 and escape level MANDELLEVEL; IOR_EARTH is 0.8*COURANT at r = 0"""
    nx, ny = 40, 20
    box = (-2.0, 2.0, -1.0, 1.0)
    xy = np.ones((nx, ny), np.int16)
    p = oracle.make_params(nx, ny)
    P = oracle.IorParams(**IOR_PARAMS)
    c, cb = IOR_PARAMS["courant"], IOR_PARAMS["courantb"]
    tcc, tc, tg = oracle.ior_tables(p, xy, 191, box, P)
    assert tcc[0, 5] == c * c and abs(tcc[-1, 5] - (c + (cb - c) * 39 / 40) ** 2) < 1e-17
    tcc, _, tg = oracle.ior_tables(p, xy, 182, box, P)
    assert tcc[20, 10] == c * c / 2.0 and tcc[0, 0] == c * c
    assert tg[20, 10] == IOR_PARAMS["gammab"] and tg[0, 0] == IOR_PARAMS["gamma"]
    for m in (1, 100):
        tcc, _, _ = oracle.ior_tables(p, xy, m, box, P)
        assert tcc[20, 10] == c * c and tcc[10, 10] == c * c  # c = 0 and c = -1 stay bounded
    tcc, _, _ = oracle.ior_tables(p, xy, 2, box, P)
    assert tcc[20, 10] == 0.8 * c and tcc[0, 0] == 0.0


def test_rotating_wells_oracle_matches_reference_fixture(oracle):
    """the centres of the rotating models, summed by the wells oracle, give
    the reference's tables at every angle bit for bit"""
    g = ior_gold()
    for tag, large, ts in (("ior_rot", 0, 0), ("ior_rot_large", 1, 1)):
        xr = g[tag + "/xy_in"]
        p = oracle.make_params(40, 30, twospeeds=ts)
        for q, ang in enumerate(ROT_ANGLES):
            xc, yc, s = oracle.rotating_wells(6, 5, large, ang, MASK_BOX)
            tcc, tc, _ = oracle.wells_tables(p, xr, xc, yc, None, s, 0, 0.1, 0.05, 0.0, MASK_BOX)
            assert bits_equal(tcc, g["%s/%d/tcc" % (tag, q)]) and bits_equal(tc, g["%s/%d/tc" % (tag, q)]), (tag, q)
            # the other lattice, or the angle's sign flipped, does not
            xc2, yc2, s2 = oracle.rotating_wells(6, 5, 1 - large, ang, MASK_BOX)
            t2 = oracle.wells_tables(p, xr, xc2, yc2, None, s2, 0, 0.1, 0.05, 0.0, MASK_BOX)
            assert not bits_equal(t2[0], g["%s/%d/tcc" % (tag, q)])
            if ang != 0.0:
                xc3, yc3, _ = oracle.rotating_wells(6, 5, large, -ang, MASK_BOX)
                t3 = oracle.wells_tables(p, xr, xc3, yc3, None, s, 0, 0.1, 0.05, 0.0, MASK_BOX)
                assert not bits_equal(t3[0], g["%s/%d/tcc" % (tag, q)])


def test_library_rotating_wells_host_code(oracle):
    """wv_rotating_wells is host code (no device): its centres equal the
    oracle's; bad arguments are refused"""
    import wavehip
    for large in (0, 1):
        for ang in ROT_ANGLES + (12.5,):
            a = wavehip.rotating_wells(7, 4, large, ang, MASK_BOX)
            b = oracle.rotating_wells(7, 4, large, ang, MASK_BOX)
            assert bits_equal(a[0], b[0]) and bits_equal(a[1], b[1]) and a[2] == b[2]
    with pytest.raises(wavehip.WaveError):
        wavehip.rotating_wells(0, 4, 0, 0.0, MASK_BOX)
    with pytest.raises(wavehip.WaveError):
        wavehip.rotating_wells(3, 4, 0, 0.0, (1.0, 1.0, 0.0, 1.0))
