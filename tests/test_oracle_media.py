"""The media oracle (orc_wells_tables) against the reference's init_ior_2d.

tests/golden/media_cases.npz holds the tables init_ior_2d (sub_wave.c:12164,
compiled from its sources) builds for IOR_PERIODIC_WELLS and IOR_RANDOM_WELLS
on a 40x30 grid, with the wells it placed; the oracle reproduces them bit for
bit from the wells."""
import os

import numpy as np

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
