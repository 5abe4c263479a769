"""The oracle's init_circular_wave / add_circular_wave (wave_common.c:89-111,
337-351) against the reference's own functions (fixtures written by
tests/golden/make_golden.py from oracle/_ref builds), bit for bit, and
closed forms the definition fixes."""
import math
import os

import numpy as np
import pytest

from gen_inputs import PULSE_INIT, PULSES

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
BOX = (-2.0, 2.0, -1.125, 1.125)
AMP, VAR, WL = 0.75, 0.05, 0.1   # INITIAL_AMP, INITIAL_VARIANCE, INITIAL_WAVELENGTH of the builds


@pytest.fixture(scope="module")
def gold():
    return np.load(os.path.join(GOLD, "pulse_cases.npz"))


def same(a, b):
    return np.array_equal(np.ascontiguousarray(a).view(np.uint64), np.ascontiguousarray(b).view(np.uint64))


@pytest.mark.parametrize("ts", [0, 1])
def test_pulses_match_reference(oracle, gold, ts):
    pre = "ts%d/" % ts
    xy = gold[pre + "xy_in"]
    p = oracle.make_params(97, 61, twospeeds=ts)
    phi, psi = oracle.circular_wave(p, 0, 0.0, *PULSE_INIT, AMP, VAR, WL, BOX, np.full((97, 61), 3.0),
                                    np.full((97, 61), 3.0), xy)
    assert same(phi, gold[pre + "phi0"]) and same(psi, gold[pre + "psi0"])
    for n, (f, x, y) in enumerate(PULSES):
        phi, psi = oracle.circular_wave(p, 1, f, x, y, AMP, VAR, WL, BOX, phi, psi, xy)
        assert same(phi, gold[pre + "phi%d" % (n + 1)]), n
        assert same(psi, gold[pre + "psi%d" % (n + 1)]), n


def test_pulse_closed_forms(oracle):
    """A drop centred on a grid corner takes amp there; the field is even
    about a centre on the grid; masked cells stay untouched (one speed) and
    are set with two speeds; planted mistakes (cos(+d) in place of the
    Gaussian's shape, a dropped factor) change the values."""
    nx, ny = 64, 48
    box = (-1.0, 1.0, -0.75, 0.75)
    h = 2.0 / nx
    xy = np.ones((nx, ny), np.int16)
    xy[10:14, 30:35] = 0
    i0, j0 = 32, 24     # corner of ij_to_xy: (0, 0)
    x0, y0 = -1.0 + i0 * h, -0.75 + j0 * 1.5 / ny
    for ts in (0, 1):
        p = oracle.make_params(nx, ny, twospeeds=ts)
        z = np.zeros((nx, ny))
        phi, psi = oracle.circular_wave(p, 0, 0.0, x0, y0, 2.5, 0.01, 0.07, box, z, z + 1.0, xy)
        assert phi[i0, j0] == 2.5 and np.all(psi == 0.0)
        assert np.allclose(phi[i0 - 5:i0 + 6, j0], phi[i0 + 5:i0 - 6:-1, j0], rtol=0, atol=1e-15)
        if ts == 0:
            assert np.all(phi[10:14, 30:35] == 0.0)
        else:
            assert np.any(phi[10:14, 30:35] != 0.0)
        # add: linear in the factor, psi untouched
        a, b = oracle.circular_wave(p, 1, 1.0, x0, y0, 2.5, 0.01, 0.07, box, z, z + 1.0, xy)
        c, _ = oracle.circular_wave(p, 1, -2.0, x0, y0, 2.5, 0.01, 0.07, box, z, z + 1.0, xy)
        assert np.allclose(c, -2.0 * a, rtol=1e-15, atol=0) and np.all(b == 1.0)
        # explicit form at one cell
        x, y = -1.0 + 40 * h, -0.75 + 20 * 1.5 / ny
        d2 = (x - x0) ** 2 + (y - y0) ** 2
        assert a[40, 20] == 2.5 * 1.0 * math.exp(-d2 / 0.01) * math.cos(-math.sqrt(d2) / 0.07)
        assert a[40, 20] != 2.5 * math.exp(-d2 / 0.01) * math.cos(math.sqrt(d2) / 0.05)
