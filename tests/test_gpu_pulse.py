"""init_circular_wave / add_circular_wave on the GPU (wv_init_circular_wave,
wv_add_circular_wave) against the reference's own functions (fixtures from
oracle/_ref builds) and the oracle: within a few ulps (exp, cos are the
device's), then stepping from the GPU-built field bit for bit against the
oracle, in fused launches and, where a drop reaches cells no step writes,
one step per launch."""
import os

import numpy as np
import pytest

from conftest import say
from gen_inputs import PULSE_INIT, PULSES

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
BOX = (-2.0, 2.0, -1.125, 1.125)
AMP, VAR, WL = 0.75, 0.05, 0.1


@pytest.fixture(scope="module")
def wh():
    import wavehip
    wavehip.load()
    return wavehip


@pytest.fixture(scope="module")
def gold():
    return np.load(os.path.join(GOLD, "pulse_cases.npz"))


def bits_equal(a, b):
    return np.array_equal(np.ascontiguousarray(a).view(np.uint64), np.ascontiguousarray(b).view(np.uint64))


def tables(xy):
    nx, ny = xy.shape
    return np.full((nx, ny), 0.01), np.full((nx, ny), 0.1), np.zeros((nx, ny))


@pytest.mark.parametrize("form", [0, 1])
@pytest.mark.parametrize("ts", [0, 1])
def test_drop_and_pulses_match_reference(wh, gold, ts, form):
    pre = "ts%d/" % ts
    xy = gold[pre + "xy_in"]
    say("pulses vs reference, twospeeds", ts, "form", form)
    W = wh.Wave(dict(nx=97, ny=61, bcond=2, twospeeds=ts), form=form)
    W.set_tables(xy, *tables(xy))
    W.init_circular_wave(PULSE_INIT[0], PULSE_INIT[1], AMP, VAR, WL, BOX)
    phi, psi = W.get_state()
    # the device's exp and cos are within a few ulps of glibc's
    tol_phi = 1e-15 * AMP if form == 0 else 1e-14 * AMP
    tol_psi = 0.0 if form == 0 else 2e-7 * AMP      # split: psi = u - v, v in binary32
    assert np.abs(phi - gold[pre + "phi0"]).max() <= tol_phi
    assert np.abs(psi - gold[pre + "psi0"]).max() <= tol_psi
    for n, (f, x, y) in enumerate(PULSES):
        W.add_circular_wave(f, x, y, AMP, VAR, WL, BOX)
        phi, psi = W.get_state()
        assert np.abs(phi - gold[pre + "phi%d" % (n + 1)]).max() <= tol_phi * (n + 2), n
        assert np.abs(psi - gold[pre + "psi%d" % (n + 1)]).max() <= tol_psi * (n + 2), n
    # cells the drawing rule does not reach stay zero
    if ts == 0:
        assert np.all(phi[xy == 0] == 0.0)
    W.close()


def test_steps_after_pulses_bit_exact(wh, oracle):
    """FP64: drop and pulses on the GPU, then fused steps; the oracle steps
    the same field (downloaded after the pulses) to the same bits.  A mask
    with ghost cells (xy_in = -1, drawn but never written by a step) makes
    the context step one step per launch with three buffers."""
    from wavehip import workloads
    for neg in (False, True):
        nx, ny = 150, 97
        rng = np.random.default_rng(7)
        xy = np.ones((nx, ny), np.int16)
        xy[rng.integers(0, nx, 200), rng.integers(0, ny, 200)] = 0
        if neg:
            xy[rng.integers(1, nx - 1, 40), rng.integers(1, ny - 1, 40)] = -1
        box = (-1.5, 1.5, -1.0, 1.0)
        say("steps after pulses, ghost cells", neg)
        W = wh.Wave(dict(nx=nx, ny=ny, bcond=2), tb_depth=4)
        W.set_tables(xy, *tables(xy))
        W.init_circular_wave(-0.3, 0.1, 1.0, 0.02, 0.05, box)
        W.add_circular_wave(-0.8, 0.5, -0.2, 1.0, 0.02, 0.05, box)
        phi, psi, tmp = W.get_state(three=True)
        W.step(21)
        W.add_circular_wave(0.6, 0.0, 0.4, 1.0, 0.02, 0.05, box)
        mid = W.get_state(three=True)
        W.step(13)
        got = W.get_state()
        W.close()
        p = oracle.make_params(nx, ny, bcond=2)
        t = tables(xy)
        a = oracle.run(p, phi, psi, 21, xy, *t, tmp=tmp)
        # the pulse changes phi only; psi (and, stepping one step per launch
        # with three buffers, tmp) are the 21 steps' own
        assert bits_equal(mid[1], a[1])
        if neg:
            assert bits_equal(mid[2], a[2])
        assert np.abs(mid[0] - a[0]).max() > 1e-3      # the pulse landed
        b = oracle.run(p, mid[0], mid[1], 13, xy, *t, tmp=mid[2])
        assert bits_equal(got[0], b[0]) and bits_equal(got[1], b[1]), neg


def test_batch_pulses_equal_single_grid(wh):
    xy = np.ones((80, 70), np.int16)
    xy[30:40, 20:30] = 0
    box = (-1.0, 1.0, -0.875, 0.875)
    W = wh.Wave(dict(nx=80, ny=70, bcond=1))
    W.set_tables(xy, *tables(xy))
    W.init_circular_wave(0.2, -0.1, 1.0, 0.02, 0.05, box)
    W.add_circular_wave(0.5, -0.4, 0.3, 1.0, 0.02, 0.05, box)
    W.step(10)
    a = W.get_state()
    W.close()
    xy2 = np.ones((50, 60), np.int16)
    B = wh.Batch([dict(nx=50, ny=60, bcond=0), dict(nx=80, ny=70, bcond=1)])
    B.load(0, dict(xy_in=xy2, tcc=np.full((50, 60), 0.01), tc=np.full((50, 60), 0.1), tgamma=np.zeros((50, 60)),
                   phi=np.zeros((50, 60)), psi=np.zeros((50, 60))))
    B.set_tables(1, xy, *tables(xy))
    B.init_circular_wave(1, 0.2, -0.1, 1.0, 0.02, 0.05, box)
    B.add_circular_wave(1, 0.5, -0.4, 0.3, 1.0, 0.02, 0.05, box)
    B.step(10)
    b = B.get_state(1)
    B.close()
    assert bits_equal(a[0], b[0]) and bits_equal(a[1], b[1])


def test_pulse_argument_checks(wh):
    W = wh.Wave(dict(nx=20, ny=20, bcond=0))
    with pytest.raises(wh.WaveError):
        W.init_circular_wave(0, 0, 1, 0.01, 0.05, (-1, 1, -1, 1))   # no mask yet
    xy = np.ones((20, 20), np.int16)
    W.set_tables(xy, *tables(xy))
    with pytest.raises(wh.WaveError):
        W.add_circular_wave(1.0, 0, 0, 1, 0.01, 0.05, (-1, 1, -1, 1))   # no state yet
    with pytest.raises(wh.WaveError):
        W.init_circular_wave(0, 0, 1, 0.0, 0.05, (-1, 1, -1, 1))   # variance 0
    with pytest.raises(wh.WaveError):
        W.init_circular_wave(0, 0, 1, 0.01, 0.05, (1, -1, -1, 1))  # empty box
    W.close()
