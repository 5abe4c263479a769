"""Probes (wv_set_probes / wv_get_probes): the time series of
wave_billiard.c:900-908 recorded on the GPU, against the oracle stepped one
step at a time.  Samples fall inside fused launches (every 3 steps with 8
steps per launch) and at cells of every kind: bulk, edge, tile borders, and
cells no step writes (which keep the buffer's value, FLOOR-clamped)."""
import numpy as np
import pytest

from conftest import say

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def wh():
    import wavehip
    wavehip.load()
    return wavehip


def bits_equal(a, b):
    return np.array_equal(np.ascontiguousarray(a).view(np.uint64), np.ascontiguousarray(b).view(np.uint64))


def oracle_series(oracle, p, w, cells, nsteps, every, tmp, **x):
    phi, psi, t = w["phi"], w["psi"], tmp
    out = []
    for g in range(1, nsteps + 1):
        if x:
            phi, psi, t = oracle.run_x(p, phi, psi, 1, w["xy_in"], w["tcc"], w["tc"], w["tgamma"], tmp=t,
                                       step0=g - 1, **x)
        else:
            phi, psi, t = oracle.run(p, phi, psi, 1, w["xy_in"], w["tcc"], w["tc"], w["tgamma"], tmp=t)
        if g % every == 0:
            out.append([phi[i, j] for (i, j) in cells])
    return np.array(out)


@pytest.mark.parametrize("exact", [False, True])
def test_probes_match_oracle_series(wh, oracle, exact):
    from wavehip import workloads
    nx, ny = 140, 120
    w = workloads.random_case(nx, ny, 31, bcond=2, kappas=True)
    w["phi"] *= 0.5
    w["psi"] *= 0.5
    tmp = np.random.default_rng(3).uniform(-0.8, 0.8, (nx, ny)) if exact else np.zeros((nx, ny))
    if not exact:
        for k in ("phi", "psi"):
            w[k][w["xy_in"] <= 0] = 0  # fused launches
    never = np.argwhere(w["xy_in"][1:-1, 1:-1] <= 0)[0] + 1
    cells = [(70, 60), (0, 50), (139, 119), (16, 51), (17, 52), (33, 103), tuple(never)]
    every, nsteps = 3, 61
    kw = dict(vmax=0.6) if exact else {}
    p = oracle.make_params(nx, ny, bcond=2, **{k: v for k, v in w["params"].items() if k.startswith(("kappa", "gamma"))})
    ref = oracle_series(oracle, p, w, cells, nsteps, every, tmp, **kw)
    say("probes", exact)
    W = wh.Wave(w["params"])
    W.set_tables(w["xy_in"], w["tcc"], w["tc"], w["tgamma"])
    if exact:
        W.set_floor(0.6)
    W.set_probes([c[0] for c in cells], [c[1] for c in cells], every=every, capacity=64)
    W.set_state(w["phi"], w["psi"], tmp)
    got = []
    for chunk in (7, 30, 24):
        W.step(chunk)
        s, lost = W.get_probes()
        assert lost == 0
        got.append(s)
    got = np.concatenate(got)
    W.close()
    assert got.shape == ref.shape, (got.shape, ref.shape)
    assert bits_equal(got, ref), np.argwhere(got != ref)[:5]


def test_probes_capacity_and_checks(wh):
    from wavehip import workloads
    w = workloads.random_case(60, 50, 2)
    for k in ("phi", "psi"):
        w[k][w["xy_in"] <= 0] = 0
    W = wh.Wave(w["params"])
    W.set_tables(w["xy_in"], w["tcc"], w["tc"], w["tgamma"])
    W.set_state(w["phi"], w["psi"])
    with pytest.raises(wh.WaveError):
        W.set_probes([60], [0])
    with pytest.raises(wh.WaveError):
        W.get_probes()  # none set
    W.set_probes([5, 6], [7, 8], every=2, capacity=4)
    W.step(13)  # 6 samples due, 4 kept
    s, lost = W.get_probes()
    assert s.shape == (4, 2) and lost == 2
    W.step(4)
    s2, lost2 = W.get_probes()
    assert s2.shape == (2, 2) and lost2 == 0  # steps 14, 16 of the run
    W.close()
