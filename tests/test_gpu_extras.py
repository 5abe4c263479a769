"""Neumann ghost points, driven cells and FLOOR on the GPU (FP64 form),
through the C ABI (wv_set_neumann, wv_set_drive, wv_set_floor).

* against the reference's own stepper: the fixtures of
  tests/golden/ext_cases.npz (wave_billiard.c:379-389, 397, 585-609 compiled
  with BC_NEUMANN, OSCILLATE_LEFT and FLOOR), one step per launch;
* against the oracle (pinned to those fixtures) with fused launches, point
  sources on top, at several tilings.
"""
import os

import numpy as np
import pytest

from conftest import say
from gen_inputs import EXT_STEPS, EXT_VMAX, KAPPAS, ext_cases, ext_inputs

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def bits_equal(a, b):
    return np.array_equal(np.ascontiguousarray(a).view(np.uint64), np.ascontiguousarray(b).view(np.uint64))


def first_diff(a, b):
    d = np.argwhere(np.ascontiguousarray(a).view(np.uint64) != np.ascontiguousarray(b).view(np.uint64))
    return "first differences at %s" % d[:4].tolist()


@pytest.fixture(scope="module")
def wh():
    import wavehip
    wavehip.load()
    return wavehip


@pytest.fixture(scope="module")
def ext():
    return np.load(os.path.join(GOLD, "ext_cases.npz"))


@pytest.mark.parametrize("split_at", [EXT_STEPS, 4])
@pytest.mark.parametrize("case", ext_cases(), ids=lambda c: c[0])
def test_extras_match_reference_fixture(wh, ext, case, split_at):
    tag, nx, ny, bc, ts = case
    say("extras fixture", tag, split_at)
    d = ext_inputs(tag, nx, ny)
    cells = d["drive_cells"]
    W = wh.Wave(dict(nx=nx, ny=ny, bcond=bc, twospeeds=ts, **KAPPAS))
    W.set_tables(d["xy_in"], d["tcc"], d["tc"], d["tgamma"])
    W.set_neumann(d["neumann"])
    W.set_drive(dict(i=[c[0] for c in cells], j=[c[1] for c in cells], val=ext[tag + "/drive_val"]))
    W.set_floor(EXT_VMAX)
    W.set_state(d["phi"], d["psi"], d["tmp"])
    W.step(split_at)
    W.step(EXT_STEPS - split_at)
    phi, psi, tmp = W.get_state(three=True)
    W.close()
    assert bits_equal(phi, ext[tag + "/phi"]), first_diff(phi, ext[tag + "/phi"])
    assert bits_equal(psi, ext[tag + "/psi"])
    assert bits_equal(tmp, ext[tag + "/tmp"])


def neumann_near_walls(xy, rng, n):
    """n ghost points on cells outside the domain next to it, each averaging
    two in-grid neighbours within one cell (weights w, 1-w)"""
    nx, ny = xy.shape
    cand = [(i, j) for i in range(nx) for j in range(ny)
            if xy[i, j] <= 0 and any(0 <= i + a < nx and 0 <= j + b < ny and xy[i + a, j + b] > 0
                                     for a in (-1, 0, 1) for b in (-1, 0, 1))]
    pick = rng.choice(len(cand), size=min(n, len(cand)), replace=False)
    out = {k: [] for k in ("i", "j", "nbi1", "nbj1", "nbi2", "nbj2", "nx", "ny")}
    for q in pick:
        i, j = cand[q]
        near = [(i + a, j + b) for a in (-1, 0, 1) for b in (-1, 0, 1) if 0 <= i + a < nx and 0 <= j + b < ny]
        p1, p2 = near[rng.integers(len(near))], near[rng.integers(len(near))]
        w = rng.uniform(0, 1)
        for k, v in zip(out, (i, j, p1[0], p1[1], p2[0], p2[1], w, 1 - w)):
            out[k].append(v)
    return {k: np.array(v) for k, v in out.items()}


@pytest.mark.parametrize("stencil", [5, 9])
@pytest.mark.parametrize("bc,ts", [(0, 0), (1, 0), (2, 1), (3, 0)])
def test_extras_fused_match_oracle(wh, oracle, bc, ts, stencil):
    from wavehip import workloads
    nx, ny, n = 150, 97, 37
    rng = np.random.default_rng(100 + 10 * bc + ts)
    w = workloads.random_case(nx, ny, 60 + bc, bcond=bc, twospeeds=ts, stencil=stencil, kappas=True)
    w["phi"] *= 0.4
    w["psi"] *= 0.4
    # cells no step writes start at +0 (fused launches)
    w["phi"][w["xy_in"] <= 0] = 0
    w["psi"][w["xy_in"] <= 0] = 0
    nb = neumann_near_walls(w["xy_in"], rng, 120)
    t = np.arange(1, n + 1)
    dcells = [(0, j) for j in range(1, ny - 1)] + [(0, 0)]
    dval = np.array([0.5 * np.cos(0.3 * t + 0.01 * j) for (_, j) in dcells])
    drive = dict(i=[c[0] for c in dcells], j=[c[1] for c in dcells], val=dval)
    act = np.argwhere(w["xy_in"][1:-1, 1:-1] > 0) + 1   # cells every step writes
    (si0, sj0), (si1, sj1) = act[len(act) // 3], act[2 * len(act) // 3]
    src = dict(i=[si0, si1], j=[sj0, sj1], amp=np.vstack([0.2 * np.sin(t / 5.0), 0.1 * np.cos(t / 3.0)]))
    p = oracle.make_params(nx, ny, bcond=bc, twospeeds=ts, stencil=stencil, **KAPPAS)
    ref = oracle.run_x(p, w["phi"], w["psi"], n, w["xy_in"], w["tcc"], w["tc"], w["tgamma"], sources=src,
                       neumann=nb, drive=drive, vmax=0.35)
    assert (np.abs(ref[0]) == 0.35).any()  # the clamp acts
    for k, L in ((4, 16), (8, 12), (3, 17), (1, 64)):
        say("extras fused", bc, ts, stencil, k, L)
        W = wh.Wave(w["params"], tb_depth=k, rows_per_tile=L)
        W.set_tables(w["xy_in"], w["tcc"], w["tc"], w["tgamma"])
        W.set_neumann(nb)
        W.set_drive(drive)
        W.set_floor(0.35)
        W.set_sources(src)
        W.set_state(w["phi"], w["psi"])
        W.step(9)
        W.step(n - 9)
        g = W.get_state()
        W.close()
        assert bits_equal(g[0], ref[0]), (k, L, first_diff(g[0], ref[0]))
        assert bits_equal(g[1], ref[1]), (k, L)


def test_extras_argument_checks(wh):
    from wavehip import workloads
    w = workloads.random_case(40, 30, 3)
    W = wh.Wave(w["params"])
    W.set_tables(w["xy_in"], w["tcc"], w["tc"], w["tgamma"])
    far = dict(i=[5], j=[5], nbi1=[7], nbj1=[5], nbi2=[5], nbj2=[5], nx=[0.5], ny=[0.5])
    with pytest.raises(wh.WaveError):
        W.set_neumann(far)  # neighbour two cells away
    out = dict(i=[0], j=[5], nbi1=[-1], nbj1=[5], nbi2=[0], nbj2=[5], nx=[0.5], ny=[0.5])
    with pytest.raises(wh.WaveError):
        W.set_neumann(out)  # neighbour outside the grid
    with pytest.raises(wh.WaveError):
        W.set_drive(dict(i=[40], j=[0], val=np.zeros((1, 3))))
    with pytest.raises(wh.WaveError):
        W.set_floor(-1.0)
    W.close()
    S = wh.Wave(w["params"], form=1)
    S.set_tables(w["xy_in"], w["tcc"], w["tc"], w["tgamma"])
    with pytest.raises(wh.WaveError):
        S.set_floor(1.0)  # FP64 form only
    S.close()
