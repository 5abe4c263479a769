"""COMPARISON on the GPU (WV_FLAG_COMPARISON): the flat twin split into two
halves (wave_3d.c:736-990), against the reference builds
(tests/golden/cmp_cases.npz, every B_COND, TWOSPEEDS on odd ones) and against
the oracle with Neumann points, drive, FLOOR and sources at several strip
heights."""
import os

import numpy as np
import pytest

from conftest import say
from gen_inputs import KAPPAS, case_inputs, cmp_cases

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def wh():
    import wavehip
    wavehip.load()
    return wavehip


def bits_equal(a, b):
    return np.array_equal(np.ascontiguousarray(a).view(np.uint64), np.ascontiguousarray(b).view(np.uint64))


@pytest.mark.parametrize("case", cmp_cases(), ids=lambda c: c[0])
def test_comparison_matches_reference_fixture(wh, case):
    tag, nx, ny, bc, ts = case
    say("comparison fixture", tag)
    g = np.load(os.path.join(GOLD, "cmp_cases.npz"))
    d = case_inputs(tag, nx, ny)
    W = wh.Wave(dict(nx=nx, ny=ny, bcond=bc, twospeeds=ts, tb_shift=ny - 2, flags=wh.FLAG_COMPARISON, **KAPPAS))
    W.set_tables(d["xy_in"], d["tcc"], d["tc"], d["tgamma"])
    W.set_state(d["phi"], d["psi"], d["tmp"])
    W.step(3)
    W.step(4)
    phi, psi, tmp = W.get_state(three=True)
    W.close()
    assert bits_equal(phi, g[tag + "/phi"]) and bits_equal(psi, g[tag + "/psi"]) and bits_equal(tmp, g[tag + "/tmp"])


@pytest.mark.parametrize("bc,ts", [(0, 0), (1, 1), (2, 0), (3, 1), (4, 0)])
def test_comparison_matches_oracle_with_extras(wh, oracle, bc, ts):
    from wavehip import workloads
    nx, ny, n = 150, 98, 21
    w = workloads.random_case(nx, ny, 70 + bc, bcond=bc, twospeeds=ts, kappas=True)
    rng = np.random.default_rng(bc)
    tmp = rng.uniform(-0.4, 0.4, (nx, ny))
    w["phi"] *= 0.4
    w["psi"] *= 0.4
    # Neumann points next to and on the interface columns
    pts = [(i, j) for i in range(3, nx - 3, 7) for j in (ny // 2 - 2, ny // 2 - 1, ny // 2, ny // 2 + 1)]
    nb = dict(i=[p[0] for p in pts], j=[p[1] for p in pts], nbi1=[p[0] + 1 for p in pts], nbj1=[p[1] for p in pts],
              nbi2=[p[0] for p in pts], nbj2=[p[1] - 1 for p in pts], nx=[0.6] * len(pts), ny=[0.4] * len(pts))
    t = np.arange(1, n + 1)
    dcells = [(0, j) for j in range(1, ny - 1)] + [(0, 0), (0, ny - 1)]
    drive = dict(i=[c[0] for c in dcells], j=[c[1] for c in dcells],
                 val=np.array([0.3 * np.cos(0.2 * t + 0.02 * j) for (_, j) in dcells]))
    act = np.argwhere(w["xy_in"][1:-1, 1:-1] > 0) + 1
    src = dict(i=[act[5][0]], j=[act[5][1]], amp=np.vstack([0.1 * np.sin(t / 4.0)]))
    kap = {k: v for k, v in w["params"].items() if k.startswith(("kappa", "gamma"))}
    p = oracle.make_params(nx, ny, bcond=bc, twospeeds=ts, tb_shift=ny - 2, comparison=1, **kap)
    ref = oracle.run_x(p, w["phi"], w["psi"], n, w["xy_in"], w["tcc"], w["tc"], w["tgamma"], tmp=tmp, sources=src,
                       neumann=nb, drive=drive, vmax=0.5)
    params = dict(w["params"], tb_shift=ny - 2, flags=wh.FLAG_COMPARISON)
    for L in (8, 16, 64):
        say("comparison vs oracle", bc, ts, L)
        W = wh.Wave(params, rows_per_tile=L)
        W.set_tables(w["xy_in"], w["tcc"], w["tc"], w["tgamma"])
        W.set_neumann(nb)
        W.set_drive(drive)
        W.set_floor(0.5)
        W.set_sources(src)
        W.set_state(w["phi"], w["psi"], tmp)
        W.step(6)
        W.step(n - 6)
        got = W.get_state(three=True)
        W.close()
        for a, b in zip(got, ref):
            assert bits_equal(a, b), (L, np.argwhere(a != b)[:4].tolist())


def test_comparison_argument_checks(wh):
    from wavehip import workloads
    w = workloads.random_case(40, 30, 3)
    for bad in (dict(form=1), dict(stencil=9), dict(ny=3, nx=40)):
        with pytest.raises(wh.WaveError):
            wh.Wave(dict(w["params"], flags=wh.FLAG_COMPARISON, **{k: v for k, v in bad.items() if k != "form"}),
                    form=bad.get("form", 0))
    W = wh.Wave(dict(w["params"], flags=wh.FLAG_COMPARISON))
    W.set_tables(w["xy_in"], w["tcc"], w["tc"], w["tgamma"])
    with pytest.raises(wh.WaveError):
        W.set_laplacian(workloads.laplacian_table(40, 30, 0))
    W.close()


def test_comparison_energy(wh, oracle):
    """wv_energy with COMPARISON takes no gradient across the halves, as
    compute_energy_mod (wave_common.c:1965-1971); same bits as the oracle."""
    from wavehip import workloads
    w = workloads.random_case(90, 64, 8, bcond=1)
    W = wh.Wave(dict(w["params"], flags=wh.FLAG_COMPARISON, e_scale=3.0))
    W.set_tables(w["xy_in"], w["tcc"], w["tc"], w["tgamma"])
    W.set_state(w["phi"], w["psi"])
    e = W.energy()
    W.close()
    p = oracle.make_params(90, 64, bcond=1, comparison=1)
    ref = oracle.energy_total(p, w["phi"], w["psi"], w["xy_in"], w["tcc"], 3.0)
    p0 = oracle.make_params(90, 64, bcond=1)
    assert e == ref and e != oracle.energy_total(p0, w["phi"], w["psi"], w["xy_in"], w["tcc"], 3.0)
