"""GPU parity: libwavehip.so against the reference fixtures and the CPU oracle.

FP64 form: bit-exact (every bit of every cell, including masked cells and
signed zeros) for every B_COND branch, both stencils, sources, every fused
depth and tile height.  Split form: bit-exact against the oracle's split
emulation (same binary32 operations in the same order), and within 1e-6
relative L2 of the FP64 path.
"""
import hashlib
import json
import os
import sys

import numpy as np
import pytest

from conftest import say
from gen_inputs import KAPPAS, case_inputs, small_cases

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def bits_equal(a, b):
    a, b = np.ascontiguousarray(a), np.ascontiguousarray(b)
    if a.dtype == np.float32:
        return np.array_equal(a.view(np.uint32), b.view(np.uint32))
    return np.array_equal(a.view(np.uint64), b.view(np.uint64))


def first_diff(a, b):
    d = np.argwhere(np.ascontiguousarray(a).view(np.uint64) != np.ascontiguousarray(b).view(np.uint64))
    return d[:5].tolist(), len(d)


@pytest.fixture(scope="module")
def wh():
    import wavehip
    wavehip.load()
    return wavehip


@pytest.fixture(scope="module")
def small():
    return np.load(os.path.join(GOLD, "small_cases.npz"))


def params_of(nx, ny, bc, ts, tb=0, tbs=None, stencil=5, form=0, **kw):
    p = dict(nx=nx, ny=ny, bcond=bc, twospeeds=ts, oscillate_topbot=tb,
             tb_shift=nx if tbs is None else tbs, stencil=stencil, form=form)
    p.update(kw)
    return p


# ------------------------------------------------------- reference fixtures --

@pytest.mark.parametrize("case", small_cases(), ids=lambda c: c[0])
def test_dropin_matches_reference_fixture(wh, small, case):
    """evolve_wave x2 + evolve_wave_half through the drop-in entry points
    (row pointers for wave_billiard.c cases, flat for wave_3d.c cases)."""
    tag, nx, ny, bc, ts, tb, tbs = case
    say("case", tag)
    d = case_inputs(tag, nx, ny)
    W = wh.Wave(params_of(nx, ny, bc, ts, tb, tbs, **KAPPAS))
    phi, psi, tmp = d["phi"].copy(), d["psi"].copy(), d["tmp"].copy()
    xy = np.ascontiguousarray(d["xy_in"])
    if tag.startswith("row"):
        for _ in range(2):
            W.evolve_wave_rows(phi, psi, tmp, xy, d["tcc"], d["tc"], d["tgamma"])
        W.evolve_wave_half_rows(phi, psi, tmp, xy, d["tcc"], d["tc"], d["tgamma"])
    else:
        for _ in range(2):
            W.evolve_wave_flat(phi, psi, tmp, xy, d["tc"], d["tcc"], d["tgamma"])
        W.evolve_wave_half_flat(phi, psi, tmp, xy, d["tc"], d["tcc"], d["tgamma"])
    W.close()
    # u^7 is in the caller's tmp, u^6 in phi, u^5 in psi
    assert bits_equal(tmp, small[tag + "/phi"]), first_diff(tmp, small[tag + "/phi"])
    assert bits_equal(phi, small[tag + "/psi"])
    assert bits_equal(psi, small[tag + "/tmp"])


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.mark.parametrize("tiling", [(1, 64), (4, 64), (8, 32), (3, 17)])
def test_c1_plain_workload_matches_reference_digests(wh, tiling):
    """1280x720 plain workload, 1,000 steps: phi after every 100 steps has the
    SHA-256 of the reference's own run."""
    from wavehip import workloads
    say("c1 tiling", tiling)
    w = workloads.c1()
    dig = json.load(open(os.path.join(GOLD, "c1_digests.json")))
    W = wh.Wave(w["params"], tb_depth=tiling[0], rows_per_tile=tiling[1])
    W.load(w)
    for n in range(100, 1001, 100):
        W.step(100)
        phi, _ = W.get_state()
        assert sha(phi) == dig["phi"][str(n)], "step %d" % n
    W.close()


# --------------------------------------------------------- oracle, random ----

CASES = []
for bc in range(5):
    for ts in (0, 1):
        for st in (5, 9):
            CASES.append((bc, ts, st))
SIZES = [(3, 3), (5, 7), (24, 17), (131, 67), (70, 300)]


@pytest.mark.parametrize("bc,ts,st", CASES)
def test_fp64_matches_oracle_every_tiling(wh, oracle, bc, ts, st):
    for (nx, ny) in SIZES:
        w = __import__("wavehip.workloads", fromlist=["x"]).random_case(nx, ny, nx * 7 + ny + bc, bcond=bc,
                                                                         twospeeds=ts, stencil=st, kappas=True)
        p = oracle.make_params(nx, ny, bcond=bc, twospeeds=ts, stencil=st, **KAPPAS)
        nsteps = 13
        # (a) arbitrary values on never-written cells: one step per launch
        ref = oracle.run(p, w["phi"], w["psi"], nsteps, w["xy_in"], w["tcc"], w["tc"], w["tgamma"])
        say("fp64 exact", bc, ts, st, nx, ny)
        W = wh.Wave(w["params"], tb_depth=4, rows_per_tile=8)
        W.load(w)
        W.step(nsteps)
        got = W.get_state(three=True)
        W.close()
        for a, b in zip(got, ref):
            assert bits_equal(a, b), (nx, ny, first_diff(a, b))
        # (b) never-written cells at +0: fused launches of every depth
        act = np.ones((nx, ny), bool)
        inner = np.zeros((nx, ny), bool)
        inner[1:-1, 1:-1] = True
        if not ts:
            act = np.where(inner, w["xy_in"] > 0, w["xy_in"] != 0)
        phi = np.where(act, w["phi"], 0.0)
        psi = np.where(act, w["psi"], 0.0)
        ref = oracle.run(p, phi, psi, nsteps, w["xy_in"], w["tcc"], w["tc"], w["tgamma"])
        for k, L in ((1, 64), (2, 5), (3, 1), (4, 64), (6, 9), (8, 3)):
            say("fp64 fused", bc, ts, st, nx, ny, k, L)
            W = wh.Wave(w["params"], tb_depth=k, rows_per_tile=L)
            W.set_tables(w["xy_in"], w["tcc"], w["tc"], w["tgamma"])
            W.set_state(phi, psi)
            W.step(nsteps)
            g = W.get_state()
            W.close()
            assert bits_equal(g[0], ref[0]), (nx, ny, k, L, first_diff(g[0], ref[0]))
            assert bits_equal(g[1], ref[1]), (nx, ny, k, L, first_diff(g[1], ref[1]))


@pytest.mark.parametrize("bc", [0, 1, 2, 3])
def test_topbot_oscillation_branch(wh, oracle, bc):
    from wavehip import workloads
    nx, ny = 40, 33
    w = workloads.random_case(nx, ny, 11 + bc, bcond=bc, kappas=True, neg_cells=False)
    for tbs in (nx, 7):
        p = oracle.make_params(nx, ny, bcond=bc, oscillate_topbot=1, tb_shift=tbs, **KAPPAS)
        ref = oracle.run(p, w["phi"], w["psi"], 9, w["xy_in"], w["tcc"], w["tc"], w["tgamma"])
        say("topbot", bc, tbs)
        W = wh.Wave(dict(w["params"], oscillate_topbot=1, tb_shift=tbs), tb_depth=1)
        W.load(w)
        W.step(9)
        g = W.get_state(three=True)
        W.close()
        for a, b in zip(g, ref):
            assert bits_equal(a, b)


def lens_case(nx, ny, bc, stencil, nsteps, seed=1):
    """two-speed lens, damping, a soft source and a Gaussian pulse"""
    from wavehip import workloads
    box = (-2.0, 2.0, -2.0 * ny / nx, 2.0 * ny / nx)
    x, y = workloads.grid_xy(nx, ny, *box)
    xy = workloads.mask_ext_ellipse(x, y, 0.6, 0.3)
    tcc, tc, tg = workloads.constant_media_tables(xy, 0.1, 0.1 / 1.5, 1e-4, 2e-4, 1)
    phi = workloads.gaussian_bumps(x, y, [(0.9, 0.1)], 0.1)
    si, sj = workloads.xy_to_ij(-1.5, 0.0, nx, ny, *box)
    t = np.arange(1, nsteps + 1)
    amp = np.vstack([np.sin(2 * np.pi * t / 40.0), 0.5 * np.cos(2 * np.pi * t / 25.0)])
    return dict(params=dict(nx=nx, ny=ny, bcond=bc, twospeeds=1, stencil=stencil), xy_in=xy, tcc=tcc,
                tc=tc, tgamma=tg, phi=phi, psi=phi.copy(),
                sources=dict(i=[si, si], j=[sj, sj + 3], amp=amp))


@pytest.mark.parametrize("stencil", [5, 9])
@pytest.mark.parametrize("bc", [0, 1, 2, 3])
def test_sources_and_energy_fp64(wh, oracle, bc, stencil):
    nx, ny, n = 150, 97, 60
    w = lens_case(nx, ny, bc, stencil, n - 10)
    p = oracle.make_params(nx, ny, bcond=bc, twospeeds=1, stencil=stencil)
    ref = oracle.run(p, w["phi"], w["psi"], n, w["xy_in"], w["tcc"], w["tc"], w["tgamma"], sources=w["sources"])
    e_ref = oracle.energy_total(p, ref[0], ref[1], w["xy_in"], w["tcc"], 1.0)
    for k, L in ((1, 64), (4, 16), (8, 64)):
        say("sources", bc, stencil, k, L)
        W = wh.Wave(w["params"], tb_depth=k, rows_per_tile=L)
        W.load(w)
        W.step(7)
        W.step(n - 7)
        g = W.get_state()
        e = W.energy()
        W.close()
        assert bits_equal(g[0], ref[0]), first_diff(g[0], ref[0])
        assert bits_equal(g[1], ref[1])
        assert e == e_ref


# ------------------------------------------------------------------ split ----

@pytest.mark.parametrize("stencil", [5, 9])
@pytest.mark.parametrize("bc", [0, 1, 2, 3, 4])
def test_split_matches_oracle_split(wh, oracle, bc, stencil):
    nx, ny, n = 133, 90, 25
    w = lens_case(nx, ny, bc, stencil, n)
    p = oracle.make_params(nx, ny, bcond=bc, twospeeds=1, stencil=stencil)
    hi, lo, v = oracle.split_from_fp64(w["phi"], w["psi"])
    hi, lo, v = oracle.split_run(p, hi, lo, v, n, w["xy_in"], w["tcc"], w["tc"], w["tgamma"],
                                 sources=w["sources"])
    u_ref, um_ref = oracle.split_to_fp64(hi, lo, v)
    e_ref = oracle.split_energy_total(p, hi, lo, v, w["xy_in"], w["tcc"], 1.0)
    for k, L in ((1, 64), (3, 7), (4, 64), (8, 16)):
        say("split", bc, stencil, k, L)
        W = wh.Wave(w["params"], form=1, tb_depth=k, rows_per_tile=L)
        W.load(w)
        W.step(n)
        g = W.get_state()
        e = W.energy()
        W.close()
        assert bits_equal(g[0], u_ref), first_diff(g[0], u_ref)
        assert bits_equal(g[1], um_ref)
        assert e == e_ref


@pytest.mark.parametrize("bc", [0, 1, 2, 3, 4])
@pytest.mark.parametrize("variant", ["plain", "damped", "topbot"])
@pytest.mark.parametrize("form", [0, 1])
def test_edge_column_tiles(wh, oracle, bc, variant, form):
    """Uniform medium, 5-point, no point lists: the tiles whose rows lie
    inside the box but whose columns reach column 0 or NY-1 take the
    interior evaluation plus the edge formula of the outer columns with the
    row tests resolved (K = 8), and the first and last strips skip the rows
    outside the box.  Random state, obstacle cells, every B_COND, damping
    and OSCILLATE_TOPBOT; bit-exact against the oracle (FP64 form against
    the reference's arithmetic, split form against its split restatement)."""
    nx, ny, n = 120, 150, 21
    rng = np.random.default_rng(40 + bc)
    xy = np.ones((nx, ny), np.int16)
    xy[rng.integers(0, nx, 300), rng.integers(0, ny, 300)] = 0
    xy[50:60, 0:5] = 0          # obstacle cells on an outer column
    g = 2e-4 if variant == "damped" else 0.0
    tcc, tc, tg = np.full((nx, ny), 0.01), np.full((nx, ny), 0.1), np.full((nx, ny), g)
    kw = dict(kappa=1e-4, kappa_sides=5e-4, gamma_sides=1e-4, kappa_topbot=2e-4, gamma_topbot=1e-7) \
        if variant == "damped" else {}
    if variant == "topbot":
        kw = dict(oscillate_topbot=1, tb_shift=nx // 2)
    phi, psi = rng.standard_normal((nx, ny)), rng.standard_normal((nx, ny))
    phi[xy == 0] = psi[xy == 0] = 0.0
    p = oracle.make_params(nx, ny, bcond=bc, **kw)
    if form == 0:
        u_ref, um_ref = oracle.run(p, phi, psi, n, xy, tcc, tc, tg)[:2]
    else:
        hi, lo, v = oracle.split_from_fp64(phi, psi)
        hi, lo, v = oracle.split_run(p, hi, lo, v, n, xy, tcc, tc, tg)
        u_ref, um_ref = oracle.split_to_fp64(hi, lo, v)
    # outer column blocks in strips of the default length, and (edge rows) of
    # 16 and 20 rows between their first and last 8 rows
    for L, edge_rows in ((8, None), (24, None), (24, "16"), (32, "20")):
        say("edge columns", form, bc, variant, L, edge_rows)
        if edge_rows:
            os.environ["WAVEHIP_EDGE_ROWS"] = edge_rows
        try:
            W = wh.Wave(dict(nx=nx, ny=ny, bcond=bc, **kw), form=form, tb_depth=8, rows_per_tile=L)
            W.set_tables(xy, tcc, tc, tg)
            W.set_state(phi, psi)
            W.step(n)
            gs = W.get_state()
            W.close()
        finally:
            os.environ.pop("WAVEHIP_EDGE_ROWS", None)
        assert bits_equal(gs[0], u_ref), first_diff(gs[0], u_ref)
        assert bits_equal(gs[1], um_ref)


def test_split_field_coefficients(wh, oracle):
    """per-cell tcc and tgamma (smooth medium): split and FP64 both exact"""
    from wavehip import workloads
    nx, ny, n = 120, 101, 20
    x, y = workloads.grid_xy(nx, ny, -1, 1, -1, 1)
    X, Y = np.meshgrid(x, y, indexing="ij")
    c = 0.75 + 0.25 * np.sin(3 * X) * np.cos(2 * Y)
    tcc = (0.1 * c) ** 2
    tc = 0.1 * c
    tg = 1e-4 * (1 + X * X)
    xy = np.ones((nx, ny), np.int16)
    phi = workloads.gaussian_bumps(x, y, [(0.1, -0.2)], 0.2)
    for bc in (2, 1):
        p = oracle.make_params(nx, ny, bcond=bc, stencil=9)
        ref = oracle.run(p, phi, phi, n, xy, tcc, tc, tg)
        hi, lo, v = oracle.split_from_fp64(phi, phi)
        hi, lo, v = oracle.split_run(p, hi, lo, v, n, xy, tcc, tc, tg)
        sref = oracle.split_to_fp64(hi, lo, v)
        for form in (0, 1):
            say("field", bc, form)
            W = wh.Wave(dict(nx=nx, ny=ny, bcond=bc, stencil=9), form=form, tb_depth=4)
            W.set_tables(xy, tcc, tc, tg)
            W.set_state(phi, phi)
            W.step(n)
            g = W.get_state()
            W.close()
            want = ref if form == 0 else sref
            assert bits_equal(g[0], want[0]) and bits_equal(g[1], want[1])


# ---------------------------------------------------------------- batches ----

def test_batch_matches_single_grids(wh, oracle):
    from wavehip import workloads
    ws = workloads.c4_instances(count=6, seed=11, nmin=40, nmax=300)
    for form in (0, 1):
        B = wh.Batch([w["params"] for w in ws], form=form)
        for k, w in enumerate(ws):
            B.load(k, w)
        say("batch form", form, B.tiling())
        for rep in range(3):  # capture once, replay
            B.step(10)
        e = B.energy()
        for k, w in enumerate(ws):
            p = oracle.make_params(w["params"]["nx"], w["params"]["ny"], bcond=w["params"]["bcond"], twospeeds=1)
            if form == 0:
                ref = oracle.run(p, w["phi"], w["psi"], 30, w["xy_in"], w["tcc"], w["tc"], w["tgamma"])
                eref = oracle.energy_total(p, ref[0], ref[1], w["xy_in"], w["tcc"], 1.0)
            else:
                hi, lo, v = oracle.split_from_fp64(w["phi"], w["psi"])
                hi, lo, v = oracle.split_run(p, hi, lo, v, 30, w["xy_in"], w["tcc"], w["tc"], w["tgamma"])
                ref = oracle.split_to_fp64(hi, lo, v)
                eref = oracle.split_energy_total(p, hi, lo, v, w["xy_in"], w["tcc"], 1.0)
            g = B.get_state(k)
            assert bits_equal(g[0], ref[0]), (form, k)
            assert bits_equal(g[1], ref[1]), (form, k)
            assert e[k] == eref
        B.close()


# ------------------------------------------------ full size, by properties ----

def test_precision_form_4096_tracks_fp64(wh):
    """4096^2 precision workload, 2,000 steps: split vs FP64 within 1e-6
    relative L2; masked cells stay 0; energies agree."""
    from wavehip import workloads
    w = workloads.c3(nsteps=2000)
    A = wh.Wave(w["params"])
    A.load(w)
    B = wh.Wave(w["params"], form=1)
    B.load(w)
    say("c3 fp64 / split", A.tiling(), B.tiling())
    e0 = A.energy()
    A.step(2000)
    B.step(2000)
    ua, _ = A.get_state()
    ub, _ = B.get_state()
    rel = np.linalg.norm(ub - ua) / np.linalg.norm(ua)
    say("rel L2", rel)
    assert rel < 1e-6
    assert np.all(ua[w["xy_in"] == 0] == 0)
    ea, eb = A.energy(), B.energy()
    assert abs(ea - eb) < 1e-6 * e0
    A.close(); B.close()


def test_tiling_invariance_large(wh):
    """bitwise identical results for every tiling at 1031 x 2050"""
    from wavehip import workloads
    w = workloads.random_case(1031, 2050, 3, bcond=1, twospeeds=1, stencil=9)
    outs = {}
    for form in (0, 1):
        for k, L in ((1, 64), (4, 64), (6, 128), (8, 256)):
            say("tiling", form, k, L)
            W = wh.Wave(w["params"], form=form, tb_depth=k, rows_per_tile=L)
            W.load(w)
            W.step(24)
            outs[(form, k, L)] = W.get_state()[0]
            W.close()
        base = outs[(form, 1, 64)]
        for key, a in outs.items():
            if key[0] == form:
                assert bits_equal(a, base), key


def test_autotune_keeps_state(wh):
    from wavehip import workloads
    w = workloads.random_case(300, 257, 5, bcond=2, neg_cells=False)
    w["phi"][w["xy_in"] == 0] = 0
    w["psi"][w["xy_in"] == 0] = 0
    W = wh.Wave(w["params"])
    W.load(w)
    W.step(5)
    before = W.get_state()
    g = W.autotune(8)
    assert g > 0
    after = W.get_state()
    assert bits_equal(before[0], after[0]) and bits_equal(before[1], after[1])
    assert W.steps_done() == 5
    W.close()


def test_batch_autotune_keeps_state_and_results(wh, oracle):
    """Tuning a batch leaves every grid's state untouched; the tuned tiling
    then steps every grid bit-identically to the oracle."""
    from wavehip import workloads
    ws = []
    for k, (nx, ny, bc) in enumerate(((90, 130, 2), (201, 77, 1), (64, 300, 0))):
        w = workloads.random_case(nx, ny, 40 + k, bcond=bc, twospeeds=1, neg_cells=False)
        w["phi"][w["xy_in"] == 0] = 0
        w["psi"][w["xy_in"] == 0] = 0
        ws.append(w)
    say("batch autotune")
    B = wh.Batch([w["params"] for w in ws])
    for k, w in enumerate(ws):
        B.load(k, w)
    B.step(3)
    before = [B.get_state(k) for k in range(len(ws))]
    assert B.autotune(8) > 0
    for k in range(len(ws)):
        a, b = B.get_state(k)
        assert bits_equal(a, before[k][0]) and bits_equal(b, before[k][1])
    B.step(20)
    for k, w in enumerate(ws):
        p = oracle.make_params(w["params"]["nx"], w["params"]["ny"], bcond=w["params"]["bcond"], twospeeds=1)
        ref = oracle.run(p, w["phi"], w["psi"], 23, w["xy_in"], w["tcc"], w["tc"], w["tgamma"])
        got = B.get_state(k)
        assert bits_equal(got[0], ref[0]) and bits_equal(got[1], ref[1]), k
    B.close()


def test_error_paths(wh):
    from wavehip import workloads
    w = workloads.random_case(20, 20, 1, neg_cells=False)
    W = wh.Wave(w["params"])
    with pytest.raises(wh.WaveError):
        W.step(1)  # no tables yet
    W.set_tables(w["xy_in"], w["tcc"], w["tc"], w["tgamma"])
    i, j = np.argwhere(w["xy_in"][1:-1, 1:-1] == 0)[0] + 1
    with pytest.raises(wh.WaveError):
        W.set_sources(dict(i=[i], j=[j], amp=np.ones((1, 3))))
    with pytest.raises(wh.WaveError):
        W.set_tiling(5, 64)
    W.close()


def test_rejected_sources_leave_state(wh, oracle):
    """A set_sources call that fails its checks (a source off the grid, or on
    a cell the step never writes) changes nothing: the grid keeps stepping
    with its old sources, bit-identical to the oracle."""
    from wavehip import workloads
    w = workloads.random_case(60, 50, 7, bcond=2, neg_cells=False)
    w["phi"][w["xy_in"] == 0] = 0
    w["psi"][w["xy_in"] == 0] = 0
    i, j = np.argwhere(w["xy_in"][1:-1, 1:-1] > 0)[5] + 1
    amp = np.sin(np.arange(1, 41) * 0.3)[None, :]
    src = dict(i=[int(i)], j=[int(j)], amp=amp)
    p = oracle.make_params(60, 50, bcond=2)
    ref = oracle.run(p, w["phi"], w["psi"], 40, w["xy_in"], w["tcc"], w["tc"], w["tgamma"], sources=src)
    zi, zj = np.argwhere(w["xy_in"][1:-1, 1:-1] == 0)[0] + 1
    for k in (1, 4):
        say("rejected sources, depth", k)
        W = wh.Wave(w["params"], tb_depth=k)
        W.load(w)
        W.set_sources(src)
        W.step(12)
        with pytest.raises(wh.WaveError):
            W.set_sources(dict(i=[70], j=[3], amp=np.ones((1, 5))))
        with pytest.raises(wh.WaveError):
            W.set_sources(dict(i=[int(zi)], j=[int(zj)], amp=np.ones((1, 5))))
        with pytest.raises(ValueError):
            W.set_sources(dict(i=[1, 2], j=[3], amp=np.ones((2, 5))))
        W.step(28)
        g = W.get_state()
        W.close()
        assert bits_equal(g[0], ref[0]) and bits_equal(g[1], ref[1]), k
