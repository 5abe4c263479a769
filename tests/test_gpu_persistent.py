"""Persistent launches (WAVEHIP_PERSISTENT=1, split form): the full-depth
launches of a step sequence run as epochs of one cooperative launch, with a
grid-wide barrier between epochs.  The tiles and their arithmetic are the
same as in separate launches, so results must equal the oracle's split
restatement and the separate launches bit for bit: single grids (both
stencils, every B_COND, sources, several depths and sequence lengths that
leave a shorter last launch) and batches.
"""
import os

import numpy as np
import pytest

from conftest import say

pytestmark = pytest.mark.gpu


def bits_equal(a, b):
    a, b = np.ascontiguousarray(a), np.ascontiguousarray(b)
    return np.array_equal(a.view(np.uint64), b.view(np.uint64))


@pytest.fixture(scope="module")
def wh():
    import wavehip
    wavehip.load()
    return wavehip


def lens_case(nx, ny, bc, stencil, nsteps):
    """two-speed lens, damping, a soft source and a Gaussian pulse"""
    from wavehip import workloads
    box = (-2.0, 2.0, -2.0 * ny / nx, 2.0 * ny / nx)
    x, y = workloads.grid_xy(nx, ny, *box)
    xy = workloads.mask_ext_ellipse(x, y, 0.6, 0.3)
    tcc, tc, tg = workloads.constant_media_tables(xy, 0.1, 0.1 / 1.5, 1e-4, 2e-4, 1)
    phi = workloads.gaussian_bumps(x, y, [(0.9, 0.1)], 0.1)
    si, sj = workloads.xy_to_ij(-1.5, 0.0, nx, ny, *box)
    t = np.arange(1, nsteps + 1)
    amp = np.vstack([np.sin(2 * np.pi * t / 40.0)])
    return dict(params=dict(nx=nx, ny=ny, bcond=bc, twospeeds=1, stencil=stencil), xy_in=xy, tcc=tcc,
                tc=tc, tgamma=tg, phi=phi, psi=phi.copy(), sources=dict(i=[si], j=[sj], amp=amp))


def wave(wh, params, persistent, **kw):
    os.environ["WAVEHIP_PERSISTENT"] = "1" if persistent else "0"
    try:
        return wh.Wave(params, form=1, **kw)
    finally:
        os.environ.pop("WAVEHIP_PERSISTENT", None)


@pytest.mark.parametrize("stencil", [5, 9])
@pytest.mark.parametrize("bc", [0, 1, 2, 3, 4])
def test_persistent_matches_oracle_split(wh, oracle, bc, stencil):
    nx, ny = 160, 230
    for k, L, n in ((8, 16, 43), (4, 32, 21), (2, 8, 9)):
        w = lens_case(nx, ny, bc, stencil, n)
        p = oracle.make_params(nx, ny, bcond=bc, twospeeds=1, stencil=stencil)
        hi, lo, v = oracle.split_from_fp64(w["phi"], w["psi"])
        hi, lo, v = oracle.split_run(p, hi, lo, v, n, w["xy_in"], w["tcc"], w["tc"], w["tgamma"],
                                     sources=w["sources"])
        u_ref, um_ref = oracle.split_to_fp64(hi, lo, v)
        say("persistent", bc, stencil, k, L, n)
        W = wave(wh, w["params"], True, tb_depth=k, rows_per_tile=L)
        try:
            W.load(w)
            W.prepare(n)
            W.step(n)
            g = W.get_state()
        finally:
            W.close()
        assert bits_equal(g[0], u_ref), (k, L, n)
        assert bits_equal(g[1], um_ref), (k, L, n)


def test_persistent_timed_run_and_repeats(wh):
    """a timed persistent run reports per-launch times; two sequences in a
    row equal one sequence of their total length with separate launches"""
    from wavehip import workloads
    w = workloads.random_case(300, 520, 9, bcond=2, twospeeds=1, stencil=5, neg_cells=False)
    W = wave(wh, w["params"], True, tb_depth=6, rows_per_tile=32)
    V = wave(wh, w["params"], False, tb_depth=6, rows_per_tile=32)
    try:
        W.load(w)
        V.load(w)
        W.prepare(60)
        ms, nl = W.step_timed(60)
        avg, nt = W.kernel_ms()
        assert ms > 0 and avg > 0 and nt == 10 and nl == 10
        W.step(37)
        V.step(97)
        a, b = W.get_state(), V.get_state()
        assert bits_equal(a[0], b[0]) and bits_equal(a[1], b[1])
        assert W.energy() == V.energy()
    finally:
        W.close()
        V.close()


def test_persistent_batch_matches_separate_launches(wh):
    from wavehip import workloads
    ws = workloads.c4_instances(count=12, seed=21, nmin=40, nmax=400)
    outs = {}
    for persistent in (False, True):
        os.environ["WAVEHIP_PERSISTENT"] = "1" if persistent else "0"
        try:
            B = wh.Batch([x["params"] for x in ws], form=1, tb_depth=6, rows_per_tile=32)
        finally:
            os.environ.pop("WAVEHIP_PERSISTENT", None)
        try:
            for q, x in enumerate(ws):
                B.load(q, x)
            B.step(50)
            B.step(25)
            outs[persistent] = ([B.get_state(q) for q in range(len(ws))], B.energy())
        finally:
            B.close()
    (sa, ea), (sb, eb) = outs[False], outs[True]
    for q in range(len(ws)):
        assert bits_equal(sa[q][0], sb[q][0]) and bits_equal(sa[q][1], sb[q][1]), q
    assert np.array_equal(ea, eb)


@pytest.mark.parametrize("form", [0, 1])
def test_kernel_times_survive_prepare(wh, form):
    """the per-launch durations of a timed run are the run's own: read after a
    later wv_prepare (which records the timing events to set them up) they
    equal those read at once on a twin context, to within run-to-run noise,
    and a timed run's total covers its launches"""
    from wavehip import workloads
    w = workloads.random_case(1024, 1024, 5, bcond=2, neg_cells=False)
    A = wh.Wave(w["params"], form=form, tb_depth=8, rows_per_tile=64)
    B = wh.Wave(w["params"], form=form, tb_depth=8, rows_per_tile=64)
    try:
        for V in (A, B):
            V.load(w)
            V.prepare(64)
            V.step_timed(64)  # warm
        ms_a, nl_a = A.step_timed(64)
        avg_a, n_a = A.kernel_ms()
        _, nl_b = B.step_timed(64)
        B.prepare(40)
        avg_b, n_b = B.kernel_ms()
        # (every launch of a 64-step run has the full depth: 64 = 8 x 8, or 64 x 1
        # where the FP64 form steps one level a launch)
        assert n_a == nl_a and n_b == nl_b and nl_a == nl_b
        assert 0.5 * avg_a < avg_b < 2.0 * avg_a, (avg_a, avg_b)
        assert n_a * avg_a <= ms_a * 1.05
    finally:
        A.close()
        B.close()
