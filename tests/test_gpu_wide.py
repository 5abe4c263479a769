"""Wide tiles of the split form: the inner columns of interior strips are
computed by workgroups of four waves that exchange their column halos
through LDS at every time level.  The exchange changes no arithmetic, so
every result must equal the oracle's split restatement bit for bit and the
one-wave tiling (the default) bit for bit, for every B_COND, both
stencils, every coefficient mode (uniform, two media, per-cell tcc with and
without its low word and a tgamma field), masked cells, damping, point
sources next to wide spans, every fused depth, and in batches.  Grids are
wide along j (the contiguous index), so each interior strip holds several
wide tiles and a remainder of one-wave tiles.  (Wide tiles are an option,
WAVEHIP_WIDE=1: measured slower than one-wave tiles on every workload.)
"""
import os

import numpy as np
import pytest

from conftest import say

pytestmark = pytest.mark.gpu


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


def lens_case(nx, ny, bc, stencil, nsteps):
    """two-speed lens, damping, two soft sources and a Gaussian pulse"""
    from wavehip import workloads
    box = (-2.0, 2.0, -2.0 * ny / nx, 2.0 * ny / nx)
    x, y = workloads.grid_xy(nx, ny, *box)
    xy = workloads.mask_ext_ellipse(x, y, 0.6, 0.3)
    tcc, tc, tg = workloads.constant_media_tables(xy, 0.1, 0.1 / 1.5, 1e-4, 2e-4, 1)
    phi = workloads.gaussian_bumps(x, y, [(0.9, 0.1), (-0.4, 3.0)], 0.1)
    si, sj = workloads.xy_to_ij(-1.5, 0.0, nx, ny, *box)
    t = np.arange(1, nsteps + 1)
    amp = np.vstack([np.sin(2 * np.pi * t / 40.0), 0.5 * np.cos(2 * np.pi * t / 25.0)])
    return dict(params=dict(nx=nx, ny=ny, bcond=bc, twospeeds=1, stencil=stencil), xy_in=xy, tcc=tcc,
                tc=tc, tgamma=tg, phi=phi, psi=phi.copy(),
                sources=dict(i=[si, si], j=[sj, sj + 3], amp=amp))


def discs_case(nx, ny, bc, stencil, seed):
    """one medium, obstacle discs (cells no step writes), no damping"""
    from wavehip import workloads
    rng = np.random.default_rng(seed)
    box = (-1.0, 1.0, -1.0 * ny / nx, 1.0 * ny / nx)
    x, y = workloads.grid_xy(nx, ny, *box)
    xy = workloads.mask_discs(x, y, workloads.random_discs(rng, 12, box, 0.02, 0.08))
    tcc, tc, tg = workloads.constant_media_tables(xy, 0.1, 0.1, 0.0, 0.0, 0)
    phi = workloads.gaussian_bumps(x, y, [(0.2, -0.3), (-0.5, 1.5)], 0.05)
    phi[xy == 0] = 0.0
    return dict(params=dict(nx=nx, ny=ny, bcond=bc, twospeeds=0, stencil=stencil), xy_in=xy, tcc=tcc,
                tc=tc, tgamma=tg, phi=phi, psi=phi.copy(), sources=None)


def field_case(nx, ny, bc, stencil, lo_word, tg_field):
    """smooth per-cell medium: tcc exactly binary32 or not, tgamma field or not"""
    from wavehip import workloads
    x, y = workloads.grid_xy(nx, ny, -1, 1, -3, 3)
    X, Y = np.meshgrid(x, y, indexing="ij")
    c = 0.75 + 0.25 * np.sin(3 * X) * np.cos(2 * Y)
    tcc = (0.1 * c) ** 2
    if not lo_word:  # (5-point: the split pair of a binary32 tcc has no low word)
        tcc = tcc.astype(np.float32).astype(np.float64)
    tc = 0.1 * c
    tg = 1e-4 * (1 + X * X) if tg_field else np.zeros((nx, ny))
    xy = np.ones((nx, ny), np.int16)
    phi = workloads.gaussian_bumps(x, y, [(0.1, -0.2), (0.3, 2.0)], 0.2)
    return dict(params=dict(nx=nx, ny=ny, bcond=bc, stencil=stencil), xy_in=xy, tcc=tcc, tc=tc, tgamma=tg,
                phi=phi, psi=phi.copy(), sources=None)


def oracle_split(oracle, w, n):
    p = oracle.make_params(w["params"]["nx"], w["params"]["ny"], bcond=w["params"]["bcond"],
                           twospeeds=w["params"].get("twospeeds", 0), stencil=w["params"]["stencil"])
    hi, lo, v = oracle.split_from_fp64(w["phi"], w["psi"])
    hi, lo, v = oracle.split_run(p, hi, lo, v, n, w["xy_in"], w["tcc"], w["tc"], w["tgamma"],
                                 sources=w["sources"])
    return oracle.split_to_fp64(hi, lo, v), oracle.split_energy_total(p, hi, lo, v, w["xy_in"], w["tcc"], 1.0)


def run_gpu(wh, w, n, k, L, wide=True):
    os.environ["WAVEHIP_WIDE"] = "1" if wide else "0"
    try:
        W = wh.Wave(w["params"], form=1, tb_depth=k, rows_per_tile=L)
    finally:
        os.environ.pop("WAVEHIP_WIDE", None)
    try:
        W.load(w)
        counts = W.tile_counts(k)
        # every output cell is covered; only the one-wave tiles ending a run
        # of wide tiles may overlap it (by less than one tile's width per row)
        nx, ny = w["params"]["nx"], w["params"]["ny"]
        assert nx * ny <= counts[3] <= nx * ny + (nx * 64 if wide else 0), counts
        W.step(n)
        g = W.get_state()
        e = W.energy()
    finally:
        W.close()
    return g, e, counts


TILINGS = ((1, 64), (2, 16), (3, 7), (4, 32), (6, 24), (8, 64))


@pytest.mark.parametrize("stencil", [5, 9])
@pytest.mark.parametrize("bc", [0, 1, 2, 3, 4])
def test_wide_lens_matches_oracle_split(wh, oracle, bc, stencil):
    """two media, damping, sources (their tiles take the boundary path and
    split a wide span into one-wave tiles); every fused depth"""
    nx, ny, n = 90, 700, 25
    w = lens_case(nx, ny, bc, stencil, n)
    (u_ref, um_ref), e_ref = oracle_split(oracle, w, n)
    for k, L in TILINGS:
        say("wide lens", bc, stencil, k, L)
        (g0, g1), e, (nw, nn, ng, _) = run_gpu(wh, w, n, k, L)
        assert nw > 0, "no wide tile in the launch"
        assert bits_equal(g0, u_ref), (k, L, first_diff(g0, u_ref))
        assert bits_equal(g1, um_ref), (k, L)
        assert e == e_ref


@pytest.mark.parametrize("stencil", [5, 9])
@pytest.mark.parametrize("bc", [0, 1, 2, 3])
def test_wide_discs_match_oracle_split(wh, oracle, bc, stencil):
    """one medium, undamped, obstacle cells inside wide tiles"""
    nx, ny, n = 100, 650, 30
    w = discs_case(nx, ny, bc, stencil, 7 + bc)
    (u_ref, um_ref), e_ref = oracle_split(oracle, w, n)
    for k, L in ((1, 32), (4, 16), (8, 24), (8, 64)):
        say("wide discs", bc, stencil, k, L)
        (g0, g1), e, (nw, _, _, _) = run_gpu(wh, w, n, k, L)
        assert nw > 0
        assert bits_equal(g0, u_ref), (k, L, first_diff(g0, u_ref))
        assert bits_equal(g1, um_ref), (k, L)
        assert e == e_ref


@pytest.mark.parametrize("tg_field", [True, False])
@pytest.mark.parametrize("stencil,lo_word", [(5, True), (5, False), (9, True)])
def test_wide_field_coefficients(wh, oracle, stencil, lo_word, tg_field):
    """per-cell tcc (with and without its binary32 low word) and tgamma
    (the 9-point pair holds tcc/6, never exactly binary32 here)"""
    nx, ny, n = 80, 600, 20
    w = field_case(nx, ny, 2, stencil, lo_word, tg_field)
    (u_ref, um_ref), e_ref = oracle_split(oracle, w, n)
    for k, L in ((3, 16), (8, 32)):
        say("wide field", stencil, tg_field, lo_word, k, L)
        (g0, g1), e, (nw, _, _, _) = run_gpu(wh, w, n, k, L)
        assert nw > 0
        assert bits_equal(g0, u_ref), (k, L, first_diff(g0, u_ref))
        assert bits_equal(g1, um_ref), (k, L)
        assert e == e_ref


def test_wide_equals_one_wave_tiles_large(wh):
    """1031 x 2050, random per-cell tables and masks: wide and one-wave
    tilings give the same bits at every depth"""
    from wavehip import workloads
    w = workloads.random_case(1031, 2050, 3, bcond=1, twospeeds=1, stencil=9)
    for k, L in ((2, 64), (6, 128), (8, 64)):
        say("wide vs narrow", k, L)
        (a0, a1), ea, (nw, _, _, _) = run_gpu(wh, w, 24, k, L, wide=True)
        (b0, b1), eb, (mw, _, _, _) = run_gpu(wh, w, 24, k, L, wide=False)
        assert nw > 0 and mw == 0
        assert bits_equal(a0, b0) and bits_equal(a1, b1), (k, L)
        assert ea == eb


@pytest.mark.parametrize("stencil", [5, 9])
def test_wide_batch_matches_single_grids(wh, oracle, stencil):
    """a batch of grids of several widths (wide tiles in some of them)"""
    ws = [lens_case(70, 720, 2, stencil, 20), discs_case(60, 400, 1, stencil, 3),
          lens_case(50, 200, 3, stencil, 20), discs_case(90, 333, 0, stencil, 4)]
    for w in ws:
        w["sources"] = None  # (batches take no point sources)
    refs = [oracle_split(oracle, w, 20) for w in ws]
    for k, L in ((4, 16), (8, 32)):
        say("wide batch", stencil, k, L)
        B = wh.Batch([w["params"] for w in ws], form=1, tb_depth=k, rows_per_tile=L)
        try:
            for q, w in enumerate(ws):
                B.load(q, w)
            B.step(20)
            outs = [B.get_state(q) for q in range(len(ws))]
            es = B.energy()
        finally:
            B.close()
        for q, ((u_ref, um_ref), e_ref) in enumerate(refs):
            assert bits_equal(outs[q][0], u_ref), (q, k, L, first_diff(outs[q][0], u_ref))
            assert bits_equal(outs[q][1], um_ref), (q, k, L)
            assert es[q] == e_ref
