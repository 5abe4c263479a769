"""Billiard domains and per-cell media models built on the GPU
(wv_build_domain_mask, wv_set_ior_tables, the rotating wells through
wv_rotating_wells + wv_set_wells_tables) against the reference's own
functions (tests/golden/domain_cases.npz, ior_cases.npz) and the oracle.

Masks are exact.  Media tables are exact except IOR_MANDELBROT, whose log is
the device library's: the escape test itself is exact, and speed = 1 +
MANDEL_IOR_SCALE*log(1 + |z|^2/MANDELLIMIT), with |z|^2 < 1e8 after escape
(log < 17) and |MANDEL_IOR_SCALE| = 0.05, is off by at most 0.05*17*2 ulp +
1 ulp < 1e-15 absolute (tcc within COURANT^2*1e-15, tc = sqrt(tcc) >=
0.1*COURANT within COURANT*1e-14); and the rotating wells, whose exp is the
device library's (relative 1e-14).  The
installed media step exactly as the oracle steps with the same tables."""
import os

import numpy as np
import pytest

from conftest import say
from gen_inputs import DOMAIN_CODES, IOR_CODES, IOR_PARAMS, MASK_BOX, ROT_ANGLES, domain_sets, ior_mask, mask_circles

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CIRCLE_CODES = (20, 201, 202)
# the box of the checked-in wave_billiard.c (64-67)
BOX_REF = (-0.75, 3.25, -1.197916667, 1.197916667)


@pytest.fixture(scope="module")
def wh():
    import wavehip
    wavehip.load()
    return wavehip


def bits_equal(a, b):
    return np.array_equal(np.asarray(a).view(np.uint64), np.asarray(b).view(np.uint64))


def mandel_close(got, want, courant):
    """IOR_MANDELBROT tables within the bound of the device log (module text)"""
    return (np.all(np.abs(got[0] - want[0]) <= courant * courant * 1e-15) and
            np.all(np.abs(got[1] - want[1]) <= courant * 1e-14) and bits_equal(got[2], want[2]))


def test_domain_masks_match_reference_fixture(wh, oracle):
    """the builds' macros: their parameter block (oracle.DOMAIN_DEFAULTS)
    with the set's own values"""
    g = np.load(os.path.join(GOLD, "domain_cases.npz"))
    circ = mask_circles()
    for tag, (nx, ny, kw0) in domain_sets().items():
        kw = dict(oracle.DOMAIN_DEFAULTS, **kw0)
        W = wh.Wave(dict(nx=nx, ny=ny))
        try:
            for dom in DOMAIN_CODES:
                say("domain fixture", tag, dom)
                got = W.build_domain_mask(dom, MASK_BOX, circles=circ if dom in CIRCLE_CODES else None, **kw)
                assert np.array_equal(got, g["%s/%d" % (tag, dom)]), (tag, dom)
        finally:
            W.close()


@pytest.mark.parametrize("shape", [(1280, 720), (333, 257)])
def test_domain_masks_match_oracle(wh, oracle, shape):
    """the checked-in configuration's box and macros (MANDELLEVEL 1000, NPOLY
    9, ...), and a second parameter set; a circle list longer than a wave"""
    nx, ny = shape
    rng = np.random.default_rng(31)
    n = 90
    circ = oracle.circle_list(rng.uniform(-0.7, 3.2, n), rng.uniform(-1.2, 1.2, n), rng.uniform(0.01, 0.2, n),
                              (rng.uniform(0, 1, n) > 0.2).astype(float))
    sets = [dict(wh.DOMAIN_DEFAULTS), dict(wh.DOMAIN_DEFAULTS, lam=0.9, mu=0.25, npoly=5, apoly=0.3, mdepth=5,
                                                 julia_scale=1.2, mandellevel=300, ngridx=6, ngridy=9)]
    W = wh.Wave(dict(nx=nx, ny=ny))
    try:
        for q, kw in enumerate(sets):
            P = oracle.domain_params(**kw)
            for dom in DOMAIN_CODES:
                say("domain", nx, ny, q, dom)
                c = circ if dom in CIRCLE_CODES else None
                got = W.build_domain_mask(dom, BOX_REF, circles=c, **kw)
                assert np.array_equal(got, oracle.domain(nx, ny, BOX_REF, dom, P, c)), (q, dom)
    finally:
        W.close()


def test_domain_mask_errors(wh):
    W = wh.Wave(dict(nx=40, ny=30))
    try:
        for dom, kw in ((53, {}), (8, dict(npoly=0)), (15, dict(mratio=3, mdepth=25)), (25, dict(julia_scale=0.0)),
                        (93, dict(ngridx=0)), (947, dict(ngridy=0)), (991, dict(ngridx=0)), (43, dict(mu=0.0)),
                        (46, dict(lam=1e-5)), (55, dict(lam=0.0)), (794, dict(julia_scale=0.0)),
                        (22, dict(mratio=3, mdepth=25)), (59, dict(npoly=0))):
            with pytest.raises(wh.WaveError):
                W.build_domain_mask(dom, MASK_BOX, **kw)
        with pytest.raises(wh.WaveError):
            W.build_domain_mask(3, (1.0, 1.0, 0.0, 1.0))
    finally:
        W.close()


def test_ior_tables_match_reference_fixture(wh):
    g = np.load(os.path.join(GOLD, "ior_cases.npz"))
    xy = ior_mask()
    for m in IOR_CODES + (0, 0):
        ts = m % 2
        W = wh.Wave(dict(nx=97, ny=61, twospeeds=ts))
        try:
            say("ior fixture", m)
            got = W.set_ior_tables(xy, m, MASK_BOX, **IOR_PARAMS)
        finally:
            W.close()
        key = "m%d" % m if m else "const_ts%d" % ts
        want = [g["%s/%s" % (key, n)] for n in ("tcc", "tc", "tgamma")]
        if m == 1:
            inside = want[0] == IOR_PARAMS["courant"] ** 2
            assert np.array_equal(got[0] == IOR_PARAMS["courant"] ** 2, inside)
            assert mandel_close(got, want, IOR_PARAMS["courant"])
        else:
            for name, a, b in zip(("tcc", "tc", "tgamma"), got, want):
                assert bits_equal(a, b), (m, name)
    # TWOSPEEDS = 1 for the constant branch too
    W = wh.Wave(dict(nx=97, ny=61, twospeeds=1))
    try:
        got = W.set_ior_tables(xy, 0, MASK_BOX, **IOR_PARAMS)
    finally:
        W.close()
    for name, a in zip(("tcc", "tc", "tgamma"), got):
        assert bits_equal(a, g["const_ts1/%s" % name]), name


@pytest.mark.parametrize("m,dom", [(17, 74), (14, 73), (16, 524), (182, 999), (1, 999), (12, 68)])
def test_ior_tables_install_and_step(wh, oracle, m, dom):
    """a model on the domain it is meant for (Michelson, tree, coated guide,
    Luneburg lens, Mandelbrot, concave lens), 640x360: tables as the oracle
    builds them (Mandelbrot to 1e-15), and 30 steps on them equal the
    oracle's 30 steps on the same tables, in both forms"""
    nx, ny = 640, 360
    kw = dict(wh.DOMAIN_DEFAULTS, lam=0.8, mu=0.3, wall_width=0.1)
    box = MASK_BOX
    ip = dict(IOR_PARAMS, mandellevel=500)
    for form in (0, 1):
        say("ior step", m, dom, form)
        W = wh.Wave(dict(nx=nx, ny=ny, bcond=2, twospeeds=1), form=form, tb_depth=4)
        try:
            xy = W.build_domain_mask(dom, box, **kw)
            got = W.set_ior_tables(xy, m, box, **ip)
            p = oracle.make_params(nx, ny, bcond=2, twospeeds=1)
            ref = oracle.ior_tables(p, xy, m, box, oracle.IorParams(**ip))
            if m == 1:
                assert mandel_close(got, ref, ip["courant"])
            else:
                for a, b in zip(got, ref):
                    assert bits_equal(a, b)
            x = box[0] + np.arange(nx) * (box[1] - box[0]) / nx
            y = box[2] + np.arange(ny) * (box[3] - box[2]) / ny
            X, Y = np.meshgrid(x, y, indexing="ij")
            phi = np.exp(-((X + 1.2) ** 2 + (Y - 0.1) ** 2) / 0.02)
            W.set_state(phi, phi)
            W.step(30)
            u = W.get_state()
        finally:
            W.close()
        if form == 0:
            r = oracle.run(p, phi, phi, 30, xy, got[0], got[1], got[2])
            assert bits_equal(u[0], r[0]) and bits_equal(u[1], r[1])
        else:
            hi, lo, v = oracle.split_from_fp64(phi, phi)
            hi, lo, v = oracle.split_run(p, hi, lo, v, 30, xy, got[0], got[1], got[2])
            ur, umr = oracle.split_to_fp64(hi, lo, v)
            assert bits_equal(u[0], ur) and bits_equal(u[1], umr)


def test_ior_errors(wh):
    W = wh.Wave(dict(nx=40, ny=30))
    xy = np.ones((40, 30), np.int16)
    try:
        for m in (3, 4, 5, 6, 7, 8, 9, 20, 77, 101):
            with pytest.raises(wh.WaveError):
                W.set_ior_tables(xy, m, MASK_BOX)
        with pytest.raises(wh.WaveError):
            W.set_ior_tables(xy, 183, MASK_BOX, mdepth=0)
    finally:
        W.close()


def test_rotating_wells_match_reference_fixture(wh):
    """the REFRESH_IOR rebuild on the GPU: centres of wv_rotating_wells, sum of
    wv_set_wells_tables, against init_ior_2d at each angle"""
    g = np.load(os.path.join(GOLD, "ior_cases.npz"))
    for tag, large, ts in (("ior_rot", 0, 0), ("ior_rot_large", 1, 1)):
        xr = g[tag + "/xy_in"]
        W = wh.Wave(dict(nx=40, ny=30, twospeeds=ts))
        try:
            for q, ang in enumerate(ROT_ANGLES):
                say("rotating wells", tag, ang)
                xc, yc, s = wh.rotating_wells(6, 5, large, ang, MASK_BOX)
                tcc, tc, _ = W.set_wells_tables(xr, xc, yc, None, s, 0, 0.1, 0.05, 0.0, MASK_BOX)
                assert np.allclose(tcc, g["%s/%d/tcc" % (tag, q)], rtol=1e-14, atol=0), (tag, q)
                assert np.allclose(tc, g["%s/%d/tc" % (tag, q)], rtol=1e-14, atol=0), (tag, q)
        finally:
            W.close()


def test_comparison_mask_matches_reference_and_oracle(wh, oracle):
    """wv_build_comparison_mask against the reference's xy_in_billiard with
    COMPARISON (dom_cmp_a, dom_cmp_b) and the oracle on a larger grid with
    other domain pairs (iterated and polygon domains in either half)"""
    from gen_inputs import COMPARISON_BUILDS, mask_circles_b
    g = np.load(os.path.join(GOLD, "domain_cases.npz"))
    kw = dict(oracle.DOMAIN_DEFAULTS, **domain_sets()["dom_a"][2])
    W = wh.Wave(dict(nx=97, ny=61))
    try:
        for tag, (d, db) in COMPARISON_BUILDS.items():
            say("comparison fixture", tag)
            got = W.build_comparison_mask(d, db, MASK_BOX, circles=mask_circles(), circles_b=mask_circles_b(), **kw)
            assert np.array_equal(got, g[tag + "/mask"]), tag
    finally:
        W.close()
    nx, ny = 640, 361
    P = oracle.domain_params(**kw)
    W = wh.Wave(dict(nx=nx, ny=ny))
    try:
        for d, db in ((24, 8), (32, 25), (20, 74), (3, 199)):
            say("comparison", d, db)
            got = W.build_comparison_mask(d, db, BOX_REF, circles=mask_circles(), circles_b=mask_circles_b(), **kw)
            ref = oracle.comparison_mask(nx, ny, BOX_REF, d, db, P, mask_circles(), mask_circles_b())
            assert np.array_equal(got, ref), (d, db)
        with pytest.raises(wh.WaveError):
            W.build_comparison_mask(3, 53, BOX_REF, **kw)
    finally:
        W.close()


def test_batch_builders_per_grid(wh, oracle):
    """the builders on grids of a batch (three sizes): each grid's mask and
    media as the oracle builds them, and 20 split-form steps of the batch
    equal the oracle's split run of each grid on the same tables"""
    box = MASK_BOX
    sizes = [(130, 77), (64, 200), (97, 61)]
    params = [dict(nx=nx, ny=ny, bcond=2, twospeeds=1) for nx, ny in sizes]
    kw = dict(wh.DOMAIN_DEFAULTS, lam=0.8, mu=0.3, wall_width=0.1)
    P = oracle.domain_params(**kw)
    B = wh.Batch(params, form=1, tb_depth=4)
    tables = []
    try:
        for k, ((nx, ny), dom) in enumerate(zip(sizes, (999, 74, 3))):
            say("batch builders", k, dom)
            xy = B.build_domain_mask(k, dom, box, **kw)
            assert np.array_equal(xy, oracle.domain(nx, ny, box, dom, P))
            p = oracle.make_params(nx, ny, bcond=2, twospeeds=1)
            if k == 1:
                got = B.set_ior_tables(k, xy, 17, box, **IOR_PARAMS)
                ref = oracle.ior_tables(p, xy, 17, box, oracle.IorParams(**IOR_PARAMS))
                assert all(bits_equal(a, b) for a, b in zip(got, ref))
            elif k == 0:
                got = B.set_ior_tables(k, xy, 182, box, **IOR_PARAMS)
                ref = oracle.ior_tables(p, xy, 182, box, oracle.IorParams(**IOR_PARAMS))
                assert all(bits_equal(a, b) for a, b in zip(got, ref))
            else:
                rng = np.random.default_rng(4)
                xc, yc = rng.uniform(-2, 2, 30), rng.uniform(-1.1, 1.1, 30)
                got = B.set_wells_tables(k, xy, xc, yc, None, 0.05, 1, 0.1, 0.05, 1e-4, box)
                ref = oracle.wells_tables(p, xy, xc, yc, None, 0.05, 1, 0.1, 0.05, 1e-4, box)
                assert all(np.allclose(a, b, rtol=1e-14, atol=0) for a, b in zip(got, ref))
            x = box[0] + np.arange(nx) * (box[1] - box[0]) / nx
            y = box[2] + np.arange(ny) * (box[3] - box[2]) / ny
            X, Y = np.meshgrid(x, y, indexing="ij")
            phi = np.exp(-((X + 0.7) ** 2 + (Y - 0.2) ** 2) / 0.05)
            B.set_state(k, phi, phi)
            tables.append((p, xy, got, phi))
        with pytest.raises(wh.WaveError):
            B.set_ior_tables(2, tables[2][1], 5, box)
        B.step(20)
        for k, (p, xy, t, phi) in enumerate(tables):
            u = B.get_state(k)
            hi, lo, v = oracle.split_from_fp64(phi, phi)
            hi, lo, v = oracle.split_run(p, hi, lo, v, 20, xy, t[0], t[1], t[2])
            ur, umr = oracle.split_to_fp64(hi, lo, v)
            assert bits_equal(u[0], ur) and bits_equal(u[1], umr), k
    finally:
        B.close()


# ---- the list-driven domains ---------------------------------------------------

def test_list_domain_masks_match_reference_fixture(wh, oracle):
    """every list-driven code on the three builds, with list_data's lists and
    the circle list: the reference's masks; a closed-form code through the
    same entry point equals wv_build_domain_mask's fixture"""
    from gen_inputs import LIST_DOMAIN_CODES, list_data
    g = np.load(os.path.join(GOLD, "domain_cases.npz"))
    for tag, (nx, ny, kw0) in domain_sets().items():
        kw = dict(oracle.DOMAIN_DEFAULTS, **kw0)
        lists = dict(list_data(kw["npoly"]), circles=mask_circles())
        W = wh.Wave(dict(nx=nx, ny=ny))
        try:
            for dom in LIST_DOMAIN_CODES:
                say("list domain fixture", tag, dom)
                got = W.build_list_domain_mask(dom, MASK_BOX, lists=lists, **kw)
                assert np.array_equal(got, g["%s/list/%d" % (tag, dom)]), (tag, dom)
            got = W.build_list_domain_mask(20, MASK_BOX, lists=lists, **kw)
            assert np.array_equal(got, g["%s/20" % tag])
        finally:
            W.close()


def long_lists(rng, npoly):
    """lists longer than a wave: 300 rectangles (every seventh reversed), 150
    rotated ones, 90 polygons of 3..16 sides, 70 circles, list_data's
    polyline"""
    from gen_inputs import list_data
    nr, nrr, npg = 300, 150, 90
    x1, y1 = rng.uniform(-1.9, 1.6, nr), rng.uniform(-1.0, 0.9, nr)
    rects = np.vstack([x1, y1, x1 + rng.uniform(0.02, 0.3, nr), y1 + rng.uniform(0.02, 0.3, nr)])
    rects[:, ::7] = rects[[2, 3, 0, 1]][:, ::7]
    ax, ay, th = rng.uniform(-1.9, 1.9, nrr), rng.uniform(-1.0, 1.0, nrr), rng.uniform(0, 6.3, nrr)
    ln = rng.uniform(0.05, 0.6, nrr)
    rrot = np.vstack([ax, ay, ax + ln * np.cos(th), ay + ln * np.sin(th), rng.uniform(0.01, 0.1, nrr)])
    pg = np.vstack([rng.uniform(-1.8, 1.8, npg), rng.uniform(-1.0, 1.0, npg), rng.uniform(0.03, 0.3, npg),
                    rng.uniform(0, 4, npg), rng.integers(3, 17, npg).astype(float),
                    (rng.uniform(0, 1, npg) > 0.2).astype(float)])
    c = np.vstack([rng.uniform(-1.9, 1.9, 70), rng.uniform(-1.1, 1.1, 70), np.zeros(70), np.ones(70)])
    return dict(rects=rects, rectrot=rrot, polygons=pg, circles=c, polyline=list_data(npoly)["polyline"])


@pytest.mark.parametrize("shape", [(1280, 720), (333, 257)])
def test_list_domain_masks_match_oracle(wh, oracle, shape):
    """the checked-in configuration's macros and a second set, with lists
    longer than a wave"""
    from gen_inputs import LIST_DOMAIN_CODES
    nx, ny = shape
    lists = long_lists(np.random.default_rng(41), 9)
    sets = [dict(wh.DOMAIN_DEFAULTS), dict(wh.DOMAIN_DEFAULTS, lam=0.9, mu=0.05, npoly=5, mdepth=3, ngridx=7, ngridy=5)]
    W = wh.Wave(dict(nx=nx, ny=ny))
    try:
        for q, kw in enumerate(sets):
            P = oracle.domain_params(**kw)
            for dom in LIST_DOMAIN_CODES:
                say("list domain", nx, ny, q, dom)
                got = W.build_list_domain_mask(dom, BOX_REF, lists=lists, **kw)
                assert np.array_equal(got, oracle.list_domain(nx, ny, BOX_REF, dom, P, **lists)), (q, dom)
    finally:
        W.close()


def test_list_domain_mask_errors(wh):
    from gen_inputs import list_data
    L = list_data(6)
    no_sides = dict(L, polygons=np.vstack([L["polygons"][:4], np.zeros((1, 12)), L["polygons"][5:]]))
    W = wh.Wave(dict(nx=40, ny=30))
    try:
        for dom, lists, kw in ((42, dict(L, polyline=L["polyline"][:, :6]), dict(npoly=6)),
                               (421, dict(L, polyline=L["polyline"][:, :9]), dict(npoly=6)),
                               (41, L, dict(mdepth=16)), (40, no_sides, {}), (95, L, dict(ngridx=0)),
                               (92, L, {})):
            say("list domain error", dom)
            with pytest.raises(wh.WaveError):
                W.build_list_domain_mask(dom, MASK_BOX, lists=lists, **kw)
        with pytest.raises(ValueError):
            W.build_list_domain_mask(53, MASK_BOX, lists=dict(rects=np.zeros((3, 4))))
    finally:
        W.close()


def test_batch_list_domain_mask(wh, oracle):
    """the batch form on the second of two grids"""
    lists = long_lists(np.random.default_rng(5), 9)
    B = wh.Batch([dict(nx=50, ny=40), dict(nx=211, ny=97)])
    try:
        for dom in (53, 40, 941, 42):
            say("batch list domain", dom)
            got = B.build_list_domain_mask(1, dom, MASK_BOX, lists=lists)
            assert np.array_equal(got, oracle.list_domain(211, 97, MASK_BOX, dom,
                                                          oracle.domain_params(**wh.DOMAIN_DEFAULTS), **lists))
    finally:
        B.close()
