"""The neighbour-table stepper on the GPU (evolve_wave_half_new,
wave_3d.c:345-597, with init_laplacian_coords / compute_laplacian,
sub_wave.c:11555-11963), through the C ABI (wv_set_laplacian).

* against the reference's own stepper: tests/golden/table_cases.npz (every
  B_COND, BC_LSHAPE, OSCILLATE_TOPBOT, TWOSPEEDS, the oscillating left edge
  and FLOOR), with the reference's exported table;
* against the oracle (pinned to those fixtures) on larger grids, with the
  L-shaped billiard, drive, FLOOR and point sources, at several strip heights.
"""
import os

import numpy as np
import pytest

from conftest import say
from gen_inputs import EXT_VMAX, KAPPAS, TABLE_STEPS, table_cases, table_inputs

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
def tbl():
    return np.load(os.path.join(GOLD, "table_cases.npz"))


@pytest.mark.parametrize("split_at", [TABLE_STEPS, 3])
@pytest.mark.parametrize("case", table_cases(), ids=lambda c: c[0])
def test_table_stepper_matches_reference_fixture(wh, tbl, case, split_at):
    tag, nx, ny, bc, ts, tb, ex = case
    say("table fixture", tag, split_at)
    d = table_inputs(tag, nx, ny, ts, ex)
    W = wh.Wave(dict(nx=nx, ny=ny, bcond=bc, twospeeds=ts, oscillate_topbot=tb, tb_shift=nx, **KAPPAS))
    W.set_tables(d["xy_in"], d["tcc"], d["tc"], d["tgamma"])
    W.set_laplacian((tbl[tag + "/nneighb"], tbl[tag + "/nghb"]))
    if ex:
        cells = d["drive_cells"]
        W.set_drive(dict(i=[c[0] for c in cells], j=[c[1] for c in cells], val=tbl[tag + "/drive_val"]))
        W.set_floor(EXT_VMAX)
    W.set_state(d["phi"], d["psi"], d["tmp"])
    W.step(split_at)
    W.step(TABLE_STEPS - split_at)
    phi, psi, tmp = W.get_state(three=True)
    W.close()
    assert bits_equal(phi, tbl[tag + "/phi"]), first_diff(phi, tbl[tag + "/phi"])
    assert bits_equal(psi, tbl[tag + "/psi"])
    assert bits_equal(tmp, tbl[tag + "/tmp"])


@pytest.mark.parametrize("bc,ts,tb", [(0, 0, 0), (1, 0, 0), (2, 1, 0), (3, 0, 1), (10, 0, 0), (10, 0, 1)])
def test_table_stepper_matches_oracle(wh, oracle, bc, ts, tb):
    from wavehip import workloads
    nx, ny, n = 150, 97, 23
    w = workloads.random_case(nx, ny, 80 + bc + 3 * tb, bcond=bc if bc < 10 else 2, twospeeds=ts, kappas=True)
    xy = w["xy_in"]
    if bc == 10:
        # the L-shaped billiard, with a few cells of every mask value inside it
        x, y = workloads.grid_xy(nx, ny, -2.0, 2.0, -1.125, 1.125)
        L = workloads.mask_lshape(x, y, 1.5)
        xy = np.where(L == 1, np.where(xy == 0, 1, xy), 0).astype(np.int16)
        lsh = workloads.lshape_lines(nx, ny, 1.5, -2.0, 2.0, -1.125, 1.125)
    else:
        lsh = None
    if ts:
        xy = np.where(xy == 0, 1, xy).astype(np.int16)  # the reference's Laplacian is unset at xy_in == 0
    table = workloads.laplacian_table(nx, ny, bc, lsh)
    rng = np.random.default_rng(bc + 10 * tb)
    phi, psi, tmp = (rng.uniform(-0.4, 0.4, (nx, ny)) for _ in range(3))
    t = np.arange(1, n + 1)
    dcells = [(0, j) for j in range(1, ny - 1)] + [(0, 0), (0, ny - 1)]
    drive = dict(i=[c[0] for c in dcells], j=[c[1] for c in dcells],
                 val=np.array([0.5 * np.cos(0.3 * t + 0.01 * j) for (_, j) in dcells]))
    act = np.argwhere(xy[1:-1, 1:-1] != 0) + 1
    (si0, sj0), (si1, sj1) = act[len(act) // 3], act[2 * len(act) // 3]
    src = dict(i=[si0, si1], j=[sj0, sj1], amp=np.vstack([0.2 * np.sin(t / 5.0), 0.1 * np.cos(t / 3.0)]))
    p = oracle.make_params(nx, ny, bcond=bc, twospeeds=ts, oscillate_topbot=tb, tb_shift=nx, **KAPPAS)
    ref = oracle.run_table(p, phi, psi, n, xy, w["tcc"], w["tc"], w["tgamma"], table[0], table[1], tmp=tmp,
                           sources=src, drive=drive, vmax=0.35)
    assert (np.abs(ref[0]) == 0.35).any()  # the clamp acts
    params = dict(w["params"], bcond=bc, twospeeds=ts, oscillate_topbot=tb, tb_shift=nx)
    for L in (8, 16, 64):
        say("table vs oracle", bc, ts, tb, L)
        W = wh.Wave(params, rows_per_tile=L)
        W.set_tables(xy, w["tcc"], w["tc"], w["tgamma"])
        W.set_laplacian(table)
        W.set_drive(drive)
        W.set_floor(0.35)
        W.set_sources(src)
        W.set_state(phi, psi, tmp)
        W.step(5)
        W.step(n - 5)
        g = W.get_state(three=True)
        W.close()
        for a, b in zip(g, ref):
            assert bits_equal(a, b), (L, first_diff(a, b))


def test_table_stepper_off_again(wh, oracle):
    """wv_set_laplacian(NULL) returns to the plain stepper (fused launches)."""
    from wavehip import workloads
    w = workloads.random_case(90, 70, 5, bcond=0)
    w["phi"][w["xy_in"] <= 0] = 0
    w["psi"][w["xy_in"] <= 0] = 0
    W = wh.Wave(w["params"])
    W.set_tables(w["xy_in"], w["tcc"], w["tc"], w["tgamma"])
    W.set_laplacian(workloads.laplacian_table(90, 70, 0))
    W.set_laplacian(None)
    W.set_state(w["phi"], w["psi"])
    W.step(17)
    g = W.get_state()
    W.close()
    p = oracle.make_params(90, 70, bcond=0)
    ref = oracle.run(p, w["phi"], w["psi"], 17, w["xy_in"], w["tcc"], w["tc"], w["tgamma"])
    assert bits_equal(g[0], ref[0]) and bits_equal(g[1], ref[1])


def test_table_argument_checks(wh):
    from wavehip import workloads
    w = workloads.random_case(40, 30, 3)
    nn, nb = workloads.laplacian_table(40, 30, 0)
    W = wh.Wave(w["params"])
    W.set_tables(w["xy_in"], w["tcc"], w["tc"], w["tgamma"])
    bad = nn.copy()
    bad[0, 5] = 5
    with pytest.raises(wh.WaveError):
        W.set_laplacian((bad, nb))  # more than four neighbours
    far = nb.copy()
    far[0, 5, 0] = 40 * 30
    with pytest.raises(wh.WaveError):
        W.set_laplacian((nn, far))  # neighbour outside the grid
    nbp = dict(i=[5], j=[5], nbi1=[5], nbj1=[6], nbi2=[5], nbj2=[4], nx=[0.5], ny=[0.5])
    W.set_neumann(nbp)
    with pytest.raises(wh.WaveError):
        W.set_laplacian((nn, nb))  # no Neumann points in the table stepper
    W.close()
    # BC_LSHAPE needs the table; the split form has no table stepper
    with pytest.raises(wh.WaveError):
        wh.Wave(dict(w["params"], bcond=10), form=1)
    L = wh.Wave(dict(w["params"], bcond=10))
    L.set_tables(w["xy_in"], w["tcc"], w["tc"], w["tgamma"])
    L.set_state(w["phi"], w["psi"])
    with pytest.raises(wh.WaveError):
        L.step(1)
    L.close()
    S = wh.Wave(w["params"], form=1)
    S.set_tables(w["xy_in"], w["tcc"], w["tc"], w["tgamma"])
    with pytest.raises(wh.WaveError):
        S.set_laplacian((nn, nb))
    S.close()
