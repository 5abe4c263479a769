"""The CPU oracle against the reference's own stepper.

Golden vectors (tests/golden/) were produced by tests/golden/make_golden.py
from the reference compiled from its sources (wave_billiard.c:302-631,
wave_3d.c:345-597 and 600-1007, sub_wave.c:11555-11963,
wave_common.c:453-472).  Comparisons are bit-exact.
"""
import hashlib
import json
import os

import numpy as np
import pytest

from gen_inputs import (EXT_STEPS, EXT_VMAX, KAPPAS, TABLE_STEPS, case_inputs, cmp_cases, ext_cases, ext_inputs,
                        small_cases, table_cases, table_inputs)

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def bits_equal(a, b):
    return np.array_equal(np.asarray(a).view(np.uint64), np.asarray(b).view(np.uint64))


@pytest.fixture(scope="module")
def small():
    return np.load(os.path.join(GOLD, "small_cases.npz"))


@pytest.mark.parametrize("case", small_cases(), ids=lambda c: c[0])
def test_oracle_matches_reference_fixture(oracle, small, case):
    tag, nx, ny, bc, ts, tb, tbs = case
    d = case_inputs(tag, nx, ny)
    p = oracle.make_params(nx, ny, bcond=bc, twospeeds=ts, oscillate_topbot=tb, tb_shift=tbs, **KAPPAS)
    phi, psi, tmp = oracle.run(p, d["phi"], d["psi"], 7, d["xy_in"], d["tcc"], d["tc"], d["tgamma"],
                               tmp=d["tmp"])
    assert bits_equal(phi, small[tag + "/phi"])
    assert bits_equal(psi, small[tag + "/psi"])
    assert bits_equal(tmp, small[tag + "/tmp"])


def test_flat_signature_matches_rowptr(oracle, small):
    """evolve_wave_flat (wave_3d.c:1010 order) gives the same three buffers."""
    tag, nx, ny, bc, ts, tb, tbs = small_cases()[2]
    d = case_inputs(tag, nx, ny)
    p = oracle.make_params(nx, ny, bcond=bc, twospeeds=ts, **KAPPAS)
    a, b, c = d["phi"].copy(), d["psi"].copy(), d["tmp"].copy()
    for _ in range(2):
        oracle.evolve_wave_flat(p, a, b, c, d["xy_in"], d["tcc"], d["tc"], d["tgamma"])
    # 6 steps with the rotation = run(6)
    phi, psi, tmp = oracle.run(p, d["phi"], d["psi"], 6, d["xy_in"], d["tcc"], d["tc"], d["tgamma"],
                               tmp=d["tmp"])
    assert bits_equal(a, phi) and bits_equal(b, psi) and bits_equal(c, tmp)


def test_fixture_detects_planted_mistakes(oracle, small):
    """A plausible mistake in the oracle (sign of KAPPA, a transposed
    B_COND, a dropped damping term) must fail the golden comparison."""
    tag, nx, ny, bc, ts, tb, tbs = ("row_24x17_bc3_ts1", 24, 17, 3, 1, 0, 24)
    d = case_inputs(tag, nx, ny)
    mutants = [
        dict(KAPPAS, kappa=-KAPPAS["kappa"]),
        dict(KAPPAS, gamma_sides=0.0),
        dict(KAPPAS, kappa_topbot=KAPPAS["kappa_sides"], kappa_sides=KAPPAS["kappa_topbot"]),
    ]
    for kw in mutants:
        p = oracle.make_params(nx, ny, bcond=bc, twospeeds=ts, **kw)
        phi, _, _ = oracle.run(p, d["phi"], d["psi"], 7, d["xy_in"], d["tcc"], d["tc"], d["tgamma"],
                               tmp=d["tmp"])
        assert not bits_equal(phi, small[tag + "/phi"])
    for wrong_bc in (0, 1, 2):
        p = oracle.make_params(nx, ny, bcond=wrong_bc, twospeeds=ts, **KAPPAS)
        phi, _, _ = oracle.run(p, d["phi"], d["psi"], 7, d["xy_in"], d["tcc"], d["tc"], d["tgamma"],
                               tmp=d["tmp"])
        assert not bits_equal(phi, small[tag + "/phi"])


@pytest.fixture(scope="module")
def ext():
    return np.load(os.path.join(GOLD, "ext_cases.npz"))


def ext_run(oracle, case, ext, **mut):
    tag, nx, ny, bc, ts = case
    d = ext_inputs(tag, nx, ny)
    p = oracle.make_params(nx, ny, bcond=bc, twospeeds=ts, **KAPPAS)
    cells = d["drive_cells"]
    drive = dict(i=[c[0] for c in cells], j=[c[1] for c in cells], val=ext[tag + "/drive_val"])
    kw = dict(neumann=d["neumann"], drive=drive, vmax=EXT_VMAX)
    kw.update(mut)
    return oracle.run_x(p, d["phi"], d["psi"], EXT_STEPS, d["xy_in"], d["tcc"], d["tc"], d["tgamma"],
                        tmp=d["tmp"], **kw)


@pytest.mark.parametrize("case", ext_cases(), ids=lambda c: c[0])
def test_oracle_extras_match_reference_fixture(oracle, ext, case):
    """Neumann ghost points (wave_billiard.c:379-389), the oscillating left
    edge and its corners (397, 585-590) and FLOOR (592-609), bit for bit."""
    tag = case[0]
    phi, psi, tmp = ext_run(oracle, case, ext)
    assert bits_equal(phi, ext[tag + "/phi"])
    assert bits_equal(psi, ext[tag + "/psi"])
    assert bits_equal(tmp, ext[tag + "/tmp"])


def test_extras_fixture_detects_planted_mistakes(oracle, ext):
    """Dropping any one of the three features, or swapping the Neumann
    weights, fails the comparison."""
    case = ext_cases()[2]
    tag, nx, ny = case[:3]
    d = ext_inputs(tag, nx, ny)
    swapped = dict(d["neumann"], nx=d["neumann"]["ny"], ny=d["neumann"]["nx"])
    for mut in (dict(neumann=None), dict(drive=None), dict(vmax=0.0), dict(neumann=swapped)):
        phi, _, _ = ext_run(oracle, case, ext, **mut)
        assert not bits_equal(phi, ext[tag + "/phi"]), mut


@pytest.fixture(scope="module")
def tbl():
    return np.load(os.path.join(GOLD, "table_cases.npz"))


def table_run(oracle, case, tbl, table=None, **kw):
    tag, nx, ny, bc, ts, tb, ex = case
    d = table_inputs(tag, nx, ny, ts, ex)
    p = oracle.make_params(nx, ny, bcond=kw.pop("bcond", bc), twospeeds=ts, oscillate_topbot=tb, tb_shift=nx,
                           **KAPPAS)
    if table is None:
        table = oracle.laplacian_table(p, tuple(tbl[tag + "/lsh"]))
    if ex:
        cells = d["drive_cells"]
        kw.setdefault("drive", dict(i=[c[0] for c in cells], j=[c[1] for c in cells], val=tbl[tag + "/drive_val"]))
        kw.setdefault("vmax", EXT_VMAX)
    return oracle.run_table(p, d["phi"], d["psi"], TABLE_STEPS, d["xy_in"], d["tcc"], d["tc"], d["tgamma"],
                            table[0], table[1], tmp=d["tmp"], **kw)


def used(nn, nb):
    """table entries with the unused slots blanked"""
    return np.where(nn[..., None] > np.arange(4), nb, -1)


@pytest.mark.parametrize("case", table_cases(), ids=lambda c: c[0])
def test_oracle_table_matches_reference_fixture(oracle, tbl, case):
    """init_laplacian_coords (sub_wave.c:11555-11892) and evolve_wave_half_new
    (wave_3d.c:345-597) with compute_laplacian: the oracle's table equals the
    reference's, and its steps equal the reference's bit for bit."""
    tag, nx, ny, bc = case[:4]
    p = oracle.make_params(nx, ny, bcond=bc)
    nn, nb = oracle.laplacian_table(p, tuple(tbl[tag + "/lsh"]))
    assert np.array_equal(nn, tbl[tag + "/nneighb"])
    assert np.array_equal(used(nn, nb), used(tbl[tag + "/nneighb"], tbl[tag + "/nghb"]))
    phi, psi, tmp = table_run(oracle, case, tbl)
    assert bits_equal(phi, tbl[tag + "/phi"])
    assert bits_equal(psi, tbl[tag + "/psi"])
    assert bits_equal(tmp, tbl[tag + "/tmp"])


def test_table_fixture_detects_planted_mistakes(oracle, tbl):
    """The fixtures fail an oracle that sums edge neighbours in another
    order, drops the L-shape wrap, uses the plain stepper's bulk rule
    (xy_in > 0 instead of != 0, direct-order sums) or drops the drive."""
    cases = {c[0]: c for c in table_cases()}
    # edge neighbours summed in reverse order
    case = cases["tbl_24x17_bc0_ts0"]
    tag = case[0]
    nn, nb = oracle.laplacian_table(oracle.make_params(24, 17, bcond=0))
    rb = nb.copy()
    edge = np.zeros((24, 17), bool)
    edge[0, :] = edge[-1, :] = edge[:, 0] = edge[:, -1] = True
    for (i, j) in np.argwhere(edge):
        rb[i, j, :nn[i, j]] = nb[i, j, :nn[i, j]][::-1]
    assert not bits_equal(table_run(oracle, case, tbl, table=(nn, rb))[0], tbl[tag + "/phi"])
    # L-shape without its wrap (the plain bulk table), and with the direct-order bulk
    case = cases["tbl_24x17_bc10_ts0"]
    tag = case[0]
    plain = oracle.laplacian_table(oracle.make_params(24, 17, bcond=2))
    assert not bits_equal(table_run(oracle, case, tbl, table=plain)[0], tbl[tag + "/phi"])
    assert not bits_equal(table_run(oracle, case, tbl, bcond=4)[0], tbl[tag + "/phi"])
    # the plain stepper on the same inputs
    case = cases["tbl_16x29_bc2_ts0"]
    tag = case[0]
    d = table_inputs(tag, 16, 29, 0, 0)
    p = oracle.make_params(16, 29, bcond=2, **KAPPAS)
    phi, _, _ = oracle.run(p, d["phi"], d["psi"], TABLE_STEPS, d["xy_in"], d["tcc"], d["tc"], d["tgamma"],
                           tmp=d["tmp"])
    assert not bits_equal(phi, tbl[tag + "/phi"])
    # no drive
    case = cases["tbl_16x29_bc2_ts0_ext"]
    assert not bits_equal(table_run(oracle, case, tbl, drive=None)[0], tbl[case[0] + "/phi"])


@pytest.mark.skipif(not __import__("pyoracle").ref_available(), reason="oracle/_ref not built (needs /root/reference)")
def test_oracle_table_matches_reference_build_random(oracle):
    """Direct cross-check of the neighbour-table stepper against oracle/_ref
    on fresh random inputs (TWOSPEEDS off: see table_inputs)."""
    import pyoracle
    rng = np.random.default_rng(8)
    for (tag, nx, ny, bc, ts, tb, ex) in table_cases():
        if ts or ex:
            continue
        R = pyoracle.RefBuild(tag)
        xy = rng.choice(np.array([-1, 0, 1, 2], np.int16), size=(nx, ny))
        t = [rng.uniform(0, 0.25, (nx, ny)) for _ in range(3)]
        s = [rng.standard_normal((nx, ny)) for _ in range(3)]
        ref = R.run(s[0].copy(), s[1].copy(), s[2].copy(), 5, xy, t[0], t[1], t[2])
        p = oracle.make_params(nx, ny, bcond=bc, oscillate_topbot=tb, tb_shift=nx, **KAPPAS)
        nn, nb = oracle.laplacian_table(p, R.lshape_lines())
        got = oracle.run_table(p, s[0], s[1], 5, xy, t[0], t[1], t[2], nn, nb, tmp=s[2])
        for a, b in zip(ref, got):
            assert bits_equal(a, b), tag


@pytest.mark.parametrize("case", cmp_cases(), ids=lambda c: c[0])
def test_oracle_comparison_matches_reference_fixture(oracle, case):
    """COMPARISON, the flat twin split into two halves (wave_3d.c:736-990):
    the interface rows of both halves with their quirks, bit for bit; the
    plain stepper on the same inputs differs."""
    tag, nx, ny, bc, ts = case
    g = np.load(os.path.join(GOLD, "cmp_cases.npz"))
    d = case_inputs(tag, nx, ny)
    p = oracle.make_params(nx, ny, bcond=bc, twospeeds=ts, tb_shift=ny - 2, comparison=1, **KAPPAS)
    phi, psi, tmp = oracle.run(p, d["phi"], d["psi"], 7, d["xy_in"], d["tcc"], d["tc"], d["tgamma"], tmp=d["tmp"])
    assert bits_equal(phi, g[tag + "/phi"]) and bits_equal(psi, g[tag + "/psi"]) and bits_equal(tmp, g[tag + "/tmp"])
    p0 = oracle.make_params(nx, ny, bcond=bc, twospeeds=ts, tb_shift=ny - 2, **KAPPAS)
    phi0, _, _ = oracle.run(p0, d["phi"], d["psi"], 7, d["xy_in"], d["tcc"], d["tc"], d["tgamma"], tmp=d["tmp"])
    assert not bits_equal(phi0, g[tag + "/phi"])


def test_c1_plain_workload_digests(oracle):
    """1280x720 plain workload: phi after every 100 of 1,000 steps equals
    the reference's bit for bit (SHA-256)."""
    from wavehip import workloads
    w = workloads.c1()
    dig = json.load(open(os.path.join(GOLD, "c1_digests.json")))
    for k in ("phi", "psi", "xy_in", "tcc", "tc", "tgamma"):
        assert sha(w[k]) == dig["inputs"][k], "workload builder changed: " + k
    p = oracle.make_params(1280, 720, bcond=0)
    phi, psi, tmp = w["phi"], w["psi"], np.zeros_like(w["phi"])
    for n in range(100, 1001, 100):
        phi, psi, tmp = oracle.run(p, phi, psi, 100, w["xy_in"], w["tcc"], w["tc"], w["tgamma"], tmp=tmp)
        assert sha(phi) == dig["phi"][str(n)], "step %d" % n
    probes = [float.fromhex(h) for h in dig["probes"]["1000"]]
    assert max(abs(v) for v in probes) > 0  # the pulse is still there


def test_energy_cell_matches_reference(oracle):
    z = np.load(os.path.join(GOLD, "energy_64x48.npz"))
    d = case_inputs("energy_64x48", 64, 48)
    p = oracle.make_params(64, 48, bcond=2, twospeeds=1)
    e = np.array([[oracle.energy_cell(p, d["phi"], d["psi"], d["xy_in"], float(z["courant"]),
                                      float(z["courantb"]), float(z["e_scale"]), i, j)
                   for j in range(48)] for i in range(64)])
    assert bits_equal(e, z["e"])


def test_energy_total_uses_table_courant(oracle):
    """Build-defined total: per-cell compute_energy with c^2 from tcc equals
    the reference's COURANT^2 form on constant media, summed in fixed order."""
    from wavehip import workloads
    z = np.load(os.path.join(GOLD, "energy_64x48.npz"))
    d = case_inputs("energy_64x48", 64, 48)
    tcc, tc, tg = workloads.constant_media_tables(d["xy_in"], float(z["courant"]), float(z["courantb"]),
                                                  0.0, 0.0, 1)
    p = oracle.make_params(64, 48, bcond=2, twospeeds=1)
    tot = oracle.energy_total(p, d["phi"], d["psi"], d["xy_in"], tcc, float(z["e_scale"]))
    assert tot == oracle.fixed_order_sum(z["e"])
    assert abs(tot - z["e"].sum()) <= 1e-12 * z["e"].sum()


@pytest.mark.skipif(not __import__("pyoracle").ref_available(), reason="oracle/_ref not built (needs /root/reference)")
def test_oracle_matches_reference_build_random(oracle):
    """Direct cross-check against oracle/_ref on fresh random inputs."""
    import pyoracle
    rng = np.random.default_rng(7)
    for (tag, nx, ny, bc, ts, tb, tbs) in small_cases():
        R = pyoracle.RefBuild(tag)
        xy = rng.choice(np.array([-1, 0, 1, 2], np.int16), size=(nx, ny))
        t = [rng.uniform(0, 0.25, (nx, ny)) for _ in range(3)]
        s = [rng.standard_normal((nx, ny)) for _ in range(3)]
        ref = R.run(s[0].copy(), s[1].copy(), s[2].copy(), 5, xy, t[0], t[1], t[2])
        p = oracle.make_params(nx, ny, bcond=bc, twospeeds=ts, oscillate_topbot=tb, tb_shift=tbs, **KAPPAS)
        got = oracle.run(p, s[0], s[1], 5, xy, t[0], t[1], t[2], tmp=s[2])
        for a, b in zip(ref, got):
            assert bits_equal(a, b), tag
