"""A C host program in the shape of a reference wave program, built against
include/wavehip_compat.h and linked with libwavehip.so (tests/compat/), calls
evolve_wave / evolve_wave_half with the reference signatures: row-pointer
tables as wave_billiard.c:302/613 take them, flat tables as wave_3d.c:600/1010
take them.  Its output must equal the oracle's bit for bit.

Each binary was built with the parameter macros of one reference program
(tests/compat/Makefile); `--params` prints them and the oracle takes its
parameters from there."""
import os
import subprocess

import numpy as np
import pytest

from conftest import say

HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(HERE, "compat", "bin")
VARIANTS = ["rows_abs", "rows_per", "rows_dir_tb", "flat_vper", "flat_abs_fixed", "rows_ext", "rows_ext_wide",
            "flat_ext", "flat_tbl_dir", "flat_tbl_lshape", "flat_cmp", "flat_cmp_ext"]


def host_params(name):
    exe = os.path.join(BIN, name)
    if not os.path.exists(exe):
        pytest.fail("%s is not built (make -C tests/compat)" % exe)
    out = subprocess.run([exe, "--params"], capture_output=True, text=True, timeout=60, check=True).stdout
    d = {}
    for kv in out.split():
        k, v = kv.split("=")
        d[k] = float(v) if "." in v or "e" in v else int(v)
    return d


@pytest.mark.parametrize("name", VARIANTS)
def test_host_program_built(name):
    d = host_params(name)
    assert d["nx"] >= 3 and d["ny"] >= 3 and (0 <= d["bcond"] <= 4 or d["bcond"] == 10)


def oracle_params(oracle, d):
    tbs = d["nx"]                    # wave_billiard.c:319 with XMIN..XMAX the whole box
    if d["flat"] and not d["table"] and tbs >= d["nx"] - 1:
        tbs = d["ny"] - 2            # wave_3d.c:613-617 (the table stepper keeps NX, 387)
    return oracle.make_params(d["nx"], d["ny"], bcond=d["bcond"], twospeeds=d["twospeeds"],
                              oscillate_topbot=d["oscillate_topbot"], tb_shift=tbs,
                              kappa=d["kappa"], kappa_sides=d["kappa_sides"],
                              gamma_sides=d["gamma_sides"], kappa_topbot=d["kappa_topbot"],
                              gamma_topbot=d["gamma_topbot"], comparison=d["comparison"])


def drive_cells(d):
    """the cells the reference drives: wave_billiard.c:336-338, 397, 585-590
    (row pointers; the corner (0, NX-1) only when it lies in row 0) or
    wave_3d.c:651-654, 989-993 (flat) or 358-376, 413, 575-579 (table stepper)"""
    nx, ny = d["nx"], d["ny"]
    cells = [(0, j) for j in range(1, ny - 1)]
    if d["comparison"]:
        # the interface rows overwrite the drive where they apply (wave_3d.c:873-982)
        xy = d["xy0"]
        cells = [(0, j) for j in range(1, ny - 1) if not (j in (ny // 2 - 1, ny // 2) and (d["twospeeds"] or xy[j]))]
    if d["flat"]:
        return cells + [(0, 0), (0, ny - 1)]
    return cells + [(0, 0)] + ([(0, nx - 1)] if nx - 1 < ny else [])


def oracle_expect(oracle, d, w, ncalls, mode):
    p = oracle_params(oracle, d)
    phi, psi, tmp = (np.array(w[k], dtype=np.float64, copy=True) for k in ("phi", "psi", "tmp"))
    xy, tcc, tc, tg = w["xy_in"], w["tcc"], w["tc"], w["tgamma"]
    if d["table"]:
        n = 3 * ncalls if mode == "wave" else ncalls
        kw = {}
        if d["oscillate_left"]:
            cells = drive_cells(d)
            val = np.array([w["table"][:n, j] for (_, j) in cells[:-2]] + [w["corner"][:n]] * 2)
            kw["drive"] = dict(i=[c[0] for c in cells], j=[c[1] for c in cells], val=val)
        if d["floor"]:
            kw["vmax"] = d["vmax"]
        return oracle.run_table(p, phi, psi, n, xy, tcc, tc, tg, w["lap"][0], w["lap"][1], tmp=tmp, **kw)
    d = dict(d, xy0=w["xy_in"][0])
    if d["neumann"] or d["oscillate_left"] or d["floor"]:
        n = 3 * ncalls if mode == "wave" else ncalls
        kw = {}
        if d["neumann"]:
            kw["neumann"] = w["nb"]
        if d["oscillate_left"]:
            cells = drive_cells(d)
            kw["drive"] = dict(i=[c[0] for c in cells], j=[c[1] for c in cells],
                               val=np.array([w["table"][:n, j] for (_, j) in cells]))
        if d["floor"]:
            kw["vmax"] = d["vmax"]
        return oracle.run_x(p, phi, psi, n, xy, tcc, tc, tg, tmp=tmp, **kw)
    if not d["flat"]:
        n = 3 * ncalls if mode == "wave" else ncalls
        return oracle.run(p, phi, psi, n, xy, tcc, tc, tg, tmp=tmp)
    for _ in range(ncalls):
        if mode == "wave":
            oracle.evolve_wave_flat(p, phi, psi, tmp, xy, tcc, tc, tg)
        else:
            oracle.evolve_wave_half_flat(p, phi, psi, tmp, xy, tc, tcc, tg)
            phi, psi, tmp = tmp, phi, psi
    return phi, psi, tmp


def neumann_points(xy, rng, n):
    """ghost points at distinct cells, each averaging two in-grid neighbours
    within one cell (weights w, 1-w)"""
    nx, ny = xy.shape
    cells = rng.choice(nx * ny, size=n, replace=False)
    out = {k: [] for k in ("i", "j", "nbi1", "nbj1", "nbi2", "nbj2", "nx", "ny")}
    for c in cells:
        i, j = divmod(int(c), ny)
        near = [(i + a, j + b) for a in (-1, 0, 1) for b in (-1, 0, 1) if 0 <= i + a < nx and 0 <= j + b < ny]
        p1, p2 = near[rng.integers(len(near))], near[rng.integers(len(near))]
        wgt = rng.uniform(0, 1)
        for k, v in zip(out, (i, j, p1[0], p1[1], p2[0], p2[1], wgt, 1 - wgt)):
            out[k].append(v)
    return {k: np.array(v) for k, v in out.items()}


@pytest.mark.gpu
@pytest.mark.parametrize("mode,ncalls", [("wave", 4), ("half", 5)])
@pytest.mark.parametrize("name", VARIANTS)
def test_host_program_matches_oracle(oracle, tmp_path, name, mode, ncalls):
    from wavehip import workloads
    d = host_params(name)
    say("compat host", name, mode, ncalls)
    w = workloads.random_case(d["nx"], d["ny"], seed=100 + VARIANTS.index(name), bcond=d["bcond"],
                              twospeeds=d["twospeeds"])
    # the third buffer holds values too: cells the step never writes keep them
    w["tmp"] = np.random.default_rng(7).standard_normal((d["nx"], d["ny"]))
    if d["floor"]:
        for k in ("phi", "psi", "tmp"):
            w[k] = w[k] * 0.5
    rng = np.random.default_rng(11)
    w["nb"] = neumann_points(w["xy_in"], rng, 40) if d["neumann"] else None
    if d["table"]:
        lsh = workloads.lshape_lines(d["nx"], d["ny"], 1.5, -2.0, 2.0, -1.125, 1.125) if d["bcond"] == 10 else None
        w["lap"] = workloads.laplacian_table(d["nx"], d["ny"], d["bcond"], lsh)
    T = 3 * ncalls
    w["table"] = np.cos(0.4 * np.arange(1, T + 1)[:, None] + 0.05 * np.arange(d["ny"])[None, :])
    inp, outp = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(inp, "wb") as f:
        for k in ("phi", "psi", "tmp", "tcc", "tc", "tgamma"):
            f.write(np.ascontiguousarray(w[k], dtype=np.float64).tobytes())
        f.write(np.ascontiguousarray(w["xy_in"], dtype=np.int16).tobytes())
        nb = w["nb"]
        f.write(np.int32(0 if nb is None else len(nb["i"])).tobytes())
        if nb is not None:
            f.write(np.ascontiguousarray(np.stack([nb[k] for k in ("i", "j", "nbi1", "nbj1", "nbi2", "nbj2")], 1),
                                         np.int32).tobytes())
            f.write(np.ascontiguousarray(nb["nx"], np.float64).tobytes())
            f.write(np.ascontiguousarray(nb["ny"], np.float64).tobytes())
        f.write(np.int32(T).tobytes())
        f.write(np.ascontiguousarray(w["table"], np.float64).tobytes())
        if d["table"]:
            f.write(np.ascontiguousarray(w["lap"][0], np.int32).tobytes())
            f.write(np.ascontiguousarray(w["lap"][1], np.int32).tobytes())
    r = subprocess.run([os.path.join(BIN, name), str(inp), str(outp), str(ncalls), mode],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    raw = np.fromfile(outp, dtype=np.float64)
    got = raw[:3 * d["nx"] * d["ny"]].reshape(3, d["nx"], d["ny"])
    w["corner"] = raw[3 * d["nx"] * d["ny"]:]  # the table stepper's corner drive, computed by the host
    exp = oracle_expect(oracle, d, w, ncalls, mode)
    for a, b, what in zip(got, exp, ("phi", "psi", "tmp")):
        assert np.array_equal(a.view(np.uint64), np.ascontiguousarray(b).view(np.uint64)), \
            "%s differs at %s" % (what, np.argwhere(a.view(np.uint64) != b.view(np.uint64))[:3])


# ---- the start-up drop-ins (WAVEHIP_INIT, tests/compat/compat_init.c) --------

INIT_VARIANTS = ["init_michelson", "init_wells", "init_flat_cmp", "init_rwells", "init_latmaze"]
# the macros compat_init.c fixes (its #define block)
INIT_BOX = (-2.0, 2.0, -1.125, 1.125)
INIT_DOMAIN = dict(lam=0.8, mu=0.3, mu_b=0.08, apoly=0.5, npoly=6, wall_width=0.1, wall_width_b=0.1,
                   julia_scale=0.8, julia_x=0.37468, julia_y=0.21115, mandellimit=10.0, mdepth=4, mratio=3,
                   mandellevel=100, ngridx=6, ngridy=5)
INIT_IOR = dict(courant=0.1, courantb=0.05, gamma=1.0e-4, gammab=2.0e-4, lam=0.8, mu=0.3, mandel_ior_scale=-0.05,
                mandellimit=10.0, mandellevel=100, mdepth=4)


@pytest.mark.parametrize("name", INIT_VARIANTS)
def test_init_program_built(name):
    d = host_params(name)
    assert d["nx"] >= 3 and d["ny"] >= 3


@pytest.mark.gpu
@pytest.mark.parametrize("name", INIT_VARIANTS)
def test_init_program_matches_oracle(oracle, tmp_path, name):
    """wavehip_init_wave_flat(_mod) and wavehip_init_ior_2d called as a
    reference program calls init_wave_flat and init_ior_2d: the mask equals
    the oracle's xy_in_billiard (with COMPARISON for the flat build), phi and
    psi are zeroed, and the media equal the oracle's tables (bit for bit; the
    wells within the GPU's exp, rtol 1e-14)"""
    from gen_inputs import mask_circles, mask_circles_b
    d = host_params(name)
    nx, ny = d["nx"], d["ny"]
    ca, cb = mask_circles(), mask_circles_b()
    inp, outp = tmp_path / "in.bin", tmp_path / "out.bin"
    if d["lists"]:  # eight discs of radius MU, and list_data's rectangles
        from gen_inputs import list_data
        ca = np.ascontiguousarray(ca[:, :8])
        lists = list_data(INIT_DOMAIN["npoly"])
    with open(inp, "wb") as f:
        for c in (ca, cb):
            f.write(np.int32(c.shape[1]).tobytes())
            f.write(np.ascontiguousarray(c, np.float64).tobytes())
        if d["lists"]:
            for key in ("rects", "rectrot"):
                f.write(np.int32(lists[key].shape[1]).tobytes())
                f.write(np.ascontiguousarray(lists[key], np.float64).tobytes())
        if d["xy_from_input"]:  # the mask and seed of the reference's IOR_RANDOM_WELLS fixture
            from gen_inputs import case_inputs
            f.write(np.ascontiguousarray(case_inputs("ior_wells_r", nx, ny)["xy_in"], np.int16).tobytes())
            f.write(np.int32(7).tobytes())
    say("compat init", name)
    r = subprocess.run([os.path.join(BIN, name), str(inp), str(outp)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    raw = open(outp, "rb").read()
    n = nx * ny
    xy = np.frombuffer(raw[:2 * n], np.int16).reshape(nx, ny)
    rest = np.frombuffer(raw[2 * n:], np.float64)
    if d["xy_from_input"]:
        # init_ior_2d's own tables for this mask after srand(7)
        # (tests/golden/media_cases.npz, reference build ior_wells_r)
        g = np.load(os.path.join(HERE, "golden", "media_cases.npz"))
        rest = np.frombuffer(raw[2 * n:], np.float64)
        tcc, tc, tg = (rest[(2 + q) * n:(3 + q) * n].reshape(nx, ny) for q in range(3))
        for a, key in ((tcc, "tcc"), (tc, "tc"), (tg, "tgamma")):
            assert np.allclose(a, g["ior_wells_r/" + key], rtol=1e-14, atol=0), key
        return
    P = oracle.domain_params(**INIT_DOMAIN)
    if d["comparison"]:
        want = oracle.comparison_mask(nx, ny, INIT_BOX, d["domain"], d["domain_b"], P, ca, cb)
    elif d["lists"]:
        want = oracle.list_domain(nx, ny, INIT_BOX, d["domain"], P, circles=ca, rects=lists["rects"],
                                  rectrot=lists["rectrot"])
        assert 0 < np.count_nonzero(want) < want.size
    else:
        want = oracle.domain(nx, ny, INIT_BOX, d["domain"], P, ca if d["domain"] == 20 else None)
    assert np.array_equal(xy, want)
    assert np.all(rest[:2 * n] == 0.0)  # phi, psi
    if rest.size == 2 * n:
        return
    tcc, tc, tg = (rest[(2 + q) * n:(3 + q) * n].reshape(nx, ny) for q in range(3))
    p = oracle.make_params(nx, ny, twospeeds=d["twospeeds"])
    if d["ior"] == 4:  # IOR_PERIODIC_WELLS, sub_wave.c:12317-12345
        dx, dy = 4.0 / 6, 2.25 / 5
        ii, jj = np.meshgrid(np.arange(6), np.arange(5), indexing="ij")
        xc = (-2.0 + dx * (ii + 0.5)).T.ravel()
        yc = (-1.125 + dy * (jj + 0.5) + np.where(jj % 2 == 1, 0.5 * dx, 0.0)).T.ravel()
        ref = oracle.wells_tables(p, xy, xc, yc, None, 0.2 * dx * dx, 0, 0.1, 0.05, 1.0e-4, INIT_BOX)
        for a, b in zip((tcc, tc, tg), ref):
            assert np.allclose(a, b, rtol=1e-14, atol=0)
    else:
        ref = oracle.ior_tables(p, xy, d["ior"], INIT_BOX, oracle.IorParams(**INIT_IOR))
        for a, b in zip((tcc, tc, tg), ref):
            assert np.array_equal(a.view(np.uint64), b.view(np.uint64))
