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
VARIANTS = ["rows_abs", "rows_per", "rows_dir_tb", "flat_vper", "flat_abs_fixed"]


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
    assert d["nx"] >= 3 and d["ny"] >= 3 and 0 <= d["bcond"] <= 4


def oracle_params(oracle, d):
    tbs = d["nx"]                    # wave_billiard.c:319 with XMIN..XMAX the whole box
    if d["flat"] and tbs >= d["nx"] - 1:
        tbs = d["ny"] - 2            # wave_3d.c:613-617
    return oracle.make_params(d["nx"], d["ny"], bcond=d["bcond"], twospeeds=d["twospeeds"],
                              oscillate_topbot=d["oscillate_topbot"], tb_shift=tbs,
                              kappa=d["kappa"], kappa_sides=d["kappa_sides"],
                              gamma_sides=d["gamma_sides"], kappa_topbot=d["kappa_topbot"],
                              gamma_topbot=d["gamma_topbot"])


def oracle_expect(oracle, d, w, ncalls, mode):
    p = oracle_params(oracle, d)
    phi, psi, tmp = (np.array(w[k], dtype=np.float64, copy=True) for k in ("phi", "psi", "tmp"))
    xy, tcc, tc, tg = w["xy_in"], w["tcc"], w["tc"], w["tgamma"]
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
    inp, outp = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(inp, "wb") as f:
        for k in ("phi", "psi", "tmp", "tcc", "tc", "tgamma"):
            f.write(np.ascontiguousarray(w[k], dtype=np.float64).tobytes())
        f.write(np.ascontiguousarray(w["xy_in"], dtype=np.int16).tobytes())
    r = subprocess.run([os.path.join(BIN, name), str(inp), str(outp), str(ncalls), mode],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    got = np.fromfile(outp, dtype=np.float64).reshape(3, d["nx"], d["ny"])
    exp = oracle_expect(oracle, d, w, ncalls, mode)
    for a, b, what in zip(got, exp, ("phi", "psi", "tmp")):
        assert np.array_equal(a.view(np.uint64), np.ascontiguousarray(b).view(np.uint64)), \
            "%s differs at %s" % (what, np.argwhere(a.view(np.uint64) != b.view(np.uint64))[:3])
