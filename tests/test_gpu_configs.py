"""GPU tests of the five benchmark workloads at their configured sizes.

The 1280x720 plain workload is pinned to the reference's digests in
test_gpu_parity.py.  Here: the 8192^2 smooth-medium workload at a reduced
size bit-exact against the oracle (FP64 form against the reference's
arithmetic, split form against its split restatement), and every workload at
full size by properties the oracle cannot afford there: the split form
against the FP64 form (relative L2 of phi, energy series), cells the step
never writes staying zero, finite fields, and batched instances equal to the
same grids stepped alone.

Split-form bounds.  The split form rounds its velocity v = u^n - u^{n-1} to
binary32 every step, an error of 2^-24 |v| per cell and step that adds up
like a random walk.  |v| is about omega*dt*|u|: tiny for the smooth pulses
of the 4096^2 workload (relative L2 below 1e-7 after 10,000 steps, bounded
here by 1e-6), but 0.08-0.3 |u| for the point sources of the lens and
smooth-medium workloads, whose periods of 20-80 steps make waves only 2-8
cells long.  Their bounds below are the measured errors with a margin.
"""
import numpy as np
import pytest

from conftest import say

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def wh():
    import wavehip
    wavehip.load()
    return wavehip


def bits_equal(a, b):
    a, b = np.ascontiguousarray(a), np.ascontiguousarray(b)
    return np.array_equal(a.view(np.uint64), b.view(np.uint64))


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def run_series(wh, w, form, nsteps, every):
    """nsteps steps, energy every `every` steps; (u, u_prev, energies)"""
    W = wh.Wave(w["params"], form=form)
    W.load(w)
    es = [W.energy()]
    for _ in range(nsteps // every):
        W.step(every)
        es.append(W.energy())
    u, um = W.get_state()
    tiling = W.tiling()
    W.close()
    return u, um, np.array(es), tiling


# ------------------------------------------ smooth medium, reduced, exact ----

def test_large_smooth_medium_reduced_matches_oracle(wh, oracle):
    """The 8192^2 workload's construction at 512^2 (9-point, smooth speed
    field c in [0.5, 1.5], absorbing edges, 20 point sources with periods
    U[20, 80]), 400 steps: FP64 bit-exact vs the oracle, split bit-exact vs
    the oracle's split restatement, energies equal."""
    from wavehip import workloads
    n, steps = 512, 400
    w = workloads.c5(n=n, nsteps=steps)
    p = oracle.make_params(n, n, bcond=2, stencil=9)
    args = (w["xy_in"], w["tcc"], w["tc"], w["tgamma"])
    ref = oracle.run(p, w["phi"], w["psi"], steps, *args, sources=w["sources"])
    e_ref = oracle.energy_total(p, ref[0], ref[1], w["xy_in"], w["tcc"], 1.0)
    hi, lo, v = oracle.split_from_fp64(w["phi"], w["psi"])
    hi, lo, v = oracle.split_run(p, hi, lo, v, steps, *args, sources=w["sources"])
    sref = oracle.split_to_fp64(hi, lo, v)
    se_ref = oracle.split_energy_total(p, hi, lo, v, w["xy_in"], w["tcc"], 1.0)
    for form, want, ew in ((0, ref, e_ref), (1, sref, se_ref)):
        for k, L in ((0, 0), (8, 64), (6, 16)):
            say("smooth 512 form", form, "tiling", k, L)
            W = wh.Wave(w["params"], form=form, tb_depth=k, rows_per_tile=L)
            W.load(w)
            W.step(steps)
            g = W.get_state()
            e = W.energy()
            W.close()
            assert bits_equal(g[0], want[0]), (form, k, L)
            assert bits_equal(g[1], want[1]), (form, k, L)
            assert e == ew


# ------------------------------------------------------- full sizes ---------

def test_large_smooth_medium_full_size(wh):
    """8192^2, 2,000 steps, energy every 100: split vs FP64."""
    from wavehip import workloads
    w = workloads.c5(nsteps=2000)
    f = run_series(wh, w, 0, 2000, 100)
    s = run_series(wh, w, 1, 2000, 100)
    rel = rel_l2(s[0], f[0])
    de = float(np.abs(s[2] - f[2]).max() / np.abs(f[2]).max())
    say("8192^2 tilings", f[3], s[3], "rel L2", rel, "energy series", de)
    assert np.isfinite(f[0]).all() and np.isfinite(s[0]).all()
    assert np.abs(f[0]).max() > 1e-3        # the sources have driven the field
    assert rel < 2e-6   # measured 9.0e-7
    assert de < 1e-6


def test_precision_form_4096_10000_steps(wh):
    """The 4096^2 precision workload for its full 10,000 steps: split within
    1e-6 relative L2 of FP64, equal energy drift, obstacle cells untouched."""
    from wavehip import workloads
    w = workloads.c3()
    A = wh.Wave(w["params"])
    A.load(w)
    B = wh.Wave(w["params"], form=1)
    B.load(w)
    e0 = A.energy()
    A.step(10000)
    B.step(10000)
    ua, _ = A.get_state()
    ub, _ = B.get_state()
    ea, eb = A.energy(), B.energy()
    A.close()
    B.close()
    rel = rel_l2(ub, ua)
    say("4096^2 10000 steps rel L2", rel, "drift", (ea - e0) / e0, (eb - e0) / e0)
    assert rel <= 1e-6
    assert np.all(ua[w["xy_in"] == 0] == 0) and np.all(ub[w["xy_in"] == 0] == 0)
    assert abs(ea - eb) < 1e-7 * e0


def test_lens_full_size(wh):
    """2560x1440 9-point lens, damping, source, 5,000 steps, energy every 50:
    split vs FP64; the energy series track each other."""
    from wavehip import workloads
    w = workloads.c2()
    f = run_series(wh, w, 0, 5000, 50)
    s = run_series(wh, w, 1, 5000, 50)
    rel = rel_l2(s[0], f[0])
    de = float(np.abs(s[2] - f[2]).max() / np.abs(f[2]).max())
    say("2560x1440 tilings", f[3], s[3], "rel L2", rel, "energy series", de)
    assert np.isfinite(f[0]).all() and np.isfinite(s[0]).all()
    assert f[2][-1] > 0 and np.all(f[2][1:] > 0)
    assert rel < 5e-6   # measured 3.1e-6
    assert de < 1e-6


def test_batch_all_instances(wh):
    """All 256 instances of the batching workload (72 M cells), 2,000 steps
    in both forms: per-instance split vs FP64, and FP64 instances equal bit
    for bit to the same grids stepped alone."""
    from wavehip import workloads
    ws = workloads.c4_instances()
    out = {}
    for form in (0, 1):
        B = wh.Batch([w["params"] for w in ws], form=form)
        for k, w in enumerate(ws):
            B.load(k, w)
        B.step(2000)
        out[form] = [B.get_state(k)[0] for k in range(len(ws))]
        say("batch form", form, B.tiling())
        B.close()
    rels = np.array([rel_l2(a, b) for a, b in zip(out[1], out[0])])
    say("batch rel L2 max", rels.max(), "median", np.median(rels))
    assert np.isfinite(rels).all()
    assert rels.max() < 5e-6 and np.median(rels) < 5e-7
    for k in (0, 17, 101, 255):
        W = wh.Wave(ws[k]["params"])
        W.load(ws[k])
        W.step(2000)
        u = W.get_state()[0]
        W.close()
        assert bits_equal(u, out[0][k]), k
