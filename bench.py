"""Benchmark of the wave stepper on one MI355X.

Workload: the precision form at 4096 x 4096 -- [-2,2]^2, 64 seeded disc
obstacles, three Gaussian pulses, courant 0.1, absorbing edges, 5-point
Laplacian.  A "step" is one leapfrog time step over the whole grid (one
evolve_wave_half of the reference, wave_billiard.c:302), so one step is
4096^2 cell-updates.  The headline value is the FP32 double-float (split)
form; the FP64 form (bit-identical to the reference) runs the same steps
beside it, and the relative L2 distance and energy drift between the two are
reported.  Inputs are resident in HBM before the timed region starts.

  python bench.py [--steps K] [--warmup W] [--gpus 1]

Prints one JSON line.  The CPU baseline is the oracle (this repository's C +
OpenMP restatement of the reference evolve_wave_half, pinned bit-exact to the
reference's own code) timed on a bounded number of steps of the same FP64
workload on the host cores.  Two extra legs (off with --no-frame) time the
frame consumer (8-bit frames of the drawing path's colours against
downloading the field) and the media builder (init_ior_2d's random wells).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "youtube-simulations_amd"))

HBM_PEAK = 8.0e12              # MI355X HBM3E, bytes/s
# wave64 VALU instructions the chip can issue per second: 256 CUs x 4 SIMDs,
# one wave64 instruction per SIMD every 4 cycles (16 lanes), 2.4 GHz
VALU_ISSUE_PEAK = 256 * 4 * 2.4e9 / 4
B_ALG = {"f32split": 24.125,   # read u_hi, u_lo, v + write them (3 x 4 B each way) + 1-bit mask
         "fp64": 24.125}       # read u^n, u^{n-1} + write u^{n+1} (3 x 8 B) + 1-bit mask
METRIC = "Gcell-updates/s at 4096^2 (FP64 plain vs FP32 split) + % of HBM roofline"
# sources a stored counter profile must match: the kernel and the host code
# that builds its tile tables (tiling, order and strip lengths shape the launch)
KERNEL_SOURCES = {"f32split": ["wv_step_split.hip", "wv_device.h", "wv_internal.h", "wv_api.cpp"],
                  "fp64": ["wv_step_fp64.hip", "wv_device.h", "wv_internal.h", "wv_api.cpp"]}


def compulsory_bytes_per_cell(form, k):
    """HBM bytes one fused launch of depth k must move per cell: the state read
    once and written once, plus the 1-bit mask.  Split: hi, lo, v in and out
    (24 B).  FP64: u^n and u^{n-1} in; u^{n+k} and u^{n+k-1} out when k > 1
    (the next launch needs both), u^{n+1} alone when k == 1."""
    if form == "f32split":
        return 24.125
    return 32.125 if k > 1 else 24.125


def kernel_source_sha(form):
    import hashlib
    h = hashlib.sha256()
    for f in KERNEL_SOURCES[form]:
        with open(os.path.join(ROOT, "youtube-simulations_amd", "csrc", f), "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()[:16]


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def run_form(wh, w, form, steps, warmup, tb_depth=0, rows=0, tune=False):
    W = wh.Wave(w["params"], form=form, tb_depth=tb_depth, rows_per_tile=rows)
    W.load(w)
    if tune:
        W.autotune(200)
    e0 = W.energy()
    if warmup:
        W.step(warmup)
        W.sync()
        W.set_state(w["phi"], w["psi"])   # the timed run starts from the initial state
    # set-up of the timed sequence outside the timed region: the tile tables of
    # every fused depth it uses and the first load of every kernel it launches
    # (no step is taken)
    W.prepare(steps)
    W.sync()
    t0 = time.perf_counter()
    # HIP events on the stepping stream: one before the first launch, one
    # after each launch and one after the step counter's update, the last
    # waited for -- when it returns every step of the sequence has completed
    ms_dev, launches = W.step_timed(steps)
    wall = time.perf_counter() - t0
    W.sync()
    avg_ms, n_timed = W.kernel_ms()          # every full-depth launch of the same run, event-timed
    e1 = W.energy()
    u, um = W.get_state()
    k, L, cols = W.tiling()
    W.close()
    return dict(wall=wall, ms_dev=ms_dev, avg_ms=avg_ms, n_timed=n_timed, launches=launches, e0=e0, e1=e1,
                u=u, k=k, L=L, cols=cols)


def frame_bench(wh, w, reps=10):
    """The frame consumer on the same grid: the drawing configuration of the
    checked-in wave_billiard.c (HIGHRES blocks of 2, P_AMPLITUDE, C_ONEDIM,
    COL_TWILIGHT_SHIFTED, SHADE_2D, FADE_IN_OBSTACLE with the fade table
    pow(tcc / courant^2, 0.1) of wave_billiard.c:1745) as an 8-bit frame,
    against downloading phi and psi for the host to colour; wall clock per
    frame, host buffers preallocated (the table is a synthetic 511-entry one:
    colours do not change the work)."""
    nx, ny = w["params"]["nx"], w["params"]["ny"]
    W = wh.Wave(w["params"], form=wh.FORM_FP64)
    W.load(w)
    W.step(100)
    W.sync()
    lut = np.random.default_rng(0).uniform(0, 1, (511, 3)).astype(np.float32)
    fade = np.where(w["xy_in"] != 0, 1.0, np.power(w["tcc"] / 0.01, 0.1))
    W.set_frame_tables(lut=lut, fade_table=fade)
    out = {}
    for size in (2, 1):
        f = W.frame_params(plot=0, scheme=3, palette=18, size=size, shade=1, fade_in_obstacle=1)
        buf = W.frame(f, u8=True)
        t0 = time.perf_counter()
        for _ in range(reps):
            W.frame(f, u8=True, out=buf)
        out["rgba8_ms_size%d" % size] = (time.perf_counter() - t0) / reps * 1e3
    phi, psi = np.empty((nx, ny)), np.empty((nx, ny))
    W.get_state(out=(phi, psi))
    t0 = time.perf_counter()
    for _ in range(reps):
        W.get_state(out=(phi, psi))
    out["get_state_ms"] = (time.perf_counter() - t0) / reps * 1e3
    W.close()
    out["bytes_rgba8_size2"] = 4 * (-(-nx // 2)) * (-(-ny // 2))
    out["bytes_state"] = 16 * nx * ny
    out["workload"] = ("%dx%d FP64 state, P_AMPLITUDE / C_ONEDIM / COL_TWILIGHT_SHIFTED, SHADE_2D, "
                       "FADE_IN_OBSTACLE, 8-bit RGBA" % (nx, ny))
    return out


def media_bench(wh, n=4096, nwells=100):
    """init_ior_2d's IOR_RANDOM_WELLS media for an n x n grid (10 x 10 wells,
    the checked-in NGRIDX, NGRIDY) built by wv_set_wells_tables, wall clock
    including the table upload, the D_CIRCLES mask of the precision workload
    built by wv_build_mask, and the IOR_MANDELBROT media and D_MANDELBROT mask
    (wv_set_ior_tables, wv_build_domain_mask); beside each the oracle (OpenMP,
    all host threads) on a band of rows of the same grid, scaled to the grid.  Also a
    1920x1080 frame saved as writetiff saves it (LZW strips coded on the
    device) beside the oracle's coder on one host thread."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    rng = np.random.default_rng(5)
    g = int(round(nwells ** 0.5))
    dx, dy = 4.0 / g, 4.0 / g
    ii, jj = np.meshgrid(np.arange(g), np.arange(g), indexing="ij")
    xc = (-2.0 + dx * (ii + 0.5 + 0.1 * rng.standard_normal(ii.shape))).ravel()
    yc = (-2.0 + dy * (jj + 0.5 + 0.1 * rng.standard_normal(jj.shape))).ravel()
    h = np.clip(0.5 + 0.5 * rng.standard_normal(xc.size), 0, 1)
    sigma = 0.2 * dx * dx
    xy = np.ones((n, n), np.int16)
    box = (-2.0, 2.0, -2.0, 2.0)
    W = wh.Wave(dict(nx=n, ny=n, bcond=2))
    t0 = time.perf_counter()
    W.set_wells_tables(xy, xc, yc, h, sigma, 1, 0.1, 0.05, 0.0, box, copy=False)
    gpu_ms = (time.perf_counter() - t0) * 1e3
    W.close()
    rows = 64
    p = pyoracle.make_params(rows, n)
    band = (-2.0, -2.0 + 4.0 * rows / n, -2.0, 2.0)
    t0 = time.perf_counter()
    pyoracle.wells_tables(p, xy[:rows], xc, yc, h, sigma, 1, 0.1, 0.05, 0.0, band)
    cpu_ms = (time.perf_counter() - t0) * 1e3 * n / rows
    # the D_CIRCLES mask of the precision workload (64 seeded discs), built by
    # wv_build_mask including the copy of the int16 mask to the host
    from wavehip import workloads
    discs = np.array(workloads.random_discs(workloads.SplitMix64(3), 64, box, 0.02, 0.1))
    circ = pyoracle.circle_list(discs[:, 0], discs[:, 1], discs[:, 2])
    W = wh.Wave(dict(nx=n, ny=n, bcond=2))
    W.build_mask(wh.D_CIRCLES, box, 0.0, 0.0, circ)  # first call: code load
    t0 = time.perf_counter()
    m = W.build_mask(wh.D_CIRCLES, box, 0.0, 0.0, circ)
    mask_ms = (time.perf_counter() - t0) * 1e3
    # a 1920x1080 frame (the glReadPixels rows) saved as writetiff saves it:
    # LZW strips coded on the device; wall clock from host rows to file bytes
    fy, fx = np.meshgrid(np.linspace(-1.125, 1.125, 1080), np.linspace(-2.0, 2.0, 1920), indexing="ij")
    r = np.hypot(fx + 1.0, fy)
    u = np.exp(-r * r) * np.cos(40.0 * r)
    frame = np.stack([127.5 + 127.0 * u, 127.5 - 90.0 * u, 127.5 + 60.0 * np.sin(3.0 * u)], -1)
    frame[np.hypot(fx - 0.5, fy - 0.3) < 0.3] = 0.0
    frame = np.floor(frame).astype(np.uint8)
    W.tiff(frame)  # first call: code load, buffers
    t0 = time.perf_counter()
    for _ in range(5):
        tf = W.tiff(frame)
    tiff_ms = (time.perf_counter() - t0) * 1e3 / 5
    W.close()
    t0 = time.perf_counter()
    tfo = pyoracle.tiff(frame)
    tiff_cpu_ms = (time.perf_counter() - t0) * 1e3
    assert tf == tfo
    t0 = time.perf_counter()
    mo = pyoracle.mask(rows, n, band, wh.D_CIRCLES, 0.0, 0.0, circ)
    mask_cpu_ms = (time.perf_counter() - t0) * 1e3 * n / rows
    assert np.array_equal(m[:rows], mo)
    # init_ior_2d's IOR_MANDELBROT media (escape time to MANDELLEVEL 1000, the
    # checked-in value; wall clock including the tables' copy to the host and
    # their install) and the D_MANDELBROT domain mask; the oracle on the band
    # of rows through the middle of the set (dyadic grid: the band's cell
    # corners are the grid's)
    i0 = n // 2 - rows // 2
    mid = (-2.0 + 4.0 * i0 / n, -2.0 + 4.0 * (i0 + rows) / n, -2.0, 2.0)
    ip = wh.ior_params(courant=0.1, courantb=0.05, lam=0.8, mu=0.3)
    dp = wh.domain_params()
    W = wh.Wave(dict(nx=n, ny=n, bcond=2))
    W.set_ior_tables(xy, 1, box, ip, copy=False)  # first call: code load
    t0 = time.perf_counter()
    tm = W.set_ior_tables(xy, 1, box, ip)
    ior_ms = (time.perf_counter() - t0) * 1e3
    W.build_domain_mask(24, box, dp)
    t0 = time.perf_counter()
    dm = W.build_domain_mask(24, box, dp)
    dom_ms = (time.perf_counter() - t0) * 1e3
    W.close()
    oip = pyoracle.IorParams(**{f: getattr(ip, f) for f, _ in ip._fields_})
    odp = pyoracle.DomainParams(**{f: getattr(dp, f) for f, _ in dp._fields_})
    t0 = time.perf_counter()
    to = pyoracle.ior_tables(pyoracle.make_params(rows, n), xy[:rows], 1, mid, oip)
    ior_cpu_ms = (time.perf_counter() - t0) * 1e3 * n / rows
    t0 = time.perf_counter()
    do = pyoracle.domain(rows, n, mid, 24, odp)
    dom_cpu_ms = (time.perf_counter() - t0) * 1e3 * n / rows
    assert np.array_equal(dm[i0:i0 + rows], do)
    assert np.all(np.abs(tm[0][i0:i0 + rows] - to[0]) <= 0.01 * 1e-15)
    return {"workload": "%dx%d IOR_RANDOM_WELLS, %d wells, tanh" % (n, n, xc.size), "gpu_ms": gpu_ms,
            "cpu_oracle_ms_scaled": cpu_ms, "cpu_threads": pyoracle.num_threads(),
            "cpu_sample": "%d of %d rows" % (rows, n),
            "mask": {"workload": "%dx%d D_CIRCLES mask, 64 discs, int16 copied to the host" % (n, n),
                     "gpu_ms": mask_ms, "cpu_oracle_ms_scaled": mask_cpu_ms},
            "ior_mandelbrot": {"workload": "%dx%d IOR_MANDELBROT media, MANDELLEVEL 1000, tables copied to the host "
                                           "and installed" % (n, n), "gpu_ms": ior_ms,
                               "cpu_oracle_ms_scaled": ior_cpu_ms, "cpu_sample": "rows %d..%d" % (i0, i0 + rows)},
            "domain_mandelbrot": {"workload": "%dx%d D_MANDELBROT mask, MANDELLEVEL 1000, int16 copied to the host"
                                              % (n, n), "gpu_ms": dom_ms, "cpu_oracle_ms_scaled": dom_cpu_ms},
            "tiff": {"workload": "1920x1080 RGB frame to an LZW TIFF file image (one row a strip), host rows in, "
                                 "file bytes out", "gpu_ms": tiff_ms, "cpu_oracle_ms": tiff_cpu_ms,
                     "cpu_threads": 1, "file_bytes": len(tf)}}


def host_cpu():
    """model name, logical CPUs, physical cores (lscpu) and the CPUs this
    process may run on"""
    import subprocess
    info = {"model": None, "nproc": os.cpu_count(), "physical_cores": None,
            "affinity": len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else None}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        kv = {}
        for line in out.splitlines():
            if ":" in line:
                a, b = line.split(":", 1)
                kv[a.strip()] = b.strip()
        info["model"] = kv.get("Model name")
        cps, sockets = kv.get("Core(s) per socket"), kv.get("Socket(s)")
        if cps and sockets and cps.isdigit() and sockets.isdigit():
            info["physical_cores"] = int(cps) * int(sockets)
    except (OSError, subprocess.SubprocessError):
        pass
    return info


def _time_steps(step, target_s, chunk):
    """run step(chunk) until target_s seconds have been timed: (steps, seconds)"""
    n, dt = 0, 0.0
    while dt < target_s:
        t0 = time.perf_counter()
        step(chunk)
        dt += time.perf_counter() - t0
        n += chunk
    return n, dt


def cpu_baseline(w, target_s=10.0, chunk=40, target_1t=4.0, chunk_1t=4):
    """the reference's OpenMP evolve_wave_half on the host cores, a bounded
    sample of the FP64 workload: chunks of steps until target_s seconds of CPU
    work have been timed; then the same on one thread.  The reference's own
    code (oracle/_ref, compiled from its sources for this grid) when it was
    built, else the oracle's restatement of it (bit-identical results).
    Threads: the OpenMP default of this process (OMP_NUM_THREADS)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    pw = w["params"]
    cells = pw["nx"] * pw["ny"]
    args = (w["xy_in"], w["tcc"], w["tc"], w["tgamma"])
    ref = pyoracle.ref_stepper(pw["nx"], pw["ny"], pw["bcond"])
    if ref is not None:
        kind = "reference"
        what = ("the reference's evolve_wave_half (wave_billiard.c:302-610, row pointers, OpenMP parallel for over i) "
                "compiled from its sources by oracle/build_ref.sh, gcc -O3 -fopenmp")
        bufs = [np.array(w["phi"]), np.array(w["psi"]), np.zeros_like(w["phi"])]

        def step(n):
            ref.run_rows(bufs, n, *args)
    else:
        kind = "port"
        what = ("the oracle's restatement of evolve_wave_half (row pointers, OpenMP parallel for over i, "
                "bit-identical to the reference)")
        p = pyoracle.make_params(pw["nx"], pw["ny"], bcond=pw["bcond"])
        st = list(pyoracle.run(p, w["phi"], w["psi"], 2, *args))

        def step(n):
            st[:] = pyoracle.run(p, st[0], st[1], n, *args, tmp=st[2])
    threads = pyoracle.num_threads()
    step(2)   # touch the buffers
    n, dt = _time_steps(step, target_s, chunk)
    prev = pyoracle.set_num_threads(1)
    try:
        n1, dt1 = _time_steps(step, target_1t, chunk_1t)
    finally:
        pyoracle.set_num_threads(prev)
    cpu = host_cpu()
    return dict(value=cells * n / dt / 1e9, unit="Gcell-updates/s", cores=threads, kind=kind,
                sample="%d steps of the %dx%d FP64 5-point workload of the timed run (%s), %.1f s on %d threads"
                       % (n, pw["nx"], pw["ny"], what, dt, threads),
                value_1thread=cells * n1 / dt1 / 1e9, sample_1thread="%d steps, %.1f s" % (n1, dt1),
                cpu_model=cpu["model"], nproc=cpu["nproc"], physical_cores=cpu["physical_cores"],
                cpus_allowed=cpu["affinity"],
                threads_note="OMP_NUM_THREADS=%s (this process's CPU share of the machine)"
                             % os.environ.get("OMP_NUM_THREADS", "unset"))


def counters_from_profile(form, k, L, size):
    """Counter figures of the stepping kernel from the committed rocprofv3
    --pmc summary (HBM bytes per launch from FETCH_SIZE + WRITE_SIZE,
    copy-calibrated; SQ_INSTS_VALU per launch), when it was measured on this
    workload with the same tiling and the same kernel sources; otherwise None.
    A profiler run of its own: these are not measured by this run."""
    name = "counters_%s_c3_4096.json" % ("split" if form == "f32split" else "fp64")
    f = os.path.join(ROOT, "profiles", name)
    if size != 4096 or not os.path.exists(f):
        return None
    d = json.load(open(f))
    if d.get("tb_depth") != k or d.get("rows_per_tile") != L or d.get("kernel_src_sha16") != kernel_source_sha(form):
        return None
    d["source"] = "profiles/" + name
    return d


def roofline(r, form, cells):
    """the roofline object of the stepping kernel of one form.  frac: the
    bytes a fused launch must move (state read once, written once) over the
    launch time, against HBM peak -- at most 1, and low when the launch is
    bound by something else.  effective: the no-reuse bytes of the same steps
    (B_alg x cells x K), which temporal blocking lets exceed the peak.
    Counter figures (measured HBM traffic, VALU issue) come from the
    committed rocprofv3 summary of the same kernel sources and tiling."""
    t = r["avg_launch_ms"] * 1e-3
    c = r["counters"]
    ro = {
        "bound": "hbm",
        "achieved": r["achieved_Bps"] / 1e9,
        "peak": HBM_PEAK / 1e9,
        "unit": "GB/s",
        "frac": r["achieved_Bps"] / HBM_PEAK,
        "traffic": c["bytes_per_launch"] if c else None,
        "traffic_source": c["source"] if c else None,
        "traffic_frac": c["bytes_per_launch"] / t / HBM_PEAK if c and t > 0 else None,
        "kernel": ("k_step_split<%d,5,0>" if form == "f32split" else "k_step_fp64<%d,5,0>") % r["k"],
        "compulsory_bytes_per_launch": r["comp_bytes_per_launch"],
        "compulsory_bytes_per_cell": compulsory_bytes_per_cell(form, r["k"]),
        "avg_launch_ms": r["avg_launch_ms"],
        "timed_launches": r["n_timed"],
        "launches": r["launches"],
        "effective": {
            "note": "B_alg x cells x fused steps per launch over the launch time: counts temporal reuse, may exceed 1",
            "b_alg_per_cell_update": B_ALG[form],
            "bytes_per_launch": r["alg_bytes_per_launch"],
            "achieved": r["effective_Bps"] / 1e9,
            "frac": r["effective_Bps"] / HBM_PEAK,
        },
    }
    if c and c.get("valu_insts_per_launch") and t > 0:
        ro["valu"] = {"insts_per_launch": c["valu_insts_per_launch"],
                      "issue_frac": c["valu_insts_per_launch"] / t / VALU_ISSUE_PEAK,
                      "peak_insts_per_s": VALU_ISSUE_PEAK, "source": c["source"]}
    return ro


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def run_single(wh, w, form, nsteps, energy_every=0, tb_depth=0, rows=0, tune=0):
    """one grid through nsteps; energy sampled every energy_every steps (the
    samples are part of the timed work); returns state, energies, time.
    tune > 0: autotune the tiling first (tune steps per candidate; the tuner
    restores the state, and its time is not counted)"""
    W = wh.Wave(w["params"], form=form, tb_depth=tb_depth, rows_per_tile=rows)
    W.load(w)
    if tune:
        W.autotune(tune)
    # tile tables and kernel loads of the timed sequence, outside the timing
    W.prepare(energy_every or nsteps)
    W.sync()
    es = [W.energy()]
    t0 = time.perf_counter()
    if energy_every:
        for _ in range(nsteps // energy_every):
            W.step(energy_every)
            es.append(W.energy())
    else:
        W.step(nsteps)
        W.sync()
    dt = time.perf_counter() - t0
    if not energy_every:
        es.append(W.energy())
    u, um = W.get_state()
    k, L, _ = W.tiling()
    W.close()
    return dict(u=u, um=um, e=np.array(es), dt=dt, k=k, L=L)


def configs_report(wh, workloads, which):
    """every configuration in both forms: throughput, rel-L2 of the split form
    against the FP64 form, energy drift (the FP64 form is pinned bit-exact to
    the reference's evolve_wave_half by the test suite)"""
    import hashlib
    rep = {}
    if "c1" in which:
        w = workloads.c1()
        dig = json.load(open(os.path.join(ROOT, "tests", "golden", "c1_digests.json")))
        W = wh.Wave(w["params"])
        W.load(w)
        ok = True
        for n in range(100, 1001, 100):
            W.step(100)
            phi, _ = W.get_state()
            ok &= hashlib.sha256(np.ascontiguousarray(phi).tobytes()).hexdigest() == dig["phi"][str(n)]
        W.close()
        f = run_single(wh, w, wh.FORM_FP64, 1000)
        s = run_single(wh, w, wh.FORM_F32SPLIT, 1000)
        cells = w["params"]["nx"] * w["params"]["ny"]
        cpu = cpu_baseline(w, target_s=4.0, chunk=100)
        rep["c1"] = dict(grid=[1280, 720], steps=1000, fp64_bitexact_vs_reference=bool(ok),
                         gcells_fp64=cells * 1000 / f["dt"] / 1e9, gcells_split=cells * 1000 / s["dt"] / 1e9,
                         rel_l2_split=rel_l2(s["u"], f["u"]),
                         energy_drift_fp64=float((f["e"][-1] - f["e"][0]) / f["e"][0]),
                         energy_drift_split=float((s["e"][-1] - s["e"][0]) / s["e"][0]),
                         cpu_gcells=cpu["value"], cpu_cores=cpu["cores"])
        log("c1", rep["c1"])
    if "c2" in which:
        w = workloads.c2()
        cells = w["params"]["nx"] * w["params"]["ny"]
        f = run_single(wh, w, wh.FORM_FP64, 5000, 50, tune=96)
        s = run_single(wh, w, wh.FORM_F32SPLIT, 5000, 50, tune=96)
        rep["c2"] = dict(grid=[2560, 1440], steps=5000, stencil=9, energy_every=50,
                         tiling_fp64=[f["k"], f["L"]], tiling_split=[s["k"], s["L"]],
                         gcells_fp64=cells * 5000 / f["dt"] / 1e9, gcells_split=cells * 5000 / s["dt"] / 1e9,
                         rel_l2_split=rel_l2(s["u"], f["u"]),
                         energy_series_max_rel_diff=float(np.abs(s["e"] - f["e"]).max() / np.abs(f["e"]).max()),
                         energy_final_fp64=float(f["e"][-1]), energy_final_split=float(s["e"][-1]))
        log("c2", rep["c2"])
    if "c3" in which:
        w = workloads.c3()
        cells = w["params"]["nx"] * w["params"]["ny"]
        f = run_single(wh, w, wh.FORM_FP64, 10000)
        s = run_single(wh, w, wh.FORM_F32SPLIT, 10000)
        rep["c3"] = dict(grid=[4096, 4096], steps=10000,
                         gcells_fp64=cells * 10000 / f["dt"] / 1e9, gcells_split=cells * 10000 / s["dt"] / 1e9,
                         rel_l2_split=rel_l2(s["u"], f["u"]),
                         energy_drift_fp64=float((f["e"][-1] - f["e"][0]) / f["e"][0]),
                         energy_drift_split=float((s["e"][-1] - s["e"][0]) / s["e"][0]))
        log("c3", rep["c3"])
    if "c4" in which:
        ws = workloads.c4_instances()
        cells = sum(w["params"]["nx"] * w["params"]["ny"] for w in ws)
        out = {}
        for name, form in (("fp64", wh.FORM_FP64), ("split", wh.FORM_F32SPLIT)):
            B = wh.Batch([w["params"] for w in ws], form=form)
            for k, w in enumerate(ws):
                B.load(k, w)
            B.autotune(48)           # restores every grid's state
            tiling = B.tiling()[:2]
            best = None
            # each pass restarts from the initial state; the rotation of the
            # buffers returns to a captured (nsteps, rotation) graph within a
            # few passes, so the fastest pass is a pure graph replay
            for _ in range(5):
                for k, w in enumerate(ws):
                    B.set_state(k, w["phi"], w["psi"])   # keeps the tables and the graphs
                e0 = B.energy()
                t0 = time.perf_counter()
                B.step(2000)
                e1 = B.energy()   # waits for the stream
                dt = time.perf_counter() - t0
                best = dt if best is None else min(best, dt)
            states = [B.get_state(k)[0] for k in range(len(ws))]
            B.close()
            out[name] = dict(e0=e0, e1=e1, u=states, dt=best, tiling=list(tiling))
        rels = [rel_l2(a, b) for a, b in zip(out["split"]["u"], out["fp64"]["u"])]
        dr = lambda o: (o["e1"] - o["e0"]) / o["e0"]
        rep["c4"] = dict(instances=len(ws), total_cells=cells, steps=2000,
                         tiling_fp64=out["fp64"]["tiling"], tiling_split=out["split"]["tiling"],
                         gcells_fp64=cells * 2000 / out["fp64"]["dt"] / 1e9,
                         gcells_split=cells * 2000 / out["split"]["dt"] / 1e9,
                         rel_l2_split_max=float(max(rels)), rel_l2_split_median=float(np.median(rels)),
                         energy_drift_fp64_median=float(np.median(dr(out["fp64"]))),
                         energy_drift_diff_max=float(np.abs(dr(out["split"]) - dr(out["fp64"])).max()))
        log("c4", rep["c4"])
    if "c5" in which:
        w = workloads.c5()
        cells = w["params"]["nx"] * w["params"]["ny"]
        f = run_single(wh, w, wh.FORM_FP64, 20000, 100, tune=48)
        s = run_single(wh, w, wh.FORM_F32SPLIT, 20000, 100, tune=48)
        rep["c5"] = dict(grid=[8192, 8192], steps=20000, stencil=9, energy_every=100, sources=20,
                         gcells_fp64=cells * 20000 / f["dt"] / 1e9, gcells_split=cells * 20000 / s["dt"] / 1e9,
                         rel_l2_split=rel_l2(s["u"], f["u"]),
                         energy_series_max_rel_diff=float(np.abs(s["e"] - f["e"]).max() / np.abs(f["e"]).max()),
                         tiling_fp64=[f["k"], f["L"]], tiling_split=[s["k"], s["L"]])
        log("c5", rep["c5"])
    return rep


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--size", type=int, default=4096)
    ap.add_argument("--tb-depth", type=int, default=0)
    ap.add_argument("--rows", type=int, default=0)
    ap.add_argument("--tune", action="store_true", help="autotune tiling before timing")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-frame", action="store_true")
    ap.add_argument("--only-fp64", action="store_true")
    ap.add_argument("--configs", default="", help="comma list of c1..c5: run those workloads in both "
                    "forms and print their report instead of the benchmark line")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the stepper runs each grid on a single MI355X and has "
              "no multi-GPU path; run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import wavehip as wh
    from wavehip import workloads

    if args.configs:
        rep = configs_report(wh, workloads, args.configs.split(","))
        print(json.dumps(dict(configs=rep)), flush=True)
        return

    log("building workload %d^2 ..." % args.size)
    w = workloads.c3(n=args.size, nsteps=args.steps)
    cells = w["params"]["nx"] * w["params"]["ny"]
    res = {}
    forms = []
    if not args.only_fp64:
        forms.append(("f32split", wh.FORM_F32SPLIT))
    if not args.no_fp64 or args.only_fp64:
        forms.append(("fp64", wh.FORM_FP64))
    for name, form in forms:
        log("running %s: %d warmup + %d timed steps" % (name, args.warmup, args.steps))
        r = run_form(wh, w, form, args.steps, args.warmup, args.tb_depth, args.rows, args.tune)
        r["gcells"] = cells * args.steps / r["wall"] / 1e9
        r["gcells_dev"] = cells * args.steps / (r["ms_dev"] * 1e-3) / 1e9
        # the stepping kernel: one launch per fused block of k steps; one launch
        # in eight of the timed run is bracketed by a HIP event pair on the
        # stepping stream
        r["avg_launch_ms"] = r["avg_ms"]
        # compulsory bytes per launch: the state read once and written once
        r["comp_bytes_per_launch"] = compulsory_bytes_per_cell(name, r["k"]) * cells
        r["achieved_Bps"] = r["comp_bytes_per_launch"] / (r["avg_launch_ms"] * 1e-3)
        # B_alg x cells x fused steps: the bytes the same steps would move
        # without temporal reuse (one sweep per step)
        r["alg_bytes_per_launch"] = B_ALG[name] * cells * r["k"]
        r["effective_Bps"] = r["alg_bytes_per_launch"] / (r["avg_launch_ms"] * 1e-3)
        r["counters"] = counters_from_profile(name, r["k"], r["L"], args.size)
        log("%s: %.1f Gcell/s (wall), %.1f (events), k=%d rows=%d, %d launches, %.4f ms/launch"
            % (name, r["gcells"], r["gcells_dev"], r["k"], r["L"], r["launches"], r["avg_launch_ms"]))
        res[name] = r

    head = res.get("f32split") or res["fp64"]
    hname = "f32split" if "f32split" in res else "fp64"
    out = {
        "metric": METRIC,
        "value": round(head["gcells"], 3),
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["wall"] * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if hname == "f32split" else "f64",
        "data": "synthetic",
        "config": {
            "workload": "precision form %dx%d: 64 seeded disc obstacles, 3 Gaussian pulses, courant 0.1, "
                        "absorbing edges, 5-point, %s state, %d leapfrog steps per run"
                        % (args.size, args.size, "FP32 double-float (u_hi, u_lo, v)" if hname == "f32split"
                           else "FP64", args.steps),
            "grid": [args.size, args.size],
            "form": hname,
            "tb_depth": head["k"],
            "rows_per_tile": head["L"],
            "tile_cols": head["cols"],
        },
        "roofline": roofline(head, hname, cells),
    }
    if "fp64" in res:
        f = res["fp64"]
        out["fp64"] = {
            "gcells": round(f["gcells"], 3),
            "gcells_events": round(f["gcells_dev"], 3),
            "ms_per_step": f["wall"] * 1e3 / args.steps,
            "tb_depth": f["k"], "rows_per_tile": f["L"],
            "roofline": roofline(f, "fp64", cells),
            "energy_drift": (f["e1"] - f["e0"]) / f["e0"],
        }
    if "fp64" in res and "f32split" in res:
        s, f = res["f32split"], res["fp64"]
        out["split_vs_fp64"] = {
            "speedup": s["gcells"] / f["gcells"],
            "rel_l2_phi": float(np.linalg.norm(s["u"] - f["u"]) / np.linalg.norm(f["u"])),
            "energy_drift_split": (s["e1"] - s["e0"]) / s["e0"],
            "energy_drift_fp64": (f["e1"] - f["e0"]) / f["e0"],
            "energy_rel_diff": (s["e1"] - f["e1"]) / f["e0"],
            "steps": args.steps,
        }
    if not args.no_frame:
        log("timing the frame consumer ...")
        try:
            out["frame"] = frame_bench(wh, w)
        except Exception as ex:  # the frame is an extra: report, do not lose the line
            out["frame"] = {"error": str(ex)}
    if not args.no_frame:
        log("timing the media builder ...")
        try:
            out["media"] = media_bench(wh)
        except Exception as ex:  # an extra: report, do not lose the line
            out["media"] = {"error": str(ex)}
    if not args.no_cpu:
        log("timing the CPU baseline ...")
        out["cpu_baseline"] = cpu_baseline(w)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
