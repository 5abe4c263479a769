"""Synthetic inputs for the wave stepper workloads.

Every builder returns a dict with the stepper parameters, the reference
tables (xy_in, tcc, tc, tgamma; flat k = i*NY + j, i slow) and the initial
state (phi = u^n, psi = u^{n-1}), built the way the reference builds them:

  ij_to_xy          sub_wave.c:469-476   x = XMIN + i*(XMAX-XMIN)/NX (cell corner)
  xy_to_ij          sub_wave.c:412-422   i = (int)((x-XMIN)/(XMAX-XMIN)*NX)
  D_SINAI           sub_wave.c:5774-5779 x^2 + y^2 > LAMBDA^2
  D_EXT_ELLIPSE     sub_wave.c:5744-5749 x^2/LAMBDA^2 + y^2/MU^2 > 1
  D_CIRCLES         sub_wave.c:6216-6227 0 inside any disc (dist^2 < r^2)
  init_ior_2d       sub_wave.c:12983-13003 constant-media tables

Random draws use SplitMix64 (seeded, recorded) so that every run and every
machine builds the same arrays.
"""
import math

import numpy as np

MASK64 = (1 << 64) - 1


class SplitMix64:
    """SplitMix64 (Steele, Lea, Flood 2014) -- 64-bit state, one output per call."""

    def __init__(self, seed):
        self.state = seed & MASK64

    def next(self):
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def uniform(self, a=0.0, b=1.0):
        return a + (b - a) * ((self.next() >> 11) * (1.0 / 9007199254740992.0))

    def randint(self, lo, hi):
        """integer in [lo, hi] inclusive"""
        return lo + self.next() % (hi - lo + 1)


def grid_xy(nx, ny, xmin, xmax, ymin, ymax):
    """x[i], y[j] per ij_to_xy (sub_wave.c:469-476), evaluated in the same order."""
    i = np.arange(nx, dtype=np.float64)
    j = np.arange(ny, dtype=np.float64)
    x = xmin + i * (xmax - xmin) / float(nx)
    y = ymin + j * (ymax - ymin) / float(ny)
    return x, y


def xy_to_ij(x, y, nx, ny, xmin, xmax, ymin, ymax):
    """sub_wave.c:412-422 (truncation toward zero)."""
    x1 = (x - xmin) / (xmax - xmin)
    y1 = (y - ymin) / (ymax - ymin)
    return int(x1 * float(nx)), int(y1 * float(ny))


def mask_sinai(x, y, lam):
    X, Y = np.meshgrid(x, y, indexing="ij")
    return (X * X + Y * Y > lam * lam).astype(np.int16)


def mask_ext_ellipse(x, y, lam, mu):
    X, Y = np.meshgrid(x, y, indexing="ij")
    return (X * X / (lam * lam) + Y * Y / (mu * mu) > 1.0).astype(np.int16)


def mask_discs(x, y, discs):
    """discs: iterable of (xc, yc, r). 0 inside any disc, 1 elsewhere."""
    m = np.ones((x.size, y.size), dtype=np.int16)
    for (xc, yc, r) in discs:
        r2 = r * r
        # restrict to the bounding box for speed; same predicate cell by cell
        i0 = np.searchsorted(x, xc - r - 1e-9) - 1
        i1 = np.searchsorted(x, xc + r + 1e-9) + 1
        j0 = np.searchsorted(y, yc - r - 1e-9) - 1
        j1 = np.searchsorted(y, yc + r + 1e-9) + 1
        i0, j0 = max(i0, 0), max(j0, 0)
        X, Y = np.meshgrid(x[i0:i1], y[j0:j1], indexing="ij")
        inside = (X - xc) * (X - xc) + (Y - yc) * (Y - yc) < r2
        m[i0:i1, j0:j1][inside] = 0
    return m


def mask_lshape(x, y, lam):
    """D_LSHAPE, sub_wave.c:6415-6421: |x| <= LAMBDA, |y| <= 1, not x > 0 and y > 0."""
    X, Y = np.meshgrid(x, y, indexing="ij")
    out = (np.abs(X) > lam) | (np.abs(Y) > 1.0) | ((X > 0.0) & (Y > 0.0))
    return (~out).astype(np.int16)


def lshape_lines(nx, ny, lam, xmin, xmax, ymin, ymax):
    """(i1, j1, i2, j2, i3, j3) of the L-shape seams, sub_wave.c:11780-11785."""
    i1, j1 = xy_to_ij(-lam, -1.0, nx, ny, xmin, xmax, ymin, ymax)
    i2, j2 = xy_to_ij(0.0, 0.0, nx, ny, xmin, xmax, ymin, ymax)
    i3, j3 = xy_to_ij(lam, 1.0, nx, ny, xmin, xmax, ymin, ymax)
    return i1 + 1, j1 + 1, i2 - 1, j2 - 1, i3 - 1, j3 - 1


def laplacian_table(nx, ny, bcond, lsh=None):
    """The neighbour table of init_laplacian_coords (sub_wave.c:11555-11892)
    as (nneighb[nx,ny], nghb[nx,ny,4]) flat indices i*ny + j (-1: unset), for
    B_COND 0-3 and BC_LSHAPE (10, seams lsh from lshape_lines); entries are
    written in the reference's order, so a cell set twice keeps the later one."""
    nn = np.zeros((nx, ny), np.int32)
    nb = np.full((nx, ny, 4), -1, np.int32)

    def put(i, j, n, pts):
        i = np.atleast_1d(i)
        j = np.atleast_1d(j)
        nn[i, j] = n
        for q, (a, b) in enumerate(pts):
            nb[i, j, q] = np.atleast_1d(a) * ny + np.atleast_1d(b)

    I, J = np.meshgrid(np.arange(1, nx - 1), np.arange(1, ny - 1), indexing="ij")
    put(I, J, 4, [(I + 1, J), (I - 1, J), (I, J + 1), (I, J - 1)])
    if 0 <= bcond <= 3:
        per = bcond == 1
        j = np.arange(1, ny - 1)
        z, r = np.zeros_like(j), np.full_like(j, nx - 1)
        if per:
            put(z, j, 4, [(z + 1, j), (r, j), (z, j + 1), (z, j - 1)])
            put(r, j, 4, [(r - 1, j), (z, j), (r, j + 1), (r, j - 1)])
        else:
            put(z, j, 3, [(z + 1, j), (z, j + 1), (z, j - 1)])
            put(r, j, 3, [(r - 1, j), (r, j + 1), (r, j - 1)])
        i = np.arange(nx)
        ip = (i + 1) % nx if per else np.minimum(i + 1, nx - 1)
        im = (i - 1) % nx if per else np.maximum(i - 1, 0)
        top, bot = np.full_like(i, ny - 1), np.zeros_like(i)
        if per or bcond == 3:
            put(i, top, 4, [(ip, top), (im, top), (i, top - 1), (i, bot)])
            put(i, bot, 4, [(ip, bot), (im, bot), (i, bot + 1), (i, top)])
        else:
            put(i, top, 3, [(ip, top), (im, top), (i, top - 1)])
            put(i, bot, 3, [(ip, bot), (im, bot), (i, bot + 1)])
    elif bcond == 10:
        i1, j1, i2, j2, i3, j3 = lsh
        ja, jb = np.arange(j1 + 1, j2), np.arange(j2, j3 - 1)
        ia, ib = np.arange(i1, i2), np.arange(i2, i3)
        c = np.full_like
        put(c(ja, i1), ja, 4, [(c(ja, i1 + 1), ja), (c(ja, i3), ja), (c(ja, i1), ja + 1), (c(ja, i1), ja - 1)])
        put(c(jb, i1), jb, 4, [(c(jb, i1 + 1), jb), (c(jb, i2 - 1), jb), (c(jb, i1), jb + 1), (c(jb, i1), jb - 1)])
        put(c(ja, i3), ja, 4, [(c(ja, i1), ja), (c(ja, i3 - 1), ja), (c(ja, i3), ja + 1), (c(ja, i3), ja - 1)])
        put(c(jb, i2), jb, 4, [(c(jb, i1), jb), (c(jb, i2 - 1), jb), (c(jb, i2), jb + 1), (c(jb, i2), jb - 1)])
        put(ia, c(ia, j3), 4, [(ia + 1, c(ia, j3)), (ia - 1, c(ia, j3)), (ia, c(ia, j1)), (ia, c(ia, j3 - 1))])
        put(ib, c(ib, j2), 4, [(ib + 1, c(ib, j2)), (ib - 1, c(ib, j2)), (ib, c(ib, j1)), (ib, c(ib, j2 - 1))])
        put(ia, c(ia, j1), 4, [(ia + 1, c(ia, j1)), (ia - 1, c(ia, j1)), (ia, c(ia, j1 + 1)), (ia, c(ia, j3))])
        put(ib, c(ib, j1), 4, [(ib + 1, c(ib, j1)), (ib - 1, c(ib, j1)), (ib, c(ib, j1 + 1)), (ib, c(ib, j2))])
        put(i1, j1, 4, [(i1 + 1, j1), (i3 - 1, j1), (i1, j1 + 1), (i1, j3 - 1)])
        put(i3, j1, 4, [(i1, j1), (i3 - 1, j1), (i3, j1 + 1), (i3, j3)])
        put(i1, j3, 4, [(i1 + 1, j3), (i3, j3), (i1, j1), (i1, j3 - 1)])
    return nn, nb


def constant_media_tables(xy_in, courant, courantb, gamma, gammab, twospeeds):
    """init_ior_2d constant-media branch, sub_wave.c:12983-13003.  Cells the
    reference leaves uninitialised (xy_in == 0 and not TWOSPEEDS) get 0."""
    courant2 = courant * courant      # wave_billiard.c:736
    courantb2 = courantb * courantb   # wave_billiard.c:737
    tc = np.zeros(xy_in.shape, np.float64)
    tcc = np.zeros(xy_in.shape, np.float64)
    tg = np.zeros(xy_in.shape, np.float64)
    ins = xy_in != 0
    tc[ins] = courant
    tcc[ins] = courant2
    tg[ins] = np.where(xy_in[ins] == 1, gamma, gammab)
    if twospeeds:
        out = ~ins
        tc[out] = courantb
        tcc[out] = courantb2
        tg[out] = gammab
    return tcc, tc, tg


def gaussian_bumps(x, y, centres, width, amps=None):
    """sum_k a_k exp(-r_k^2 / width^2) on the grid"""
    X, Y = np.meshgrid(x, y, indexing="ij")
    f = np.zeros(X.shape, np.float64)
    for k, (xc, yc) in enumerate(centres):
        a = 1.0 if amps is None else amps[k]
        f += a * np.exp(-((X - xc) * (X - xc) + (Y - yc) * (Y - yc)) / (width * width))
    return f


def _params(nx, ny, **kw):
    p = dict(nx=nx, ny=ny, bcond=0, twospeeds=0, stencil=5, oscillate_topbot=0, tb_shift=nx,
             kappa=0.0, kappa_sides=0.0, gamma_sides=0.0, kappa_topbot=0.0, gamma_topbot=0.0,
             e_scale=1.0)
    p.update(kw)
    return p


def c1(nx=1280, ny=720):
    """Plain method, CPU-runnable: 1280x720 on [-2,2]x[-1.125,1.125], 5-point,
    FP64, courant 0.1, no damping, B_COND 0 (zero-flux outer edges), Sinai
    mask (disc of radius 0.3 removed), phi = psi = exp(-r^2/0.05^2) at
    (-1,0) in the domain, 1,000 steps."""
    box = (-2.0, 2.0, -1.125, 1.125)
    x, y = grid_xy(nx, ny, *box)
    xy = mask_sinai(x, y, 0.3)
    tcc, tc, tg = constant_media_tables(xy, 0.1, 0.1, 0.0, 0.0, 0)
    phi = gaussian_bumps(x, y, [(-1.0, 0.0)], 0.05)
    phi[xy == 0] = 0.0
    return dict(name="c1_plain_1280x720", params=_params(nx, ny, bcond=0), xy_in=xy, tcc=tcc,
                tc=tc, tgamma=tg, phi=phi, psi=phi.copy(), sources=None, nsteps=1000,
                check_every=100, box=box, courant=0.1, courantb=0.1)


def c2(nx=2560, ny=1440, nsteps=5000):
    """Plain method on GPU: 2560x1440, 9-point bulk Laplacian, TWOSPEEDS lens:
    c = 1 outside and 1/1.5 inside an ellipse (semi-axes 0.6, 0.3), damping
    1e-4, absorbing edges (KAPPA_SIDES = GAMMA_SIDES = KAPPA_TOPBOT =
    GAMMA_TOPBOT = 0), soft source sin(2 pi n / 40) at (-1.5, 0)."""
    box = (-2.0, 2.0, -1.125, 1.125)
    x, y = grid_xy(nx, ny, *box)
    xy = mask_ext_ellipse(x, y, 0.6, 0.3)
    courant, courantb = 0.1, 0.1 / 1.5
    tcc, tc, tg = constant_media_tables(xy, courant, courantb, 1e-4, 1e-4, 1)
    si, sj = xy_to_ij(-1.5, 0.0, nx, ny, *box)
    n = np.arange(1, nsteps + 1, dtype=np.float64)
    amp = np.sin(2.0 * math.pi * n / 40.0)[None, :]
    phi = np.zeros((nx, ny))
    return dict(name="c2_plain_9pt_2560x1440", params=_params(nx, ny, bcond=2, twospeeds=1, stencil=9),
                xy_in=xy, tcc=tcc, tc=tc, tgamma=tg, phi=phi, psi=phi.copy(),
                sources=dict(i=[si], j=[sj], amp=amp), nsteps=nsteps, energy_every=50, box=box,
                courant=courant, courantb=courantb)


def random_discs(rng, n, box, rmin, rmax):
    xmin, xmax, ymin, ymax = box
    out = []
    for _ in range(n):
        xc = rng.uniform(xmin, xmax)
        yc = rng.uniform(ymin, ymax)
        r = rng.uniform(rmin, rmax)
        out.append((xc, yc, r))
    return out


def c3(n=4096, nsteps=10000, seed=3):
    """Precision form: 4096^2 on [-2,2]^2, 64 random disc obstacles (radii
    U[0.02,0.1], seeded centres), three seeded Gaussian pulses (width 0.05)
    outside the discs, psi = phi, courant 0.1, no damping, absorbing edges."""
    box = (-2.0, 2.0, -2.0, 2.0)
    rng = SplitMix64(seed)
    discs = random_discs(rng, 64, box, 0.02, 0.1)
    x, y = grid_xy(n, n, *box)
    xy = mask_discs(x, y, discs)
    centres = []
    while len(centres) < 3:
        xc, yc = rng.uniform(-1.5, 1.5), rng.uniform(-1.5, 1.5)
        if all((xc - a) ** 2 + (yc - b) ** 2 >= (r + 0.1) ** 2 for (a, b, r) in discs):
            centres.append((xc, yc))
    phi = gaussian_bumps(x, y, centres, 0.05)
    phi[xy == 0] = 0.0
    tcc, tc, tg = constant_media_tables(xy, 0.1, 0.1, 0.0, 0.0, 0)
    return dict(name="c3_precision_4096", params=_params(n, n, bcond=2), xy_in=xy, tcc=tcc, tc=tc,
                tgamma=tg, phi=phi, psi=phi.copy(), sources=None, nsteps=nsteps, box=box,
                discs=discs, pulses=centres, courant=0.1, courantb=0.1)


def c4_instances(count=256, seed=4, nmin=64, nmax=1024):
    """Batching form: independent instances with NX, NY ~ U{64..1024}, 1-16
    discs of medium B (TWOSPEEDS, c_B = c_A / r, r ~ U[1,2]), damping
    U[0,1e-3], B_COND ~ U{0,1,2}, one Gaussian pulse each, courant 0.1,
    domain [-2,2] x [-2 NY/NX, 2 NY/NX] (dx = dy)."""
    rng = SplitMix64(seed)
    out = []
    for k in range(count):
        nx, ny = rng.randint(nmin, nmax), rng.randint(nmin, nmax)
        bc = rng.randint(0, 2)
        ratio = rng.uniform(1.0, 2.0)
        gamma = rng.uniform(0.0, 1e-3)
        nd = rng.randint(1, 16)
        hy = 2.0 * ny / nx
        box = (-2.0, 2.0, -hy, hy)
        discs = random_discs(rng, nd, box, 0.05, 0.3)
        x, y = grid_xy(nx, ny, *box)
        xy = mask_discs(x, y, discs)
        courant = 0.1
        tcc, tc, tg = constant_media_tables(xy, courant, courant / ratio, gamma, gamma, 1)
        xc, yc = rng.uniform(-1.0, 1.0), rng.uniform(-0.5 * hy, 0.5 * hy)
        phi = gaussian_bumps(x, y, [(xc, yc)], 0.1)
        out.append(dict(name="c4_inst%03d_%dx%d" % (k, nx, ny),
                        params=_params(nx, ny, bcond=bc, twospeeds=1), xy_in=xy, tcc=tcc, tc=tc,
                        tgamma=tg, phi=phi, psi=phi.copy(), sources=None, box=box,
                        courant=courant, courantb=courant / ratio))
    return out


def c5(n=8192, nsteps=20000, seed=5, nsrc=20):
    """Large single grid: 8192^2 on [-2,2]^2, 9-point, smooth speed field
    c = 0.5 + normalised sum of 32 seeded Gaussian bumps (c in [0.5, 1.5]),
    tcc = (0.1 c)^2 rounded to binary32 (so the FP64 and FP32 forms see the
    same medium), tc = 0.1 c, absorbing edges, 20 point sources with periods
    U[20,80] steps."""
    box = (-2.0, 2.0, -2.0, 2.0)
    rng = SplitMix64(seed)
    x, y = grid_xy(n, n, *box)
    s = np.zeros((n, n), np.float64)
    for _ in range(32):
        xc, yc = rng.uniform(-2, 2), rng.uniform(-2, 2)
        w = rng.uniform(0.2, 0.8)
        a = rng.uniform(-1.0, 1.0)
        # separable: exp(-(dx^2 + dy^2)/w^2) = exp(-dx^2/w^2) * exp(-dy^2/w^2)
        s += np.outer(a * np.exp(-((x - xc) ** 2) / (w * w)), np.exp(-((y - yc) ** 2) / (w * w)))
    s = (s - s.min()) / (s.max() - s.min())
    c = 0.5 + s
    del s
    tcc = (0.1 * c) ** 2
    tcc = tcc.astype(np.float32).astype(np.float64)
    tc = 0.1 * c
    del c
    xy = np.ones((n, n), np.int16)
    tg = np.zeros((n, n), np.float64)
    si, sj, periods = [], [], []
    for _ in range(nsrc):
        i, j = xy_to_ij(rng.uniform(-1.8, 1.8), rng.uniform(-1.8, 1.8), n, n, *box)
        si.append(i); sj.append(j); periods.append(rng.uniform(20.0, 80.0))
    t = np.arange(1, nsteps + 1, dtype=np.float64)
    amp = np.stack([0.01 * np.sin(2.0 * math.pi * t / P) for P in periods])
    phi = np.zeros((n, n))
    return dict(name="c5_large_9pt_8192", params=_params(n, n, bcond=2, stencil=9), xy_in=xy,
                tcc=tcc, tc=tc, tgamma=tg, phi=phi, psi=phi.copy(),
                sources=dict(i=si, j=sj, amp=amp), nsteps=nsteps, energy_every=100, box=box)


def random_case(nx, ny, seed, bcond=0, twospeeds=0, stencil=5, kappas=False, neg_cells=True):
    """Small random state/tables/mask for parity tests (values of xy_in in
    {-1,0,1,2}; random tables), every cell class and every edge branch."""
    rng = np.random.default_rng(seed)
    vals = np.array([-1, 0, 1, 2], np.int16) if neg_cells else np.array([0, 1, 2], np.int16)
    pr = [0.1, 0.2, 0.5, 0.2] if neg_cells else [0.2, 0.6, 0.2]
    xy = rng.choice(vals, size=(nx, ny), p=pr).astype(np.int16)
    tcc = rng.uniform(0, 0.2, (nx, ny))
    tc = rng.uniform(0, 1, (nx, ny))
    tg = rng.uniform(0, 1e-3, (nx, ny))
    phi = rng.standard_normal((nx, ny))
    psi = rng.standard_normal((nx, ny))
    kw = dict(bcond=bcond, twospeeds=twospeeds, stencil=stencil)
    if kappas:
        kw.update(kappa=1.0e-3, kappa_sides=5.0e-4, gamma_sides=1.0e-4, kappa_topbot=2.0e-4,
                  gamma_topbot=1.0e-7)
    return dict(name="random_%dx%d_s%d" % (nx, ny, seed), params=_params(nx, ny, **kw), xy_in=xy,
                tcc=tcc, tc=tc, tgamma=tg, phi=phi, psi=psi, sources=None)
