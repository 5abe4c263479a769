"""Deterministic inputs for the golden fixtures (SplitMix64 counter form, so
the same arrays are rebuilt on any machine without storing them)."""
import zlib

import numpy as np

GOLD = np.uint64(0x9E3779B97F4A7C15)


def splitmix64_stream(seed, n):
    """n successive SplitMix64 outputs from state `seed` (Steele et al. 2014)."""
    with np.errstate(over="ignore"):
        k = np.arange(1, n + 1, dtype=np.uint64)
        z = np.uint64(seed) + k * GOLD
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def uniform(seed, shape, a, b):
    n = int(np.prod(shape))
    u = (splitmix64_stream(seed, n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    return (a + (b - a) * u).reshape(shape)


def case_seed(tag):
    return zlib.crc32(tag.encode())


def case_inputs(tag, nx, ny):
    """xy_in in {-1,0,1,2} (10/20/50/20 %), tcc U[0,0.2], tc U[0,1],
    tgamma U[0,1e-3], phi/psi/tmp U[-1,1]."""
    s = case_seed(tag)
    r = uniform(s, (nx, ny), 0.0, 1.0)
    xy = np.where(r < 0.1, -1, np.where(r < 0.3, 0, np.where(r < 0.8, 1, 2))).astype(np.int16)
    return dict(
        xy_in=xy,
        tcc=uniform(s + 1, (nx, ny), 0.0, 0.2),
        tc=uniform(s + 2, (nx, ny), 0.0, 1.0),
        tgamma=uniform(s + 3, (nx, ny), 0.0, 1e-3),
        phi=uniform(s + 4, (nx, ny), -1.0, 1.0),
        psi=uniform(s + 5, (nx, ny), -1.0, 1.0),
        tmp=uniform(s + 6, (nx, ny), -1.0, 1.0),
    )


KAPPAS = dict(kappa=1.0e-3, kappa_sides=5.0e-4, gamma_sides=1.0e-4, kappa_topbot=2.0e-4,
              gamma_topbot=1.0e-7)


def small_cases():
    """(tag, nx, ny, bcond, twospeeds, oscillate_topbot, tb_shift) for every
    small reference build used as a fixture."""
    out = []
    for nx, ny in ((24, 17), (16, 29)):
        for bc in range(4):
            for ts in range(2):
                out.append(("row_%dx%d_bc%d_ts%d" % (nx, ny, bc, ts), nx, ny, bc, ts, 0, nx))
        for ts in range(2):
            out.append(("flat_%dx%d_bc4_ts%d" % (nx, ny, ts), nx, ny, 4, ts, 0, ny - 2))
        out.append(("row_%dx%d_bc0_ts0_tb" % (nx, ny), nx, ny, 0, 0, 1, nx))
        # wave_3d.c:615-616: tb_shift = NX, then NY-2 once it reaches NX-1
        out.append(("flat_%dx%d_bc2_ts1_tb" % (nx, ny), nx, ny, 2, 1, 1, ny - 2))
    return out


def ext_cases():
    """(tag, nx, ny, bcond, twospeeds) of the reference builds with Neumann
    ghost points, the oscillating left edge and FLOOR (VMAX 0.9)."""
    return [("row_16x29_bc%d_ts%d_ext" % (bc, ts), 16, 29, bc, ts) for bc, ts in ((0, 0), (1, 0), (2, 0), (2, 1))]


EXT_VMAX = 0.9
EXT_STEPS = 9


def ext_inputs(tag, nx, ny):
    """case_inputs with the fields halved (inside +-VMAX) and 12 Neumann ghost
    points at distinct random cells, each with two neighbours among its 3x3
    block inside the grid and weights (w, 1-w)."""
    d = case_inputs(tag, nx, ny)
    for k in ("phi", "psi", "tmp"):
        d[k] = d[k] * 0.5
    u = uniform(case_seed(tag) + 7, (4096,), 0.0, 1.0)
    pos = 0

    def nxt():
        nonlocal pos
        pos += 1
        return u[pos - 1]

    taken, nb = set(), {k: [] for k in ("i", "j", "nbi1", "nbj1", "nbi2", "nbj2", "nx", "ny")}
    while len(nb["i"]) < 12:
        i, j = int(nxt() * nx), int(nxt() * ny)
        if (i, j) in taken:
            continue
        taken.add((i, j))
        near = [(i + a, j + b) for a in (-1, 0, 1) for b in (-1, 0, 1) if 0 <= i + a < nx and 0 <= j + b < ny]
        p1, p2 = near[int(nxt() * len(near))], near[int(nxt() * len(near))]
        w = nxt()
        for k, v in zip(("i", "j", "nbi1", "nbj1", "nbi2", "nbj2", "nx", "ny"), (i, j, p1[0], p1[1], p2[0], p2[1], w, 1.0 - w)):
            nb[k].append(v)
    d["neumann"] = {k: np.array(v, np.float64 if k in ("nx", "ny") else np.int32) for k, v in nb.items()}
    # the cells the reference drives: (0, j) for 0 < j < NY-1, then the corners (0, 0) and (0, NX-1)
    d["drive_cells"] = [(0, j) for j in range(1, ny - 1)] + [(0, 0), (0, nx - 1)]
    return d


TABLE_STEPS = 7


def table_cases():
    """(tag, nx, ny, bcond, twospeeds, oscillate_topbot, extras) of the
    reference builds of the neighbour-table stepper (evolve_wave_half_new);
    extras: the oscillating left edge and corners plus FLOOR (VMAX 0.9)."""
    out = []
    for nx, ny in ((24, 17), (16, 29)):
        for bc in (0, 1, 2, 3, 4, 10):
            out.append(("tbl_%dx%d_bc%d_ts0" % (nx, ny, bc), nx, ny, bc, 0, 0, 0))
    out += [("tbl_24x17_bc1_ts1", 24, 17, 1, 1, 0, 0), ("tbl_24x17_bc0_ts0_tb", 24, 17, 0, 0, 1, 0),
            ("tbl_16x29_bc3_ts0_tb", 16, 29, 3, 0, 1, 0), ("tbl_16x29_bc2_ts0_ext", 16, 29, 2, 0, 0, 1)]
    return out


def table_inputs(tag, nx, ny, twospeeds, extras):
    """case_inputs; with TWOSPEEDS no cell has xy_in == 0 (the reference reads
    an unset Laplacian there); with extras the fields are halved and the
    driven cells listed: (0, j) for 0 < j < NY-1, then (0, 0) and (0, NY-1)."""
    d = case_inputs(tag, nx, ny)
    if twospeeds:
        d["xy_in"] = np.where(d["xy_in"] == 0, 1, d["xy_in"]).astype(np.int16)
    if extras:
        for k in ("phi", "psi", "tmp"):
            d[k] = d[k] * 0.5
        d["drive_cells"] = [(0, j) for j in range(1, ny - 1)] + [(0, 0), (0, ny - 1)]
    return d


def color_builds():
    """(tag, frame parameters) of the reference colour builds (oracle/build_ref.sh:
    24x17, E_SCALE 50, COURANT 0.1, COURANTB 0.05, NSTEPS 500, COLORDRIFT 30)."""
    base = dict(e_scale=50.0, courant=0.1, courantb=0.05, nsteps=500, colordrift=30.0)
    out = []
    for sc in (0, 1, 3, 4, 5):
        out.append(("col_s%d" % sc, dict(base, scheme=sc, asym=0, twospeeds=0)))
    out.append(("col_s1_asym", dict(base, scheme=1, asym=1, twospeeds=1)))
    out.append(("col_s3_asym", dict(base, scheme=3, asym=1, twospeeds=0)))
    out.append(("col_s3_att", dict(base, scheme=3, asym=0, twospeeds=1, color_range=1.0, attenuation=0.01,
                                   rescale_center=1)))
    return out


COLOR_FRAMES = 3


def color_inputs(tag, nx=24, ny=17):
    """mask, color_scale and the fields of frames 0..COLOR_FRAMES-1: phi
    U[-0.05, 0.05], psi = phi + U[-0.01, 0.01]"""
    s = case_seed(tag)
    r = uniform(s, (nx, ny), 0.0, 1.0)
    xy = np.where(r < 0.1, -1, np.where(r < 0.3, 0, np.where(r < 0.8, 1, 2))).astype(np.int16)
    frames = []
    for t in range(COLOR_FRAMES):
        phi = uniform(s + 10 + 2 * t, (nx, ny), -0.05, 0.05)
        frames.append((phi, phi + uniform(s + 11 + 2 * t, (nx, ny), -0.01, 0.01)))
    return dict(xy_in=xy, color_scale=uniform(s + 1, (nx, ny), 0.2, 1.0), frames=frames)


def color_cases():
    """(tag, params, palette, plot) with a colour for every cell: the schemes
    C_ONEDIM_LINEAR / C_BASIC_LINEAR set none for the asymmetric scheme
    (P_ENERGY_FLUX); palettes COL_JET, COL_HSLUV, COL_GRAY"""
    out = []
    for tag, prm in color_builds():
        for pal in (0, 1, 2):
            for plot in range(10):
                if plot == 6 and prm["scheme"] in (4, 5):
                    continue
                out.append((tag, prm, pal, plot))
    return out


def cmp_cases():
    """(tag, nx, ny, bcond, twospeeds) of the flat reference builds with
    COMPARISON (two halves; TWOSPEEDS on for odd B_COND)"""
    return [("flat_%dx%d_bc%d_cmp" % (nx, ny, bc), nx, ny, bc, bc % 2) for nx, ny in ((24, 17), (16, 29))
            for bc in range(5)]


# billiard masks (xy_in_billiard_single_domain): reference builds of a
# 97x61 and a 1280x720 grid on the box of oracle/ref_params.h, with their
# LAMBDA and MU; the circle list of D_CIRCLES is drawn here
MASK_BOX = (-2.0, 2.0, -1.125, 1.125)


def mask_cases():
    """(build tag, lambda, mu)"""
    return [("mask_97x61_l03", 0.3, 0.1), ("mask_97x61_l06_m03", 0.6, 0.3), ("mask_1280x720_l03", 0.3, 0.1)]


def mask_circles(n=40):
    """(4, n): centres in the box, radii 0.02..0.3, every fifth circle inactive"""
    xc = uniform(case_seed("mask_xc"), n, -1.9, 1.9)
    yc = uniform(case_seed("mask_yc"), n, -1.1, 1.1)
    r = uniform(case_seed("mask_r"), n, 0.02, 0.3)
    act = np.array([0.0 if k % 5 == 4 else 1.0 for k in range(n)])
    return np.ascontiguousarray(np.vstack([xc, yc, r, act]))


TIFF_DESC = b"Wave equation in a planar domain"


def tiff_cases():
    """(tag, height, width, content, description, compression): widths around
    the LZW code-width steps, the SHORT/LONG StripByteCounts switch (2184/2185
    pixels), the 10000-byte ratio checkpoint and the table reset (4096 and
    more), one to three rows, inline and out-of-line descriptions"""
    out = []
    for w in (1, 2, 85, 86, 171, 1365, 2184, 2185, 3334, 4096):
        for content in ("rand", "low", "const", "smooth"):
            if content == "rand" and w not in (1, 2, 85, 86, 171, 1365, 4096):
                continue
            rows = [(1, TIFF_DESC, 5), (2, b"abc", 5)]
            if w <= 171 or content in ("const", "low"):
                rows += [(3, b"Wave", 5), (3, b"x", 1)]
            for h, desc, comp in rows:
                out.append(("tiff_%s_%dx%d_c%d_d%d" % (content, h, w, comp, len(desc)), h, w, content, desc, comp))
    return out


def tiff_rgb(tag, h, w, content):
    """uint8 (h, w, 3) rows of a frame: uniform bytes, bytes 0..2, a constant,
    or a smooth colour ramp"""
    if content == "rand":
        return np.floor(uniform(case_seed(tag), (h, w, 3), 0.0, 256.0)).astype(np.uint8)
    if content == "low":
        return np.floor(uniform(case_seed(tag), (h, w, 3), 0.0, 3.0)).astype(np.uint8)
    if content == "const":
        return np.full((h, w, 3), 7, np.uint8)
    x = np.arange(w)[None, :] * (18.0 / max(w, 1))
    y = np.arange(h)[:, None] * 0.5
    v = np.floor(127.0 + 120.0 * np.sin(x + y)).astype(np.uint8)
    return np.ascontiguousarray(np.stack([v, 255 - v, v // 2], -1))


# init_circular_wave / add_circular_wave fixtures (oracle/build_ref.sh builds
# pulse_97x61_sinai_ts0/1): the centre of the initial drop and the (factor,
# x, y) of the pulses added after it
PULSE_INIT = (-0.5, 0.2)
PULSES = ((1.0, 0.7, -0.4), (-1.3, -1.6, 0.9), (0.35, 0.05, 0.02))


# every closed-form domain of xy_in_billiard_single_domain (the codes of
# global_pdes.c:14-236), on three reference builds (oracle/build_ref.sh dom_a,
# dom_b, dom_c) whose parameter blocks are given here; other macros keep the
# values of oracle/ref_params.h
DOMAIN_CODES = (999, 0, 1, 196, 199, 198, 197, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 15, 16, 17, 18, 19, 191, 195,
                1951, 1952, 20, 201, 202, 24, 25, 26, 31, 32, 33, 34, 50, 51, 522, 523, 524, 525, 61, 62, 63, 64, 67,
                68, 69, 73, 74, 76, 79, 791, 795, 796, 90, 91,
                35, 43, 45, 44, 441, 46, 461, 47, 48, 49, 52, 521, 526, 55, 56, 57, 58, 581, 59, 65, 66, 731, 75,
                751, 78, 781, 792, 793, 794, 797, 798, 93, 934, 935, 936, 937, 94, 942, 947, 98, 981, 982, 983, 99,
                991, 21, 22, 23, 77, 741)
# the domains of xy_in_billiard_single_domain that read lists other reference
# functions fill (polyrect, polyrectrot, polygons, polyline, the global
# circle list); their fixtures use list_data's lists on the dom_* builds
LIST_DOMAIN_CODES = (40, 41, 42, 421, 53, 531, 54, 541, 192, 193, 194, 931, 95, 97, 96, 932, 933, 941, 27)


def list_data(npoly):
    """the lists of the list-driven domains for a build with NPOLY npoly:
    rects (4, 24), every fourth with its corners in reverse order; rectrot
    (5, 16), one of length 0; polygons (6, 12), every fourth inactive;
    polyline (2, 64): npoly star points counter-clockwise, the centre at
    index npoly, then further points on a wavy circle"""
    x1 = uniform(case_seed("rect_x"), 24, -1.9, 1.5)
    y1 = uniform(case_seed("rect_y"), 24, -1.05, 0.8)
    w = uniform(case_seed("rect_w"), 24, 0.05, 0.6)
    h = uniform(case_seed("rect_h"), 24, 0.05, 0.4)
    rect = np.vstack([x1, y1, x1 + w, y1 + h])
    rect[:, 3::4] = rect[[2, 3, 0, 1]][:, 3::4]
    ax = uniform(case_seed("rrot_ax"), 16, -1.9, 1.9)
    ay = uniform(case_seed("rrot_ay"), 16, -1.0, 1.0)
    th = uniform(case_seed("rrot_th"), 16, 0.0, 6.0)
    ln = uniform(case_seed("rrot_l"), 16, 0.1, 0.9)
    ln[5] = 0.0
    rrot = np.vstack([ax, ay, ax + ln * np.cos(th), ay + ln * np.sin(th), uniform(case_seed("rrot_w"), 16, 0.02, 0.2)])
    pg = np.vstack([uniform(case_seed("pg_x"), 12, -1.7, 1.7), uniform(case_seed("pg_y"), 12, -0.9, 0.9),
                    uniform(case_seed("pg_r"), 12, 0.1, 0.5), uniform(case_seed("pg_a"), 12, 0.0, 4.0),
                    np.array([3, 4, 5, 6, 7, 8, 3, 5, 9, 4, 6, 12], np.float64),
                    np.array([0.0 if k % 4 == 3 else 1.0 for k in range(12)])])
    t = 2.0 * np.pi * np.arange(npoly) / npoly
    r = np.where(np.arange(npoly) % 2 == 0, 0.95, 0.45)
    star = np.vstack([r * np.cos(t), r * np.sin(t)])
    rest = 64 - npoly - 1
    t2 = 2.0 * np.pi * np.arange(rest) / rest
    wavy = np.vstack([(0.8 + 0.25 * np.sin(5 * t2)) * np.cos(t2), (0.8 + 0.25 * np.sin(5 * t2)) * np.sin(t2)])
    pl = np.hstack([star, np.zeros((2, 1)), wavy])
    return dict(rects=np.ascontiguousarray(rect), rectrot=np.ascontiguousarray(rrot), polygons=np.ascontiguousarray(pg),
                polyline=np.ascontiguousarray(pl))


# codes whose one-time constants the reference keeps in a static `first`
# flag shared by the function (sub_wave.c:5717): their fixtures come from a
# fresh process each, as a program that uses one domain would compute them
DOMAIN_FIRST_CODES = (56, 57, 59)


def mask_circles_b(n=25):
    """(4, n): the second circle list of a COMPARISON mask (circles_b), drawn
    in the lower half, every third circle inactive"""
    xc = uniform(case_seed("mask_b_xc"), n, -1.9, 1.9)
    yc = uniform(case_seed("mask_b_yc"), n, -1.1, 0.2)
    r = uniform(case_seed("mask_b_r"), n, 0.05, 0.35)
    act = np.array([0.0 if k % 3 == 2 else 1.0 for k in range(n)])
    return np.ascontiguousarray(np.vstack([xc, yc, r, act]))


# xy_in_billiard with COMPARISON (oracle/build_ref.sh dom_cmp_a, dom_cmp_b:
# the parameter block of dom_a): (B_DOMAIN, B_DOMAIN_B)
COMPARISON_BUILDS = {"dom_cmp_a": (20, 20), "dom_cmp_b": (73, 8)}


def domain_sets():
    """build tag -> (nx, ny, parameters for pyoracle.domain_params)"""
    return {
        "dom_a": (97, 61, dict(lam=0.6, mu=0.3, npoly=6, apoly=0.5)),
        "dom_b": (160, 90, dict(lam=1.2, mu=0.15, npoly=9, apoly=1.0, mdepth=3, julia_scale=1.0, mandellevel=200,
                                wall_width=0.075, mu_b=0.3, wall_width_b=0.15, ngridx=8, ngridy=6,
                                wall_width_asym=1.25, wall_width_asym_b=0.5, oscil_ymid=-0.3,
                                michelson_position=0.13)),
        "dom_c": (121, 77, dict(lam=0.45, mu=0.05, npoly=12, apoly=0.25, mdepth=5, mratio=5, julia_scale=1.3,
                                mandellevel=500, mandellimit=4.0, wall_width=0.04, mu_b=0.15, wall_width_b=0.05,
                                ngridx=12, ngridy=8, wall_width_asym=0.6, wall_width_asym_b=1.4, oscil_ymid=0.45,
                                michelson_position=-0.071)),
    }


# init_ior_2d's per-cell models (oracle/build_ref.sh ior_m<code>, ior_const_*):
# 97 x 61 on MASK_BOX, LAMBDA 0.8, MU 0.3, COURANT 0.1, COURANTB 0.05, GAMMA
# 1e-4, GAMMAB 2e-4, MDEPTH 6, TWOSPEEDS = code % 2; the mask they read is
# drawn here with every value the models distinguish (-1, 0, 1, 2, 3)
IOR_CODES = (1, 100, 2, 10, 11, 12, 13, 14, 15, 151, 16, 17, 18, 181, 182, 183, 19, 191)
IOR_PARAMS = dict(courant=0.1, courantb=0.05, gamma=1.0e-4, gammab=2.0e-4, lam=0.8, mu=0.3, mandel_ior_scale=-0.05,
                  mandellimit=10.0, mandellevel=100, mdepth=6)


def ior_mask(nx=97, ny=61):
    return (np.floor(uniform(case_seed("ior_mask"), (nx, ny), 0.0, 5.0)) - 1.0).astype(np.int16)


# rotation angles of the rotating-well rebuilds (ior_rot, ior_rot_large)
ROT_ANGLES = (0.0, 0.3, -1.7, 4.1)
