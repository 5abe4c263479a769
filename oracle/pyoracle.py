"""ctypes binding of the CPU oracle (oracle/liboracle.so) and of the reference
builds in oracle/_ref/.

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the
CPU-baseline leg of bench.py, never by the product library.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_DIR = os.path.join(HERE, "_ref")

_D = ctypes.POINTER(ctypes.c_double)
_F = ctypes.POINTER(ctypes.c_float)
_S = ctypes.POINTER(ctypes.c_short)
_I = ctypes.POINTER(ctypes.c_int)


class OrcParams(ctypes.Structure):
    """orc_params of wave_oracle.h."""
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("bcond", ctypes.c_int),
        ("twospeeds", ctypes.c_int), ("stencil", ctypes.c_int),
        ("oscillate_topbot", ctypes.c_int), ("tb_shift", ctypes.c_int),
        ("kappa", ctypes.c_double), ("kappa_sides", ctypes.c_double),
        ("gamma_sides", ctypes.c_double), ("kappa_topbot", ctypes.c_double),
        ("gamma_topbot", ctypes.c_double), ("comparison", ctypes.c_int),
    ]


def make_params(nx, ny, bcond=0, twospeeds=0, stencil=5, oscillate_topbot=0, tb_shift=None,
                kappa=0.0, kappa_sides=0.0, gamma_sides=0.0, kappa_topbot=0.0, gamma_topbot=0.0, comparison=0):
    if tb_shift is None:
        tb_shift = nx
    return OrcParams(nx, ny, bcond, twospeeds, stencil, oscillate_topbot, tb_shift,
                     kappa, kappa_sides, gamma_sides, kappa_topbot, gamma_topbot, comparison)


class OrcFrameParams(ctypes.Structure):
    """orc_frame_params of wave_oracle.h."""
    _fields_ = [(n, ctypes.c_int) for n in ("plot", "scheme", "palette", "asym", "size", "time", "shade",
                                            "rescale_center", "fade_in_obstacle", "fade")] + \
               [("nsteps", ctypes.c_long)] + \
               [(n, ctypes.c_double) for n in ("scale", "fade_value", "slope", "color_range", "attenuation",
                                               "vshift_amplitude", "vscale_amplitude", "log_shift", "log_scale",
                                               "flux_scale", "avrg_e_factor", "colorhue", "colordrift",
                                               "lummean", "lumamp", "huemean", "hueamp", "shade_scale",
                                               "light_x", "light_y", "courant", "courantb", "e_scale")]


# the drawing macros of the checked-in wave_billiard.c:200-241, 287
FRAME_DEFAULTS = dict(plot=0, scheme=3, palette=0, asym=0, size=1, time=0, shade=0, rescale_center=0,
                      fade_in_obstacle=0, fade=0, nsteps=500, scale=1.0, fade_value=1.0, slope=1.0,
                      color_range=0.8, attenuation=0.0, vshift_amplitude=-0.2, vscale_amplitude=1.0,
                      log_shift=0.75, log_scale=0.75, flux_scale=250.0, avrg_e_factor=0.95, colorhue=260.0,
                      colordrift=0.0, lummean=0.5, lumamp=0.3, huemean=180.0, hueamp=-180.0, shade_scale=0.5,
                      light_x=0.40824829, light_y=0.816496581, courant=0.1, courantb=0.05, e_scale=300.0)


def frame_params(**kw):
    d = dict(FRAME_DEFAULTS)
    d.update(kw)
    return OrcFrameParams(**d)


class OrcExtras(ctypes.Structure):
    """orc_extras of wave_oracle.h (Neumann ghost points, driven cells, FLOOR)."""
    _fields_ = [
        ("n_neu", ctypes.c_int),
        ("ni", _I), ("nj", _I), ("nbi1", _I), ("nbj1", _I), ("nbi2", _I), ("nbj2", _I),
        ("nnx", _D), ("nny", _D),
        ("n_drv", ctypes.c_int), ("di", _I), ("dj", _I), ("dval", _D), ("dlen", ctypes.c_long),
        ("vmax", ctypes.c_double),
    ]


def _p(a, t):
    return a.ctypes.data_as(t)


def _c64(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a


_lib = None


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (make -C oracle)")
        L = ctypes.CDLL(path)
        P = ctypes.POINTER(OrcParams)
        L.orc_evolve_wave_half_flat.argtypes = [P, _D, _D, _D, _S, _D, _D, _D]
        L.orc_evolve_wave_flat.argtypes = [P, _D, _D, _D, _S, _D, _D, _D]
        L.orc_run.argtypes = [P, _D, _D, _D, _I, ctypes.c_long, _S, _D, _D, _D, ctypes.c_int,
                              _I, _I, _D, ctypes.c_long, ctypes.c_long]
        L.orc_energy_cell.restype = ctypes.c_double
        L.orc_energy_cell.argtypes = [P, _D, _D, _S, ctypes.c_double, ctypes.c_double,
                                      ctypes.c_double, ctypes.c_int, ctypes.c_int]
        L.orc_energy_total.restype = ctypes.c_double
        L.orc_energy_total.argtypes = [P, _D, _D, _S, _D, ctypes.c_double]
        L.orc_fixed_order_sum.restype = ctypes.c_double
        L.orc_fixed_order_sum.argtypes = [ctypes.c_int, ctypes.c_int, _D]
        L.orc_split_from_fp64.argtypes = [ctypes.c_long, _D, _D, _F, _F, _F]
        L.orc_split_to_fp64.argtypes = [ctypes.c_long, _F, _F, _F, _D, _D]
        L.orc_split_run.argtypes = [P, _F, _F, _F, ctypes.c_long, _S, _D, _D, _D, ctypes.c_int,
                                    _I, _I, _D, ctypes.c_long, ctypes.c_long]
        L.orc_run_x.argtypes = [P, _D, _D, _D, _I, ctypes.c_long, _S, _D, _D, _D, ctypes.c_int,
                                _I, _I, _D, ctypes.c_long, ctypes.c_long, ctypes.POINTER(OrcExtras)]
        L.orc_laplacian_table.argtypes = [P, _I, _I, _I]
        L.orc_run_table.argtypes = [P, _D, _D, _D, _I, ctypes.c_long, _S, _D, _D, _D, _I, _I,
                                    ctypes.c_int, _I, _I, _D, ctypes.c_long, ctypes.c_long,
                                    ctypes.POINTER(OrcExtras)]
        L.orc_frame.argtypes = [P, ctypes.POINTER(OrcFrameParams), _D, _D, _S, _D, _D, _D, _D, _D, _F,
                                ctypes.c_int, _D, _D, ctypes.POINTER(ctypes.c_ubyte)]
        L.orc_color.argtypes = [ctypes.POINTER(OrcFrameParams), ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                ctypes.c_double, ctypes.c_int, _F, ctypes.c_int, _D]
        L.orc_mask.argtypes = [ctypes.c_int, ctypes.c_int, _D, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                               ctypes.c_int, _D, _S]
        L.orc_mask.restype = ctypes.c_int
        L.orc_wells_tables.argtypes = [P, _S, ctypes.c_int, _D, _D, _D, ctypes.c_double, ctypes.c_int,
                                       ctypes.c_double, ctypes.c_double, ctypes.c_double, _D, _D, _D, _D]
        L.orc_lzw_strip.restype = ctypes.c_long
        L.orc_lzw_strip.argtypes = [ctypes.c_void_p, ctypes.c_long, ctypes.c_void_p]
        L.orc_tiff_bound.restype = ctypes.c_long
        L.orc_tiff_bound.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_char_p]
        L.orc_tiff.restype = ctypes.c_long
        L.orc_tiff.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_char_p, ctypes.c_int,
                               ctypes.c_void_p, ctypes.c_long]
        L.orc_num_threads.restype = ctypes.c_int
        L.orc_split_energy_total.restype = ctypes.c_double
        L.orc_split_energy_total.argtypes = [P, _F, _F, _F, _S, _D, ctypes.c_double]
        _lib = L
    return _lib


def num_threads():
    return lib().orc_num_threads()


def set_num_threads(n):
    """OpenMP threads of the oracle and of the reference builds loaded in this
    process (one OpenMP runtime); returns the previous count"""
    L = lib()
    L.orc_set_num_threads.restype = ctypes.c_int
    L.orc_set_num_threads.argtypes = [ctypes.c_int]
    return L.orc_set_num_threads(int(n))


def evolve_wave_half_flat(p, phi_in, psi_in, phi_out, xy_in, tc, tcc, tgamma):
    """One step in place on phi_out (flat twin signature, wave_3d.c:600)."""
    lib().orc_evolve_wave_half_flat(ctypes.byref(p), _p(phi_in, _D), _p(psi_in, _D),
                                    _p(phi_out, _D), _p(xy_in, _S), _p(tc, _D), _p(tcc, _D),
                                    _p(tgamma, _D))


def evolve_wave_flat(p, phi, psi, tmp, xy_in, tcc, tc, tgamma):
    lib().orc_evolve_wave_flat(ctypes.byref(p), _p(phi, _D), _p(psi, _D), _p(tmp, _D),
                               _p(xy_in, _S), _p(tcc, _D), _p(tc, _D), _p(tgamma, _D))


def run(p, phi, psi, nsteps, xy_in, tcc, tc, tgamma, tmp=None, sources=None, step0=0):
    """Advance nsteps with the evolve_wave rotation. Returns (phi, psi, tmp)
    views of the three buffers after the run (new arrays; inputs untouched)."""
    b = [np.array(phi, dtype=np.float64, copy=True), np.array(psi, dtype=np.float64, copy=True),
         np.zeros_like(phi, dtype=np.float64) if tmp is None else np.array(tmp, dtype=np.float64)]
    st = ctypes.c_int(0)
    nsrc, si, sj, amp, alen = _sources(sources)
    lib().orc_run(ctypes.byref(p), _p(b[0], _D), _p(b[1], _D), _p(b[2], _D), ctypes.byref(st),
                  nsteps, _p(xy_in, _S), _p(tcc, _D), _p(tc, _D), _p(tgamma, _D), nsrc,
                  _p(si, _I), _p(sj, _I), _p(amp, _D), alen, step0)
    s = st.value
    return b[s], b[(s + 1) % 3], b[(s + 2) % 3]


def run_x(p, phi, psi, nsteps, xy_in, tcc, tc, tgamma, tmp=None, sources=None, step0=0,
          neumann=None, drive=None, vmax=0.0):
    """orc_run with the reference's Neumann ghost points, driven cells and
    FLOOR clamp.  neumann: dict of int arrays i, j, nbi1, nbj1, nbi2, nbj2 and
    float arrays nx, ny; drive: dict i, j, val (n x len)."""
    b = [np.array(phi, dtype=np.float64, copy=True), np.array(psi, dtype=np.float64, copy=True),
         np.zeros_like(phi, dtype=np.float64) if tmp is None else np.array(tmp, dtype=np.float64)]
    st = ctypes.c_int(0)
    nsrc, si, sj, amp, alen = _sources(sources)
    keep = []

    def ia(a):
        a = np.ascontiguousarray(a, np.int32)
        keep.append(a)
        return _p(a, _I)

    def da(a):
        a = np.ascontiguousarray(a, np.float64)
        keep.append(a)
        return _p(a, _D)

    x = OrcExtras()
    z = np.zeros(1, np.int32)
    if neumann:
        x.n_neu = len(neumann["i"])
        x.ni, x.nj = ia(neumann["i"]), ia(neumann["j"])
        x.nbi1, x.nbj1 = ia(neumann["nbi1"]), ia(neumann["nbj1"])
        x.nbi2, x.nbj2 = ia(neumann["nbi2"]), ia(neumann["nbj2"])
        x.nnx, x.nny = da(neumann["nx"]), da(neumann["ny"])
    else:
        x.n_neu = 0
        x.ni = x.nj = x.nbi1 = x.nbj1 = x.nbi2 = x.nbj2 = ia(z)
        x.nnx = x.nny = da(np.zeros(1))
    if drive:
        val = np.ascontiguousarray(drive["val"], np.float64)
        x.n_drv, x.dlen = val.shape[0], val.shape[1]
        x.di, x.dj, x.dval = ia(drive["i"]), ia(drive["j"]), da(val.reshape(-1))
    else:
        x.n_drv, x.dlen = 0, 0
        x.di = x.dj = ia(z)
        x.dval = da(np.zeros(1))
    x.vmax = vmax
    lib().orc_run_x(ctypes.byref(p), _p(b[0], _D), _p(b[1], _D), _p(b[2], _D), ctypes.byref(st),
                    nsteps, _p(xy_in, _S), _p(tcc, _D), _p(tc, _D), _p(tgamma, _D), nsrc,
                    _p(si, _I), _p(sj, _I), _p(amp, _D), alen, step0, ctypes.byref(x))
    s = st.value
    return b[s], b[(s + 1) % 3], b[(s + 2) % 3]


def laplacian_table(p, lsh=(0, 0, 0, 0, 0, 0)):
    """init_laplacian_coords as (nneighb[NX,NY], nghb[NX,NY,4]) flat indices;
    lsh = (i1, j1, i2, j2, i3, j3) for BC_LSHAPE."""
    nn = np.zeros((p.nx, p.ny), np.int32)
    nb = np.zeros((p.nx, p.ny, 4), np.int32)
    ls = np.ascontiguousarray(lsh, np.int32)
    lib().orc_laplacian_table(ctypes.byref(p), _p(ls, _I), _p(nn, _I), _p(nb, _I))
    return nn, nb


def _extras(drive, vmax, keep):
    def ia(a):
        a = np.ascontiguousarray(a, np.int32)
        keep.append(a)
        return _p(a, _I)

    def da(a):
        a = np.ascontiguousarray(a, np.float64)
        keep.append(a)
        return _p(a, _D)

    x = OrcExtras()
    z = np.zeros(1, np.int32)
    x.n_neu = 0
    x.ni = x.nj = x.nbi1 = x.nbj1 = x.nbi2 = x.nbj2 = ia(z)
    x.nnx = x.nny = da(np.zeros(1))
    if drive:
        val = np.ascontiguousarray(drive["val"], np.float64)
        x.n_drv, x.dlen = val.shape[0], val.shape[1]
        x.di, x.dj, x.dval = ia(drive["i"]), ia(drive["j"]), da(val.reshape(-1))
    else:
        x.n_drv, x.dlen = 0, 0
        x.di = x.dj = ia(z)
        x.dval = da(np.zeros(1))
    x.vmax = vmax
    return x


def run_table(p, phi, psi, nsteps, xy_in, tcc, tc, tgamma, nneighb, nghb, tmp=None, sources=None,
              step0=0, drive=None, vmax=0.0):
    """nsteps of the neighbour-table stepper (evolve_wave_half_new) with the
    evolve_wave rotation, driven cells and FLOOR as run_x."""
    b = [np.array(phi, dtype=np.float64, copy=True), np.array(psi, dtype=np.float64, copy=True),
         np.zeros_like(phi, dtype=np.float64) if tmp is None else np.array(tmp, dtype=np.float64)]
    st = ctypes.c_int(0)
    nsrc, si, sj, amp, alen = _sources(sources)
    keep = []
    x = _extras(drive, vmax, keep)
    nn = np.ascontiguousarray(nneighb, np.int32)
    nb = np.ascontiguousarray(nghb, np.int32)
    lib().orc_run_table(ctypes.byref(p), _p(b[0], _D), _p(b[1], _D), _p(b[2], _D), ctypes.byref(st),
                        nsteps, _p(xy_in, _S), _p(tcc, _D), _p(tc, _D), _p(tgamma, _D), _p(nn, _I),
                        _p(nb, _I), nsrc, _p(si, _I), _p(sj, _I), _p(amp, _D), alen, step0,
                        ctypes.byref(x))
    s = st.value
    return b[s], b[(s + 1) % 3], b[(s + 2) % 3]


class FrameAccumulators:
    """the running sums of the mean / average / flux plots"""

    def __init__(self, nx, ny):
        self.te = np.zeros((nx, ny))
        self.ae = np.zeros((nx, ny))
        self.tf = np.zeros(2 * nx * ny)


def frame(p, f, phi, psi, xy_in, acc, color_scale=None, fade_table=None, lut=None):
    """orc_frame: (values, rgb[nx, ny, 3], drawn) at block cells (others 0)"""
    nx, ny = p.nx, p.ny
    phi = np.ascontiguousarray(phi, np.float64)
    psi = np.ascontiguousarray(psi, np.float64)
    xy = np.ascontiguousarray(xy_in, np.int16)
    cs = np.ascontiguousarray(np.ones((nx, ny)) if color_scale is None else color_scale, np.float64)
    ft = np.ascontiguousarray(np.ones((nx, ny)) if fade_table is None else fade_table, np.float64)
    lt = np.ascontiguousarray(np.zeros((1, 3)) if lut is None else lut, np.float32)
    values = np.zeros((nx, ny))
    rgb = np.zeros((nx, ny, 3))
    drawn = np.zeros((nx, ny), np.uint8)
    lib().orc_frame(ctypes.byref(p), ctypes.byref(f), _p(phi, _D), _p(psi, _D), _p(xy, _S), _p(acc.te, _D),
                    _p(acc.ae, _D), _p(acc.tf, _D), _p(cs, _D), _p(ft, _D), _p(lt, _F), lt.shape[0],
                    _p(values, _D), _p(rgb, _D), drawn.ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte)))
    return values, rgb, drawn.astype(bool)


def wells_tables(p, xy_in, xc, yc, h, sigma, use_tanh, courant, courantb, gamma, box):
    """orc_wells_tables: (tcc, tc, tgamma); tc is 0 where the rule leaves it"""
    nx, ny = p.nx, p.ny
    xy = np.ascontiguousarray(xy_in, np.int16)
    xc = np.ascontiguousarray(xc, np.float64)
    yc = np.ascontiguousarray(yc, np.float64)
    hh = None if h is None else np.ascontiguousarray(h, np.float64)
    bx = np.ascontiguousarray(box, np.float64)
    tcc, tc, tg = np.zeros((nx, ny)), np.zeros((nx, ny)), np.zeros((nx, ny))
    lib().orc_wells_tables(ctypes.byref(p), _p(xy, _S), len(xc), _p(xc, _D), _p(yc, _D),
                           None if hh is None else _p(hh, _D), sigma, int(use_tanh), courant, courantb, gamma,
                           _p(bx, _D), _p(tcc, _D), _p(tc, _D), _p(tg, _D))
    return tcc, tc, tg


def circular_wave(p, mode, factor, x0, y0, amp, variance, wavelength, box, phi, psi, xy_in):
    """init_circular_wave (mode 0) / add_circular_wave (mode 1) on copies of
    phi, psi; returns (phi, psi)"""
    L = lib()
    L.orc_circular_wave.argtypes = [ctypes.POINTER(OrcParams), ctypes.c_int] + [ctypes.c_double] * 6 + \
        [_D, _D, _D, _S]
    a = np.array(phi, dtype=np.float64, copy=True)
    b = np.array(psi, dtype=np.float64, copy=True)
    bx = np.ascontiguousarray(box, np.float64)
    xy = np.ascontiguousarray(xy_in, np.int16)
    L.orc_circular_wave(ctypes.byref(p), mode, factor, x0, y0, amp, variance, wavelength, _p(bx, _D), _p(a, _D),
                        _p(b, _D), _p(xy, _S))
    return a, b


def circle_list(xc, yc, r, active=None):
    """the (4, n) circle array of the mask builders: xc, yc, radius, active"""
    xc = np.asarray(xc, np.float64)
    a = np.ones(len(xc)) if active is None else np.asarray(active, np.float64)
    return np.ascontiguousarray(np.vstack([xc, np.asarray(yc, np.float64), np.asarray(r, np.float64), a]))


def mask(nx, ny, box, domain, lam, mu=0.0, circles=None):
    """orc_mask: init_wave_flat's xy_in (int16 (nx, ny)) for D_SINAI (3),
    D_EXT_ELLIPSE (199), D_CIRCLES (20, circles from circle_list)"""
    c = np.zeros((4, 1)) if circles is None else np.ascontiguousarray(circles, np.float64)
    n = 0 if circles is None else c.shape[1]
    out = np.zeros((nx, ny), np.int16)
    bx = np.ascontiguousarray(box, np.float64)
    rc = lib().orc_mask(nx, ny, _p(bx, _D), domain, lam, mu, n, _p(c, _D), _p(out, _S))
    if rc:
        raise ValueError("domain %d not provided" % domain)
    return out


class DomainParams(ctypes.Structure):
    """orc_domain_params: the macros the closed-form domains read"""
    _fields_ = [(n, ctypes.c_double) for n in ("lam", "mu", "mu_b", "apoly", "wall_width", "wall_width_b",
                                                "julia_scale", "julia_x", "julia_y", "mandellimit")] + \
               [(n, ctypes.c_int) for n in ("npoly", "mdepth", "mratio", "mandellevel", "ngridx", "ngridy")] + \
               [(n, ctypes.c_double) for n in ("wall_width_asym", "wall_width_asym_b", "oscil_ymid",
                                                "michelson_position")]


# the values of the fixture builds' parameter block (oracle/ref_params.h) and
# julia_x, julia_y of global_pdes.c:567
DOMAIN_DEFAULTS = dict(lam=1.5, mu=0.1, mu_b=0.08, apoly=0.0, wall_width=0.1, wall_width_b=0.1, julia_scale=0.8,
                       julia_x=0.37468, julia_y=0.21115, mandellimit=10.0, npoly=6, mdepth=4, mratio=3,
                       mandellevel=100, ngridx=10, ngridy=10, wall_width_asym=0.75, wall_width_asym_b=0.75,
                       oscil_ymid=0.0, michelson_position=0.0)


def domain_params(**kw):
    d = dict(DOMAIN_DEFAULTS)
    bad = set(kw) - set(d)
    if bad:
        raise KeyError(bad)
    d.update(kw)
    return DomainParams(**d)


class DomainLists(ctypes.Structure):
    """orc_domain_lists: the lists of the list-driven domains, each a count
    and a flat array of that many values per field"""
    _fields_ = [("ncircles", ctypes.c_int), ("circles", _D), ("nrects", ctypes.c_int), ("rects", _D),
                ("nrectrot", ctypes.c_int), ("rectrot", _D), ("npolygons", ctypes.c_int), ("polygons", _D),
                ("npolyline", ctypes.c_int), ("polyline", _D)]


# rows of each list: circles xc, yc, radius, active; rects x1, y1, x2, y2;
# rectrot x1, y1, x2, y2, width; polygons xc, yc, radius, angle, nsides,
# active; polyline x, y
LIST_ROWS = dict(circles=4, rects=4, rectrot=5, polygons=6, polyline=2)


def domain_lists(**lists):
    """DomainLists from (rows, n) arrays; the arrays stay alive on the struct"""
    bad = set(lists) - set(LIST_ROWS)
    if bad:
        raise KeyError(bad)
    s = DomainLists()
    keep = []
    for name, rows in LIST_ROWS.items():
        a = lists.get(name)
        a = np.zeros((rows, 0)) if a is None else np.ascontiguousarray(a, np.float64)
        if a.ndim != 2 or a.shape[0] != rows:
            raise ValueError("%s: (%d, n) expected" % (name, rows))
        keep.append(a)
        setattr(s, "n" + name, a.shape[1])
        setattr(s, name, _p(a, _D) if a.shape[1] else None)
    s._keep = keep
    return s


def list_domain(nx, ny, box, dom, params, **lists):
    """orc_list_domain: xy_in_billiard_single_domain over the grid for the
    domains that read polyrect, polyrectrot, polygons, polyline or the global
    circle list; lists as for domain_lists"""
    L = lib()
    L.orc_list_domain.argtypes = [ctypes.c_int, ctypes.c_int, _D, ctypes.c_int, ctypes.POINTER(DomainParams),
                                  ctypes.POINTER(DomainLists), _S]
    dl = domain_lists(**lists)
    out = np.zeros((nx, ny), np.int16)
    bx = np.ascontiguousarray(box, np.float64)
    if L.orc_list_domain(nx, ny, _p(bx, _D), dom, ctypes.byref(params), ctypes.byref(dl), _p(out, _S)):
        raise ValueError("domain %d not provided, or its lists too short" % dom)
    return out


def domain(nx, ny, box, dom, params, circles=None):
    """orc_domain: xy_in_billiard_single_domain over the grid (int16 (nx, ny))
    for the closed-form domains; params from domain_params"""
    L = lib()
    L.orc_domain.argtypes = [ctypes.c_int, ctypes.c_int, _D, ctypes.c_int, ctypes.POINTER(DomainParams),
                             ctypes.c_int, _D, _S]
    c = np.zeros((4, 1)) if circles is None else np.ascontiguousarray(circles, np.float64)
    n = 0 if circles is None else c.shape[1]
    out = np.zeros((nx, ny), np.int16)
    bx = np.ascontiguousarray(box, np.float64)
    if L.orc_domain(nx, ny, _p(bx, _D), dom, ctypes.byref(params), n, _p(c, _D), _p(out, _S)):
        raise ValueError("domain %d not provided" % dom)
    return out


def comparison_mask(nx, ny, box, dom, dom_b, params, circles=None, circles_b=None):
    """orc_comparison_mask: xy_in_billiard with COMPARISON (dom and circles
    where y > 0, dom_b and circles_b where y <= 0)"""
    L = lib()
    L.orc_comparison_mask.argtypes = [ctypes.c_int, ctypes.c_int, _D, ctypes.c_int, ctypes.c_int,
                                      ctypes.POINTER(DomainParams), ctypes.c_int, _D, ctypes.c_int, _D, _S]
    cs = [np.zeros((4, 1)) if c is None else np.ascontiguousarray(c, np.float64) for c in (circles, circles_b)]
    ns = [0 if c is None else cs[q].shape[1] for q, c in enumerate((circles, circles_b))]
    out = np.zeros((nx, ny), np.int16)
    bx = np.ascontiguousarray(box, np.float64)
    if L.orc_comparison_mask(nx, ny, _p(bx, _D), dom, dom_b, ctypes.byref(params), ns[0], _p(cs[0], _D), ns[1],
                             _p(cs[1], _D), _p(out, _S)):
        raise ValueError("domain %d or %d not provided" % (dom, dom_b))
    return out


class IorParams(ctypes.Structure):
    """orc_ior_params: the macros init_ior_2d reads"""
    _fields_ = [(n, ctypes.c_double) for n in ("courant", "courantb", "gamma", "gammab", "lam", "mu",
                                                "mandel_ior_scale", "mandellimit")] + \
               [(n, ctypes.c_int) for n in ("mandellevel", "mdepth")]


def ior_tables(p, xy_in, ior, box, params):
    """orc_ior_tables: (tcc, tc, tgamma) of init_ior_2d's per-cell model ior"""
    L = lib()
    L.orc_ior_tables.argtypes = [ctypes.POINTER(OrcParams), _S, ctypes.c_int, _D, ctypes.POINTER(IorParams),
                                 _D, _D, _D]
    xy = np.ascontiguousarray(xy_in, np.int16)
    bx = np.ascontiguousarray(box, np.float64)
    t = [np.zeros((p.nx, p.ny)) for _ in range(3)]
    if L.orc_ior_tables(ctypes.byref(p), _p(xy, _S), ior, _p(bx, _D), ctypes.byref(params), *[_p(a, _D) for a in t]):
        raise ValueError("IOR %d not provided" % ior)
    return tuple(t)


def rotating_wells(ngridx, ngridy, large, angle, box):
    """orc_rotating_wells: (xc, yc, sigma)"""
    L = lib()
    L.orc_rotating_wells.restype = ctypes.c_double
    L.orc_rotating_wells.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, _D, _D, _D]
    xc, yc = np.zeros(ngridx * ngridy), np.zeros(ngridx * ngridy)
    bx = np.ascontiguousarray(box, np.float64)
    s = L.orc_rotating_wells(ngridx, ngridy, int(large), angle, _p(bx, _D), _p(xc, _D), _p(yc, _D))
    return xc, yc, s


def lzw_strip(data):
    """orc_lzw_strip: one strip coded as libtiff's LZW coder codes it"""
    src = np.frombuffer(bytes(data), np.uint8).copy()
    out = np.zeros(len(src) + len(src) // 2 + 64, np.uint8)
    n = lib().orc_lzw_strip(src.ctypes.data if len(src) else None, len(src), out.ctypes.data)
    return out[:n].tobytes()


def tiff(rgb, description=b"Wave equation in a planar domain", compression=5):
    """orc_tiff: the file writetiff writes for the glReadPixels rows rgb
    (uint8 (height, width, 3), bottom row first)"""
    rgb = np.ascontiguousarray(rgb, np.uint8)
    h, w = rgb.shape[0], rgb.shape[1]
    d = description.encode() if isinstance(description, str) else bytes(description)
    cap = lib().orc_tiff_bound(w, h, d)
    out = np.zeros(cap, np.uint8)
    n = lib().orc_tiff(rgb.ctypes.data, w, h, d, compression, out.ctypes.data, cap)
    if n < 0:
        raise ValueError("frame %dx%d, compression %d not provided" % (w, h, compression))
    return out[:n].tobytes()


def _sources(sources):
    if not sources:
        z = np.zeros(1, np.int32)
        return 0, z, z, np.zeros(1), 0
    si = np.ascontiguousarray(sources["i"], np.int32)
    sj = np.ascontiguousarray(sources["j"], np.int32)
    amp = np.ascontiguousarray(sources["amp"], np.float64)
    return len(si), si, sj, amp.reshape(-1), amp.shape[1]


def energy_cell(p, phi, psi, xy_in, courant, courantb, e_scale, i, j):
    return lib().orc_energy_cell(ctypes.byref(p), _p(phi, _D), _p(psi, _D), _p(xy_in, _S),
                                 courant, courantb, e_scale, i, j)


def energy_total(p, phi, psi, xy_in, tcc, e_scale=1.0):
    return lib().orc_energy_total(ctypes.byref(p), _p(phi, _D), _p(psi, _D), _p(xy_in, _S),
                                  _p(tcc, _D), e_scale)


def fixed_order_sum(e):
    e = np.ascontiguousarray(e, np.float64)
    return lib().orc_fixed_order_sum(e.shape[0], e.shape[1], _p(e, _D))


def split_from_fp64(u_n, u_nm1):
    n = u_n.size
    hi = np.empty(u_n.shape, np.float32); lo = np.empty_like(hi); v = np.empty_like(hi)
    lib().orc_split_from_fp64(n, _p(_c64(u_n), _D), _p(_c64(u_nm1), _D), _p(hi, _F), _p(lo, _F),
                              _p(v, _F))
    return hi, lo, v


def split_to_fp64(hi, lo, v):
    u = np.empty(hi.shape, np.float64); um = np.empty_like(u)
    lib().orc_split_to_fp64(hi.size, _p(hi, _F), _p(lo, _F), _p(v, _F), _p(u, _D), _p(um, _D))
    return u, um


def split_run(p, hi, lo, v, nsteps, xy_in, tcc, tc, tgamma, sources=None, step0=0):
    hi = np.array(hi, np.float32, copy=True); lo = np.array(lo, np.float32, copy=True)
    v = np.array(v, np.float32, copy=True)
    nsrc, si, sj, amp, alen = _sources(sources)
    lib().orc_split_run(ctypes.byref(p), _p(hi, _F), _p(lo, _F), _p(v, _F), nsteps, _p(xy_in, _S),
                        _p(tcc, _D), _p(tc, _D), _p(tgamma, _D), nsrc, _p(si, _I), _p(sj, _I),
                        _p(amp, _D), alen, step0)
    return hi, lo, v


def split_energy_total(p, hi, lo, v, xy_in, tcc, e_scale=1.0):
    return lib().orc_split_energy_total(ctypes.byref(p), _p(hi, _F), _p(lo, _F), _p(v, _F),
                                        _p(xy_in, _S), _p(tcc, _D), e_scale)


# ---------------------------------------------------------------- reference --

def ref_stepper(nx, ny, bcond):
    """the reference's row-pointer stepper built for an nx x ny grid with this
    B_COND (no damping, one speed: the precision workload's build), or None"""
    try:
        return RefBuild("row_%dx%d_bc%d_ts0" % (nx, ny, bcond))
    except FileNotFoundError:
        return None


def ref_available():
    return os.path.isdir(REF_DIR) and any(f.endswith(".so") for f in os.listdir(REF_DIR))


class RefBuild:
    """One reference build from oracle/build_ref.sh (compile-time NX, NY, B_COND...)."""

    def __init__(self, tag):
        path = os.path.join(REF_DIR, "libref_%s.so" % tag)
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        self.L = ctypes.CDLL(path)
        self.nx = self.L.ref_nx()
        self.ny = self.L.ref_ny()
        self.flat = hasattr(self.L, "ref_evolve_wave_half_flat")
        self.table = hasattr(self.L, "ref_evolve_wave_half_new")
        if hasattr(self.L, "ref_init_circular_wave"):
            self.L.ref_init_circular_wave.argtypes = [ctypes.c_double, ctypes.c_double, _D, _D, _S]
            self.L.ref_add_circular_wave.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double, _D, _D, _S]
            return
        if hasattr(self.L, "ref_mask"):
            self.L.ref_mask.argtypes = [ctypes.c_int, ctypes.c_int, _D, _D, _D, _I, _S]
            self.L.ref_mask.restype = ctypes.c_int
            if hasattr(self.L, "ref_mask_billiard"):
                self.L.ref_mask_billiard.argtypes = [ctypes.c_int, _D, _D, _D, _I, ctypes.c_int, _D, _D, _D, _I, _S]
                self.L.ref_mask_billiard.restype = ctypes.c_int
            if hasattr(self.L, "ref_set_michelson"):
                self.L.ref_set_michelson.argtypes = [ctypes.c_double]
                self.L.ref_set_michelson.restype = None
            if hasattr(self.L, "ref_set_lists"):
                self.L.ref_set_lists.argtypes = [ctypes.c_int, _D, ctypes.c_int, _D, ctypes.c_int, _D, ctypes.c_int,
                                                 _D]
                self.L.ref_set_lists.restype = ctypes.c_int
            return
        if hasattr(self.L, "ref_init_ior"):
            self.L.ref_init_ior.argtypes = [_S, _D, _D, _D, ctypes.c_uint]
            self.L.ref_init_ior_at.argtypes = [_S, _D, _D, _D, ctypes.c_uint, ctypes.c_double]
            self.L.ref_wells.restype = ctypes.c_double
            self.L.ref_wells.argtypes = [ctypes.c_uint, _D, _D, _D]
            return
        self.color = hasattr(self.L, "ref_wave_value")
        if self.color:
            self.L.ref_wave_value.restype = ctypes.c_double
            self.L.ref_wave_value.argtypes = [_D, _D, _D, _D, _D, _D, _S, ctypes.c_double, ctypes.c_int,
                                              ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _D]
            self.L.ref_palette_table.argtypes = [ctypes.c_int, _F]
            self.L.ref_palette_table.restype = ctypes.c_int
            return
        if self.table:
            self.L.ref_evolve_wave_half_new.argtypes = [_D, _D, _D, _S, _D, _D, _D]
            self.L.ref_laplacian_table.argtypes = [_I, _I]
            self.L.ref_lshape_lines.argtypes = [_I]
            self.L.ref_corner_drive.restype = ctypes.c_double
            self.L.ref_corner_drive.argtypes = [ctypes.c_int]
        elif self.flat:
            self.L.ref_evolve_wave_half_flat.argtypes = [_D, _D, _D, _S, _D, _D, _D]
        else:
            self.L.ref_evolve_wave_half.argtypes = [_D, _D, _D, _S, _D, _D, _D]
            self.L.ref_evolve_wave.argtypes = [_D, _D, _D, _S, _D, _D, _D]
            self.L.ref_compute_energy.restype = ctypes.c_double
            self.L.ref_compute_energy.argtypes = [_D, _D, _S, ctypes.c_int, ctypes.c_int]
        if hasattr(self.L, "ref_set_neumann"):
            self.L.ref_set_neumann.argtypes = [ctypes.c_int] + [_I] * 6 + [_D, _D]
            self.L.ref_oscillating_bc.restype = ctypes.c_double
            self.L.ref_oscillating_bc.argtypes = [ctypes.c_int, ctypes.c_int]

    def set_neumann(self, nb):
        """hand Neumann ghost points to the reference stepper (dict as run_x)"""
        arrs = [np.ascontiguousarray(nb[k], np.int32) for k in ("i", "j", "nbi1", "nbj1", "nbi2", "nbj2")]
        fl = [np.ascontiguousarray(nb[k], np.float64) for k in ("nx", "ny")]
        rc = self.L.ref_set_neumann(len(arrs[0]), *[_p(a, _I) for a in arrs], *[_p(a, _D) for a in fl])
        assert rc == 0

    def oscillating_bc(self, time, j):
        return self.L.ref_oscillating_bc(time, j)

    def wave_value(self, phi, psi, acc, cs, xy_in, scale, time, plot, palette, i, j):
        """wave_value at (i, j): (value, rgb)"""
        rgb = np.zeros(3)
        v = self.L.ref_wave_value(_p(phi, _D), _p(psi, _D), _p(acc.te, _D), _p(acc.ae, _D), _p(acc.tf, _D),
                                  _p(cs, _D), _p(xy_in, _S), scale, time, plot, palette, i, j, _p(rgb, _D))
        return v, rgb

    def init_ior(self, xy_in, seed):
        """init_ior_2d's tables (tcc, tc, tgamma), tables zeroed first"""
        xy = np.ascontiguousarray(xy_in, np.int16)
        t = [np.zeros((self.nx, self.ny)) for _ in range(3)]
        self.L.ref_init_ior(_p(xy, _S), _p(t[0], _D), _p(t[1], _D), _p(t[2], _D), seed)
        return tuple(t)

    def init_ior_at(self, xy_in, seed, angle):
        """init_ior_2d at a rotation angle (its first call in the process
        fixes the rotating models' lattice), tables zeroed first"""
        xy = np.ascontiguousarray(xy_in, np.int16)
        t = [np.zeros((self.nx, self.ny)) for _ in range(3)]
        self.L.ref_init_ior_at(_p(xy, _S), _p(t[0], _D), _p(t[1], _D), _p(t[2], _D), seed, angle)
        return tuple(t)

    def init_circular_wave(self, x, y):
        """init_circular_wave(x, y, phi, psi, xy_in): (phi, psi, xy_in)"""
        phi, psi = np.zeros((self.nx, self.ny)), np.zeros((self.nx, self.ny))
        xy = np.zeros((self.nx, self.ny), np.int16)
        self.L.ref_init_circular_wave(x, y, _p(phi, _D), _p(psi, _D), _p(xy, _S))
        return phi, psi, xy

    def add_circular_wave(self, factor, x, y, phi, psi, xy):
        """add_circular_wave(factor, x, y, phi, psi, xy_in) on copies"""
        a, b = np.array(phi, copy=True), np.array(psi, copy=True)
        xy = np.ascontiguousarray(xy, np.int16)
        self.L.ref_add_circular_wave(factor, x, y, _p(a, _D), _p(b, _D), _p(xy, _S))
        return a, b

    def mask(self, domain, circles=None):
        """xy_in_billiard_single_domain over the build's grid (its LAMBDA, MU,
        XMIN..YMAX), circles from circle_list for D_CIRCLES"""
        c = np.zeros((4, 1)) if circles is None else np.ascontiguousarray(circles, np.float64)
        n = 0 if circles is None else c.shape[1]
        a = np.ascontiguousarray(c[3], np.int32)
        out = np.zeros((self.nx, self.ny), np.int16)
        cc = [np.ascontiguousarray(c[k]) for k in range(3)]
        rc = self.L.ref_mask(domain, n, _p(cc[0], _D), _p(cc[1], _D), _p(cc[2], _D), _p(a, _I), _p(out, _S))
        assert rc == 0
        return out

    def set_michelson(self, position):
        """the build's michelson_position global"""
        self.L.ref_set_michelson(float(position))

    def set_lists(self, rects=None, rectrot=None, polygons=None, polyline=None):
        """fill the build's polyrect, polyrectrot, polygons and polyline
        globals from (rows, n) arrays (rows as LIST_ROWS)"""
        args, keep = [], []
        for name, a in (("rects", rects), ("rectrot", rectrot), ("polygons", polygons), ("polyline", polyline)):
            n = 0 if a is None else np.asarray(a).shape[1]
            a = np.zeros((LIST_ROWS[name], 1)) if a is None else np.ascontiguousarray(a, np.float64)
            keep.append(a)
            args += [n, _p(a, _D)]
        assert self.L.ref_set_lists(*args) == 0

    def mask_billiard(self, circles=None, circles_b=None):
        """xy_in_billiard over the build's grid (with COMPARISON: B_DOMAIN and
        circles where y > 0, B_DOMAIN_B and circles_b where y <= 0)"""
        args = []
        for c in (circles, circles_b):
            c = np.zeros((4, 1)) if c is None else np.ascontiguousarray(c, np.float64)
            n = 0 if c is None else c.shape[1]
            args += [n] + [np.ascontiguousarray(c[k]) for k in range(3)] + [np.ascontiguousarray(c[3], np.int32)]
        out = np.zeros((self.nx, self.ny), np.int16)
        rc = self.L.ref_mask_billiard(args[0], _p(args[1], _D), _p(args[2], _D), _p(args[3], _D), _p(args[4], _I),
                                      args[5], _p(args[6], _D), _p(args[7], _D), _p(args[8], _D), _p(args[9], _I),
                                      _p(out, _S))
        assert rc == 0
        return out

    def wells(self, seed):
        """the well centres, heights and sigma init_ior_2d uses"""
        n = self.L.ref_nwells()
        xc, yc, h = np.zeros(n), np.zeros(n), np.zeros(n)
        sigma = self.L.ref_wells(seed, _p(xc, _D), _p(yc, _D), _p(h, _D))
        return xc, yc, h, sigma

    def palette_table(self, palette):
        out = np.zeros((511, 3), np.float32)
        n = self.L.ref_palette_table(palette, _p(out, _F))
        return out[:n] if n else None

    def laplacian_table(self):
        """the reference's init_laplacian_coords table as counts and flat indices"""
        nn = np.zeros((self.nx, self.ny), np.int32)
        nb = np.zeros((self.nx, self.ny, 4), np.int32)
        self.L.ref_laplacian_table(_p(nn, _I), _p(nb, _I))
        return nn, nb

    def lshape_lines(self):
        a = np.zeros(6, np.int32)
        self.L.ref_lshape_lines(_p(a, _I))
        return tuple(int(v) for v in a)

    def corner_drive(self, time):
        return self.L.ref_corner_drive(time)

    def half(self, phi_in, psi_in, phi_out, xy_in, tcc, tc, tgamma):
        if self.table:
            self.L.ref_evolve_wave_half_new(_p(phi_in, _D), _p(psi_in, _D), _p(phi_out, _D),
                                            _p(xy_in, _S), _p(tc, _D), _p(tcc, _D), _p(tgamma, _D))
        elif self.flat:
            self.L.ref_evolve_wave_half_flat(_p(phi_in, _D), _p(psi_in, _D), _p(phi_out, _D),
                                             _p(xy_in, _S), _p(tc, _D), _p(tcc, _D), _p(tgamma, _D))
        else:
            self.L.ref_evolve_wave_half(_p(phi_in, _D), _p(psi_in, _D), _p(phi_out, _D),
                                        _p(xy_in, _S), _p(tcc, _D), _p(tc, _D), _p(tgamma, _D))

    def evolve_wave(self, phi, psi, tmp, xy_in, tcc, tc, tgamma):
        self.L.ref_evolve_wave(_p(phi, _D), _p(psi, _D), _p(tmp, _D), _p(xy_in, _S), _p(tcc, _D),
                               _p(tc, _D), _p(tgamma, _D))

    def run_rows(self, bufs, nsteps, xy_in, tcc, tc, tgamma):
        """nsteps evolve_wave_half calls with the evolve_wave rotation on the
        three C-contiguous buffers of bufs (phi, psi, tmp), rotated in place"""
        for _ in range(nsteps):
            self.L.ref_evolve_wave_half(_p(bufs[0], _D), _p(bufs[1], _D), _p(bufs[2], _D), _p(xy_in, _S),
                                        _p(tcc, _D), _p(tc, _D), _p(tgamma, _D))
            bufs[:] = [bufs[2], bufs[0], bufs[1]]

    def run(self, phi, psi, tmp, nsteps, xy_in, tcc, tc, tgamma):
        """nsteps half-steps with the evolve_wave rotation; returns (phi, psi, tmp)."""
        b = [phi, psi, tmp]
        s = 0
        for _ in range(nsteps):
            self.half(b[s], b[(s + 1) % 3], b[(s + 2) % 3], xy_in, tcc, tc, tgamma)
            s = (s + 2) % 3
        return b[s], b[(s + 1) % 3], b[(s + 2) % 3]

    def energy(self, phi, psi, xy_in, i, j):
        return self.L.ref_compute_energy(_p(phi, _D), _p(psi, _D), _p(xy_in, _S), i, j)
