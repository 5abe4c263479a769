"""wavehip -- Python mirror of the C ABI in include/wavehip.h (libwavehip.so).

The stepper itself is the HIP library; this module only moves numpy arrays
across the ABI with ctypes.  There is no CPU fallback: if libwavehip.so is
missing, or no GPU is visible, calls raise.

Names follow the reference (nilsberglund-orleans/YouTube-simulations):
evolve_wave_half / evolve_wave (wave_billiard.c:302-631, wave_3d.c:600-1031),
phi = u^n, psi = u^{n-1}, tmp the third rotation buffer, xy_in the domain
mask, tcc / tc / tgamma the tables of init_ior_2d (sub_wave.c:12164).
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libwavehip.so")

WV_OK, WV_EINVAL, WV_EHIP, WV_ESTATE, WV_EUNSUPPORTED, WV_ENOMEM, WV_EIO = 0, -1, -2, -3, -4, -5, -6
BC_DIRICHLET, BC_PERIODIC, BC_ABSORBING, BC_VPER_HABS, BC_ABS_REFLECT = 0, 1, 2, 3, 4
BC_LSHAPE = 10
FORM_FP64, FORM_F32SPLIT = 0, 1
FLAG_TABLES_FIXED = 1
FLAG_FULL_FORMULA = 2
FLAG_COMPARISON = 4
D_SINAI, D_EXT_ELLIPSE, D_CIRCLES = 3, 199, 20
TIFF_NONE, TIFF_LZW, TIFF_MAX_WIDTH = 1, 5, 8192

# every symbol include/wavehip.h declares
EXPORTS = [
    "wv_params_default", "wv_create", "wv_destroy", "wv_last_error", "wv_set_tables",
    "wv_set_state", "wv_set_state3", "wv_get_state", "wv_get_state3", "wv_set_sources",
    "wv_set_neumann", "wv_set_drive", "wv_set_floor", "wv_set_laplacian", "wv_set_probes", "wv_get_probes",
    "wv_format_time_series", "wv_set_wells_tables", "wv_build_mask", "wv_build_mask_static", "wv_build_domain_mask",
    "wv_build_list_domain_mask", "wv_batch_build_list_domain_mask", "wv_build_comparison_mask", "wv_set_ior_tables", "wv_rotating_wells", "wv_init_circular_wave",
    "wv_add_circular_wave", "wv_batch_init_circular_wave", "wv_batch_add_circular_wave",
    "wv_batch_set_wells_tables", "wv_batch_set_ior_tables", "wv_batch_build_domain_mask", "wv_tiff_bound", "wv_tiff_encode", "wv_tiff_write",
    "wv_step", "wv_prepare", "wv_step_timed", "wv_sync", "wv_steps_done", "wv_last_kernel_ms", "wv_energy",
    "wv_get_tiling", "wv_get_tile_counts", "wv_frame_params_default", "wv_frame_set_tables", "wv_frame_reset", "wv_frame",
    "wv_frame_rgba8",
    "wv_set_tiling", "wv_autotune", "wv_evolve_wave_half_rows", "wv_evolve_wave_rows",
    "wv_evolve_wave_half_flat", "wv_evolve_wave_flat", "wv_batch_create", "wv_batch_destroy",
    "wv_batch_last_error", "wv_batch_set_tables", "wv_batch_set_state", "wv_batch_get_state",
    "wv_batch_step", "wv_batch_prepare", "wv_batch_step_timed", "wv_batch_last_kernel_ms", "wv_batch_energy",
    "wv_batch_set_tiling",
    "wv_batch_get_tiling", "wv_batch_autotune",
]


class WvParams(ctypes.Structure):
    """struct wv_params of include/wavehip.h"""
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("bcond", ctypes.c_int),
        ("twospeeds", ctypes.c_int), ("stencil", ctypes.c_int), ("form", ctypes.c_int),
        ("oscillate_topbot", ctypes.c_int), ("tb_shift", ctypes.c_int),
        ("kappa", ctypes.c_double), ("kappa_sides", ctypes.c_double),
        ("gamma_sides", ctypes.c_double), ("kappa_topbot", ctypes.c_double),
        ("gamma_topbot", ctypes.c_double), ("e_scale", ctypes.c_double),
        ("tb_depth", ctypes.c_int), ("rows_per_tile", ctypes.c_int), ("device", ctypes.c_int),
        ("flags", ctypes.c_int),
    ]


class WvFrameParams(ctypes.Structure):
    """wv_frame_params of include/wavehip.h"""
    _fields_ = [(n, ctypes.c_int) for n in ("plot", "scheme", "palette", "asym", "size", "time", "shade",
                                            "rescale_center", "fade_in_obstacle", "fade")] + \
               [("nsteps", ctypes.c_long)] + \
               [(n, ctypes.c_double) for n in ("scale", "fade_value", "slope", "color_range", "attenuation",
                                               "vshift_amplitude", "vscale_amplitude", "log_shift", "log_scale",
                                               "flux_scale", "avrg_e_factor", "colorhue", "colordrift",
                                               "lummean", "lumamp", "huemean", "hueamp", "shade_scale",
                                               "light_x", "light_y", "courant", "courantb")]


class WaveError(RuntimeError):
    pass


_D = ctypes.POINTER(ctypes.c_double)
_S = ctypes.POINTER(ctypes.c_short)
_I = ctypes.POINTER(ctypes.c_int)
_DD = ctypes.POINTER(_D)
_SS = ctypes.POINTER(_S)
_lib = None


class WvDomainParams(ctypes.Structure):
    """struct wv_domain_params of include/wavehip.h"""
    _fields_ = [(n, ctypes.c_double) for n in ("lam", "mu", "mu_b", "apoly", "wall_width", "wall_width_b",
                                                "julia_scale", "julia_x", "julia_y", "mandellimit")] + \
               [(n, ctypes.c_int) for n in ("npoly", "mdepth", "mratio", "mandellevel", "ngridx", "ngridy")] + \
               [(n, ctypes.c_double) for n in ("wall_width_asym", "wall_width_asym_b", "oscil_ymid",
                                                "michelson_position")]


class WvDomainLists(ctypes.Structure):
    """struct wv_domain_lists of include/wavehip.h"""
    _fields_ = [("ncircles", ctypes.c_int), ("circles", _D), ("nrects", ctypes.c_int), ("rects", _D),
                ("nrectrot", ctypes.c_int), ("rectrot", _D), ("npolygons", ctypes.c_int), ("polygons", _D),
                ("npolyline", ctypes.c_int), ("polyline", _D)]


# rows of each list of wv_domain_lists
LIST_ROWS = dict(circles=4, rects=4, rectrot=5, polygons=6, polyline=2)


def domain_lists(**lists):
    """WvDomainLists from (rows, n) arrays (rows as LIST_ROWS: circles xc, yc,
    radius, active; rects x1, y1, x2, y2; rectrot x1, y1, x2, y2, width;
    polygons xc, yc, radius, angle, nsides, active; polyline x, y); the
    arrays stay alive on the struct"""
    bad = set(lists) - set(LIST_ROWS)
    if bad:
        raise ValueError("unknown lists %s" % sorted(bad))
    s = WvDomainLists()
    keep = []
    for name, rows in LIST_ROWS.items():
        a = lists.get(name)
        a = np.zeros((rows, 0)) if a is None else np.ascontiguousarray(a, np.float64)
        if a.ndim != 2 or a.shape[0] != rows:
            raise ValueError("%s must be (%d, n)" % (name, rows))
        keep.append(a)
        setattr(s, "n" + name, a.shape[1])
        setattr(s, name, _d(a) if a.shape[1] else None)
    s._keep = keep
    return s


class WvIorParams(ctypes.Structure):
    """struct wv_ior_params of include/wavehip.h"""
    _fields_ = [(n, ctypes.c_double) for n in ("courant", "courantb", "gamma", "gammab", "lam", "mu",
                                                "mandel_ior_scale", "mandellimit")] + \
               [(n, ctypes.c_int) for n in ("mandellevel", "mdepth")]


# the domain macros of the checked-in wave_billiard.c (73-131) and julia_x,
# julia_y of global_pdes.c:567
DOMAIN_DEFAULTS = dict(lam=0.2, mu=0.008, mu_b=0.08, apoly=1.0, wall_width=0.075, wall_width_b=0.1,
                       julia_scale=0.8, julia_x=0.37468, julia_y=0.21115, mandellimit=10.0, npoly=9, mdepth=6,
                       mratio=3, mandellevel=1000, ngridx=10, ngridy=10, wall_width_asym=0.75,
                       wall_width_asym_b=0.75, oscil_ymid=0.0, michelson_position=0.0)
# init_ior_2d's macros of the checked-in wave_billiard.c (54, 92-100, 140-143)
IOR_DEFAULTS = dict(courant=0.09, courantb=0.25, gamma=0.0, gammab=0.0, lam=0.2, mu=0.008, mandel_ior_scale=-0.05,
                    mandellimit=10.0, mandellevel=1000, mdepth=6)


def _struct(cls, defaults, kw):
    d = dict(defaults)
    bad = set(kw) - set(d)
    if bad:
        raise KeyError("unknown parameters: %s" % sorted(bad))
    d.update(kw)
    return cls(**d)


def domain_params(**kw):
    """wv_domain_params from DOMAIN_DEFAULTS updated with kw"""
    return _struct(WvDomainParams, DOMAIN_DEFAULTS, kw)


def ior_params(**kw):
    """wv_ior_params from IOR_DEFAULTS updated with kw"""
    return _struct(WvIorParams, IOR_DEFAULTS, kw)


def rotating_wells(ngridx, ngridy, large, angle, box):
    """the rotating-well centres at angle (wv_rotating_wells): (xc, yc, sigma)"""
    bx = np.ascontiguousarray(box, np.float64).ravel()
    if bx.size != 4:
        raise ValueError("box must hold XMIN, XMAX, YMIN, YMAX")
    xc, yc = np.empty(ngridx * ngridy), np.empty(ngridx * ngridy)
    s = ctypes.c_double(0)
    rc = load().wv_rotating_wells(ngridx, ngridy, int(large), angle, _d(bx), _d(xc), _d(yc), ctypes.byref(s))
    if rc:
        raise WaveError("wv_rotating_wells: %d" % rc)
    return xc, yc, s.value


def format_time_series(values):
    """the SAVE_TIME_SERIES text of one probe's samples (wave_billiard.c:902-905)"""
    v = np.ascontiguousarray(values, dtype=np.float64).ravel()
    cap = 21 * max(len(v), 1)
    buf = ctypes.create_string_buffer(cap)
    n = ctypes.c_long(0)
    L = load()
    rc = L.wv_format_time_series(_d(v), len(v), 1, buf, cap, ctypes.byref(n))
    if rc:
        raise WaveError("wv_format_time_series: %d" % rc)
    return buf.raw[:n.value].decode()


def load():
    """Load libwavehip.so (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise WaveError("libwavehip.so is not built at %s (run __graft_entry__.build())" % LIB_PATH)
    L = ctypes.CDLL(LIB_PATH)
    P = ctypes.POINTER(WvParams)
    vp = ctypes.c_void_p
    sig = {
        "wv_params_default": (None, [P, ctypes.c_int, ctypes.c_int]),
        "wv_create": (ctypes.c_int, [ctypes.POINTER(vp), P]),
        "wv_destroy": (None, [vp]),
        "wv_last_error": (ctypes.c_char_p, [vp]),
        "wv_set_tables": (ctypes.c_int, [vp, _S, _D, _D, _D]),
        "wv_set_state": (ctypes.c_int, [vp, _D, _D]),
        "wv_set_state3": (ctypes.c_int, [vp, _D, _D, _D]),
        "wv_get_state": (ctypes.c_int, [vp, _D, _D]),
        "wv_get_state3": (ctypes.c_int, [vp, _D, _D, _D]),
        "wv_set_sources": (ctypes.c_int, [vp, ctypes.c_int, _I, _I, _D, ctypes.c_long]),
        "wv_set_neumann": (ctypes.c_int, [vp, ctypes.c_int, _I, _I, _I, _I, _I, _I, _D, _D]),
        "wv_set_drive": (ctypes.c_int, [vp, ctypes.c_int, _I, _I, _D, ctypes.c_long]),
        "wv_set_floor": (ctypes.c_int, [vp, ctypes.c_double]),
        "wv_set_laplacian": (ctypes.c_int, [vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]),
        "wv_set_probes": (ctypes.c_int, [vp, ctypes.c_int, _I, _I, ctypes.c_long, ctypes.c_long]),
        "wv_set_wells_tables": (ctypes.c_int, [vp, _S, ctypes.c_int, _D, _D, _D, ctypes.c_double, ctypes.c_int,
                                               ctypes.c_double, ctypes.c_double, ctypes.c_double, _D, _D, _D, _D]),
        "wv_build_mask": (ctypes.c_int, [vp, ctypes.c_int, _D, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                         _D, _S]),
        "wv_build_mask_static": (ctypes.c_int, [vp, ctypes.c_int, _D, ctypes.c_double, ctypes.c_double,
                                                ctypes.c_int, _D, _S]),
        "wv_build_domain_mask": (ctypes.c_int, [vp, ctypes.c_int, _D, ctypes.POINTER(WvDomainParams), ctypes.c_int,
                                                _D, _S]),
        "wv_build_list_domain_mask": (ctypes.c_int, [vp, ctypes.c_int, _D, ctypes.POINTER(WvDomainParams),
                                                     ctypes.POINTER(WvDomainLists), _S]),
        "wv_batch_build_list_domain_mask": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int, _D,
                                                           ctypes.POINTER(WvDomainParams),
                                                           ctypes.POINTER(WvDomainLists), _S]),
        "wv_build_comparison_mask": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int, _D, ctypes.POINTER(WvDomainParams),
                                                    ctypes.c_int, _D, ctypes.c_int, _D, _S]),
        "wv_set_ior_tables": (ctypes.c_int, [vp, _S, ctypes.c_int, _D, ctypes.POINTER(WvIorParams), _D, _D, _D]),
        "wv_rotating_wells": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, _D, _D, _D,
                                             _D]),
        "wv_init_circular_wave": (ctypes.c_int, [vp] + [ctypes.c_double] * 5 + [_D]),
        "wv_add_circular_wave": (ctypes.c_int, [vp] + [ctypes.c_double] * 6 + [_D]),
        "wv_batch_init_circular_wave": (ctypes.c_int, [vp, ctypes.c_int] + [ctypes.c_double] * 5 + [_D]),
        "wv_batch_set_wells_tables": (ctypes.c_int, [vp, ctypes.c_int, _S, ctypes.c_int, _D, _D, _D, ctypes.c_double,
                                                     ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                                     _D, _D, _D, _D]),
        "wv_batch_set_ior_tables": (ctypes.c_int, [vp, ctypes.c_int, _S, ctypes.c_int, _D,
                                                   ctypes.POINTER(WvIorParams), _D, _D, _D]),
        "wv_batch_build_domain_mask": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int, _D,
                                                      ctypes.POINTER(WvDomainParams), ctypes.c_int, _D, _S]),
        "wv_batch_add_circular_wave": (ctypes.c_int, [vp, ctypes.c_int] + [ctypes.c_double] * 6 + [_D]),
        "wv_tiff_bound": (ctypes.c_size_t, [ctypes.c_int, ctypes.c_int, ctypes.c_char_p]),
        "wv_tiff_encode": (ctypes.c_int, [vp, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_char_p,
                                          ctypes.c_int, ctypes.c_void_p, ctypes.c_size_t,
                                          ctypes.POINTER(ctypes.c_size_t)]),
        "wv_tiff_write": (ctypes.c_int, [vp, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_void_p, ctypes.c_int,
                                         ctypes.c_int, ctypes.c_int]),
        "wv_get_probes": (ctypes.c_int, [vp, _D, ctypes.c_long, ctypes.POINTER(ctypes.c_long),
                                         ctypes.POINTER(ctypes.c_long)]),
        "wv_step": (ctypes.c_int, [vp, ctypes.c_long]),
        "wv_prepare": (ctypes.c_int, [vp, ctypes.c_long]),
        "wv_step_timed": (ctypes.c_int, [vp, ctypes.c_long, _D, ctypes.POINTER(ctypes.c_long)]),
        "wv_sync": (ctypes.c_int, [vp]),
        "wv_steps_done": (ctypes.c_long, [vp]),
        "wv_last_kernel_ms": (ctypes.c_int, [vp, _D, _D]),
        "wv_batch_last_kernel_ms": (ctypes.c_int, [vp, _D, _D]),
        "wv_energy": (ctypes.c_int, [vp, _D]),
        "wv_frame_params_default": (None, [ctypes.POINTER(WvFrameParams)]),
        "wv_frame_set_tables": (ctypes.c_int, [vp, ctypes.POINTER(ctypes.c_float), ctypes.c_int, _D, _D]),
        "wv_frame_reset": (ctypes.c_int, [vp]),
        "wv_frame": (ctypes.c_int, [vp, ctypes.POINTER(WvFrameParams), ctypes.POINTER(ctypes.c_float)]),
        "wv_frame_rgba8": (ctypes.c_int, [vp, ctypes.POINTER(WvFrameParams), ctypes.POINTER(ctypes.c_ubyte)]),
        "wv_get_tiling": (ctypes.c_int, [vp, _I, _I, _I]),
        "wv_format_time_series": (ctypes.c_int, [_D, ctypes.c_long, ctypes.c_long, ctypes.c_char_p, ctypes.c_long,
                                                 ctypes.POINTER(ctypes.c_long)]),
        "wv_get_tile_counts": (ctypes.c_int, [vp, ctypes.c_int, _I, _I, _I, ctypes.POINTER(ctypes.c_long)]),
        "wv_set_tiling": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int]),
        "wv_autotune": (ctypes.c_int, [vp, ctypes.c_long, _D]),
        "wv_evolve_wave_half_rows": (ctypes.c_int, [vp, _DD, _DD, _DD, _SS, _DD, _DD, _DD]),
        "wv_evolve_wave_rows": (ctypes.c_int, [vp, _DD, _DD, _DD, _SS, _DD, _DD, _DD]),
        "wv_evolve_wave_half_flat": (ctypes.c_int, [vp, _D, _D, _D, _S, _D, _D, _D]),
        "wv_evolve_wave_flat": (ctypes.c_int, [vp, _D, _D, _D, _S, _D, _D, _D]),
        "wv_batch_create": (ctypes.c_int, [ctypes.POINTER(vp), ctypes.c_int, P]),
        "wv_batch_destroy": (None, [vp]),
        "wv_batch_last_error": (ctypes.c_char_p, [vp]),
        "wv_batch_set_tables": (ctypes.c_int, [vp, ctypes.c_int, _S, _D, _D, _D]),
        "wv_batch_set_state": (ctypes.c_int, [vp, ctypes.c_int, _D, _D]),
        "wv_batch_get_state": (ctypes.c_int, [vp, ctypes.c_int, _D, _D]),
        "wv_batch_step": (ctypes.c_int, [vp, ctypes.c_long]),
        "wv_batch_prepare": (ctypes.c_int, [vp, ctypes.c_long]),
        "wv_batch_step_timed": (ctypes.c_int, [vp, ctypes.c_long, _D, ctypes.POINTER(ctypes.c_long)]),
        "wv_batch_energy": (ctypes.c_int, [vp, _D]),
        "wv_batch_set_tiling": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int]),
        "wv_batch_get_tiling": (ctypes.c_int, [vp, _I, _I, _I]),
        "wv_batch_autotune": (ctypes.c_int, [vp, ctypes.c_long, _D]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def make_params(p):
    """wv_params from a dict (the 'params' of a workload) with defaults."""
    s = WvParams()
    load().wv_params_default(ctypes.byref(s), int(p["nx"]), int(p["ny"]))
    for k, v in p.items():
        if hasattr(s, k) and not k.startswith("_"):
            setattr(s, k, type(getattr(s, k))(v))
    return s


def _d(a):
    return a.ctypes.data_as(_D)


def _f64(a, shape):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if a.shape != shape:
        raise ValueError("expected shape %s, got %s" % (shape, a.shape))
    return a


def _i16(a, shape):
    a = np.ascontiguousarray(a, dtype=np.int16)
    if a.shape != shape:
        raise ValueError("expected shape %s, got %s" % (shape, a.shape))
    return a


def _out(a, shape, dtype):
    """a caller's output array the library writes in place: exact shape,
    dtype and C-contiguity, never a converted copy"""
    if not isinstance(a, np.ndarray) or a.shape != tuple(shape) or a.dtype != dtype or not a.flags.c_contiguous \
            or not a.flags.writeable:
        raise ValueError("output must be a writeable C-contiguous %s array of shape %s" % (np.dtype(dtype), shape))
    return a


def _inplace(a, shape, dtype, what):
    """an array the library reads and writes in place (drop-in forms)"""
    if not isinstance(a, np.ndarray) or a.dtype != dtype or not a.flags.c_contiguous:
        raise ValueError("%s must be a C-contiguous %s array" % (what, np.dtype(dtype)))
    if a.size != shape[0] * shape[1] or (a.ndim == 2 and a.shape != tuple(shape)):
        raise ValueError("%s must hold %d x %d values, got shape %s" % (what, shape[0], shape[1], a.shape))
    return a


class Wave:
    """One grid on the GPU (a wv_ctx)."""

    def __init__(self, params, **overrides):
        p = dict(params)
        p.update(overrides)
        self.p = make_params(p)
        self.shape = (self.p.nx, self.p.ny)
        self._L = load()
        self._h = ctypes.c_void_p()
        rc = self._L.wv_create(ctypes.byref(self._h), ctypes.byref(self.p))
        if rc:
            msg = self._L.wv_last_error(self._h).decode()
            self._L.wv_destroy(self._h)
            self._h = None
            raise WaveError("wv_create: %d %s" % (rc, msg))
        self._keep = []

    def close(self):
        if self._h:
            self._L.wv_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc, what):
        if rc:
            raise WaveError("%s: %d %s" % (what, rc, self._L.wv_last_error(self._h).decode()))

    def set_tables(self, xy_in, tcc, tc, tgamma):
        xy = _i16(xy_in, self.shape)
        a, b, c = (_f64(t, self.shape) for t in (tcc, tc, tgamma))
        self._chk(self._L.wv_set_tables(self._h, xy.ctypes.data_as(_S), _d(a), _d(b), _d(c)),
                  "wv_set_tables")

    def set_state(self, phi, psi, tmp=None):
        a, b = _f64(phi, self.shape), _f64(psi, self.shape)
        if tmp is None:
            self._chk(self._L.wv_set_state(self._h, _d(a), _d(b)), "wv_set_state")
        else:
            t = _f64(tmp, self.shape)
            self._chk(self._L.wv_set_state3(self._h, _d(a), _d(b), _d(t)), "wv_set_state3")

    def get_state(self, three=False, out=None):
        """(phi, psi) [, tmp] as new arrays, or into the C-contiguous float64
        arrays of out"""
        if out is None:
            a, b = np.empty(self.shape), np.empty(self.shape)
        else:
            a, b = (_out(x, self.shape, np.float64) for x in out[:2])
        if three:
            t = np.empty(self.shape) if out is None or len(out) < 3 else _out(out[2], self.shape, np.float64)
            self._chk(self._L.wv_get_state3(self._h, _d(a), _d(b), _d(t)), "wv_get_state3")
            return a, b, t
        self._chk(self._L.wv_get_state(self._h, _d(a), _d(b)), "wv_get_state")
        return a, b

    def set_sources(self, sources):
        if not sources:
            self._chk(self._L.wv_set_sources(self._h, 0, None, None, None, 0), "wv_set_sources")
            return
        si = np.ascontiguousarray(sources["i"], np.int32).ravel()
        sj = np.ascontiguousarray(sources["j"], np.int32).ravel()
        amp = np.ascontiguousarray(sources["amp"], np.float64)
        if sj.size != si.size or amp.ndim != 2 or amp.shape[0] != si.size or amp.shape[1] < 1:
            raise ValueError("sources: i and j of one length n, amp of shape (n, steps)")
        self._chk(self._L.wv_set_sources(self._h, len(si), si.ctypes.data_as(_I), sj.ctypes.data_as(_I),
                                         _d(amp), amp.shape[1]), "wv_set_sources")

    def set_neumann(self, nb):
        """Neumann ghost points: dict of int arrays i, j, nbi1, nbj1, nbi2, nbj2 and
        float arrays nx, ny (or None to remove them)"""
        if not nb:
            self._chk(self._L.wv_set_neumann(self._h, 0, None, None, None, None, None, None, None, None),
                      "wv_set_neumann")
            return
        ia = [np.ascontiguousarray(nb[k], np.int32).ravel() for k in ("i", "j", "nbi1", "nbj1", "nbi2", "nbj2")]
        fa = [np.ascontiguousarray(nb[k], np.float64).ravel() for k in ("nx", "ny")]
        if any(a.size != ia[0].size for a in ia + fa):
            raise ValueError("Neumann points: every array needs one entry per point")
        self._chk(self._L.wv_set_neumann(self._h, len(ia[0]), *[a.ctypes.data_as(_I) for a in ia],
                                         *[_d(a) for a in fa]), "wv_set_neumann")

    def set_drive(self, drive):
        """driven cells: dict i, j, val (n x len), or None"""
        if not drive:
            self._chk(self._L.wv_set_drive(self._h, 0, None, None, None, 0), "wv_set_drive")
            return
        i = np.ascontiguousarray(drive["i"], np.int32).ravel()
        j = np.ascontiguousarray(drive["j"], np.int32).ravel()
        val = np.ascontiguousarray(drive["val"], np.float64)
        if j.size != i.size or val.ndim != 2 or val.shape[0] != i.size or val.shape[1] < 1:
            raise ValueError("driven cells: i and j of one length n, val of shape (n, steps)")
        self._chk(self._L.wv_set_drive(self._h, len(i), i.ctypes.data_as(_I), j.ctypes.data_as(_I), _d(val),
                                       val.shape[1]), "wv_set_drive")

    def set_floor(self, vmax):
        self._chk(self._L.wv_set_floor(self._h, float(vmax)), "wv_set_floor")

    def frame_params(self, **kw):
        """wv_frame_params with the reference defaults, fields overridden by kw"""
        f = WvFrameParams()
        self._L.wv_frame_params_default(ctypes.byref(f))
        for k, v in kw.items():
            setattr(f, k, v)
        return f

    def set_frame_tables(self, lut=None, color_scale=None, fade_table=None):
        F = ctypes.POINTER(ctypes.c_float)
        lt = None if lut is None else np.ascontiguousarray(lut, np.float32)
        if lt is not None and (lt.ndim != 2 or lt.shape[1] != 3):
            raise ValueError("palette table must be (n, 3)")
        cs = None if color_scale is None else _f64(color_scale, self.shape)
        ft = None if fade_table is None else _f64(fade_table, self.shape)
        self._chk(self._L.wv_frame_set_tables(self._h, None if lt is None else lt.ctypes.data_as(F),
                                              0 if lt is None else lt.shape[0],
                                              None if cs is None else cs.ctypes.data_as(_D),
                                              None if ft is None else ft.ctypes.data_as(_D)), "wv_frame_set_tables")

    def frame_reset(self):
        self._chk(self._L.wv_frame_reset(self._h), "wv_frame_reset")

    def frame(self, f, u8=False, out=None):
        """one frame: rgba[nbx, nby, 4] float32 (alpha 1 where drawn), or
        uint8 (alpha 255 where drawn) with u8; into out if given"""
        z = max(f.size, 1)  # the library rejects size < 1
        nbx = -(-self.p.nx // z)
        nby = -(-self.p.ny // z)
        if out is not None:
            _out(out, (nbx, nby, 4), np.uint8 if u8 else np.float32)
        if u8:
            out = np.empty((nbx, nby, 4), np.uint8) if out is None else out
            rc = self._L.wv_frame_rgba8(self._h, ctypes.byref(f), out.ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte)))
            self._chk(rc, "wv_frame_rgba8")
            return out
        out = np.empty((nbx, nby, 4), np.float32) if out is None else out
        self._chk(self._L.wv_frame(self._h, ctypes.byref(f), out.ctypes.data_as(ctypes.POINTER(ctypes.c_float))),
                  "wv_frame")
        return out

    def set_wells_tables(self, xy_in, xc, yc, height, sigma, use_tanh, courant, courantb, gamma, box, copy=True):
        """Gaussian-well media built on the GPU; returns (tcc, tc, tgamma), or
        None without copy"""
        xy = _i16(xy_in, self.shape)
        xc = np.ascontiguousarray(xc, np.float64).ravel()
        yc = np.ascontiguousarray(yc, np.float64).ravel()
        hh = None if height is None else np.ascontiguousarray(height, np.float64).ravel()
        bx = np.ascontiguousarray(box, np.float64).ravel()
        if yc.size != xc.size or (hh is not None and hh.size != xc.size) or bx.size != 4:
            raise ValueError("wells: xc, yc (and height) of one length, box of 4 values")
        t = [np.empty(self.shape) for _ in range(3)] if copy else None
        outs = [_d(a) for a in t] if copy else [None, None, None]
        self._chk(self._L.wv_set_wells_tables(self._h, xy.ctypes.data_as(_S), len(xc), _d(xc), _d(yc),
                                              None if hh is None else _d(hh), sigma, int(use_tanh), courant,
                                              courantb, gamma, _d(bx), *outs), "wv_set_wells_tables")
        return tuple(t) if copy else None

    def build_mask(self, domain, box, lam, mu=0.0, circles=None, static=False):
        """init_wave_flat's xy_in on the GPU (D_SINAI 3, D_EXT_ELLIPSE 199,
        D_CIRCLES 20 with circles = (4, n): xc, yc, radius, active); returns
        int16 (nx, ny).  static: wv_build_mask_static (no device allocation;
        the library's fixed scratch array)"""
        c = np.zeros((4, 1)) if circles is None else np.ascontiguousarray(circles, np.float64)
        if c.ndim != 2 or c.shape[0] != 4:
            raise ValueError("circles must be (4, n): xc, yc, radius, active")
        n = 0 if circles is None else c.shape[1]
        bx = np.ascontiguousarray(box, np.float64).ravel()
        if bx.size != 4:
            raise ValueError("box must hold XMIN, XMAX, YMIN, YMAX")
        out = np.empty(self.shape, np.int16)
        f = self._L.wv_build_mask_static if static else self._L.wv_build_mask
        self._chk(f(self._h, int(domain), _d(bx), lam, mu, n, _d(c), out.ctypes.data_as(_S)),
                  "wv_build_mask_static" if static else "wv_build_mask")
        return out

    def build_domain_mask(self, domain, box, params=None, circles=None, **kw):
        """xy_in_billiard_single_domain over the grid on the GPU for any
        closed-form B_DOMAIN code (wv_build_domain_mask); params: a
        WvDomainParams, or keywords over DOMAIN_DEFAULTS; circles (4, n) for
        the D_CIRCLES family.  Returns int16 (nx, ny)."""
        P = params if params is not None else domain_params(**kw)
        c = np.zeros((4, 1)) if circles is None else np.ascontiguousarray(circles, np.float64)
        if c.ndim != 2 or c.shape[0] != 4:
            raise ValueError("circles must be (4, n): xc, yc, radius, active")
        n = 0 if circles is None else c.shape[1]
        bx = np.ascontiguousarray(box, np.float64).ravel()
        if bx.size != 4:
            raise ValueError("box must hold XMIN, XMAX, YMIN, YMAX")
        out = np.empty(self.shape, np.int16)
        self._chk(self._L.wv_build_domain_mask(self._h, int(domain), _d(bx), ctypes.byref(P), n, _d(c),
                                               out.ctypes.data_as(_S)), "wv_build_domain_mask")
        return out

    def build_list_domain_mask(self, domain, box, params=None, lists=None, **kw):
        """xy_in_billiard_single_domain over the grid on the GPU for the
        domains that read polyrect, polyrectrot, polygons, polyline or the
        circle list, and every closed-form code (wv_build_list_domain_mask);
        lists: a dict of (rows, n) arrays as domain_lists takes them.
        Returns int16 (nx, ny)."""
        P = params if params is not None else domain_params(**kw)
        dl = domain_lists(**(lists or {}))
        bx = np.ascontiguousarray(box, np.float64).ravel()
        if bx.size != 4:
            raise ValueError("box must hold XMIN, XMAX, YMIN, YMAX")
        out = np.empty(self.shape, np.int16)
        self._chk(self._L.wv_build_list_domain_mask(self._h, int(domain), _d(bx), ctypes.byref(P), ctypes.byref(dl),
                                                    out.ctypes.data_as(_S)), "wv_build_list_domain_mask")
        return out

    def build_comparison_mask(self, domain, domain_b, box, params=None, circles=None, circles_b=None, **kw):
        """xy_in_billiard with COMPARISON on the GPU (wv_build_comparison_mask):
        domain and circles where y > 0, domain_b and circles_b where y <= 0"""
        P = params if params is not None else domain_params(**kw)
        cs = []
        for c in (circles, circles_b):
            a = np.zeros((4, 1)) if c is None else np.ascontiguousarray(c, np.float64)
            if a.ndim != 2 or a.shape[0] != 4:
                raise ValueError("circles must be (4, n): xc, yc, radius, active")
            cs.append((0 if c is None else a.shape[1], a))
        bx = np.ascontiguousarray(box, np.float64).ravel()
        if bx.size != 4:
            raise ValueError("box must hold XMIN, XMAX, YMIN, YMAX")
        out = np.empty(self.shape, np.int16)
        self._chk(self._L.wv_build_comparison_mask(self._h, int(domain), int(domain_b), _d(bx), ctypes.byref(P),
                                                   cs[0][0], _d(cs[0][1]), cs[1][0], _d(cs[1][1]),
                                                   out.ctypes.data_as(_S)), "wv_build_comparison_mask")
        return out

    def set_ior_tables(self, xy_in, ior, box, params=None, copy=True, **kw):
        """init_ior_2d's per-cell media model ior built on the GPU and
        installed (wv_set_ior_tables); params: a WvIorParams, or keywords over
        IOR_DEFAULTS.  Returns (tcc, tc, tgamma), or None without copy."""
        P = params if params is not None else ior_params(**kw)
        xy = _i16(xy_in, self.shape)
        bx = np.ascontiguousarray(box, np.float64).ravel()
        if bx.size != 4:
            raise ValueError("box must hold XMIN, XMAX, YMIN, YMAX")
        t = [np.empty(self.shape) for _ in range(3)] if copy else None
        outs = [_d(a) for a in t] if copy else [None, None, None]
        self._chk(self._L.wv_set_ior_tables(self._h, xy.ctypes.data_as(_S), int(ior), _d(bx), ctypes.byref(P),
                                            *outs), "wv_set_ior_tables")
        return tuple(t) if copy else None

    def init_circular_wave(self, x0, y0, amp, variance, wavelength, box):
        """init_circular_wave (wave_common.c:89) on the GPU state"""
        bx = np.ascontiguousarray(box, np.float64).ravel()
        if bx.size != 4:
            raise ValueError("box must hold XMIN, XMAX, YMIN, YMAX")
        self._chk(self._L.wv_init_circular_wave(self._h, x0, y0, amp, variance, wavelength, _d(bx)),
                  "wv_init_circular_wave")

    def add_circular_wave(self, factor, x0, y0, amp, variance, wavelength, box):
        """add_circular_wave (wave_common.c:337) on the GPU state"""
        bx = np.ascontiguousarray(box, np.float64).ravel()
        if bx.size != 4:
            raise ValueError("box must hold XMIN, XMAX, YMIN, YMAX")
        self._chk(self._L.wv_add_circular_wave(self._h, factor, x0, y0, amp, variance, wavelength, _d(bx)),
                  "wv_add_circular_wave")

    def tiff(self, rgb, description="Wave equation in a planar domain", compression=TIFF_LZW):
        """the TIFF file (bytes) writetiff writes for the glReadPixels rows
        rgb, uint8 (height, width, 3) with the bottom row first"""
        rgb = np.ascontiguousarray(rgb, np.uint8)
        if rgb.ndim != 3 or rgb.shape[2] != 3:
            raise ValueError("rgb must be (height, width, 3)")
        h, w = rgb.shape[0], rgb.shape[1]
        d = description.encode() if isinstance(description, str) else bytes(description)
        cap = self._L.wv_tiff_bound(w, h, d)
        if cap == 0:
            raise ValueError("frame %dx%d out of range (width 1..%d, height 1..65535)" % (w, h, TIFF_MAX_WIDTH))
        out = np.empty(cap, np.uint8)
        n = ctypes.c_size_t(0)
        self._chk(self._L.wv_tiff_encode(self._h, rgb.ctypes.data, w, h, d, int(compression), out.ctypes.data, cap,
                                         ctypes.byref(n)), "wv_tiff_encode")
        return out[:n.value].tobytes()

    def write_tiff(self, filename, rgb, description="Wave equation in a planar domain", compression=TIFF_LZW):
        """writetiff(filename, description, 0, 0, width, height, compression)
        for the glReadPixels rows rgb (bottom row first)"""
        rgb = np.ascontiguousarray(rgb, np.uint8)
        if rgb.ndim != 3 or rgb.shape[2] != 3:
            raise ValueError("rgb must be (height, width, 3)")
        d = description.encode() if isinstance(description, str) else bytes(description)
        self._chk(self._L.wv_tiff_write(self._h, str(filename).encode(), d, rgb.ctypes.data, rgb.shape[1],
                                        rgb.shape[0], int(compression)), "wv_tiff_write")

    def set_probes(self, i, j, every=3, capacity=4096):
        """record u at cells (i[k], j[k]) after every `every`-th step"""
        ii = np.ascontiguousarray(i, np.int32).ravel()
        jj = np.ascontiguousarray(j, np.int32).ravel()
        if jj.size != ii.size:
            raise ValueError("probes: i and j of one length")
        self._keep_probe_n = len(ii)
        self._chk(self._L.wv_set_probes(self._h, len(ii), ii.ctypes.data_as(_I), jj.ctypes.data_as(_I), every,
                                        capacity), "wv_set_probes")

    def get_probes(self, max_samples=1 << 20):
        """(samples[count, n], lost) recorded since the probes were set or last read"""
        n = getattr(self, "_keep_probe_n", 0)
        out = np.empty(max(max_samples * n, 1))
        cnt, lost = ctypes.c_long(0), ctypes.c_long(0)
        self._chk(self._L.wv_get_probes(self._h, _d(out), max_samples, ctypes.byref(cnt), ctypes.byref(lost)),
                  "wv_get_probes")
        return out[:cnt.value * n].reshape(cnt.value, n), lost.value

    def set_laplacian(self, table):
        """neighbour-table stepping: table = (nneighb[NX,NY], nghb[NX,NY,4]) of
        flat indices i*NY + j, or None for the plain stepper"""
        if table is None:
            self._chk(self._L.wv_set_laplacian(self._h, None, None), "wv_set_laplacian")
            return
        I = ctypes.POINTER(ctypes.c_int)
        nn = np.ascontiguousarray(table[0], np.int32)
        nb = np.ascontiguousarray(table[1], np.int32)
        if nn.size != self.shape[0] * self.shape[1] or nb.size != 4 * nn.size:
            raise ValueError("neighbour table: nneighb (NX, NY) and nghb (NX, NY, 4)")
        self._chk(self._L.wv_set_laplacian(self._h, nn.ctypes.data_as(I), nb.ctypes.data_as(I)),
                  "wv_set_laplacian")

    def load(self, w):
        """tables, state and sources of a workload dict"""
        self.set_tables(w["xy_in"], w["tcc"], w["tc"], w["tgamma"])
        self.set_state(w["phi"], w["psi"])
        if w.get("sources"):
            self.set_sources(w["sources"])

    def step(self, n):
        self._chk(self._L.wv_step(self._h, int(n)), "wv_step")

    def prepare(self, n):
        """tile tables and kernel loads of a later step(n), done now"""
        self._chk(self._L.wv_prepare(self._h, int(n)), "wv_prepare")

    def step_timed(self, n):
        ms = ctypes.c_double()
        nl = ctypes.c_long()
        self._chk(self._L.wv_step_timed(self._h, int(n), ctypes.byref(ms), ctypes.byref(nl)),
                  "wv_step_timed")
        return ms.value, nl.value

    def kernel_ms(self):
        """(average duration in ms, number) of the launches of the last
        step_timed at the full fused depth (every launch is event-timed)"""
        a, b = ctypes.c_double(), ctypes.c_double()
        self._chk(self._L.wv_last_kernel_ms(self._h, ctypes.byref(a), ctypes.byref(b)), "wv_last_kernel_ms")
        return a.value, int(b.value)

    def sync(self):
        self._chk(self._L.wv_sync(self._h), "wv_sync")

    def steps_done(self):
        return self._L.wv_steps_done(self._h)

    def energy(self):
        e = ctypes.c_double()
        self._chk(self._L.wv_energy(self._h, ctypes.byref(e)), "wv_energy")
        return e.value

    def tiling(self):
        k, L, w = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        self._chk(self._L.wv_get_tiling(self._h, ctypes.byref(k), ctypes.byref(L), ctypes.byref(w)),
                  "wv_get_tiling")
        return k.value, L.value, w.value

    def tile_counts(self, tb_depth=0):
        """(wide tiles, one-wave tiles, workgroups, output cells covered) of
        one launch of fused depth tb_depth (0: the current depth)"""
        a, b, c, n = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_long()
        self._chk(self._L.wv_get_tile_counts(self._h, int(tb_depth), ctypes.byref(a), ctypes.byref(b),
                                             ctypes.byref(c), ctypes.byref(n)), "wv_get_tile_counts")
        return a.value, b.value, c.value, n.value

    def set_tiling(self, tb_depth, rows_per_tile):
        self._chk(self._L.wv_set_tiling(self._h, int(tb_depth), int(rows_per_tile)), "wv_set_tiling")

    def autotune(self, steps_per_trial=64):
        g = ctypes.c_double()
        self._chk(self._L.wv_autotune(self._h, int(steps_per_trial), ctypes.byref(g)), "wv_autotune")
        return g.value

    # ---- drop-in forms (reference signatures) --------------------------------
    def evolve_wave_half_flat(self, phi_in, psi_in, phi_out, xy_in, tc, tcc, tgamma):
        """evolve_wave_half of wave_3d.c:600 -- in place on phi_out."""
        for a, nm in ((phi_in, "phi_in"), (psi_in, "psi_in"), (phi_out, "phi_out"), (tc, "tc"), (tcc, "tcc"),
                      (tgamma, "tgamma")):
            _inplace(a, self.shape, np.float64, nm)
        _inplace(xy_in, self.shape, np.int16, "xy_in")
        self._chk(self._L.wv_evolve_wave_half_flat(self._h, _d(phi_in), _d(psi_in), _d(phi_out),
                                                   xy_in.ctypes.data_as(_S), _d(tc), _d(tcc),
                                                   _d(tgamma)), "wv_evolve_wave_half_flat")

    def evolve_wave_flat(self, phi, psi, tmp, xy_in, tc, tcc, tgamma):
        """evolve_wave of wave_3d.c:1010 -- three steps in place."""
        for a, nm in ((phi, "phi"), (psi, "psi"), (tmp, "tmp"), (tc, "tc"), (tcc, "tcc"), (tgamma, "tgamma")):
            _inplace(a, self.shape, np.float64, nm)
        _inplace(xy_in, self.shape, np.int16, "xy_in")
        self._chk(self._L.wv_evolve_wave_flat(self._h, _d(phi), _d(psi), _d(tmp),
                                              xy_in.ctypes.data_as(_S), _d(tc), _d(tcc), _d(tgamma)),
                  "wv_evolve_wave_flat")

    def _rows(self, a, t):
        """array of row pointers into a C-contiguous 2D array (double *a[NX])"""
        dt = np.int16 if t is ctypes.c_short else np.float64
        if not isinstance(a, np.ndarray) or a.dtype != dt or a.shape != self.shape or not a.flags.c_contiguous:
            raise ValueError("row-pointer arrays must be C-contiguous %s of shape %s" % (np.dtype(dt), self.shape))
        n = a.shape[0]
        base = a.ctypes.data
        stride = a.strides[0]
        arr = (ctypes.POINTER(t) * n)(*[ctypes.cast(base + i * stride, ctypes.POINTER(t)) for i in range(n)])
        return arr

    def evolve_wave_half_rows(self, phi_in, psi_in, phi_out, xy_in, tcc, tc, tgamma):
        """evolve_wave_half of wave_billiard.c:302 on row pointers -- in place on phi_out."""
        R = [self._rows(a, ctypes.c_double) for a in (phi_in, psi_in, phi_out)]
        M = self._rows(xy_in, ctypes.c_short)
        T = [self._rows(a, ctypes.c_double) for a in (tcc, tc, tgamma)]
        self._chk(self._L.wv_evolve_wave_half_rows(self._h, R[0], R[1], R[2], M, T[0], T[1], T[2]),
                  "wv_evolve_wave_half_rows")

    def evolve_wave_rows(self, phi, psi, tmp, xy_in, tcc, tc, tgamma):
        """evolve_wave of wave_billiard.c:613 on row pointers (three steps, in place)."""
        R = [self._rows(a, ctypes.c_double) for a in (phi, psi, tmp)]
        M = self._rows(xy_in, ctypes.c_short)
        T = [self._rows(a, ctypes.c_double) for a in (tcc, tc, tgamma)]
        self._chk(self._L.wv_evolve_wave_rows(self._h, R[0], R[1], R[2], M, T[0], T[1], T[2]),
                  "wv_evolve_wave_rows")


class Batch:
    """Many independent grids stepped together (a wv_batch)."""

    def __init__(self, params_list, **overrides):
        ps = []
        for p in params_list:
            q = dict(p)
            q.update(overrides)
            ps.append(make_params(q))
        self.n = len(ps)
        self.shapes = [(p.nx, p.ny) for p in ps]
        arr = (WvParams * self.n)(*ps)
        self._L = load()
        self._h = ctypes.c_void_p()
        rc = self._L.wv_batch_create(ctypes.byref(self._h), self.n, arr)
        if rc:
            msg = self._L.wv_batch_last_error(self._h).decode()
            self._L.wv_batch_destroy(self._h)
            self._h = None
            raise WaveError("wv_batch_create: %d %s" % (rc, msg))

    def close(self):
        if self._h:
            self._L.wv_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc, what):
        if rc:
            raise WaveError("%s: %d %s" % (what, rc, self._L.wv_batch_last_error(self._h).decode()))

    def set_tables(self, k, xy_in, tcc, tc, tgamma):
        xy = _i16(xy_in, self.shapes[k])
        a, b, c = (_f64(t, self.shapes[k]) for t in (tcc, tc, tgamma))
        self._chk(self._L.wv_batch_set_tables(self._h, k, xy.ctypes.data_as(_S), _d(a), _d(b), _d(c)),
                  "wv_batch_set_tables")

    def set_state(self, k, phi, psi):
        a, b = _f64(phi, self.shapes[k]), _f64(psi, self.shapes[k])
        self._chk(self._L.wv_batch_set_state(self._h, k, _d(a), _d(b)), "wv_batch_set_state")

    def get_state(self, k):
        a = np.empty(self.shapes[k])
        b = np.empty(self.shapes[k])
        self._chk(self._L.wv_batch_get_state(self._h, k, _d(a), _d(b)), "wv_batch_get_state")
        return a, b

    def load(self, k, w):
        self.set_tables(k, w["xy_in"], w["tcc"], w["tc"], w["tgamma"])
        self.set_state(k, w["phi"], w["psi"])

    def init_circular_wave(self, k, x0, y0, amp, variance, wavelength, box):
        bx = np.ascontiguousarray(box, np.float64).ravel()
        self._chk(self._L.wv_batch_init_circular_wave(self._h, k, x0, y0, amp, variance, wavelength, _d(bx)),
                  "wv_batch_init_circular_wave")

    def add_circular_wave(self, k, factor, x0, y0, amp, variance, wavelength, box):
        bx = np.ascontiguousarray(box, np.float64).ravel()
        self._chk(self._L.wv_batch_add_circular_wave(self._h, k, factor, x0, y0, amp, variance, wavelength,
                                                     _d(bx)), "wv_batch_add_circular_wave")

    def _box(self, box):
        bx = np.ascontiguousarray(box, np.float64).ravel()
        if bx.size != 4:
            raise ValueError("box must hold XMIN, XMAX, YMIN, YMAX")
        return bx

    def set_wells_tables(self, k, xy_in, xc, yc, height, sigma, use_tanh, courant, courantb, gamma, box, copy=True):
        """Gaussian-well media of grid k built on the GPU (wv_batch_set_wells_tables)"""
        xy = _i16(xy_in, self.shapes[k])
        xc = np.ascontiguousarray(xc, np.float64).ravel()
        yc = np.ascontiguousarray(yc, np.float64).ravel()
        hh = None if height is None else np.ascontiguousarray(height, np.float64).ravel()
        if yc.size != xc.size or (hh is not None and hh.size != xc.size):
            raise ValueError("wells: xc, yc (and height) of one length")
        t = [np.empty(self.shapes[k]) for _ in range(3)] if copy else None
        outs = [_d(a) for a in t] if copy else [None, None, None]
        self._chk(self._L.wv_batch_set_wells_tables(self._h, k, xy.ctypes.data_as(_S), len(xc), _d(xc), _d(yc),
                                                    None if hh is None else _d(hh), sigma, int(use_tanh), courant,
                                                    courantb, gamma, _d(self._box(box)), *outs),
                  "wv_batch_set_wells_tables")
        return tuple(t) if copy else None

    def set_ior_tables(self, k, xy_in, ior, box, params=None, copy=True, **kw):
        """init_ior_2d's per-cell model ior for grid k (wv_batch_set_ior_tables)"""
        P = params if params is not None else ior_params(**kw)
        xy = _i16(xy_in, self.shapes[k])
        t = [np.empty(self.shapes[k]) for _ in range(3)] if copy else None
        outs = [_d(a) for a in t] if copy else [None, None, None]
        self._chk(self._L.wv_batch_set_ior_tables(self._h, k, xy.ctypes.data_as(_S), int(ior), _d(self._box(box)),
                                                  ctypes.byref(P), *outs), "wv_batch_set_ior_tables")
        return tuple(t) if copy else None

    def build_domain_mask(self, k, domain, box, params=None, circles=None, **kw):
        """xy_in_billiard_single_domain over grid k (wv_batch_build_domain_mask)"""
        P = params if params is not None else domain_params(**kw)
        c = np.zeros((4, 1)) if circles is None else np.ascontiguousarray(circles, np.float64)
        if c.ndim != 2 or c.shape[0] != 4:
            raise ValueError("circles must be (4, n): xc, yc, radius, active")
        out = np.empty(self.shapes[k], np.int16)
        self._chk(self._L.wv_batch_build_domain_mask(self._h, k, int(domain), _d(self._box(box)), ctypes.byref(P),
                                                     0 if circles is None else c.shape[1], _d(c),
                                                     out.ctypes.data_as(_S)), "wv_batch_build_domain_mask")
        return out

    def build_list_domain_mask(self, k, domain, box, params=None, lists=None, **kw):
        """wv_build_list_domain_mask over grid k (wv_batch_build_list_domain_mask)"""
        P = params if params is not None else domain_params(**kw)
        dl = domain_lists(**(lists or {}))
        out = np.empty(self.shapes[k], np.int16)
        self._chk(self._L.wv_batch_build_list_domain_mask(self._h, k, int(domain), _d(self._box(box)),
                                                          ctypes.byref(P), ctypes.byref(dl), out.ctypes.data_as(_S)),
                  "wv_batch_build_list_domain_mask")
        return out

    def step(self, n):
        self._chk(self._L.wv_batch_step(self._h, int(n)), "wv_batch_step")

    def prepare(self, n):
        """tile tables, kernel loads and the graph of a later step(n), done now"""
        self._chk(self._L.wv_batch_prepare(self._h, int(n)), "wv_batch_prepare")

    def step_timed(self, n):
        ms = ctypes.c_double()
        nl = ctypes.c_long()
        self._chk(self._L.wv_batch_step_timed(self._h, int(n), ctypes.byref(ms), ctypes.byref(nl)),
                  "wv_batch_step_timed")
        return ms.value, nl.value

    def kernel_ms(self):
        """(average duration in ms, number) of the event-timed launches of the
        last step_timed"""
        a, b = ctypes.c_double(), ctypes.c_double()
        self._chk(self._L.wv_batch_last_kernel_ms(self._h, ctypes.byref(a), ctypes.byref(b)),
                  "wv_batch_last_kernel_ms")
        return a.value, int(b.value)

    def energy(self):
        e = np.empty(self.n)
        self._chk(self._L.wv_batch_energy(self._h, _d(e)), "wv_batch_energy")
        return e

    def set_tiling(self, tb_depth, rows_per_tile):
        self._chk(self._L.wv_batch_set_tiling(self._h, int(tb_depth), int(rows_per_tile)),
                  "wv_batch_set_tiling")

    def tiling(self):
        k, L, w = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        self._chk(self._L.wv_batch_get_tiling(self._h, ctypes.byref(k), ctypes.byref(L), ctypes.byref(w)),
                  "wv_batch_get_tiling")
        return k.value, L.value, w.value

    def autotune(self, steps_per_trial=64):
        g = ctypes.c_double()
        self._chk(self._L.wv_batch_autotune(self._h, int(steps_per_trial), ctypes.byref(g)), "wv_batch_autotune")
        return g.value
