"""The frame-colour oracle (orc_frame) against the reference's own colour path.

tests/golden/frame_digests.json holds SHA-256 digests of the values and
colours wave_value (wave_common.c:1053-1171) computes cell by cell, through
color_scheme_palette / color_scheme_asym_palette (colors_waves.c) and the
palettes COL_JET, COL_HSLUV and COL_GRAY (colormaps.c; COL_HSLUV through
hsluv2rgb of the hsluv-c library the reference vendors), for 8 reference builds (every
colour scheme, the asymmetric energy scheme, attenuation and
RESCALE_COLOR_IN_CENTER), 10 plot types and 3 frames of running sums.  The
table palettes are checked against the reference build directly when
oracle/_ref is present (their tables are the program's, read at run time).
The drawing function's shading and fading (wave_common.c:1772-1820) call
OpenGL, so they are pinned by closed forms here.
"""
import hashlib
import json
import os

import numpy as np
import pytest

from gen_inputs import color_builds, color_cases, color_inputs

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.fixture(scope="module")
def digests():
    with open(os.path.join(GOLD, "frame_digests.json")) as f:
        return json.load(f)


def run_frames(oracle, prm, pal, plot, d, lut=None, **over):
    kw = {k: v for k, v in prm.items() if k != "twospeeds"}
    kw.update(over)
    p = oracle.make_params(24, 17, twospeeds=prm["twospeeds"])
    acc = oracle.FrameAccumulators(24, 17)
    out = []
    for t, (phi, psi) in enumerate(d["frames"]):
        f = oracle.frame_params(plot=plot, palette=pal, time=t, **kw)
        vals, rgb, drawn = oracle.frame(p, f, phi, psi, d["xy_in"], acc, color_scale=d["color_scale"], lut=lut)
        out.append((vals, rgb, drawn))
    return out


@pytest.mark.parametrize("case", color_cases(), ids=lambda c: "%s-pal%d-plot%d" % (c[0], c[2], c[3]))
def test_frame_oracle_matches_reference_digests(oracle, digests, case):
    tag, prm, pal, plot = case
    d = color_inputs(tag)
    ref = digests["%s/pal%d/plot%d" % (tag, pal, plot)]
    for t, (vals, rgb, _) in enumerate(run_frames(oracle, prm, pal, plot, d)):
        assert sha(vals) == ref[t]["values"], (t, float(vals[5, 7]).hex(), ref[t]["probe"])
        assert sha(rgb) == ref[t]["rgb"], t


def test_frame_digests_detect_planted_mistakes(oracle, digests):
    """A wrong slope, the symmetric scheme where the asymmetric one applies,
    a dropped attenuation or a dropped running sum fail the digests."""
    b = dict(color_builds())
    for tag, pal, plot, over in (("col_s3", 0, 1, dict(slope=1.0000001)),
                                 ("col_s1_asym", 0, 1, dict(asym=0)),
                                 ("col_s3_att", 0, 0, dict(attenuation=0.0)),
                                 ("col_s0", 2, 3, dict(plot=1))):
        d = color_inputs(tag)
        frames = run_frames(oracle, b[tag], pal, over.pop("plot", plot), d, **over)
        ref = digests["%s/pal%d/plot%d" % (tag, pal, plot)]
        assert any(sha(v) != r["values"] or sha(c) != r["rgb"] for (v, c, _), r in zip(frames, ref)), (tag, over)


@pytest.mark.skipif(not __import__("pyoracle").ref_available(), reason="oracle/_ref not built (needs /root/reference)")
def test_table_palettes_match_reference_build(oracle):
    """The table palettes (COL_TURBO .. COL_TWILIGHT_SHIFTED, COL_TURBO_CYCLIC)
    with the reference's own tables, read from the build at run time."""
    import pyoracle
    for tag, prm in color_builds():
        R = pyoracle.RefBuild(tag)
        d = color_inputs(tag)
        for pal in (10, 11, 12, 13, 14, 15, 16, 17, 18, 101):
            lut = R.palette_table(pal)
            for plot in range(10):
                if prm["scheme"] in (4, 5) and plot in (4, 5, 6, 9):
                    continue  # negative table indices: the reference reads outside its table
                frames = run_frames(oracle, prm, pal, plot, d, lut=lut)
                acc = pyoracle.FrameAccumulators(24, 17)
                for t, (phi, psi) in enumerate(d["frames"]):
                    vals, rgb, _ = frames[t]
                    for i in range(0, 24, 3):
                        for j in range(17):
                            v, c = R.wave_value(phi, psi, acc, d["color_scale"], d["xy_in"], 1.0, t, plot, pal, i, j)
                            assert np.float64(v).tobytes() == vals[i, j].tobytes(), (tag, pal, plot, t, i, j)
                            assert c.tobytes() == rgb[i, j].tobytes(), (tag, pal, plot, t, i, j)


def flat_case(oracle, nx=20, ny=16, **kw):
    p = oracle.make_params(nx, ny)
    xy = np.ones((nx, ny), np.int16)
    xy[3, 4] = 0
    return p, xy


def test_shading_closed_forms(oracle):
    """SHADE_2D, wave_common.c:1772-1791: a flat value field gives ca = 1
    (colours unchanged); a ramp of slope g along i gives
    ca = ((S - g*lx)/sqrt(S^2 + g^2) + 1)*0.4 + 0.2; the last block row and
    column stay unshaded."""
    p, xy = flat_case(oracle)
    phi = np.full((20, 16), 0.3)
    acc = oracle.FrameAccumulators(20, 16)
    f0 = oracle.frame_params(plot=0, shade=0)
    f1 = oracle.frame_params(plot=0, shade=1)
    _, rgb0, _ = oracle.frame(p, f0, phi, phi, xy, acc)
    _, rgb1, _ = oracle.frame(p, f1, phi, phi, xy, acc)
    assert np.array_equal(rgb0, rgb1)
    g = 0.01
    phi = 0.3 + g * np.arange(20)[:, None] * np.ones((1, 16))
    vals, rgb0, _ = oracle.frame(p, f0, phi, phi, xy, acc)
    _, rgb1, _ = oracle.frame(p, f1, phi, phi, xy, acc)
    S, lx = 0.5, 0.40824829
    gx = vals[1:, :-1] - vals[:-1, :-1]
    ca = ((-gx * lx + S) / np.sqrt(S * S + gx * gx) + 1.0) * 0.4 + 0.2
    assert np.allclose(rgb1[:-1, :-1], rgb0[:-1, :-1] * ca[..., None], rtol=1e-14, atol=0)
    assert np.array_equal(rgb1[-1], rgb0[-1]) and np.array_equal(rgb1[:, -1], rgb0[:, -1])
    assert abs(ca[0, 0] - ((S - g * lx) / np.hypot(S, g) + 1) * 0.4 - 0.2) < 1e-9


def test_fade_and_draw_rule(oracle):
    """FADE_IN_OBSTACLE and fade multiply drawn cells only; a cell is drawn
    when TWOSPEEDS or xy_in != 0 (wave_common.c:1797-1820)."""
    p, xy = flat_case(oracle)
    rng = np.random.default_rng(1)
    phi = rng.uniform(-0.1, 0.1, (20, 16))
    ft = rng.uniform(0.2, 1.0, (20, 16))
    acc = oracle.FrameAccumulators(20, 16)
    _, rgb0, drawn = oracle.frame(p, oracle.frame_params(plot=0), phi, phi, xy, acc, fade_table=ft)
    _, rgb1, _ = oracle.frame(p, oracle.frame_params(plot=0, fade_in_obstacle=1, fade=1, fade_value=0.5), phi, phi,
                              xy, acc, fade_table=ft)
    assert not drawn[3, 4] and drawn.sum() == 20 * 16 - 1
    assert np.array_equal(rgb1[3, 4], rgb0[3, 4])
    m = drawn[..., None]
    assert np.array_equal(np.where(m, rgb0 * ft[..., None] * 0.5, rgb0), rgb1)
    p2 = oracle.make_params(20, 16, twospeeds=1)
    _, _, drawn2 = oracle.frame(p2, oracle.frame_params(plot=0), phi, phi, xy, acc)
    assert drawn2.all()


def test_block_size_uses_block_cells(oracle):
    """With size 2 the values and unshaded colours are those of size 1 at
    the block cells, and shading takes the gradient to the next block cell."""
    p, xy = flat_case(oracle)
    rng = np.random.default_rng(2)
    phi = rng.uniform(-0.05, 0.05, (20, 16))
    acc = oracle.FrameAccumulators(20, 16)
    v1, c1, _ = oracle.frame(p, oracle.frame_params(plot=0), phi, phi, xy, acc)
    v2, c2, _ = oracle.frame(p, oracle.frame_params(plot=0, size=2), phi, phi, xy, acc)
    assert np.array_equal(v2[::2, ::2], v1[::2, ::2]) and np.array_equal(c2[::2, ::2], c1[::2, ::2])
    _, s2, _ = oracle.frame(p, oracle.frame_params(plot=0, size=2, shade=1), phi, phi, xy, acc)
    b = v1[::2, ::2]
    gx, gy = b[1:, :-1] - b[:-1, :-1], b[:-1, 1:] - b[:-1, :-1]
    S, lx, ly = 0.5, 0.40824829, 0.816496581
    ca = ((-gx * lx - gy * ly + S) / np.sqrt(S * S + gx * gx + gy * gy) + 1.0) * 0.4 + 0.2
    assert np.allclose(s2[::2, ::2][:-1, :-1], c1[::2, ::2][:-1, :-1] * ca[..., None], rtol=1e-14, atol=0)


def hsluv(oracle, h, s, l):
    import ctypes
    L = oracle.lib()
    L.orc_hsluv_rgb.argtypes = [ctypes.c_double] * 3 + [ctypes.POINTER(ctypes.c_double)]
    out = (ctypes.c_double * 3)()
    L.orc_hsluv_rgb(h, s, l, out)
    return np.array(out[:])


def test_hsluv_closed_forms(oracle):
    """HSLuv's defining property: at saturation 100 the colour lies on the
    sRGB gamut boundary (the largest in-gamut chroma), so a channel is 0 or
    1; below it the colour is inside.  Lightness 0 is black, 100 is white
    (the D65 white point through the sRGB matrix), saturation 0 is grey."""
    rng = np.random.default_rng(4)
    for _ in range(400):
        h, l = rng.uniform(0, 360), rng.uniform(1, 99)
        c = hsluv(oracle, h, 100.0, l)
        assert min(abs(c.min()), abs(c.max() - 1.0)) < 1e-9, (h, l, c)
        inner = hsluv(oracle, h, 60.0, l)
        assert inner.min() > 1e-9 and inner.max() < 1.0 - 1e-9, (h, l, inner)
    for h in (0.0, 77.0, 300.0):
        assert np.array_equal(hsluv(oracle, h, 50.0, 0.0), np.zeros(3))
        assert np.allclose(hsluv(oracle, h, 50.0, 100.0), 1.0, atol=1e-12)
        g = hsluv(oracle, h, 0.0, 37.0)
        assert np.allclose(g, g[0], atol=1e-14) and 0.0 < g[0] < 1.0
    # lightness is monotone along a grey ramp and the hue turns the colour
    ramp = [hsluv(oracle, 0.0, 0.0, l)[0] for l in np.linspace(1, 99, 50)]
    assert np.all(np.diff(ramp) > 0)
    assert not np.allclose(hsluv(oracle, 10.0, 80.0, 50.0), hsluv(oracle, 130.0, 80.0, 50.0))


@pytest.mark.skipif(not __import__("pyoracle").ref_available(), reason="oracle/_ref not built (needs /root/reference)")
def test_hsluv_matches_reference_build(oracle):
    """orc_hsluv_rgb against hsluv2rgb of the reference's colour build, bit for
    bit, over hues beyond [0, 360) and the lightness edges of the library"""
    import ctypes
    import pyoracle
    R = pyoracle.RefBuild("col_s3")
    D = ctypes.c_double
    R.L.hsluv2rgb.argtypes = [D, D, D, ctypes.POINTER(D), ctypes.POINTER(D), ctypes.POINTER(D)]
    rng = np.random.default_rng(6)
    for _ in range(3000):
        h = rng.uniform(-20, 380)
        s = rng.choice([0.0, 1e-9, 90.0, 100.0, rng.uniform(0, 100)])
        l = rng.choice([0.0, 1e-9, 0.5, 7.99, 8.0, 8.01, 99.99999995, 100.0, rng.uniform(0, 1), rng.uniform(0, 100)])
        r, g, b = D(), D(), D()
        R.L.hsluv2rgb(h, s, l, ctypes.byref(r), ctypes.byref(g), ctypes.byref(b))
        assert np.array_equal(np.array([r.value, g.value, b.value]).view(np.uint64),
                              hsluv(oracle, h, s, l).view(np.uint64)), (h, s, l)
