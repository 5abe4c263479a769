"""TIFF frames coded on the GPU (wv_tiff_encode / wv_tiff_write) against the
files libtiff wrote for writetiff's calls (sub_wave.c:62-123,
tests/golden/tiff_cases.npz) and against the oracle (orc_tiff) on full-size
frames: byte for byte, LZW strips and layout alike."""
import ctypes
import os

import numpy as np
import pytest

from conftest import say
from gen_inputs import TIFF_DESC, tiff_cases, tiff_rgb

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def W():
    import wavehip
    wavehip.load()
    w = wavehip.Wave(dict(nx=64, ny=64))
    yield w
    w.close()


def test_tiff_matches_libtiff_fixtures(W):
    g = np.load(os.path.join(GOLD, "tiff_cases.npz"))
    say("tiff fixtures", len(tiff_cases()))
    for tag, h, w, content, desc, comp in tiff_cases():
        assert W.tiff(tiff_rgb(tag, h, w, content), desc, comp) == g[tag].tobytes(), tag


@pytest.mark.parametrize("h,w,content", [(1080, 1920, "smooth"), (720, 1280, "rand"), (16, 8192, "rand"),
                                         (9, 8192, "smooth"), (300, 2185, "low"), (1, 1, "const"),
                                         (2, 3334, "rand")])
def test_tiff_matches_oracle(W, oracle, h, w, content):
    say("tiff", h, w, content)
    rgb = tiff_rgb("gpu_%d_%d_%s" % (h, w, content), h, w, content)
    for comp in (5, 1):
        assert W.tiff(rgb, TIFF_DESC, comp) == oracle.tiff(rgb, TIFF_DESC, comp), comp


def test_tiff_buffers_serve_smaller_frames(W, oracle):
    # buffers grown for a large frame serve smaller ones after it
    for h, w in ((700, 4000), (33, 50), (1, 8192), (480, 640)):
        say("tiff sequence", h, w)
        rgb = tiff_rgb("seq_%d_%d" % (h, w), h, w, "smooth")
        rgb[h // 2:, : w // 3] = 0
        assert W.tiff(rgb) == oracle.tiff(rgb)


def test_tiff_write_file(W, oracle, tmp_path):
    rgb = tiff_rgb("file", 120, 160, "smooth")
    path = tmp_path / "wave.00001.tif"
    W.write_tiff(path, rgb)
    assert path.read_bytes() == oracle.tiff(rgb)


def test_tiff_errors(W):
    import wavehip
    L = wavehip.load()
    rgb = np.zeros((4, 8, 3), np.uint8)
    cap = L.wv_tiff_bound(8, 4, TIFF_DESC)
    assert cap > 0 and L.wv_tiff_bound(0, 4, TIFF_DESC) == 0 and L.wv_tiff_bound(8193, 4, TIFF_DESC) == 0
    out = np.zeros(cap, np.uint8)
    n = ctypes.c_size_t(0)
    args = (rgb.ctypes.data, 8, 4, TIFF_DESC)
    assert L.wv_tiff_encode(W._h, *args, 7, out.ctypes.data, cap, ctypes.byref(n)) == wavehip.WV_EUNSUPPORTED
    assert L.wv_tiff_encode(W._h, *args, 5, out.ctypes.data, cap - 1, ctypes.byref(n)) == wavehip.WV_EINVAL
    assert L.wv_tiff_encode(W._h, rgb.ctypes.data, 8193, 4, TIFF_DESC, 5, out.ctypes.data, cap,
                            ctypes.byref(n)) == wavehip.WV_EINVAL
    assert L.wv_tiff_write(W._h, b"/nonexistent-dir/x.tif", TIFF_DESC, rgb.ctypes.data, 8, 4, 5) == wavehip.WV_EIO
    with pytest.raises(ValueError):
        W.tiff(np.zeros((2, 9000, 3), np.uint8))
