"""The allocation-free entry point wv_build_mask_static against
wv_build_mask, bit for bit, through the C ABI: two calls in a row (the second
with a smaller input), a grid that needs several row bands of the fixed
scratch array, and the size limits."""
import numpy as np
import pytest

from conftest import say

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def wh():
    import wavehip
    wavehip.load()
    return wavehip


def circles(n, seed):
    import pyoracle
    rng = np.random.default_rng(seed)
    act = (rng.uniform(0, 1, n) > 0.1).astype(np.float64)
    return pyoracle.circle_list(rng.uniform(-2, 2, n), rng.uniform(-2, 2, n), rng.uniform(0.02, 0.3, n), act)


def test_static_mask_matches_allocating_builder(wh):
    box = (-2.0, 2.0, -2.0, 2.0)
    # large grid first (4096 x 4096: five bands of the 8 MiB scratch), then
    # a smaller grid and a shorter circle list; then the other domains
    cases = [((4096, 4096), wh.D_CIRCLES, 0.0, 0.0, circles(300, 1)),
             ((700, 333), wh.D_CIRCLES, 0.0, 0.0, circles(40, 2)),
             ((513, 257), wh.D_SINAI, 0.3, 0.0, None),
             ((300, 200), wh.D_EXT_ELLIPSE, 0.6, 0.3, None)]
    for (nx, ny), dom, lam, mu, circ in cases:
        say("static mask", nx, ny, dom)
        W = wh.Wave(dict(nx=nx, ny=ny, bcond=2))
        a = W.build_mask(dom, box, lam, mu, circ)
        b = W.build_mask(dom, box, lam, mu, circ, static=True)
        c = W.build_mask(dom, box, lam, mu, circ, static=True)   # second call, same scratch
        W.close()
        assert np.array_equal(a, b) and np.array_equal(a, c), (nx, ny, dom)
        assert 0 < int(a.sum()) < nx * ny


def test_static_mask_limits(wh):
    W = wh.Wave(dict(nx=16, ny=16, bcond=0))
    with pytest.raises(wh.WaveError):
        W.build_mask(wh.D_CIRCLES, (-1, 1, -1, 1), 0, 0, circles(40000, 3), static=True)
    ok = W.build_mask(wh.D_CIRCLES, (-1, 1, -1, 1), 0, 0, circles(5, 4), static=True)
    W.close()
    assert ok.shape == (16, 16)
