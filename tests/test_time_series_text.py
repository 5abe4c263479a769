"""The SAVE_TIME_SERIES text of wave_billiard.c:902-905: each sample as
(long)(phi*1.0e16) printed with "%019ld\\n".  Host code of the library (no
device): checked against Python's own formatting of the same truncated
binary64 product (int() truncates toward zero like the C conversion, and
"%019d" pads after the sign like "%019ld"), including negatives, zero, values
whose scaled product sits next to an integer, and the error paths.
"""
import ctypes

import numpy as np

import wavehip


def expected(values, stride=1):
    return "".join("%019d\n" % int(float(x) * 1.0e16) for x in np.asarray(values)[::stride])


def test_lines_match_c_formatting():
    rng = np.random.default_rng(5)
    v = np.concatenate([
        [0.0, -0.0, 1.0, -1.0, 0.5, -0.25, 1e-16, -1e-16, 3e-17, 0.9999999999999999, -0.1234567890123456,
         2.5e-3, 7.0, -7.0, 9.2e2],
        rng.standard_normal(200), rng.standard_normal(50) * 1e-12, np.nextafter(1.0, 0.0) * np.arange(1, 11) / 7.0,
    ])
    assert wavehip.format_time_series(v) == expected(v)


def test_interleaved_probes_and_errors():
    L = wavehip.load()
    n, count = 3, 40
    v = np.random.default_rng(2).standard_normal(n * count)
    for k in range(n):
        cap = 21 * count
        buf = ctypes.create_string_buffer(cap)
        w = ctypes.c_long(-1)
        ptr = v[k:].ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        assert L.wv_format_time_series(ptr, count, n, buf, cap, ctypes.byref(w)) == 0
        assert buf.raw[:w.value].decode() == expected(v[k:], n)
    # too small a buffer: an error, nothing claimed written
    buf = ctypes.create_string_buffer(10)
    w = ctypes.c_long(-1)
    assert L.wv_format_time_series(v.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), 2, 1, buf, 10,
                                   ctypes.byref(w)) != 0
    assert w.value == 0
    # zero samples: nothing written
    assert L.wv_format_time_series(None, 0, 1, None, 0, ctypes.byref(w)) == 0 and w.value == 0
    # samples but no buffers: an error
    assert L.wv_format_time_series(None, 3, 1, None, 0, None) != 0
