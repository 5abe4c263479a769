"""writetiff's libtiff calls (sub_wave.c:62-123) through ctypes: the tag set
TIFFSetField installs and the TIFFWriteScanline order (bottom glReadPixels row
first, into line height-1).  libtiff is a third-party library (4.2.0 under
/opt/conda/lib in this image); it writes the fixture files and, where it can
be loaded, checks the writers directly."""
import ctypes
import os

import numpy as np

_CANDIDATES = ("/opt/conda/lib/libtiff.so.5", "libtiff.so.5", "libtiff.so.6", "libtiff.so")
_lib = None


def load():
    """the libtiff shared library, or None"""
    global _lib
    if _lib is None:
        for name in _CANDIDATES:
            if name.startswith("/") and not os.path.exists(name):
                continue
            try:
                L = ctypes.CDLL(name)
            except OSError:
                continue
            L.TIFFOpen.restype = ctypes.c_void_p
            L.TIFFOpen.argtypes = [ctypes.c_char_p, ctypes.c_char_p]
            L.TIFFWriteScanline.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint16]
            L.TIFFClose.argtypes = [ctypes.c_void_p]
            L.TIFFGetVersion.restype = ctypes.c_char_p
            _lib = L
            break
    return _lib


def version():
    L = load()
    return None if L is None else L.TIFFGetVersion().decode().splitlines()[0]


def write(path, rgb, description=b"Wave equation in a planar domain", compression=5):
    """write rgb (uint8 (height, width, 3), bottom row first) as writetiff does"""
    L = load()
    if L is None:
        raise RuntimeError("libtiff is not available")
    rgb = np.ascontiguousarray(rgb, np.uint8)
    h, w = rgb.shape[0], rgb.shape[1]
    t = ctypes.c_void_p(L.TIFFOpen(str(path).encode(), b"w"))
    if not t.value:
        raise OSError("TIFFOpen failed for %s" % path)

    def field(tag, v):
        L.TIFFSetField(t, ctypes.c_uint32(tag), v)

    field(256, ctypes.c_uint32(w))        # IMAGEWIDTH
    field(257, ctypes.c_uint32(h))        # IMAGELENGTH
    field(258, ctypes.c_int(8))           # BITSPERSAMPLE
    field(259, ctypes.c_int(compression))
    field(262, ctypes.c_int(2))           # PHOTOMETRIC_RGB
    field(274, ctypes.c_int(4))           # ORIENTATION_BOTLEFT
    field(277, ctypes.c_int(3))           # SAMPLESPERPIXEL
    field(284, ctypes.c_int(1))           # PLANARCONFIG_CONTIG
    field(278, ctypes.c_int(1))           # ROWSPERSTRIP
    field(270, ctypes.c_char_p(bytes(description)))
    for i in range(h - 1, -1, -1):
        if L.TIFFWriteScanline(t, rgb[h - 1 - i].ctypes.data, i, 0) < 0:
            L.TIFFClose(t)
            raise OSError("TIFFWriteScanline failed")
    L.TIFFClose(t)


def encode(rgb, description=b"Wave equation in a planar domain", compression=5, tmpdir="/tmp"):
    """the bytes of the file write() produces"""
    path = os.path.join(tmpdir, "libtiff_writer_%d.tif" % os.getpid())
    try:
        write(path, rgb, description, compression)
        with open(path, "rb") as f:
            return f.read()
    finally:
        if os.path.exists(path):
            os.remove(path)
