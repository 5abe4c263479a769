"""The TIFF frame writer's oracle (orc_tiff, orc_lzw_strip) against libtiff,
the library writetiff (sub_wave.c:62-123) calls.

tests/golden/tiff_cases.npz holds the files libtiff 4.2.0 wrote for
writetiff's TIFFSetField / TIFFWriteScanline sequence over 134 frames: widths
around the LZW code-width steps, the SHORT/LONG StripByteCounts switch
(2184/2185 pixels), the 10000-byte ratio checkpoint and the 4094-code table
reset; one to three rows; LZW and no compression; descriptions stored inline
and out of line.  The oracle reproduces every file byte for byte, so a slip in
the coder (a late code-width change, LSB-first packing, no reset at 4094) or
in the layout (top-down strip order, a wrong StripByteCounts type) fails them.
An LZW decoder written from TIFF 6.0 section 13 and PIL's TIFF reader check
the round trip; where libtiff loads, larger frames are compared with it live.
"""
import io
import os

import numpy as np
import pytest

import libtiff_writer
from gen_inputs import tiff_cases, tiff_rgb

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def gold():
    return np.load(os.path.join(GOLD, "tiff_cases.npz"))


def lzw_decode(data):
    """TIFF 6.0 section 13 LZW decoder: MSB-first codes of 9..12 bits, the
    width growing once the table holds 2^width - 1 entries"""
    total = len(data) * 8
    bits = int.from_bytes(bytes(data), "big")
    pos, width = 0, 9
    table = [bytes([k]) for k in range(256)] + [b"", b""]
    prev, out = None, bytearray()
    while True:
        assert pos + width <= total, "stream ends without EndOfInformation"
        code = (bits >> (total - pos - width)) & ((1 << width) - 1)
        pos += width
        if code == 257:
            break
        if code == 256:
            del table[258:]
            width, prev = 9, None
            continue
        if prev is None:
            entry = table[code]
        else:
            if code < len(table):
                entry = table[code]
            else:
                assert code == len(table), "code %d beyond the table (%d)" % (code, len(table))
                entry = prev + prev[:1]
            table.append(prev + entry[:1])
            if len(table) + 1 >= (1 << width) and width < 12:
                width += 1
        out += entry
        prev = entry
    assert total - pos < 8, "bytes after EndOfInformation"
    return bytes(out)


@pytest.mark.parametrize("case", tiff_cases(), ids=lambda c: c[0])
def test_tiff_oracle_matches_libtiff_fixture(oracle, gold, case):
    tag, h, w, content, desc, comp = case
    got = oracle.tiff(tiff_rgb(tag, h, w, content), desc, comp)
    assert got == gold[tag].tobytes()


@pytest.mark.parametrize("n,content", [(1, "rand"), (2, "const"), (3, "rand"), (600, "low"), (5000, "rand"),
                                       (24576, "rand"), (24576, "smooth"), (9000, "const")])
def test_lzw_strip_decodes_to_its_input(oracle, n, content):
    rgb = tiff_rgb("lzw_rt_%d_%s" % (n, content), 1, max(n // 3, 1), content).reshape(-1)
    src = np.resize(rgb, n).tobytes()
    code = oracle.lzw_strip(src)
    assert code[:2] != b"", "empty strip"
    assert lzw_decode(code) == src


def test_lzw_known_answer(oracle):
    # "aaaa": Clear, 'a', 258 ("aa"), 'a', EndOfInformation at 9 bits each
    # = 100000000 001100001 100000010 001100001 100000001, then 3 zero bits
    want = int("100000000" "001100001" "100000010" "001100001" "100000001" "000", 2).to_bytes(6, "big")
    assert oracle.lzw_strip(b"aaaa") == want


@pytest.mark.parametrize("comp", [5, 1])
def test_pil_reads_oracle_file(oracle, comp):
    Image = pytest.importorskip("PIL.Image")
    rgb = tiff_rgb("pil_rt", 37, 211, "smooth")
    img = Image.open(io.BytesIO(oracle.tiff(rgb, b"Wave equation in a planar domain", comp)))
    assert img.size == (211, 37) and img.mode == "RGB"
    # line 0 holds the last glReadPixels row; Orientation BOTLEFT says line 0
    # is the bottom of the picture, so a reader that honours it (PIL does, for
    # libtiff's own files too) returns the rows in glReadPixels order
    assert np.array_equal(np.asarray(img), rgb)


@pytest.mark.parametrize("h,w,content", [(5, 2500, "rand"), (4, 8192, "rand"), (3, 8192, "smooth"),
                                         (64, 1920, "smooth"), (2, 6667, "low")])
def test_tiff_oracle_matches_libtiff_live(oracle, tmp_path, h, w, content):
    if libtiff_writer.load() is None:
        pytest.skip("libtiff is not loadable here")
    rgb = tiff_rgb("live_%d_%d_%s" % (h, w, content), h, w, content)
    want = libtiff_writer.encode(rgb, b"Wave equation in a planar domain", 5, str(tmp_path))
    assert oracle.tiff(rgb) == want
