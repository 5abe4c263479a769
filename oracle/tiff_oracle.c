/*
 * tiff_oracle.c -- CPU oracle of the TIFF frame writer (test infrastructure,
 * see wave_oracle.h).
 *
 * The reference saves frames with writetiff (sub_wave.c:62-123): libtiff
 * TIFFSetField for ImageWidth, ImageLength, BitsPerSample 8, Compression
 * (COMPRESSION_LZW at every call site, sub_wave.c:210,226), Photometric RGB,
 * Orientation BOTLEFT, SamplesPerPixel 3, PlanarConfig CONTIG, RowsPerStrip 1
 * and ImageDescription, then TIFFWriteScanline of the glReadPixels rows
 * (bottom row first) into lines height-1 .. 0.  The coder and the file layout
 * are libtiff's (a third-party library, version 4.2.0 in this image); this
 * file restates the published algorithm it implements:
 *  - LZW of TIFF 6.0 section 13: 8-bit roots, Clear 256, EndOfInformation
 *    257, codes from 258, 9..12-bit codes packed most significant bit
 *    first, the width growing one code early ("early change": after the
 *    entry that makes the next free code exceed 2^bits - 1), each strip
 *    starting with Clear and ending with EndOfInformation, the last byte
 *    padded with zero bits;
 *  - libtiff's encoder resets the table (a Clear at the current width) once
 *    the next free code would be 4094, and every 10000 input bytes compares
 *    the compression ratio (input bytes * 256 / output bits) with the last
 *    one and resets when it has not improved;
 *  - the file: little-endian header, strips in the order they were written,
 *    the IFD at the next even offset (12 entries in tag order), then the
 *    values that do not fit in an entry: BitsPerSample, StripByteCounts,
 *    StripOffsets (LONG), ImageDescription.  StripByteCounts are LONG for
 *    one strip or when the uncompressed strip is at least 65535/10 bytes
 *    under LZW (over 65535 bytes without compression), else SHORT.
 * Both were checked byte for byte against libtiff's own output for the
 * reference's call sequence (tests/test_oracle_tiff.py).
 */
#include <stdlib.h>
#include <string.h>

#include "wave_oracle.h"

#define LZW_CLEAR 256
#define LZW_EOI 257
#define LZW_FIRST 258
#define LZW_FULL 4094
#define LZW_GAP 10000

typedef struct {
    unsigned char *op;
    unsigned long data;
    int bits, nbits;
    long outbits;
} bitw;

static void put_code(bitw *w, int code)
{
    w->data = (w->data << w->nbits) | (unsigned long)code;
    w->bits += w->nbits;
    while (w->bits >= 8) {
        *w->op++ = (unsigned char)(w->data >> (w->bits - 8));
        w->bits -= 8;
    }
    w->outbits += w->nbits;
}

/* table of (prefix code, byte) -> code, cleared by bumping a generation */
typedef struct {
    unsigned *gen;
    unsigned short *code;
    unsigned cur;
} lzw_tab;

static int tab_get(const lzw_tab *t, int ent, int c)
{
    const size_t k = (size_t)ent * 256 + (size_t)c;
    return t->gen[k] == t->cur ? t->code[k] : -1;
}

static void tab_put(lzw_tab *t, int ent, int c, int code)
{
    const size_t k = (size_t)ent * 256 + (size_t)c;
    t->gen[k] = t->cur;
    t->code[k] = (unsigned short)code;
}

static long lzw_strip(lzw_tab *t, const unsigned char *in, long n, unsigned char *out)
{
    bitw w = {out, 0, 0, 9, 0};
    int free_ent = LZW_FIRST, maxcode = 511;
    long incount = 0, checkpoint = LZW_GAP, ratio = 0;
    t->cur++;
    if (n <= 0) {
        put_code(&w, LZW_EOI);
    } else {
        put_code(&w, LZW_CLEAR);
        int ent = in[0];
        incount = 1;
        for (long k = 1; k < n; k++) {
            const int c = in[k];
            incount++;
            const int hit = tab_get(t, ent, c);
            if (hit >= 0) {
                ent = hit;
                continue;
            }
            put_code(&w, ent);
            tab_put(t, ent, c, free_ent++);
            ent = c;
            int reset = 0;
            if (free_ent == LZW_FULL) {
                reset = 1;
            } else if (free_ent > maxcode) {
                w.nbits++;
                maxcode = (1 << w.nbits) - 1;
            } else if (incount >= checkpoint) {
                checkpoint = incount + LZW_GAP;
                const long rat = (incount << 8) / w.outbits;
                if (rat <= ratio)
                    reset = 1;
                else
                    ratio = rat;
            }
            if (reset) {
                t->cur++;
                ratio = 0;
                incount = 0;
                w.outbits = 0;
                free_ent = LZW_FIRST;
                put_code(&w, LZW_CLEAR);
                w.nbits = 9;
                maxcode = 511;
            }
        }
        put_code(&w, ent);
        free_ent++;
        if (free_ent == LZW_FULL) {
            put_code(&w, LZW_CLEAR);
            w.nbits = 9;
        } else if (free_ent > maxcode) {
            w.nbits++;
        }
        put_code(&w, LZW_EOI);
    }
    if (w.bits > 0) *w.op++ = (unsigned char)((w.data << (8 - w.bits)) & 0xff);
    return (long)(w.op - out);
}

long orc_lzw_strip(const unsigned char *in, long n, unsigned char *out)
{
    lzw_tab t;
    t.gen = (unsigned *)calloc(4096 * 256, sizeof(unsigned));
    t.code = (unsigned short *)malloc(4096 * 256 * sizeof(unsigned short));
    t.cur = 0;
    if (!t.gen || !t.code) {
        free(t.gen);
        free(t.code);
        return -1;
    }
    const long len = lzw_strip(&t, in, n, out);
    free(t.gen);
    free(t.code);
    return len;
}

static unsigned char *le16(unsigned char *p, unsigned v)
{
    p[0] = (unsigned char)v;
    p[1] = (unsigned char)(v >> 8);
    return p + 2;
}

static unsigned char *le32(unsigned char *p, unsigned long v)
{
    p[0] = (unsigned char)v;
    p[1] = (unsigned char)(v >> 8);
    p[2] = (unsigned char)(v >> 16);
    p[3] = (unsigned char)(v >> 24);
    return p + 4;
}

/* one IFD entry; inline values are left-justified in the 4-byte field */
static unsigned char *entry(unsigned char *p, unsigned tag, unsigned type, unsigned long count,
                            unsigned long value)
{
    p = le16(p, tag);
    p = le16(p, type);
    p = le32(p, count);
    return le32(p, value);
}

long orc_tiff_bound(int width, int height, const char *description)
{
    const long n = 3L * width;
    const long strip = n + (n + 2) / 2 + 16;
    return 8 + (long)height * strip + 2 + 12 * 12 + 4 + 6 + 8L * height + (long)strlen(description) + 1 + 16;
}

long orc_tiff(const unsigned char *rgb, int width, int height, const char *description, int compression,
              unsigned char *out, long cap)
{
    if (width < 1 || height < 1 || width > 65535 || height > 65535) return -1;
    if (compression != 1 && compression != 5) return -1;
    if (cap < orc_tiff_bound(width, height, description)) return -1;
    const long n = 3L * width;
    unsigned long *off = (unsigned long *)malloc(sizeof(unsigned long) * (size_t)height);
    unsigned long *cnt = (unsigned long *)malloc(sizeof(unsigned long) * (size_t)height);
    lzw_tab t = {NULL, NULL, 0};
    if (compression == 5) {
        t.gen = (unsigned *)calloc(4096 * 256, sizeof(unsigned));
        t.code = (unsigned short *)malloc(4096 * 256 * sizeof(unsigned short));
    }
    if (!off || !cnt || (compression == 5 && (!t.gen || !t.code))) {
        free(off);
        free(cnt);
        free(t.gen);
        free(t.code);
        return -1;
    }
    unsigned char *p = out;
    memcpy(p, "II*\0", 4);
    p += 8;
    /* glReadPixels row r (r = 0 at the bottom) is TIFF line height-1-r,
       written in the order r = 0, 1, ... (sub_wave.c:113-122) */
    for (int r = 0; r < height; r++) {
        const int line = height - 1 - r;
        const unsigned char *src = rgb + (size_t)r * (size_t)n;
        off[line] = (unsigned long)(p - out);
        long len;
        if (compression == 5) {
            len = lzw_strip(&t, src, n, p);
        } else {
            memcpy(p, src, (size_t)n);
            len = n;
        }
        cnt[line] = (unsigned long)len;
        p += len;
    }
    free(t.gen);
    free(t.code);
    if ((p - out) & 1) *p++ = 0;
    const unsigned long ifd = (unsigned long)(p - out);
    le32(out + 4, ifd);
    const size_t dlen = strlen(description) + 1;
    unsigned long data = ifd + 2 + 12 * 12 + 4;
    const unsigned long bps_at = data;
    data += 6;
    const int one = (height == 1);
    /* LONG counts for one strip, or when the uncompressed strip could code
       to 65535 bytes or more (allowing LZW ten times its input), else SHORT */
    const int sbc_long = one || (compression == 5 ? n >= 65535 / 10 : n > 65535);
    const unsigned long sbc_bytes = (sbc_long ? 4UL : 2UL) * (unsigned long)height;
    const unsigned long sbc_at = data;
    if (sbc_bytes > 4) data += sbc_bytes;
    const unsigned long so_at = data;
    if (!one) data += 4UL * height;
    const unsigned long desc_at = data;
    unsigned char *e = le16(p, 12);
    e = entry(e, 256, 3, 1, (unsigned long)width);
    e = entry(e, 257, 3, 1, (unsigned long)height);
    e = entry(e, 258, 3, 3, bps_at);
    e = entry(e, 259, 3, 1, (unsigned long)compression);
    e = entry(e, 262, 3, 1, 2);
    unsigned long dval = desc_at;
    if (dlen <= 4) {
        unsigned char b[4] = {0, 0, 0, 0};
        memcpy(b, description, dlen);
        dval = b[0] | (unsigned long)b[1] << 8 | (unsigned long)b[2] << 16 | (unsigned long)b[3] << 24;
    }
    e = entry(e, 270, 2, dlen, dval);
    e = entry(e, 273, 4, (unsigned long)height, one ? off[0] : so_at);
    e = entry(e, 274, 3, 1, 4);
    e = entry(e, 277, 3, 1, 3);
    e = entry(e, 278, 3, 1, 1);
    if (sbc_bytes <= 4)
        e = entry(e, 279, sbc_long ? 4 : 3, (unsigned long)height, sbc_long ? cnt[0] : cnt[0] | cnt[1] << 16);
    else
        e = entry(e, 279, sbc_long ? 4 : 3, (unsigned long)height, sbc_at);
    e = entry(e, 284, 3, 1, 1);
    e = le32(e, 0);
    e = le16(e, 8);
    e = le16(e, 8);
    e = le16(e, 8);
    if (sbc_bytes > 4)
        for (int r = 0; r < height; r++) e = sbc_long ? le32(e, cnt[r]) : le16(e, (unsigned)cnt[r]);
    if (!one)
        for (int r = 0; r < height; r++) e = le32(e, off[r]);
    if (dlen > 4) {
        memcpy(e, description, dlen);
        e += dlen;
    }
    free(off);
    free(cnt);
    return (long)(e - out);
}
