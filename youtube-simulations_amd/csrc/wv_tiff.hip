// LZW coding of TIFF frames on the device: the coder behind writetiff
// (sub_wave.c:62-123), which stores RGB frames as LZW strips of one row each
// (TIFFTAG_ROWSPERSTRIP 1, COMPRESSION_LZW at every call site, sub_wave.c:210,
// 226).  Strips are coded independently, so each is one workgroup: a single
// wave that stages the row in LDS and clears the code table there, then runs
// the (inherently serial) coder in step on wave-uniform values.  The code stream is libtiff's (version 4.2.0): TIFF 6.0 section 13
// LZW, codes MSB-first, the width growing one code early, Clear 256 at the
// start and EndOfInformation 257 at the end; a Clear at the current width
// when the next free code would be 4094, and every 10000 input bytes a
// compression-ratio check (input bytes * 256 / output bits) that resets the
// table when the ratio has not improved.  A second kernel packs the coded
// strips back to back in file order.
#include "wv_internal.h"

namespace {

// table slots: at most 3836 live codes (62% full); 6144 words plus a
// 1920-pixel row keep a workgroup under 32 KiB of LDS, five to a CU
constexpr uint32_t HSIZE = 6144;
constexpr uint32_t EMPTY = 0xffffffffu;  // no live entry has key 0xfffff (prefix 4095)

__device__ __forceinline__ uint32_t uni(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }

__device__ __forceinline__ uint32_t slot_of(uint32_t key)
{
    return (uint32_t)(((uint64_t)(key * 2654435761u) * HSIZE) >> 32);
}

// one strip per workgroup (one wave); n = 3*width bytes per row; strip r is
// the r-th row of rgb; out + r*stride receives its code bytes, lens[r] their
// count.  Every lane runs the coder on the same wave-uniform values (LDS
// reads of one address broadcast, readfirstlane moves them to scalar
// registers), so control flow is scalar; lane 0 alone writes.
__global__ __launch_bounds__(64) void k_lzw(const uint8_t *__restrict__ rgb, int n, uint8_t *__restrict__ strips,
                                            long stride, int *__restrict__ lens)
{
    extern __shared__ uint32_t sh[];
    uint32_t *tab = sh;
    uint32_t *roww = sh + HSIZE;  // the row, four bytes a word
    const int r = blockIdx.x;
    const uint32_t lane = threadIdx.x;
    const bool w0 = lane == 0;
    const uint8_t *src = rgb + (size_t)r * (size_t)n;
    uint8_t *rowb = reinterpret_cast<uint8_t *>(roww);
    for (int k = (int)lane; k < n; k += 64) rowb[k] = src[k];
    for (uint32_t k = lane; k < HSIZE; k += 64) tab[k] = EMPTY;
    __syncthreads();

    // codes accumulate MSB-first in acc (nacc < 32 bits pending between
    // calls); each full 32 bits go out as one word, byte-swapped so that the
    // stream's first byte lands at the lowest address
    uint32_t *out = reinterpret_cast<uint32_t *>(strips + (size_t)r * (size_t)stride);
    uint64_t acc = 0;
    int nacc = 0, nbits = 9, nword = 0, outbits = 0;
    auto put = [&](int code) {
        acc = (acc << nbits) | (uint32_t)code;
        nacc += nbits;
        if (nacc >= 32) {
            nacc -= 32;
            if (w0) out[nword] = __builtin_bswap32((uint32_t)(acc >> nacc));
            nword++;
        }
        outbits += nbits;
    };

    int free_ent = 258, maxcode = 511;
    int incount = 1, checkpoint = 10000;
    long ratio = 0;
    put(256);
    int ent = (int)(uni(roww[0]) & 0xffu);
    auto code_byte = [&](int c) {
        incount++;
        const uint32_t key = ((uint32_t)c << 12) | (uint32_t)ent;
        uint32_t h = slot_of(key);
        uint32_t e = uni(tab[h]);
        while (e != EMPTY && (e >> 12) != key) {
            h = (h + 1 == HSIZE) ? 0 : h + 1;
            e = uni(tab[h]);
        }
        if (e != EMPTY) {
            ent = (int)(e & 0xfffu);
            return;
        }
        put(ent);
        if (w0) tab[h] = (key << 12) | (uint32_t)free_ent;
        free_ent++;
        ent = c;
        bool reset = false;
        if (free_ent == 4094) {
            reset = true;
        } else if (free_ent > maxcode) {
            nbits++;
            maxcode = (1 << nbits) - 1;
        } else if (incount >= checkpoint) {
            checkpoint = incount + 10000;
            const long rat = ((long)incount << 8) / outbits;
            if (rat <= ratio)
                reset = true;
            else
                ratio = rat;
        }
        if (reset) {
            for (uint32_t k = lane; k < HSIZE; k += 64) tab[k] = EMPTY;
            __syncthreads();
            ratio = 0;
            incount = 0;
            outbits = 0;
            free_ent = 258;
            put(256);
            nbits = 9;
            maxcode = 511;
        }
        // the entry written by lane 0 is read by every lane from here on
        __builtin_amdgcn_s_waitcnt(0xc07f);
    };
    // bytes 1..n-1, a word of the row at a time
    uint32_t cur = uni(roww[0]) >> 8;
    for (int k = 1; k < n; k++) {
        if ((k & 3) == 0) cur = uni(roww[k >> 2]);
        code_byte((int)(cur & 0xffu));
        cur >>= 8;
    }
    put(ent);
    free_ent++;
    if (free_ent == 4094) {
        put(256);
        nbits = 9;
    } else if (free_ent > maxcode) {
        nbits++;
    }
    put(257);
    // the pending bits, zero-padded to a byte; the word may run up to three
    // bytes past the strip, inside its stride
    if (nacc > 0 && w0) out[nword] = __builtin_bswap32((uint32_t)(acc << (32 - nacc)));
    if (w0) lens[r] = 4 * nword + (nacc + 7) / 8;
}

// strip r to body + off[r] (off[r] = bytes of the strips before it)
__global__ __launch_bounds__(256) void k_pack(const uint8_t *__restrict__ strips, long stride,
                                              const int *__restrict__ lens, const long *__restrict__ off,
                                              uint8_t *__restrict__ body)
{
    const int r = blockIdx.x;
    const uint8_t *s = strips + (size_t)r * (size_t)stride;
    uint8_t *d = body + off[r];
    const int len = lens[r];
    for (int k = threadIdx.x; k < len; k += 256) d[k] = s[k];
}

}  // namespace

long wv_lzw_stride(int n) { return ((long)n + n / 2 + n / 512 + 16 + 15) / 16 * 16; }

int wv_launch_lzw(const uint8_t *rgb, int n, int rows, uint8_t *strips, long stride, int *lens, hipStream_t s)
{
    const size_t lds = sizeof(uint32_t) * HSIZE + (size_t)(n + 15) / 16 * 16;  // <= 48 KiB
    hipLaunchKernelGGL(k_lzw, dim3(rows), dim3(64), lds, s, rgb, n, strips, stride, lens);
    return (int)hipGetLastError();
}

int wv_launch_tiff_pack(const uint8_t *strips, long stride, const int *lens, const long *off, int rows, uint8_t *body,
                        hipStream_t s)
{
    hipLaunchKernelGGL(k_pack, dim3(rows), dim3(256), 0, s, strips, stride, lens, off, body);
    return (int)hipGetLastError();
}
