// LZW coding of TIFF frames on the device: the coder behind writetiff
// (sub_wave.c:62-123), which stores RGB frames as LZW strips of one row each
// (TIFFTAG_ROWSPERSTRIP 1, COMPRESSION_LZW at every call site, sub_wave.c:210,
// 226).  Strips are coded independently, so each is one workgroup: a single
// wave whose lane 0 runs the (inherently serial) coder against a code table
// kept in LDS, after the whole wave has staged the row in LDS and cleared the
// table.  The code stream is libtiff's (version 4.2.0): TIFF 6.0 section 13
// LZW, codes MSB-first, the width growing one code early, Clear 256 at the
// start and EndOfInformation 257 at the end; a Clear at the current width
// when the next free code would be 4094, and every 10000 input bytes a
// compression-ratio check (input bytes * 256 / output bits) that resets the
// table when the ratio has not improved.  A second kernel packs the coded
// strips back to back in file order.
#include "wv_internal.h"

namespace {

constexpr int HBITS = 13;
constexpr int HSIZE = 1 << HBITS;        // table slots; at most 3836 live codes
constexpr uint32_t EMPTY = 0xffffffffu;  // no live entry has key 0xfffff (prefix 4095)

struct BitOut {
    uint8_t *out;  // strip base, 4-byte aligned
    uint32_t data = 0, word = 0;
    int bits = 0, nbits = 9, pos = 0;
    long outbits = 0;
    __device__ void byte(uint32_t b)
    {
        word |= (b & 0xffu) << (8 * (pos & 3));
        if ((pos & 3) == 3) {
            reinterpret_cast<uint32_t *>(out)[pos >> 2] = word;
            word = 0;
        }
        pos++;
    }
    __device__ void put(int code)
    {
        data = (data << nbits) | (uint32_t)code;  // at most 7 + 12 live bits
        bits += nbits;
        while (bits >= 8) {
            byte(data >> (bits - 8));
            bits -= 8;
        }
        outbits += nbits;
    }
    __device__ void finish()
    {
        if (bits > 0) byte(data << (8 - bits));
        if (pos & 3) reinterpret_cast<uint32_t *>(out)[pos >> 2] = word;
    }
};

__device__ __forceinline__ uint32_t slot_of(uint32_t key) { return (key * 2654435761u) >> (32 - HBITS); }

__device__ void clear_table(uint32_t *tab)
{
    uint4 *t4 = reinterpret_cast<uint4 *>(tab);
    const uint4 e = make_uint4(EMPTY, EMPTY, EMPTY, EMPTY);
    for (int k = 0; k < HSIZE / 4; k++) t4[k] = e;
}

// one strip per workgroup; n = 3*width bytes per row; strip r is the r-th row
// of rgb; out + r*stride receives its code bytes and lens[r] their count
__global__ __launch_bounds__(64) void k_lzw(const uint8_t *__restrict__ rgb, int n, uint8_t *__restrict__ strips,
                                            long stride, int *__restrict__ lens)
{
    extern __shared__ uint32_t sh[];
    uint32_t *tab = sh;
    uint8_t *row = reinterpret_cast<uint8_t *>(sh + HSIZE);
    const int r = blockIdx.x, lane = threadIdx.x;
    const uint8_t *src = rgb + (size_t)r * (size_t)n;
    for (int k = lane; k < n; k += 64) row[k] = src[k];
    for (int k = lane; k < HSIZE; k += 64) tab[k] = EMPTY;
    __syncthreads();
    if (lane != 0) return;

    BitOut w;
    w.out = strips + (size_t)r * (size_t)stride;
    int free_ent = 258, maxcode = 511;
    long incount = 1, checkpoint = 10000, ratio = 0;
    w.put(256);
    int ent = row[0];
    for (int k = 1; k < n; k++) {
        const int c = row[k];
        incount++;
        const uint32_t key = ((uint32_t)c << 12) | (uint32_t)ent;
        uint32_t h = slot_of(key);
        uint32_t e = tab[h];
        while (e != EMPTY && (e >> 12) != key) {
            h = (h + 1) & (HSIZE - 1);
            e = tab[h];
        }
        if (e != EMPTY) {
            ent = (int)(e & 0xfffu);
            continue;
        }
        w.put(ent);
        tab[h] = (key << 12) | (uint32_t)free_ent;
        free_ent++;
        ent = c;
        bool reset = false;
        if (free_ent == 4094) {
            reset = true;
        } else if (free_ent > maxcode) {
            w.nbits++;
            maxcode = (1 << w.nbits) - 1;
        } else if (incount >= checkpoint) {
            checkpoint = incount + 10000;
            const long rat = (incount << 8) / w.outbits;
            if (rat <= ratio)
                reset = true;
            else
                ratio = rat;
        }
        if (reset) {
            clear_table(tab);
            ratio = 0;
            incount = 0;
            w.outbits = 0;
            free_ent = 258;
            w.put(256);
            w.nbits = 9;
            maxcode = 511;
        }
    }
    w.put(ent);
    free_ent++;
    if (free_ent == 4094) {
        w.put(256);
        w.nbits = 9;
    } else if (free_ent > maxcode) {
        w.nbits++;
    }
    w.put(257);
    w.finish();
    lens[r] = w.pos;
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
    const size_t lds = sizeof(uint32_t) * HSIZE + (size_t)(n + 15) / 16 * 16;
    hipLaunchKernelGGL(k_lzw, dim3(rows), dim3(64), lds, s, rgb, n, strips, stride, lens);
    return (int)hipGetLastError();
}

int wv_launch_tiff_pack(const uint8_t *strips, long stride, const int *lens, const long *off, int rows, uint8_t *body,
                        hipStream_t s)
{
    hipLaunchKernelGGL(k_pack, dim3(rows), dim3(256), 0, s, strips, stride, lens, off, body);
    return (int)hipGetLastError();
}
