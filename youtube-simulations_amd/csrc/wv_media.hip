// wv_media.hip -- media tables and billiard masks built on the GPU.
//
// init_ior_2d's Gaussian-well models (sub_wave.c:12315-12555:
// IOR_PERIODIC_WELLS, IOR_RANDOM_WELLS, IOR_POISSON_WELLS, IOR_PPP_WELLS) sum
// one Gaussian per well at every cell: O(NX*NY*wells) exponentials, done by
// the reference on one host thread before the first frame.  Here one thread
// per cell walks the well list (wave-uniform loads, summed in the
// reference's order) in FP64; exp and tanh are the device library's.
#include <mutex>

#include "wv_device.h"

using namespace wv;

namespace {

// tcc and tc of cell (i, j) into the engine's per-cell tables (row stride
// ldy); tc = sqrt(tcc) where xy_in != 0 (nz plane), courantb where TWOSPEEDS,
// 0 elsewhere (sub_wave.c:12970-12981); the split form's binary32 pair too
__global__ void __launch_bounds__(256) k_wells(WvWells a)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y;
    if (j >= a.ny || i >= a.nx) return;
    // ij_to_xy, sub_wave.c:469-476
    const double x = a.xmin + ((double)i) * (a.xmax - a.xmin) / ((double)a.nx);
    const double y = a.ymin + ((double)j) * (a.ymax - a.ymin) / ((double)a.ny);
    double sum = 0.0;
    for (int n = 0; n < a.nw; n++) {
        const double dx = x - gc(a.xc)[n], dy = y - gc(a.yc)[n];
        const double r2 = dx * dx + dy * dy;
        const double e = exp(-r2 / a.sigma);
        sum = a.h ? sum + e * gc(a.h)[n] : sum + e;
    }
    if (a.use_tanh) sum = tanh(sum);
    const double tcc = a.courant * sum + a.courantb * (1.0 - sum);
    const size_t k = (size_t)i * a.ldy + j;
    gm(a.tcc)[k] = tcc;
    const bool nz = (gc(a.nz)[(size_t)i * a.ldw + (j >> 6)] >> (j & 63)) & 1u;
    gm(a.tc)[k] = nz ? sqrt(tcc) : (a.twospeeds ? a.courantb : 0.0);
    if (a.tcc_hi) {
        const double ts = tcc * a.split_weight;
        const float hi = (float)ts;
        gm(a.tcc_hi)[k] = hi;
        gm(a.tcc_lo)[k] = (float)(ts - (double)hi);
    }
}

// init_circular_wave (wave_common.c:89-111) and add_circular_wave (337-351)
// on the device state, one thread per cell, the reference's expressions in
// its order (exp, cos and sqrt are the device library's)
__global__ void __launch_bounds__(256) k_pulse(WvPulse a)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y;
    if (j >= a.ny || i >= a.nx) return;
    // ij_to_xy, sub_wave.c:469-476
    const double x = a.xmin + ((double)i) * (a.xmax - a.xmin) / ((double)a.nx);
    const double y = a.ymin + ((double)j) * (a.ymax - a.ymin) / ((double)a.ny);
    const double dist2 = (x - a.x0) * (x - a.x0) + (y - a.y0) * (y - a.y0);
    const bool drawn = a.twospeeds || ((gc(a.nz)[(size_t)i * a.ldw + (j >> 6)] >> (j & 63)) & 1u);
    const size_t k = (size_t)i * a.ldy + j;
    double term = 0.0;
    if (drawn) {
        const double pre = a.mode == 0 ? a.amp : a.amp * a.factor;
        term = pre * exp(-dist2 / a.variance) * cos(-sqrt(dist2) / a.wavelength);
    }
    if (a.phi) {
        if (a.mode == 0) {
            gm(a.phi)[k] = term;   // 0 on cells not drawn
            gm(a.psi)[k] = 0.0;
        } else if (drawn) {
            gm(a.phi)[k] = gc(a.phi)[k] + term;
        }
        return;
    }
    // split form: u = hi + lo, v = u - psi; both sets
    for (int s = 0; s < 2; s++) {
        if (a.mode == 0) {
            const float h = (float)term;
            gm(a.hi[s])[k] = h;
            gm(a.lo[s])[k] = (float)(term - (double)h);
            gm(a.v[s])[k] = (float)term;
        } else if (drawn) {
            const double u = ((double)gc(a.hi[s])[k] + (double)gc(a.lo[s])[k]) + term;
            const float h = (float)u;
            gm(a.hi[s])[k] = h;
            gm(a.lo[s])[k] = (float)(u - (double)h);
            gm(a.v[s])[k] = (float)((double)gc(a.v[s])[k] + term);   // psi unchanged
        }
    }
}

// xy_in_billiard_single_domain (sub_wave.c:5712) at the cell corner of
// ij_to_xy (sub_wave.c:469-476), for the domains of the plain workloads:
// D_SINAI (5774-5779), D_EXT_ELLIPSE (5744-5749) and D_CIRCLES (6216-6227:
// 0 inside any active circle, tested in list order).  The circle list is
// read with wave-uniform loads.
__global__ void __launch_bounds__(256) k_mask(WvMask a)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x, i = a.i0 + (int)blockIdx.y;
    if (j >= a.ny || i >= a.nx || i >= a.i0 + a.rows) return;
    const double x = a.xmin + ((double)i) * (a.xmax - a.xmin) / ((double)a.nx);
    const double y = a.ymin + ((double)j) * (a.ymax - a.ymin) / ((double)a.ny);
    short v = 1;
    if (a.kind == 0) {  // D_SINAI
        v = x * x + y * y > a.lambda * a.lambda;
    } else if (a.kind == 1) {  // D_EXT_ELLIPSE
        v = x * x / (a.lambda * a.lambda) + y * y / (a.mu * a.mu) > 1.0;
    } else {  // D_CIRCLES
        const int n = a.ncirc;
        for (int k = 0; k < n; k++) {
            if (gc(a.circ)[3 * n + k] == 0.0) continue;  // inactive
            const double x1 = gc(a.circ)[k], y1 = gc(a.circ)[n + k], r = gc(a.circ)[2 * n + k];
            const double r2 = r * r;
            if ((x - x1) * (x - x1) + (y - y1) * (y - y1) < r2) {
                v = 0;
                break;
            }
        }
    }
    gm(a.out)[(size_t)(i - a.i0) * a.ny + j] = v;
}

// Fixed device scratch of the allocation-free mask builder: the circle list
// (at most WV_MASK_SCRATCH_MAX_CIRCLES x 4 doubles) and then bands of mask
// rows.  One copy per device, alive until the process ends; every call on a
// device shares it.
static __device__ __align__(256) unsigned char scratch[WV_MASK_SCRATCH_BYTES];

}  // namespace

// device address of the scratch array, looked up once per device with that
// device current (the array's name is a host-side stand-in, not an address)
int wv_mask_scratch(int device, void **ptr)
{
    constexpr int MAXDEV = 64;
    static void *addr[MAXDEV];
    static std::mutex mu;
    if (device < 0 || device >= MAXDEV) return (int)hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(mu);
    if (!addr[device]) {
        int cur = 0;
        hipError_t err = hipGetDevice(&cur);
        if (err == hipSuccess && cur != device) err = hipSetDevice(device);
        void *p = nullptr;
        if (err == hipSuccess) err = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != device) (void)hipSetDevice(cur);
        if (err != hipSuccess) return (int)err;
        addr[device] = p;
    }
    *ptr = addr[device];
    return 0;
}

int wv_launch_pulse(const WvPulse &a, hipStream_t s)
{
    const dim3 block(256), grid((a.ny + 255) / 256, a.nx);
    hipLaunchKernelGGL(k_pulse, grid, block, 0, s, a);
    return (int)hipGetLastError();
}

int wv_launch_mask(const WvMask &a, hipStream_t s)
{
    const dim3 block(256), grid((a.ny + 255) / 256, a.rows);
    hipLaunchKernelGGL(k_mask, grid, block, 0, s, a);
    return (int)hipGetLastError();
}

int wv_launch_wells(const WvWells &a, hipStream_t s)
{
    const dim3 block(256), grid((a.ny + 255) / 256, a.nx);
    hipLaunchKernelGGL(k_wells, grid, block, 0, s, a);
    return (int)hipGetLastError();
}
