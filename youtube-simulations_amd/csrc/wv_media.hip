// wv_media.hip -- media tables built on the GPU.
//
// init_ior_2d's Gaussian-well models (sub_wave.c:12315-12555:
// IOR_PERIODIC_WELLS, IOR_RANDOM_WELLS, IOR_POISSON_WELLS, IOR_PPP_WELLS) sum
// one Gaussian per well at every cell: O(NX*NY*wells) exponentials, done by
// the reference on one host thread before the first frame.  Here one thread
// per cell walks the well list (wave-uniform loads, summed in the
// reference's order) in FP64; exp and tanh are the device library's.
#include "wv_device.h"

using namespace wv;

namespace {

__global__ void __launch_bounds__(256) k_wells(int nx, int ny, double xmin, double xmax, double ymin, double ymax,
                                               int nw, const double *__restrict__ xc, const double *__restrict__ yc,
                                               const double *__restrict__ h, double sigma, int use_tanh,
                                               double courant, double courantb, double *__restrict__ tcc)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y;
    if (j >= ny || i >= nx) return;
    // ij_to_xy, sub_wave.c:469-476
    const double x = xmin + ((double)i) * (xmax - xmin) / ((double)nx);
    const double y = ymin + ((double)j) * (ymax - ymin) / ((double)ny);
    double sum = 0.0;
    for (int n = 0; n < nw; n++) {
        const double dx = x - gc(xc)[n], dy = y - gc(yc)[n];
        const double r2 = dx * dx + dy * dy;
        const double e = exp(-r2 / sigma);
        sum = h ? sum + e * gc(h)[n] : sum + e;
    }
    if (use_tanh) sum = tanh(sum);
    gm(tcc)[(size_t)i * ny + j] = courant * sum + courantb * (1.0 - sum);
}

}  // namespace

int wv_launch_wells(int nx, int ny, const double box[4], int nw, const double *xc, const double *yc, const double *h,
                    double sigma, int use_tanh, double courant, double courantb, double *tcc, hipStream_t s)
{
    const dim3 block(256), grid((ny + 255) / 256, nx);
    hipLaunchKernelGGL(k_wells, grid, block, 0, s, nx, ny, box[0], box[1], box[2], box[3], nw, xc, yc, h, sigma,
                       use_tanh, courant, courantb, tcc);
    return (int)hipGetLastError();
}
