// noalloc.hip -- the file-scope device scratch of hmgUpdateCNoAlloc and the
// two small kernels that move its seam table in and its column out.
#include <hip/hip_runtime.h>
#include <stdio.h>
#include "noalloc.h"

namespace chfem {

// One copy per GPU, held until the process ends.  Every call on a GPU shares
// it: calls that overlap on one GPU race on it.
static __device__ __align__(256) double scratch[NOALLOC_SCRATCH_DOUBLES];

static constexpr int NOALLOC_MAX_DEVICES = 64;
static double *scratch_addr[NOALLOC_MAX_DEVICES] = {};

double *noalloc_scratch()
{
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= NOALLOC_MAX_DEVICES) {
        fprintf(stderr, "chfem: no current device for the scratch array\n");
        return nullptr;
    }
    if (!scratch_addr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            fprintf(stderr, "chfem: hipGetSymbolAddress(scratch) failed on device %d: %s\n", dev,
                    hipGetErrorString(e));
            return nullptr;
        }
        scratch_addr[dev] = (double *)p;
    }
    return scratch_addr[dev];
}

struct SeamTable {
    double v[6 * NOALLOC_MAX_MAT];
};

__global__ void k_noalloc_seam(double *dst, SeamTable t, int n)
{
    for (int i = threadIdx.x; i < n; i += blockDim.x) dst[i] = t.v[i];
}

__global__ void k_noalloc_column(const double *sums, double *out, int n, double v, int divide)
{
    const int k = threadIdx.x;
    if (k < n) out[k] = divide ? sums[k] / v : sums[k] * v;
}

void launch_noalloc_seam(double *dst, const double *seam, int n, hipStream_t s)
{
    SeamTable t;
    for (int i = 0; i < n; i++) t.v[i] = seam[i];
    for (int i = n; i < 6 * NOALLOC_MAX_MAT; i++) t.v[i] = 0.0;
    k_noalloc_seam<<<1, 256, 0, s>>>(dst, t, n);
}

void launch_noalloc_column(const double *sums, double *out, int n, double v, int divide, hipStream_t s)
{
    k_noalloc_column<<<1, 64, 0, s>>>(sums, out, n, v, divide);
}

}  // namespace chfem
