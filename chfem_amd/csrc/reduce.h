// reduce.h -- deterministic FP64 reductions shared by the kernel files:
// per-block tree sums written as block partials, summed in a fixed order by a
// separate single-block kernel (kernels.h: Reduce, Finalize,
// launch_reduce_finish).
#pragma once
#include <hip/hip_runtime.h>
#include "kernels.h"

namespace chfem {

// ---------------------------------------------------------------------------
// deterministic reductions
// ---------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// block sum; result valid in thread 0
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV])
{
    __shared__ double lds[NV][16];  // up to 1024 work-items
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < NV; k++) {
        double t = wave_sum(v[k]);
        if (lane == 0) lds[k][w] = t;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < NV; k++) {
            double t = 0;
            for (int i = 0; i < (int)(blockDim.x + 63) / 64; i++) t += lds[k][i];
            v[k] = t;
        }
    }
    __syncthreads();
}

// a guarded launch does nothing once its Krylov loop has converged
__device__ __forceinline__ bool stopped(const Reduce &red) { return red.guard && red.S[S_DONE] != 0.0; }

// Each block writes its NV partial sums to red.partials[k * nb + block]; the
// launcher then enqueues the single-block finishing kernel, which adds the
// partials in a fixed order.  No atomics or device-scope fences: on a
// multi-XCD part a per-block release fence writes back the whole L2 of the
// XCD, and same-address atomics from thousands of blocks serialise.
template <int NV>
__device__ __forceinline__ void store_partials(const Reduce &red, double (&v)[NV])
{
    // linear block id / count: the launching grid may be 1-D, 2-D or 3-D
    const unsigned nb = gridDim.x * gridDim.y * gridDim.z;
    const unsigned bid = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
    block_sum<NV>(v);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < NV; k++) red.partials[(size_t)k * nb + bid] = v[k];
    }
}

// element scale of a scalar density field (1 without one); same operation
// order as the reference (femhmg_thermal_3D.c:265)
__device__ __forceinline__ double sdf_scale(const Sdf &f, size_t e)
{
    return f.d ? (double)f.d[e] * (1.0 / 65535.0) * (f.fmax - f.fmin) + f.fmin : 1.0;
}

}  // namespace chfem
