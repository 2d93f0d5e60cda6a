// reduce.h -- deterministic FP64 reductions shared by the kernel files:
// per-block tree sums, then the last block to finish sums the block partials
// in a fixed order and publishes the total (kernels.h: Reduce, Finalize).
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

__device__ __forceinline__ double load_coherent(const double *p)
{
    unsigned long long u = __hip_atomic_load((const unsigned long long *)p, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT);
    return __longlong_as_double((long long)u);
}

template <int NV>
__device__ __forceinline__ void reduce_finalize(const Reduce &red, double (&v)[NV])
{
    __shared__ int amLast;
    // linear block id / count: the launching grid may be 1-D, 2-D or 3-D
    const unsigned nb = gridDim.x * gridDim.y * gridDim.z;
    const unsigned bid = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
    block_sum<NV>(v);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < NV; k++) red.partials[(size_t)k * nb + bid] = v[k];
        __threadfence();
        unsigned prev = atomicAdd(red.counter, 1u);
        amLast = (prev == nb - 1);
    }
    __syncthreads();
    if (amLast) {
        __threadfence();
        double s[NV];
#pragma unroll
        for (int k = 0; k < NV; k++) {
            s[k] = 0;
            for (unsigned b = threadIdx.x; b < nb; b += blockDim.x) s[k] += load_coherent(red.partials + (size_t)k * nb + b);
        }
        block_sum<NV>(s);
        if (threadIdx.x == 0) {
            double *S = red.S;
            if (red.fin == FIN_ALPHA) {
                S[S_DQ] = s[0];
                S[S_ALPHA] = S[S_DELTA] / s[0];
            } else if (red.fin == FIN_DELTA) {
                S[S_DELTA_OLD] = S[S_DELTA];
                S[S_DELTA] = s[0];
            } else {
#pragma unroll
                for (int k = 0; k < NV; k++) S[red.slot + k] = s[k];
            }
            __hip_atomic_store(red.counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

}  // namespace chfem
