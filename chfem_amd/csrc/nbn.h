// nbn.h -- node-by-node helpers shared by kernels.hip and the batched
// solver (batch.hip): neighbourhood offsets, wave-uniform material loops and
// the thermal node product.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace chfem {

// compile-time copies for index arithmetic that must fold after unrolling
#define LDR(k) ((k) == 0 || (k) == 1 || (k) == 4 || (k) == 5 ? 1 : 0)
#define LDC(k) ((k) == 1 || (k) == 2 || (k) == 5 || (k) == 6 ? 1 : 0)
#define LDL(k) ((k) >= 4 ? 1 : 0)

// ---------------------------------------------------------------------------
// neighbourhood helpers
// ---------------------------------------------------------------------------
struct Nbh {
    uint32_t ro[3];   // row index r-1, r, r+1 (periodic)
    uint32_t co[3];   // column offset (c-1, c, c+1) * ny
    size_t lo[3];     // layer offset (l-1, l, l+1) * nx*ny
};

__device__ __forceinline__ void node_rcl(uint32_t i, int ny, int nx, int &r, int &c, int &l)
{
    uint32_t t = i / (uint32_t)ny;
    r = (int)(i - t * (uint32_t)ny);
    l = (int)(t / (uint32_t)nx);
    c = (int)(t - (uint32_t)l * (uint32_t)nx);
}

__device__ __forceinline__ Nbh make_nbh(int r, int c, int l, int ny, int nx, int nz)
{
    Nbh h;
    const size_t nxy = (size_t)nx * ny;
    h.ro[0] = r == 0 ? ny - 1 : r - 1; h.ro[1] = r; h.ro[2] = r == ny - 1 ? 0 : r + 1;
    h.co[0] = (uint32_t)(c == 0 ? nx - 1 : c - 1) * ny; h.co[1] = (uint32_t)c * ny;
    h.co[2] = (uint32_t)(c == nx - 1 ? 0 : c + 1) * ny;
    h.lo[0] = (size_t)(l == 0 ? nz - 1 : l - 1) * nxy; h.lo[1] = (size_t)l * nxy;
    h.lo[2] = (size_t)(l == nz - 1 ? 0 : l + 1) * nxy;
    return h;
}

// View of read-only device data in the constant address space: loads with a
// wave-uniform address become scalar (SMEM) loads.
template <typename T>
using cptr = const __attribute__((address_space(4))) T *;
template <typename T>
__device__ __forceinline__ cptr<T> as_const(const T *p) { return (cptr<T>)p; }

// Visit the distinct values of m across the active lanes of the wavefront;
// body(mu) runs with mu wave-uniform and only the lanes with m == mu active,
// so data indexed by mu is read with scalar loads.
template <typename F>
__device__ __forceinline__ void for_each_uniform(const int m, F &&body)
{
    unsigned long long pend = __ballot(1);
    while (pend) {
        const int lead = __ffsll((long long)pend) - 1;
        const int mu = __builtin_amdgcn_readlane(m, lead);
        if (m == mu) {
            // re-read as a wave-uniform value inside the branch: the compiler
            // may otherwise rewrite mu as the per-lane m (equal on the active
            // lanes) and address the local matrix with vector loads
            body(__builtin_amdgcn_readfirstlane(mu));
        }
        pend &= ~__ballot(m == mu);
    }
}

// thermal node-by-node product at node i: K layout [mat][8][8], keys packed
// per node (chfem/src/femhmg/cudapcg/kernels/matrix_free.cu:191-323)
template <typename T>
__device__ __forceinline__ T thermal_node(const T *__restrict__ v, const uint16_t *__restrict__ keys,
                                          const T *__restrict__ K, int ny, int nx, int nz, uint32_t i)
{
    int r, c, l;
    node_rcl(i, ny, nx, r, c, l);
    const Nbh h = make_nbh(r, c, l, ny, nx, nz);
    T u[3][3][3];
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
        for (int b = 0; b < 3; b++)
#pragma unroll
            for (int d = 0; d < 3; d++) u[a][b][d] = v[h.lo[a] + h.co[b] + h.ro[d]];
    const unsigned key = keys[i];
    T res = 0;
#pragma unroll
    for (int e = 0; e < 8; e++) {
        const int me = (key >> (2 * e)) & 3;
        for_each_uniform(me, [&](const int mu) {
            cptr<T> Km = as_const(K) + mu * 64 + e * 8;
#pragma unroll
            for (int j = 0; j < 8; j++)
                res = fma(Km[j], u[LDL(j) - LDL(e) + 1][LDC(j) - LDC(e) + 1][LDR(j) - LDR(e) + 1], res);
        });
    }
    return res;
}

}  // namespace chfem
