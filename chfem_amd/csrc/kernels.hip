// kernels.hip -- gfx950 kernels of the chfem matrix-free solve path.
//
// Operators (q = A v) for the periodic voxel mesh:
//   thermal / elastic node-by-node (NBN): one work-item per node, the 27-node
//     neighbourhood gathered into registers once, and for each of the node's 8
//     elements the rows of that element's local matrix applied to the element's
//     nodal values (chfem/src/femhmg/cudapcg/kernels/matrix_free.cu:191-323,
//     :416-459).  The material of an element varies per lane; the kernel walks
//     the distinct materials present in the wavefront (ballot + readlane), so
//     every local-matrix entry is a wave-uniform scalar load instead of a
//     divergent constant-memory read.
//   thermal / elastic element-by-element (EBE): one work-item per element,
//     K_e u_e scattered with FP64/FP32 atomics (matrix_free.cu:2187-2268,
//     :3164-3262).
//   Stokes DOF-number pore mapping: one work-item per pore/border DOF id
//     (matrix_free.cu:674-732, :927-977).
// Reductions are FP64 and deterministic: per-block tree sums, then the last
// block to finish sums the block partials in a fixed order.
#include <hip/hip_runtime.h>
#include <stdlib.h>
#include "kernels.h"
#include "nbn.h"
#include "reduce.h"

namespace chfem {

#define BS 256
#define MAX_RED_BLOCKS MAX_RED_GRID

__constant__ int cDR[8] = {1, 1, 0, 0, 1, 1, 0, 0};
__constant__ int cDC[8] = {0, 1, 1, 0, 0, 1, 1, 0};
__constant__ int cDL[8] = {0, 0, 0, 0, 1, 1, 1, 1};
int reduce_blocks(size_t n)
{
    size_t b = (n + BS - 1) / BS;
    if (b > MAX_RED_BLOCKS) b = MAX_RED_BLOCKS;
    if (b < 1) b = 1;
    return (int)b;
}

// ---------------------------------------------------------------------------
// thermal NBN: K layout [mat][8][8]
// ---------------------------------------------------------------------------
template <typename T, bool DOT>
__global__ void __launch_bounds__(BS) k_thermal_nbn(const T *__restrict__ v, T *__restrict__ q,
                                                    const uint16_t *__restrict__ keys, const T *__restrict__ K,
                                                    int ny, int nx, int nz, uint32_t N, Reduce red)
{
    if (DOT && stopped(red)) return;
    double acc = 0.0;
    for (uint32_t i = blockIdx.x * BS + threadIdx.x; i < N; i += gridDim.x * BS) {
        const T res = thermal_node<T>(v, keys, K, ny, nx, nz, i);
        q[i] = res;
        if (DOT) acc += (double)v[i] * (double)res;
    }
    if (DOT) {
        double vv[1] = {acc};
        store_partials<1>(red, vv);
    }
}

// ---------------------------------------------------------------------------
// elastic NBN: K layout [mat][24][24], DOFs node-major (x,y,z)
// ---------------------------------------------------------------------------
template <typename T, bool DOT>
__global__ void __launch_bounds__(BS) k_elastic_nbn(const T *__restrict__ v, T *__restrict__ q,
                                                    const uint16_t *__restrict__ keys, const T *__restrict__ K,
                                                    int ny, int nx, int nz, uint32_t N, Reduce red)
{
    if (DOT && stopped(red)) return;
    double acc = 0.0;
    for (uint32_t i = blockIdx.x * BS + threadIdx.x; i < N; i += gridDim.x * BS) {
        int r, c, l;
        node_rcl(i, ny, nx, r, c, l);
        const Nbh h = make_nbh(r, c, l, ny, nx, nz);
        T u[3][3][3][3];
#pragma unroll
        for (int a = 0; a < 3; a++)
#pragma unroll
            for (int b = 0; b < 3; b++)
#pragma unroll
                for (int d = 0; d < 3; d++) {
                    const T *p = v + 3 * (h.lo[a] + h.co[b] + h.ro[d]);
                    u[a][b][d][0] = p[0]; u[a][b][d][1] = p[1]; u[a][b][d][2] = p[2];
                }
        const unsigned key = keys[i];
        T rx = 0, ry = 0, rz = 0;
#pragma unroll
        for (int e = 0; e < 8; e++) {
            const int me = (key >> (2 * e)) & 3;
            for_each_uniform(me, [&](const int mu) {
                cptr<T> Km = as_const(K) + mu * 576 + e * 72;
#pragma unroll
                for (int j = 0; j < 8; j++) {
                    const T *uu = u[LDL(j) - LDL(e) + 1][LDC(j) - LDC(e) + 1][LDR(j) - LDR(e) + 1];
#pragma unroll
                    for (int b = 0; b < 3; b++) {
                        rx = fma(Km[3 * j + b], uu[b], rx);
                        ry = fma(Km[24 + 3 * j + b], uu[b], ry);
                        rz = fma(Km[48 + 3 * j + b], uu[b], rz);
                    }
                }
            });
        }
        T *o = q + 3 * (size_t)i;
        o[0] = rx; o[1] = ry; o[2] = rz;
        if (DOT) acc += (double)u[1][1][1][0] * rx + (double)u[1][1][1][1] * ry + (double)u[1][1][1][2] * rz;
    }
    if (DOT) {
        double vv[1] = {acc};
        store_partials<1>(red, vv);
    }
}

// ---------------------------------------------------------------------------
// EBE: one element per work-item; element map gives the material; the
// element's 8*ND outputs are added to q with atomics (q zeroed beforehand).
// ---------------------------------------------------------------------------
template <typename T, int ND>
__global__ void __launch_bounds__(BS) k_ebe(const T *__restrict__ v, T *__restrict__ q,
                                            const uint16_t *__restrict__ emap, const T *__restrict__ K,
                                            int ny, int nx, int nz, uint32_t N, Sdf sdf)
{
    constexpr int KE = 8 * ND;
    for (uint32_t i = blockIdx.x * BS + threadIdx.x; i < N; i += gridDim.x * BS) {
        int r, c, l;
        node_rcl(i, ny, nx, r, c, l);
        const Nbh h = make_nbh(r, c, l, ny, nx, nz);
        size_t idx[8];
#pragma unroll
        for (int k = 0; k < 8; k++) idx[k] = h.lo[1 + LDL(k)] + h.co[1 + LDC(k)] + h.ro[1 + LDR(k)];
        T ue[KE];
#pragma unroll
        for (int k = 0; k < 8; k++)
#pragma unroll
            for (int a = 0; a < ND; a++) ue[ND * k + a] = v[ND * idx[k] + a];
        const int me = emap[i];
        T fe[KE];
#pragma unroll
        for (int t = 0; t < KE; t++) fe[t] = 0;
        for_each_uniform(me, [&](const int mu) {
            cptr<T> Km = as_const(K) + mu * KE * KE;
#pragma unroll
            for (int t = 0; t < KE; t++) {
                T s = 0;
#pragma unroll
                for (int j = 0; j < KE; j++) s = fma(Km[t * KE + j], ue[j], s);
                fe[t] = s;
            }
        });
        const T sc = (T)sdf_scale(sdf, i);
#pragma unroll
        for (int k = 0; k < 8; k++)
#pragma unroll
            for (int a = 0; a < ND; a++) atomicAdd(q + ND * idx[k] + a, sc * fe[ND * k + a]);
    }
}

// ---------------------------------------------------------------------------
// Stokes, DOF-number pore mapping.  Vector layout [3*nvel velocities |
// nvel+nborder pressures]; K is the single 32x32 [K G; G^T S] matrix; pressure
// rows use -S (matrix_free.cu:674-732).
// ---------------------------------------------------------------------------
template <typename T, bool DOT>
__global__ void __launch_bounds__(BS) k_fluid(const T *__restrict__ v, T *__restrict__ q,
                                              const uint32_t *__restrict__ dof2node,
                                              const uint32_t *__restrict__ node2dof,
                                              const uint8_t *__restrict__ bkeys, const T *__restrict__ K,
                                              int ny, int nx, int nz, uint32_t nvel, uint32_t nborder, Reduce red)
{
    if (DOT && stopped(red)) return;
    const uint32_t ntot = nvel + nborder;
    double acc = 0.0;
    for (uint32_t t = blockIdx.x * BS + threadIdx.x; t < ntot; t += gridDim.x * BS) {
        const uint32_t node = dof2node[t];
        int r, c, l;
        node_rcl(node, ny, nx, r, c, l);
        const Nbh h = make_nbh(r, c, l, ny, nx, nz);
        const bool pore = t < nvel;
        const unsigned fkey = pore ? 255u : bkeys[t - nvel];
        T uv[3][3][3][3], up[3][3][3];
        bool hv[3][3][3];
#pragma unroll
        for (int a = 0; a < 3; a++)
#pragma unroll
            for (int b = 0; b < 3; b++)
#pragma unroll
                for (int d = 0; d < 3; d++) {
                    const uint32_t id = node2dof[h.lo[a] + h.co[b] + h.ro[d]];
                    const bool isv = id < nvel;
                    hv[a][b][d] = isv;
                    const size_t vi = isv ? 3 * (size_t)id : 0;
                    uv[a][b][d][0] = isv ? v[vi] : T(0);
                    uv[a][b][d][1] = isv ? v[vi + 1] : T(0);
                    uv[a][b][d][2] = isv ? v[vi + 2] : T(0);
                    up[a][b][d] = v[3 * (size_t)nvel + (id < ntot ? id : 0)];
                }
        T rx = 0, ry = 0, rz = 0, rp = 0;
#pragma unroll
        for (int e = 0; e < 8; e++) {
            if ((fkey >> e) & 1u) {
#pragma unroll
                for (int j = 0; j < 8; j++) {
                    const int A = LDL(j) - LDL(e) + 1, B = LDC(j) - LDC(e) + 1, D = LDR(j) - LDR(e) + 1;
                    const T *uu = uv[A][B][D];
                    const T p = up[A][B][D];
                    if (pore) {
                        const T *k0 = K + (3 * e) * 32, *k1 = k0 + 32, *k2 = k0 + 64;
                        rx += k0[3 * j] * uu[0] + k0[3 * j + 1] * uu[1] + k0[3 * j + 2] * uu[2] + k0[24 + j] * p;
                        ry += k1[3 * j] * uu[0] + k1[3 * j + 1] * uu[1] + k1[3 * j + 2] * uu[2] + k1[24 + j] * p;
                        rz += k2[3 * j] * uu[0] + k2[3 * j + 1] * uu[1] + k2[3 * j + 2] * uu[2] + k2[24 + j] * p;
                    }
                    const T *kp = K + (24 + e) * 32;
                    rp += kp[3 * j] * uu[0] + kp[3 * j + 1] * uu[1] + kp[3 * j + 2] * uu[2] - kp[24 + j] * p;
                }
            }
        }
        (void)hv;
        if (pore) {
            T *o = q + 3 * (size_t)t;
            o[0] = rx; o[1] = ry; o[2] = rz;
            if (DOT) acc += (double)uv[1][1][1][0] * rx + (double)uv[1][1][1][1] * ry + (double)uv[1][1][1][2] * rz;
        }
        q[3 * (size_t)nvel + t] = rp;
        if (DOT) acc += (double)up[1][1][1] * rp;
    }
    if (DOT) {
        double vv[1] = {acc};
        store_partials<1>(red, vv);
    }
}

// ---------------------------------------------------------------------------
// vector kernels
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(BS) k_fill(T *x, T val, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)BS + threadIdx.x; i < n; i += (size_t)gridDim.x * BS) x[i] = val;
}

template <typename Tin, typename Tout>
__global__ void __launch_bounds__(BS) k_convert(const Tin *in, Tout *out, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)BS + threadIdx.x; i < n; i += (size_t)gridDim.x * BS) out[i] = (Tout)in[i];
}

// Streaming vector kernels.  Arrays are processed as 16-byte vectors
// (2 doubles / 4 floats); block b owns one contiguous range of vectors and its
// work-items walk it with a stride of one block, so every wavefront reads and
// writes whole, consecutive 1-KiB segments.  The ragged tail (n not a
// multiple of the vector width) is handled by the last block entry by entry.
template <typename T>
struct V16 {
    static constexpr int W = 16 / sizeof(T);
    T e[W];
};
template <typename T>
__device__ __forceinline__ V16<T> ld16(const T *p)
{
    V16<T> r;
    const uint4 q = *reinterpret_cast<const uint4 *>(p);
    __builtin_memcpy(&r, &q, 16);
    return r;
}
template <typename T>
__device__ __forceinline__ void st16(T *p, const V16<T> &r)
{
    uint4 q;
    __builtin_memcpy(&q, &r, 16);
    *reinterpret_cast<uint4 *>(p) = q;
}

// range of 16-byte vectors owned by this block; tail entries [nv*W, n)
#define STREAM_RANGE(T, n)                                                                        \
    constexpr int W_ = V16<T>::W;                                                                 \
    const size_t nv_ = (n) / W_;                                                                  \
    const size_t per_ = (nv_ + gridDim.x - 1) / gridDim.x;                                        \
    const size_t vb_ = (size_t)blockIdx.x * per_;                                                 \
    const size_t ve_ = vb_ + per_ < nv_ ? vb_ + per_ : nv_;
#define FOR_VECTORS for (size_t j_ = vb_ + threadIdx.x; j_ < ve_; j_ += BS)
#define FOR_TAIL(n)                                                                               \
    if (blockIdx.x == gridDim.x - 1)                                                              \
        for (size_t i = nv_ * W_ + threadIdx.x; i < (n); i += BS)

template <typename T, bool RED>
__global__ void __launch_bounds__(BS) k_precond(const T *__restrict__ dinv, const T *v, T *out, size_t n,
                                                Reduce red)
{
    double acc = 0;
    STREAM_RANGE(T, n)
    FOR_VECTORS
    {
        const V16<T> a = ld16(v + j_ * W_), m = ld16(dinv + j_ * W_);
        V16<T> o;
#pragma unroll
        for (int k = 0; k < W_; k++) {
            o.e[k] = m.e[k] * a.e[k];
            if (RED) acc += (double)a.e[k] * (double)o.e[k];
        }
        if (out) st16(out + j_ * W_, o);
    }
    FOR_TAIL(n)
    {
        const T o = dinv[i] * v[i];
        if (RED) acc += (double)v[i] * (double)o;
        if (out) out[i] = o;
    }
    if (RED) {
        double vv[1] = {acc};
        store_partials<1>(red, vv);
    }
}

template <typename T>
__global__ void __launch_bounds__(BS) k_dot(const T *__restrict__ a, const T *__restrict__ b, size_t n, Reduce red)
{
    if (stopped(red)) return;
    double acc = 0;
    STREAM_RANGE(T, n)
    FOR_VECTORS
    {
        const V16<T> x = ld16(a + j_ * W_), y = ld16(b + j_ * W_);
#pragma unroll
        for (int k = 0; k < W_; k++) acc += (double)x.e[k] * (double)y.e[k];
    }
    FOR_TAIL(n) acc += (double)a[i] * (double)b[i];
    double vv[1] = {acc};
    store_partials<1>(red, vv);
}

template <typename T>
__global__ void __launch_bounds__(BS) k_pcg_dir(T *__restrict__ d, const T *__restrict__ z,
                                                const double *__restrict__ S, int first, size_t n, int guard)
{
    if (guard && S[S_DONE] != 0.0) return;
    const T beta = first ? T(0) : (T)(S[S_DELTA] / S[S_DELTA_OLD]);
    STREAM_RANGE(T, n)
    FOR_VECTORS
    {
        const V16<T> zz = ld16(z + j_ * W_);
        V16<T> dd;
        if (first) dd = zz;
        else {
            dd = ld16(d + j_ * W_);
#pragma unroll
            for (int k = 0; k < W_; k++) dd.e[k] = zz.e[k] + beta * dd.e[k];
        }
        st16(d + j_ * W_, dd);
    }
    FOR_TAIL(n) d[i] = first ? z[i] : z[i] + beta * d[i];
}

// x += a d ; r -= a q ; z = dinv r (z holds q = A d on entry) ; reduce r.z
template <typename T>
__global__ void __launch_bounds__(BS) k_pcg_update(T *__restrict__ x, T *__restrict__ r, const T *__restrict__ d,
                                                   T *__restrict__ z, const T *__restrict__ dinv, size_t n,
                                                   Reduce red)
{
    if (stopped(red)) return;
    const T a = (T)red.S[S_ALPHA];
    double acc = 0;
    STREAM_RANGE(T, n)
    FOR_VECTORS
    {
        const size_t o = j_ * W_;
        V16<T> xv = ld16(x + o), rv = ld16(r + o);
        const V16<T> dv = ld16(d + o), qv = ld16(z + o), mv = ld16(dinv + o);
        V16<T> zv;
#pragma unroll
        for (int k = 0; k < W_; k++) {
            xv.e[k] += a * dv.e[k];
            rv.e[k] -= a * qv.e[k];
            zv.e[k] = mv.e[k] * rv.e[k];
            acc += (double)rv.e[k] * (double)zv.e[k];
        }
        st16(x + o, xv);
        st16(r + o, rv);
        st16(z + o, zv);
    }
    FOR_TAIL(n)
    {
        x[i] += a * d[i];
        const T rn = r[i] - a * z[i];
        r[i] = rn;
        const T zn = dinv[i] * rn;
        z[i] = zn;
        acc += (double)rn * (double)zn;
    }
    double vv[1] = {acc};
    store_partials<1>(red, vv);
}

template <typename T>
__global__ void __launch_bounds__(BS) k_sub_precond(T *__restrict__ r, const T *__restrict__ q,
                                                    const T *__restrict__ dinv, T *__restrict__ z, size_t n,
                                                    Reduce red)
{
    double acc = 0;
    STREAM_RANGE(T, n)
    FOR_VECTORS
    {
        const size_t o = j_ * W_;
        V16<T> rv = ld16(r + o);
        const V16<T> qv = ld16(q + o), mv = ld16(dinv + o);
        V16<T> zv;
#pragma unroll
        for (int k = 0; k < W_; k++) {
            rv.e[k] -= qv.e[k];
            zv.e[k] = mv.e[k] * rv.e[k];
            acc += (double)rv.e[k] * (double)zv.e[k];
        }
        st16(r + o, rv);
        st16(z + o, zv);
    }
    FOR_TAIL(n)
    {
        const T rn = r[i] - q[i];
        r[i] = rn;
        const T zn = dinv[i] * rn;
        z[i] = zn;
        acc += (double)rn * (double)zn;
    }
    double vv[1] = {acc};
    store_partials<1>(red, vv);
}

// MINRES: d = rt + b dold (written to dnew, which may be dold) ; q = s + b q ;
// reduce q.(dinv q)
template <typename T>
__global__ void __launch_bounds__(BS) k_minres_dir(const T *dold, T *dnew, T *__restrict__ q,
                                                   const T *__restrict__ rt, const T *__restrict__ sv,
                                                   const T *__restrict__ dinv, const double *__restrict__ S, int first,
                                                   size_t n, Reduce red)
{
    if (stopped(red)) return;
    const T b = first ? T(0) : (T)(S[S_DELTA] / S[S_DELTA_OLD]);
    double acc = 0;
    STREAM_RANGE(T, n)
    FOR_VECTORS
    {
        const size_t o = j_ * W_;
        const V16<T> rv = ld16(rt + o), s_ = ld16(sv + o), mv = ld16(dinv + o);
        V16<T> dv = ld16(dold + o), qv = ld16(q + o);
#pragma unroll
        for (int k = 0; k < W_; k++) {
            dv.e[k] = first ? rv.e[k] : rv.e[k] + b * dv.e[k];
            qv.e[k] = first ? s_.e[k] : s_.e[k] + b * qv.e[k];
            acc += (double)qv.e[k] * (double)(mv.e[k] * qv.e[k]);
        }
        st16(dnew + o, dv);
        st16(q + o, qv);
    }
    FOR_TAIL(n)
    {
        const T dn = first ? rt[i] : rt[i] + b * dold[i];
        const T qn = first ? sv[i] : sv[i] + b * q[i];
        dnew[i] = dn;
        q[i] = qn;
        acc += (double)qn * (double)(dinv[i] * qn);
    }
    double vv[1] = {acc};
    store_partials<1>(red, vv);
}

// MINRES: rn = rt/dinv - a q ; rt = dinv rn ; (XU) x += a d ; reduce rn.rt.
// Without XU the iterate is left to the direction ring's flush.
template <typename T, bool XU>
__global__ void __launch_bounds__(BS) k_minres_update(T *__restrict__ x, T *__restrict__ rt, const T *__restrict__ d,
                                                      const T *__restrict__ q, const T *__restrict__ dinv, size_t n,
                                                      Reduce red)
{
    if (stopped(red)) return;
    const T a = (T)red.S[S_ALPHA];
    double acc = 0;
    STREAM_RANGE(T, n)
    FOR_VECTORS
    {
        const size_t o = j_ * W_;
        V16<T> rv = ld16(rt + o);
        const V16<T> qv = ld16(q + o), mv = ld16(dinv + o);
#pragma unroll
        for (int k = 0; k < W_; k++) {
            const T rn = rv.e[k] / mv.e[k] - a * qv.e[k];
            rv.e[k] = mv.e[k] * rn;
            acc += (double)rn * (double)rv.e[k];
        }
        if constexpr (XU) {
            V16<T> xv = ld16(x + o);
            const V16<T> dv = ld16(d + o);
#pragma unroll
            for (int k = 0; k < W_; k++) xv.e[k] += a * dv.e[k];
            st16(x + o, xv);
        }
        st16(rt + o, rv);
    }
    FOR_TAIL(n)
    {
        const T rn = rt[i] / dinv[i] - a * q[i];
        const T rtn = dinv[i] * rn;
        rt[i] = rtn;
        if constexpr (XU) x[i] += a * d[i];
        acc += (double)rn * (double)rtn;
    }
    double vv[1] = {acc};
    store_partials<1>(red, vv);
}

// ---------------------------------------------------------------------------
// Key-Jacobi vector kernels (thermal/elastic NBN): the Jacobi diagonal is
// taken from the node's packed key instead of a stored vector, as the
// reference's preconditioner rebuilds it from the key (matrix_free.cu:167-188,
// 388-408), and the preconditioned residual z = dinv r is never stored.
// kinv[key * ND + a] (one entry per possible 16-bit key) holds
// 1 / sum_e K[m_e][e][a] summed over e = 0..7 in FP64 and rounded to T, the
// same value k_jacobi_keys computes; a two-phase image touches 256 entries.
// Work is done in groups of 4 nodes (4*ND entries = whole 16-byte vectors).
// ---------------------------------------------------------------------------
template <typename T, int ND>
struct Group {
    static constexpr int E = 4 * ND;
    static constexpr int NV = E * (int)sizeof(T) / 16;
    T v[E];
};
template <typename T, int ND>
__device__ __forceinline__ void gload(const T *p, Group<T, ND> &g)
{
    uint4 t[Group<T, ND>::NV];
    const uint4 *q = reinterpret_cast<const uint4 *>(p);
#pragma unroll
    for (int i = 0; i < Group<T, ND>::NV; i++) t[i] = q[i];
    __builtin_memcpy(g.v, t, sizeof(t));
}
template <typename T, int ND>
__device__ __forceinline__ void gstore(T *p, const Group<T, ND> &g)
{
    uint4 t[Group<T, ND>::NV];
    __builtin_memcpy(t, g.v, sizeof(t));
    uint4 *q = reinterpret_cast<uint4 *>(p);
#pragma unroll
    for (int i = 0; i < Group<T, ND>::NV; i++) q[i] = t[i];
}

// dinv of the ND DOFs of a node from its packed key: kinv[key * ND + a]
template <typename T, int ND>
__device__ __forceinline__ void node_dinv(const T *__restrict__ kinv, unsigned key, T *dv)
{
#pragma unroll
    for (int a = 0; a < ND; a++) dv[a] = kinv[(size_t)key * ND + a];
}

template <typename T, int ND>
__device__ __forceinline__ void group_dinv(const T *__restrict__ kinv, const uint16_t *keys, size_t g,
                                           Group<T, ND> &dv)
{
    const uint2 k = *reinterpret_cast<const uint2 *>(keys + 4 * g);
    node_dinv<T, ND>(kinv, k.x & 0xffffu, dv.v);
    node_dinv<T, ND>(kinv, k.x >> 16, dv.v + ND);
    node_dinv<T, ND>(kinv, k.y & 0xffffu, dv.v + 2 * ND);
    node_dinv<T, ND>(kinv, k.y >> 16, dv.v + 3 * ND);
}

// block b owns groups [gb_, ge_); the last block also takes the N % 4 tail nodes
#define GROUP_RANGE(N)                                                                                \
    const size_t ng_ = (N) / 4;                                                                       \
    const size_t per_ = (ng_ + gridDim.x - 1) / gridDim.x;                                            \
    const size_t gb_ = (size_t)blockIdx.x * per_;                                                     \
    const size_t ge_ = gb_ + per_ < ng_ ? gb_ + per_ : ng_;
#define FOR_GROUPS for (size_t g_ = gb_ + threadIdx.x; g_ < ge_; g_ += BS)
#define FOR_TAIL_NODES(N)                                                                             \
    if (blockIdx.x == gridDim.x - 1)                                                                  \
        for (size_t i = ng_ * 4 + threadIdx.x; i < (N); i += BS)

// out = dinv v ; reduce v.out
template <typename T, int ND>
__global__ void __launch_bounds__(BS) k_kj_precond(const uint16_t *__restrict__ keys, const T *__restrict__ kinv, const T *__restrict__ v, T *__restrict__ out,
                                                   size_t N, Reduce red)
{
    double acc = 0;
    GROUP_RANGE(N)
    FOR_GROUPS
    {
        Group<T, ND> a, m, o;
        gload<T, ND>(v + g_ * 4 * ND, a);
        group_dinv<T, ND>(kinv, keys, g_, m);
#pragma unroll
        for (int k = 0; k < 4 * ND; k++) {
            o.v[k] = m.v[k] * a.v[k];
            acc += (double)a.v[k] * (double)o.v[k];
        }
        gstore<T, ND>(out + g_ * 4 * ND, o);
    }
    FOR_TAIL_NODES(N)
    {
        T dv[ND];
        node_dinv<T, ND>(kinv, keys[i], dv);
#pragma unroll
        for (int a = 0; a < ND; a++) {
            const T o = dv[a] * v[ND * i + a];
            out[ND * i + a] = o;
            acc += (double)v[ND * i + a] * (double)o;
        }
    }
    double vv[1] = {acc};
    store_partials<1>(red, vv);
}

// d = dinv r + beta d
template <typename T, int ND>
__global__ void __launch_bounds__(BS) k_kj_pcg_dir(T *__restrict__ d, const T *__restrict__ r,
                                                   const uint16_t *__restrict__ keys, const T *__restrict__ kinv, const double *__restrict__ S, size_t N,
                                                   int guard)
{
    if (guard && S[S_DONE] != 0.0) return;
    const T beta = (T)(S[S_DELTA] / S[S_DELTA_OLD]);
    GROUP_RANGE(N)
    FOR_GROUPS
    {
        Group<T, ND> rv, dv, m;
        gload<T, ND>(r + g_ * 4 * ND, rv);
        gload<T, ND>(d + g_ * 4 * ND, dv);
        group_dinv<T, ND>(kinv, keys, g_, m);
#pragma unroll
        for (int k = 0; k < 4 * ND; k++) dv.v[k] = m.v[k] * rv.v[k] + beta * dv.v[k];
        gstore<T, ND>(d + g_ * 4 * ND, dv);
    }
    FOR_TAIL_NODES(N)
    {
        T dvn[ND];
        node_dinv<T, ND>(kinv, keys[i], dvn);
#pragma unroll
        for (int a = 0; a < ND; a++) d[ND * i + a] = dvn[a] * r[ND * i + a] + beta * d[ND * i + a];
    }
}

// x += a d ; r -= a q ; reduce r.(dinv r)
template <typename T, int ND, bool WITHX = true>
__global__ void __launch_bounds__(BS) k_kj_pcg_update(T *__restrict__ x, T *__restrict__ r, const T *__restrict__ d,
                                                      const T *__restrict__ q, const uint16_t *__restrict__ keys,
                                                      const T *__restrict__ kinv, size_t N,
                                                      Reduce red)
{
    if (stopped(red)) return;
    const T a = (T)red.S[S_ALPHA];
    double acc = 0;
    GROUP_RANGE(N)
    FOR_GROUPS
    {
        const size_t o = g_ * 4 * ND;
        Group<T, ND> xv, rv, dv, qv, m;
        if (WITHX) {
            gload<T, ND>(x + o, xv);
            gload<T, ND>(d + o, dv);
        }
        gload<T, ND>(r + o, rv);
        gload<T, ND>(q + o, qv);
        group_dinv<T, ND>(kinv, keys, g_, m);
#pragma unroll
        for (int k = 0; k < 4 * ND; k++) {
            if (WITHX) xv.v[k] += a * dv.v[k];
            rv.v[k] -= a * qv.v[k];
            acc += (double)rv.v[k] * (double)(m.v[k] * rv.v[k]);
        }
        if (WITHX) gstore<T, ND>(x + o, xv);
        gstore<T, ND>(r + o, rv);
    }
    FOR_TAIL_NODES(N)
    {
        T dvn[ND];
        node_dinv<T, ND>(kinv, keys[i], dvn);
#pragma unroll
        for (int c = 0; c < ND; c++) {
            const size_t j = ND * i + c;
            if (WITHX) x[j] += a * d[j];
            const T rn = r[j] - a * q[j];
            r[j] = rn;
            acc += (double)rn * (double)(dvn[c] * rn);
        }
    }
    double vv[1] = {acc};
    store_partials<1>(red, vv);
}

// PCG on the preconditioned residual z = dinv r (the CG loop of the fused
// elastic operator): z -= a dinv q ; reduce r.z = z.(diag z).  kinv holds
// dinv per key, then diag per key at offset 65536 * ND.
template <typename T, int ND>
__global__ void __launch_bounds__(BS) k_kj_zupdate(T *__restrict__ z, const T *__restrict__ q,
                                                   const uint16_t *__restrict__ keys, const T *__restrict__ kinv,
                                                   size_t N, Reduce red)
{
    if (stopped(red)) return;
    const T a = (T)red.S[S_ALPHA];
    const T *kdiag = kinv + (size_t)65536 * ND;
    double acc = 0;
    GROUP_RANGE(N)
    FOR_GROUPS
    {
        const size_t o = g_ * 4 * ND;
        Group<T, ND> zv, qv, m, dg;
        gload<T, ND>(z + o, zv);
        gload<T, ND>(q + o, qv);
        group_dinv<T, ND>(kinv, keys, g_, m);
        group_dinv<T, ND>(kdiag, keys, g_, dg);
#pragma unroll
        for (int k = 0; k < 4 * ND; k++) {
            zv.v[k] -= a * (m.v[k] * qv.v[k]);
            acc += (double)zv.v[k] * (double)(dg.v[k] * zv.v[k]);
        }
        gstore<T, ND>(z + o, zv);
    }
    FOR_TAIL_NODES(N)
    {
        T mv[ND], dv[ND];
        node_dinv<T, ND>(kinv, keys[i], mv);
        node_dinv<T, ND>(kdiag, keys[i], dv);
#pragma unroll
        for (int c = 0; c < ND; c++) {
            const size_t j = ND * i + c;
            const T zn = z[j] - a * (mv[c] * q[j]);
            z[j] = zn;
            acc += (double)zn * (double)(dv[c] * zn);
        }
    }
    double vv[1] = {acc};
    store_partials<1>(red, vv);
}

// r <- z = dinv r in place ; reduce r.z
template <typename T, int ND>
__global__ void __launch_bounds__(BS) k_kj_to_z(T *r, const uint16_t *__restrict__ keys,
                                                const T *__restrict__ kinv, size_t N, Reduce red)
{
    double acc = 0;
    GROUP_RANGE(N)
    FOR_GROUPS
    {
        const size_t o = g_ * 4 * ND;
        Group<T, ND> rv, m;
        gload<T, ND>(r + o, rv);
        group_dinv<T, ND>(kinv, keys, g_, m);
#pragma unroll
        for (int k = 0; k < 4 * ND; k++) {
            const T zk = m.v[k] * rv.v[k];
            acc += (double)rv.v[k] * (double)zk;
            rv.v[k] = zk;
        }
        gstore<T, ND>(r + o, rv);
    }
    FOR_TAIL_NODES(N)
    {
        T mv[ND];
        node_dinv<T, ND>(kinv, keys[i], mv);
#pragma unroll
        for (int c = 0; c < ND; c++) {
            const T zk = mv[c] * r[ND * i + c];
            acc += (double)r[ND * i + c] * (double)zk;
            r[ND * i + c] = zk;
        }
    }
    double vv[1] = {acc};
    store_partials<1>(red, vv);
}

// r -= q ; z = dinv r ; reduce r.z
template <typename T, int ND>
__global__ void __launch_bounds__(BS) k_kj_sub_precond(T *__restrict__ r, const T *__restrict__ q, T *__restrict__ z,
                                                       const uint16_t *__restrict__ keys,
                                                       const T *__restrict__ kinv, size_t N,
                                                       Reduce red)
{
    double acc = 0;
    GROUP_RANGE(N)
    FOR_GROUPS
    {
        const size_t o = g_ * 4 * ND;
        Group<T, ND> rv, qv, m, zv;
        gload<T, ND>(r + o, rv);
        gload<T, ND>(q + o, qv);
        group_dinv<T, ND>(kinv, keys, g_, m);
#pragma unroll
        for (int k = 0; k < 4 * ND; k++) {
            rv.v[k] -= qv.v[k];
            zv.v[k] = m.v[k] * rv.v[k];
            acc += (double)rv.v[k] * (double)zv.v[k];
        }
        gstore<T, ND>(r + o, rv);
        gstore<T, ND>(z + o, zv);
    }
    FOR_TAIL_NODES(N)
    {
        T dvn[ND];
        node_dinv<T, ND>(kinv, keys[i], dvn);
#pragma unroll
        for (int c = 0; c < ND; c++) {
            const size_t j = ND * i + c;
            const T rn = r[j] - q[j];
            r[j] = rn;
            const T zn = dvn[c] * rn;
            z[j] = zn;
            acc += (double)rn * (double)zn;
        }
    }
    double vv[1] = {acc};
    store_partials<1>(red, vv);
}

template <typename T>
__global__ void __launch_bounds__(BS) k_add(T *y, const T *__restrict__ x, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * BS + threadIdx.x; i < n; i += (size_t)gridDim.x * BS) y[i] += x[i];
}

__global__ void __launch_bounds__(BS) k_add_correction(double *x, const float *__restrict__ e, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)BS + threadIdx.x; i < n; i += (size_t)gridDim.x * BS) x[i] += (double)e[i];
}

// ---------------------------------------------------------------------------
// Jacobi diagonal
// ---------------------------------------------------------------------------
template <typename T, int ND>
__global__ void __launch_bounds__(BS) k_jacobi_keys(const uint16_t *__restrict__ keys, const T *__restrict__ K,
                                                    T *dinv, uint32_t N, int ebe, int ny, int nx, int nz, Sdf sdf)
{
    constexpr int KE = 8 * ND;
    for (uint32_t i = blockIdx.x * BS + threadIdx.x; i < N; i += gridDim.x * BS) {
        double dsum[ND];
#pragma unroll
        for (int a = 0; a < ND; a++) dsum[a] = 0;
        unsigned key = 0;
        double esc[8] = {1, 1, 1, 1, 1, 1, 1, 1};
        if (ebe) {
            // EBE keeps the element map: the material of the element with this
            // node as local node e sits at the element origin.
            int r, c, l;
            node_rcl(i, ny, nx, r, c, l);
            for (int e = 0; e < 8; e++) {
                int re = r - cDR[e], ce = c - cDC[e], le = l - cDL[e];
                re += re < 0 ? ny : 0; ce += ce < 0 ? nx : 0; le += le < 0 ? nz : 0;
                const size_t el = (size_t)re + (size_t)ce * ny + (size_t)le * nx * ny;
                key |= (unsigned)(keys[el] & 3u) << (2 * e);
                esc[e] = sdf_scale(sdf, el);
            }
        } else key = keys[i];
        for (int e = 0; e < 8; e++) {
            const int m = (key >> (2 * e)) & 3;
#pragma unroll
            for (int a = 0; a < ND; a++) {
                const double kd = (double)K[m * KE * KE + (ND * e + a) * KE + ND * e + a];
                dsum[a] += sdf.d ? esc[e] * kd : kd;
            }
        }
#pragma unroll
        for (int a = 0; a < ND; a++) dinv[(size_t)ND * i + a] = (T)(1.0 / dsum[a]);
    }
}

template <typename T>
__global__ void __launch_bounds__(BS) k_jacobi_fluid(const uint8_t *__restrict__ bkeys, const T *__restrict__ K,
                                                     T *dinv, uint32_t nvel, uint32_t nborder)
{
    double dv = 0, dp = 0;
    for (int e = 0; e < 8; e++) {
        dv += (double)K[(3 * e) * 32 + 3 * e];
        dp -= (double)K[(24 + e) * 32 + 24 + e];
    }
    const uint32_t ntot = nvel + nborder;
    for (uint32_t t = blockIdx.x * BS + threadIdx.x; t < ntot; t += gridDim.x * BS) {
        if (t < nvel) {
            dinv[3 * (size_t)t] = (T)(1.0 / dv);
            dinv[3 * (size_t)t + 1] = (T)(1.0 / dv);
            dinv[3 * (size_t)t + 2] = (T)(1.0 / dv);
            dinv[3 * (size_t)nvel + t] = (T)(1.0 / dp);
        } else {
            const unsigned fk = bkeys[t - nvel];
            double d = 0;
            for (int e = 0; e < 8; e++)
                if ((fk >> e) & 1u) d -= (double)K[(24 + e) * 32 + 24 + e];
            dinv[3 * (size_t)nvel + t] = (T)(1.0 / d);
        }
    }
}

// ---------------------------------------------------------------------------
// right-hand side of a seam-jump load case (femhmg_thermal_3D.c:122-236,
// femhmg_elastic_3D.c:130-462), gathered node by node; Stokes body force
// (femhmg_fluid_3D.c:112-181).
// ---------------------------------------------------------------------------
template <int ND>
__global__ void __launch_bounds__(BS) k_rhs_seam(double *b, const uint16_t *__restrict__ keys,
                                                 const double *__restrict__ K, int ebe, int ny, int nx, int nz,
                                                 uint32_t N, LoadCase lc, Sdf sdf)
{
    constexpr int KE = 8 * ND;
    for (uint32_t i = blockIdx.x * BS + threadIdx.x; i < N; i += gridDim.x * BS) {
        int r, c, l;
        node_rcl(i, ny, nx, r, c, l);
        double acc[ND];
#pragma unroll
        for (int a = 0; a < ND; a++) acc[a] = 0;
        const bool plane = lc.axis == 0 ? (c == nx - 1 || c == 0) : (lc.axis == 1 ? (r == 0 || r == 1 % ny) :
                                                                     (l == 0 || l == 1 % nz));
        if (plane) {
            unsigned key = ebe ? 0u : keys[i];
            for (int e = 0; e < 8; e++) {
                int re = r - cDR[e], ce = c - cDC[e], le = l - cDL[e];
                re += re < 0 ? ny : 0; ce += ce < 0 ? nx : 0; le += le < 0 ? nz : 0;
                const bool seam = lc.axis == 0 ? ce == nx - 1 : (lc.axis == 1 ? re == 0 : le == 0);
                if (!seam) continue;
                const size_t el = (size_t)re + (size_t)ce * ny + (size_t)le * nx * ny;
                const int m = ebe ? (int)keys[el] : (int)((key >> (2 * e)) & 3u);
                const double *Km = K + (size_t)m * KE * KE;
                const double sc = sdf_scale(sdf, el);
#pragma unroll
                for (int a = 0; a < ND; a++) {
                    double s = 0;
                    for (int t = 0; t < 4; t++) s += Km[(ND * e + a) * KE + ND * lc.S[t] + lc.comp];
                    acc[a] -= (sdf.d ? sc * s : s) * lc.L;
                }
            }
        }
#pragma unroll
        for (int a = 0; a < ND; a++) b[(size_t)ND * i + a] = acc[a];
    }
}

__global__ void __launch_bounds__(BS) k_rhs_fluid(double *b, uint32_t nvel, uint32_t ntot, int dir, double f)
{
    for (uint32_t t = blockIdx.x * BS + threadIdx.x; t < ntot; t += gridDim.x * BS) {
        if (t < nvel) {
            b[3 * (size_t)t] = dir == 0 ? f : 0.0;
            b[3 * (size_t)t + 1] = dir == 1 ? f : 0.0;
            b[3 * (size_t)t + 2] = dir == 2 ? f : 0.0;
        }
        b[3 * (size_t)nvel + t] = 0.0;
    }
}

// ---------------------------------------------------------------------------
// effective tensor column: sum_e CB_e u_e (+ seam term), FP64
// (femhmg_thermal_3D.c:355-500, femhmg_elastic_3D.c:1075-1342)
// ---------------------------------------------------------------------------
template <int ND>
__global__ void __launch_bounds__(BS) k_updateC(const double *__restrict__ x, const uint16_t *__restrict__ keys,
                                                const double *__restrict__ CB, const double *__restrict__ seamCB,
                                                int ebe, int ny, int nx, int nz, uint32_t N, LoadCase lc, Reduce red,
                                                Sdf sdf)
{
    constexpr int KE = 8 * ND, NR = ND == 3 ? 6 : 3;
    double acc[6] = {0, 0, 0, 0, 0, 0};
    for (uint32_t i = blockIdx.x * BS + threadIdx.x; i < N; i += gridDim.x * BS) {
        int r, c, l;
        node_rcl(i, ny, nx, r, c, l);
        const Nbh h = make_nbh(r, c, l, ny, nx, nz);
        const int m = ebe ? (int)keys[i] : (int)((keys[i] >> 6) & 3u);  // node i is local node 3 of element i
        double ue[KE];
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const size_t id = h.lo[1 + LDL(k)] + h.co[1 + LDC(k)] + h.ro[1 + LDR(k)];
#pragma unroll
            for (int a = 0; a < ND; a++) ue[ND * k + a] = x[ND * id + a];
        }
        const bool seam = lc.axis == 0 ? c == nx - 1 : (lc.axis == 1 ? r == 0 : l == 0);
        const double sc = sdf_scale(sdf, i);
        for_each_uniform(m, [&](const int mu) {
            cptr<double> CBm = as_const(CB) + (size_t)mu * NR * KE;
#pragma unroll
            for (int row = 0; row < NR; row++) {
                double s = 0;
#pragma unroll
                for (int j = 0; j < KE; j++) s = fma(CBm[row * KE + j], ue[j], s);
                if (seam) s += seamCB[mu * 6 + row];
                acc[row] += sdf.d ? sc * s : s;
            }
        });
    }
    store_partials<6>(red, acc);
}

template <typename T>
__global__ void __launch_bounds__(BS) k_fluid_vsum(const T *__restrict__ x, uint32_t nvel, Reduce red)
{
    double acc[3] = {0, 0, 0};
    for (uint32_t t = blockIdx.x * BS + threadIdx.x; t < nvel; t += gridDim.x * BS) {
        acc[0] += (double)x[3 * (size_t)t];
        acc[1] += (double)x[3 * (size_t)t + 1];
        acc[2] += (double)x[3 * (size_t)t + 2];
    }
    store_partials<3>(red, acc);
}

// ---------------------------------------------------------------------------
// thermal expansion load (femhmg_elastic_3D.c:464-508), gathered node by node:
// node i receives, from each element e that has it as local node k,
// B_c(k)^T D_e a_e with the eigenstrain a_e = alpha_e (1,1,1,0,0,0) and B_c
// the strain matrix at the element centre.  Only the normal rows of B_c meet
// column (k, a), so the term is +-1/4 (D_e a_e)_a; da[m] holds
// (D_m a_m)_xx = (D_m a_m)_yy = (D_m a_m)_zz per material.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(BS) k_rhs_texp(double *b, const uint16_t *__restrict__ keys,
                                                 const double *__restrict__ da, int ebe, int ny, int nx, int nz,
                                                 uint32_t N)
{
    for (uint32_t i = blockIdx.x * BS + threadIdx.x; i < N; i += gridDim.x * BS) {
        int r, c, l;
        node_rcl(i, ny, nx, r, c, l);
        const unsigned key = ebe ? 0u : keys[i];
        double acc[3] = {0, 0, 0};
        for (int k = 0; k < 8; k++) {
            int m;
            if (ebe) {
                int re = r - cDR[k], ce = c - cDC[k], le = l - cDL[k];
                re += re < 0 ? ny : 0; ce += ce < 0 ? nx : 0; le += le < 0 ? nz : 0;
                m = (int)keys[(size_t)re + (size_t)ce * ny + (size_t)le * nx * ny];
            } else {
                m = (int)((key >> (2 * k)) & 3u);
            }
            const double f = 0.25 * da[m];
            // element frame: x = column offset, y = 1 - row offset, z = 1 - layer offset
            acc[0] += cDC[k] ? f : -f;
            acc[1] += cDR[k] ? -f : f;
            acc[2] += cDL[k] ? -f : f;
        }
        b[3 * (size_t)i] = acc[0];
        b[3 * (size_t)i + 1] = acc[1];
        b[3 * (size_t)i + 2] = acc[2];
    }
}

// ---------------------------------------------------------------------------
// nodal fields for export (femhmg_thermal_3D.c:694-1105,
// femhmg_elastic_3D.c:1800-2645): node i gets 1/8 of the flux k grad T
// (thermal, NF = 3) or stress D B u (elastic, NF = 6) of each element that
// has it as a corner, evaluated at that corner.  The element's values carry
// the load case's seam jump (L on component comp of the seam-side corners S
// of every seam element), which is the reference's seam compensation up to a
// rigid translation.  F: per material, per corner k, the NF x 8*ND matrix
// mapping element values to the corner field.
// ---------------------------------------------------------------------------
template <int ND>
__global__ void __launch_bounds__(BS) k_fields(const double *__restrict__ x, const uint16_t *__restrict__ keys,
                                               const double *__restrict__ F, int ebe, int ny, int nx, int nz,
                                               uint32_t N, LoadCase lc, int jump, double *out)
{
    constexpr int KE = 8 * ND, NF = ND == 3 ? 6 : 3;
    for (uint32_t i = blockIdx.x * BS + threadIdx.x; i < N; i += gridDim.x * BS) {
        int r, c, l;
        node_rcl(i, ny, nx, r, c, l);
        const unsigned key = ebe ? 0u : keys[i];
        double acc[NF];
#pragma unroll
        for (int q = 0; q < NF; q++) acc[q] = 0;
        for (int k = 0; k < 8; k++) {
            int re = r - cDR[k], ce = c - cDC[k], le = l - cDL[k];
            re += re < 0 ? ny : 0; ce += ce < 0 ? nx : 0; le += le < 0 ? nz : 0;
            const size_t el = (size_t)re + (size_t)ce * ny + (size_t)le * nx * ny;
            const int m = ebe ? (int)keys[el] : (int)((key >> (2 * k)) & 3u);
            double ue[KE];
#pragma unroll
            for (int j = 0; j < 8; j++) {
                int rj = re + cDR[j], cj = ce + cDC[j], lj = le + cDL[j];
                rj -= rj >= ny ? ny : 0; cj -= cj >= nx ? nx : 0; lj -= lj >= nz ? nz : 0;
                const size_t nj = (size_t)rj + (size_t)cj * ny + (size_t)lj * nx * ny;
#pragma unroll
                for (int a = 0; a < ND; a++) ue[ND * j + a] = x[ND * nj + a];
            }
            if (jump && (lc.axis == 0 ? ce == nx - 1 : (lc.axis == 1 ? re == 0 : le == 0))) {
#pragma unroll
                for (int t = 0; t < 4; t++) ue[ND * lc.S[t] + lc.comp] += lc.L;
            }
            const double *Fk = F + ((size_t)m * 8 + k) * NF * KE;
#pragma unroll
            for (int q = 0; q < NF; q++) {
                double s = 0;
#pragma unroll
                for (int j = 0; j < KE; j++) s = fma(Fk[q * KE + j], ue[j], s);
                acc[q] += 0.125 * s;
            }
        }
#pragma unroll
        for (int q = 0; q < NF; q++) out[(size_t)NF * i + q] = acc[q];
    }
}

// ---------------------------------------------------------------------------
// coarse -> fine initial guess (chfem/src/femhmg/cudapcg/kernels/image.cu:19-121)
// ---------------------------------------------------------------------------
__global__ void k_project2(double *res, const double *v, uint32_t nx, uint32_t ny, uint32_t nz, int stride)
{
    const uint32_t nxc = nx / 2, nyc = ny / 2, nzc = nz / 2;
    const size_t nc = (size_t)nxc * nyc * (nzc + (nzc < 1));
    for (size_t i = blockIdx.x * (size_t)BS + threadIdx.x; i < nc; i += (size_t)gridDim.x * BS) {
        size_t ii = i % nyc, jj = (i % ((size_t)nxc * nyc)) / nyc, kk = i / ((size_t)nxc * nyc);
        size_t idx = 2 * ii + 2 * jj * ny + 2 * kk * (size_t)nx * ny;
        for (int j = 0; j < stride; j++) res[stride * idx + j] = 2.0 * v[stride * i + j];
    }
}

__global__ void k_interp(double *v, uint32_t nx, uint32_t ny, uint32_t nz, int stride, int pass)
{
    const size_t nxy = (size_t)nx * ny, N = nxy * nz;
    for (size_t i = blockIdx.x * (size_t)BS + threadIdx.x; i < N; i += (size_t)gridDim.x * BS) {
        size_t ii = i % ny, jj = (i % nxy) / ny, kk = i / nxy;
        size_t a, b;
        if (pass == 0) {
            if (!((ii & 1) && !(jj & 1) && !(kk & 1))) continue;
            a = ii == 0 ? i - 1 + ny : i - 1;
            b = (ii + 1) == ny ? i + 1 - ny : i + 1;
        } else if (pass == 1) {
            if (!((jj & 1) && !(kk & 1))) continue;
            size_t loc = i % nxy;
            a = (loc + nxy - ny) % nxy + kk * nxy;
            b = (loc + ny) % nxy + kk * nxy;
        } else {
            if (!(kk & 1)) continue;
            a = (i + N - nxy) % N;
            b = (i + nxy) % N;
        }
        for (int j = 0; j < stride; j++) v[stride * i + j] = 0.5 * v[stride * a + j] + 0.5 * v[stride * b + j];
    }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static inline int grid_for(size_t n) { return reduce_blocks(n); }

// Streaming vector kernels: one contiguous range of 16-byte vectors per
// block, at most stream_cap() blocks (env CHFEM_STREAM_GRID, default 8192:
// measured 3 % faster than 2048 on the 256^3 elastic solve).
static int stream_cap()
{
    static int cap = 0;
    if (!cap) {
        const char *e = getenv("CHFEM_STREAM_GRID");
        cap = e ? atoi(e) : 8192;
        if (cap < 1) cap = 1;
        if (cap > MAX_RED_GRID) cap = MAX_RED_GRID;
    }
    return cap;
}
static inline int stream_grid(size_t n, size_t elem_bytes)
{
    const size_t nv = n * elem_bytes / 16;
    size_t b = (nv + BS - 1) / BS;
    if (b > (size_t)stream_cap()) b = stream_cap();
    return b < 1 ? 1 : (int)b;
}

static Reduce noreduce() { Reduce r; r.partials = nullptr; r.S = nullptr; r.fin = 0; r.slot = 0; r.guard = 0; return r; }

template <int NV>
__global__ void __launch_bounds__(1024) k_reduce_finish(Reduce red, unsigned nb)
{
    if (stopped(red)) return;
    double s[NV];
#pragma unroll
    for (int k = 0; k < NV; k++) {
        s[k] = 0;
        for (unsigned b = threadIdx.x; b < nb; b += blockDim.x) s[k] += red.partials[(size_t)k * nb + b];
    }
    block_sum<NV>(s);
    if (threadIdx.x == 0) {
        double *S = red.S;
        if (red.fin == FIN_ALPHA) {
            S[S_DQ] = s[0];
            S[S_ALPHA] = S[S_DELTA] / s[0];
            if (red.guard) S[S_AHIST + (unsigned)S[S_ITER] % S_AHIST_LEN] = S[S_ALPHA];
        } else if (red.fin == FIN_DELTA) {
            S[S_DELTA_OLD] = S[S_DELTA];
            S[S_DELTA] = s[0];
            if (red.guard) {
                S[S_ITER] += 1.0;
                if (fabs(s[0]) <= S[S_THRESH]) S[S_DONE] = 1.0;
            }
        } else if (red.fin == FIN_CG1 && NV == 4) {
            // delta of this iteration from its own z (s[3]), not carried
            // over from the previous recurrence step: rounding then reaches
            // beta through one step only
            const double delta = s[3 % NV], alpha = delta / s[0];
            const double dnew = delta - 2.0 * alpha * s[1 % NV] + alpha * alpha * s[2 % NV];
            // the beta this iteration's launch formed its direction with
            // (the same quotient it computed), for the next launch that
            // recovers z from the directions
            S[S_BPREV] = S[S_DELTA] / S[S_DELTA_OLD];
            S[S_DQ] = s[0];
            S[S_ALPHA] = alpha;
            S[S_DELTA_OLD] = delta;
            S[S_DELTA] = dnew;
            if (red.guard) {
                S[S_AHIST + (unsigned)S[S_ITER] % S_AHIST_LEN] = alpha;
                S[S_ITER] += 1.0;
                if (fabs(dnew) <= S[S_THRESH]) S[S_DONE] = 1.0;
            }
        } else if (red.fin == FIN_STOP) {
            S[S_STOP] = s[0];
            if (red.guard) {
                S[S_ITER] += 1.0;
                if (fabs(s[0]) <= S[S_THRESH]) S[S_DONE] = 1.0;
            }
        } else if (red.fin == FIN_DELTA_NOCOUNT) {
            S[S_DELTA_OLD] = S[S_DELTA];
            S[S_DELTA] = s[0];
        } else {
#pragma unroll
            for (int k = 0; k < NV; k++) S[red.slot + k] = s[k];
        }
    }
}

void launch_reduce_finish(const Reduce &red, unsigned nb, int nv, hipStream_t s)
{
    if (nv == 1) k_reduce_finish<1><<<1, 1024, 0, s>>>(red, nb);
    else if (nv == 3) k_reduce_finish<3><<<1, 1024, 0, s>>>(red, nb);
    else if (nv == 4) k_reduce_finish<4><<<1, 1024, 0, s>>>(red, nb);
    else k_reduce_finish<6><<<1, 1024, 0, s>>>(red, nb);
}

// reducing launch over a 1-D grid, followed by its finishing kernel
#define LAUNCH_RED(NV, KERNEL, GRID, s, red, ...)                                                  \
    do {                                                                                           \
        const int g_ = (GRID);                                                                     \
        KERNEL<<<g_, BS, 0, s>>>(__VA_ARGS__);                                                     \
        launch_reduce_finish(red, (unsigned)g_, NV, s);                                            \
    } while (0)

template <typename T>
void launch_apply(const OpDesc &op, const T *v, T *q, const Reduce *red, hipStream_t s)
{
    const Reduce rr = red ? *red : noreduce();
    const T *K = (const T *)op.K;
    if (op.physics == PHY_FLUID && op.fkeys) {
        launch_stokes_walsh<T>(op, v, q, red, s);
        return;
    }
    if (op.physics == PHY_FLUID) {
        const uint32_t nt = op.nvel + op.nborder;
        if (red) LAUNCH_RED(1, (k_fluid<T, true>), grid_for(nt), s, rr, v, q, op.dof2node, op.node2dof, op.bkeys, K, op.ny, op.nx, op.nz, op.nvel, op.nborder, rr);
        else k_fluid<T, false><<<grid_for(nt), BS, 0, s>>>(v, q, op.dof2node, op.node2dof, op.bkeys, K, op.ny, op.nx, op.nz, op.nvel, op.nborder, rr);
        return;
    }
    if (op.ebe && op.nmat <= 4 && !op.ebe_atomic) {
        launch_ebe_mfma<T>(op, v, q, red, s);
        return;
    }
    if (op.lamu) {
        // Walsh-basis elastic (isotropic) or conduction operator, NBN or EBE
        WalshTiles t = {op.tile_tr, op.tile_tc, op.tile_lz};
        launch_walsh<T>(op, (const T *)op.lamu, v, q, red, t, s);
        return;
    }
    if (op.ebe) {
        k_fill<T><<<grid_for(op.ndof), BS, 0, s>>>(q, T(0), op.ndof);
        if (op.physics == PHY_THERMAL) k_ebe<T, 1><<<grid_for(op.N), BS, 0, s>>>(v, q, op.keys, K, op.ny, op.nx, op.nz, op.N, op.sdf);
        else k_ebe<T, 3><<<grid_for(op.N), BS, 0, s>>>(v, q, op.keys, K, op.ny, op.nx, op.nz, op.N, op.sdf);
        if (red) LAUNCH_RED(1, k_dot<T>, stream_grid(op.ndof, sizeof(T)), s, rr, v, q, op.ndof, rr);
        return;
    }
    if (op.physics == PHY_THERMAL) {
        if (red) LAUNCH_RED(1, (k_thermal_nbn<T, true>), grid_for(op.N), s, rr, v, q, op.keys, K, op.ny, op.nx, op.nz, op.N, rr);
        else k_thermal_nbn<T, false><<<grid_for(op.N), BS, 0, s>>>(v, q, op.keys, K, op.ny, op.nx, op.nz, op.N, rr);
    } else {
        if (red) LAUNCH_RED(1, (k_elastic_nbn<T, true>), grid_for(op.N), s, rr, v, q, op.keys, K, op.ny, op.nx, op.nz, op.N, rr);
        else k_elastic_nbn<T, false><<<grid_for(op.N), BS, 0, s>>>(v, q, op.keys, K, op.ny, op.nx, op.nz, op.N, rr);
    }
}

template <typename T>
void launch_jacobi(const OpDesc &op, T *dinv, hipStream_t s)
{
    const T *K = (const T *)op.K;
    if (op.physics == PHY_FLUID)
        k_jacobi_fluid<T><<<grid_for(op.nvel + op.nborder), BS, 0, s>>>(op.bkeys, K, dinv, op.nvel, op.nborder);
    else if (op.physics == PHY_THERMAL)
        k_jacobi_keys<T, 1><<<grid_for(op.N), BS, 0, s>>>(op.keys, K, dinv, op.N, op.ebe, op.ny, op.nx, op.nz, op.sdf);
    else
        k_jacobi_keys<T, 3><<<grid_for(op.N), BS, 0, s>>>(op.keys, K, dinv, op.N, op.ebe, op.ny, op.nx, op.nz, op.sdf);
}

template <typename T>
void launch_fill(T *x, double val, size_t n, hipStream_t s) { k_fill<T><<<grid_for(n), BS, 0, s>>>(x, (T)val, n); }

template <typename Tin, typename Tout>
void launch_convert(const Tin *in, Tout *out, size_t n, hipStream_t s) { k_convert<Tin, Tout><<<grid_for(n), BS, 0, s>>>(in, out, n); }

template <typename T>
void launch_precond(const T *dinv, const T *v, T *out, size_t n, const Reduce *red, hipStream_t s)
{
    if (red) LAUNCH_RED(1, (k_precond<T, true>), stream_grid(n, sizeof(T)), s, *red, dinv, v, out, n, *red);
    else k_precond<T, false><<<stream_grid(n, sizeof(T)), BS, 0, s>>>(dinv, v, out, n, noreduce());
}

template <typename T>
void launch_pcg_dir(T *d, const T *z, const double *S, int first, size_t n, hipStream_t s, int guard)
{
    k_pcg_dir<T><<<stream_grid(n, sizeof(T)), BS, 0, s>>>(d, z, S, first, n, guard);
}

template <typename T>
void launch_pcg_update(T *x, T *r, const T *d, T *z, const T *dinv, const Reduce *red, size_t n, hipStream_t s)
{
    LAUNCH_RED(1, k_pcg_update<T>, stream_grid(n, sizeof(T)), s, *red, x, r, d, z, dinv, n, *red);
}

template <typename T>
void launch_sub_precond(T *r, const T *q, const T *dinv, T *z, const Reduce *red, size_t n, hipStream_t s)
{
    LAUNCH_RED(1, k_sub_precond<T>, stream_grid(n, sizeof(T)), s, *red, r, q, dinv, z, n, *red);
}

template <typename T>
void launch_minres_dir(const T *dold, T *dnew, T *q, const T *rt, const T *sv, const T *dinv, const double *S,
                       int first, const Reduce *red, size_t n, hipStream_t s)
{
    LAUNCH_RED(1, k_minres_dir<T>, stream_grid(n, sizeof(T)), s, *red, dold, dnew, q, rt, sv, dinv, S, first, n,
               *red);
}

template <typename T>
void launch_minres_update(T *x, T *rt, const T *d, const T *q, const T *dinv, const Reduce *red, size_t n,
                          hipStream_t s)
{
    if (x) LAUNCH_RED(1, (k_minres_update<T, true>), stream_grid(n, sizeof(T)), s, *red, x, rt, d, q, dinv, n, *red);
    else LAUNCH_RED(1, (k_minres_update<T, false>), stream_grid(n, sizeof(T)), s, *red, x, rt, d, q, dinv, n, *red);
}

// Key-Jacobi launchers: one block per contiguous range of 4-node groups
static inline int group_grid(const OpDesc &op, size_t)
{
    size_t b = ((size_t)op.N / 4 + BS - 1) / BS;
    if (b > (size_t)stream_cap()) b = stream_cap();
    return b < 1 ? 1 : (int)b;
}
#define KJ_DISPATCH(KERNEL, GRIDEXPR, ...)                                                           \
    do {                                                                                             \
        if (op.physics == PHY_ELASTIC) KERNEL<T, 3><<<(GRIDEXPR), BS, 0, s>>>(__VA_ARGS__);          \
        else KERNEL<T, 1><<<(GRIDEXPR), BS, 0, s>>>(__VA_ARGS__);                                    \
    } while (0)

template <typename T>
void launch_kj_precond(const OpDesc &op, const T *v, T *out, const Reduce *red, hipStream_t s)
{
    const int gr = group_grid(op, sizeof(T));
    KJ_DISPATCH(k_kj_precond, gr, op.keys, (const T *)op.kinv, v, out, (size_t)op.N, *red);
    launch_reduce_finish(*red, gr, 1, s);
}

template <typename T>
void launch_kj_pcg_dir(const OpDesc &op, T *d, const T *r, const double *S, hipStream_t s, int guard)
{
    KJ_DISPATCH(k_kj_pcg_dir, group_grid(op, sizeof(T)), d, r, op.keys, (const T *)op.kinv, S, (size_t)op.N, guard);
}

template <typename T>
void launch_kj_pcg_update(const OpDesc &op, T *x, T *r, const T *d, const T *q, const Reduce *red, hipStream_t s)
{
    const int gr = group_grid(op, sizeof(T));
    const T *kinv = (const T *)op.kinv;
    const size_t N = op.N;
    if (op.physics == PHY_ELASTIC) k_kj_pcg_update<T, 3, true><<<gr, BS, 0, s>>>(x, r, d, q, op.keys, kinv, N, *red);
    else k_kj_pcg_update<T, 1, true><<<gr, BS, 0, s>>>(x, r, d, q, op.keys, kinv, N, *red);
    launch_reduce_finish(*red, gr, 1, s);
}

template <typename T>
void launch_kj_zupdate(const OpDesc &op, T *z, const T *q, const Reduce *red, hipStream_t s)
{
    const int gr = group_grid(op, sizeof(T));
    KJ_DISPATCH(k_kj_zupdate, gr, z, q, op.keys, (const T *)op.kinv, (size_t)op.N, *red);
    launch_reduce_finish(*red, gr, 1, s);
}

template <typename T>
void launch_kj_to_z(const OpDesc &op, T *r, const Reduce *red, hipStream_t s)
{
    const int gr = group_grid(op, sizeof(T));
    KJ_DISPATCH(k_kj_to_z, gr, r, op.keys, (const T *)op.kinv, (size_t)op.N, *red);
    launch_reduce_finish(*red, gr, 1, s);
}

// x += sum_j alpha_j ring[j % w] over the window's iterations that ran; the
// coefficients are read once per block into LDS
template <typename T>
__global__ void __launch_bounds__(BS) k_ring_flush(T *x, Ring<T> ring, unsigned j0, unsigned j1, unsigned base,
                                                  const double *__restrict__ S, size_t n)
{
    __shared__ T coef[RING_MAX];
    __shared__ const T *src[RING_MAX];
    __shared__ int cnt;
    if (threadIdx.x == 0) {
        const unsigned ran = base + (unsigned)S[S_ITER];
        const unsigned end = j1 < ran ? j1 : ran;
        int c = 0;
        for (unsigned j = j0; j < end && c < RING_MAX; j++, c++) {
            coef[c] = (T)S[S_AHIST + (j - base) % S_AHIST_LEN];
            src[c] = ring.p[j % (unsigned)ring.w];
        }
        cnt = c;
    }
    __syncthreads();
    const int c = cnt;
    if (c == 0) return;
    // FP32: 16-byte vectors (4 values per lane and load); FP64: one value
    // per lane (measured 4 % faster than 16-byte vectors at 256^3 elastic).
    // The loads of several directions are issued before their products are
    // added.
    constexpr int V = sizeof(T) == 4 ? 4 : 1;
    using GT = T __attribute__((address_space(1)));
    const size_t nv = n / V;
    const size_t stride = (size_t)gridDim.x * BS;
    for (size_t iv = blockIdx.x * (size_t)BS + threadIdx.x; iv < nv; iv += stride) {
        T acc[V];
#pragma unroll
        for (int e = 0; e < V; e++) acc[e] = x[iv * V + e];
#pragma unroll 4
        for (int k = 0; k < c; k++) {
            const T ck = coef[k];
            // (the directions are device buffers: global loads, not flat)
            const GT *sk = (const GT *)(src[k] + iv * V);
#pragma unroll
            for (int e = 0; e < V; e++) acc[e] += ck * sk[e];
        }
#pragma unroll
        for (int e = 0; e < V; e++) x[iv * V + e] = acc[e];
    }
    // the last n % V values
    for (size_t i = nv * V + blockIdx.x * (size_t)BS + threadIdx.x; i < n; i += stride) {
        T acc = x[i];
        for (int k = 0; k < c; k++) acc += coef[k] * src[k][i];
        x[i] = acc;
    }
}

template <typename T>
void launch_ring_flush(T *x, const Ring<T> &ring, unsigned j0, unsigned j1, unsigned base, const double *S, size_t n,
                       hipStream_t s)
{
    if (j1 <= j0) return;
    k_ring_flush<T><<<stream_grid(n, sizeof(T)), BS, 0, s>>>(x, ring, j0, j1, base, S, n);
}

template <typename T>
void launch_kj_sub_precond(const OpDesc &op, T *r, const T *q, T *z, const Reduce *red, hipStream_t s)
{
    const int gr = group_grid(op, sizeof(T));
    KJ_DISPATCH(k_kj_sub_precond, gr, r, q, z, op.keys, (const T *)op.kinv, (size_t)op.N, *red);
    launch_reduce_finish(*red, gr, 1, s);
}

template <typename T>
void launch_add(T *y, const T *x, size_t n, hipStream_t s)
{
    k_add<T><<<grid_for(n), BS, 0, s>>>(y, x, n);
}

void launch_add_correction(double *x, const float *e, size_t n, hipStream_t s)
{
    k_add_correction<<<grid_for(n), BS, 0, s>>>(x, e, n);
}

template <typename T>
void launch_dot(const T *a, const T *b, size_t n, const Reduce *red, hipStream_t s)
{
    LAUNCH_RED(1, k_dot<T>, stream_grid(n, sizeof(T)), s, *red, a, b, n, *red);
}

void launch_rhs(const OpDesc &op, const double *K64, const LoadCase &lc, int dir, double *b, hipStream_t s)
{
    if (op.physics == PHY_FLUID) {
        k_rhs_fluid<<<grid_for(op.nvel + op.nborder), BS, 0, s>>>(b, op.nvel, op.nvel + op.nborder, dir, lc.L);
        return;
    }
    if (op.physics == PHY_THERMAL)
        k_rhs_seam<1><<<grid_for(op.N), BS, 0, s>>>(b, op.keys, K64, op.ebe, op.ny, op.nx, op.nz, op.N, lc, op.sdf);
    else
        k_rhs_seam<3><<<grid_for(op.N), BS, 0, s>>>(b, op.keys, K64, op.ebe, op.ny, op.nx, op.nz, op.N, lc, op.sdf);
}

void launch_updateC(const OpDesc &op, const double *CB64, const double *seamCB, const LoadCase &lc, const double *x,
                    const Reduce *red, hipStream_t s)
{
    if (op.physics == PHY_THERMAL)
        LAUNCH_RED(6, k_updateC<1>, grid_for(op.N), s, *red, x, op.keys, CB64, seamCB, op.ebe, op.ny, op.nx, op.nz, op.N, lc, *red, op.sdf);
    else
        LAUNCH_RED(6, k_updateC<3>, grid_for(op.N), s, *red, x, op.keys, CB64, seamCB, op.ebe, op.ny, op.nx, op.nz, op.N, lc, *red, op.sdf);
}

void launch_rhs_thermal_expansion(const OpDesc &op, const double *da, double *b, hipStream_t s)
{
    k_rhs_texp<<<grid_for(op.N), BS, 0, s>>>(b, op.keys, da, op.ebe, op.ny, op.nx, op.nz, op.N);
}

void launch_fields(const OpDesc &op, const double *F, const LoadCase &lc, int jump, const double *x, double *out,
                   hipStream_t s)
{
    if (op.physics == PHY_THERMAL)
        k_fields<1><<<grid_for(op.N), BS, 0, s>>>(x, op.keys, F, op.ebe, op.ny, op.nx, op.nz, op.N, lc, jump, out);
    else
        k_fields<3><<<grid_for(op.N), BS, 0, s>>>(x, op.keys, F, op.ebe, op.ny, op.nx, op.nz, op.N, lc, jump, out);
}

template <typename T>
void launch_fluid_velocity_sum(const OpDesc &op, const T *x, const Reduce *red, hipStream_t s)
{
    LAUNCH_RED(3, k_fluid_vsum<T>, grid_for(op.nvel ? op.nvel : 1), s, *red, x, op.nvel, *red);
}

void launch_interpl2(const double *coarse, double *fine, int rows, int cols, int layers, int stride, hipStream_t s)
{
    const bool is2d = layers == 0;
    const int lay = layers < 1 ? 1 : layers;
    const size_t n = (size_t)rows * cols * lay;
    k_fill<double><<<grid_for(n * stride), BS, 0, s>>>(fine, 0.0, n * stride);
    k_project2<<<grid_for(n), BS, 0, s>>>(fine, coarse, cols, rows, lay, stride);
    k_interp<<<grid_for(n), BS, 0, s>>>(fine, cols, rows, lay, stride, 0);
    k_interp<<<grid_for(n), BS, 0, s>>>(fine, cols, rows, lay, stride, 1);
    if (!is2d) k_interp<<<grid_for(n), BS, 0, s>>>(fine, cols, rows, lay, stride, 2);
}

// explicit instantiations
#define INST(T)                                                                                                \
    template void launch_apply<T>(const OpDesc &, const T *, T *, const Reduce *, hipStream_t);               \
    template void launch_jacobi<T>(const OpDesc &, T *, hipStream_t);                                         \
    template void launch_fill<T>(T *, double, size_t, hipStream_t);                                            \
    template void launch_add<T>(T *, const T *, size_t, hipStream_t);                                         \
    template void launch_precond<T>(const T *, const T *, T *, size_t, const Reduce *, hipStream_t);          \
    template void launch_pcg_dir<T>(T *, const T *, const double *, int, size_t, hipStream_t, int);           \
    template void launch_pcg_update<T>(T *, T *, const T *, T *, const T *, const Reduce *, size_t, hipStream_t); \
    template void launch_sub_precond<T>(T *, const T *, const T *, T *, const Reduce *, size_t, hipStream_t); \
    template void launch_minres_dir<T>(const T *, T *, T *, const T *, const T *, const T *, const double *, int, \
                                       const Reduce *, size_t, hipStream_t);                                  \
    template void launch_minres_update<T>(T *, T *, const T *, const T *, const T *, const Reduce *, size_t,  \
                                          hipStream_t);                                                       \
    template void launch_dot<T>(const T *, const T *, size_t, const Reduce *, hipStream_t);                   \
    template void launch_fluid_velocity_sum<T>(const OpDesc &, const T *, const Reduce *, hipStream_t);       \
    template void launch_kj_precond<T>(const OpDesc &, const T *, T *, const Reduce *, hipStream_t);          \
    template void launch_kj_pcg_dir<T>(const OpDesc &, T *, const T *, const double *, hipStream_t, int);     \
    template void launch_kj_pcg_update<T>(const OpDesc &, T *, T *, const T *, const T *, const Reduce *,     \
                                          hipStream_t);                                                       \
    template void launch_kj_sub_precond<T>(const OpDesc &, T *, const T *, T *, const Reduce *, hipStream_t); \
    template void launch_ring_flush<T>(T *, const Ring<T> &, unsigned, unsigned, unsigned, const double *, size_t, \
                                       hipStream_t);                                                                \
    template void launch_kj_zupdate<T>(const OpDesc &, T *, const T *, const Reduce *, hipStream_t);          \
    template void launch_kj_to_z<T>(const OpDesc &, T *, const Reduce *, hipStream_t);
INST(float)
INST(double)
template void launch_convert<double, float>(const double *, float *, size_t, hipStream_t);
template void launch_convert<float, double>(const float *, double *, size_t, hipStream_t);
template void launch_convert<double, double>(const double *, double *, size_t, hipStream_t);

}  // namespace chfem

// ---------------------------------------------------------------------------
// calibration: plain grid-stride FP64 copy (known traffic: 8 B read + 8 B
// written per entry) for scaling hardware byte counters
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(BS) k_copy_calibration(const double *__restrict__ a, double *__restrict__ b, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)BS + threadIdx.x; i < n; i += (size_t)gridDim.x * BS) b[i] = a[i];
}

extern "C" double chfemCopyKernelBench(size_t n, unsigned int reps)
{
    double *a = nullptr, *b = nullptr;
    if (hipMalloc(&a, n * sizeof(double)) != hipSuccess) return -1.0;
    if (hipMalloc(&b, n * sizeof(double)) != hipSuccess) { hipFree(a); return -1.0; }
    hipMemset(a, 0, n * sizeof(double));
    hipEvent_t e0, e1;
    hipEventCreate(&e0);
    hipEventCreate(&e1);
    const int grid = (int)((n + BS - 1) / BS < 65536 ? (n + BS - 1) / BS : 65536);
    k_copy_calibration<<<grid, BS>>>(a, b, n);
    hipEventRecord(e0, 0);
    for (unsigned r = 0; r < reps; r++) k_copy_calibration<<<grid, BS>>>(a, b, n);
    hipEventRecord(e1, 0);
    hipEventSynchronize(e1);
    float ms = 0;
    hipEventElapsedTime(&ms, e0, e1);
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    hipFree(a);
    hipFree(b);
    return reps ? ms / reps : 0.0;
}
