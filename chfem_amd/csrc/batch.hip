// batch.hip -- kernels of the batched solver: many independent systems
// (every load case of many samples of different sizes) advanced together,
// one launch per kernel and iteration for the whole batch.
//
// Work is described by tables built on the host: operator tiles
// (system, Walsh tile coordinates) and vector chunks (system, node range).  Vectors of all
// systems are concatenated; each system keeps its own FP64 scalars
// (delta, alpha, stop flags) in a row of Sb, and its block partials sit in a
// contiguous range of the partials buffer, summed in a fixed order by one
// finishing block per system.  Systems whose flags say so skip their tiles
// and chunks, so converged systems cost only the launch.
#include <hip/hip_runtime.h>
#include "batch.h"
#include "walsh.h"

namespace chfem {
namespace batch {

#define BB 256

__device__ __forceinline__ bool flag(const double *Sb, int sys, int slot) { return Sb[sys * SB_STRIDE + slot] != 0.0; }

// block sum of one double; valid in thread 0
__device__ __forceinline__ double bsum(double v)
{
    double a[1] = {v};
    block_sum<1>(a);
    return a[0];
}

// ---------------------------------------------------------------------------
// operator: q = A v (+ per-block partials of v.q) for the systems selected
// by `mode`: OP_ITER runs systems with ACTIVE set and DONE clear, OP_REPL
// those with REPL set.
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ bool selected(const double *Sb, int sys, int mode)
{
    return mode == OP_ITER ? (flag(Sb, sys, SB_ACTIVE) && !flag(Sb, sys, SB_DONE)) : flag(Sb, sys, SB_REPL);
}

// Walsh-basis operator over the tile table (ND = 3 elastic, 1 conduction).
// Blocks are dispatched to the 8 XCDs round-robin; block b takes table entry
// start(b % 8) + b / 8, so each XCD walks one contiguous range of the table.
// The table holds, for each XCD in turn, one contiguous eighth of every
// system's tiles (rows of tiles fastest): neighbouring tiles share an XCD's
// L2, and every XCD keeps work while any system is still iterating.  Block
// partials go to the tile's slot in its system's contiguous range.
template <typename T, int ND, int TR, int TC>
__global__ void __launch_bounds__(TR *TC + (TILE_HALO ? 64 : 0)) k_apply_walsh(const Sys *__restrict__ sys, const Tile *__restrict__ tiles,
                                                         const T *v, T *q, const double *__restrict__ Sb,
                                                         double *__restrict__ part, int mode)
{
    const unsigned b = blockIdx.x, nb = gridDim.x, per = nb / 8u, rem = nb % 8u, xcd = b % 8u, k = b / 8u;
    const unsigned ti = (xcd < rem ? xcd * (per + 1u) : rem * (per + 1u) + (xcd - rem) * per) + k;
    const Tile t = tiles[ti];
    if (!selected<T>(Sb, t.sys, mode)) return;
    const Sys &S = sys[t.sys];
    const T *lamu = (const T *)(sizeof(T) == 4 ? S.mat32 : S.mat64);
    const double d = walsh_tile<T, ND, TR, TC, true, 0, TILE_HALO>(v + S.off, q + S.off, S.keys, lamu, 0, S.ny, S.nx, S.nz, S.LZ,
                                                     t.bx, t.by, t.bz).dq;
    const double s = bsum(d);
    if (threadIdx.x == 0) part[t.slot] = s;
}

// Fused step of the z-form loop: d = z + beta d_old (or d = z on a fresh
// step) formed while the planes are fetched, q = A d and the partial d.q of
// each tile.  A system's iteration j (its own count, SB_ITER) writes ring
// slot j % w and reads slot (j - 1) % w: the direction carries over the
// steps a system sits out and its residual replacements.
template <int ND, int TR, int TC>
__global__ void __launch_bounds__(TR *TC + (TILE_HALO ? 64 : 0)) k_apply_dir(const Sys *__restrict__ sys, const Tile *__restrict__ tiles,
                                                       const float *z, float *ring, size_t total, int w, float *q,
                                                       const double *__restrict__ Sb, double *__restrict__ part)
{
    const unsigned b = blockIdx.x, nb = gridDim.x, per = nb / 8u, rem = nb % 8u, xcd = b % 8u, k = b / 8u;
    const unsigned ti = (xcd < rem ? xcd * (per + 1u) : rem * (per + 1u) + (xcd - rem) * per) + k;
    const Tile t = tiles[ti];
    if (!selected<float>(Sb, t.sys, OP_ITER)) return;
    const Sys &S = sys[t.sys];
    const double *R = Sb + t.sys * SB_STRIDE;
    const unsigned j = (unsigned)R[SB_ITER], uw = (unsigned)w;
    const float *dold = ring + (size_t)((j + uw - 1) % uw) * total;
    DirFuse<float> fz{};
    fz.z = z + S.off;
    fz.dnew = ring + (size_t)(j % uw) * total + S.off;
    fz.first = R[SB_FRESH] != 0.0 ? 1 : 0;
    fz.beta = fz.first ? 0.f : (float)(R[SB_DELTA] / R[SB_DELTA_OLD]);
    const double d = walsh_tile<float, ND, TR, TC, true, 1, TILE_HALO>(dold + S.off, q + S.off, S.keys, (const float *)S.mat32, 0,
                                                            S.ny, S.nx, S.nz, S.LZ, t.bx, t.by, t.bz, fz).dq;
    const double s = bsum(d);
    if (threadIdx.x == 0) part[t.slot] = s;
}

// Single-kernel step of the z-form loop (see apply_cg1 in batch.h); 4
// partials per tile at part[k * ntiles + tile].
template <int ND, int TR, int TC>
__global__ void __launch_bounds__(TR *TC + (TILE_HALO ? 64 : 0)) k_apply_cg1(const Sys *__restrict__ sys, const Tile *__restrict__ tiles,
                                                       int ntiles, float *z1, float *z2, float *q1, float *q2,
                                                       float *ring, size_t total, int w,
                                                       const double *__restrict__ Sb, double *__restrict__ part)
{
    const unsigned b = blockIdx.x, nb = gridDim.x, per = nb / 8u, rem = nb % 8u, xcd = b % 8u, k = b / 8u;
    const unsigned ti = (xcd < rem ? xcd * (per + 1u) : rem * (per + 1u) + (xcd - rem) * per) + k;
    const Tile t = tiles[ti];
    if (!selected<float>(Sb, t.sys, OP_ITER)) return;
    const Sys &S = sys[t.sys];
    const double *R = Sb + t.sys * SB_STRIDE;
    const unsigned j = (unsigned)R[SB_ITER], base = (unsigned)R[SB_BASE], uw = (unsigned)w;
    float *zb[2] = {z2 + S.off, z1 + S.off}, *qb[2] = {q1 + S.off, q2 + S.off};
    const bool first = R[SB_FRESH] != 0.0;
    const bool pend = !first && j > base;
    const bool zr = uw >= 3 && pend && j >= base + 2;
    const float *dold = first ? nullptr : ring + (size_t)((j + uw - 1) % uw) * total + S.off;
    DirFuse<float> fz{};
    fz.z = zr ? ring + (size_t)((j + uw - 2) % uw) * total + S.off : zb[(j + 1) & 1];
    fz.dnew = ring + (size_t)(j % uw) * total + S.off;
    fz.first = first ? 1 : 0;
    fz.beta = first ? 0.f : (float)(R[SB_DELTA] / R[SB_DELTA_OLD]);
    fz.zpend = pend ? 1 : 0;
    fz.qold = pend ? qb[(j + 1) & 1] : fz.z;
    fz.alpha = pend ? (float)R[SB_ALPHA] : 0.f;
    fz.znew = (uw >= 3 && j >= base + 1) ? nullptr : zb[j & 1];
    fz.kinv = (const float *)S.kinv32;
    fz.zring = zr ? 1 : 0;
    fz.bprev = zr ? (float)R[SB_BPREV] : 0.f;
    fz.lean = 0;
    const TileSums ts = walsh_tile<float, ND, TR, TC, true, 2, TILE_HALO>(dold, qb[j & 1], S.keys, (const float *)S.mat32, 0,
                                                             S.ny, S.nx, S.nz, S.LZ, t.bx, t.by, t.bz, fz);
    double v[4] = {ts.dq, ts.zq, ts.qmq, ts.zdz};
    block_sum<4>(v);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int c = 0; c < 4; c++) part[(size_t)c * ntiles + t.slot] = v[c];
    }
}

// ---------------------------------------------------------------------------
// finishing kernels: one block per system sums its partial range
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(BB) k_finish_cg1(const Sys *__restrict__ sys, double *__restrict__ Sb,
                                                   const double *__restrict__ part, int ntiles, int w)
{
    const int s = blockIdx.x;
    double *R = Sb + s * SB_STRIDE;
    if (R[SB_ACTIVE] == 0.0 || R[SB_DONE] != 0.0) return;
    const Sys &S = sys[s];
    double v[4];
#pragma unroll
    for (int c = 0; c < 4; c++) {
        v[c] = 0.0;
        for (int b = S.tb0 + (int)threadIdx.x; b < S.tb1; b += blockDim.x) v[c] += part[(size_t)c * ntiles + b];
    }
    block_sum<4>(v);
    if (threadIdx.x != 0) return;
    // as FIN_CG1 of the single-system loop: this step's delta from its own
    // z, alpha, the next r.z by the recurrence
    const double delta = v[3], alpha = delta / v[0];
    const double dnew = delta - 2.0 * alpha * v[1] + alpha * alpha * v[2];
    R[SB_BPREV] = R[SB_DELTA] / R[SB_DELTA_OLD];
    R[SB_DQ] = v[0];
    R[SB_ALPHA] = alpha;
    R[SB_DELTA_OLD] = delta;
    R[SB_DELTA] = dnew;
    if (w > 0) R[SB_AHIST + (unsigned)R[SB_ITER] % (unsigned)w] = alpha;
    R[SB_ITER] += 1.0;
    R[SB_FRESH] = 0.0;
    if (fabs(dnew) <= R[SB_THRESH]) R[SB_DONE] = 1.0;
}
__global__ void __launch_bounds__(BB) k_finish(const Sys *__restrict__ sys, double *__restrict__ Sb,
                                               const double *__restrict__ part, int what, int w)
{
    const int s = blockIdx.x;
    double *R = Sb + s * SB_STRIDE;
    const Sys &S = sys[s];
    bool run;
    if (what == FIN_B_ALPHA || what == FIN_B_DELTA) run = R[SB_ACTIVE] != 0.0 && R[SB_DONE] == 0.0;
    else if (what == FIN_B_TRUE) run = R[SB_REPL] != 0.0;
    else if (what == FIN_B_START) run = R[SB_START] != 0.0;
    else run = true;
    if (!run) return;
    const bool op = what == FIN_B_ALPHA;
    const int b0 = op ? S.tb0 : S.vb0, b1 = op ? S.tb1 : S.vb1;
    double acc = 0;
    for (int b = b0 + (int)threadIdx.x; b < b1; b += blockDim.x) acc += part[b];
    const double tot = bsum(acc);
    if (threadIdx.x != 0) return;
    if (what == FIN_B_ALPHA) {
        R[SB_DQ] = tot;
        R[SB_ALPHA] = R[SB_DELTA] / tot;
        if (w > 0) R[SB_AHIST + (unsigned)R[SB_ITER] % (unsigned)w] = R[SB_ALPHA];
    } else if (what == FIN_B_DELTA) {
        R[SB_DELTA_OLD] = R[SB_DELTA];
        R[SB_DELTA] = tot;
        R[SB_ITER] += 1.0;
        R[SB_FRESH] = 0.0;
        if (fabs(tot) <= R[SB_THRESH]) R[SB_DONE] = 1.0;
    } else if (what == FIN_B_TRUE) {
        R[SB_DELTA] = tot;
    } else if (what == FIN_B_DELTA0) {
        R[SB_DELTA0] = tot;
    } else {  // FIN_B_START: FP32 r.z of a fresh start
        R[SB_DELTA] = tot;
    }
}

// ---------------------------------------------------------------------------
// vector kernels over chunks of BB nodes
// ---------------------------------------------------------------------------
template <typename T, int ND>
__device__ __forceinline__ void node_dinv(const Sys &S, uint32_t i, T *dv)
{
    const T *kinv = (const T *)(sizeof(T) == 4 ? S.kinv32 : S.kinv64);
    const unsigned key = S.keys[i];
#pragma unroll
    for (int a = 0; a < ND; a++) dv[a] = kinv[(size_t)key * ND + a];
}

// d = dinv r + beta d (beta = delta / delta_old; not on a system's first iteration)
template <typename T, int ND>
__global__ void __launch_bounds__(BB) k_dir(const Sys *__restrict__ sys, const Chunk *__restrict__ chunks,
                                            const double *__restrict__ Sb, T *d, const T *__restrict__ r)
{
    const Chunk ch = chunks[blockIdx.x];
    const double *R = Sb + ch.sys * SB_STRIDE;
    if (R[SB_ACTIVE] == 0.0 || R[SB_DONE] != 0.0 || R[SB_FRESH] != 0.0) return;
    const Sys &S = sys[ch.sys];
    const uint32_t i = ch.i0 + threadIdx.x;
    if (i >= ch.i1) return;
    const T beta = (T)(R[SB_DELTA] / R[SB_DELTA_OLD]);
    T dv[ND];
    node_dinv<T, ND>(S, i, dv);
    const size_t o = S.off + (size_t)ND * i;
#pragma unroll
    for (int a = 0; a < ND; a++) d[o + a] = dv[a] * r[o + a] + beta * d[o + a];
}

// x += alpha d ; r -= alpha q ; partial r.(dinv r)
template <typename T, int ND>
__global__ void __launch_bounds__(BB) k_update(const Sys *__restrict__ sys, const Chunk *__restrict__ chunks,
                                               const double *__restrict__ Sb, T *x, T *r, const T *__restrict__ d,
                                               const T *__restrict__ q, double *__restrict__ part)
{
    const Chunk ch = chunks[blockIdx.x];
    const double *R = Sb + ch.sys * SB_STRIDE;
    if (R[SB_ACTIVE] == 0.0 || R[SB_DONE] != 0.0) return;
    const Sys &S = sys[ch.sys];
    const uint32_t i = ch.i0 + threadIdx.x;
    double acc = 0;
    if (i < ch.i1) {
        const T alpha = (T)R[SB_ALPHA];
        T dv[ND];
        node_dinv<T, ND>(S, i, dv);
        const size_t o = S.off + (size_t)ND * i;
#pragma unroll
        for (int a = 0; a < ND; a++) {
            x[o + a] += alpha * d[o + a];
            const T rn = r[o + a] - alpha * q[o + a];
            r[o + a] = rn;
            acc += (double)rn * (double)(dv[a] * rn);
        }
    }
    const double s = bsum(acc);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// z -= alpha dinv q ; partial z.(diag z) (= r.z)
template <int ND>
__global__ void __launch_bounds__(BB) k_zupdate(const Sys *__restrict__ sys, const Chunk *__restrict__ chunks,
                                                const double *__restrict__ Sb, float *z, const float *__restrict__ q,
                                                double *__restrict__ part)
{
    const Chunk ch = chunks[blockIdx.x];
    const double *R = Sb + ch.sys * SB_STRIDE;
    if (R[SB_ACTIVE] == 0.0 || R[SB_DONE] != 0.0) return;
    const Sys &S = sys[ch.sys];
    const uint32_t i = ch.i0 + threadIdx.x;
    double acc = 0;
    if (i < ch.i1) {
        const float alpha = (float)R[SB_ALPHA];
        const float *kinv = (const float *)S.kinv32;
        const unsigned key = S.keys[i];
        const size_t o = S.off + (size_t)ND * i;
#pragma unroll
        for (int a = 0; a < ND; a++) {
            const float m = kinv[(size_t)key * ND + a], dg = kinv[(size_t)65536 * ND + (size_t)key * ND + a];
            const float zn = z[o + a] - alpha * (m * q[o + a]);
            z[o + a] = zn;
            acc += (double)zn * (double)(dg * zn);
        }
    }
    const double s = bsum(acc);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// x32 += sum_k alpha_k ring[k] over the window (every system; alpha 0 for
// steps a system did not run)
__global__ void __launch_bounds__(BB) k_ring_flush(const Sys *__restrict__ sys, const Chunk *__restrict__ chunks,
                                                   const double *__restrict__ Sb, float *x32,
                                                   const float *__restrict__ ring, size_t total, int w, int nd)
{
    __shared__ float coef[SB_RING_MAX];
    const Chunk ch = chunks[blockIdx.x];
    const double *R = Sb + ch.sys * SB_STRIDE;
    if (R[SB_ACTIVE] == 0.0) return;  // ran no step since its alphas were cleared (block-uniform)
    if (threadIdx.x < (unsigned)w) coef[threadIdx.x] = (float)R[SB_AHIST + threadIdx.x];
    __syncthreads();
    const Sys &S = sys[ch.sys];
    // the chunk's nodes as one contiguous run of nd values each: consecutive
    // work-items take consecutive values (coalesced; a node per work-item
    // strided its loads by nd values)
    const size_t o0 = S.off + (size_t)nd * ch.i0;
    const uint32_t len = (uint32_t)nd * (ch.i1 - ch.i0);
    for (uint32_t f = threadIdx.x; f < len; f += BB) {
        const size_t o = o0 + f;
        float acc = x32[o];
        for (int k = 0; k < w; k++) acc += coef[k] * ring[(size_t)k * total + o];
        x32[o] = acc;
    }
}

// zeroes the window's alphas after its flush (one block per system)
__global__ void k_clear_alpha(double *Sb, int w)
{
    if (threadIdx.x < (unsigned)w) Sb[blockIdx.x * SB_STRIDE + SB_AHIST + threadIdx.x] = 0.0;
}

// z-form start of an FP32 segment: z32 = dinv (float) r64, x32 = 0, partial
// r.z (systems with `start` set)
__global__ void __launch_bounds__(BB) k_start_z(const Sys *__restrict__ sys, const Chunk *__restrict__ chunks,
                                                const double *__restrict__ Sb, const double *__restrict__ r64,
                                                float *z32, float *x32, double *__restrict__ part, int nd, float *z2)
{
    const Chunk ch = chunks[blockIdx.x];
    const double *R = Sb + ch.sys * SB_STRIDE;
    if (R[SB_START] == 0.0) return;
    if (z2 && ((unsigned)R[SB_ITER] & 1u)) z32 = z2;
    const Sys &S = sys[ch.sys];
    const uint32_t i = ch.i0 + threadIdx.x;
    double acc = 0;
    if (i < ch.i1) {
        const float *kinv = (const float *)S.kinv32;
        const unsigned key = S.keys[i];
        for (int a = 0; a < nd; a++) {
            const size_t o = S.off + (size_t)nd * i + a;
            const float rv = (float)r64[o];
            const float z = kinv[(size_t)key * nd + a] * rv;
            z32[o] = z;
            x32[o] = 0.f;
            acc += (double)rv * (double)z;
        }
    }
    const double s = bsum(acc);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// start of an FP32 segment: r32 = (float) r64, d32 = dinv r32, x32 = 0,
// partial r32.d32 (systems with `start` set)
__global__ void __launch_bounds__(BB) k_start(const Sys *__restrict__ sys, const Chunk *__restrict__ chunks,
                                              const double *__restrict__ Sb, const double *__restrict__ r64,
                                              float *r32, float *d32, float *x32, double *__restrict__ part, int nd)
{
    const Chunk ch = chunks[blockIdx.x];
    const double *R = Sb + ch.sys * SB_STRIDE;
    if (R[SB_START] == 0.0) return;
    const Sys &S = sys[ch.sys];
    const uint32_t i = ch.i0 + threadIdx.x;
    double acc = 0;
    if (i < ch.i1) {
        const float *kinv = (const float *)S.kinv32;
        const unsigned key = S.keys[i];
        for (int a = 0; a < nd; a++) {
            const size_t o = S.off + (size_t)nd * i + a;
            const float rv = (float)r64[o];
            const float z = kinv[(size_t)key * nd + a] * rv;
            r32[o] = rv;
            d32[o] = z;
            x32[o] = 0.f;
            acc += (double)rv * (double)z;
        }
    }
    const double s = bsum(acc);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// FP64 r.(dinv r) of every system (initial delta0), partials per chunk
__global__ void __launch_bounds__(BB) k_delta0(const Sys *__restrict__ sys, const Chunk *__restrict__ chunks,
                                               const double *__restrict__ r64, double *__restrict__ part, int nd)
{
    const Chunk ch = chunks[blockIdx.x];
    const Sys &S = sys[ch.sys];
    const uint32_t i = ch.i0 + threadIdx.x;
    double acc = 0;
    if (i < ch.i1) {
        const double *kinv = (const double *)S.kinv64;
        const unsigned key = S.keys[i];
        for (int a = 0; a < nd; a++) {
            const double rv = r64[S.off + (size_t)nd * i + a];
            acc += rv * (kinv[(size_t)key * nd + a] * rv);
        }
    }
    const double s = bsum(acc);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// residual replacement, part 1 (systems with REPL set): x64 += x32,
// e64 = x32, x32 = 0
__global__ void __launch_bounds__(BB) k_flush(const Sys *__restrict__ sys, const Chunk *__restrict__ chunks,
                                              const double *__restrict__ Sb, double *x64, double *e64, float *x32,
                                              int nd)
{
    const Chunk ch = chunks[blockIdx.x];
    if (Sb[ch.sys * SB_STRIDE + SB_REPL] == 0.0) return;
    const Sys &S = sys[ch.sys];
    const uint32_t i = ch.i0 + threadIdx.x;
    if (i >= ch.i1) return;
    for (int a = 0; a < nd; a++) {
        const size_t o = S.off + (size_t)nd * i + a;
        const double e = (double)x32[o];
        x64[o] += e;
        e64[o] = e;
        x32[o] = 0.f;
    }
}

// part 2: r64 -= q64 ; partial r64.(dinv r64) ; r32 = (float) r64 (the next
// FP32 segment continues from the FP64 residual)
__global__ void __launch_bounds__(BB) k_resid(const Sys *__restrict__ sys, const Chunk *__restrict__ chunks,
                                              const double *__restrict__ Sb, double *r64, const double *__restrict__ q64,
                                              float *r32, double *__restrict__ part, int nd)
{
    const Chunk ch = chunks[blockIdx.x];
    if (Sb[ch.sys * SB_STRIDE + SB_REPL] == 0.0) return;
    const Sys &S = sys[ch.sys];
    const uint32_t i = ch.i0 + threadIdx.x;
    double acc = 0;
    if (i < ch.i1) {
        const double *kinv = (const double *)S.kinv64;
        const unsigned key = S.keys[i];
        for (int a = 0; a < nd; a++) {
            const size_t o = S.off + (size_t)nd * i + a;
            const double rn = r64[o] - q64[o];
            r64[o] = rn;
            r32[o] = (float)rn;
            acc += rn * (kinv[(size_t)key * nd + a] * rn);
        }
    }
    const double s = bsum(acc);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// z-form replacement, part 2: r64 -= q64 ; partial r64.(dinv r64) ;
// z32 = dinv32 (float) r64
__global__ void __launch_bounds__(BB) k_resid_z(const Sys *__restrict__ sys, const Chunk *__restrict__ chunks,
                                                const double *__restrict__ Sb, double *r64,
                                                const double *__restrict__ q64, float *z32, double *__restrict__ part,
                                                int nd, float *z2)
{
    const Chunk ch = chunks[blockIdx.x];
    if (Sb[ch.sys * SB_STRIDE + SB_REPL] == 0.0) return;
    if (z2 && ((unsigned)Sb[ch.sys * SB_STRIDE + SB_ITER] & 1u)) z32 = z2;
    const Sys &S = sys[ch.sys];
    const uint32_t i = ch.i0 + threadIdx.x;
    double acc = 0;
    if (i < ch.i1) {
        const double *kinv = (const double *)S.kinv64;
        const float *kinv32 = (const float *)S.kinv32;
        const unsigned key = S.keys[i];
        for (int a = 0; a < nd; a++) {
            const size_t o = S.off + (size_t)nd * i + a;
            const double rn = r64[o] - q64[o];
            r64[o] = rn;
            z32[o] = kinv32[(size_t)key * nd + a] * (float)rn;
            acc += rn * (kinv[(size_t)key * nd + a] * rn);
        }
    }
    const double s = bsum(acc);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
template <typename T>
void apply(const Batch &b, const T *v, T *q, int mode, hipStream_t s)
{
    if (b.ntiles == 0) return;
    if (b.nd == 3)
        k_apply_walsh<T, 3, TILE_R, TILE_C><<<b.ntiles, TILE_THREADS, 0, s>>>(b.sys, b.tiles, v, q, b.Sb, b.part, mode);
    else
        k_apply_walsh<T, 1, TILE_R, TILE_C><<<b.ntiles, TILE_THREADS, 0, s>>>(b.sys, b.tiles, v, q, b.Sb, b.part, mode);
}
template void apply<float>(const Batch &, const float *, float *, int, hipStream_t);
template void apply<double>(const Batch &, const double *, double *, int, hipStream_t);

void finish(const Batch &b, int what, hipStream_t s)
{
    k_finish<<<b.nsys, BB, 0, s>>>(b.sys, b.Sb, b.part, what, 0);
}

void finish_alpha(const Batch &b, int w, hipStream_t s)
{
    k_finish<<<b.nsys, BB, 0, s>>>(b.sys, b.Sb, b.part, FIN_B_ALPHA, w);
}

void apply_dir(const Batch &b, const float *z, float *ring, size_t total, int w, float *q, hipStream_t s)
{
    if (b.ntiles == 0) return;
    if (b.nd == 3)
        k_apply_dir<3, TILE_R, TILE_C><<<b.ntiles, TILE_THREADS, 0, s>>>(b.sys, b.tiles, z, ring, total, w, q, b.Sb,
                                                                          b.part);
    else
        k_apply_dir<1, TILE_R, TILE_C><<<b.ntiles, TILE_THREADS, 0, s>>>(b.sys, b.tiles, z, ring, total, w, q, b.Sb,
                                                                          b.part);
}

void zupdate(const Batch &b, float *z, const float *q, hipStream_t s)
{
    if (b.nd == 3) k_zupdate<3><<<b.nchunks, BB, 0, s>>>(b.sys, b.chunks, b.Sb, z, q, b.part);
    else k_zupdate<1><<<b.nchunks, BB, 0, s>>>(b.sys, b.chunks, b.Sb, z, q, b.part);
}

void ring_flush(const Batch &b, float *x32, const float *ring, size_t total, int w, hipStream_t s)
{
    k_ring_flush<<<b.nchunks, BB, 0, s>>>(b.sys, b.chunks, b.Sb, x32, ring, total, w, b.nd);
    k_clear_alpha<<<b.nsys, SB_RING_MAX, 0, s>>>(b.Sb, w);
}

void start_z(const Batch &b, const double *r64, float *z32, float *x32, hipStream_t s, float *z2)
{
    k_start_z<<<b.nchunks, BB, 0, s>>>(b.sys, b.chunks, b.Sb, r64, z32, x32, b.part, b.nd, z2);
}

void resid_z(const Batch &b, double *r64, const double *q64, float *z32, hipStream_t s, float *z2)
{
    k_resid_z<<<b.nchunks, BB, 0, s>>>(b.sys, b.chunks, b.Sb, r64, q64, z32, b.part, b.nd, z2);
}

void apply_cg1(const Batch &b, float *z, float *z2, float *q, float *q2, float *ring, size_t total, int w,
               hipStream_t s)
{
    if (b.ntiles == 0) return;
    if (b.nd == 3)
        k_apply_cg1<3, TILE_R, TILE_C><<<b.ntiles, TILE_THREADS, 0, s>>>(b.sys, b.tiles, b.ntiles, z, z2, q, q2, ring,
                                                                           total, w, b.Sb, b.part);
    else
        k_apply_cg1<1, TILE_R, TILE_C><<<b.ntiles, TILE_THREADS, 0, s>>>(b.sys, b.tiles, b.ntiles, z, z2, q, q2, ring,
                                                                           total, w, b.Sb, b.part);
    k_finish_cg1<<<b.nsys, BB, 0, s>>>(b.sys, b.Sb, b.part, b.ntiles, w);
}

void dir(const Batch &b, float *d, const float *r, hipStream_t s)
{
    if (b.nd == 3) k_dir<float, 3><<<b.nchunks, BB, 0, s>>>(b.sys, b.chunks, b.Sb, d, r);
    else k_dir<float, 1><<<b.nchunks, BB, 0, s>>>(b.sys, b.chunks, b.Sb, d, r);
}

void update(const Batch &b, float *x, float *r, const float *d, const float *q, hipStream_t s)
{
    if (b.nd == 3) k_update<float, 3><<<b.nchunks, BB, 0, s>>>(b.sys, b.chunks, b.Sb, x, r, d, q, b.part);
    else k_update<float, 1><<<b.nchunks, BB, 0, s>>>(b.sys, b.chunks, b.Sb, x, r, d, q, b.part);
}

void start(const Batch &b, const double *r64, float *r32, float *d32, float *x32, hipStream_t s)
{
    k_start<<<b.nchunks, BB, 0, s>>>(b.sys, b.chunks, b.Sb, r64, r32, d32, x32, b.part, b.nd);
}

void delta0(const Batch &b, const double *r64, hipStream_t s)
{
    k_delta0<<<b.nchunks, BB, 0, s>>>(b.sys, b.chunks, r64, b.part, b.nd);
}

void flush(const Batch &b, double *x64, double *e64, float *x32, hipStream_t s)
{
    k_flush<<<b.nchunks, BB, 0, s>>>(b.sys, b.chunks, b.Sb, x64, e64, x32, b.nd);
}

void resid(const Batch &b, double *r64, const double *q64, float *r32, hipStream_t s)
{
    k_resid<<<b.nchunks, BB, 0, s>>>(b.sys, b.chunks, b.Sb, r64, q64, r32, b.part, b.nd);
}

}  // namespace batch
}  // namespace chfem
