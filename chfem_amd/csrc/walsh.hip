// walsh.hip -- kernel and launcher of the Walsh-basis elastic and conduction operators
// (device code in walsh.h).
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>
#include "walsh.h"

namespace chfem {

template <typename T, int ND, int TR, int TC, bool DOT, int FM, bool HALO>
__global__ void __launch_bounds__(TR *TC + (HALO ? 64 : 0)) k_walsh(const T *__restrict__ v, T *__restrict__ q,
                                                   const uint16_t *__restrict__ keys, const T *__restrict__ lamu,
                                                   int ebe, int ny, int nx, int nz, int LZ, Reduce red, DirFuse<T> fz)
{
    constexpr bool FUSE = FM == 1 || FM == 2;
    if (DOT && stopped(red)) return;
    // beta of the fused direction update, as k_kj_pcg_dir forms it
    if (FUSE && !fz.first) fz.beta = (T)(red.S[S_DELTA] / red.S[S_DELTA_OLD]);
    // single-kernel CG: alpha of the previous iteration for the deferred z
    // update
    if (FM == 2) fz.alpha = fz.zpend ? (T)red.S[S_ALPHA] : T(0);
    // single-kernel CG with z recovered from the ring: the beta that formed
    // d_old
    if (FM == 2 && fz.zring) fz.bprev = (T)red.S[S_BPREV];
    // lean z update: alpha of this iteration
    if (FM == 3) fz.alpha = fz.lean == LEAN_ZUPD ? (T)red.S[S_ALPHA] : T(0);
    // 1-D grid; blocks are dispatched to the 8 XCDs round-robin (block b to
    // XCD b % 8), so each XCD is given one contiguous range of tiles (rows of
    // tiles fastest): tiles that share halo nodes then run on the same XCD at
    // about the same time and meet each other's halo lines in its L2.
    // Single-kernel CG iterations of a guarded loop walk each XCD's range
    // backwards on odd iterations: an iteration then starts with the tiles
    // whose d and q the previous one wrote last, some of which the
    // memory-side cache still holds (1.7 % off the 256^3 FP32 launch).
    constexpr int OR = HALO ? TR : TR - 1;  // node rows per tile
    const unsigned gx = (unsigned)(ny + OR - 1) / OR, gy = (unsigned)(nx + TC - 2) / (TC - 1);
    const unsigned b = blockIdx.x, nb = gridDim.x, per = nb / 8u, rem = nb % 8u, xcd = b % 8u;
    unsigned k = b / 8u;
    if (FM == 2 && DOT && red.guard && (((unsigned)red.S[S_ITER]) & 1u)) k = (xcd < rem ? per + 1u : per) - 1u - k;
    const unsigned t = (xcd < rem ? xcd * (per + 1u) : rem * (per + 1u) + (xcd - rem) * per) + k;
    const TileSums ts = walsh_tile<T, ND, TR, TC, DOT, FM, HALO>(v, q, keys, lamu, ebe, ny, nx, nz, LZ, (int)(t % gx),
                                                           (int)((t / gx) % gy), (int)(t / (gx * gy)), fz);
    if (DOT && FM == 2) {
        double vv[4] = {ts.dq, ts.zq, ts.qmq, ts.zdz};
        store_partials<4>(red, vv);
    } else if (DOT) {
        double vv[1] = {ts.dq};
        store_partials<1>(red, vv);
    }
}

template <typename T, int ND>
static void launch_walsh_nd(const OpDesc &op, const T *lamu, const T *v, T *q, const Reduce *red, WalshTiles t,
                            hipStream_t s, const DirFuse<T> *fz, const hipEvent_t *ev)
{
    Reduce rr = {nullptr, nullptr, 0, 0, 0};
    if (red) rr = *red;
    DirFuse<T> f = {nullptr, nullptr, T(0), T(0), 1, 0, nullptr, nullptr, nullptr, 0, T(0), 0};
    if (fz) f = *fz;
    // lean forms of the memory-lean solvers: one tile shape (32 x 8)
    const bool lean = fz && fz->lean;
    // single-kernel CG form: the dinv table given (only launch_apply_cg1
    // sets it; its z output may be absent)
    const bool cg1 = fz && fz->kinv;
    int LZ = t.LZ > 0 ? t.LZ : 16;
    {
        // keep the block count within the reduction partials buffer
        const int tr = t.TR > 0 ? t.TR : 32, tc = t.TC > 0 ? t.TC : 8;
        const long cols = (long)((op.ny + tr - 2) / (tr - 1)) * ((op.nx + tc - 2) / (tc - 1));
        while (cols * ((op.nz + LZ - 1) / LZ) > MAX_PARTIAL_BLOCKS) LZ *= 2;
    }
    // ev: a start and a stop event bound to the k_walsh dispatch itself
    // (hipExtLaunchKernel), so their interval is the kernel's own run time
#define WALSH_GO(TRv, TCv, H, DOTv, FMv)                                                                \
    if (ev) hipExtLaunchKernelGGL(k_walsh<T, ND, TRv, TCv, DOTv, FMv, H>, dim3(grid), dim3(nt), 0, s, ev[0], ev[1], \
                                  0, v, q, op.keys, lamu, op.ebe, op.ny, op.nx, op.nz, LZ, rr, f);               \
    else k_walsh<T, ND, TRv, TCv, DOTv, FMv, H><<<grid, nt, 0, s>>>(v, q, op.keys, lamu, op.ebe, op.ny, op.nx,   \
                                                                     op.nz, LZ, rr, f);
    // H: the halo-wavefront variant (tiles own all TRv rows; one more
    // wavefront per block)
#define WALSH_LAUNCH(TRv, TCv, H)                                                                       \
    {                                                                                                   \
        constexpr int orows = H ? TRv : TRv - 1, nt = TRv * TCv + (H ? 64 : 0);                         \
        const unsigned grid = (unsigned)((op.ny + orows - 1) / orows) * ((op.nx + TCv - 2) / (TCv - 1)) * \
                              ((op.nz + LZ - 1) / LZ);                                                  \
        if (cg1) {                                                                                      \
            WALSH_GO(TRv, TCv, H, true, 2)                                                              \
            launch_reduce_finish(rr, grid, 4, s);                                                       \
        } else if (fz) {                                                                                \
            WALSH_GO(TRv, TCv, H, true, 1)                                                              \
            launch_reduce_finish(rr, grid, 1, s);                                                       \
        } else if (red) {                                                                               \
            WALSH_GO(TRv, TCv, H, true, 0)                                                              \
            launch_reduce_finish(rr, grid, 1, s);                                                       \
        } else {                                                                                        \
            WALSH_GO(TRv, TCv, H, false, 0)                                                             \
        }                                                                                               \
        return;                                                                                         \
    }
    // tile rows 65: 65 element rows per column (64 owned + the row above),
    // the halo-wavefront variant of 64-row tiles.  15 columns (1024
    // work-items) only in FP32: the FP64 planes of the single-kernel CG
    // iteration would not fit in LDS.
    if (lean) {
#define WALSH_LAUNCH_LEAN(TRv, TCv)                                                                     \
    {                                                                                                   \
        constexpr int nt = TRv * TCv;                                                                   \
        const unsigned grid = (unsigned)((op.ny + TRv - 2) / (TRv - 1)) * ((op.nx + TCv - 2) / (TCv - 1)) * \
                              ((op.nz + LZ - 1) / LZ);                                                  \
        WALSH_GO(TRv, TCv, false, true, 3)                                                              \
        launch_reduce_finish(rr, grid, 1, s);                                                           \
        return;                                                                                         \
    }
        WALSH_LAUNCH_LEAN(32, 8)
#undef WALSH_LAUNCH_LEAN
    }
    if constexpr (sizeof(T) == 4) {
        if (t.TR == 65 && t.TC == 15) WALSH_LAUNCH(64, 15, true)
    }
    if (t.TR == 65) WALSH_LAUNCH(64, 7, true)
    if (t.TR == 16 && t.TC == 16) WALSH_LAUNCH(16, 16, false)
    if (t.TR == 64 && t.TC == 4) WALSH_LAUNCH(64, 4, false)
    if (t.TR == 64 && t.TC == 8) WALSH_LAUNCH(64, 8, false)
    if (t.TR == 32 && t.TC == 16) WALSH_LAUNCH(32, 16, false)
    WALSH_LAUNCH(32, 8, false)
#undef WALSH_LAUNCH
#undef WALSH_GO
}

template <typename T>
void launch_walsh(const OpDesc &op, const T *lamu, const T *v, T *q, const Reduce *red, WalshTiles t, hipStream_t s,
                  const DirFuse<T> *fz, const hipEvent_t *ev)
{
    if (op.physics == PHY_ELASTIC) launch_walsh_nd<T, 3>(op, lamu, v, q, red, t, s, fz, ev);
    else launch_walsh_nd<T, 1>(op, lamu, v, q, red, t, s, fz, ev);
}

template <typename T>
void launch_walsh_lean(const OpDesc &op, int form, const T *v, T *z, const void *kinv, const Reduce *red,
                       hipStream_t s)
{
    DirFuse<T> f = {nullptr, nullptr, T(0), T(0), 1, 0, nullptr, nullptr, (const T *)kinv, 0, T(0), form};
    const WalshTiles t = {32, 8, op.tile_lz};
    launch_walsh<T>(op, (const T *)op.lamu, v, z, red, t, s, &f, nullptr);
}
template void launch_walsh_lean<float>(const OpDesc &, int, const float *, float *, const void *, const Reduce *,
                                       hipStream_t);
template void launch_walsh_lean<double>(const OpDesc &, int, const double *, double *, const void *,
                                        const Reduce *, hipStream_t);

bool apply_dir_ok(const OpDesc &op)
{
    // the operators launch_apply sends to the Walsh kernel, with key-Jacobi on
    return op.physics != PHY_FLUID && !op.ebe && op.lamu && op.kinv;
}

template <typename T>
void launch_apply_dir(const OpDesc &op, const T *z, const T *dold, T *dnew, T *q, const Reduce *red, hipStream_t s,
                      const hipEvent_t *ev)
{
    const DirFuse<T> f = {z, dnew, T(0), T(0), dold ? 0 : 1, 0, nullptr, nullptr, nullptr, 0, T(0), 0};
    const WalshTiles t = {op.tile_tr, op.tile_tc, op.tile_lz};
    launch_walsh<T>(op, (const T *)op.lamu, dold, q, red, t, s, &f, ev);
}

template <typename T>
void launch_apply_cg1(const OpDesc &op, const T *zold, const T *qold, const T *dold, T *znew, T *dnew, T *qnew,
                      int pend, const Reduce *red, hipStream_t s, const hipEvent_t *ev, const T *dold2)
{
    const int first = dold ? 0 : 1, p = dold && pend ? 1 : 0;
    // dold2: z_old recovered from the directions (needs d_old and the
    // pending update)
    const int zr = dold2 && p ? 1 : 0;
    const DirFuse<T> f = {zr ? dold2 : zold, dnew, T(0), T(0), first, p, p ? qold : zold, znew, (const T *)op.kinv,
                          zr, T(0), 0};
    const WalshTiles t = {op.tile_tr, op.tile_tc, op.tile_lz};
    launch_walsh<T>(op, (const T *)op.lamu, dold, qnew, red, t, s, &f, ev);
}

template void launch_apply_cg1<float>(const OpDesc &, const float *, const float *, const float *, float *, float *,
                                      float *, int, const Reduce *, hipStream_t, const hipEvent_t *, const float *);
template void launch_apply_cg1<double>(const OpDesc &, const double *, const double *, const double *, double *,
                                       double *, double *, int, const Reduce *, hipStream_t, const hipEvent_t *,
                                       const double *);

template void launch_apply_dir<float>(const OpDesc &, const float *, const float *, float *, float *, const Reduce *,
                                      hipStream_t, const hipEvent_t *);
template void launch_apply_dir<double>(const OpDesc &, const double *, const double *, double *, double *,
                                       const Reduce *, hipStream_t, const hipEvent_t *);
template void launch_walsh<float>(const OpDesc &, const float *, const float *, float *, const Reduce *, WalshTiles,
                                  hipStream_t, const DirFuse<float> *, const hipEvent_t *);
template void launch_walsh<double>(const OpDesc &, const double *, const double *, double *, const Reduce *,
                                   WalshTiles, hipStream_t, const DirFuse<double> *, const hipEvent_t *);

}  // namespace chfem
