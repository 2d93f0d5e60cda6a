// walsh.hip -- kernel and launcher of the Walsh-basis elastic operator
// (device code in walsh.h).
#include <hip/hip_runtime.h>
#include "walsh.h"

namespace chfem {

template <typename T, int TR, int TC, bool DOT>
__global__ void __launch_bounds__(TR *TC) k_elastic_walsh(const T *__restrict__ v, T *__restrict__ q,
                                                           const uint16_t *__restrict__ keys,
                                                           const T *__restrict__ lamu, int ebe, int ny, int nx,
                                                           int nz, int LZ, Reduce red)
{
    if (DOT && stopped(red)) return;
    const double dsum = walsh_tile<T, TR, TC, DOT>(v, q, keys, lamu, ebe, ny, nx, nz, LZ, blockIdx.x, blockIdx.y,
                                                    blockIdx.z);
    if (DOT) {
        double vv[1] = {dsum};
        store_partials<1>(red, vv);
    }
}

template <typename T>
void launch_elastic_walsh(const OpDesc &op, const T *lamu, const T *v, T *q, const Reduce *red, WalshTiles t,
                          hipStream_t s)
{
    Reduce rr = {nullptr, nullptr, 0, 0, 0};
    if (red) rr = *red;
    int LZ = t.LZ > 0 ? t.LZ : 16;
    {
        // keep the block count within the reduction partials buffer
        const int tr = t.TR > 0 ? t.TR : 32, tc = t.TC > 0 ? t.TC : 8;
        const long cols = (long)((op.ny + tr - 2) / (tr - 1)) * ((op.nx + tc - 2) / (tc - 1));
        while (cols * ((op.nz + LZ - 1) / LZ) > MAX_PARTIAL_BLOCKS) LZ *= 2;
    }
#define WALSH_LAUNCH(TRv, TCv)                                                                          \
    {                                                                                                   \
        dim3 grid((op.ny + TRv - 2) / (TRv - 1), (op.nx + TCv - 2) / (TCv - 1), (op.nz + LZ - 1) / LZ);             \
        if (red) {                                                                                      \
            k_elastic_walsh<T, TRv, TCv, true><<<grid, TRv * TCv, 0, s>>>(v, q, op.keys, lamu, op.ebe,     \
                                                                          op.ny, op.nx, op.nz, LZ, rr);    \
            launch_reduce_finish(rr, grid.x * grid.y * grid.z, 1, s);                                   \
        } else k_elastic_walsh<T, TRv, TCv, false><<<grid, TRv * TCv, 0, s>>>(v, q, op.keys, lamu, op.ebe,  \
                                                                            op.ny, op.nx, op.nz, LZ, rr);      \
        return;                                                                                         \
    }
    if (t.TR == 16 && t.TC == 16) WALSH_LAUNCH(16, 16)
    if (t.TR == 64 && t.TC == 4) WALSH_LAUNCH(64, 4)
    if (t.TR == 32 && t.TC == 16) WALSH_LAUNCH(32, 16)
    WALSH_LAUNCH(32, 8)
#undef WALSH_LAUNCH
}

template void launch_elastic_walsh<float>(const OpDesc &, const float *, const float *, float *, const Reduce *,
                                          WalshTiles, hipStream_t);
template void launch_elastic_walsh<double>(const OpDesc &, const double *, const double *, double *, const Reduce *,
                                           WalshTiles, hipStream_t);

}  // namespace chfem
