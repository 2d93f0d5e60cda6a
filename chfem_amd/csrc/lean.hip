// lean.hip -- streaming kernels of the memory-lean Krylov solvers (CG3 and
// MINRES3 keep x, z and d; CG2 and MINRES2 keep z and d and reduce the
// solution on the fly, chfem/src/femhmg/cudapcg/solvers/{pcg3,pminres3,pcg2,
// pminres2,xreduce}.cu), the absolute-maximum reductions of the INF and
// error stopping rules (stopping_criteria.cu, pcg.cu:130-146), and the
// dispatch of the lean operator forms.
//
// The solvers iterate on the preconditioned residual z = M^-1 r, so the
// Jacobi diagonal is needed only where an operator result meets z: in the
// operator's epilogue (Walsh-basis NBN and Stokes operators) or, for the
// other operators, in k_lean_epi over a work vector holding A v.
#include <hip/hip_runtime.h>
#include "kernels.h"
#include "reduce.h"

namespace chfem {

#define LBS 256

template <typename T>
__device__ __forceinline__ T jac_inv(const Jac<T> &j, size_t i)
{
    if (j.kind == JAC_KEYS) {
        const size_t node = j.nd == 3 ? i / 3 : i;
        return j.kinv[(size_t)j.nd * j.keys[node] + (i - node * j.nd)];
    }
    if (j.kind == JAC_FLUID) {
        const size_t nv3 = 3 * (size_t)j.nvel;
        if (i < nv3) return j.vinv;
        const size_t p = i - nv3;
        return p < j.nvel ? j.pinv : j.btab[j.bkeys[p - j.nvel]];
    }
    if (j.kind == JAC_STORED) return j.dinv[i];
    return T(1);
}

// z.diag(M) z for one entry, in FP64 (diag from the key table, or 1 / M^-1)
template <typename T>
__device__ __forceinline__ double jac_zdz(const Jac<T> &j, size_t i, T z)
{
    if (j.kind == JAC_KEYS) {
        const size_t node = j.nd == 3 ? i / 3 : i;
        const T dg = j.kinv[(size_t)65536 * j.nd + (size_t)j.nd * j.keys[node] + (i - node * j.nd)];
        return (double)z * (double)(dg * z);
    }
    if (j.kind == JAC_NONE) return (double)z * (double)z;
    return (double)z * (double)z / (double)jac_inv(j, i);
}

template <typename T>
__global__ void __launch_bounds__(LBS) k_lean_to_z(Jac<T> jac, T *z, size_t n, Reduce red)
{
    double acc = 0.0;
    for (size_t i = (size_t)blockIdx.x * LBS + threadIdx.x; i < n; i += (size_t)gridDim.x * LBS) {
        const T r = z[i], zi = jac_inv(jac, i) * r;
        z[i] = zi;
        acc += (double)r * (double)zi;
    }
    double v[1] = {acc};
    store_partials<1>(red, v);
}

template <typename T>
__global__ void __launch_bounds__(LBS) k_lean_dir(T *x, T *d, const T *__restrict__ z, const double *S, int first,
                                                  int guard, size_t n)
{
    if (guard && S[S_DONE] != 0.0) return;
    const T a = first ? T(0) : (T)S[S_ALPHA];
    const T b = first ? T(0) : (T)(S[S_DELTA] / S[S_DELTA_OLD]);
    for (size_t i = (size_t)blockIdx.x * LBS + threadIdx.x; i < n; i += (size_t)gridDim.x * LBS) {
        const T di = d[i];
        if (!first && x) x[i] += a * di;
        d[i] = first ? z[i] : z[i] + b * di;
    }
}

template <typename T>
__global__ void __launch_bounds__(LBS) k_lean_axpy(T *x, const T *__restrict__ d, const double *S, size_t n)
{
    const T a = (T)S[S_ALPHA];
    for (size_t i = (size_t)blockIdx.x * LBS + threadIdx.x; i < n; i += (size_t)gridDim.x * LBS) x[i] += a * d[i];
}

// the lean forms over a work vector q = A v (operators without a lean
// epilogue)
template <typename T>
__global__ void __launch_bounds__(LBS) k_lean_epi(int form, const T *__restrict__ v, const T *__restrict__ q, T *z,
                                                  Jac<T> jac, size_t n, Reduce red)
{
    if (stopped(red)) return;
    const T a = form == LEAN_ZUPD ? (T)red.S[S_ALPHA] : T(0);
    double acc = 0.0;
    for (size_t i = (size_t)blockIdx.x * LBS + threadIdx.x; i < n; i += (size_t)gridDim.x * LBS) {
        const T qi = q[i];
        if (form == LEAN_VAV) {
            acc += (double)v[i] * (double)qi;
        } else if (form == LEAN_QMQ) {
            acc += (double)qi * (double)(jac_inv(jac, i) * qi);
        } else {
            const T zn = z[i] - a * (jac_inv(jac, i) * qi);
            z[i] = zn;
            acc += jac_zdz(jac, i, zn);
        }
    }
    double vv[1] = {acc};
    store_partials<1>(red, vv);
}

// absolute maximum (optionally of z / M^-1, the residual of a
// preconditioned z); block maxima, then one block takes their maximum
template <typename T>
__global__ void __launch_bounds__(LBS) k_absmax(const T *__restrict__ v, size_t n, Jac<T> jac, int scaled,
                                                double *partials)
{
    double m = 0.0;
    for (size_t i = (size_t)blockIdx.x * LBS + threadIdx.x; i < n; i += (size_t)gridDim.x * LBS) {
        const double a = scaled ? fabs((double)v[i] / (double)jac_inv(jac, i)) : fabs((double)v[i]);
        m = fmax(m, a);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
    __shared__ double w[LBS / 64];
    if ((threadIdx.x & 63) == 0) w[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < LBS / 64; k++) m = fmax(m, w[k]);
        partials[blockIdx.x] = fmax(m, w[0]);
    }
}

__global__ void __launch_bounds__(1024) k_max_finish(const double *partials, unsigned nb, double *S, int slot)
{
    double m = 0.0;
    for (unsigned b = threadIdx.x; b < nb; b += blockDim.x) m = fmax(m, partials[b]);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
    __shared__ double w[16];
    if ((threadIdx.x & 63) == 0) w[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 0; k < (int)(blockDim.x + 63) / 64; k++) m = fmax(m, w[k]);
        S[slot] = m;
    }
}

// xreduce step (see launch_xreduce); init: x0 is the initial guess
template <typename T>
__global__ void __launch_bounds__(LBS) k_xreduce(T *xr, const T *__restrict__ d, const T *__restrict__ x0, int mode,
                                                 int shift, int p, size_t nh, Reduce red)
{
    if (stopped(red)) return;
    const bool init = x0 != nullptr;
    const T a = init ? T(0) : (T)red.S[S_ALPHA];
    double s[4] = {0.0, 0.0, 0.0, 0.0};
    const size_t nn = nh / (size_t)p;
    if (mode == 1 && !init) {
        for (size_t j = (size_t)blockIdx.x * LBS + threadIdx.x; j < nn; j += (size_t)gridDim.x * LBS) {
            const T xn = xr[j] + a * d[(size_t)p * j + shift];
            xr[j] = xn;
            s[0] += (double)xn;
        }
    } else {
        for (size_t i = (size_t)blockIdx.x * LBS + threadIdx.x; i < nn * p; i += (size_t)gridDim.x * LBS) {
            const int c = (int)(i % (size_t)p);
            T val;
            if (init) {
                val = x0[i];
                if (mode == 2) xr[i] = val;
                if (mode == 1 && c == shift) xr[i / p] = val;
            } else if (mode == 2) {
                val = xr[i] + a * d[i];
                xr[i] = val;
            } else {
                val = d[i];
            }
            s[c] += (double)val;
        }
    }
    store_partials<4>(red, s);
}

__global__ void __launch_bounds__(1024) k_xreduce_finish(Reduce red, unsigned nb, double *hist, size_t hist_len,
                                                         int mode, int shift, int p, int init)
{
    if (stopped(red)) return;
    double s[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        s[k] = 0;
        for (unsigned b = threadIdx.x; b < nb; b += blockDim.x) s[k] += red.partials[(size_t)k * nb + b];
    }
    block_sum<4>(s);
    if (threadIdx.x == 0) {
        const size_t j = init ? 0 : (size_t)red.S[S_ITER] + 1;
        if (j >= hist_len) return;
        double *h = hist + (size_t)p * j;
        if (init) {
            for (int c = 0; c < p; c++) h[c] = s[c];
        } else if (mode == 0) {
            for (int c = 0; c < p; c++) h[c] = red.S[S_ALPHA] * s[c];
        } else if (mode == 1) {
            h[shift] = s[0];
        } else {
            for (int c = 0; c < p; c++) h[c] = s[c];
        }
    }
}

static inline int lean_grid(size_t n)
{
    size_t b = (n + LBS - 1) / LBS;
    if (b > MAX_RED_GRID) b = MAX_RED_GRID;
    return b < 1 ? 1 : (int)b;
}

template <typename T>
void launch_lean_to_z(const Jac<T> &jac, T *z, size_t n, const Reduce *red, hipStream_t s)
{
    const int g = lean_grid(n);
    k_lean_to_z<T><<<g, LBS, 0, s>>>(jac, z, n, *red);
    launch_reduce_finish(*red, (unsigned)g, 1, s);
}

template <typename T>
void launch_lean_dir(T *x, T *d, const T *z, const double *S, int first, int guard, size_t n, hipStream_t s)
{
    k_lean_dir<T><<<lean_grid(n), LBS, 0, s>>>(x, d, z, S, first, guard, n);
}

template <typename T>
void launch_lean_axpy(T *x, const T *d, const double *S, size_t n, hipStream_t s)
{
    k_lean_axpy<T><<<lean_grid(n), LBS, 0, s>>>(x, d, S, n);
}

template <typename T>
void launch_xreduce(T *xr, const T *d, const T *x0, double *hist, size_t hist_len, int mode, int shift, int nvpn,
                    size_t nh, const Reduce &red, hipStream_t s)
{
    const int g = lean_grid(nh);
    k_xreduce<T><<<g, LBS, 0, s>>>(xr, d, x0, mode, shift, nvpn, nh, red);
    k_xreduce_finish<<<1, 1024, 0, s>>>(red, (unsigned)g, hist, hist_len, mode, shift, nvpn, x0 ? 1 : 0);
}

template <typename T>
void launch_absmax(const T *v, size_t n, const Jac<T> *scaled, double *partials, double *S, int slot, hipStream_t s)
{
    const int g = lean_grid(n);
    Jac<T> j = {};
    if (scaled) j = *scaled;
    k_absmax<T><<<g, LBS, 0, s>>>(v, n, j, scaled ? 1 : 0, partials);
    k_max_finish<<<1, 1024, 0, s>>>(partials, (unsigned)g, S, slot);
}

bool lean_needs_qtmp(const OpDesc &op)
{
    if (op.physics == PHY_FLUID) return op.fkeys == nullptr;
    return !(op.lamu && !op.ebe);
}

template <typename T>
void launch_lean_apply(const OpDesc &op, const Jac<T> &jac, int form, const T *v, T *z, T *qtmp, const Reduce *red,
                       hipStream_t s)
{
    if (op.physics == PHY_FLUID && op.fkeys && (jac.kind == JAC_FLUID || jac.kind == JAC_NONE)) {
        launch_stokes_lean<T>(op, jac, form, v, z, red, s);
        return;
    }
    if (!lean_needs_qtmp(op) && op.physics != PHY_FLUID && (jac.kind == JAC_KEYS || jac.kind == JAC_NONE)) {
        launch_walsh_lean<T>(op, form, v, z, jac.kind == JAC_KEYS ? jac.kinv : nullptr, red, s);
        return;
    }
    launch_apply<T>(op, v, qtmp, nullptr, s);
    const size_t n = op.ndof;
    const int g = lean_grid(n);
    k_lean_epi<T><<<g, LBS, 0, s>>>(form, v, qtmp, z, jac, n, *red);
    launch_reduce_finish(*red, (unsigned)g, 1, s);
}

#define LEAN_INST(T)                                                                                                \
    template void launch_lean_to_z<T>(const Jac<T> &, T *, size_t, const Reduce *, hipStream_t);                  \
    template void launch_lean_dir<T>(T *, T *, const T *, const double *, int, int, size_t, hipStream_t);         \
    template void launch_lean_axpy<T>(T *, const T *, const double *, size_t, hipStream_t);                       \
    template void launch_xreduce<T>(T *, const T *, const T *, double *, size_t, int, int, int, size_t,           \
                                    const Reduce &, hipStream_t);                                                 \
    template void launch_absmax<T>(const T *, size_t, const Jac<T> *, double *, double *, int, hipStream_t);      \
    template void launch_lean_apply<T>(const OpDesc &, const Jac<T> &, int, const T *, T *, T *, const Reduce *, \
                                       hipStream_t);
LEAN_INST(float)
LEAN_INST(double)

}  // namespace chfem
