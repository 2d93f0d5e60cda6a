// guess.hip -- kernels of the learned initial guess: a periodic 3D
// convolutional encoder-decoder over the voxel grid whose products run on the
// matrix cores (v_mfma_f32_16x16x16_bf16: BF16 operands, FP32 accumulation),
// and the kernels that train it (input and weight gradients, loss, Adam).
//
// Convolutions are implicit GEMMs: rows are output voxels, columns output
// channels, and the reduction runs over (tap, input channel).  One wavefront
// owns 4 groups of 16 output voxels and all (<= 64) output channels; its A
// operands are gathered straight from the channels-last input (8 bytes = 4
// channels per lane), its B operand is the packed BF16 weight row, loaded
// once for the 4 groups (37-47 % faster than one group per wavefront for
// the full-resolution layers at 256^3).  The products are the 16x16x16 BF16
// shape: these layers are bound by their operand gathers, not by the matrix
// cores (a 16 -> 16 channel layer at 256^3 runs about 230 GFLOP in 2.2 ms),
// so the 16x16x32 shape of gfx950 would not shorten them.
#include <hip/hip_runtime.h>
#include <stdio.h>
#include "guess.h"

namespace chfem {
namespace nn {

typedef short short4_t __attribute__((ext_vector_type(4)));
typedef float float4_t __attribute__((ext_vector_type(4)));

#define NB 256

__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// FP32 -> BF16, round to nearest even
__device__ __forceinline__ uint16_t f2bf(float f)
{
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__device__ __forceinline__ float4_t mfma16(short4_t a, short4_t b, float4_t c)
{
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ int wrap(int a, int n) { return a < 0 ? a + n : (a >= n ? a - n : a); }

// input voxel feeding tap t of the output row at (r, c, l)
template <int MODE>
__device__ __forceinline__ size_t in_index(int r, int c, int l, int t, const Grid3 &gi, const Grid3 &go)
{
    if (MODE == CONV_SAME) {
        const int rr = wrap(r + t % 3 - 1, gi.ny), cc = wrap(c + (t / 3) % 3 - 1, gi.nx), ll = wrap(l + t / 9 - 1, gi.nz);
        return (size_t)rr + (size_t)cc * gi.ny + (size_t)ll * gi.ny * gi.nx;
    } else if (MODE == CONV_DOWN) {
        const int rr = 2 * r + (t & 1), cc = 2 * c + ((t >> 1) & 1), ll = 2 * l + (t >> 2);
        return (size_t)rr + (size_t)cc * gi.ny + (size_t)ll * gi.ny * gi.nx;
    } else if (MODE == CONV_UP) {
        const int rr = wrap(r + t % 3 - 1, go.ny), cc = wrap(c + (t / 3) % 3 - 1, go.nx), ll = wrap(l + t / 9 - 1, go.nz);
        return (size_t)(rr >> 1) + (size_t)(cc >> 1) * gi.ny + (size_t)(ll >> 1) * gi.ny * gi.nx;
    } else {
        return (size_t)r + (size_t)c * gi.ny + (size_t)l * gi.ny * gi.nx;
    }
}

// Each wavefront owns CG groups of 16 consecutive output voxels (64): a
// weight operand, loaded once per (tap, 16 input channels, 16 output
// channels), feeds CG products, and the CG input gathers of a step are in
// flight together.
constexpr int CG = 4;

template <int MODE, int NT, bool OUTF>
__global__ void __launch_bounds__(NB) k_conv(const uint16_t *__restrict__ in, Grid3 gi, int Ci,
                                             const uint16_t *__restrict__ W, int taps, const float *__restrict__ bias,
                                             const uint16_t *__restrict__ skip, int relu, void *out, int accumulate,
                                             Grid3 go)
{
    constexpr int Co = NT * 16;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int i = lane & 15, kg = lane >> 4;
    const Grid3 rg = MODE == CONV_PARITY ? gi : go;  // grid of the output rows
    const size_t V = (size_t)rg.ny * rg.nx * rg.nz;
    const size_t vb = ((size_t)blockIdx.x * 4 + wave) * 16 * CG;
    if (vb >= V) return;  // no barriers in this kernel
    const int par = MODE == CONV_PARITY ? (int)blockIdx.y : 0;
    const uint16_t *Wp = W + (size_t)par * Co * taps * Ci;
    bool okA[CG];
    int r[CG], c[CG], l[CG];
#pragma unroll
    for (int gq = 0; gq < CG; gq++) {
        size_t va = vb + 16 * gq + i;
        okA[gq] = va < V;
        if (!okA[gq]) va = V - 1;
        r[gq] = (int)(va % rg.ny);
        c[gq] = (int)((va / rg.ny) % rg.nx);
        l[gq] = (int)(va / ((size_t)rg.ny * rg.nx));
    }
    float4_t acc[CG][NT];
#pragma unroll
    for (int gq = 0; gq < CG; gq++)
#pragma unroll
        for (int nt = 0; nt < NT; nt++) acc[gq][nt] = float4_t{0.f, 0.f, 0.f, 0.f};
    const size_t Ktot = (size_t)taps * Ci;
    for (int t = 0; t < taps; t++) {
        const uint16_t *src[CG];
#pragma unroll
        for (int gq = 0; gq < CG; gq++) src[gq] = in + in_index<MODE>(r[gq], c[gq], l[gq], t, gi, go) * Ci + 4 * kg;
        const uint16_t *wsrc = Wp + (size_t)i * Ktot + (size_t)t * Ci + 4 * kg;
        for (int c0 = 0; c0 < Ci; c0 += 16) {
            short4_t a[CG];
#pragma unroll
            for (int gq = 0; gq < CG; gq++) {
                a[gq] = *(const short4_t *)(src[gq] + c0);
                if (!okA[gq]) a[gq] = short4_t{0, 0, 0, 0};
            }
#pragma unroll
            for (int nt = 0; nt < NT; nt++) {
                const short4_t bw = *(const short4_t *)(wsrc + (size_t)nt * 16 * Ktot + c0);
#pragma unroll
                for (int gq = 0; gq < CG; gq++) acc[gq][nt] = mfma16(a[gq], bw, acc[gq][nt]);
            }
        }
    }
    // D[row][col]: this lane holds column i, rows 4*kg + e of each group
#pragma unroll
    for (int gq = 0; gq < CG; gq++)
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const size_t vr = vb + 16 * gq + 4 * kg + e;
            if (vr >= V) continue;
            size_t vo = vr;
            if (MODE == CONV_PARITY) {
                const int rr = (int)(vr % rg.ny), cc = (int)((vr / rg.ny) % rg.nx),
                          ll = (int)(vr / ((size_t)rg.ny * rg.nx));
                vo = (size_t)(2 * rr + (par & 1)) + (size_t)(2 * cc + ((par >> 1) & 1)) * go.ny +
                     (size_t)(2 * ll + (par >> 2)) * go.ny * go.nx;
            }
#pragma unroll
            for (int nt = 0; nt < NT; nt++) {
                const int co = nt * 16 + i;
                float val = acc[gq][nt][e];
                if (bias) val += bias[co];
                if (skip) val += bf2f(skip[vo * Co + co]);
                if (relu) val = fmaxf(val, 0.f);
                if (OUTF) {
                    float *o = (float *)out + vo * Co + co;
                    *o = accumulate ? *o + val : val;
                } else {
                    ((uint16_t *)out)[vo * Co + co] = f2bf(val);
                }
            }
        }
}

void conv(int mode, const uint16_t *in, Grid3 gi, int Ci, const uint16_t *W, const float *bias, const uint16_t *skip,
          int relu, void *out, int out_f32, int accumulate, Grid3 go, int Co, hipStream_t s)
{
    const int taps = mode == CONV_DOWN ? 8 : (mode == CONV_PARITY ? 1 : 27);
    const Grid3 rg = mode == CONV_PARITY ? gi : go;
    const size_t V = (size_t)rg.ny * rg.nx * rg.nz;
    dim3 grid((unsigned)((V + 64 * CG - 1) / (64 * CG)), mode == CONV_PARITY ? 8 : 1);
#define CONV_CASE(M, NTv)                                                                              \
    if (mode == M && Co == NTv * 16) {                                                                 \
        if (out_f32)                                                                                   \
            k_conv<M, NTv, true><<<grid, NB, 0, s>>>(in, gi, Ci, W, taps, bias, skip, relu, out, accumulate, go); \
        else                                                                                           \
            k_conv<M, NTv, false><<<grid, NB, 0, s>>>(in, gi, Ci, W, taps, bias, skip, relu, out, accumulate, go); \
        return;                                                                                        \
    }
#define CONV_MODE(M) CONV_CASE(M, 1) CONV_CASE(M, 2) CONV_CASE(M, 3) CONV_CASE(M, 4)
    CONV_MODE(CONV_SAME)
    CONV_MODE(CONV_DOWN)
    CONV_MODE(CONV_UP)
    CONV_MODE(CONV_PARITY)
#undef CONV_MODE
#undef CONV_CASE
    fprintf(stderr, "chfem guess: no convolution kernel for mode %d, %d -> %d channels\n", mode, Ci, Co);
}

// ---------------------------------------------------------------------------
// weight gradients: block (chunk of output voxels, tap); the Co x Ci tile of
// the tap is split into 16x16 products spread over the 4 wavefronts
// ---------------------------------------------------------------------------
// output voxels per block: 2048, halved (down to 256) while the launch
// would have fewer than 2048 blocks, so the coarse levels' few voxels are
// still spread over the GPU
static int wchunk(size_t V, int taps)
{
    int c = 2048;
    while (c > 256 && ((V + c - 1) / c) * (size_t)taps < 2048) c /= 2;
    return c;
}

template <int MODE>
__global__ void __launch_bounds__(NB) k_convw(const uint16_t *__restrict__ in, Grid3 gi, int Ci,
                                              const uint16_t *__restrict__ dz, Grid3 go, int Co, int taps,
                                              int chunk, float *__restrict__ part)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int i = lane & 15, kg = lane >> 4;
    const int t = blockIdx.y;
    const size_t V = (size_t)go.ny * go.nx * go.nz;
    const size_t v0 = (size_t)blockIdx.x * chunk;
    const size_t v1 = v0 + chunk < V ? v0 + chunk : V;
    const int nci = Ci / 16, ntiles = (Co / 16) * nci;
    // With 4 tiles or more, wavefront w takes tiles w, w + 4, .. over all the
    // chunk's voxels.  With fewer, the wavefronts split the voxels as well:
    // wavefront w takes tile w % ntiles on every nsplit-th step of 16 voxels
    // from step w / ntiles, and the shares are added through LDS.
    const bool few = ntiles < 4;
    const int nsplit = few ? 4 / ntiles : 1;
    const int grp = few ? wave / ntiles : 0;
    const int wt = few ? wave % ntiles : wave;
    const int mmax = few ? 1 : 4;
    float4_t acc[4];
#pragma unroll
    for (int m = 0; m < 4; m++) acc[m] = float4_t{0.f, 0.f, 0.f, 0.f};
    if (grp < nsplit) {
        for (size_t vk = v0 + 16 * (size_t)grp; vk < v1; vk += 16 * (size_t)nsplit) {
            size_t u[4];
            bool ok[4];
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const size_t v = vk + 4 * kg + e;
                ok[e] = v < v1;
                const size_t vv = ok[e] ? v : v0;
                const int r = (int)(vv % go.ny), c = (int)((vv / go.ny) % go.nx), l = (int)(vv / ((size_t)go.ny * go.nx));
                u[e] = in_index<MODE>(r, c, l, t, gi, go);
            }
#pragma unroll
            for (int m = 0; m < 4; m++) {
                const int tt = wt + 4 * m;
                if (m >= mmax || tt >= ntiles) break;
                const int cot = tt / nci, cit = tt % nci;
                short4_t a, b;
#pragma unroll
                for (int e = 0; e < 4; e++) {
                    const size_t v = vk + 4 * kg + e;
                    a[e] = ok[e] ? (short)dz[v * Co + cot * 16 + i] : (short)0;
                    b[e] = ok[e] ? (short)in[u[e] * Ci + cit * 16 + i] : (short)0;
                }
                acc[m] = mfma16(a, b, acc[m]);
            }
        }
    }
    if (few && nsplit > 1) {
        __shared__ float4_t red[4][64];
        red[wave][lane] = acc[0];
        __syncthreads();
        if (wave >= ntiles) return;
        for (int k = 1; k < nsplit; k++) acc[0] += red[wave + ntiles * k][lane];
    }
    if (few && wave >= ntiles) return;
#pragma unroll
    for (int m = 0; m < 4; m++) {
        const int tt = wt + 4 * m;
        if (m >= mmax || tt >= ntiles) break;
        const int cot = tt / nci, cit = tt % nci;
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const int co = cot * 16 + 4 * kg + e, ci = cit * 16 + i;
            part[(((size_t)blockIdx.x * taps + t) * Co + co) * Ci + ci] = acc[m][e];
        }
    }
}

__global__ void __launch_bounds__(NB) k_wsum(const float *__restrict__ part, int nchunk, int taps, int Co, int Ci,
                                             float *__restrict__ grad)
{
    const size_t n = (size_t)taps * Co * Ci;
    for (size_t j = (size_t)blockIdx.x * NB + threadIdx.x; j < n; j += (size_t)gridDim.x * NB) {
        // j enumerates (co, t, ci) in the master layout
        const int ci = (int)(j % Ci), t = (int)((j / Ci) % taps), co = (int)(j / ((size_t)Ci * taps));
        float sum = 0.f;
        for (int ch = 0; ch < nchunk; ch++) sum += part[(((size_t)ch * taps + t) * Co + co) * Ci + ci];
        grad[j] += sum;
    }
}

size_t wpart_size(Grid3 go, int Co, int Ci, int taps)
{
    const size_t V = (size_t)go.ny * go.nx * go.nz;
    const int c = wchunk(V, taps);
    return ((V + c - 1) / c) * (size_t)taps * Co * Ci;
}

void conv_wgrad(int mode, const uint16_t *in, Grid3 gi, int Ci, const uint16_t *dz, Grid3 go, int Co, float *part,
                float *grad, hipStream_t s)
{
    const int taps = mode == CONV_DOWN ? 8 : 27;
    const size_t V = (size_t)go.ny * go.nx * go.nz;
    const int chunk = wchunk(V, taps);
    const int nchunk = (int)((V + chunk - 1) / chunk);
    dim3 grid(nchunk, taps);
    if (mode == CONV_SAME) k_convw<CONV_SAME><<<grid, NB, 0, s>>>(in, gi, Ci, dz, go, Co, taps, chunk, part);
    else if (mode == CONV_DOWN) k_convw<CONV_DOWN><<<grid, NB, 0, s>>>(in, gi, Ci, dz, go, Co, taps, chunk, part);
    else k_convw<CONV_UP><<<grid, NB, 0, s>>>(in, gi, Ci, dz, go, Co, taps, chunk, part);
    const size_t n = (size_t)taps * Co * Ci;
    k_wsum<<<(unsigned)((n + NB - 1) / NB), NB, 0, s>>>(part, nchunk, taps, Co, Ci, grad);
}

// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(NB) k_pack(const float *__restrict__ Wm, uint16_t *__restrict__ out, int Co, int Ci,
                                             int taps, int how)
{
    const size_t n = (size_t)Co * Ci * taps;
    for (size_t j = (size_t)blockIdx.x * NB + threadIdx.x; j < n; j += (size_t)gridDim.x * NB) {
        const int ci = (int)(j % Ci), t = (int)((j / Ci) % taps), co = (int)(j / ((size_t)Ci * taps));
        const uint16_t w = f2bf(Wm[j]);
        size_t o;
        if (how == 0) o = j;
        else if (how == 1) o = ((size_t)ci * taps + (taps - 1 - t)) * Co + co;
        else o = ((size_t)t * Ci + ci) * Co + co;
        out[o] = w;
    }
}

void pack_weights(const float *Wm, uint16_t *out, int Co, int Ci, int taps, int how, hipStream_t s)
{
    const size_t n = (size_t)Co * Ci * taps;
    k_pack<<<(unsigned)((n + NB - 1) / NB), NB, 0, s>>>(Wm, out, Co, Ci, taps, how);
}

// ---------------------------------------------------------------------------
// voxels per block of the activation backward (enough blocks at 64^3 and
// below, where training runs)
#define ACHUNK 256

__global__ void __launch_bounds__(NB) k_act_bwd(const float *__restrict__ dy, const uint16_t *__restrict__ y, int relu,
                                                uint16_t *__restrict__ dz, float *__restrict__ skip_grad,
                                                float *__restrict__ bpart, size_t V, int C)
{
    __shared__ float red[NB];
    const int c = threadIdx.x % C, lanes = NB / C, vl = threadIdx.x / C;
    const size_t v0 = (size_t)blockIdx.x * ACHUNK, v1 = v0 + ACHUNK < V ? v0 + ACHUNK : V;
    float bsum = 0.f;
    if (vl < lanes) {
        for (size_t v = v0 + vl; v < v1; v += lanes) {
            const size_t j = v * C + c;
            float g = dy[j];
            if (relu && !(bf2f(y[j]) > 0.f)) g = 0.f;
            const uint16_t gb = f2bf(g);
            dz[j] = gb;
            if (skip_grad) skip_grad[j] += g;
            bsum += bf2f(gb);
        }
    }
    red[threadIdx.x] = (vl < lanes) ? bsum : 0.f;
    __syncthreads();
    if (threadIdx.x < C) {
        float s = 0.f;
        for (int k = 0; k < lanes; k++) s += red[k * C + threadIdx.x];
        bpart[(size_t)blockIdx.x * C + threadIdx.x] = s;
    }
}

__global__ void k_bsum(const float *__restrict__ bpart, int nb, int C, float *__restrict__ bias_grad)
{
    const int c = threadIdx.x;
    if (c >= C) return;
    float s = 0.f;
    for (int b = 0; b < nb; b++) s += bpart[(size_t)b * C + c];
    bias_grad[c] += s;
}

size_t bias_part_size(size_t V, int C) { return ((V + ACHUNK - 1) / ACHUNK) * (size_t)C; }

void act_backward(const float *dy, const uint16_t *y, int relu, uint16_t *dz, float *skip_grad, float *bias_part,
                  float *bias_grad, size_t V, int C, hipStream_t s)
{
    const int nb = (int)((V + ACHUNK - 1) / ACHUNK);
    k_act_bwd<<<nb, NB, 0, s>>>(dy, y, relu, dz, skip_grad, bias_part, V, C);
    k_bsum<<<1, 64, 0, s>>>(bias_part, nb, C, bias_grad);
}

// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(NB) k_pool(const float *__restrict__ fine, Grid3 gf, float *__restrict__ coarse, int C,
                                             int accumulate)
{
    const int ny = gf.ny / 2, nx = gf.nx / 2, nz = gf.nz / 2;
    const size_t n = (size_t)ny * nx * nz * C;
    for (size_t j = (size_t)blockIdx.x * NB + threadIdx.x; j < n; j += (size_t)gridDim.x * NB) {
        const int c = (int)(j % C);
        const size_t w = j / C;
        const int r = (int)(w % ny), cc = (int)((w / ny) % nx), l = (int)(w / ((size_t)ny * nx));
        float s = 0.f;
#pragma unroll
        for (int t = 0; t < 8; t++) {
            const size_t u = (size_t)(2 * r + (t & 1)) + (size_t)(2 * cc + ((t >> 1) & 1)) * gf.ny +
                             (size_t)(2 * l + (t >> 2)) * gf.ny * gf.nx;
            s += fine[u * C + c];
        }
        coarse[j] = accumulate ? coarse[j] + s : s;
    }
}

void pool_children(const float *fine, Grid3 gf, float *coarse, int C, int accumulate, hipStream_t s)
{
    const size_t n = (size_t)(gf.ny / 2) * (gf.nx / 2) * (gf.nz / 2) * C;
    unsigned b = (unsigned)((n + NB - 1) / NB);
    if (b > 65536) b = 65536;
    k_pool<<<b, NB, 0, s>>>(fine, gf, coarse, C, accumulate);
}

// ---------------------------------------------------------------------------
#define LCHUNK 512

__global__ void __launch_bounds__(NB) k_mse(const float *__restrict__ y, int Cp, const float *__restrict__ target, int nt,
                                            size_t V, double scale, float *__restrict__ dy, double *__restrict__ part)
{
    __shared__ double red[NB];
    const size_t j0 = (size_t)blockIdx.x * LCHUNK * Cp, j1 = j0 + (size_t)LCHUNK * Cp < V * Cp ? j0 + (size_t)LCHUNK * Cp : V * Cp;
    double acc = 0;
    for (size_t j = j0 + threadIdx.x; j < j1; j += NB) {
        const int c = (int)(j % Cp);
        const size_t v = j / Cp;
        float g = 0.f;
        if (c < nt) {
            const float d = y[j] - target[v * nt + c];
            acc += (double)d * (double)d;
            g = (float)(2.0 * scale) * d;
        }
        dy[j] = g;
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = NB / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0] * scale;
}

__global__ void k_dsum(const double *__restrict__ part, int nb, double *__restrict__ out)
{
    if (threadIdx.x != 0) return;
    double s = 0;
    for (int b = 0; b < nb; b++) s += part[b];
    *out += s;
}

size_t loss_part_size(size_t V) { return (V + LCHUNK - 1) / LCHUNK; }

void mse(const float *y, int Cp, const float *target, int nt, size_t V, double scale, float *dy, double *part,
         double *loss, hipStream_t s)
{
    const int nb = (int)loss_part_size(V);
    k_mse<<<nb, NB, 0, s>>>(y, Cp, target, nt, V, scale, dy, part);
    k_dsum<<<1, 64, 0, s>>>(part, nb, loss);
}

// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(NB) k_adam(float *__restrict__ w, const float *__restrict__ g, float *__restrict__ m,
                                             float *__restrict__ v, size_t n, float lr, float b1, float b2, float eps,
                                             float c1, float c2, float gscale)
{
    for (size_t j = (size_t)blockIdx.x * NB + threadIdx.x; j < n; j += (size_t)gridDim.x * NB) {
        const float gj = g[j] * gscale;
        const float mj = b1 * m[j] + (1.f - b1) * gj;
        const float vj = b2 * v[j] + (1.f - b2) * gj * gj;
        m[j] = mj;
        v[j] = vj;
        w[j] -= lr * (mj / c1) / (sqrtf(vj / c2) + eps);
    }
}

void adam(float *w, const float *g, float *m, float *v, size_t n, float lr, float b1, float b2, float eps, int step,
          float gscale, hipStream_t s)
{
    const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
    unsigned b = (unsigned)((n + NB - 1) / NB);
    if (b > 4096) b = 4096;
    k_adam<<<b, NB, 0, s>>>(w, g, m, v, n, lr, b1, b2, eps, c1, c2, gscale);
}

// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(NB) k_features(const uint16_t *__restrict__ elem, const float *__restrict__ table,
                                                 int nf, uint16_t *__restrict__ in, int C, size_t V)
{
    const size_t n = V * C;
    for (size_t j = (size_t)blockIdx.x * NB + threadIdx.x; j < n; j += (size_t)gridDim.x * NB) {
        const int c = (int)(j % C);
        const size_t v = j / C;
        in[j] = c < nf ? f2bf(table[(size_t)elem[v] * nf + c]) : (uint16_t)0;
    }
}

void features(const uint16_t *elem, const float *table, int nf, uint16_t *in, int C, size_t V, hipStream_t s)
{
    unsigned b = (unsigned)((V * C + NB - 1) / NB);
    if (b > 65536) b = 65536;
    k_features<<<b, NB, 0, s>>>(elem, table, nf, in, C, V);
}

// load cases (seam axis, jumping component): X, Y, Z, YZ, XZ, XY
__constant__ int cLoadAxis[6] = {0, 1, 2, 1, 2, 1};
__constant__ int cLoadComp[6] = {0, 1, 2, 2, 0, 0};

__global__ void __launch_bounds__(NB) k_guess_x0(const float *__restrict__ y, int Cp, int load, int nd, Grid3 g,
                                                 float scale, double *__restrict__ x0)
{
    const size_t N = (size_t)g.ny * g.nx * g.nz;
    const int axis = cLoadAxis[load], comp = nd == 3 ? cLoadComp[load] : 0;
    for (size_t i = (size_t)blockIdx.x * NB + threadIdx.x; i < N; i += (size_t)gridDim.x * NB) {
        const int r = (int)(i % g.ny), c = (int)((i / g.ny) % g.nx), l = (int)(i / ((size_t)g.ny * g.nx));
        // total field of the homogeneous medium: ramp with the jump on the seam
        const double ramp = axis == 0 ? (double)c : (axis == 1 ? (double)((g.ny - r) % g.ny) : (double)((g.nz - l) % g.nz));
        for (int a = 0; a < nd; a++)
            x0[nd * i + a] = (a == comp ? ramp : 0.0) + (double)(scale * y[i * Cp + load * nd + a]);
    }
}

void guess_to_x0(const float *y, int Cp, int load, int nd, Grid3 g, float scale, double *x0, hipStream_t s)
{
    const size_t N = (size_t)g.ny * g.nx * g.nz;
    unsigned b = (unsigned)((N + NB - 1) / NB);
    if (b > 65536) b = 65536;
    k_guess_x0<<<b, NB, 0, s>>>(y, Cp, load, nd, g, scale, x0);
}

struct X0Set {
    double *x[6];
};

__global__ void __launch_bounds__(NB) k_guess_x0_all(const float *__restrict__ y, int Cp, int nd, Grid3 g, float scale,
                                                     X0Set xs)
{
    const size_t N = (size_t)g.ny * g.nx * g.nz;
    const int nloads = nd == 3 ? 6 : 3;
    for (size_t i = (size_t)blockIdx.x * NB + threadIdx.x; i < N; i += (size_t)gridDim.x * NB) {
        const int r = (int)(i % g.ny), c = (int)((i / g.ny) % g.nx), l = (int)(i / ((size_t)g.ny * g.nx));
        const double ramps[3] = {(double)c, (double)((g.ny - r) % g.ny), (double)((g.nz - l) % g.nz)};
        const float *yi = y + i * Cp;
        for (int load = 0; load < nloads; load++) {
            double *x0 = xs.x[load];
            if (!x0) continue;
            const int comp = nd == 3 ? cLoadComp[load] : 0;
            const double ramp = ramps[cLoadAxis[load]];
            for (int a = 0; a < nd; a++)
                x0[nd * i + a] = (a == comp ? ramp : 0.0) + (double)(scale * yi[load * nd + a]);
        }
    }
}

void guess_to_x0_all(const float *y, int Cp, int nd, Grid3 g, float scale, double *const x0[6], hipStream_t s)
{
    const size_t N = (size_t)g.ny * g.nx * g.nz;
    unsigned b = (unsigned)((N + NB - 1) / NB);
    if (b > 65536) b = 65536;
    X0Set xs;
    for (int k = 0; k < 6; k++) xs.x[k] = x0[k];
    k_guess_x0_all<<<b, NB, 0, s>>>(y, Cp, nd, g, scale, xs);
}

__global__ void k_fill_f32(float *p, float v, size_t n)
{
    for (size_t j = (size_t)blockIdx.x * NB + threadIdx.x; j < n; j += (size_t)gridDim.x * NB) p[j] = v;
}
__global__ void k_add_f32(float *d, const float *src, size_t n)
{
    for (size_t j = (size_t)blockIdx.x * NB + threadIdx.x; j < n; j += (size_t)gridDim.x * NB) d[j] += src[j];
}
void fill_f32(float *p, float v, size_t n, hipStream_t s)
{
    unsigned b = (unsigned)((n + NB - 1) / NB);
    if (b > 65536) b = 65536;
    if (b) k_fill_f32<<<b, NB, 0, s>>>(p, v, n);
}
void add_f32(float *dst, const float *src, size_t n, hipStream_t s)
{
    unsigned b = (unsigned)((n + NB - 1) / NB);
    if (b > 65536) b = 65536;
    if (b) k_add_f32<<<b, NB, 0, s>>>(dst, src, n);
}

__global__ void k_mfma_check(const uint16_t *__restrict__ A, const uint16_t *__restrict__ B, float *__restrict__ D)
{
    const int lane = threadIdx.x, i = lane & 15, kg = lane >> 4;
    short4_t a, b;
#pragma unroll
    for (int e = 0; e < 4; e++) {
        a[e] = (short)A[i * 16 + 4 * kg + e];  // A[row i][k]
        b[e] = (short)B[(4 * kg + e) * 16 + i];  // B[k][col i]
    }
    const float4_t d = mfma16(a, b, float4_t{0.f, 0.f, 0.f, 0.f});
#pragma unroll
    for (int e = 0; e < 4; e++) D[(4 * kg + e) * 16 + i] = d[e];
}

void mfma_check(const uint16_t *A, const uint16_t *B, float *D, hipStream_t s)
{
    k_mfma_check<<<1, 64, 0, s>>>(A, B, D);
}

}  // namespace nn
}  // namespace chfem
