// ebe_mfma.hip -- element-by-element operator q = A v for thermal (8x8) and
// elastic (24x24) voxel elements on the matrix cores, for any dense local
// matrices of up to 4 materials (chfem/src/femhmg/cudapcg/kernels/
// matrix_free.cu:1097-1335 computes the same products one element per
// thread and scatters them with atomics).
//
// Per element layer a block of TR x TC work-items (4 wavefronts for 16x16
// or 32x8) owns TR x TC elements.  FP64 (and FP32 with CHFEM_EBE=mfma16): each
// wavefront groups its 64 elements by material into column tiles of 16 and
// multiplies each tile by its material's local matrix (KE x KE) with
// v_mfma_{f32,f64}_16x16x4: row tiles of 16 local DOFs (24 rows fill two),
// 4 DOFs of the reduction per step.  FP32 by default: v_mfma_f32_4x4x1
// runs 16 independent 4x4x1 products per instruction; the wavefront groups
// its elements by material into groups of 4 (one per 4-lane block), and
// each block multiplies its group by 4 rows of its material's matrix, one
// reduction DOF per instruction: no row padding, and only the groups past
// the 16th (at most one per extra material) cost more instructions.
// Element forces land in LDS and every interior node sums its 8 element
// contributions there, as in the Walsh-basis kernel (walsh.hip): tiles
// overlap by one node row/column, nothing is scattered with atomics, and the
// result is deterministic.
#include <hip/hip_runtime.h>
#include <stdlib.h>
#include <algorithm>
#include <string.h>
#include "kernels.h"
#include "reduce.h"

namespace chfem {

typedef float f4_t __attribute__((ext_vector_type(4)));
typedef double d4_t __attribute__((ext_vector_type(4)));

template <typename T> struct MfmaAcc;
// Operand layout of both products: lane l supplies A[l % 16][l / 16] and
// B[l / 16][l % 16].  Result item i of lane l holds D[row(l / 16, i)][l % 16]
// with row = 4 (l/16) + i for FP32 and (l/16) + 4 i for FP64 (checked by
// chfemMfmaCheckF).
template <> struct MfmaAcc<float> {
    typedef f4_t type;
    static __device__ __forceinline__ type mac(float a, float b, type c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int kg, int i) { return 4 * kg + i; }
};
template <> struct MfmaAcc<double> {
    typedef d4_t type;
    static __device__ __forceinline__ type mac(double a, double b, type c)
    {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int kg, int i) { return kg + 4 * i; }
};

#define EDR(k) ((k) == 0 || (k) == 1 || (k) == 4 || (k) == 5 ? 1 : 0)
#define EDC(k) ((k) == 1 || (k) == 2 || (k) == 5 || (k) == 6 ? 1 : 0)
#define EDL(k) ((k) >= 4 ? 1 : 0)

__device__ __forceinline__ int ewrap(int a, int n) { return a < 0 ? a + n : (a >= n ? a - n : a); }
__device__ __forceinline__ int edr(int k) { return (0x33 >> k) & 1; }
__device__ __forceinline__ int edc(int k) { return (0x66 >> k) & 1; }
__device__ __forceinline__ int edl(int k) { return k >> 2; }

template <typename T, int ND, int TR, int TC, bool DOT, bool B4>
__global__ void __launch_bounds__(TR *TC) k_ebe_mfma(const T *__restrict__ v, T *__restrict__ q,
                                                      const uint16_t *__restrict__ keys, const T *__restrict__ K,
                                                      int nmat, int ebe, int ny, int nx, int nz, int LZ, Reduce red,
                                                      Sdf sdf)
{
    static_assert(TR * TC == 256, "four wavefronts per block");
    static_assert(!B4 || sizeof(T) == 4, "the 4x4 block form is FP32");
    if (DOT && stopped(red)) return;
    constexpr int KE = 8 * ND;                 // local DOFs
    constexpr int MT = (KE + 15) / 16;         // row tiles
    constexpr int NT = TR * TC, NR = TR + 1, NC = TC + 1, NV = NR * NC * ND;
    constexpr int SLOTS = (NV + NT - 1) / NT;
    constexpr int SF = KE + 1;                 // padded element stride of the force buffer
    // padded row stride of the local matrices: the A operand's 16 rows per
    // step then fall on distinct LDS banks (a stride of KE = 24 or 8 put
    // every fourth row on the same bank)
    // 4x4 block form: rows at stride KE (16-byte row reads of 4 reduction
    // DOFs), materials at KE*KE + 4 so that two materials' rows read by one
    // instruction fall on different banks
    constexpr int KP = B4 ? KE : KE + 1;
    constexpr int MS = B4 ? KE * KE + 4 : KE * KP;
    __shared__ T su[2][NV];
    __shared__ T sf[NT * SF];
    __shared__ __align__(16) T sK[4 * MS];     // K[m][row][col], row stride KP, material stride MS
    __shared__ int smat[NT];
    // per wavefront: block element at each position of the material-grouped
    // column tiles (at most 64 elements in at most 7 tiles)
    __shared__ int sperm[NT / 64][112];
    __shared__ T sesc[NT];  // scalar density field: element scale
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6, li = lane & 15, kg = lane >> 4;
    const int tr = tid % TR, tc = tid / TR;
    const int r0 = blockIdx.x * (TR - 1), c0 = blockIdx.y * (TC - 1), l0 = blockIdx.z * LZ;
    const int vr = min(TR - 1, ny - r0), vc = min(TC - 1, nx - c0), vl = min(LZ, nz - l0);
    const size_t nxy = (size_t)nx * ny;
    for (int i = tid; i < nmat * KE * KE; i += NT) {
        const int m = i / (KE * KE), rc = i % (KE * KE);
        sK[m * MS + (rc / KE) * KP + rc % KE] = K[i];
    }

    uint32_t goff[SLOTS];
    bool gval[SLOTS];
#pragma unroll
    for (int k = 0; k < SLOTS; k++) {
        const int idx = tid + k * NT;
        const int comp = idx % ND, t2 = idx / ND, nr = t2 % NR, nc = t2 / NR;
        gval[k] = idx < NV && nr <= vr + 1 && nc <= vc + 1;
        const int gr = ewrap(r0 - 1 + nr, ny), gc = ewrap(c0 - 1 + nc, nx);
        goff[k] = gval[k] ? (uint32_t)ND * ((uint32_t)gc * (uint32_t)ny + (uint32_t)gr) + (uint32_t)comp : 0u;
    }
    T pre[SLOTS];
    auto fetch = [&](int lg) {
        const T *base = v + (size_t)ND * (size_t)ewrap(lg, nz) * nxy;
#pragma unroll
        for (int k = 0; k < SLOTS; k++) pre[k] = gval[k] ? base[goff[k]] : T(0);
    };
    auto store = [&](T *dst) {
#pragma unroll
        for (int k = 0; k < SLOTS; k++)
            if (gval[k]) dst[tid + k * NT] = pre[k];
    };
    const bool elem_ok = tr <= vr && tc <= vc;
    const uint32_t eoff = (uint32_t)ewrap(c0 - 1 + tc, nx) * (uint32_t)ny + (uint32_t)ewrap(r0 - 1 + tr, ny);
    // material of this work-item's element in layer le; -1: outside the patch
    auto material = [&](int le) -> int {
        if (!elem_ok) return -1;
        const unsigned kv = keys[(size_t)ewrap(le, nz) * nxy + eoff];
        return ebe ? (int)(kv & 3u) : (int)((kv >> 6) & 3u);
    };
    auto escale = [&](int le) -> T {
        return (sdf.d && elem_ok) ? (T)sdf_scale(sdf, (size_t)ewrap(le, nz) * nxy + eoff) : T(1);
    };

    // element forces of the wavefront's 64 elements (block elements 64*wave ..)
    // The elements are grouped by material into column tiles of 16: material
    // m's c_m elements fill ceil(c_m / 16) tiles (lane order), so every
    // element is multiplied by its own material's matrix only.  sperm holds,
    // per wavefront, the block element at each tile position.
    auto elements = [&](const T *cur, const T *nxt) {
        const int me = smat[tid];
        const unsigned long long below = (1ull << lane) - 1ull;
        int cnt[4], tb[4];
        int ntiles = 0;
#pragma unroll
        for (int m = 0; m < 4; m++) {
            const unsigned long long bal = __ballot(me == m);
            cnt[m] = __popcll(bal);
            tb[m] = ntiles;
            if (me == m) sperm[wave][16 * ntiles + __popcll(bal & below)] = tid;
            ntiles += (cnt[m] + 15) >> 4;
        }
        __syncthreads();  // sperm written before any lane reads it
#pragma unroll
        for (int m = 0; m < 4; m++) {
            if (!cnt[m]) continue;
            const T *Km = sK + m * MS;
            for (int t = 0; t < ((cnt[m] + 15) >> 4); t++) {
                // column li of this tile: the material's element 16 t + li, if any
                const bool valid = 16 * t + li < cnt[m];
                const int e = valid ? sperm[wave][16 * (tb[m] + t) + li] : 0;
                const int eb = ((e / TR) * NR + (e % TR)) * ND;
                const T esc = valid ? sesc[e] : T(0);
                typename MfmaAcc<T>::type acc[MT];
#pragma unroll
                for (int mt = 0; mt < MT; mt++) acc[mt] = typename MfmaAcc<T>::type{0, 0, 0, 0};
#pragma unroll
                for (int s = 0; s < KE / 4; s++) {
                    const int d = 4 * s + kg;           // local DOF of this lane in the step
                    const int kn = d / ND, comp = d % ND;
                    const T *pl = edl(kn) ? nxt : cur;
                    const int doff = (edc(kn) * NR + edr(kn)) * ND + comp;
                    const T b = esc * pl[eb + doff];
#pragma unroll
                    for (int mt = 0; mt < MT; mt++) {
                        const int row = 16 * mt + li;
                        const T a = row < KE ? Km[row * KP + d] : T(0);
                        acc[mt] = MfmaAcc<T>::mac(a, b, acc[mt]);
                    }
                }
                // D[row][col]: this lane holds column li (element e) and 4 rows
                // of each row tile
                if (valid) {
                    T *dst = sf + e * SF;
#pragma unroll
                    for (int mt = 0; mt < MT; mt++)
#pragma unroll
                        for (int i = 0; i < 4; i++) {
                            const int row = 16 * mt + MfmaAcc<T>::row(kg, i);
                            if (row < KE) dst[row] = acc[mt][i];
                        }
                }
            }
        }
    };
    // 4x4 block form (FP32).  Material m's c_m elements of the wavefront fill
    // groups gb[m] .. gb[m] + ceil(c_m / 4) - 1 (lane order; sperm[wave][4 g +
    // j] is the block element in column j of group g).  Lane l = 4 p + j
    // takes column j of group p: it supplies row 4 rb + j of group p's
    // material matrix (A) and its element's local DOF values (B), and
    // receives rows 4 rb .. 4 rb + 3 of its element's forces.  Groups past the
    // 16th are spread as (group, row block) jobs over the 16 blocks, in as
    // few further rounds as they need.
    auto elements4 = [&](const T *cur, const T *nxt) {
        constexpr int NRB = KE / 4;  // row blocks of 4 local DOFs
        const int me = smat[tid];
        const unsigned long long below = (1ull << lane) - 1ull;
        int cnt[4], gb[4];
        int G = 0;
#pragma unroll
        for (int m = 0; m < 4; m++) {
            const unsigned long long bal = __ballot(me == m);
            cnt[m] = __popcll(bal);
            gb[m] = G;
            if (me == m) sperm[wave][4 * G + __popcll(bal & below)] = tid;
            G += (cnt[m] + 3) >> 2;
        }
        __syncthreads();  // sperm written before any lane reads it
        const int p = lane >> 2, j = lane & 3;
        // material and element of column j of group g (g < G: the materials
        // with elements cover 0 .. G - 1 in order)
        auto locate = [&](int g, int &m, int &e) -> bool {
            m = 0;
#pragma unroll
            for (int k = 1; k < 4; k++)
                if (cnt[k] && g >= gb[k]) m = k;
            const bool ok = g < G && 4 * (g - gb[m]) + j < cnt[m];
            e = ok ? sperm[wave][4 * g + j] : 0;
            return ok;
        };
        auto gather = [&](int e, bool ok, T (&u)[KE]) {
            const int eb = ((e / TR) * NR + (e % TR)) * ND;
            const T esc = ok ? sesc[e] : T(0);
#pragma unroll
            for (int d = 0; d < KE; d++) {
                const int kn = d / ND, comp = d % ND;
                const T *pl = edl(kn) ? nxt : cur;
                u[d] = esc * pl[eb + (edc(kn) * NR + edr(kn)) * ND + comp];
            }
        };
        auto put = [&](int e, int rb, f4_t acc) {
            T *dst = sf + e * SF + 4 * rb;
#pragma unroll
            for (int i = 0; i < 4; i++) dst[i] = acc[i];
        };
        {
            int m, e;
            const bool ok = locate(p, m, e);
            T u[KE];
            gather(e, ok, u);
            const T *Kb = sK + m * MS + j * KP;
#pragma unroll
            for (int rb = 0; rb < NRB; rb += 2) {
                // two row blocks at a time: independent accumulation chains
                const T *K0 = Kb + 4 * rb * KP, *K1 = K0 + 4 * KP;
                f4_t a0 = {0, 0, 0, 0}, a1 = {0, 0, 0, 0};
#pragma unroll
                for (int k4 = 0; k4 < KE / 4; k4++) {
                    const f4_t x0 = *(const f4_t *)(K0 + 4 * k4), x1 = *(const f4_t *)(K1 + 4 * k4);
#pragma unroll
                    for (int kk = 0; kk < 4; kk++) {
                        a0 = __builtin_amdgcn_mfma_f32_4x4x1f32(x0[kk], u[4 * k4 + kk], a0, 0, 0, 0);
                        a1 = __builtin_amdgcn_mfma_f32_4x4x1f32(x1[kk], u[4 * k4 + kk], a1, 0, 0, 0);
                    }
                }
                if (ok) {
                    put(e, rb, a0);
                    put(e, rb + 1, a1);
                }
            }
        }
        const int O = (G - 16) * NRB;  // row-block jobs of the groups past the 16th
        for (int x = 0; 16 * x < O; x++) {
            const int qj = 16 * x + p;
            const int rb = qj % NRB;
            int m = 0, e = 0;
            const bool ok = qj < O && locate(16 + qj / NRB, m, e);
            T u[KE];
            gather(e, ok, u);
            const T *Kr = sK + m * MS + (4 * rb + j) * KP;
            f4_t a0 = {0, 0, 0, 0};
#pragma unroll
            for (int k4 = 0; k4 < KE / 4; k4++) {
                const f4_t x0 = *(const f4_t *)(Kr + 4 * k4);
#pragma unroll
                for (int kk = 0; kk < 4; kk++)
                    a0 = __builtin_amdgcn_mfma_f32_4x4x1f32(x0[kk], u[4 * k4 + kk], a0, 0, 0, 0);
            }
            if (ok) put(e, rb, a0);
        }
    };
    const bool own = tr >= 1 && tc >= 1 && tr - 1 < vr && tc - 1 < vc;
    auto fslot = [&](int k) { return ((tc - EDC(k)) * TR + (tr - EDR(k))) * SF + ND * k; };

    T *cur = su[0], *nxt = su[1];
    fetch(l0 - 1);
    store(cur);
    fetch(l0);
    store(nxt);
    smat[tid] = material(l0 - 1);
    sesc[tid] = escale(l0 - 1);
    __syncthreads();
    fetch(l0 + 1);
    if constexpr (B4) elements4(cur, nxt);
    else elements(cur, nxt);
    const int mnext = material(l0);
    const T enext = escale(l0);
    __syncthreads();
    smat[tid] = mnext;
    sesc[tid] = enext;
    T acc[ND];
#pragma unroll
    for (int a = 0; a < ND; a++) acc[a] = 0;
    if (own) {
#pragma unroll
        for (int k = 4; k < 8; k++) {
            const T *sp = sf + fslot(k);
#pragma unroll
            for (int a = 0; a < ND; a++) acc[a] += sp[a];
        }
    }
    store(cur);
    {
        T *t = cur; cur = nxt; nxt = t;
    }
    __syncthreads();
    double dsum = 0.0;
    for (int s = 0; s < vl; s++) {
        const int l = l0 + s;
        const bool more = s + 1 < vl;
        if (more) fetch(l + 2);
        T own_u[ND];
#pragma unroll
        for (int a = 0; a < ND; a++) own_u[a] = 0;
        if (DOT && own) {
            const T *uu = cur + (tc * NR + tr) * ND;
#pragma unroll
            for (int a = 0; a < ND; a++) own_u[a] = uu[a];
        }
        if constexpr (B4) elements4(cur, nxt);
        else elements(cur, nxt);
        const int mn = more ? material(l + 1) : -1;
        const T en = more ? escale(l + 1) : T(1);
        __syncthreads();
        if (own) {
            T o[ND];
#pragma unroll
            for (int a = 0; a < ND; a++) o[a] = acc[a];
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const T *sp = sf + fslot(k);
#pragma unroll
                for (int a = 0; a < ND; a++) o[a] += sp[a];
            }
#pragma unroll
            for (int a = 0; a < ND; a++) acc[a] = 0;
#pragma unroll
            for (int k = 4; k < 8; k++) {
                const T *sp = sf + fslot(k);
#pragma unroll
                for (int a = 0; a < ND; a++) acc[a] += sp[a];
            }
            T *qo = q + ND * ((size_t)l * nxy + (size_t)(c0 + tc - 1) * ny + (r0 + tr - 1));
#pragma unroll
            for (int a = 0; a < ND; a++) {
                qo[a] = o[a];
                if (DOT) dsum += (double)own_u[a] * (double)o[a];
            }
        }
        if (more) {
            store(cur);
            T *t = cur; cur = nxt; nxt = t;
            smat[tid] = mn;
            sesc[tid] = en;
        }
        __syncthreads();
    }
    if (DOT) {
        double vv[1] = {dsum};
        store_partials<1>(red, vv);
    }
}

template <typename T, int TR, int TC>
static void ebe_mfma_tiles(const OpDesc &op, const T *v, T *q, const Reduce *red, hipStream_t s)
{
    Reduce rr = {nullptr, nullptr, 0, 0, 0};
    if (red) rr = *red;
    // element layers per block (each block also evaluates the layer below
    // its first one): 32 for FP64 elastic, 16 otherwise (at 256^3: FP64
    // elastic 1.660 -> 1.627 ms with 32; FP32 elastic the same; conduction
    // 1-4 % slower with 32, 4-10 % with 64); CHFEM_EBE_LZ overrides
    const char *lz_env = getenv("CHFEM_EBE_LZ");
    int LZ = lz_env ? std::max(1, atoi(lz_env)) : (sizeof(T) == 8 && op.physics == PHY_ELASTIC ? 32 : 16);
    const long cols = (long)((op.ny + TR - 2) / (TR - 1)) * ((op.nx + TC - 2) / (TC - 1));
    while (cols * ((op.nz + LZ - 1) / LZ) > MAX_PARTIAL_BLOCKS) LZ *= 2;
    dim3 grid((op.ny + TR - 2) / (TR - 1), (op.nx + TC - 2) / (TC - 1), (op.nz + LZ - 1) / LZ);
    const T *K = (const T *)op.K;
#define EBE_LAUNCH(ND_, DOT_, B4_)                                                                                  \
    k_ebe_mfma<T, ND_, TR, TC, DOT_, B4_><<<grid, TR * TC, 0, s>>>(v, q, op.keys, K, op.nmat, op.ebe, op.ny, op.nx, \
                                                                   op.nz, LZ, rr, op.sdf)
    if constexpr (sizeof(T) == 4) {
        if (!op.ebe_tiles16) {
            if (op.physics == PHY_ELASTIC) {
                if (red) EBE_LAUNCH(3, true, true);
                else EBE_LAUNCH(3, false, true);
            } else {
                if (red) EBE_LAUNCH(1, true, true);
                else EBE_LAUNCH(1, false, true);
            }
            if (red) launch_reduce_finish(rr, grid.x * grid.y * grid.z, 1, s);
            return;
        }
    }
    if (op.physics == PHY_ELASTIC) {
        if (red) EBE_LAUNCH(3, true, false);
        else EBE_LAUNCH(3, false, false);
    } else {
        if (red) EBE_LAUNCH(1, true, false);
        else EBE_LAUNCH(1, false, false);
    }
#undef EBE_LAUNCH
    if (red) launch_reduce_finish(rr, grid.x * grid.y * grid.z, 1, s);
}

// element tiles of 16 x 16 (a wavefront: 16 rows x 4 columns), or 32 x 8
// (32 rows x 2 columns) with CHFEM_EBE_TILE=32x8.  At 256^3 the square tile
// evaluates 1.27 element slots per element against 1.30 (halo and grid
// tail), its compact wavefront patches hold fewer two-material groups, and
// it ran 3 % faster in both precisions
template <typename T>
void launch_ebe_mfma(const OpDesc &op, const T *v, T *q, const Reduce *red, hipStream_t s)
{
    const char *e = getenv("CHFEM_EBE_TILE");
    if (e && !strcmp(e, "32x8")) ebe_mfma_tiles<T, 32, 8>(op, v, q, red, s);
    else ebe_mfma_tiles<T, 16, 16>(op, v, q, red, s);
}

template void launch_ebe_mfma<float>(const OpDesc &, const float *, float *, const Reduce *, hipStream_t);
template void launch_ebe_mfma<double>(const OpDesc &, const double *, double *, const Reduce *, hipStream_t);

// lane-layout check of the FP32 / FP64 16x16x4 matrix-core products used
// above: D = A (16x4) * B (4x16), row-major inputs and output
__global__ void k_mfma_check_f(const float *A, const float *B, float *D, const double *A2, const double *B2, double *D2)
{
    const int l = threadIdx.x, li = l & 15, kg = l >> 4;
    f4_t d = __builtin_amdgcn_mfma_f32_16x16x4f32(A[li * 4 + kg], B[kg * 16 + li], f4_t{0, 0, 0, 0}, 0, 0, 0);
    d4_t d2 = __builtin_amdgcn_mfma_f64_16x16x4f64(A2[li * 4 + kg], B2[kg * 16 + li], d4_t{0, 0, 0, 0}, 0, 0, 0);
#pragma unroll
    for (int i = 0; i < 4; i++) {
        D[MfmaAcc<float>::row(kg, i) * 16 + li] = d[i];
        D2[MfmaAcc<double>::row(kg, i) * 16 + li] = d2[i];
    }
}

// lane-layout check of the FP32 4x4x1 16-block product: lane l supplies
// A[l] and B[l] and receives its 4 result items at D[4 l .. 4 l + 3]
__global__ void k_mfma_check4(const float *A, const float *B, float *D)
{
    const int l = threadIdx.x;
    const f4_t d = __builtin_amdgcn_mfma_f32_4x4x1f32(A[l], B[l], f4_t{0, 0, 0, 0}, 0, 0, 0);
#pragma unroll
    for (int i = 0; i < 4; i++) D[4 * l + i] = d[i];
}

}  // namespace chfem

extern "C" int chfemMfmaCheck4(const float *A, const float *B, float *D)
{
    float *dA = nullptr, *dB = nullptr, *dD = nullptr;
    int ok = hipMalloc(&dA, 64 * 4) == hipSuccess && hipMalloc(&dB, 64 * 4) == hipSuccess &&
             hipMalloc(&dD, 256 * 4) == hipSuccess;
    ok = ok && hipMemcpy(dA, A, 64 * 4, hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(dB, B, 64 * 4, hipMemcpyHostToDevice) == hipSuccess;
    if (ok) {
        chfem::k_mfma_check4<<<1, 64>>>(dA, dB, dD);
        ok = hipDeviceSynchronize() == hipSuccess && hipMemcpy(D, dD, 256 * 4, hipMemcpyDeviceToHost) == hipSuccess;
    }
    if (dA) (void)hipFree(dA);
    if (dB) (void)hipFree(dB);
    if (dD) (void)hipFree(dD);
    return ok;
}

extern "C" int chfemMfmaCheckF(const float *A, const float *B, float *D, const double *A2, const double *B2,
                               double *D2)
{
    float *dA = nullptr, *dB = nullptr, *dD = nullptr;
    double *dA2 = nullptr, *dB2 = nullptr, *dD2 = nullptr;
    int ok = hipMalloc(&dA, 64 * 4) == hipSuccess && hipMalloc(&dB, 64 * 4) == hipSuccess &&
             hipMalloc(&dD, 256 * 4) == hipSuccess && hipMalloc(&dA2, 64 * 8) == hipSuccess &&
             hipMalloc(&dB2, 64 * 8) == hipSuccess && hipMalloc(&dD2, 256 * 8) == hipSuccess;
    ok = ok && hipMemcpy(dA, A, 64 * 4, hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(dB, B, 64 * 4, hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(dA2, A2, 64 * 8, hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(dB2, B2, 64 * 8, hipMemcpyHostToDevice) == hipSuccess;
    if (ok) {
        chfem::k_mfma_check_f<<<1, 64>>>(dA, dB, dD, dA2, dB2, dD2);
        ok = hipDeviceSynchronize() == hipSuccess && hipMemcpy(D, dD, 256 * 4, hipMemcpyDeviceToHost) == hipSuccess &&
             hipMemcpy(D2, dD2, 256 * 8, hipMemcpyDeviceToHost) == hipSuccess;
    }
    if (dA) (void)hipFree(dA);
    if (dB) (void)hipFree(dB);
    if (dD) (void)hipFree(dD);
    if (dA2) (void)hipFree(dA2);
    if (dB2) (void)hipFree(dB2);
    if (dD2) (void)hipFree(dD2);
    return ok;
}
