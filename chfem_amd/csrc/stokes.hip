// stokes.hip -- the stabilised Stokes operator on the voxel grid, element by
// element in the Walsh basis, over the same layer-march tiles as the
// elastic/conduction operators (walsh.h).
//
// Element matrix (unit voxel, one fluid, chfem/src/femhmg/physical_phenomena/
// femhmg_fluid_3D.c:41-110): velocity rows [K G], pressure rows [G^T  -S]
// with K = 2 eps(u):eps(v), G_ij = int dN_i/dx_a N_j, S = lap / 6.  Every
// block is invariant under the reflections of the cube, so in the Walsh
// basis of each field (u, v, w, p) it is sparse: 64 non-zeros of 1024, plus
// the pressure's constant mode.  With W = wht(field) (P1 = sum of the 8
// pressures) and F the table below, f = iwht(F) (+ Fp1 on every pressure
// node).  The table was derived from the dense matrix in the frame of wht()
// (x = col offset, y = 1 - row offset, z = 1 - layer offset) and reproduces
// it to rounding.
//
// Vectors: [u,v,w interleaved for DOF < nVel | p for DOF < nVel + nBorder]
// (matrix_free.cu:711,4978).  Per node the fluid key says what the node is:
// 255 pore (4 DOFs), 1..254 border (pressure only; its velocity is the wall's
// zero), 0 solid (no DOF); bit k is set iff the element for which the node is
// local node k is pore.  Only pore elements contribute.
#include <hip/hip_runtime.h>
#include "walsh.h"

namespace chfem {

// Element forces f = K_e u (scale: 1, or 0 for an element that contributes
// nothing, which then yields zero forces without a branch around the
// evaluation)
template <typename T>
__device__ __forceinline__ void stokes_forces(const T (&u)[8][4], T (&f)[8][4], T scale = T(1))
{
    T c[8];
    Walsh<T> Wu, Wv, Ww, Wp;
    T P1;
#pragma unroll
    for (int k = 0; k < 8; k++) c[k] = u[k][0];
    Wu = wht(c);
#pragma unroll
    for (int k = 0; k < 8; k++) c[k] = u[k][1];
    Wv = wht(c);
#pragma unroll
    for (int k = 0; k < 8; k++) c[k] = u[k][2];
    Ww = wht(c);
#pragma unroll
    for (int k = 0; k < 8; k++) c[k] = u[k][3];
    Wp = wht(c, &P1);
    const T h8 = scale * T(1.0 / 8.0), h16 = scale * T(1.0 / 16.0), h24 = scale * T(1.0 / 24.0),
            h32 = scale * T(1.0 / 32.0), h36 = scale * T(1.0 / 36.0), h48 = scale * T(1.0 / 48.0),
            h96 = scale * T(1.0 / 96.0), h144 = scale * T(1.0 / 144.0), h288 = scale * T(1.0 / 288.0);
    Walsh<T> Fu, Fv, Fw, Fp;
    Fu.x = h8 * Wu.x + h32 * P1;
    Fu.y = h16 * (Wu.y + Wv.x);
    Fu.z = h16 * (Wu.z - Ww.x);
    Fu.xy = h16 * Wu.xy + h96 * Wp.y;
    Fu.yz = h24 * Wu.yz + h48 * (Wv.xz - Ww.xy);
    Fu.xz = h16 * Wu.xz + h96 * Wp.z;
    Fu.xyz = h36 * Wu.xyz + h288 * Wp.yz;
    Fv.x = Fu.y;
    Fv.y = h8 * Wv.y + h32 * P1;
    Fv.z = h16 * (Wv.z - Ww.y);
    Fv.xy = h16 * Wv.xy + h96 * Wp.x;
    Fv.yz = h16 * Wv.yz + h96 * Wp.z;
    Fv.xz = h48 * (Wu.yz - Ww.xy) + h24 * Wv.xz;
    Fv.xyz = h36 * Wv.xyz + h288 * Wp.xz;
    Fw.x = -Fu.z;
    Fw.y = -Fv.z;
    Fw.z = h8 * Ww.z - h32 * P1;
    Fw.xy = h24 * Ww.xy - h48 * (Wu.yz + Wv.xz);
    Fw.yz = h16 * Ww.yz - h96 * Wp.y;
    Fw.xz = h16 * Ww.xz - h96 * Wp.x;
    Fw.xyz = h36 * Ww.xyz - h288 * Wp.xy;
    const T Fp1 = h32 * (Wu.x + Wv.y - Ww.z);
    Fp.x = h96 * (Wv.xy - Ww.xz - Wp.x);
    Fp.y = h96 * (Wu.xy - Ww.yz - Wp.y);
    Fp.z = h96 * (Wu.xz + Wv.yz - Wp.z);
    Fp.xy = -h288 * Ww.xyz - h144 * Wp.xy;
    Fp.yz = h288 * Wu.xyz - h144 * Wp.yz;
    Fp.xz = h288 * Wv.xyz - h144 * Wp.xz;
    Fp.xyz = -h288 * Wp.xyz;
    iwht(Fu, c);
#pragma unroll
    for (int k = 0; k < 8; k++) f[k][0] = c[k];
    iwht(Fv, c);
#pragma unroll
    for (int k = 0; k < 8; k++) f[k][1] = c[k];
    iwht(Fw, c);
#pragma unroll
    for (int k = 0; k < 8; k++) f[k][2] = c[k];
    iwht(Fp, c, &Fp1);
#pragma unroll
    for (int k = 0; k < 8; k++) f[k][3] = c[k];
}

// LDS plane entries of 4 values (one node's u, v, w, p), 16-byte aligned,
// moved as 16-byte vectors: a lane stride of 16 bytes read or written one
// value at a time falls on a quarter of the LDS banks (measured: 42 M
// bank-conflict cycles per 256^3 FP32 launch, none with the vectors)
typedef float f4v __attribute__((ext_vector_type(4)));
typedef double d2v __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void lds_get4(const float *p, float (&o)[4])
{
    const f4v v = *reinterpret_cast<const f4v *>(p);
    o[0] = v.x; o[1] = v.y; o[2] = v.z; o[3] = v.w;
}
__device__ __forceinline__ void lds_get4(const double *p, double (&o)[4])
{
    const d2v a = *reinterpret_cast<const d2v *>(p), b = *reinterpret_cast<const d2v *>(p + 2);
    o[0] = a.x; o[1] = a.y; o[2] = b.x; o[3] = b.y;
}
__device__ __forceinline__ void lds_put4(float *p, const float (&o)[4])
{
    *reinterpret_cast<f4v *>(p) = f4v{o[0], o[1], o[2], o[3]};
}
__device__ __forceinline__ void lds_put4(double *p, const double (&o)[4])
{
    *reinterpret_cast<d2v *>(p) = d2v{o[0], o[1]};
    *reinterpret_cast<d2v *>(p + 2) = d2v{o[2], o[3]};
}

// Per-node fluid keys from the DOF maps: a node is pore if its DOF is below
// nVel and maps back to it, border if its DOF is in [nVel, nVel + nBorder)
// and maps back to it (key from the border keys), solid otherwise (the
// node-to-DOF map holds 0 for solid nodes; femhmg_fluid_3D.c:370-424).
__global__ void k_fluid_node_keys(const uint32_t *__restrict__ node2dof, const uint32_t *__restrict__ dof2node,
                                  const uint8_t *__restrict__ bkeys, uint32_t nvel, uint32_t nborder, uint32_t N,
                                  uint8_t *__restrict__ out)
{
    for (uint32_t n = blockIdx.x * blockDim.x + threadIdx.x; n < N; n += gridDim.x * blockDim.x) {
        const uint32_t id = node2dof[n];
        uint8_t k = 0;
        if (id < nvel + nborder && dof2node[id] == n) k = id < nvel ? 255 : bkeys[id - nvel];
        out[n] = k;
    }
}

// Lean operator form of the memory-lean solvers (LeanForm; form 0: plain
// apply) with the Jacobi inverse diagonal per DOF class: velocity vinv, pore
// pressure pinv, border pressure btab[key] (btab == nullptr and vinv = pinv
// = 1: no preconditioner).  Form 3 reads and writes z through q.
template <typename T>
struct StokesLean {
    int form;
    T alpha, vinv, pinv;
    const T *btab;
};

// one DOF of a lean form: the operator result o at a DOF whose input value
// was u, M^-1 entry inv, z (form 3) at zq
template <typename T>
__device__ __forceinline__ void lean_dof(const StokesLean<T> &L, T u, T o, T inv, T *zq, double &acc)
{
    if (L.form == LEAN_VAV) {
        acc += (double)u * o;
    } else if (L.form == LEAN_QMQ) {
        acc += (double)o * (double)(inv * o);
    } else {
        const T zn = *zq - L.alpha * (inv * o);
        *zq = zn;
        acc += (double)zn * (double)zn / (double)inv;
    }
}

// One tile of the Stokes operator (see walsh_tile for the march, the tile
// order and the load pipelining; here a plane node holds (u, v, w, p)
// gathered through the node-to-DOF map, whose entries and keys are loaded
// one layer ahead of the value gathers).
template <typename T, int TR, int TC, bool DOT, bool LEAN = false>
__device__ __forceinline__ double stokes_tile(const T *__restrict__ v, T *__restrict__ q,
                                              const uint32_t *__restrict__ node2dof, const uint8_t *__restrict__ fk,
                                              uint32_t nvel, int ny, int nx, int nz, int LZ, int bx, int by, int bz,
                                              StokesLean<T> L = StokesLean<T>{})
{
    constexpr int NT = TR * TC, NR = TR + 1, NC = TC + 1, NN = NR * NC;
    constexpr int NS = (NN + NT - 1) / NT;  // plane nodes per work-item
    constexpr int SB = 9;                   // 2 x 4 row-combined forces + pad
    static_assert(TR == 16 || TR == 32 || TR == 64, "rows per tile must tile a 64-wide wavefront");
    __shared__ __align__(16) T su[2][NS * NT * 4];
    __shared__ T sb[NT * SB];
    const int tid = threadIdx.x;
    const int tr = tid % TR, tc = tid / TR;
    const int r0 = bx * (TR - 1), c0 = by * (TC - 1), l0 = bz * LZ;
    const int vr = min(TR - 1, ny - r0), vc = min(TC - 1, nx - c0), vl = min(LZ, nz - l0);
    const size_t nxy = (size_t)nx * ny;
    const T *vp = v + 3 * (size_t)nvel;  // pressures
    auto wrapl = [&](int lg) {
        int lw = lg % nz;
        return lw < 0 ? lw + nz : lw;
    };
    // plane nodes of this work-item: LDS node index j = tid + k * NT is node
    // (r0 - 1 + j % NR, c0 - 1 + j / NR); nodes outside the tile read node 0
    uint32_t noff[NS];
#pragma unroll
    for (int k = 0; k < NS; k++) {
        const int j = tid + k * NT, nr = j % NR, nc = j / NR;
        const bool ok = j < NN && nr <= vr + 1 && nc <= vc + 1;
        noff[k] = ok ? (uint32_t)wrapi(c0 - 1 + nc, nx) * (uint32_t)ny + (uint32_t)wrapi(r0 - 1 + nr, ny) : 0u;
    }
    uint32_t nid[NS], nkey[NS];  // DOF ids and keys of the next plane to gather
    auto load_map = [&](int lg) {
        const size_t lo = (size_t)wrapl(lg) * nxy;
#pragma unroll
        for (int k = 0; k < NS; k++) {
            nid[k] = node2dof[lo + noff[k]];
            nkey[k] = fk[lo + noff[k]];
        }
    };
    T pre[NS][4];
    auto fetch = [&](int lg) {
#pragma unroll
        for (int k = 0; k < NS; k++) {
            const bool pore = nkey[k] == 255u, fluid = nkey[k] != 0u;
            const size_t iv = pore ? 3 * (size_t)nid[k] : 0, ip = fluid ? nid[k] : 0;
            const T a0 = v[iv], a1 = v[iv + 1], a2 = v[iv + 2], a3 = vp[ip];
            pre[k][0] = pore ? a0 : T(0);
            pre[k][1] = pore ? a1 : T(0);
            pre[k][2] = pore ? a2 : T(0);
            pre[k][3] = fluid ? a3 : T(0);
        }
        load_map(lg + 1);
    };
    auto store = [&](T *dst) {
#pragma unroll
        for (int k = 0; k < NS; k++) lds_put4(dst + (tid + k * NT) * 4, pre[k]);
    };
    const bool elem_ok = tr <= vr && tc <= vc;
    const uint32_t eoff = elem_ok ? (uint32_t)wrapi(c0 - 1 + tc, nx) * (uint32_t)ny + (uint32_t)wrapi(r0 - 1 + tr, ny)
                                  : 0u;
    // the element with origin node n is pore iff bit 3 of n's key is set
    auto elem_key = [&](int le) -> uint32_t { return fk[(size_t)wrapl(le) * nxy + eoff]; };
    auto is_pore = [&](uint32_t kv) -> bool { return elem_ok && ((kv >> 3) & 1u); };
    auto element = [&](const T *cur, const T *nxt, bool pore_e, T (&A)[4], T (&Af)[4]) {
        T *dst = sb + tid * SB;
        auto row_sums = [&](const T (&f)[8][4]) {
#pragma unroll
            for (int a = 0; a < 4; a++) {
                A[a] = f[3][a] + lane_up1(f[0][a]);
                Af[a] = f[7][a] + lane_up1(f[4][a]);
                dst[a] = f[2][a] + lane_up1(f[1][a]);
                dst[4 + a] = f[6][a] + lane_up1(f[5][a]);
            }
        };
        if constexpr (sizeof(T) == 4) {
            // FP32: a wavefront-uniform branch; a wavefront without a pore
            // element skips the evaluation, in the others the non-pore
            // elements evaluate with scale 0 (a per-lane branch's join
            // zeroed all 32 forces in every wavefront: 7 % slower at 256^3)
            if (__ballot(pore_e)) {
                T u[8][4], f[8][4];
#pragma unroll
                for (int k = 0; k < 8; k++)
                    lds_get4((WDL(k) ? nxt : cur) + ((tc + WDC(k)) * NR + tr + WDR(k)) * 4, u[k]);
                stokes_forces(u, f, pore_e ? T(1) : T(0));
                row_sums(f);
            } else {
#pragma unroll
                for (int a = 0; a < 4; a++) {
                    A[a] = Af[a] = T(0);
                    dst[a] = dst[4 + a] = T(0);
                }
            }
        } else {
            // FP64: the per-lane branch (the scaled evaluation of non-pore
            // elements was 3 % slower)
            T f[8][4];
            if (pore_e) {
                T u[8][4];
#pragma unroll
                for (int k = 0; k < 8; k++)
                    lds_get4((WDL(k) ? nxt : cur) + ((tc + WDC(k)) * NR + tr + WDR(k)) * 4, u[k]);
                stokes_forces(u, f);
            } else {
#pragma unroll
                for (int k = 0; k < 8; k++)
#pragma unroll
                    for (int a = 0; a < 4; a++) f[k][a] = T(0);
            }
            row_sums(f);
        }
    };
    const bool own = tr >= 1 && tc >= 1 && tr - 1 < vr && tc - 1 < vc;
    const T *bleft = sb + ((tc >= 1 ? tc - 1 : 0) * TR + tr) * SB;
    const size_t own_off = own ? (size_t)(c0 + tc - 1) * ny + (r0 + tr - 1) : 0;

    T *cur = su[0], *nxt = su[1];
    load_map(l0 - 1);
    fetch(l0 - 1);
    uint32_t ek = elem_key(l0 - 1);
    store(cur);
    fetch(l0);
    store(nxt);
    bool pe = is_pore(ek);
    ek = elem_key(l0);
    // owned node's DOF id and key, one layer ahead
    uint32_t oid = node2dof[(size_t)l0 * nxy + own_off], okey = fk[(size_t)l0 * nxy + own_off];
    __syncthreads();
    fetch(l0 + 1);
    T A[4], Af[4];
    element(cur, nxt, pe, A, Af);
    __syncthreads();
    T acc[4];
#pragma unroll
    for (int a = 0; a < 4; a++) acc[a] = own ? Af[a] + bleft[4 + a] : T(0);
    store(cur);
    pe = is_pore(ek);
    {
        T *t = cur; cur = nxt; nxt = t;
    }
    __syncthreads();
    double dsum = 0.0;
    for (int s = 0; s < vl; s++) {
        const int l = l0 + s;
        const uint32_t id = oid, key = okey;
        oid = node2dof[(size_t)wrapl(l + 1) * nxy + own_off];
        okey = fk[(size_t)wrapl(l + 1) * nxy + own_off];
        ek = elem_key(l + 1);
        fetch(l + 2);
        T own_u[4] = {0, 0, 0, 0};
        if (DOT && own) lds_get4(cur + (tc * NR + tr) * 4, own_u);
        element(cur, nxt, pe, A, Af);
        __syncthreads();
        store(cur);  // layer l + 2 replaces layer l
        pe = is_pore(ek);
        if (own) {
            T o[4];
#pragma unroll
            for (int a = 0; a < 4; a++) {
                o[a] = acc[a] + A[a] + bleft[a];
                acc[a] = Af[a] + bleft[4 + a];
            }
            if constexpr (LEAN) {
                if (key == 255u) {
                    T *qo = q + 3 * (size_t)id;
#pragma unroll
                    for (int a = 0; a < 3; a++) lean_dof(L, own_u[a], o[a], L.vinv, qo + a, dsum);
                }
                if (key != 0u) {
                    const T inv = key == 255u ? L.pinv : (L.btab ? L.btab[key] : T(1));
                    lean_dof(L, own_u[3], o[3], inv, q + 3 * (size_t)nvel + id, dsum);
                }
            } else {
                if (key == 255u) {
                    T *qo = q + 3 * (size_t)id;
                    qo[0] = o[0]; qo[1] = o[1]; qo[2] = o[2];
                    if (DOT) dsum += (double)own_u[0] * o[0] + (double)own_u[1] * o[1] + (double)own_u[2] * o[2];
                }
                if (key != 0u) {
                    q[3 * (size_t)nvel + id] = o[3];
                    if (DOT) dsum += (double)own_u[3] * o[3];
                }
            }
        }
        {
            T *t = cur; cur = nxt; nxt = t;
        }
        __syncthreads();
    }
    return dsum;
}

template <typename T, int TR, int TC, bool DOT, bool LEAN = false>
__global__ void __launch_bounds__(TR *TC) k_stokes_walsh(const T *__restrict__ v, T *__restrict__ q,
                                                          const uint32_t *__restrict__ node2dof,
                                                          const uint8_t *__restrict__ fk, uint32_t nvel, int ny, int nx,
                                                          int nz, int LZ, Reduce red,
                                                          StokesLean<T> L = StokesLean<T>{})
{
    if (DOT && stopped(red)) return;
    if (LEAN) L.alpha = L.form == LEAN_ZUPD ? (T)red.S[S_ALPHA] : T(0);
    // XCD-contiguous tile order (see k_walsh)
    const unsigned gx = (unsigned)(ny + TR - 2) / (TR - 1), gy = (unsigned)(nx + TC - 2) / (TC - 1);
    const unsigned b = blockIdx.x, nb = gridDim.x, per = nb / 8u, rem = nb % 8u, xcd = b % 8u, k = b / 8u;
    const unsigned t = (xcd < rem ? xcd * (per + 1u) : rem * (per + 1u) + (xcd - rem) * per) + k;
    const double dsum = stokes_tile<T, TR, TC, DOT, LEAN>(v, q, node2dof, fk, nvel, ny, nx, nz, LZ, (int)(t % gx),
                                                          (int)((t / gx) % gy), (int)(t / (gx * gy)), L);
    if (DOT) {
        double vv[1] = {dsum};
        store_partials<1>(red, vv);
    }
}

void launch_fluid_node_keys(const OpDesc &op, uint8_t *out, hipStream_t s)
{
    const uint32_t N = op.N;
    k_fluid_node_keys<<<(N + 255) / 256 < 65536 ? (N + 255) / 256 : 65536, 256, 0, s>>>(
        op.node2dof, op.dof2node, op.bkeys, op.nvel, op.nborder, N, out);
}

template <typename T>
void launch_stokes_walsh(const OpDesc &op, const T *v, T *q, const Reduce *red, hipStream_t s)
{
    Reduce rr = {nullptr, nullptr, 0, 0, 0};
    if (red) rr = *red;
    constexpr int TR = 32, TC = 8;
    int LZ = op.tile_lz > 0 ? op.tile_lz : 16;
    const long cols = (long)((op.ny + TR - 2) / (TR - 1)) * ((op.nx + TC - 2) / (TC - 1));
    while (cols * ((op.nz + LZ - 1) / LZ) > MAX_PARTIAL_BLOCKS) LZ *= 2;
    const unsigned grid = (unsigned)cols * (unsigned)((op.nz + LZ - 1) / LZ);
    if (red) {
        k_stokes_walsh<T, TR, TC, true><<<grid, TR * TC, 0, s>>>(v, q, op.node2dof, op.fkeys, op.nvel, op.ny, op.nx,
                                                                 op.nz, LZ, rr);
        launch_reduce_finish(rr, grid, 1, s);
    } else {
        k_stokes_walsh<T, TR, TC, false><<<grid, TR * TC, 0, s>>>(v, q, op.node2dof, op.fkeys, op.nvel, op.ny, op.nx,
                                                                  op.nz, LZ, rr);
    }
}

template <typename T>
void launch_stokes_lean(const OpDesc &op, const Jac<T> &jac, int form, const T *v, T *z, const Reduce *red,
                        hipStream_t s)
{
    constexpr int TR = 32, TC = 8;
    int LZ = op.tile_lz > 0 ? op.tile_lz : 16;
    const long cols = (long)((op.ny + TR - 2) / (TR - 1)) * ((op.nx + TC - 2) / (TC - 1));
    while (cols * ((op.nz + LZ - 1) / LZ) > MAX_PARTIAL_BLOCKS) LZ *= 2;
    const unsigned grid = (unsigned)cols * (unsigned)((op.nz + LZ - 1) / LZ);
    const bool pc = jac.kind == JAC_FLUID;
    const StokesLean<T> L = {form, T(0), pc ? jac.vinv : T(1), pc ? jac.pinv : T(1), pc ? jac.btab : nullptr};
    k_stokes_walsh<T, TR, TC, true, true><<<grid, TR * TC, 0, s>>>(v, z, op.node2dof, op.fkeys, op.nvel, op.ny, op.nx,
                                                                   op.nz, LZ, *red, L);
    launch_reduce_finish(*red, grid, 1, s);
}

template void launch_stokes_lean<float>(const OpDesc &, const Jac<float> &, int, const float *, float *,
                                        const Reduce *, hipStream_t);
template void launch_stokes_lean<double>(const OpDesc &, const Jac<double> &, int, const double *, double *,
                                         const Reduce *, hipStream_t);
template void launch_stokes_walsh<float>(const OpDesc &, const float *, float *, const Reduce *, hipStream_t);
template void launch_stokes_walsh<double>(const OpDesc &, const double *, double *, const Reduce *, hipStream_t);

}  // namespace chfem
