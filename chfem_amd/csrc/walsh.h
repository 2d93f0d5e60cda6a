// walsh.h -- operators q = A v for isotropic elastic and for conduction
// materials, computed element by element in the Walsh (hierarchical) basis of
// the trilinear voxel (device code shared by walsh.hip and the batched
// solver, batch.hip).
//
// For an element with nodal displacements u^a_k (component a, local node k)
// the Walsh sums W^a_s = sum_k s_k u^a_k over the sign patterns
// s in {x, y, z, xy, yz, xz, xyz} of the local nodes determine the
// displacement gradient polynomial exactly.  Because 1, x-1/2, y-1/2, ... are
// orthogonal on the unit cube, the element force vector int sigma(u):grad N_k
// has Walsh coefficients F^a_s that are short linear combinations of the W^a_s
// (lambda, mu weighted), and f^a_k = sum_s F^a_s s_k.  Each element costs
// about 180 flops instead of the 24x24 product's 1152, with the same result
// (the local matrix is lambda*A + mu*B for the unit-voxel element;
// chfem/src/femhmg/physical_phenomena/femhmg_elastic_3D.c:51-128).
//
// No atomics; node results are deterministic.
#pragma once
#include <hip/hip_runtime.h>
#include "kernels.h"
#include "reduce.h"

namespace chfem {

// local node k of an element sits at (row + DR, col + DC, layer + DL)
#define WDR(k) ((k) == 0 || (k) == 1 || (k) == 4 || (k) == 5 ? 1 : 0)
#define WDC(k) ((k) == 1 || (k) == 2 || (k) == 5 || (k) == 6 ? 1 : 0)
#define WDL(k) ((k) >= 4 ? 1 : 0)
// elastic element frame: x = col offset, y = 1 - row offset, z = 1 - layer offset
// node index inside the 2x2x2 cube [zb][yb][xb]
#define CUBE(xb, yb, zb) ((zb) == 1 ? ((yb) == 1 ? ((xb) == 1 ? 2 : 3) : ((xb) == 1 ? 1 : 0)) \
                                    : ((yb) == 1 ? ((xb) == 1 ? 6 : 7) : ((xb) == 1 ? 5 : 4)))

template <typename T>
struct Walsh { T x, y, z, xy, yz, xz, xyz; };

// forward: Walsh sums of 8 nodal values (local node order); total (if
// given) receives the constant mode, the plain sum of the 8 values
template <typename T>
__device__ __forceinline__ Walsh<T> wht(const T (&u)[8], T *total = nullptr)
{
    // x stage: pairs (xb=1) - (xb=0) and (xb=1) + (xb=0) for each (yb, zb)
    T a[2][2], b[2][2];
#pragma unroll
    for (int zb = 0; zb < 2; zb++)
#pragma unroll
        for (int yb = 0; yb < 2; yb++) {
            const T u1 = u[CUBE(1, yb, zb)], u0 = u[CUBE(0, yb, zb)];
            a[zb][yb] = u1 + u0;
            b[zb][yb] = u1 - u0;
        }
    T as[2], ad[2], bs[2], bd[2];
#pragma unroll
    for (int zb = 0; zb < 2; zb++) {
        as[zb] = a[zb][1] + a[zb][0];
        ad[zb] = a[zb][1] - a[zb][0];
        bs[zb] = b[zb][1] + b[zb][0];
        bd[zb] = b[zb][1] - b[zb][0];
    }
    Walsh<T> w;
    w.z = as[1] - as[0];
    w.y = ad[1] + ad[0];
    w.yz = ad[1] - ad[0];
    w.x = bs[1] + bs[0];
    w.xz = bs[1] - bs[0];
    w.xy = bd[1] + bd[0];
    w.xyz = bd[1] - bd[0];
    if (total) *total = as[1] + as[0];
    return w;
}

// inverse: f_k = sum_s F_s s_k, plus *c0 (a constant mode, if given)
template <typename T>
__device__ __forceinline__ void iwht(const Walsh<T> &F, T (&f)[8], const T *c0 = nullptr)
{
#pragma unroll
    for (int yb = 0; yb < 2; yb++) {
        const T sy = yb ? T(1) : T(-1);
        const T A = F.x + sy * F.xy, B = F.z + sy * F.yz, C = F.xz + sy * F.xyz;
        const T Y = c0 ? sy * F.y + *c0 : sy * F.y;
#pragma unroll
        for (int xb = 0; xb < 2; xb++) {
            const T sx = xb ? T(1) : T(-1);
            const T P = sx * A + Y, Q = B + sx * C;
            f[CUBE(xb, yb, 1)] = P + Q;
            f[CUBE(xb, yb, 0)] = P - Q;
        }
    }
}

// Per-material coefficients of the element force maps, with the Walsh
// normalisations folded in (set up once per block by material_coefficients):
// elastic  c = {lam/16, mu/16, mu/8, lam/48, (lam+3mu)/48, mu/48, (lam+4mu)/144}
// conduction c = {k/16, k/24, k/48}
constexpr int WCOEF = 8;

template <typename T>
__device__ __forceinline__ void material_coefficients(int nd, T p0, T p1, T *c)
{
    if (nd == 3) {
        c[0] = p0 * T(1.0 / 16.0);
        c[1] = p1 * T(1.0 / 16.0);
        c[2] = p1 * T(1.0 / 8.0);
        c[3] = p0 * T(1.0 / 48.0);
        c[4] = (p0 + 3 * p1) * T(1.0 / 48.0);
        c[5] = p1 * T(1.0 / 48.0);
        c[6] = (p0 + 4 * p1) * T(1.0 / 144.0);
    } else {
        c[0] = p0 * T(1.0 / 16.0);
        c[1] = p0 * T(1.0 / 24.0);
        c[2] = p0 * T(1.0 / 48.0);
    }
}

// element forces from nodal displacements u[k][a]; c: the element's material
// coefficients (material_coefficients, elastic)
template <typename T>
__device__ __forceinline__ void element_forces(const T (&u)[8][3], const T *c_, T (&f)[8][3])
{
    T c[8];
    Walsh<T> W[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
#pragma unroll
        for (int k = 0; k < 8; k++) c[k] = u[k][a];
        W[a] = wht(c);
    }
    const T l16 = c_[0], m16 = c_[1], m8 = c_[2], l48 = c_[3], l3 = c_[4], m48 = c_[5], l4 = c_[6];
    const T ltr = l16 * (W[0].x + W[1].y + W[2].z);
    const T S = W[0].yz + W[1].xz + W[2].xy;
    Walsh<T> F[3];
    // component x
    F[0].x = ltr + m8 * W[0].x;
    F[0].y = m16 * (W[0].y + W[1].x);
    F[0].z = m16 * (W[0].z + W[2].x);
    F[0].xy = l3 * W[0].xy + l48 * W[2].yz;
    F[0].xz = l3 * W[0].xz + l48 * W[1].yz;
    F[0].yz = m48 * (S + W[0].yz);
    F[0].xyz = l4 * W[0].xyz;
    // component y
    F[1].x = F[0].y;
    F[1].y = ltr + m8 * W[1].y;
    F[1].z = m16 * (W[1].z + W[2].y);
    F[1].xy = l3 * W[1].xy + l48 * W[2].xz;
    F[1].yz = l3 * W[1].yz + l48 * W[0].xz;
    F[1].xz = m48 * (S + W[1].xz);
    F[1].xyz = l4 * W[1].xyz;
    // component z
    F[2].x = F[0].z;
    F[2].y = F[1].z;
    F[2].z = ltr + m8 * W[2].z;
    F[2].xz = l3 * W[2].xz + l48 * W[1].xy;
    F[2].yz = l3 * W[2].yz + l48 * W[0].xy;
    F[2].xy = m48 * (S + W[2].xy);
    F[2].xyz = l4 * W[2].xyz;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        iwht(F[a], c);
#pragma unroll
        for (int k = 0; k < 8; k++) f[k][a] = c[k];
    }
}

// Conduction element forces f = K u for conductivity k.  The Walsh modes
// diagonalise the unit-voxel trilinear conduction matrix: with u written as
// (1/8) sum_s W_s prod_{d in s} 2(x_d - 1/2), the gradients of the modes are
// orthogonal on the cube and u^T K u = k sum_s a_s W_s^2 with a_s = 1/16,
// 1/24, 1/48 for modes of 1, 2, 3 axes, so F_s = k a_s W_s and f = iwht(F).
// This reproduces k * {1/3 diagonal, 0 between nodes that differ in one
// coordinate, -1/12 across face and body diagonals}
// (chfem/src/femhmg/physical_phenomena/femhmg_thermal_3D.c:42-120).
// (c: the material coefficients {k/16, k/24, k/48})
template <typename T>
__device__ __forceinline__ void conduction_forces(const T (&u)[8], const T *c, T (&f)[8])
{
    const Walsh<T> W = wht(u);
    const T a1 = c[0], a2 = c[1], a3 = c[2];
    Walsh<T> F;
    F.x = a1 * W.x;
    F.y = a1 * W.y;
    F.z = a1 * W.z;
    F.xy = a2 * W.xy;
    F.yz = a2 * W.yz;
    F.xz = a2 * W.xz;
    F.xyz = a3 * W.xyz;
    iwht(F, f);
}

__device__ __forceinline__ int wrapi(int a, int n) { return a < 0 ? a + n : (a >= n ? a - n : a); }

// value of lane - 1 of the wavefront (a DPP whole-wave shift right by one
// lane; lane 0 receives 0).  Same as __shfl_up(v, 1) for lanes 1..63, without
// the LDS permute.
__device__ __forceinline__ float lane_up1(float v)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x138, 0xf, 0xf, false));
}
__device__ __forceinline__ double lane_up1(double v)
{
    const long long b = __builtin_bit_cast(long long, v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffll), 0x138, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), 0x138, 0xf, 0xf, false);
    return __builtin_bit_cast(double, ((long long)hi << 32) | (long long)(unsigned)lo);
}

// Keeps a value's computation at this point of the instruction stream (an
// empty asm that consumes it; no instruction is emitted).  Used where the
// compiler would sink a decode of loaded data below later stores, whose wait
// (vmcnt counts stores too) would then also wait for the stores.
__device__ __forceinline__ void pin_here(int v) { asm volatile("" ::"v"(v)); }

// Direction update fused into the operator.  The CG loop that launches this
// form iterates on the preconditioned residual z = M^-1 r (Jacobi from the
// packed keys in its update kernel), so the operator's input
// d = z + beta d_old is formed from two streamed vectors while the node
// planes are fetched; q = A d is computed from it, and d is written to dnew
// for the nodes the block owns.  The iterate is not touched: the loop keeps
// its directions in a ring of buffers and adds sum alpha_j d_j to x once per
// window (launch_ring_flush).  v is d_old; first: d = z (v unused).
template <typename T>
struct DirFuse {
    const T *z;
    T *dnew;
    T beta;
    T alpha;
    int first;
    // single-kernel CG (FM 2): the z update of the previous iteration is
    // deferred here, z = z_old - alpha qt_old (alpha = S[ALPHA]), written to
    // znew on the owned nodes; this launch's q output is qt = dinv q (dinv
    // from the packed keys of the owned nodes, kinv[key * ND + a]; the
    // diagonal itself at kinv[65536 * ND + key * ND + a]), the only form of
    // q the loop reads.  With zpend clear the z update is skipped (qt_old
    // reads z).  znew == nullptr: z is not written.
    int zpend;
    const T *qold;
    T *znew;
    const T *kinv;
    // zring: z_old is not read but recovered from the two previous
    // directions, z_old = d_old - bprev d_old2 (the update that formed d_old,
    // bprev = S[BPREV], recorded by that iteration's finishing kernel); z
    // then points at d_old2
    int zring;
    T bprev;
    // lean operator form (FM 3, LeanForm): 1 v.Av; 2 Av.M^-1 Av; 3 z -= alpha
    // M^-1 A v in place in q (z) with z.diag z; M^-1 from kinv by the owned
    // node's packed key (kinv == nullptr: no preconditioner)
    int lean;
};

// Sums a tile contributes to the launch's reductions: d.q, and for the
// single-kernel CG z.q, q.(dinv q) and z.(diag z) (the r.z of the z it
// formed)
struct TileSums {
    double dq, zq, qmq, zdz;
};

// Work decomposition: a block of TR x TC work-items evaluates a TR x TC
// patch of elements per element layer and outputs the (TR-1) x (TC-1) nodes
// interior to the patch (work-item (tr, tc), tr, tc >= 1, owns node
// (r0 + tr - 1, c0 + tc - 1)).  Every work-item evaluates exactly one element
// per layer step.  HALO (TR = 64): the TR x TC work-items evaluate the
// elements whose origin is a node of the tile's TR rows, and one more
// wavefront evaluates the TC elements of the row above (origin row r0 - 1),
// whose far-row forces reach row 0 of the tile through LDS; the block then
// outputs TR x (TC-1) nodes (work-item (tr, tc), tc >= 1, owns node
// (r0 + tr, c0 + tc - 1)), so tiles of 64 rows cover 256 rows exactly.  The block marches through LZ node layers; node layers stream
// through two LDS buffers, the next layer's values are fetched into registers
// while the current layer's elements are evaluated, and the far-face forces of
// each node are carried to the next step in registers.
//
// ND = 3: isotropic elastic (lamu holds {lambda, mu} per material); ND = 1:
// conduction (lamu holds {k, 0} per material).  Vectors hold ND values per
// node, [ND * node + a].
//
// walsh_tile computes the tile (bx, by, bz) of one operator application and
// returns this work-item's share of v.q (DOT); k_walsh maps tiles to blocks
// one to one, the batched solver (batch.hip) through a tile table.  FUSE: the
// input is the fused direction update of DirFuse (the dot is then d.q with
// the updated d).
template <typename T, int ND, int TR, int TC, bool DOT, int FM = 0, bool HALO = false>
__device__ __forceinline__ TileSums walsh_tile(const T *__restrict__ v, T *__restrict__ q,
                                               const uint16_t *__restrict__ keys, const T *__restrict__ lamu, int ebe,
                                               int ny, int nx, int nz, int LZ, int bx, int by, int bz,
                                               DirFuse<T> fz = DirFuse<T>{})
{
    static_assert(ND == 1 || ND == 3, "1 (conduction) or 3 (elastic) values per node");
    // FM 0: plain apply; 1: direction update inside (FUSE); 2: single-kernel
    // CG, the z update inside as well (CG1); 3: a lean form of the
    // memory-lean solvers (fz.lean), the operator's result reduced or folded
    // into z instead of stored
    constexpr bool FUSE = FM == 1 || FM == 2, CG1 = FM == 2, LEAN = FM == 3;
    static_assert(!HALO || TR == 64, "the halo wavefront variant takes 64-row columns");
    // NCOL column work-items (+ the halo wavefront); OR node rows owned per
    // tile; the node planes hold rows r0 - 1 .. r0 + OR
    constexpr int NCOL = TR * TC, NT = NCOL + (HALO ? 64 : 0), OR = HALO ? TR : TR - 1;
    constexpr int NR = OR + 2, NC = TC + 1, NV = NR * NC * ND;
    constexpr int SLOTS = (NV + NT - 1) / NT;
    constexpr int SB = 2 * ND + 1;  // padded stride of the row-combined forces in LDS (no bank conflicts)
    // node planes padded to SLOTS * NT entries: every work-item stores every
    // slot, so the plane loads and stores need no per-slot branches (a
    // branch around a load makes the compiler wait for all loads in flight
    // at its join)
    __shared__ T su[2][SLOTS * NT];
    // CG1: the updated z planes, for the owned nodes' z output and z.q
    __shared__ T szp[CG1 ? 2 : 1][CG1 ? SLOTS * NT : 1];
    __shared__ T sb[NCOL * SB];
    // HALO: far-row forces f0, f4, f1, f5 of the row-above elements, per
    // column (one slot per work-item of the halo wavefront)
    __shared__ T shal[HALO ? 64 * 4 * ND : 1];
    // material coefficients of materials 0..3, and material 4 (all zero) for
    // the work-items without an element: they evaluate a zero element instead
    // of branching around the evaluation
    __shared__ T smat[5 * WCOEF];
    const int tid = threadIdx.x;
    // hl: work-item of the halo wavefront (wavefront-uniform), evaluating the
    // element of column hc (hc < TC) of the row above the tile
    const bool hl = HALO && tid >= NCOL;
    const int hc = tid - NCOL;
    const int tr = hl ? 0 : tid % TR, tc = hl ? 0 : tid / TR;
    const int r0 = bx * OR, c0 = by * (TC - 1), l0 = bz * LZ;
    const int vr = min(OR, ny - r0), vc = min(TC - 1, nx - c0), vl = min(LZ, nz - l0);
    // this work-item's element: origin at plane row er, column ec (node
    // (r0 - 1 + er, c0 - 1 + ec)); halo work-items past the tile's columns
    // point at column 0 and evaluate the zero material
    const int er = hl ? 0 : tr + (HALO ? 1 : 0);
    const int ec = hl ? (hc < TC ? hc : 0) : tc;
    const size_t nxy = (size_t)nx * ny;
    if (tid < 5) material_coefficients<T>(ND, tid < 4 ? lamu[2 * tid] : T(0), tid < 4 ? lamu[2 * tid + 1] : T(0),
                                          smat + WCOEF * tid);

    // node-layer load pattern: LDS index idx = (nc * NR + nr) * ND + comp holds node
    // (r0 - 1 + nr, c0 - 1 + nc); the in-layer global offset is the same for every
    // layer.  Slots outside the tile's nodes load node 0 of the layer (in bounds);
    // what they store to LDS is never read.
    uint32_t goff[SLOTS];
#pragma unroll
    for (int k = 0; k < SLOTS; k++) {
        const int idx = tid + k * NT;
        const int comp = idx % ND, t2 = idx / ND, nr = t2 % NR, nc = t2 / NR;
        const bool ok = idx < NV && nr <= vr + 1 && nc <= vc + 1;
        const int gr = wrapi(r0 - 1 + nr, ny), gc = wrapi(c0 - 1 + nc, nx);
        goff[k] = ok ? (uint32_t)ND * ((uint32_t)gc * (uint32_t)ny + (uint32_t)gr) + (uint32_t)comp : 0u;
    }
    // layer index modulo nz for any lg (the prologue; the loop steps its
    // wrapped layer indices incrementally)
    auto wrapl = [&](int lg) {
        int lw = lg % nz;
        return lw < 0 ? lw + nz : lw;
    };
    auto next_layer = [&](int lw) { return lw + 1 == nz ? 0 : lw + 1; };
    T pre[SLOTS];
    // FUSE: the direction update's operands (z and d_old) are fetched with
    // the plane and combined when the plane is stored to LDS.  On the first
    // iteration (no d_old) the d_old operand reads z and beta is 0.
    T pz[FUSE ? SLOTS : 1];
    const T *dsrc = v;
    T beta = T(0);
    if constexpr (FUSE) {
        dsrc = fz.first ? fz.z : v;
        beta = fz.first ? T(0) : fz.beta;
    }
    // CG1: qt_old (= dinv q_old) of every plane slot
    T pq[CG1 ? SLOTS : 1];
    const T *qsrc = CG1 && fz.zpend ? fz.qold : fz.z;
    // fetch(lw): the plane of wrapped layer lw (uniform base pointer, 32-bit
    // per-slot offsets)
    auto fetch = [&](int lw) {
        const size_t lo = ND * (size_t)lw * nxy;
        const T *bq = qsrc + lo, *bz = fz.z + lo, *bd = dsrc + lo;
#pragma unroll
        for (int k = 0; k < SLOTS; k++) {
            if constexpr (CG1) pq[k] = bq[goff[k]];
            if constexpr (FUSE) pz[k] = bz[goff[k]];
            pre[k] = bd[goff[k]];
        }
    };
    auto store = [&](T *dst, T *zdst) {
#pragma unroll
        for (int k = 0; k < SLOTS; k++) {
            if constexpr (CG1) {
                const T zo = fz.zring ? pre[k] - fz.bprev * pz[k] : pz[k];
                const T zk = zo - fz.alpha * pq[k];
                dst[tid + k * NT] = zk + beta * pre[k];
                zdst[tid + k * NT] = zk;
            } else if constexpr (FUSE) {
                dst[tid + k * NT] = pz[k] + beta * pre[k];
            } else {
                dst[tid + k * NT] = pre[k];
            }
        }
    };
    // element of this work-item: origin (r0 - 1 + tr, c0 - 1 + tc).  Its
    // material is read from the packed keys one layer ahead (mat_key) and
    // decoded when needed (material).
    const bool elem_ok = hl ? (hc < TC && hc <= vc) : ((HALO ? tr < vr : tr <= vr) && tc <= vc);
    // (work-items without an element read node 0 of the layer: the key load
    // is unconditional)
    const uint32_t eoff = elem_ok ? (uint32_t)wrapi(c0 - 1 + ec, nx) * (uint32_t)ny + (uint32_t)wrapi(r0 - 1 + er, ny)
                                  : 0u;
    // (mat_key takes the wrapped layer index)
    auto mat_key = [&](int lw) -> uint16_t { return (keys + (size_t)lw * nxy)[eoff]; };
    auto material = [&](uint32_t kv) -> int {
        // NBN keys: node (r,c,l) is local node 3 of element (r,c,l); EBE: element map
        if (!elem_ok) return 4;
        return ebe ? (int)(kv & 3u) : (int)((kv >> 6) & 3u);
    };
    // Node (tr, tc) of a layer receives local node k's force from the element
    // (tr - WDR(k), tc - WDC(k)): from element layer l (k = 0..3)
    //   E(tr,tc).f3 + E(tr-1,tc).f0  +  E(tr,tc-1).f2 + E(tr-1,tc-1).f1
    //   =        A(tr, tc)           +            B(tr, tc-1)
    // and from element layer l - 1 (k = 4..7) the same with f7, f4 (Af) and
    // f6, f5 (Bf).  E(tr-1, tc) is lane - 1 of the same wavefront (TR is 16,
    // 32 or 64, and row 0 only feeds nodes no work-item owns), so A and Af
    // are formed with cross-lane moves; B and Bf go through LDS (2 ND values
    // per element).
    static_assert(TR == 16 || TR == 32 || TR == 64, "rows per tile must tile a 64-wide wavefront");
    // (work-items without an element read in-tile LDS entries, all written
    // each step from finite vector values, and evaluate them with the zero
    // material 4)
    auto element = [&](const T *cur, const T *nxt, int m, T (&A)[ND], T (&Af)[ND]) {
        T f[8][ND];
        T u[8][ND];
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const T *src = WDL(k) ? nxt : cur;
            const int o = ((ec + WDC(k)) * NR + er + WDR(k)) * ND;
#pragma unroll
            for (int a = 0; a < ND; a++) u[k][a] = src[o + a];
        }
        if constexpr (ND == 3) {
            element_forces(u, smat + WCOEF * m, f);
        } else {
            T u1[8], f1[8];
#pragma unroll
            for (int k = 0; k < 8; k++) u1[k] = u[k][0];
            conduction_forces(u1, smat + WCOEF * m, f1);
#pragma unroll
            for (int k = 0; k < 8; k++) f[k][0] = f1[k];
        }
        if (hl) {
            // row-above element: its forces on the tile's row 0 (local
            // nodes 0, 4 in its column, 1, 5 in the next)
            T *h = shal + hc * 4 * ND;
#pragma unroll
            for (int a = 0; a < ND; a++) {
                h[a] = f[0][a];
                h[ND + a] = f[4][a];
                h[2 * ND + a] = f[1][a];
                h[3 * ND + a] = f[5][a];
                A[a] = Af[a] = T(0);
            }
            return;
        }
        T *dst = sb + tid * SB;
#pragma unroll
        for (int a = 0; a < ND; a++) {
            A[a] = f[3][a] + lane_up1(f[0][a]);
            Af[a] = f[7][a] + lane_up1(f[4][a]);
            dst[a] = f[2][a] + lane_up1(f[1][a]);
            dst[ND + a] = f[6][a] + lane_up1(f[5][a]);
        }
    };
    // HALO: row 0 of the tile (lane 0 of each column wavefront, where the
    // cross-lane move brings 0) adds the row-above elements' forces, after
    // the barrier that follows the elements
    auto halo_row = [&](T (&A)[ND], T (&Af)[ND]) {
        if (HALO && tr == 0) {
            const T *h0 = shal + tc * 4 * ND, *h1 = shal + (tc - 1) * 4 * ND;
#pragma unroll
            for (int a = 0; a < ND; a++) {
                A[a] += h0[a] + h1[2 * ND + a];
                Af[a] += h0[ND + a] + h1[3 * ND + a];
            }
        }
    };
    // owned node (r0 + orow, c0 + tc - 1)
    const int orow = HALO ? tr : tr - 1;
    const bool own = !hl && orow >= 0 && tc >= 1 && orow < vr && tc - 1 < vc;
    // B and Bf of the element left of this node's column, same row
    const T *bleft = sb + ((tc >= 1 ? tc - 1 : 0) * TR + tr) * SB;

    T *cur = su[0], *nxt = su[1];
    T *zcur = szp[0], *znxt = szp[CG1 ? 1 : 0];   // z planes (CG1), swapped with cur / nxt
    const int lwm = wrapl(l0 - 1), lw0 = wrapl(l0);
    int lw1 = next_layer(lw0), lw2 = next_layer(lw1);  // wrapped l + 1, l + 2 of the step
    fetch(lwm);
    uint16_t mk = mat_key(lwm);
    store(cur, zcur);
    fetch(lw0);
    store(nxt, znxt);
    int m = material(mk);
    mk = mat_key(lw0);
    __syncthreads();
    fetch(lw1);
    T A[ND], Af[ND];
    element(cur, nxt, m, A, Af);
    __syncthreads();
    T acc[ND];
    if (own) halo_row(A, Af);
#pragma unroll
    for (int a = 0; a < ND; a++) acc[a] = own ? Af[a] + bleft[ND + a] : T(0);
    store(cur, zcur);  // layer l0 + 1 replaces layer l0 - 1
    m = material(mk);
    uint32_t kraw = mk;  // packed key of the owned node at the current layer (CG1 dinv)
    {
        T *t = cur; cur = nxt; nxt = t;  // cur = layer l0, nxt = layer l0 + 1
        t = zcur; zcur = znxt; znxt = t;
    }
    __syncthreads();
    double dsum = 0.0, zq = 0.0, qmq = 0.0, zdz = 0.0;
    // this work-item's node offset in a layer (work-items that own no node
    // use node 0 for their unconditional loads; only owners store)
    const size_t own_off = own ? (size_t)(c0 + tc - 1) * ny + (r0 + orow) : 0;
    for (int s = 0; s < vl; s++) {
        const int l = l0 + s;
        const size_t on = ND * ((size_t)l * nxy + own_off);
        // loads for the next steps are issued first and consumed after this
        // layer's elements: element materials of layer l + 1, plane l + 2
        // (materials first: the wait for the plane then covers them, and no
        // wait is left for after this layer's stores).  The last step issues
        // them too (in bounds, unused): every step then has the same loads
        // and none is pending across the loop's back edge.
        const uint32_t kcur = kraw;
        mk = mat_key(lw1);
        // CG1: dinv and diag of the owned node (its key is the element's,
        // loaded a step ahead), read after this layer's elements
        T mo[CG1 || LEAN ? ND : 1], dgo[CG1 || LEAN ? ND : 1];
        // LEAN: M^-1 and diag(M) of the owned node, and (form 3) its z
        T zo[LEAN ? ND : 1];
        if constexpr (LEAN) {
#pragma unroll
            for (int a = 0; a < ND; a++) {
                mo[a] = fz.kinv ? fz.kinv[ND * kcur + a] : T(1);
                dgo[a] = fz.kinv ? fz.kinv[(size_t)65536 * ND + ND * kcur + a] : T(1);
                zo[a] = fz.lean == 3 ? q[on + a] : T(0);
            }
        }
        // owned node's d (and, CG1, its updated z), read from the planes of
        // layer l here, before any work-item replaces them (after the
        // barrier that follows this layer's elements).  Every work-item reads
        // its in-tile entry; only owners use it.
        T own_u[ND], own_z[CG1 ? ND : 1];
        {
            const T *uu = cur + (ec * NR + er) * ND;
#pragma unroll
            for (int a = 0; a < ND; a++) own_u[a] = uu[a];
            if constexpr (CG1) {
                const T *zz = zcur + (ec * NR + er) * ND;
#pragma unroll
                for (int a = 0; a < ND; a++) own_z[a] = zz[a];
            }
        }
        fetch(lw2);
        lw1 = lw2;
        lw2 = next_layer(lw2);
        element(cur, nxt, m, A, Af);
        __syncthreads();
        // the prefetched plane goes to LDS before this layer's node results
        // are stored, so waiting for the loads does not wait for the stores
        store(cur, zcur);  // layer l + 2 replaces layer l
        // first use of the key loaded at the top of the step: without this
        // the compiler consumes it (and reuses its register) right after the
        // load, exposing the load's latency at the top of every step
        asm volatile("" : "+v"(mk));
        m = material(mk);
        kraw = mk;
        pin_here(m);
        // (loaded here rather than before the elements: six fewer registers
        // live across them, 3 % faster at 256^3; the table rows of the few
        // keys in use stay in the caches)
        if constexpr (CG1) {
#pragma unroll
            for (int a = 0; a < ND; a++) {
                mo[a] = fz.kinv[ND * kcur + a];
                dgo[a] = fz.kinv[(size_t)65536 * ND + ND * kcur + a];
            }
        }
        if (own) {
            halo_row(A, Af);
            T o[ND];
#pragma unroll
            for (int a = 0; a < ND; a++) {
                o[a] = acc[a] + A[a] + bleft[a];
                acc[a] = Af[a] + bleft[ND + a];
            }
            T *qo = q + on;
            if constexpr (LEAN) {
                if (fz.lean == 1) {
#pragma unroll
                    for (int a = 0; a < ND; a++) dsum += (double)own_u[a] * o[a];
                } else if (fz.lean == 2) {
#pragma unroll
                    for (int a = 0; a < ND; a++) dsum += (double)o[a] * (double)(mo[a] * o[a]);
                } else {
#pragma unroll
                    for (int a = 0; a < ND; a++) {
                        const T zn = zo[a] - fz.alpha * (mo[a] * o[a]);
                        qo[a] = zn;
                        dsum += (double)zn * (double)(dgo[a] * zn);
                    }
                }
            } else if constexpr (CG1) {
                // (FP32: q and d of this iteration are next read an
                // iteration, about 1 GB of traffic, later; non-temporal
                // stores keep them from displacing the plane halos in L2,
                // 2 % faster at 256^3; in FP64 they were slower)
                if constexpr (sizeof(T) == 4) {
#pragma unroll
                    for (int a = 0; a < ND; a++) __builtin_nontemporal_store(mo[a] * o[a], &qo[a]);
                } else {
#pragma unroll
                    for (int a = 0; a < ND; a++) qo[a] = mo[a] * o[a];
                }
            } else {
#pragma unroll
                for (int a = 0; a < ND; a++) qo[a] = o[a];
            }
            if constexpr (FUSE) {
                T *dn = fz.dnew + on;
#pragma unroll
                for (int a = 0; a < ND; a++) {
                    if constexpr (CG1 && sizeof(T) == 4) __builtin_nontemporal_store(own_u[a], &dn[a]);
                    else dn[a] = own_u[a];
                }
            }
            if constexpr (CG1) {
                if (fz.znew) {
                    T *zn = fz.znew + on;
#pragma unroll
                    for (int a = 0; a < ND; a++) zn[a] = own_z[a];
                }
#pragma unroll
                for (int a = 0; a < ND; a++) {
                    zq += (double)own_z[a] * o[a];
                    qmq += (double)o[a] * (double)(mo[a] * o[a]);
                    zdz += (double)own_z[a] * (double)(dgo[a] * own_z[a]);
                }
            }
            if (DOT && !LEAN) {
#pragma unroll
                for (int a = 0; a < ND; a++) dsum += (double)own_u[a] * o[a];
            }
        }
        {
            T *t = cur; cur = nxt; nxt = t;
                t = zcur; zcur = znxt; znxt = t;
        }
        __syncthreads();
    }
    return TileSums{dsum, zq, qmq, zdz};
}

}  // namespace chfem
