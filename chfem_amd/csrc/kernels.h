// kernels.h -- host-side launchers of the gfx950 kernels (kernels.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace chfem {

// upper bound on the blocks of any reducing launch: 1-D vector kernels use at
// most MAX_RED_GRID blocks; tiled operators at most MAX_PARTIAL_BLOCKS (the
// partials buffer holds 8 values per block)
#define MAX_RED_GRID 8192
#define MAX_PARTIAL_BLOCKS 65536

enum Physics { PHY_THERMAL = 0, PHY_ELASTIC = 1, PHY_FLUID = 2 };

// Device-resident scalars of one Krylov solve (FP64).
enum ScalarSlot {
    S_DELTA = 0, S_DELTA_OLD, S_DELTA0, S_DQ, S_ALPHA, S_STOP, S_RES0, S_TMP0, S_TMP1,
    S_TMP2, S_TMP3, S_TMP4, S_TMP5, S_TMP6, S_TMP7,
    // guarded Krylov loops: done flag, iterations counted on the device,
    // stop threshold on r.z
    S_DONE = 16, S_ITER, S_THRESH,
    // single-kernel CG: beta of the last finished iteration (the one that
    // formed the newest direction)
    S_BPREV,
    // alpha of guarded iteration j (j counted in S[S_ITER]) at
    // S[S_AHIST + j % S_AHIST_LEN], for the ring flush of the iterate
    S_AHIST = 24, S_AHIST_LEN = 32, S_COUNT = S_AHIST + S_AHIST_LEN
};

// How the finishing kernel of a fused reduction publishes its total.
enum Finalize {
    FIN_STORE = 0,      // S[slot] = sum
    FIN_ALPHA = 1,      // S[S_DQ] = sum; S[S_ALPHA] = S[S_DELTA] / sum
    FIN_DELTA = 2,      // S[S_DELTA_OLD] = S[S_DELTA]; S[S_DELTA] = sum
    // single-kernel CG, sums (d.q, z.q, q.dinv q, z.diag z): delta = z.diag z
    // (r.z of this iteration's z), alpha = delta / d.q, delta_new = delta -
    // 2 alpha z.q + alpha^2 q.dinv q (r.z of the next z), S[DELTA_OLD] =
    // delta, S[DELTA] = delta_new; guarded: counts the iteration and sets
    // S[DONE] when |delta_new| <= S[THRESH]
    FIN_CG1 = 3,
    // memory-lean MINRES: S[S_STOP] = sum (the residual measure r.M^-1 r);
    // guarded: counts the iteration and sets S[DONE] when |sum| <= S[THRESH]
    FIN_STOP = 4,
    // S[S_DELTA_OLD] = S[S_DELTA]; S[S_DELTA] = sum, never counted
    FIN_DELTA_NOCOUNT = 5,
};

struct Reduce {
    double *partials;    // >= (blocks of the launch) * (values reduced) doubles
    double *S;           // scalar slots
    int fin;
    int slot;
    // guarded: the launch and its finishing kernel do nothing once S[S_DONE]
    // is set; a guarded FIN_DELTA finish counts the iteration in S[S_ITER]
    // and sets S[S_DONE] when |r.z| <= S[S_THRESH].  The host can then queue
    // several iterations between synchronisations.
    int guard;
};

// Scalar density field (EBE only): element e's local matrices are scaled by
// d[e] * (1/65535) * (fmax - fmin) + fmin (chfem/src/femhmg/femhmg.c:1760-1835);
// d == nullptr: no field.
struct Sdf {
    const uint16_t *d;
    double fmin, fmax;
};

// Geometry and material data of one operator, all device pointers.
struct OpDesc {
    int physics;         // Physics
    int ebe;             // 0 NBN (keys), 1 EBE (element map)
    int ny, nx, nz;      // rows, cols, layers (elements)
    uint32_t N;          // nodes = elements
    size_t ndof;
    const uint16_t *keys;      // NBN packed keys, or EBE element material map
    const void *K;             // per-material dense local matrices in the operator precision
    // Stokes
    uint32_t nvel, nborder;
    const uint32_t *dof2node;  // DOF id -> periodic node
    const uint32_t *node2dof;  // periodic node -> DOF id
    const uint8_t *bkeys;      // border node fluid keys
    // per-node fluid keys (255 pore, 1..254 border, 0 solid); non-null
    // selects the Walsh-basis Stokes operator (stokes.hip)
    const uint8_t *fkeys;
    // isotropic elastic materials: {lambda, mu} per material, or conduction:
    // {k, 0} per material (precision of K); non-null selects the Walsh-basis
    // operator (walsh.hip)
    const void *lamu;
    int tile_tr, tile_tc, tile_lz;
    // inverse Jacobi diagonal per packed key, kinv[key * nd + a] (65536 keys,
    // precision of K), followed by the diagonal itself at kinv[65536 * nd +
    // key * nd + a]: non-null selects the key-Jacobi vector kernels
    // (thermal/elastic NBN)
    const void *kinv;
    int nmat;
    // EBE: 1 selects the reference-style element kernel with atomics instead
    // of the matrix-core kernel (ebe_mfma.hip; up to 4 materials)
    int ebe_atomic;
    // EBE matrix-core kernel in FP32: 1 selects the 16x16x4 column-tile form
    // (the FP64 form) instead of the 4x4x1 block form
    int ebe_tiles16;
    Sdf sdf;
};

template <typename T> void launch_apply(const OpDesc &op, const T *v, T *q, const Reduce *red, hipStream_t s);

// Jacobi preconditioner of the memory-lean solvers, evaluated per DOF
// without a stored diagonal where the model allows it:
//   JAC_NONE   no preconditioner (M = I)
//   JAC_KEYS   thermal/elastic NBN: kinv[key * nd + a], diagonal at
//              kinv[65536 * nd + key * nd + a], key = keys[dof / nd]
//   JAC_FLUID  Stokes: velocity DOFs vinv, pore pressures pinv, border
//              pressures btab[border key] (256 entries: 1 / -(sum of the
//              pressure diagonal over the key's pore elements))
//   JAC_STORED a stored inverse diagonal dinv[dof] (EBE, generic operators)
enum JacKind { JAC_NONE = 0, JAC_KEYS = 1, JAC_FLUID = 2, JAC_STORED = 3 };
template <typename T>
struct Jac {
    int kind;
    int nd;
    const T *kinv;
    const uint16_t *keys;
    const T *dinv;
    T vinv, pinv;
    const T *btab;
    const uint8_t *bkeys;
    uint32_t nvel;
};

// Operator forms of the memory-lean solvers (3 or 2 solver vectors, no
// stored q = A v):
//   LEAN_VAV  reduce v.(A v), nothing stored
//   LEAN_QMQ  reduce (A v).M^-1 (A v), nothing stored
//   LEAN_ZUPD z <- z - S[S_ALPHA] M^-1 A v in place, reduce z.diag(M) z
//             (= r.M^-1 r for the residual r = diag(M) z)
enum LeanForm { LEAN_VAV = 1, LEAN_QMQ = 2, LEAN_ZUPD = 3 };
// One lean operator launch (red required, finished as red->fin says).  The
// Walsh-basis NBN operators and the Walsh-basis Stokes operator evaluate the
// form in their epilogue; every other operator writes A v to qtmp (a work
// vector the solver then holds) and a streaming kernel evaluates the form.
template <typename T>
void launch_lean_apply(const OpDesc &op, const Jac<T> &jac, int form, const T *v, T *z, T *qtmp, const Reduce *red,
                       hipStream_t s);
// whether launch_lean_apply needs qtmp for this operator
bool lean_needs_qtmp(const OpDesc &op);
// the lean forms in the epilogues of the Walsh-basis NBN operator (kinv:
// key-Jacobi table, nullptr: no preconditioner) and of the Walsh-basis
// Stokes operator (Jacobi per DOF class, jac.kind JAC_FLUID or JAC_NONE)
template <typename T>
void launch_walsh_lean(const OpDesc &op, int form, const T *v, T *z, const void *kinv, const Reduce *red,
                       hipStream_t s);
template <typename T>
void launch_stokes_lean(const OpDesc &op, const Jac<T> &jac, int form, const T *v, T *z, const Reduce *red,
                        hipStream_t s);
// z = M^-1 z in place, red <- z.diag(M) z (= r.M^-1 r of the r z held)
template <typename T> void launch_lean_to_z(const Jac<T> &jac, T *z, size_t n, const Reduce *red, hipStream_t s);
// first: d = z; else (x non-null: x += S[S_ALPHA] d) then d = z + beta d
// with beta = S[S_DELTA] / S[S_DELTA_OLD] (guarded by red.guard via S)
template <typename T> void launch_lean_dir(T *x, T *d, const T *z, const double *S, int first, int guard, size_t n,
                                           hipStream_t s);
// x += S[S_ALPHA] d
template <typename T> void launch_lean_axpy(T *x, const T *d, const double *S, size_t n, hipStream_t s);
// On-the-fly reduction of the solution (chfem/src/femhmg/cudapcg/solvers/
// xreduce.cu:4-26) for the 2-vector solvers, after iteration j's alpha, with
// p = nvpn values per node (1..4) over the first nh entries:
//   mode 0 (NONE):    hist[p (j+1) + a] = alpha * sum_i d[p i + a]
//   mode 1 (ONLYDIR): xr[i] += alpha d[p i + shift]; hist[p (j+1) + shift] = sum_i xr[i]
//   mode 2 (FULL):    xr[i] += alpha d[i]; hist[p (j+1) + a] = sum_i xr[p i + a]
// with j = S[S_ITER] (the guarded count of finished iterations) and alpha =
// S[S_ALPHA]; entries past hist_len iterations are not recorded.  x0
// non-null: the initial entry instead, hist[a] = sum_i x0[p i + a], and xr
// seeded from x0 (ONLYDIR: xr[i] = x0[p i + shift]; FULL: xr = x0).
template <typename T> void launch_xreduce(T *xr, const T *d, const T *x0, double *hist, size_t hist_len, int mode,
                                          int shift, int nvpn, size_t nh, const Reduce &red, hipStream_t s);
// S[slot] = max |v_i| over n entries (scaled: max |v_i / dinv_i|, the
// residual of a preconditioned z), deterministic
template <typename T> void launch_absmax(const T *v, size_t n, const Jac<T> *scaled, double *partials, double *S,
                                         int slot, hipStream_t s);

// EBE operator on the matrix cores (ebe_mfma.hip), thermal/elastic, nmat <= 4
template <typename T> void launch_ebe_mfma(const OpDesc &op, const T *v, T *q, const Reduce *red, hipStream_t s);

// Sums the nv x nb partials a reducing launch of nb blocks left in
// red.partials (fixed order) and publishes them as red.fin says.  Every
// launcher that takes a Reduce calls it after its kernel.
void launch_reduce_finish(const Reduce &red, unsigned nb, int nv, hipStream_t s);

// dinv[i] = 1 / diag(A)_i in precision T, from the keys (thermal/elastic) or the
// Stokes pore/border structure.
template <typename T> void launch_jacobi(const OpDesc &op, T *dinv, hipStream_t s);

// Vector kernels of the solvers (n entries, precision T, reductions FP64).
template <typename T> void launch_fill(T *x, double val, size_t n, hipStream_t s);
template <typename Tin, typename Tout> void launch_convert(const Tin *in, Tout *out, size_t n, hipStream_t s);
// out = dinv*v ; optional reduction of v.out
template <typename T> void launch_precond(const T *dinv, const T *v, T *out, size_t n, const Reduce *red, hipStream_t s);
// d = z + beta*d, beta = S[DELTA]/S[DELTA_OLD] (first: d = z)
template <typename T> void launch_pcg_dir(T *d, const T *z, const double *S, int first, size_t n, hipStream_t s,
                                          int guard = 0);
// x += a d ; r -= a q ; z = dinv r ; reduce r.z  (a = S[ALPHA])
template <typename T> void launch_pcg_update(T *x, T *r, const T *d, T *z, const T *dinv, const Reduce *red,
                                             size_t n, hipStream_t s);
// r -= q ; z = dinv r ; reduce r.z
template <typename T> void launch_sub_precond(T *r, const T *q, const T *dinv, T *z, const Reduce *red, size_t n,
                                              hipStream_t s);
// MINRES: dnew = rt + b dold (dnew may be dold) ; q = s + b q ; reduce q.(dinv q) ;
// b = S[DELTA]/S[DELTA_OLD] (first: d = rt, q = s)
template <typename T> void launch_minres_dir(const T *dold, T *dnew, T *q, const T *rt, const T *sv, const T *dinv,
                                             const double *S, int first, const Reduce *red, size_t n, hipStream_t s);
// MINRES: rn = rt/dinv - a q ; rt = dinv rn ; x += a d (x == nullptr: the
// iterate is left to the direction ring's flush) ; reduce rn.rt
template <typename T> void launch_minres_update(T *x, T *rt, const T *d, const T *q, const T *dinv, const Reduce *red,
                                                size_t n, hipStream_t s);
// FP64 outer refinement: x64 += e32
void launch_add_correction(double *x, const float *e, size_t n, hipStream_t s);
// y += x
template <typename T> void launch_add(T *y, const T *x, size_t n, hipStream_t s);

// Key-Jacobi PCG vector kernels (op.kinv non-null; thermal/elastic NBN):
// dinv is looked up by key and z = dinv r is never stored.
//   kj_precond:    out = dinv v ; red <- v.out
//   kj_pcg_dir:    d = dinv r + (S[delta]/S[delta_old]) d
//   kj_pcg_update: x += alpha d ; r -= alpha q ; red <- r.(dinv r)
//   kj_to_z:       r <- z = dinv r in place ; red <- r.z
//   kj_zupdate:    z -= alpha dinv q ; red <- r.z = z.(diag z)
// op.kinv holds dinv per key, then diag per key at offset 65536 * nd.
//   kj_sub_precond: r -= q ; z = dinv r ; red <- r.z
template <typename T> void launch_kj_precond(const OpDesc &op, const T *v, T *out, const Reduce *red, hipStream_t s);
template <typename T> void launch_kj_pcg_dir(const OpDesc &op, T *d, const T *r, const double *S, hipStream_t s,
                                             int guard = 0);
template <typename T> void launch_kj_pcg_update(const OpDesc &op, T *x, T *r, const T *d, const T *q,
                                                const Reduce *red, hipStream_t s);
template <typename T> void launch_kj_sub_precond(const OpDesc &op, T *r, const T *q, T *z, const Reduce *red,
                                                 hipStream_t s);
template <typename T> void launch_kj_to_z(const OpDesc &op, T *r, const Reduce *red, hipStream_t s);
template <typename T> void launch_kj_zupdate(const OpDesc &op, T *z, const T *q, const Reduce *red, hipStream_t s);
// Walsh-basis Stokes operator over layer-march tiles (stokes.hip; needs
// op.fkeys), and the per-node fluid keys it reads, built from the DOF maps
template <typename T> void launch_stokes_walsh(const OpDesc &op, const T *v, T *q, const Reduce *red, hipStream_t s);
void launch_fluid_node_keys(const OpDesc &op, uint8_t *out, hipStream_t s);
// dot reduction of two vectors
template <typename T> void launch_dot(const T *a, const T *b, size_t n, const Reduce *red, hipStream_t s);

// Right-hand side of load case dir (FP64), and effective-tensor column.
struct LoadCase { int axis; int S[4]; int comp; double L; };
void launch_rhs(const OpDesc &op, const double *K64, const LoadCase &lc, int dir, double *b, hipStream_t s);
// seamCB: per material, per C row: sum over seam columns of CB (already x L)
void launch_updateC(const OpDesc &op, const double *CB64, const double *seamCB, const LoadCase &lc,
                    const double *x, const Reduce *red, hipStream_t s);
template <typename T> void launch_fluid_velocity_sum(const OpDesc &op, const T *x, const Reduce *red, hipStream_t s);
// thermal expansion load (elastic); da[m] = (D_m alpha_m (1,1,1,0,0,0))_xx per material
void launch_rhs_thermal_expansion(const OpDesc &op, const double *da, double *b, hipStream_t s);
// nodal flux (thermal, 3 per node) or stress (elastic, 6 per node) for field
// export; F: per material, per corner, the field matrix (NF x 8*nd); jump = 0
// leaves the seam jump out (thermal expansion load case)
void launch_fields(const OpDesc &op, const double *F, const LoadCase &lc, int jump, const double *x, double *out,
                   hipStream_t s);

// x0 from a half-resolution solution (rows/cols/layers = fine element counts)
void launch_interpl2(const double *coarse, double *fine, int rows, int cols, int layers, int stride, hipStream_t s);

int reduce_blocks(size_t n);

}  // namespace chfem

namespace chfem {
// Walsh-basis operators (walsh.hip) for isotropic elastic materials (lamu:
// {lambda, mu} per material, when every local matrix equals lambda*A + mu*B
// for the unit-voxel trilinear element) and for conduction (lamu: {k, 0} per
// material, when every local matrix is k times the unit conduction matrix).
struct WalshTiles { int TR, TC, LZ; };
template <typename T> struct DirFuse;
// fz non-null: v is d_old and the launch forms d = z + beta d_old
// (walsh.h DirFuse), writes it to fz->dnew and reduces d.q (red required).
// ev non-null: {start, stop} events bound to the operator kernel's dispatch
// (hipExtLaunchKernel), timing that kernel alone.
template <typename T>
void launch_walsh(const OpDesc &op, const T *lamu, const T *v, T *q, const Reduce *red, WalshTiles t, hipStream_t s,
                  const DirFuse<T> *fz = nullptr, const hipEvent_t *ev = nullptr);

// Jacobi-PCG step on the preconditioned residual z = dinv r with the
// direction update inside the operator (Walsh-basis thermal or isotropic
// elastic NBN with key-Jacobi, see apply_dir_ok): d = z + beta d_old (beta =
// S[DELTA]/S[DELTA_OLD]; dold == nullptr: d = z), dnew <- d, q = A d,
// red <- d.q.  dnew must not alias dold, z or q.  The loop's other kernels:
// launch_kj_to_z (r -> z), launch_kj_zupdate, and launch_ring_flush for the
// iterate.
bool apply_dir_ok(const OpDesc &op);
template <typename T>
void launch_apply_dir(const OpDesc &op, const T *z, const T *dold, T *dnew, T *q, const Reduce *red, hipStream_t s,
                      const hipEvent_t *ev = nullptr);
// Single-kernel Jacobi-PCG iteration (same operators as launch_apply_dir):
// z = zold - S[ALPHA] qold (qold holds dinv q of the previous iteration;
// skipped unless pend), d = z + beta dold (dold == nullptr: d = z), znew <- z,
// dnew <- d, q = A d with qnew <- dinv q, red (FIN_CG1) <- d.q, z.q,
// q.dinv q, z.diag z.  The new buffers must not alias the old ones or each
// other.  znew == nullptr: z is not written.  dold2 non-null (with dold and
// pend): zold is not read; z_old = dold - S[BPREV] dold2 (dold2: the
// direction before dold, which the previous iteration updated to dold with
// beta S[BPREV]).
template <typename T>
void launch_apply_cg1(const OpDesc &op, const T *zold, const T *qold, const T *dold, T *znew, T *dnew, T *qnew,
                      int pend, const Reduce *red, hipStream_t s, const hipEvent_t *ev = nullptr,
                      const T *dold2 = nullptr);
// The iterate of the fused CG loops: iteration j writes its direction to
// ring slot j % w; x += sum of S[S_AHIST + (j - base) % S_AHIST_LEN] *
// ring[j % w] over the iterations j in [j0, j1) that ran (j - base <
// S[S_ITER], the loop's guarded count since base).  w <= S_AHIST_LEN.
constexpr int RING_MAX = 32;
template <typename T>
struct Ring {
    const T *p[RING_MAX];
    int w;
};
template <typename T>
void launch_ring_flush(T *x, const Ring<T> &ring, unsigned j0, unsigned j1, unsigned base, const double *S, size_t n,
                       hipStream_t s);
}  // namespace chfem
