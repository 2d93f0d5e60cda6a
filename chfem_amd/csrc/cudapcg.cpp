// cudapcg.cpp -- the solver plugin C ABI (include/cudapcg.h) over the gfx950
// kernels.  One global solver per process, as in the reference
// (chfem/src/femhmg/cudapcg/cudapcg.cu:37); all work is queued on one HIP
// stream owned by the solver.
//
// Krylov loops follow the reference's iteration order and L2 stopping rule
// (chfem/src/femhmg/cudapcg/solvers/pcg.cu:28-243, pminres.cu:28-247,
// stopping_criteria.cu:7-14).  Differences by design: alpha/beta and the dot
// products stay on the device (fused into the vector kernels), and the host
// reads one 128-byte scalar block per iteration for the stopping test.
#include <hip/hip_runtime.h>
#include <math.h>
#include <omp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <type_traits>
#include <vector>

#include "solver.h"
#include "voxel.h"

namespace chfem {

struct Solver {
    cudapcgFlag_t analysis_flag = CUDAPCG_THERMAL_3D;
    cudapcgFlag_t parStrategy_flag = CUDAPCG_NBN;
    cudapcgModel_t *model = nullptr;
    unsigned long count = 0;
    char *header = nullptr;
    cudapcgTol_t num_tol = CUDAPCG_TOLERANCE;
    unsigned max_iterations = CUDAPCG_MAX_ITERATIONS;
    cudapcgFlag_t xreduce_flag = CUDAPCG_XREDUCE_FULL;
    unsigned xreduce_shift = 0;
    double xreduce_scale = 1.0, reduce_stab = 0.0001;
    unsigned solver_flag = CUDAPCG_CG_SOLVER;
    cudapcgFlag_t precond = CUDAPCG_TRUE;
    cudapcgFlag_t resnorm = CUDAPCG_L2_NORM;
    cudapcgFlag_t precision = CHFEM_PREC_FP64;
    // results
    double residual = 1.0;
    unsigned iteration = 0, sweeps = 0;
    // initial guess of the last solve: relative residual, kept or replaced
    double x0_residual = 0;
    unsigned x0_used = 0;
    double total_time = 0, mean_time = 0, solve_seconds = 0;
    cudapcgFlag_t found = CUDAPCG_FALSE;
    // device state
    DeviceModel dm;
    double *r64 = nullptr, *x64 = nullptr, *d64 = nullptr, *q64 = nullptr, *s64 = nullptr, *dinv64 = nullptr;
    float *x32 = nullptr, *r32 = nullptr, *d32 = nullptr, *q32 = nullptr, *s32 = nullptr, *dinv32 = nullptr;
    // second z and q buffers of the single-kernel CG loop
    double *z2_64 = nullptr, *q2_64 = nullptr;
    float *z2_32 = nullptr, *q2_32 = nullptr;
    // direction ring of the fused CG loops beyond d and s (slots 2 .. w - 1;
    // w = 0: not set up yet), per precision
    double *ring64[RING_MAX] = {};
    float *ring32[RING_MAX] = {};
    int ring_w64 = 0, ring_w32 = 0;
    unsigned ring_try64 = 0, ring_try32 = 0;  // widest ring asked for so far
    bool arrays = false, user_arrays = false, x0_set = false, dinv_valid = false;
    double *partials = nullptr, *S = nullptr, *Sh = nullptr;
    size_t partials_n = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // inverse Jacobi diagonal per packed key (65536 x nd) for the key-Jacobi
    // vector kernels, FP64 and FP32
    double *kinv64 = nullptr;
    float *kinv32 = nullptr;
    // operator tiles chosen by the autotuner (rows, columns, layers per block)
    bool autotune = false, tuned = false;
    int tile64[3] = {32, 8, 16}, tile32[3] = {32, 8, 16};
    // CG loop form per precision (see fuse_mode_env), set by the autotuner
    int mode64 = 1, mode32 = 1;
    // isotropic elastic materials for the Walsh-basis operator
    double *lamu64 = nullptr;
    float *lamu32 = nullptr;
    bool walsh_ok = false;
    // per-node fluid keys for the Walsh-basis Stokes operator (N bytes)
    uint8_t *fkeys = nullptr;
    size_t fkeys_n = 0;
    // operator-launch timing
    bool timing = false;
    // operator descriptors built (make_ops) for the current model
    bool ops_ready = false;
    // Stokes Jacobi per DOF class for the memory-lean solvers: velocity and
    // pore-pressure inverse diagonals, and the border-pressure inverse per
    // fluid key (256 entries, device)
    double fl_vinv = 1.0, fl_pinv = 1.0;
    double *btab64 = nullptr;
    // two-vector solvers: the reduced solution (xreduce FULL: nhmgvars
    // entries, ONLYDIR: nhmgvars / nvarspernode) and its per-iteration
    // history ((max_iterations + 1) x nvarspernode), and the values the last
    // solve returns
    double *xr64 = nullptr, *hist = nullptr;
    size_t xr_n = 0, hist_n = 0;
    double xred_out[4] = {0, 0, 0, 0};
    int xred_count = 0;
    hipEvent_t opA = nullptr, opB = nullptr;
    double op_ms = 0;
    unsigned op_count = 0;
};

static Solver *g = nullptr;

static cudapcgFlag_t ctor_simple(cudapcgModel_t **ptr, const void *data)
{
    if (!data) return CUDAPCG_FALSE;
    *ptr = (cudapcgModel_t *)data;
    return CUDAPCG_TRUE;
}
static cudapcgFlag_t (*g_ctor)(cudapcgModel_t **, const void *) = ctor_simple;

static int make_ops();

DeviceModel *active_model()
{
    if (!g) return nullptr;
    if (g->model) make_ops();
    return &g->dm;
}
const DeviceModel *built_model() { return g && g->model && g->ops_ready ? &g->dm : nullptr; }
static bool is_two_vector();
int reduced_solution(double *out)
{
    if (!g || !is_two_vector() || g->xred_count == 0) return 0;
    for (int a = 0; a < g->xred_count; a++) out[a] = g->xred_out[a];
    return g->xred_count;
}
hipStream_t active_stream() { return g ? g->stream : nullptr; }
double *rhs_buffer() { return g ? g->r64 : nullptr; }
double *solution_buffer() { return g ? g->x64 : nullptr; }
double *scalars_host() { return g ? g->Sh : nullptr; }

Reduce make_reduce(int fin, int slot)
{
    Reduce r;
    r.partials = g->partials;
    r.S = g->S;
    r.fin = fin;
    r.slot = slot;
    r.guard = 0;
    return r;
}

static Reduce make_guarded(int fin, int slot)
{
    Reduce r = make_reduce(fin, slot);
    r.guard = 1;
    return r;
}

// iterations queued between host synchronisations in the CG loops
// (CHFEM_CG_BATCH, default 16; 1 synchronises every iteration)
static unsigned cg_batch()
{
    static unsigned b = 0;
    if (!b) {
        const char *e = getenv("CHFEM_CG_BATCH");
        const int v = e ? atoi(e) : 16;
        b = v < 1 ? 1u : (unsigned)v;
    }
    return b;
}

// arms the device-side stop test of a guarded loop: S_DONE = 0,
// S_ITER = 0, S_THRESH = thresh
static int arm_guard(double thresh)
{
    g->Sh[S_DONE] = 0.0;
    g->Sh[S_ITER] = 0.0;
    g->Sh[S_THRESH] = thresh;
    if (hipMemcpyAsync(g->S + S_DONE, g->Sh + S_DONE, 3 * sizeof(double), hipMemcpyHostToDevice, g->stream) !=
        hipSuccess)
        return 0;
    return hipStreamSynchronize(g->stream) == hipSuccess;
}

int sync_scalars()
{
    if (hipMemcpyAsync(g->Sh, g->S, sizeof(double) * S_COUNT, hipMemcpyDeviceToHost, g->stream) != hipSuccess) return 0;
    if (hipStreamSynchronize(g->stream) != hipSuccess) return 0;
    return 1;
}

static int set_scalar(int slot, double v)
{
    g->Sh[slot] = v;
    if (hipMemcpyAsync(g->S + slot, g->Sh + slot, sizeof(double), hipMemcpyHostToDevice, g->stream) != hipSuccess)
        return 0;
    return hipStreamSynchronize(g->stream) == hipSuccess;
}

// q = A v inside a Krylov loop; with timing on the launch is bracketed by
// events whose interval is collected at the next host synchronisation (the
// fused CG forms below bind the events to the operator kernel's dispatch
// instead, so their interval is that kernel's run time alone).
static bool op_pending = false;
template <typename T>
static void timed_apply(const OpDesc &op, const T *v, T *q, const Reduce *red)
{
    if (g->timing) hipEventRecord(g->opA, g->stream);
    launch_apply<T>(op, v, q, red, g->stream);
    if (g->timing) {
        hipEventRecord(g->opB, g->stream);
        op_pending = true;
    }
}

// form of the CG loops on the Walsh-basis operators (CHFEM_FUSE_DIR): 0
// separate operator, direction and update kernels; 1 direction and iterate
// updates inside the operator, z update kernel; 2 single-kernel iteration,
// the z update inside the operator too.  Unset or "auto": the autotuner
// times forms 1 and 2 with every tile and keeps the faster per precision
// (form 1 when the tiles are fixed).
static int fuse_mode_env()
{
    const char *e = getenv("CHFEM_FUSE_DIR");
    if (!e || !strcmp(e, "auto")) return -1;
    const int m = atoi(e);
    return m < 0 ? 0 : (m > 2 ? 2 : m);
}
template <typename T> static int loop_mode();
static bool fuse_dir_enabled() { return fuse_mode_env() != 0; }

template <typename T>
static void timed_apply_cg1(const OpDesc &op, const T *zold, const T *qold, const T *dold, T *znew, T *dnew, T *qnew,
                            int pend, const Reduce *red, bool timed, const T *dold2 = nullptr)
{
    const hipEvent_t ev[2] = {g->opA, g->opB};
    const bool on = timed && g->timing;
    launch_apply_cg1<T>(op, zold, qold, dold, znew, dnew, qnew, pend, red, g->stream, on ? ev : nullptr, dold2);
    if (on) op_pending = true;
}

template <typename T>
static void timed_apply_dir(const OpDesc &op, const T *r, const T *dold, T *dnew, T *q, const Reduce *red)
{
    const hipEvent_t ev[2] = {g->opA, g->opB};
    launch_apply_dir<T>(op, r, dold, dnew, q, red, g->stream, g->timing ? ev : nullptr);
    if (g->timing) op_pending = true;
}

static int sync_loop()
{
    if (!sync_scalars()) return 0;
    if (op_pending) {
        float ms = 0;
        if (hipEventElapsedTime(&ms, g->opA, g->opB) == hipSuccess) {
            g->op_ms += ms;
            g->op_count++;
        }
        op_pending = false;
    }
    return 1;
}

static int phys(cudapcgFlag_t a)
{
    return a == CUDAPCG_THERMAL_3D ? PHY_THERMAL : (a == CUDAPCG_ELASTIC_3D ? PHY_ELASTIC : PHY_FLUID);
}

static size_t nvars() { return g && g->model ? (size_t)g->model->nvars : 0; }

template <typename T> static void dfree(T *&p)
{
    if (p) hipFree(p);
    p = nullptr;
}

static cudapcgFlag_t free_solver_arrays()
{
    if (!g) return CUDAPCG_FALSE;
    if (!g->x0_set) dfree(g->x64);
    dfree(g->d64); dfree(g->q64); dfree(g->s64); dfree(g->dinv64);
    dfree(g->x32); dfree(g->r32); dfree(g->d32); dfree(g->q32); dfree(g->s32); dfree(g->dinv32);
    dfree(g->z2_64); dfree(g->q2_64); dfree(g->z2_32); dfree(g->q2_32);
    dfree(g->xr64); dfree(g->hist);
    g->xr_n = g->hist_n = 0;
    for (int k = 0; k < RING_MAX; k++) { dfree(g->ring64[k]); dfree(g->ring32[k]); }
    g->ring_w64 = g->ring_w32 = 0;
    g->ring_try64 = g->ring_try32 = 0;
    dfree(g->partials); dfree(g->S);
    if (g->Sh) hipHostFree(g->Sh);
    g->Sh = nullptr;
    g->arrays = false;
    g->user_arrays = false;
    g->dinv_valid = false;
    return CUDAPCG_TRUE;
}

static bool is_lean();
static bool is_two_vector();
static cudapcgFlag_t alloc_solver_arrays()
{
    if (!g || !g->model) return CUDAPCG_FALSE;
    const size_t n = nvars();
    const size_t b8 = n * sizeof(double), b4 = n * sizeof(float);
    if (is_lean()) {
        // memory-lean solvers: z (the RHS buffer) and d, plus x for CG3 and
        // MINRES3; a stored diagonal or an operator work vector only where
        // the model needs one (allocated by the solver on first use)
        if (!is_two_vector() && !g->x64) CHECK_HIP(hipMalloc(&g->x64, b8));
        if (!g->d64) CHECK_HIP(hipMalloc(&g->d64, b8));
    } else {
    if (!g->x64) CHECK_HIP(hipMalloc(&g->x64, b8));
    if (!g->d64) CHECK_HIP(hipMalloc(&g->d64, b8));
    if (!g->q64) CHECK_HIP(hipMalloc(&g->q64, b8));
    // s: MINRES's A M^-1 r, or CG's second direction buffer (fused direction update)
    if (!g->s64) CHECK_HIP(hipMalloc(&g->s64, b8));
    if (!g->z2_64) CHECK_HIP(hipMalloc(&g->z2_64, b8));
    if (!g->q2_64) CHECK_HIP(hipMalloc(&g->q2_64, b8));
    if (!g->dinv64) CHECK_HIP(hipMalloc(&g->dinv64, b8));
    if (g->precision == CHFEM_PREC_MIXED) {
        if (!g->x32) CHECK_HIP(hipMalloc(&g->x32, b4));
        if (!g->r32) CHECK_HIP(hipMalloc(&g->r32, b4));
        if (!g->d32) CHECK_HIP(hipMalloc(&g->d32, b4));
        if (!g->q32) CHECK_HIP(hipMalloc(&g->q32, b4));
        if (!g->s32) CHECK_HIP(hipMalloc(&g->s32, b4));
        if (!g->z2_32) CHECK_HIP(hipMalloc(&g->z2_32, b4));
        if (!g->q2_32) CHECK_HIP(hipMalloc(&g->q2_32, b4));
        if (!g->dinv32) CHECK_HIP(hipMalloc(&g->dinv32, b4));
    }
    }
    const size_t pn = (size_t)MAX_PARTIAL_BLOCKS * 8;
    if (!g->partials) CHECK_HIP(hipMalloc(&g->partials, pn * sizeof(double)));
    if (!g->S) CHECK_HIP(hipMalloc(&g->S, S_COUNT * sizeof(double)));
    if (!g->Sh) CHECK_HIP(hipHostMalloc((void **)&g->Sh, S_COUNT * sizeof(double), hipHostMallocDefault));
    CHECK_HIP(hipMemsetAsync(g->S, 0, S_COUNT * sizeof(double), g->stream));
    g->arrays = true;
    g->dinv_valid = false;
    return CUDAPCG_TRUE;
}

// Build the operator descriptors from the model and the local matrices.
static int make_ops()
{
    cudapcgModel_t *m = g->model;
    OpDesc o;
    memset(&o, 0, sizeof(o));
    o.physics = phys(g->analysis_flag);
    o.ebe = g->parStrategy_flag == CUDAPCG_EBE;
    o.ny = (int)m->nrows;
    o.nx = (int)m->ncols;
    o.nz = (int)m->nlayers;
    o.N = m->nelem;
    o.ndof = m->nvars;
    o.keys = m->image;
    o.nvel = m->nporenodes;
    o.nborder = m->nbordernodes;
    o.dof2node = m->DOF2periodic_map;
    o.node2dof = m->periodic2DOF_map;
    o.bkeys = m->border_pore_map;
    o.fkeys = nullptr;
    o.lamu = nullptr;
    o.tile_tr = 32;
    o.tile_tc = 8;
    o.tile_lz = 16;
    const char *tile_env = getenv("CHFEM_WALSH_TILE");
    const bool fixed_tile = tile_env && strcmp(tile_env, "auto") != 0;
    if (fixed_tile) sscanf(tile_env, "%dx%dx%d", &o.tile_tr, &o.tile_tc, &o.tile_lz);
    const char *opsel = getenv("CHFEM_OPERATOR");
    bool use_walsh = g->walsh_ok && !(opsel && !strcmp(opsel, "generic"));
    // Stokes: Walsh-basis operator over tiles, with per-node fluid keys built
    // from the DOF maps (the generic node kernel for CHFEM_OPERATOR=generic,
    // and for pore spaces without a fully fluid node or under 3 DOFs)
    if (o.physics == PHY_FLUID && !(opsel && !strcmp(opsel, "generic")) && o.nvel > 0 && o.nvel + o.nborder >= 3 &&
        o.node2dof && o.dof2node && (o.bkeys || o.nborder == 0)) {
        if (g->fkeys_n != (size_t)o.N) {
            dfree(g->fkeys);
            if (hipMalloc(&g->fkeys, o.N) != hipSuccess) {
                g->fkeys = nullptr;
                g->fkeys_n = 0;
                printf("ERROR: could not allocate the fluid node keys.\n");
                return 0;
            }
            g->fkeys_n = o.N;
        }
        launch_fluid_node_keys(o, g->fkeys, g->stream);
        o.fkeys = g->fkeys;
    }
    // key-Jacobi vector kernels: NBN thermal/elastic with the Jacobi
    // preconditioner on (CHFEM_VECTOR=stored keeps the stored-dinv kernels)
    // scalar density field: element-by-element operator only (the reference
    // forces EBE for it, femhmg.c:196-203)
    o.sdf.d = nullptr;
    o.sdf.fmin = o.sdf.fmax = 0.0;
    if (m->parametric_density_field_flag && m->parametric_density_field && o.physics != PHY_FLUID) {
        if (!o.ebe) {
            printf("ERROR: the scalar density field needs the element-by-element strategy.\n");
            return 0;
        }
        o.sdf.d = m->parametric_density_field;
        o.sdf.fmin = m->limits_density_field[0];
        o.sdf.fmax = m->limits_density_field[1];
        use_walsh = false;  // the Walsh operator has no per-element scale
    }
    const char *vsel = getenv("CHFEM_VECTOR");
    const bool kj = g->kinv64 && g->precond && !o.ebe && o.physics != PHY_FLUID && !(vsel && !strcmp(vsel, "stored"));
    o.kinv = nullptr;
    o.nmat = (int)g->dm.nmat;
    const char *esel = getenv("CHFEM_EBE");
    o.ebe_atomic = (esel && !strcmp(esel, "atomic")) || o.physics == PHY_FLUID;
    o.ebe_tiles16 = esel && !strcmp(esel, "mfma16");
    g->dm.op64 = o;
    g->dm.op64.K = g->dm.K64;
    g->dm.op64.lamu = use_walsh ? g->lamu64 : nullptr;
    g->dm.op64.kinv = kj ? g->kinv64 : nullptr;
    g->dm.op32 = o;
    g->dm.op32.K = g->dm.K32;
    g->dm.op32.lamu = use_walsh ? g->lamu32 : nullptr;
    g->dm.op32.kinv = kj ? g->kinv32 : nullptr;
    if (!fixed_tile && g->tuned) {
        // tiles picked by the autotuner for this model
        g->dm.op64.tile_tr = g->tile64[0]; g->dm.op64.tile_tc = g->tile64[1]; g->dm.op64.tile_lz = g->tile64[2];
        g->dm.op32.tile_tr = g->tile32[0]; g->dm.op32.tile_tc = g->tile32[1]; g->dm.op32.tile_lz = g->tile32[2];
    }
    g->autotune = !fixed_tile && use_walsh && !o.ebe;
    g->ops_ready = true;
    return 1;
}

// Operator tile autotuner: times every Walsh tile shape (rows x columns per
// block, node layers per block) on the model's own grid in each precision the
// solve uses and keeps the fastest.  Runs once per model, before its first
// solve; CHFEM_WALSH_TILE=TRxTCxLZ fixes the tile instead.
static bool is_minres();
template <typename T> static Ring<T> dir_ring(T *d, T *s, unsigned cap = RING_MAX);
static void autotune_tiles()
{
    if (!g->autotune || g->tuned) return;
    // (tile rows 65: 64-row tiles with the row above on an extra wavefront)
    static const int cand[][3] = {{32, 8, 16}, {32, 8, 32}, {16, 16, 16}, {32, 16, 16}, {32, 16, 32}, {64, 4, 16},
                                  {64, 8, 16}, {65, 7, 16}, {65, 7, 32}, {65, 15, 16}};
    const int nc = (int)(sizeof(cand) / sizeof(cand[0]));
    const size_t n = nvars();
    // the CG loops' kernels per iteration: form 1 = operator with the
    // direction update + z update kernel, form 2 = single-kernel iteration,
    // form 0 (MINRES, CHFEM_FUSE_DIR=0) = the plain apply
    const int env = fuse_mode_env();
    const int m0 = is_minres() ? 0 : (env >= 0 ? env : 1), m1 = is_minres() ? 0 : (env >= 0 ? env : 2);
    Reduce rt = make_reduce(FIN_STORE, S_TMP0);
    g->Sh[S_ALPHA] = 0.0;  // the deferred z updates of the timed launches change nothing
    g->Sh[S_BPREV] = 0.0;
    if (hipMemcpyAsync(g->S + S_ALPHA, g->Sh + S_ALPHA, sizeof(double), hipMemcpyHostToDevice, g->stream) != hipSuccess ||
        hipMemcpyAsync(g->S + S_BPREV, g->Sh + S_BPREV, sizeof(double), hipMemcpyHostToDevice, g->stream) != hipSuccess)
        return;
    auto tune = [&](OpDesc &op, auto *d, auto *q, auto *d2, int *best, int &best_mode) {
        using T = typename std::remove_pointer<decltype(d)>::type;
        // z2, q2: scratch
        const bool dbl = std::is_same<T, double>::value;
        T *z2 = dbl ? (T *)g->z2_64 : (T *)g->z2_32;
        T *q2 = dbl ? (T *)g->q2_64 : (T *)g->q2_32;
        const bool fz = d2 && z2 && q2 && apply_dir_ok(op);
        // the ring width decides the single-kernel form the loop runs: with
        // 3 or more slots z is recovered from the two previous directions,
        // with 2 it is stored and reloaded
        const bool zrec = fz && dir_ring<T>(d, d2).w >= 3;
        launch_fill<T>(z2, 1.0, n, g->stream);
        launch_fill<T>(q2, 0.5, n, g->stream);
        auto run = [&](int mode) {
            if (fz && mode == 2) {
                // the loop's steady iteration, with the deferred updates
                // pending (alpha = 0, S[BPREV] = 0): z recovered from the
                // directions d and z2 (or, with a 2-slot ring, z2 read and
                // rewritten), q_old streamed, d2 and q written
                if (zrec) launch_apply_cg1<T>(op, z2, q2, d, nullptr, d2, q, 1, &rt, g->stream, nullptr, (const T *)z2);
                // (z2 is read and written in place: with alpha = 0 the value
                // written is the value read, bit for bit)
                else launch_apply_cg1<T>(op, z2, q2, d, z2, d2, q, 1, &rt, g->stream, nullptr, (const T *)nullptr);
            } else if (fz && mode == 1) {
                launch_apply_dir<T>(op, z2, d, d2, q, &rt, g->stream);
                launch_kj_zupdate<T>(op, z2, q, &rt, g->stream);
            } else {
                launch_apply<T>(op, d, q, nullptr, g->stream);
            }
        };
        // two stages: every (form, tile) timed over 2 launches after one
        // untimed, then the 3 fastest over 8 launches (about half the
        // launches of timing all of them over 8)
        auto timed = [&](int mode, int c, int reps, float &ms) -> bool {
            op.tile_tr = cand[c][0]; op.tile_tc = cand[c][1]; op.tile_lz = cand[c][2];
            hipEventRecord(g->ev0, g->stream);
            for (int k = 0; k < reps; k++) run(mode);
            hipEventRecord(g->ev1, g->stream);
            if (hipEventSynchronize(g->ev1) != hipSuccess) return false;
            ms = 0;
            hipEventElapsedTime(&ms, g->ev0, g->ev1);
            return true;
        };
        struct Cand { float ms; int mode, c; };
        Cand all[2 * 16];
        int na = 0;
        for (int mode = m0; mode <= m1; mode++) {
            if (mode > 0 && !fz) continue;
            for (int c = 0; c < nc && na < 32; c++) {
                op.tile_tr = cand[c][0]; op.tile_tc = cand[c][1]; op.tile_lz = cand[c][2];
                run(mode);
                float ms;
                if (!timed(mode, c, 2, ms)) return;
                all[na++] = {ms, mode, c};
            }
        }
        if (na == 0) return;
        std::sort(all, all + na, [](const Cand &a, const Cand &b) { return a.ms < b.ms; });
        float best_ms = 1e30f;
        for (int i = 0; i < std::min(na, 3); i++) {
            float ms;
            if (!timed(all[i].mode, all[i].c, 8, ms)) return;
            if (ms < best_ms) {
                best_ms = ms;
                best_mode = all[i].mode > 0 ? all[i].mode : 1;
                best[0] = cand[all[i].c][0]; best[1] = cand[all[i].c][1]; best[2] = cand[all[i].c][2];
            }
        }
        op.tile_tr = best[0]; op.tile_tc = best[1]; op.tile_lz = best[2];
    };
    launch_fill<double>(g->d64, 1.0, n, g->stream);
    tune(g->dm.op64, g->d64, g->q64, g->s64, g->tile64, g->mode64);
    if (g->precision == CHFEM_PREC_MIXED && g->d32) {
        launch_fill<float>(g->d32, 1.0, n, g->stream);
        tune(g->dm.op32, g->d32, g->q32, g->s32, g->tile32, g->mode32);
    } else {
        memcpy(g->tile32, g->tile64, sizeof(g->tile32));
        g->mode32 = g->mode64;
    }
    g->tuned = true;
}

template <typename T> static int loop_mode()
{
    const int env = fuse_mode_env();
    if (env >= 0) return env;
    return std::is_same<T, double>::value ? g->mode64 : g->mode32;
}

static bool above(double delta, double delta0, double tol) { return fabs(delta) > tol * tol * fabs(delta0); }

// Records the relative residual of an initial guess; the guess is kept unless
// its residual is larger than the zero guess's (CHFEM_X0_FALLBACK=0 keeps it
// always, as the reference does).
static void note_x0(double delta, double delta0)
{
    static int fallback = -1;
    if (fallback < 0) {
        const char *e = getenv("CHFEM_X0_FALLBACK");
        fallback = e ? atoi(e) != 0 : 1;
    }
    g->x0_residual = sqrt(fabs(delta) / fabs(delta0));
    g->x0_used = !(fallback && fabs(delta) > fabs(delta0));
}

// Direction ring of the fused CG loops: d, s and up to ring_width() - 2 more
// buffers, allocated on first use.  The ring's flush reads every direction
// once and the iterate once per window, so a wider ring moves fewer bytes
// per iteration (32 slots: 1.0625 vector passes per iteration against 1.125
// for 16, 1 % off the 256^3 elastic solve).  By default the extra slots stay
// within a quarter of the free device memory; CHFEM_RING sets the width.
// Slots that do not fit are left out (the iterate is then flushed every two
// iterations at the least).
static unsigned ring_width()
{
    const char *e = getenv("CHFEM_RING");
    const int v = e ? atoi(e) : (int)RING_MAX;
    return (unsigned)std::min(std::max(v, 2), (int)RING_MAX);
}

// The ring's extra buffers (slots 2 ..) are allocated on first use, up to
// cap slots; a later call with a larger cap (a CG solve after a MINRES one)
// grows them.
template <typename T>
static Ring<T> dir_ring(T *d, T *s, unsigned cap)
{
    const bool dbl = std::is_same<T, double>::value;
    int &w = dbl ? g->ring_w64 : g->ring_w32;
    unsigned &tried = dbl ? g->ring_try64 : g->ring_try32;
    T **extra = dbl ? (T **)g->ring64 : (T **)g->ring32;
    unsigned want = std::min(ring_width(), std::max(cap, 2u));
    if (want > tried) {
        tried = want;
        const size_t vb = nvars() * sizeof(T);
        size_t freeb = 0, totb = 0;
        if (!getenv("CHFEM_RING") && vb && hipMemGetInfo(&freeb, &totb) == hipSuccess)
            want = (unsigned)std::min<size_t>(want, std::max(w, 2) + freeb / 4 / vb);
        if (w < 2) w = 2;
        for (unsigned k = (unsigned)w; k < want; k++) {
            if (hipMalloc(&extra[k], vb) != hipSuccess) {
                (void)hipGetLastError();
                break;
            }
            w = (int)k + 1;
        }
    }
    Ring<T> r{};
    r.w = w;
    r.p[0] = d;
    r.p[1] = s;
    for (int k = 2; k < w; k++) r.p[k] = extra[k];
    return r;
}

// Jacobi-PCG in precision T (pcg.cu:28-243).  r holds b on entry; returns 1 if converged.
// With op.kinv (thermal/elastic NBN) the Jacobi diagonal comes from the keys
// and z = M^-1 r is folded into the direction update (d = M^-1 r + beta d);
// otherwise z is stored in q and the diagonal read from dinv.
// With d2 and an operator that supports it (apply_dir_ok), the loop runs on
// the preconditioned residual z = M^-1 r (held in r's buffer from the loop's
// start) and the direction update runs inside the operator: iteration j
// writes its direction to ring slot j % w from z and slot (j - 1) % w; the
// update kernel only updates z.  x += alpha_j d_j is applied once per w
// iterations for the whole window (launch_ring_flush) and after the loop.
template <typename T>
static int run_pcg(const OpDesc &op, T *x, T *r, T *d, T *q, const T *dinv, bool x0set, double tol, unsigned maxit,
                   unsigned &it, double &residual, double *delta0_out, T *d2)
{
    const size_t n = op.ndof;
    hipStream_t st = g->stream;
    const bool kj = op.kinv != nullptr;
    const bool fused = kj && d2 && fuse_dir_enabled() && apply_dir_ok(op);
    const Ring<T> ring = fused ? dir_ring<T>(d, d2) : Ring<T>{};
    const unsigned W = fused ? (unsigned)ring.w : 1u;
    auto slot = [&](unsigned j) { return (T *)ring.p[j % W]; };
    // single-kernel iterations: iteration j reads z, q from buffer (j + 1) % 2
    // and writes buffer j % 2; the loop's first z is r's buffer
    T *z2 = std::is_same<T, double>::value ? (T *)g->z2_64 : (T *)g->z2_32;
    T *q2 = std::is_same<T, double>::value ? (T *)g->q2_64 : (T *)g->q2_32;
    const bool cg1 = fused && loop_mode<T>() == 2 && z2 && q2;
    T *zb[2] = {z2, r}, *qb[2] = {q, q2};
    if (!x0set) launch_fill<T>(x, 0.0, n, st);
    Reduce r0 = make_reduce(FIN_STORE, S_DELTA0);
    if (kj) launch_kj_precond<T>(op, r, d, &r0, st);
    else launch_precond<T>(dinv, r, d, n, &r0, st);
    if (!sync_scalars()) return -1;
    const double delta0 = g->Sh[S_DELTA0];
    if (delta0_out) *delta0_out = delta0;
    it = 0;
    residual = 0;
    if (fabs(delta0) < 1e-12) {
        launch_fill<T>(x, 0.0, n, st);
        return 1;
    }
    double delta = delta0;
    if (x0set) {
        launch_apply<T>(op, x, q, nullptr, st);
        Reduce rd = make_reduce(FIN_STORE, S_DELTA);
        if (kj) launch_kj_sub_precond<T>(op, r, q, d, &rd, st);
        else launch_sub_precond<T>(r, q, dinv, d, &rd, n, st);
        if (!sync_scalars()) return -1;
        delta = g->Sh[S_DELTA];
        note_x0(delta, delta0);
        if (!g->x0_used) {
            // the guess is worse than zero: back to x = 0, r = b, d = M^-1 b
            launch_add<T>(r, q, n, st);
            launch_fill<T>(x, 0.0, n, st);
            Reduce r0b = make_reduce(FIN_STORE, S_DELTA);
            if (kj) launch_kj_precond<T>(op, r, d, &r0b, st);
            else launch_precond<T>(dinv, r, d, n, &r0b, st);
            if (!sync_scalars()) return -1;
            delta = g->Sh[S_DELTA];
        } else if (!above(delta, delta0, tol)) {
            residual = sqrt(fabs(delta) / fabs(delta0));
            return 1;
        }
    } else if (!set_scalar(S_DELTA, delta0)) return -1;
    // Iterations are queued cg_batch() at a time: the finishing kernel of
    // each update counts the iteration and raises S_DONE once delta <=
    // tol^2 delta0, after which every queued launch returns at once, so the
    // iterate and the iteration count are those of a loop that synchronises
    // every iteration.
    Reduce ra = make_guarded(FIN_ALPHA, 0), rdl = make_guarded(FIN_DELTA, 0), rc = make_guarded(FIN_CG1, 0);
    if (fused) {
        Reduce rz = make_reduce(FIN_STORE, S_TMP0);
        launch_kj_to_z<T>(op, r, &rz, st);
    }
    if (!arm_guard(tol * tol * fabs(delta0))) return -1;
    it = 0;
    bool done = false;
    unsigned queued = 0, flushed = 0;  // fused: iterations queued, and folded into x
    while (!done && it < maxit) {
        const unsigned nb = std::min(cg_batch(), maxit - it);
        for (unsigned b = 0; b < nb; b++) {
            const unsigned j = it + b;
            if (fused) {
                if (cg1) {
                    // with 3 or more ring slots, iterations from the third on
                    // recover z from the two previous directions, and only
                    // the first writes z (the second reads it)
                    const bool zr = W >= 3 && j >= 2;
                    T *zw = W >= 3 && j >= 1 ? nullptr : zb[j & 1];
                    // (the timed launch: the third of the batch, a steady
                    // iteration)
                    timed_apply_cg1<T>(op, zb[(j + 1) & 1], qb[(j + 1) & 1], j > 0 ? slot(j - 1) : nullptr, zw,
                                       slot(j), qb[j & 1], j > 0 ? 1 : 0, &rc, b == (nb > 2 ? 2u : 0u),
                                       zr ? slot(j - 2) : nullptr);
                } else {
                    const T *dold = j > 0 ? slot(j - 1) : nullptr;
                    if (b == 0) timed_apply_dir<T>(op, r, dold, slot(j), q, &ra);
                    else launch_apply_dir<T>(op, r, dold, slot(j), q, &ra, st);
                    launch_kj_zupdate<T>(op, r, q, &rdl, st);
                }
                queued = j + 1;
                if (queued % W == 0) {
                    launch_ring_flush<T>(x, ring, flushed, queued, 0, g->S, n, st);
                    flushed = queued;
                }
                continue;
            }
            if (j > 0) {
                if (kj) launch_kj_pcg_dir<T>(op, d, r, g->S, st, 1);
                else launch_pcg_dir<T>(d, q, g->S, 0, n, st, 1);
            }
            if (b == 0) timed_apply<T>(op, d, q, &ra);
            else launch_apply<T>(op, d, q, &ra, st);
            if (kj) launch_kj_pcg_update<T>(op, x, r, d, q, &rdl, st);
            else launch_pcg_update<T>(x, r, d, q, dinv, &rdl, n, st);
        }
        if (!sync_loop()) return -1;
        it = (unsigned)g->Sh[S_ITER];
        done = g->Sh[S_DONE] != 0.0;
    }
    // the iterations since the last window flush
    if (fused) launch_ring_flush<T>(x, ring, flushed, queued, 0, g->S, n, st);
    delta = g->Sh[S_DELTA];
    residual = sqrt(fabs(delta) / fabs(delta0));
    return !above(delta, delta0, tol);
}

// Jacobi-PCG (FP64) with the INF or error stopping rule (pcg.cu:28-243, the
// switch at :130-146 and :176-190): separate direction, operator and update
// kernels with the stored diagonal, the rule's metric evaluated after every
// iteration.  INF: max|r| / max|b| <= tol.  Error: |alpha| max|d| / max|x|
// over the first nhmgvars entries <= tol, or the L2 rule met.
static int absmax_of(const double *v, size_t n, const Jac<double> *scaled, int slot);
static size_t hmg_vars();
static int run_pcg_rule(bool x0set, unsigned &it, double &residual)
{
    const OpDesc &op = g->dm.op64;
    const size_t n = op.ndof, nh = hmg_vars();
    hipStream_t st = g->stream;
    const double tol = g->num_tol;
    const bool inf = g->resnorm == CUDAPCG_INF_NORM;
    double *x = g->x64, *r = g->r64, *d = g->d64, *q = g->q64;
    const double *dinv = g->dinv64;
    double res0_inf = 0;
    if (inf) {
        if (!absmax_of(r, n, nullptr, S_TMP2)) return -1;
        res0_inf = g->Sh[S_TMP2];
    }
    Reduce r0 = make_reduce(FIN_STORE, S_DELTA0);
    launch_precond<double>(dinv, r, d, n, &r0, st);
    if (!sync_scalars()) return -1;
    const double delta0 = g->Sh[S_DELTA0];
    const double res0 = inf ? res0_inf : delta0;
    it = 0;
    residual = 0;
    if (fabs(res0) < 1e-12) {
        launch_fill<double>(x, 0.0, n, st);
        return 1;
    }
    double delta = delta0;
    if (x0set) {
        launch_apply<double>(op, x, q, nullptr, st);
        Reduce rd = make_reduce(FIN_STORE, S_DELTA);
        launch_sub_precond<double>(r, q, dinv, d, &rd, n, st);
        if (!sync_scalars()) return -1;
        delta = g->Sh[S_DELTA];
        note_x0(delta, delta0);
        if (!above(delta, delta0, tol)) {
            residual = sqrt(fabs(delta) / fabs(delta0));
            return 1;
        }
    } else {
        launch_fill<double>(x, 0.0, n, st);
    }
    if (!set_scalar(S_DELTA, delta) || !arm_guard(-1.0)) return -1;
    Reduce ra = make_reduce(FIN_ALPHA, 0), rdl = make_reduce(FIN_DELTA, 0);
    double metric = 1.0;
    bool done = false;
    while (!done && it < g->max_iterations) {
        if (it > 0) launch_pcg_dir<double>(d, q, g->S, 0, n, st);
        timed_apply<double>(op, d, q, &ra);
        launch_pcg_update<double>(x, r, d, q, dinv, &rdl, n, st);
        if (!sync_loop()) return -1;
        it++;
        if (inf) {
            if (!absmax_of(r, n, nullptr, S_TMP2)) return -1;
            metric = g->Sh[S_TMP2] / res0;
            done = !(metric > tol);
        } else {
            if (!absmax_of(d, nh, nullptr, S_TMP3) || !absmax_of(x, nh, nullptr, S_TMP4)) return -1;
            metric = fabs(g->Sh[S_ALPHA]) * g->Sh[S_TMP3] / g->Sh[S_TMP4];
            done = metric <= tol || !above(g->Sh[S_DELTA], delta0, tol);
        }
    }
    residual = metric;
    return metric <= tol;
}

// Jacobi-MINRES in precision T (pminres.cu:28-247).  r holds b on entry.
// stagnation > 0 (the FP32 inner solves of mixed precision): the loop also
// ends once the residual measure has not fallen below 0.9 of its best value
// for that many iterations plus a tenth of those done, i.e. when the solve
// has reached what its precision can attain; the FP64 refinement then
// restarts it from the FP64 residual.
template <typename T>
static int run_minres(const OpDesc &op, T *x, T *r, T *d, T *q, T *s, const T *dinv, bool x0set, double tol,
                      unsigned maxit, unsigned &it, double &residual, unsigned stagnation = 0)
{
    const size_t n = op.ndof;
    hipStream_t st = g->stream;
    // INF and error rules (pminres.cu:188-200): FP64 solves only (the FP32
    // inner solves of mixed precision stop on the L2 measure)
    constexpr bool dbl = std::is_same<T, double>::value;
    const cudapcgFlag_t rule = dbl && !stagnation ? g->resnorm : (cudapcgFlag_t)CUDAPCG_L2_NORM;
    double res0_inf = 0;
    if constexpr (dbl) {
        if (rule == CUDAPCG_INF_NORM) {
            if (!absmax_of(r, n, nullptr, S_TMP2)) return -1;
            res0_inf = g->Sh[S_TMP2];
        }
    }
    if (!x0set) launch_fill<T>(x, 0.0, n, st);
    Reduce rr0 = make_reduce(FIN_STORE, S_RES0), rd0 = make_reduce(FIN_STORE, S_DELTA0);
    launch_precond<T>(dinv, r, q, n, &rr0, st);        // q = M^-1 r ; res0 = r.q
    launch_apply<T>(op, q, s, &rd0, st);               // s = A q ; delta0 = q.s
    if (!sync_scalars()) return -1;
    const double res0 = rule == CUDAPCG_INF_NORM ? res0_inf : g->Sh[S_RES0], delta0 = g->Sh[S_DELTA0];
    // the rule's metric after an iteration (r holds M^-1 r by then)
    auto rule_metric = [&](double &m) -> bool {
        if constexpr (dbl) {
            if (rule == CUDAPCG_INF_NORM) {
                Jac<double> jd;
                memset(&jd, 0, sizeof(jd));
                jd.kind = JAC_STORED;
                jd.dinv = (const double *)dinv;
                if (!absmax_of((const double *)r, n, &jd, S_TMP2)) return false;
                m = g->Sh[S_TMP2] / res0;
            } else {
                const size_t nh = hmg_vars();
                if (!absmax_of((const double *)d, nh, nullptr, S_TMP3) ||
                    !absmax_of((const double *)x, nh, nullptr, S_TMP4))
                    return false;
                m = fabs(g->Sh[S_ALPHA]) * g->Sh[S_TMP3] / g->Sh[S_TMP4];
            }
        }
        return true;
    };
    it = 0;
    residual = 0;
    if (fabs(res0) < 1e-12) {
        launch_fill<T>(x, 0.0, n, st);
        return 1;
    }
    double delta = delta0;
    if (x0set) {
        Reduce rt0 = make_reduce(FIN_STORE, S_TMP0), rdl = make_reduce(FIN_STORE, S_DELTA);
        launch_apply<T>(op, x, d, nullptr, st);
        launch_sub_precond<T>(r, d, dinv, q, &rt0, n, st);
        launch_apply<T>(op, q, s, &rdl, st);
        if (!sync_scalars()) return -1;
        delta = g->Sh[S_DELTA];
        if (!above(delta, delta0, tol)) {
            residual = sqrt(fabs(delta) / fabs(delta0));
            return 1;
        }
    } else if (!set_scalar(S_DELTA, delta0)) return -1;
    // first iteration: d = M^-1 r ; q = A M^-1 r ; r <- M^-1 r kept from here on
    hipMemcpyAsync(d, q, n * sizeof(T), hipMemcpyDeviceToDevice, st);
    hipMemcpyAsync(q, s, n * sizeof(T), hipMemcpyDeviceToDevice, st);
    launch_precond<T>(dinv, r, r, n, nullptr, st);
    // L2 rule: iterations are queued cg_batch() at a time, as in the CG
    // loops: the update's finishing kernel counts the iteration and raises
    // S_DONE once stop <= tol^2 res0, after which the queued launches return
    // at once (the iterate and the count are those of a loop that
    // synchronises every iteration).  The INF and error rules evaluate their
    // metric on the host after every iteration.
    const bool queued = rule == CUDAPCG_L2_NORM;
    Reduce ra = make_reduce(FIN_ALPHA, 0), rs = make_reduce(FIN_STORE, S_STOP), rdl = make_reduce(FIN_DELTA, 0);
    if (queued) {
        ra = make_guarded(FIN_ALPHA, 0);
        rs = make_guarded(FIN_STOP, 0);
        rdl = make_guarded(FIN_DELTA_NOCOUNT, 0);
        if (!arm_guard(tol * tol * fabs(res0))) return -1;
    }
    // Queued loops keep their directions in a ring (d and up to 7 of the
    // CG loops' extra direction buffers): iteration j writes d_j to slot
    // j % W and reads d_{j-1}; x += alpha_j d_j is applied once per window
    // (the guarded alpha finish records alpha_j at S_AHIST + j % 32), so the
    // update kernel neither reads d nor reads and writes x: about 12.25
    // instead of 14 vector passes per iteration.  A wider ring would save
    // under 1 % more for a vector of device memory per slot.
    constexpr int MINRES_RING = 8;
    const Ring<T> cg_ring = queued ? dir_ring<T>(d, d, MINRES_RING + 1) : Ring<T>{};
    Ring<T> ring{};
    ring.p[0] = d;
    ring.w = 1;
    for (int k = 2; k < cg_ring.w && ring.w < MINRES_RING; k++) ring.p[ring.w++] = cg_ring.p[k];
    const bool rng = queued && ring.w >= 2;
    const unsigned W = rng ? (unsigned)ring.w : 1u;
    auto slot = [&](unsigned j) { return (T *)ring.p[j % W]; };
    unsigned issued = 1, flushed = 0;  // iterations queued, and folded into x
    launch_precond<T>(dinv, q, nullptr, n, &ra, st);
    it = 1;
    launch_minres_update<T>(rng ? nullptr : x, r, d, q, dinv, &rs, n, st);
    timed_apply<T>(op, r, s, &rdl);
    if (!sync_loop()) return -1;
    double stop = g->Sh[S_STOP], metric = 1.0;
    bool cont = above(stop, res0, tol);
    if (rule != CUDAPCG_L2_NORM) {
        if (!rule_metric(metric)) return -1;
        cont = metric > tol;
    }
    double best = fabs(stop);
    unsigned best_it = it;
    while (cont && it < maxit) {
        const unsigned nb = queued ? std::min(cg_batch(), maxit - it) : 1u;
        for (unsigned b = 0; b < nb; b++) {
            if (rng) {
                const unsigned j = issued;
                launch_minres_dir<T>(slot(j - 1), slot(j), q, r, s, dinv, g->S, 0, &ra, n, st);
                launch_minres_update<T>(nullptr, r, slot(j), q, dinv, &rs, n, st);
                timed_apply<T>(op, r, s, &rdl);
                issued = j + 1;
                if (issued - flushed == W) {
                    launch_ring_flush<T>(x, ring, flushed, issued, 0, g->S, n, st);
                    flushed = issued;
                }
            } else {
                launch_minres_dir<T>(d, d, q, r, s, dinv, g->S, 0, &ra, n, st);
                launch_minres_update<T>(x, r, d, q, dinv, &rs, n, st);
                timed_apply<T>(op, r, s, &rdl);
            }
        }
        if (!sync_loop()) return -1;
        it = queued ? (unsigned)g->Sh[S_ITER] : it + 1;
        stop = g->Sh[S_STOP];
        cont = queued ? g->Sh[S_DONE] == 0.0 : above(stop, res0, tol);
        if (rule != CUDAPCG_L2_NORM) {
            if (!rule_metric(metric)) return -1;
            cont = metric > tol;
        }
        if (stagnation) {
            if (fabs(stop) < 0.9 * best) {
                best = fabs(stop);
                best_it = it;
            } else if (it - best_it > stagnation + it / 10) {
                break;
            }
        }
    }
    // the iterations since the last window flush (those past the stop, which
    // the guard skipped, add nothing)
    if (rng && issued > flushed) launch_ring_flush<T>(x, ring, flushed, issued, 0, g->S, n, st);
    residual = rule != CUDAPCG_L2_NORM ? metric : sqrt(fabs(stop) / fabs(res0));
    return residual <= tol;
}

static bool is_minres()
{
    return g->solver_flag == CUDAPCG_MINRES_SOLVER || g->solver_flag == CUDAPCG_MINRES3_SOLVER ||
           g->solver_flag == CUDAPCG_MINRES2_SOLVER;
}

// ---------------------------------------------------------------------------
// Memory-lean solvers: CG3 / MINRES3 keep x, z and d (pcg3.cu, pminres3.cu),
// CG2 / MINRES2 keep z and d and reduce the solution on the fly (pcg2.cu,
// pminres2.cu, xreduce.cu).  The same iterations as CG and MINRES, with A d
// applied again instead of stored: CG3 and CG2 apply the operator twice per
// iteration (d.Ad, then z -= alpha M^-1 A d), MINRES3 and MINRES2 three
// times (Ad.M^-1 Ad, the z update, z.Az).  They iterate on z = M^-1 r, so
// the Jacobi diagonal meets only operator results; the Walsh-basis NBN and
// Stokes operators fold each form into their epilogue, other operators go
// through one work vector.  FP64 in every precision mode.
// ---------------------------------------------------------------------------
static bool is_lean()
{
    return g->solver_flag >= CUDAPCG_CG3_SOLVER && g->solver_flag <= CUDAPCG_MINRES2_SOLVER;
}
static bool is_two_vector() { return g->solver_flag >= CUDAPCG_CG2_SOLVER && g->solver_flag <= CUDAPCG_MINRES2_SOLVER; }

// entries the error rule and the on-the-fly reduction cover (nhmgvars: the
// velocity DOFs for Stokes), and values per node of the reduction
static size_t hmg_vars()
{
    const size_t n = nvars(), h = g->model->nhmgvars;
    if (h && h <= n) return h;
    // unset: the velocity DOFs for Stokes, every DOF otherwise (femhmg.c:82)
    return g->analysis_flag == CUDAPCG_FLUID_3D ? 3 * (size_t)g->model->nporenodes : n;
}
static int vars_per_node()
{
    const unsigned p = g->model->nvarspernode;
    return p < 1 ? 1 : (p > 4 ? 4 : (int)p);
}

static Jac<double> make_jac()
{
    Jac<double> j;
    memset(&j, 0, sizeof(j));
    const OpDesc &op = g->dm.op64;
    j.nd = op.physics == PHY_ELASTIC ? 3 : 1;
    j.vinv = j.pinv = 1.0;
    if (!g->precond) {
        j.kind = JAC_NONE;
    } else if (op.physics == PHY_FLUID) {
        j.kind = JAC_FLUID;
        j.vinv = g->fl_vinv;
        j.pinv = g->fl_pinv;
        j.btab = g->btab64;
        j.bkeys = op.bkeys;
        j.nvel = op.nvel;
    } else if (g->kinv64 && !op.ebe) {
        j.kind = JAC_KEYS;
        j.kinv = g->kinv64;
        j.keys = op.keys;
    } else {
        j.kind = JAC_STORED;
    }
    return j;
}

static int ensure_dinv();

// absolute maxima of the INF and error rules into S[slot]
static int absmax_of(const double *v, size_t n, const Jac<double> *scaled, int slot)
{
    launch_absmax<double>(v, n, scaled, g->partials, g->S, slot, g->stream);
    return sync_scalars();
}

// The lean solve on the RHS in r64 (returns 1 converged, 0 not, -1 error).
static int run_lean(bool x0set, unsigned &it, double &residual)
{
    const OpDesc &op = g->dm.op64;
    const size_t n = op.ndof, nh = hmg_vars();
    const int p = vars_per_node();
    hipStream_t st = g->stream;
    const bool minres = is_minres(), two = is_two_vector();
    const cudapcgFlag_t rule = g->resnorm;
    const bool l2 = rule == CUDAPCG_L2_NORM;
    const double tol = g->num_tol;
    const unsigned maxit = g->max_iterations;
    Jac<double> jac = make_jac();
    if (jac.kind == JAC_STORED) {
        if (!g->dinv64) {
            if (hipMalloc(&g->dinv64, n * sizeof(double)) != hipSuccess) {
                (void)hipGetLastError();
                fprintf(stderr, "chfem: could not allocate the Jacobi diagonal\n");
                return -1;
            }
            g->dinv_valid = false;
        }
        ensure_dinv();
        jac.dinv = g->dinv64;
    }
    double *qtmp = nullptr;
    if (lean_needs_qtmp(op)) {
        if (!g->q64 && hipMalloc(&g->q64, n * sizeof(double)) != hipSuccess) {
            (void)hipGetLastError();
            fprintf(stderr, "chfem: could not allocate the operator work vector\n");
            return -1;
        }
        qtmp = g->q64;
    }
    double *z = g->r64, *d = g->d64, *x = two ? nullptr : g->x64;
    const double *x0 = x0set ? g->x64 : nullptr;
    // two-vector solvers: reduced solution and its history, per solve
    const int xmode = g->xreduce_flag;
    const int shift = (int)std::min<unsigned>(g->xreduce_shift, (unsigned)p - 1);
    const size_t xn = two ? (xmode == CUDAPCG_XREDUCE_FULL ? nh : (xmode == CUDAPCG_XREDUCE_ONLYDIR ? nh / p : 0)) : 0;
    const size_t hlen = (size_t)maxit + 1;
    if (two) {
        if (g->xr_n != xn) {
            dfree(g->xr64);
            g->xr_n = 0;
            if (xn && hipMalloc(&g->xr64, xn * sizeof(double)) != hipSuccess) return -1;
            g->xr_n = xn;
        }
        if (g->hist_n != hlen * p) {
            dfree(g->hist);
            g->hist_n = 0;
            if (hipMalloc(&g->hist, hlen * p * sizeof(double)) != hipSuccess) return -1;
            g->hist_n = hlen * p;
        }
        if (xn) launch_fill<double>(g->xr64, 0.0, xn, st);
        launch_fill<double>(g->hist, 0.0, hlen * p, st);
    }
    double *xr = g->xr64;
    // INF rule: absolute maximum of the initial residual (before any guess)
    double res0_inf = 0;
    if (rule == CUDAPCG_INF_NORM) {
        if (!absmax_of(z, n, nullptr, S_TMP2)) return -1;
        res0_inf = g->Sh[S_TMP2];
    }
    // z = M^-1 b (r.M^-1 r of b to S_RES0); MINRES: delta0 = z.Az
    Reduce r0 = make_reduce(FIN_STORE, S_RES0);
    launch_lean_to_z<double>(jac, z, n, &r0, st);
    if (minres) {
        Reduce rd0 = make_reduce(FIN_STORE, S_DELTA0);
        launch_lean_apply<double>(op, jac, LEAN_VAV, z, z, qtmp, &rd0, st);
    }
    if (!sync_scalars()) return -1;
    const double rz0 = g->Sh[S_RES0];
    const double delta0 = minres ? g->Sh[S_DELTA0] : rz0;
    const double res0 = rule == CUDAPCG_INF_NORM ? res0_inf : rz0;
    it = 0;
    residual = 0;
    // x (CG3, MINRES3) starts at the guess x0 (the same buffer) or zero
    if (x && !x0) launch_fill<double>(x, 0.0, n, st);
    if (fabs(res0) < 1e-12) {
        // (MINRES zeroes a guess here, pminres3.cu; CG keeps it, pcg3.cu)
        if (x && minres) launch_fill<double>(x, 0.0, n, st);
        return 1;
    }
    double delta = delta0;
    if (x0) {
        // z = M^-1 (b - A x0); the reduced solution (CG2, MINRES2) starts
        // from x0
        if (!set_scalar(S_ALPHA, 1.0)) return -1;
        Reduce rx = make_reduce(FIN_STORE, S_TMP0);
        launch_lean_apply<double>(op, jac, LEAN_ZUPD, x0, z, qtmp, &rx, st);
        if (minres) {
            Reduce rdx = make_reduce(FIN_STORE, S_DELTA);
            launch_lean_apply<double>(op, jac, LEAN_VAV, z, z, qtmp, &rdx, st);
        }
        if (two) {
            Reduce rxr = make_reduce(FIN_STORE, 0);
            launch_xreduce<double>(xr, nullptr, x0, g->hist, hlen, xmode, shift, p, nh, rxr, st);
        }
        if (!sync_scalars()) return -1;
        delta = minres ? g->Sh[S_DELTA] : g->Sh[S_TMP0];
        if (!above(delta, delta0, tol)) {
            residual = sqrt(fabs(delta) / fabs(delta0));
            return 1;
        }
    }
    if (!set_scalar(S_DELTA, delta)) return -1;
    // L2: the stop is decided on the device (CG on r.M^-1 r, MINRES on its
    // residual measure) and iterations are queued in batches; the INF and
    // error rules are evaluated on the host after every iteration
    if (!arm_guard(l2 ? tol * tol * fabs(minres ? res0 : delta0) : -1.0)) return -1;
    Reduce ra = make_guarded(FIN_ALPHA, 0);
    Reduce rz = make_guarded(minres ? FIN_STOP : FIN_DELTA, 0);
    Reduce rv = make_guarded(FIN_DELTA_NOCOUNT, 0);
    Reduce rxr = make_guarded(FIN_STORE, 0);
    bool done = false;
    double metric = 1.0;
    while (!done && it < maxit) {
        const unsigned nb = l2 ? std::min(cg_batch(), maxit - it) : 1u;
        for (unsigned b = 0; b < nb; b++) {
            const unsigned j = it + b;
            // L2: x += alpha d of the previous iteration rides on the
            // direction update; the other rules need x current every step
            launch_lean_dir<double>(l2 ? x : nullptr, d, z, g->S, j == 0 ? 1 : 0, 1, n, st);
            launch_lean_apply<double>(op, jac, minres ? LEAN_QMQ : LEAN_VAV, d, z, qtmp, &ra, st);
            if (!l2 && x) launch_lean_axpy<double>(x, d, g->S, n, st);
            if (two) launch_xreduce<double>(xr, d, nullptr, g->hist, hlen, xmode, shift, p, nh, rxr, st);
            launch_lean_apply<double>(op, jac, LEAN_ZUPD, d, z, qtmp, &rz, st);
            if (minres) launch_lean_apply<double>(op, jac, LEAN_VAV, z, z, qtmp, &rv, st);
        }
        if (!sync_loop()) return -1;
        it = (unsigned)g->Sh[S_ITER];
        done = g->Sh[S_DONE] != 0.0;
        if (!l2) {
            const double a = g->Sh[S_ALPHA], dl = g->Sh[S_DELTA];
            if (rule == CUDAPCG_INF_NORM) {
                if (!absmax_of(z, n, jac.kind == JAC_NONE ? nullptr : &jac, S_TMP2)) return -1;
                metric = g->Sh[S_TMP2] / res0;
                done = !(metric > tol);
            } else {
                if (!absmax_of(d, nh, nullptr, S_TMP3)) return -1;
                double ax = 1.0;
                // denominators as the reference forms them: x for CG3 and
                // MINRES3, the reduced solution for CG2, 1 for MINRES2
                if (x) {
                    if (!absmax_of(x, nh, nullptr, S_TMP4)) return -1;
                    ax = g->Sh[S_TMP4];
                } else if (!minres && xn) {
                    if (!absmax_of(xr, xn, nullptr, S_TMP4)) return -1;
                    ax = g->Sh[S_TMP4];
                }
                metric = fabs(a) * g->Sh[S_TMP3] / ax;
                done = minres ? !(metric > tol) : (metric <= tol || !above(dl, delta0, tol));
            }
        }
    }
    if (l2 && x && it > 0) launch_lean_axpy<double>(x, d, g->S, n, st);
    if (l2) {
        if (minres) {
            residual = sqrt(fabs(g->Sh[S_STOP]) / fabs(res0));
            return residual <= tol;
        }
        residual = sqrt(fabs(g->Sh[S_DELTA]) / fabs(delta0));
        return !above(g->Sh[S_DELTA], delta0, tol);
    }
    residual = metric;
    return metric <= tol;
}

// The values a two-vector solve returns (xreduce.cu:4-26, pcg2.cu:246-262):
// the history scaled by the xreduce scale and written to
// <model name>_xreduce_metrics_<count>.bin, then per node value the total
// (NONE: the per-iteration terms summed from the last back to the initial
// one; ONLYDIR: the reduced direction only; FULL: the last entry).
static int finish_two_vector(unsigned it)
{
    const int p = vars_per_node();
    const size_t hl = (size_t)it + 1;
    std::vector<double> h(hl * p, 0.0);
    if (g->hist && hl * p <= g->hist_n &&
        hipMemcpy(h.data(), g->hist, hl * p * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
        return 0;
    const int mode = g->xreduce_flag;
    const int shift = (int)std::min<unsigned>(g->xreduce_shift, (unsigned)p - 1);
    const double scl = g->xreduce_scale;
    // per node value ii: res_per_it[ii][0 .. it]
    std::vector<std::vector<double>> per(p, std::vector<double>(hl));
    for (int ii = 0; ii < p; ii++)
        for (size_t jj = 0; jj < hl; jj++) per[ii][jj] = h[jj * p + ii];
    for (int ii = 0; ii < p; ii++) {
        if (mode == CUDAPCG_XREDUCE_ONLYDIR && ii != shift) continue;
        for (size_t jj = 0; jj < hl; jj++) per[ii][jj] *= scl;
    }
    if (g->model->name) {
        char fn[4096];
        snprintf(fn, sizeof(fn), "%s_xreduce_metrics_%lu.bin", g->model->name, g->count);
        FILE *f = fopen(fn, "wb");
        if (f) {
            for (int ii = 0; ii < p; ii++) {
                if (mode == CUDAPCG_XREDUCE_ONLYDIR && ii != shift) continue;
                fwrite(per[ii].data(), sizeof(double), hl, f);
            }
            fclose(f);
        }
    }
    for (int ii = 0; ii < p; ii++) {
        double v;
        if (mode == CUDAPCG_XREDUCE_NONE) {
            for (size_t jj = hl - 1; jj > 0; jj--) per[ii][jj - 1] += per[ii][jj];
            v = per[ii][0];
        } else if (mode == CUDAPCG_XREDUCE_ONLYDIR) {
            v = ii == shift ? per[ii][hl - 1] : 0.0;
        } else {
            v = per[ii][hl - 1];
        }
        g->xred_out[ii] = v;
    }
    g->xred_count = p;
    return 1;
}

static int ensure_dinv()
{
    if (g->dinv_valid) return 1;
    const size_t n = nvars();
    if (!g->dinv64) {
        if (is_lean()) return 1;  // the lean solvers allocate it only when they need it
        return 0;
    }
    if (g->precond) launch_jacobi<double>(g->dm.op64, g->dinv64, g->stream);
    else launch_fill<double>(g->dinv64, 1.0, n, g->stream);
    if (g->precision == CHFEM_PREC_MIXED) launch_convert<double, float>(g->dinv64, g->dinv32, n, g->stream);
    g->dinv_valid = true;
    return 1;
}

static double inner_tol_floor()
{
    const char *e = getenv("CHFEM_INNER_TOL");
    return e ? atof(e) : 1e-5;
}

static double reliable_drop()
{
    const char *e = getenv("CHFEM_RELIABLE_DROP");
    return e ? atof(e) : 1e-2;
}

// Mixed-precision Jacobi-PCG with reliable updates: one FP32 CG iteration
// sequence (FP32 operator and vectors, FP64 dots/alpha/beta) whose solution
// is flushed into the FP64 x and whose residual is replaced by the FP64
// residual b - A x whenever the iterated r.z has dropped by CHFEM_RELIABLE_DROP
// (default 1e-2) since the last such update, or looks converged.  The search
// direction is kept across updates, so the Krylov history is not lost as it
// is with restarted refinement.  Convergence is judged on the FP64 residual
// with the same rule as the FP64 solver (delta <= tol^2 delta0).
static int run_mixed_reliable(bool x0set, unsigned &it_total, unsigned &sweeps, double &residual)
{
    const OpDesc &o64 = g->dm.op64, &o32 = g->dm.op32;
    const size_t n = o64.ndof;
    hipStream_t st = g->stream;
    const double tol = g->num_tol;
    const bool kj = o32.kinv != nullptr;
    float *x = g->x32, *r = g->r32, *d = g->d32, *q = g->q32;
    const float *dinv = g->dinv32;
    // direction update inside the FP32 operator, on the preconditioned
    // residual z = M^-1 r (kept in r's buffer); iteration j (counted over the
    // whole solve, the direction is kept across replacements) writes ring
    // slot j % w, and x += alpha_j d_j is applied per window of w iterations
    // and at each segment end
    const bool fused = kj && g->s32 && fuse_dir_enabled() && apply_dir_ok(o32);
    const Ring<float> ring = fused ? dir_ring<float>(d, g->s32) : Ring<float>{};
    const unsigned W = fused ? (unsigned)ring.w : 1u;
    auto slot = [&](unsigned j) { return (float *)ring.p[j % W]; };
    Reduce rz = make_reduce(FIN_STORE, S_TMP0);
    // single-kernel iterations: iteration j reads z, q from buffer (j + 1) % 2
    // and writes buffer j % 2; a segment's fresh z goes to the buffer its
    // first iteration reads (r's buffer for the first segment)
    const bool cg1 = fused && loop_mode<float>() == 2 && g->z2_32 && g->q2_32;
    float *zb[2] = {g->z2_32, r}, *qb[2] = {q, g->q2_32};
    Reduce rc = make_guarded(FIN_CG1, 0);
    Reduce r0 = make_reduce(FIN_STORE, S_DELTA0), rd = make_reduce(FIN_STORE, S_DELTA);
    launch_precond<double>(g->dinv64, g->r64, nullptr, n, &r0, st);
    if (!sync_scalars()) return -1;
    const double delta0 = g->Sh[S_DELTA0];
    it_total = 0;
    sweeps = 0;
    residual = 0;
    if (fabs(delta0) < 1e-12) {
        launch_fill<double>(g->x64, 0.0, n, st);
        return 1;
    }
    double delta = delta0;
    if (x0set) {
        launch_apply<double>(o64, g->x64, g->q64, nullptr, st);
        launch_sub_precond<double>(g->r64, g->q64, g->dinv64, g->d64, &rd, n, st);
        if (!sync_scalars()) return -1;
        delta = g->Sh[S_DELTA];
        note_x0(delta, delta0);
        if (!g->x0_used) {
            // the guess is worse than zero: back to x = 0, r = b
            launch_add<double>(g->r64, g->q64, n, st);
            launch_fill<double>(g->x64, 0.0, n, st);
            delta = delta0;
        } else if (!above(delta, delta0, tol)) {
            residual = sqrt(fabs(delta) / fabs(delta0));
            return 1;
        }
    } else launch_fill<double>(g->x64, 0.0, n, st);
    const double drop = reliable_drop();
    double delta_rel = delta;
    // FP32 start: r = (float) r64, x = 0, d = M^-1 r, S[delta] = r.d
    launch_convert<double, float>(g->r64, r, n, st);
    launch_fill<float>(x, 0.0, n, st);
    if (fused) launch_kj_to_z<float>(o32, r, &rd, st);
    else if (kj) launch_kj_precond<float>(o32, r, d, &rd, st);
    else launch_precond<float>(dinv, r, d, n, &rd, st);
    // FP32 iterations are queued cg_batch() at a time; the device stops the
    // queue once r.z falls to the next replacement threshold
    // max(tol^2 delta0, drop * delta at the last replacement).
    Reduce ra = make_guarded(FIN_ALPHA, 0), rdl = make_guarded(FIN_DELTA, 0);
    const unsigned maxit = g->max_iterations;
    auto thresh = [&]() { return std::max(tol * tol * fabs(delta0), drop * fabs(delta_rel)); };
    if (!arm_guard(thresh())) return -1;
    bool converged = false;
    unsigned base = 0;  // iterations before the current guarded segment
    unsigned queued = 0, flushed = 0;  // fused: iterations queued, and folded into x
    while (true) {
        unsigned seg = 0;
        bool done = false;
        while (!done && base + seg < maxit) {
            const unsigned nb = std::min(cg_batch(), maxit - base - seg);
            for (unsigned b = 0; b < nb; b++) {
                const unsigned j = base + seg + b;
                if (fused) {
                    if (cg1) {
                        // z recovered from the directions from the segment's
                        // third iteration on (3 or more ring slots); only the
                        // segment's first iteration writes z
                        const bool zr = W >= 3 && j >= base + 2;
                        float *zw = W >= 3 && j >= base + 1 ? nullptr : zb[j & 1];
                        timed_apply_cg1<float>(o32, zb[(j + 1) & 1], qb[(j + 1) & 1], j > 0 ? slot(j - 1) : nullptr,
                                               zw, slot(j), qb[j & 1], j > base ? 1 : 0, &rc,
                                               b == (nb > 2 ? 2u : 0u), zr ? slot(j - 2) : nullptr);
                    } else {
                        const float *dold = j > 0 ? slot(j - 1) : nullptr;
                        if (b == 0) timed_apply_dir<float>(o32, r, dold, slot(j), q, &ra);
                        else launch_apply_dir<float>(o32, r, dold, slot(j), q, &ra, st);
                        launch_kj_zupdate<float>(o32, r, q, &rdl, st);
                    }
                    queued = j + 1;
                    if (queued % W == 0) {
                        launch_ring_flush<float>(x, ring, flushed, queued, base, g->S, n, st);
                        flushed = queued;
                    }
                    continue;
                }
                if (j > 0) {
                    if (kj) launch_kj_pcg_dir<float>(o32, d, r, g->S, st, 1);
                    else launch_pcg_dir<float>(d, q, g->S, 0, n, st, 1);
                }
                if (b == 0) timed_apply<float>(o32, d, q, &ra);
                else launch_apply<float>(o32, d, q, &ra, st);
                if (kj) launch_kj_pcg_update<float>(o32, x, r, d, q, &rdl, st);
                else launch_pcg_update<float>(x, r, d, q, dinv, &rdl, n, st);
            }
            if (!sync_loop()) return -1;
            seg = (unsigned)g->Sh[S_ITER];
            done = g->Sh[S_DONE] != 0.0;
        }
        it_total = base + seg;
        const bool last = it_total >= maxit;
        // the segment's iterations since the last window flush; the next
        // segment continues the ring from slot it_total % w
        if (fused) {
            launch_ring_flush<float>(x, ring, flushed, queued, base, g->S, n, st);
            flushed = queued = it_total;
        }
        // reliable update: x64 += x ; r64 -= A64 x ; x = 0 ; r = (float) r64
        launch_add_correction(g->x64, x, n, st);
        launch_convert<float, double>(x, g->d64, n, st);
        launch_fill<float>(x, 0.0, n, st);
        launch_apply<double>(o64, g->d64, g->q64, nullptr, st);
        launch_sub_precond<double>(g->r64, g->q64, g->dinv64, g->d64, &rd, n, st);  // S[delta] = FP64 r.z
        if (!sync_scalars()) return -1;
        sweeps++;
        const double prev = delta_rel;
        delta = g->Sh[S_DELTA];
        if (!above(delta, delta0, tol)) {
            converged = true;
            break;
        }
        if (last || !(fabs(delta) < fabs(prev))) break;  // no progress left in FP32
        delta_rel = delta;
        float *rs = cg1 ? zb[(it_total + 1) & 1] : r;  // the next segment's first z
        launch_convert<double, float>(g->r64, rs, n, st);
        if (fused) launch_kj_to_z<float>(o32, rs, &rz, st);
        else if (!kj) launch_precond<float>(dinv, r, q, n, nullptr, st);  // stored-dinv direction reads z from q
        base = it_total;
        if (!arm_guard(thresh())) return -1;
    }
    residual = sqrt(fabs(delta) / fabs(delta0));
    return converged ? 1 : 0;
}

// iterations without a 10 % drop of the FP32 inner MINRES residual after
// which a sweep ends (CHFEM_MINRES_STAGNATION, default 100)
static unsigned minres_stagnation()
{
    const char *e = getenv("CHFEM_MINRES_STAGNATION");
    const int v = e ? atoi(e) : 100;
    return v < 1 ? 1u : (unsigned)v;
}

// FP32 inner Krylov + FP64 outer iterative refinement.
static int run_mixed(bool x0set, unsigned &it_total, unsigned &sweeps, double &residual)
{
    if (!is_minres()) return run_mixed_reliable(x0set, it_total, sweeps, residual);
    const OpDesc &o64 = g->dm.op64, &o32 = g->dm.op32;
    const size_t n = o64.ndof;
    hipStream_t st = g->stream;
    const double tol = g->num_tol;
    Reduce r0 = make_reduce(FIN_STORE, S_DELTA0), rd = make_reduce(FIN_STORE, S_DELTA);
    launch_precond<double>(g->dinv64, g->r64, nullptr, n, &r0, st);
    if (!sync_scalars()) return -1;
    const double delta0 = g->Sh[S_DELTA0];
    it_total = 0;
    sweeps = 0;
    residual = 0;
    if (fabs(delta0) < 1e-12) {
        launch_fill<double>(g->x64, 0.0, n, st);
        return 1;
    }
    double delta = delta0;
    if (x0set) {
        launch_apply<double>(o64, g->x64, g->q64, nullptr, st);
        launch_sub_precond<double>(g->r64, g->q64, g->dinv64, g->d64, &rd, n, st);
        if (!sync_scalars()) return -1;
        delta = g->Sh[S_DELTA];
    } else launch_fill<double>(g->x64, 0.0, n, st);
    const double floor_tol = inner_tol_floor();
    while (above(delta, delta0, tol) && it_total < g->max_iterations) {
        double tau = 0.5 * tol * sqrt(fabs(delta0) / fabs(delta));
        if (tau < floor_tol) tau = floor_tol;
        if (tau > 0.5) tau = 0.5;
        launch_convert<double, float>(g->r64, g->r32, n, st);
        unsigned it = 0;
        double res_in = 0;
        int rc;
        if (is_minres())
            rc = run_minres<float>(o32, g->x32, g->r32, g->d32, g->q32, g->s32, g->dinv32, false, tau,
                                   g->max_iterations - it_total, it, res_in, minres_stagnation());
        else
            rc = run_pcg<float>(o32, g->x32, g->r32, g->d32, g->q32, g->dinv32, false, tau,
                                g->max_iterations - it_total, it, res_in, nullptr, g->s32);
        if (rc < 0) return -1;
        it_total += it;
        sweeps++;
        launch_add_correction(g->x64, g->x32, n, st);
        launch_convert<float, double>(g->x32, g->d64, n, st);
        launch_apply<double>(o64, g->d64, g->q64, nullptr, st);
        launch_sub_precond<double>(g->r64, g->q64, g->dinv64, g->d64, &rd, n, st);
        if (!sync_scalars()) return -1;
        const double prev = delta;
        delta = g->Sh[S_DELTA];
        if (it == 0 || !(fabs(delta) < fabs(prev))) {
            // FP32 sweeps no longer lower the FP64 residual (the saddle-point
            // system at tight tolerances): finish in FP64 from the refined
            // iterate.  r64 holds b - A x; b = r64 + A x restores the
            // right-hand side the FP64 solver measures its residual against.
            launch_apply<double>(o64, g->x64, g->q64, nullptr, st);
            launch_add<double>(g->r64, g->q64, n, st);
            unsigned it64 = 0;
            double res64 = 0;
            const int rc64 = run_minres<double>(o64, g->x64, g->r64, g->d64, g->q64, g->s64, g->dinv64, true, tol,
                                                g->max_iterations - it_total, it64, res64);
            if (rc64 < 0) return -1;
            it_total += it64;
            sweeps++;
            residual = res64;
            return rc64;
        }
    }
    residual = sqrt(fabs(delta) / fabs(delta0));
    return !above(delta, delta0, tol);
}

int solve_device(const double *x0_dev)
{
    if (!g || !g->model) return -1;
    if (!g->arrays && alloc_solver_arrays() != CUDAPCG_TRUE) return -1;
    if (!make_ops()) return -1;
    // (the lean solvers keep no buffers for the autotuner's trial launches:
    // they run one fixed tile)
    if (!is_lean()) autotune_tiles();
    ensure_dinv();
    const size_t n = nvars();
    bool x0set = g->x0_set;
    if (x0_dev) {
        // (two-vector solvers hold a guess in x64 only while they need it)
        if (!g->x64 && hipMalloc(&g->x64, n * sizeof(double)) != hipSuccess) return -1;
        hipMemcpyAsync(g->x64, x0_dev, n * sizeof(double), hipMemcpyDeviceToDevice, g->stream);
        x0set = true;
    }
    g->count++;
    g->x0_residual = 0;
    g->x0_used = 0;
    hipEventRecord(g->ev0, g->stream);
    unsigned it = 0, sweeps = 1;
    double res = 0;
    int rc;
    const bool rule = g->resnorm != CUDAPCG_L2_NORM;
    g->xred_count = 0;
    if (is_lean()) {
        rc = run_lean(x0set, it, res);
        if (rc >= 0 && is_two_vector() && !finish_two_vector(it)) rc = -1;
    } else if (g->precision == CHFEM_PREC_MIXED && !rule) {
        rc = run_mixed(x0set, it, sweeps, res);
    } else if (is_minres()) {
        rc = run_minres<double>(g->dm.op64, g->x64, g->r64, g->d64, g->q64, g->s64, g->dinv64, x0set, g->num_tol,
                                g->max_iterations, it, res);
    } else if (rule) {
        rc = run_pcg_rule(x0set, it, res);
    } else {
        rc = run_pcg<double>(g->dm.op64, g->x64, g->r64, g->d64, g->q64, g->dinv64, x0set, g->num_tol,
                             g->max_iterations, it, res, nullptr, g->s64);
    }
    hipEventRecord(g->ev1, g->stream);
    hipEventSynchronize(g->ev1);
    float ms = 0;
    hipEventElapsedTime(&ms, g->ev0, g->ev1);
    g->solve_seconds = ms * 1e-3;
    g->iteration = it;
    g->sweeps = sweeps;
    g->residual = res;
    g->mean_time = it ? g->solve_seconds / it : 0.0;
    g->found = rc == 1 ? CUDAPCG_TRUE : CUDAPCG_FALSE;
    if (rc < 0) {
        fprintf(stderr, "chfem: solver failed on the device\n");
        return -1;
    }
    printf("%siteration: %u, residual: %.3e\n", g->header ? g->header : "", it, res);
    return rc;
}

}  // namespace chfem

using namespace chfem;

extern "C" {

cudapcgFlag_t cudapcgSetNumTol(cudapcgTol_t t) { if (!g) return CUDAPCG_FALSE; g->num_tol = t; return CUDAPCG_TRUE; }
cudapcgFlag_t cudapcgSetMaxIterations(unsigned int n) { if (!g) return CUDAPCG_FALSE; g->max_iterations = n; return CUDAPCG_TRUE; }
cudapcgFlag_t cudapcgSetXReduceMode(cudapcgFlag_t flag)
{
    if (!g) return CUDAPCG_FALSE;
    g->xreduce_flag = flag <= CUDAPCG_XREDUCE_FULL ? flag : CUDAPCG_XREDUCE_FULL;
    return CUDAPCG_TRUE;
}
cudapcgFlag_t cudapcgSetXReduceShift(unsigned int shift) { if (!g) return CUDAPCG_FALSE; g->xreduce_shift = shift; return CUDAPCG_TRUE; }
cudapcgFlag_t cudapcgSetXReduceScale(double scl) { if (!g) return CUDAPCG_FALSE; g->xreduce_scale = scl; return CUDAPCG_TRUE; }
cudapcgFlag_t cudapcgSetReduceStabFactor(double scl) { if (!g) return CUDAPCG_FALSE; g->reduce_stab = scl; return CUDAPCG_TRUE; }
unsigned int cudapcgGetNumIterations(void) { return g ? g->iteration : 0; }
unsigned int cudapcgGetMaxNumIterations(void) { return g ? g->max_iterations : 0; }
cudapcgVar_t cudapcgGetResidual(void) { return g ? g->residual : 0; }

cudapcgFlag_t cudapcgSetPrecision(cudapcgFlag_t flag)
{
    if (!g) return CUDAPCG_FALSE;
    if (flag > CHFEM_PREC_MIXED) return CUDAPCG_FALSE;
    if (flag != g->precision && g->arrays) {
        g->precision = flag;
        return alloc_solver_arrays();
    }
    g->precision = flag;
    return CUDAPCG_TRUE;
}

cudapcgFlag_t cudapcgSetKernelTiming(cudapcgFlag_t on)
{
    if (!g) return CUDAPCG_FALSE;
    g->timing = on != 0;
    g->op_ms = 0;
    g->op_count = 0;
    return CUDAPCG_TRUE;
}

cudapcgFlag_t cudapcgGetKernelTiming(double *op_ms_total, unsigned int *op_launches)
{
    if (!g) return CUDAPCG_FALSE;
    if (op_ms_total) *op_ms_total = g->op_ms;
    if (op_launches) *op_launches = g->op_count;
    return CUDAPCG_TRUE;
}

cudapcgFlag_t cudapcgGetSolveStats(chfemSolveStats_t *s)
{
    if (!g || !s) return CUDAPCG_FALSE;
    s->iterations = g->iteration;
    s->sweeps = g->sweeps;
    s->solve_seconds = g->solve_seconds;
    s->residual = g->residual;
    s->x0_residual = g->x0_residual;
    s->x0_used = g->x0_used;
    return CUDAPCG_TRUE;
}

cudapcgFlag_t cudapcgSetSolver(cudapcgFlag_t flag)
{
    if (!g) return CUDAPCG_FALSE;
    if (flag > CUDAPCG_MINRES2_SOLVER) {
        printf("ERROR: Invalid \"solverFlag\" (%d).\n", flag);
        return CUDAPCG_FALSE;
    }
    const bool was_minres = is_minres(), was_lean = is_lean(), was_two = is_two_vector();
    g->solver_flag = flag;
    if (g->arrays && (was_lean != is_lean() || was_two != is_two_vector())) {
        free_solver_arrays();
        return alloc_solver_arrays();
    }
    if (g->arrays && is_minres() && !was_minres) return alloc_solver_arrays();
    return CUDAPCG_TRUE;
}

cudapcgFlag_t cudapcgSetPreconditioner(cudapcgFlag_t flag)
{
    if (!g) return CUDAPCG_FALSE;
    g->precond = flag > 0 ? CUDAPCG_TRUE : CUDAPCG_FALSE;
    g->dinv_valid = false;
    return CUDAPCG_TRUE;
}

cudapcgFlag_t cudapcgSetResNorm(cudapcgFlag_t flag)
{
    if (!g) return CUDAPCG_FALSE;
    g->resnorm = flag > CUDAPCG_ERROR_NORM ? CUDAPCG_L2_NORM : flag;
    return CUDAPCG_TRUE;
}

static void free_model_device(cudapcgModel_t *m)
{
    if (!m) return;
    dfree(m->image); dfree(m->parametric_density_field); dfree(m->pore_map); dfree(m->border_pore_map);
    dfree(m->periodic2DOF_map); dfree(m->DOF2periodic_map);
    if (m->name) free(m->name);
    m->name = nullptr;
}

cudapcgFlag_t cudapcgEnd(void)
{
    if (!g) return CUDAPCG_FALSE;
    if (g->stream) hipStreamSynchronize(g->stream);
    free(g->header);
    g->x0_set = false;
    free_solver_arrays();
    dfree(g->x64);
    dfree(g->r64);
    dfree(g->dm.K64);
    dfree(g->dm.K32);
    dfree(g->lamu64);
    dfree(g->lamu32);
    dfree(g->fkeys);
    g->fkeys_n = 0;
    dfree(g->kinv64);
    dfree(g->kinv32);
    dfree(g->btab64);
    if (g->model && g->model->freeAllowed_flag) {
        free_model_device(g->model);
        free(g->model);
    }
    if (g->ev0) hipEventDestroy(g->ev0);
    if (g->ev1) hipEventDestroy(g->ev1);
    if (g->opA) hipEventDestroy(g->opA);
    if (g->opB) hipEventDestroy(g->opB);
    if (g->stream) hipStreamDestroy(g->stream);
    delete g;
    g = nullptr;
    return CUDAPCG_TRUE;
}

cudapcgFlag_t cudapcgInit(cudapcgFlag_t analysis_flag, cudapcgFlag_t parStrategy_flag)
{
    cudapcgEnd();
    if (analysis_flag != CUDAPCG_THERMAL_3D && analysis_flag != CUDAPCG_ELASTIC_3D && analysis_flag != CUDAPCG_FLUID_3D) {
        if (analysis_flag <= CUDAPCG_FLUID_3D)
            printf("ERROR: 2D analyses are not provided by this library.\n");
        return CUDAPCG_FALSE;
    }
    if (analysis_flag == CUDAPCG_FLUID_3D && parStrategy_flag != CUDAPCG_NBN) {
        if (parStrategy_flag == CUDAPCG_EBE) printf("ERROR: EBE solver is unavailable for FLUID analysis.\n");
        else printf("ERROR: Invalid \"parStrategy_flag\" (%d) for FLUID analysis.\n", parStrategy_flag);
        return CUDAPCG_FALSE;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        fprintf(stderr, "chfem: no HIP device available\n");
        return CUDAPCG_FALSE;
    }
    g = new Solver();
    g->analysis_flag = analysis_flag;
    g->parStrategy_flag = parStrategy_flag;
    g->header = (char *)malloc(1);
    g->header[0] = 0;
    if (hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&g->ev0) != hipSuccess || hipEventCreate(&g->ev1) != hipSuccess ||
        hipEventCreate(&g->opA) != hipSuccess || hipEventCreate(&g->opB) != hipSuccess) {
        fprintf(stderr, "chfem: failed to create the solver stream\n");
        delete g;
        g = nullptr;
        return CUDAPCG_FALSE;
    }
    return CUDAPCG_TRUE;
}

cudapcgModel_t *cudapcgNewModel(void)
{
    cudapcgModel_t *m = (cudapcgModel_t *)calloc(1, sizeof(cudapcgModel_t));
    if (!m) return nullptr;
    m->freeAllowed_flag = CUDAPCG_TRUE;
    m->parStrategy_flag = CUDAPCG_NBN;
    m->poremap_flag = CUDAPCG_POREMAP_NUM;
    return m;
}

static cudapcgFlag_t alloc_model_device()
{
    cudapcgModel_t *m = g->model;
    const size_t n = m->nvars;
    dfree(g->r64);
    free_solver_arrays();
    CHECK_HIP(hipMalloc(&g->r64, n * sizeof(double)));
    if (g->analysis_flag != CUDAPCG_FLUID_3D) {
        if (!m->image) CHECK_HIP(hipMalloc(&m->image, sizeof(cudapcgMap_t) * m->nelem));
    } else {
        if (!m->periodic2DOF_map) CHECK_HIP(hipMalloc(&m->periodic2DOF_map, sizeof(cudapcgIdMap_t) * m->nelem));
        if (m->poremap_flag == CUDAPCG_POREMAP_NUM) {
            if (!m->border_pore_map)
                CHECK_HIP(hipMalloc(&m->border_pore_map, sizeof(cudapcgFlag_t) * (m->nbordernodes + 1)));
            if (!m->DOF2periodic_map)
                CHECK_HIP(hipMalloc(&m->DOF2periodic_map,
                                    sizeof(cudapcgIdMap_t) * ((size_t)m->nporenodes + m->nbordernodes + 1)));
        } else {
            printf("ERROR: image pore mapping is not provided by this library; use DOF-number mapping.\n");
            return CUDAPCG_FALSE;
        }
    }
    return CUDAPCG_TRUE;
}

cudapcgFlag_t cudapcgSetModel(cudapcgModel_t *model)
{
    if (!g || !model) return CUDAPCG_FALSE;
    if (g->model && g->model != model && g->model->freeAllowed_flag) {
        free_model_device(g->model);
        free(g->model);
    }
    g->model = model;
    g->tuned = false;
    g->ops_ready = false;
    return alloc_model_device();
}

cudapcgFlag_t cudapcgSetModelConstructorFcn(cudapcgFlag_t (*fcn)(cudapcgModel_t **, const void *))
{
    if (!fcn) return CUDAPCG_FALSE;
    g_ctor = fcn;
    return CUDAPCG_TRUE;
}

cudapcgFlag_t cudapcgBuildModel(const void *data)
{
    if (!g) return CUDAPCG_FALSE;
    if (!g->model) g->model = cudapcgNewModel();
    else if (g->model->freeAllowed_flag) {
        free_model_device(g->model);
        free(g->model);
        g->model = cudapcgNewModel();
    }
    if (!g_ctor(&g->model, data)) return CUDAPCG_FALSE;
    return alloc_model_device();
}

cudapcgFlag_t cudapcgAllocateArrays(void)
{
    if (!g || !g->model) return CUDAPCG_FALSE;
    cudapcgFlag_t ok = alloc_solver_arrays();
    g->user_arrays = ok == CUDAPCG_TRUE;
    return ok;
}

cudapcgFlag_t cudapcgFreeArrays(void)
{
    if (!g) return CUDAPCG_FALSE;
    return free_solver_arrays();
}

cudapcgFlag_t cudapcgSetRHS(cudapcgVar_t *RHS)
{
    if (!g || !g->model || !RHS) return CUDAPCG_FALSE;
    CHECK_HIP(hipMemcpyAsync(g->r64, RHS, nvars() * sizeof(double), hipMemcpyHostToDevice, g->stream));
    CHECK_HIP(hipStreamSynchronize(g->stream));
    return CUDAPCG_TRUE;
}

cudapcgFlag_t cudapcgSetX0(cudapcgVar_t *x0, cudapcgFlag_t mustInterpolate)
{
    if (!g || !g->model) return CUDAPCG_FALSE;
    const size_t n = nvars();
    if (!x0) {
        g->x0_set = false;
        return CUDAPCG_TRUE;
    }
    if (!g->x64) CHECK_HIP(hipMalloc(&g->x64, n * sizeof(double)));
    if (mustInterpolate) {
        cudapcgModel_t *m = g->model;
        const unsigned stride = m->nvars / m->nelem;
        const size_t nc = (size_t)stride * (m->ncols / 2) * (m->nrows / 2) * (m->nlayers / 2 + (m->nlayers < 2));
        double *tmp = nullptr;
        CHECK_HIP(hipMalloc(&tmp, nc * sizeof(double)));
        CHECK_HIP(hipMemcpyAsync(tmp, x0, nc * sizeof(double), hipMemcpyHostToDevice, g->stream));
        launch_interpl2(tmp, g->x64, m->nrows, m->ncols, m->nlayers, stride, g->stream);
        CHECK_HIP(hipStreamSynchronize(g->stream));
        hipFree(tmp);
    } else {
        CHECK_HIP(hipMemcpyAsync(g->x64, x0, n * sizeof(double), hipMemcpyHostToDevice, g->stream));
        CHECK_HIP(hipStreamSynchronize(g->stream));
    }
    g->x0_set = true;
    return CUDAPCG_TRUE;
}

cudapcgFlag_t cudapcgSetImage(cudapcgMap_t *img)
{
    if (!g || !g->model || !img || !g->model->image) return CUDAPCG_FALSE;
    CHECK_HIP(hipMemcpyAsync(g->model->image, img, g->model->nelem * sizeof(cudapcgMap_t), hipMemcpyHostToDevice,
                             g->stream));
    CHECK_HIP(hipStreamSynchronize(g->stream));
    g->dinv_valid = false;
    return CUDAPCG_TRUE;
}

// Scalar density field (chfem/src/femhmg/cudapcg/cudapcg.cu:339-352): uint16
// per element, element scale field * (fmax - fmin) / 65535 + fmin applied to
// the local matrices of the element-by-element operator.
cudapcgFlag_t cudapcgSetParametricDensityField(parametricScalarField_t *field, double fmin, double fmax)
{
    if (!g || !g->model || !field) return CUDAPCG_FALSE;
    cudapcgModel_t *m = g->model;
    if (g->analysis_flag == CUDAPCG_FLUID_3D) {
        printf("ERROR: scalar density field input is not supported for permeability analysis.\n");
        return CUDAPCG_FALSE;
    }
    if (!m->parametric_density_field)
        CHECK_HIP(hipMalloc(&m->parametric_density_field, sizeof(parametricScalarField_t) * m->nelem));
    CHECK_HIP(hipMemcpyAsync(m->parametric_density_field, field, sizeof(parametricScalarField_t) * m->nelem,
                             hipMemcpyHostToDevice, g->stream));
    CHECK_HIP(hipStreamSynchronize(g->stream));
    m->limits_density_field[0] = fmin;
    m->limits_density_field[1] = fmax;
    m->parametric_density_field_flag = CUDAPCG_TRUE;
    g->dinv_valid = false;
    return CUDAPCG_TRUE;
}

cudapcgFlag_t cudapcgSetPoreMap(cudapcgFlag_t *pores)
{
    if (!g || !g->model || !pores) return CUDAPCG_FALSE;
    if (g->model->poremap_flag != CUDAPCG_POREMAP_NUM) return CUDAPCG_FALSE;
    if (g->model->nbordernodes)
        CHECK_HIP(hipMemcpyAsync(g->model->border_pore_map, pores, g->model->nbordernodes, hipMemcpyHostToDevice,
                                 g->stream));
    CHECK_HIP(hipStreamSynchronize(g->stream));
    g->dinv_valid = false;
    return CUDAPCG_TRUE;
}

cudapcgFlag_t cudapcgSetPeriodic2DOFMap(cudapcgIdMap_t *vars)
{
    if (!g || !g->model || !vars) return CUDAPCG_FALSE;
    CHECK_HIP(hipMemcpyAsync(g->model->periodic2DOF_map, vars, g->model->nelem * sizeof(cudapcgIdMap_t),
                             hipMemcpyHostToDevice, g->stream));
    CHECK_HIP(hipStreamSynchronize(g->stream));
    return CUDAPCG_TRUE;
}

cudapcgFlag_t cudapcgSetDOF2PeriodicMap(cudapcgIdMap_t *nodes)
{
    if (!g || !g->model || !nodes) return CUDAPCG_FALSE;
    if (g->model->poremap_flag != CUDAPCG_POREMAP_NUM) return CUDAPCG_FALSE;
    const size_t n = (size_t)g->model->nporenodes + g->model->nbordernodes;
    if (n) CHECK_HIP(hipMemcpyAsync(g->model->DOF2periodic_map, nodes, n * sizeof(cudapcgIdMap_t),
                                    hipMemcpyHostToDevice, g->stream));
    CHECK_HIP(hipStreamSynchronize(g->stream));
    return CUDAPCG_TRUE;
}

// Local matrices in the reference layout (chfem/src/femhmg/cudapcg/cudapcg.cu:365-375):
// thermal 40 stored non-zeros per material, elastic 24x24, Stokes one 32x32.
// They are expanded to dense blocks and kept in FP64 and FP32.
cudapcgFlag_t cudapcgSetLclMtxs(cudapcgVar_t *L)
{
    if (!g || !g->model || !L) return CUDAPCG_FALSE;
    const unsigned nk = g->model->nkeys ? g->model->nkeys : 1;
    const int ph = phys(g->analysis_flag);
    const size_t ksz = ph == PHY_THERMAL ? 64 : (ph == PHY_ELASTIC ? 576 : 1024);
    double *K = (double *)calloc(ksz * nk, sizeof(double));
    float *Kf = (float *)calloc(ksz * nk, sizeof(float));
    static const int dr[8] = {1, 1, 0, 0, 1, 1, 0, 0}, dc[8] = {0, 1, 1, 0, 0, 1, 1, 0}, dl[8] = {0, 0, 0, 0, 1, 1, 1, 1};
    for (unsigned m = 0; m < nk; m++) {
        if (ph == PHY_THERMAL) {
            const double *src = L + (size_t)m * g->model->localmtxdim;
            int t = 0;
            for (int i = 0; i < 8; i++)
                for (int j = 0; j < 8; j++) {
                    // a conduction matrix of the voxel vanishes between nodes that differ in one coordinate
                    const int diff = (dr[i] != dr[j]) + (dc[i] != dc[j]) + (dl[i] != dl[j]);
                    K[m * 64 + i * 8 + j] = diff == 1 ? 0.0 : src[t++];
                }
        } else {
            memcpy(K + m * ksz, L + (size_t)m * ksz, ksz * sizeof(double));
        }
    }
    for (size_t i = 0; i < ksz * nk; i++) Kf[i] = (float)K[i];
    // inverse Jacobi diagonal of every 16-bit key (2 bits per local node: at
    // most 4 materials) for the key-Jacobi vector kernels; sums in the order
    // k_jacobi_keys uses, so the values equal the stored dinv
    dfree(g->kinv64);
    dfree(g->kinv32);
    if (ph != PHY_FLUID && nk <= 4) {
        const int nd = ph == PHY_ELASTIC ? 3 : 1;
        std::vector<double> t64;
        std::vector<float> t32;
        voxel::inverse_diagonal_table(K, nk, nd, t64, t32);
        if (hipMalloc(&g->kinv64, t64.size() * sizeof(double)) != hipSuccess ||
            hipMalloc(&g->kinv32, t32.size() * sizeof(float)) != hipSuccess ||
            hipMemcpy(g->kinv64, t64.data(), t64.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(g->kinv32, t32.data(), t32.size() * sizeof(float), hipMemcpyHostToDevice) != hipSuccess) {
            fprintf(stderr, "chfem: failed to upload the Jacobi table\n");
            free(K);
            free(Kf);
            return CUDAPCG_FALSE;
        }
    }
    // Stokes Jacobi per DOF class for the memory-lean solvers, summed in the
    // order k_jacobi_fluid uses, so the values equal the stored diagonal's
    if (ph == PHY_FLUID) {
        double dv = 0, dp = 0, bt[256];
        for (int e = 0; e < 8; e++) {
            dv += K[(3 * e) * 32 + 3 * e];
            dp -= K[(24 + e) * 32 + 24 + e];
        }
        g->fl_vinv = 1.0 / dv;
        g->fl_pinv = 1.0 / dp;
        for (int k = 0; k < 256; k++) {
            double dd = 0;
            for (int e = 0; e < 8; e++)
                if ((k >> e) & 1) dd -= K[(24 + e) * 32 + 24 + e];
            bt[k] = dd != 0.0 ? 1.0 / dd : 0.0;
        }
        if ((!g->btab64 && hipMalloc(&g->btab64, sizeof(bt)) != hipSuccess) ||
            hipMemcpy(g->btab64, bt, sizeof(bt), hipMemcpyHostToDevice) != hipSuccess) {
            fprintf(stderr, "chfem: failed to upload the Stokes Jacobi table\n");
            free(K);
            free(Kf);
            return CUDAPCG_FALSE;
        }
    }
    dfree(g->dm.K64);
    dfree(g->dm.K32);
    hipError_t e1 = hipMalloc(&g->dm.K64, ksz * nk * sizeof(double));
    hipError_t e2 = hipMalloc(&g->dm.K32, ksz * nk * sizeof(float));
    if (e1 == hipSuccess && e2 == hipSuccess) {
        e1 = hipMemcpy(g->dm.K64, K, ksz * nk * sizeof(double), hipMemcpyHostToDevice);
        e2 = hipMemcpy(g->dm.K32, Kf, ksz * nk * sizeof(float), hipMemcpyHostToDevice);
    }
    free(Kf);
    if (e1 != hipSuccess || e2 != hipSuccess) {
        free(K);
        fprintf(stderr, "chfem: failed to upload local matrices\n");
        return CUDAPCG_FALSE;
    }
    g->dm.ksz = ksz;
    g->dm.nmat = nk;
    g->dinv_valid = false;
    // isotropic elastic or conduction materials: the operator can use the
    // Walsh-basis kernel (per material {lambda, mu} or {k, 0})
    g->walsh_ok = false;
    g->tuned = false;
    g->mode64 = g->mode32 = 1;
    if ((ph == PHY_ELASTIC || ph == PHY_THERMAL) && nk <= 4) {
        double lm[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        float lmf[8];
        bool ok = true;
        for (unsigned m = 0; m < nk && ok; m++) {
            if (ph == PHY_ELASTIC) ok = voxel::isotropic_params(L + (size_t)m * 576, lm[2 * m], lm[2 * m + 1]);
            else ok = voxel::conduction_param(K + (size_t)m * 64, lm[2 * m]);
        }
        for (int i = 0; i < 8; i++) lmf[i] = (float)lm[i];
        if (ok) {
            if (!g->lamu64) CHECK_HIP(hipMalloc(&g->lamu64, 8 * sizeof(double)));
            if (!g->lamu32) CHECK_HIP(hipMalloc(&g->lamu32, 8 * sizeof(float)));
            CHECK_HIP(hipMemcpy(g->lamu64, lm, sizeof(lm), hipMemcpyHostToDevice));
            CHECK_HIP(hipMemcpy(g->lamu32, lmf, sizeof(lmf), hipMemcpyHostToDevice));
            g->walsh_ok = true;
        }
    }
    free(K);
    return CUDAPCG_TRUE;
}

cudapcgFlag_t cudapcgSolve(cudapcgVar_t *x)
{
    if (!g || !g->model || !x) return CUDAPCG_FALSE;
    if (g->model->poremap_flag != CUDAPCG_POREMAP_NUM && g->analysis_flag == CUDAPCG_FLUID_3D) return CUDAPCG_FALSE;
    const bool temp = !g->arrays;
    double t = omp_get_wtime();
    int rc = solve_device(nullptr);
    if (rc < 0) return CUDAPCG_FALSE;
    const size_t n = nvars();
    if (is_two_vector()) {
        // two-vector solvers return the reduced solution, one value per
        // node value (chfem/src/femhmg/cudapcg/solvers/pcg2.cu:250-262)
        for (int a = 0; a < g->xred_count; a++) x[a] = g->xred_out[a];
    } else {
        CHECK_HIP(hipMemcpyAsync(x, g->x64, n * sizeof(double), hipMemcpyDeviceToHost, g->stream));
        CHECK_HIP(hipStreamSynchronize(g->stream));
    }
    if (temp && !g->user_arrays) free_solver_arrays();
    g->total_time = omp_get_wtime() - t;
    return g->found;
}

cudapcgFlag_t cudapcgApplyOperator(const cudapcgVar_t *v, cudapcgVar_t *q)
{
    if (g && is_lean()) {
        fprintf(stderr, "chfem: this entry point needs the CG or MINRES solver's vectors\n");
        return CUDAPCG_FALSE;
    }
    if (!g || !g->model || !v || !q) return CUDAPCG_FALSE;
    if (!g->arrays && alloc_solver_arrays() != CUDAPCG_TRUE) return CUDAPCG_FALSE;
    make_ops();
    const size_t n = nvars();
    CHECK_HIP(hipMemcpyAsync(g->d64, v, n * sizeof(double), hipMemcpyHostToDevice, g->stream));
    launch_apply<double>(g->dm.op64, g->d64, g->q64, nullptr, g->stream);
    CHECK_HIP(hipGetLastError());
    CHECK_HIP(hipMemcpyAsync(q, g->q64, n * sizeof(double), hipMemcpyDeviceToHost, g->stream));
    CHECK_HIP(hipStreamSynchronize(g->stream));
    return CUDAPCG_TRUE;
}

cudapcgFlag_t cudapcgApplyOperatorF32(const float *v, float *q)
{
    if (g && is_lean()) {
        fprintf(stderr, "chfem: this entry point needs the CG or MINRES solver's vectors\n");
        return CUDAPCG_FALSE;
    }
    if (!g || !g->model || !v || !q) return CUDAPCG_FALSE;
    if (!g->arrays && alloc_solver_arrays() != CUDAPCG_TRUE) return CUDAPCG_FALSE;
    make_ops();
    const size_t n = nvars();
    float *dv = nullptr, *dq = nullptr;
    if (hipMalloc(&dv, n * sizeof(float)) != hipSuccess || hipMalloc(&dq, n * sizeof(float)) != hipSuccess) {
        dfree(dv);
        return CUDAPCG_FALSE;
    }
    bool ok = hipMemcpyAsync(dv, v, n * sizeof(float), hipMemcpyHostToDevice, g->stream) == hipSuccess;
    if (ok) {
        launch_apply<float>(g->dm.op32, dv, dq, nullptr, g->stream);
        ok = hipGetLastError() == hipSuccess &&
             hipMemcpyAsync(q, dq, n * sizeof(float), hipMemcpyDeviceToHost, g->stream) == hipSuccess &&
             hipStreamSynchronize(g->stream) == hipSuccess;
    }
    dfree(dv);
    dfree(dq);
    return ok ? CUDAPCG_TRUE : CUDAPCG_FALSE;
}

cudapcgFlag_t cudapcgApplyPreconditioner(const cudapcgVar_t *v, cudapcgVar_t *z)
{
    if (g && is_lean()) {
        fprintf(stderr, "chfem: this entry point needs the CG or MINRES solver's vectors\n");
        return CUDAPCG_FALSE;
    }
    if (!g || !g->model || !v || !z) return CUDAPCG_FALSE;
    if (!g->arrays && alloc_solver_arrays() != CUDAPCG_TRUE) return CUDAPCG_FALSE;
    make_ops();
    ensure_dinv();
    const size_t n = nvars();
    CHECK_HIP(hipMemcpyAsync(g->d64, v, n * sizeof(double), hipMemcpyHostToDevice, g->stream));
    launch_precond<double>(g->dinv64, g->d64, g->q64, n, nullptr, g->stream);
    CHECK_HIP(hipMemcpyAsync(z, g->q64, n * sizeof(double), hipMemcpyDeviceToHost, g->stream));
    CHECK_HIP(hipStreamSynchronize(g->stream));
    return CUDAPCG_TRUE;
}

double cudapcgOperatorBench(cudapcgFlag_t precision, unsigned int reps, cudapcgFlag_t fused_dot)
{
    if (g && is_lean()) {
        fprintf(stderr, "chfem: this entry point needs the CG or MINRES solver's vectors\n");
        return -1.0;
    }
    if (!g || !g->model) return -1.0;
    if (!g->arrays && alloc_solver_arrays() != CUDAPCG_TRUE) return -1.0;
    if (precision == CHFEM_PREC_MIXED && !g->d32) {
        g->precision = CHFEM_PREC_MIXED;
        if (alloc_solver_arrays() != CUDAPCG_TRUE) return -1.0;
    }
    make_ops();
    const size_t n = nvars();
    // two inputs (d and x) so that consecutive launches do not re-read one
    // cache-resident vector
    launch_fill<double>(g->d64, 1.0, n, g->stream);
    launch_fill<double>(g->x64, 0.5, n, g->stream);
    if (precision == CHFEM_PREC_MIXED) {
        launch_convert<double, float>(g->d64, g->d32, n, g->stream);
        launch_convert<double, float>(g->x64, g->x32, n, g->stream);
    }
    Reduce red = make_reduce(FIN_STORE, S_TMP0), red3 = make_reduce(FIN_STORE, S_TMP0);
    const Reduce *rp = fused_dot ? &red : nullptr;
    const bool dir = fused_dot == 2 || fused_dot == 3, cg1 = fused_dot == 3;
    if (dir) {
        // CG-loop form: r = x, directions alternate between d and s
        if (!g->precond || !apply_dir_ok(precision == CHFEM_PREC_MIXED ? g->dm.op32 : g->dm.op64)) return -1.0;
        if (!set_scalar(S_DELTA_OLD, 1.0) || !set_scalar(S_DELTA, 0.5) || !set_scalar(S_ALPHA, 0.0) ||
            !set_scalar(S_BPREV, 0.5))
            return -1.0;
        if (precision == CHFEM_PREC_MIXED && !g->dinv32) return -1.0;
        launch_fill<double>(g->s64, 1.0, n, g->stream);
        launch_fill<double>(g->q64, 0.25, n, g->stream);
        launch_fill<double>(g->q2_64, 0.25, n, g->stream);
        launch_fill<double>(g->z2_64, 1.0, n, g->stream);
        if (precision == CHFEM_PREC_MIXED) {
            launch_fill<float>(g->s32, 1.0, n, g->stream);
            launch_fill<float>(g->q32, 0.25, n, g->stream);
            launch_fill<float>(g->q2_32, 0.25, n, g->stream);
            launch_fill<float>(g->z2_32, 1.0, n, g->stream);
        }
    }
    auto once = [&](unsigned k) {
        if (cg1) {
            // the solver's steady iteration: z recovered from the two
            // previous directions (d rotates over d / s / z2), not written;
            // q alternates between q / q2; the deferred z update runs with
            // S[ALPHA] = 0
            if (precision == CHFEM_PREC_MIXED) {
                float *qb[2] = {g->q32, g->q2_32}, *db[3] = {g->d32, g->s32, g->z2_32};
                launch_apply_cg1<float>(g->dm.op32, g->x32, qb[k & 1], db[(k + 1) % 3], nullptr, db[(k + 2) % 3],
                                        qb[(k + 1) & 1], 1, &red3, g->stream, nullptr, db[k % 3]);
            } else {
                double *qb[2] = {g->q64, g->q2_64}, *db[3] = {g->d64, g->s64, g->z2_64};
                launch_apply_cg1<double>(g->dm.op64, g->x64, qb[k & 1], db[(k + 1) % 3], nullptr, db[(k + 2) % 3],
                                         qb[(k + 1) & 1], 1, &red3, g->stream, nullptr, db[k % 3]);
            }
            return;
        }
        if (dir) {
            // z = x
            if (precision == CHFEM_PREC_MIXED)
                launch_apply_dir<float>(g->dm.op32, g->x32, k & 1 ? g->s32 : g->d32, k & 1 ? g->d32 : g->s32, g->q32,
                                        &red, g->stream);
            else
                launch_apply_dir<double>(g->dm.op64, g->x64, k & 1 ? g->s64 : g->d64, k & 1 ? g->d64 : g->s64, g->q64,
                                         &red, g->stream);
            return;
        }
        if (precision == CHFEM_PREC_MIXED) launch_apply<float>(g->dm.op32, k & 1 ? g->x32 : g->d32, g->q32, rp, g->stream);
        else launch_apply<double>(g->dm.op64, k & 1 ? g->x64 : g->d64, g->q64, rp, g->stream);
    };
    once(0);
    once(1);
    hipEventRecord(g->ev0, g->stream);
    for (unsigned r = 0; r < reps; r++) once(r);
    hipEventRecord(g->ev1, g->stream);
    if (hipEventSynchronize(g->ev1) != hipSuccess) return -1.0;
    float ms = 0;
    hipEventElapsedTime(&ms, g->ev0, g->ev1);
    return reps ? ms / reps : 0.0;
}

int cudapcgGetLoopForm(cudapcgFlag_t precision)
{
    if (!g) return -1;
    return precision == CHFEM_PREC_MIXED ? loop_mode<float>() : loop_mode<double>();
}

cudapcgFlag_t cudapcgGetOperatorTile(cudapcgFlag_t precision, int *tile)
{
    if (!g || !g->model || !tile) return CUDAPCG_FALSE;
    const OpDesc &o = precision == CHFEM_PREC_MIXED ? g->dm.op32 : g->dm.op64;
    tile[0] = o.tile_tr;
    tile[1] = o.tile_tc;
    tile[2] = o.tile_lz;
    return CUDAPCG_TRUE;
}

cudapcgFlag_t cudapcgSetHeaderString(char *header)
{
    if (!g) return CUDAPCG_FALSE;
    free(g->header);
    const char *h = header ? header : "";
    g->header = (char *)malloc(strlen(h) + 1);
    strcpy(g->header, h);
    return CUDAPCG_TRUE;
}

cudapcgFlag_t cudapcgPrintSolverMetrics(void)
{
    if (!g) return CUDAPCG_FALSE;
    printf("%sMean time per iteration: %.2e s\n", g->header, g->mean_time);
    printf("%sElapsed time: %.2e s\n", g->header, g->total_time);
    return CUDAPCG_TRUE;
}

cudapcgFlag_t cudapcgPrintSolverReport(void)
{
    if (!g) return CUDAPCG_FALSE;
    if (g->found) printf("%sPCG solver found a solution in %d iterations.\n", g->header, g->iteration);
    else printf("%sPCG solver did not converge within provided max number of iterations (%d).\n", g->header,
                g->max_iterations);
    printf("%sResidual: %e\n", g->header, g->residual);
    return cudapcgPrintSolverMetrics();
}

cudapcgFlag_t cudapcgPrintSolverReport2(char *dest)
{
    if (!g || !dest) return CUDAPCG_FALSE;
    int sz = sprintf(dest,
                     "%sPCG solver ran through %d iterations.\n%sResidual: %e\n%sMean time per iteration: %.2e s\n"
                     "%sElapsed time: %.2e s\n",
                     g->header, g->iteration, g->header, g->residual, g->header, g->mean_time, g->header,
                     g->total_time);
    return sz < 0 ? CUDAPCG_FALSE : CUDAPCG_TRUE;
}

cudapcgFlag_t cudapcgPrintSolverMetrics2(char *dest)
{
    if (!g || !dest) return CUDAPCG_FALSE;
    int sz = sprintf(dest, "%sMean time per iteration: %.2e s\n%sElapsed time: %.2e s\n", g->header, g->mean_time,
                     g->header, g->total_time);
    return sz < 0 ? CUDAPCG_FALSE : CUDAPCG_TRUE;
}

}  // extern "C"
