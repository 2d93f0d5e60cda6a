/*
 * chfem_amd.h -- entry points of the MI355X chfem library that have no
 * counterpart in the reference, or that replace its CPython extension with a
 * plain C ABI.
 */
#ifndef CHFEM_AMD_EXT_H
#define CHFEM_AMD_EXT_H

#include <stddef.h>
#include <stdint.h>
#include "cudapcg.h"
#include "femhmg.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces the CPython extension chfem.wrapper.run
 * (chfem/src/chfem_wrapper.c:7-89): same arguments (image flattened in file
 * order, .nf path, analysis, direction, solver, precondition, output_fields);
 * the n x n effective matrix (row-major) and any thermal-expansion row go to
 * eff_out (>= 42 doubles).  Returns 0 on success, -1 on failure, and writes
 * n (3 or 6) to *n_out and the thermal-expansion count to *alpha_out. */
int chfem_run(const uint8_t *img, size_t img_size, const char *nf_file, int analysis, int direction,
              int solver, int precondition, int output_fields, double *eff_out, unsigned *n_out,
              unsigned *alpha_out);

/* Arithmetic of the iterative solve.  No reference counterpart (the reference
 * fixes it at compile time with CUDAPCG_VAR_32BIT).
 *   CHFEM_PREC_FP64  : operator, vectors and reductions in FP64 (default).
 *   CHFEM_PREC_MIXED : FP32 operator and inner Krylov vectors, FP64 dot
 *                      products / alpha / beta, FP64 outer residual and
 *                      iterative refinement to the requested tolerance. */
#define CHFEM_PREC_FP64 0
#define CHFEM_PREC_MIXED 1
cudapcgFlag_t cudapcgSetPrecision(cudapcgFlag_t flag);
void hmgSetPrecisionFlag(cudapcgFlag_t flag);

/* Statistics of the last cudapcgSolve: total Krylov iterations, refinement
 * sweeps (mixed precision: FP64 residual replacements), device time in seconds
 * of the solve, final relative residual sqrt(delta / delta0); for a solve
 * given an initial guess, the guess's relative residual sqrt(delta(x0) /
 * delta0) and whether it was kept (1) or replaced by the zero guess because
 * its residual was larger (0). */
typedef struct {
    unsigned int iterations;
    unsigned int sweeps;
    double solve_seconds;
    double residual;
    double x0_residual;
    unsigned int x0_used;
} chfemSolveStats_t;
cudapcgFlag_t cudapcgGetSolveStats(chfemSolveStats_t *stats);

/* Per-direction statistics of the last hmgSolveHomogenization (up to 6). */
cudapcgFlag_t hmgGetSolveStats(chfemSolveStats_t *stats, unsigned int *count);

/* Test hooks on the current solver model: q = A v and z = M^-1 v on host
 * arrays of nvars entries, through the same device kernels the solver uses. */
cudapcgFlag_t cudapcgApplyOperator(const cudapcgVar_t *v, cudapcgVar_t *q);
/* q = A v with the FP32 operator the mixed-precision solver uses (float host
 * arrays of nvars entries). */
cudapcgFlag_t cudapcgApplyOperatorF32(const float *v, float *q);
cudapcgFlag_t cudapcgApplyPreconditioner(const cudapcgVar_t *v, cudapcgVar_t *z);

/* Right-hand side and effective-tensor column computed on the device for the
 * current hmg model: b for load case dir (host array of nvars entries), and
 * the column dir of C from solution x (writes C[row*dim + dir]). */
logical hmgAssembleRHS(unsigned int dir, cudapcgVar_t *b);
logical hmgUpdateC(unsigned int dir, const cudapcgVar_t *x, var *C);

/* Allocation-free variant of hmgUpdateC.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops).  For the model hmgPrepare made resident, it
 * writes column dir of C (3 or 6 doubles, the values hmgUpdateC gives) from
 * the device solution x_dev (ndof doubles) to the device array Ccol_dev, and
 * returns 1 when the column is there.  It calls no allocator: its reduction
 * partials, sums and seam table live in one static device array per GPU
 * (49,600 doubles, 387.5 KiB, held until the process ends; its address is
 * looked up once per device by hmgPrepare).  Limit: at most 64 materials
 * (thermal/elastic); larger models are rejected with a message.  It runs on
 * the library's stream and synchronises it before returning; x_dev must be
 * complete before the call.  Calls that overlap on one GPU race on the
 * shared array. */
logical hmgUpdateCNoAlloc(unsigned int dir, const double *x_dev, double *Ccol_dev);

/* Model maps as built by hmgInit (host copies): NBN keys (uint16 per node),
 * or for Stokes the node fluid keys, node->DOF ids, DOF->node map and border
 * keys.  Any pointer may be NULL. */
logical hmgGetMaps(uint16_t *keys, uint8_t *fluid_keys, uint32_t *dof_id, uint32_t *dof2node,
                   uint8_t *border_keys);
/* Local matrices as hmgInit built them, in the reference layouts
 * (thermal 40 stored entries per material, elastic 576 + CB 144 per material,
 * Stokes 1024 + CB 80).  Either pointer may be NULL. */
logical hmgGetLocalMatrices(double *Mtxs, double *CB);
/* Sizes: nelem, ndof, velocity nodes, border nodes. */
logical hmgGetSizes(unsigned int *nelem, unsigned int *ndof, unsigned int *nvel, unsigned int *nborder);

/* Overrides the solver tolerance / iteration cap read from the .nf file. */
void hmgSetSolverTolerance(double tol);
void hmgSetMaxIterationsOverride(unsigned int maxit);

/* Resident solves: hmgPrepare uploads the current model (maps, local
 * matrices) and allocates the solver once; hmgSolveResident then runs every
 * configured load case (device RHS -> solve -> device C column) against the
 * HBM-resident model and writes C (C_dim entries); hmgRelease frees the
 * solver.  hmgSolveHomogenization is prepare + one resident solve + release. */
logical hmgPrepare(void);
logical hmgSolveResident(var *C);
void hmgRelease(void);
/* One load case against the resident model: device RHS of direction dir,
 * solve from x0 (host, NULL: zero guess), solution copied to x (host, ndof
 * entries, nullable) and the effective-tensor column to Ccol (3 or 6 entries,
 * nullable).  Returns 1 when the solver converged. */
logical hmgSolveLoadCase(unsigned int dir, const var *x0, var *x, var *Ccol);

/* Learned initial guess (no reference counterpart; the reference seeds PCG
 * with zeros, a coarse-grid solution or a guess read from file,
 * chfem/src/femhmg/femhmg.c:1053-1084).  A periodic 3D convolutional
 * encoder-decoder (about 332k parameters, BF16 matrix-core kernels, FP32
 * master weights) maps the material properties of a voxel image to the
 * fluctuation field of every load case (elastic 6 x 3 outputs, thermal 3).
 * Grids must be multiples of 8 in every direction.
 *   chfemGuessCreate(physics 0 thermal / 1 elastic, seed) -> network or NULL
 *   chfemGuessPredict: out[v * nout + k], v = row + col*ny + layer*nx*ny,
 *     k = load * ndof_per_node + component (nout = 18 elastic, 3 thermal)
 *   chfemGuessAccumulate: forward + backward of the mean squared error to
 *     target (same layout), gradients summed into the network; *loss = MSE
 *   chfemGuessStep: Adam step (lr, gradients scaled by grad_scale), then
 *     zeroes the gradients
 *   parameters: FP32, per layer W[Co][taps][Ci] then b[Co]
 *   hmgSetInitialGuessNetwork: subsequent thermal/elastic homogenizations
 *     seed every load case with ramp + prediction (NULL: zero guess).  A
 *     guess whose initial residual exceeds the zero guess's is replaced by
 *     the zero guess (chfemSolveStats_t.x0_used = 0). */
void *chfemGuessCreate(int physics, unsigned int seed);
void chfemGuessDestroy(void *net);
size_t chfemGuessNumParams(void *net);
unsigned int chfemGuessNumOutputs(void *net);
int chfemGuessGetParams(void *net, float *params);
int chfemGuessSetParams(void *net, const float *params);
int chfemGuessGetGradients(void *net, float *grad);
int chfemGuessZeroGradients(void *net);
int chfemGuessPredict(void *net, const uint16_t *elem, const double *props, unsigned int nmat, unsigned int nx,
                      unsigned int ny, unsigned int nz, float *out);
int chfemGuessAccumulate(void *net, const uint16_t *elem, const double *props, unsigned int nmat, unsigned int nx,
                         unsigned int ny, unsigned int nz, const float *target, double *loss);
int chfemGuessStep(void *net, float lr, float grad_scale);
int hmgSetInitialGuessNetwork(void *net);
/* Batched solver (no reference counterpart; the reference solves one model
 * and one load case at a time, chfem/src/femhmg/femhmg.c:848-1198).  Every
 * load case of every sample added is one system; all systems advance
 * together, each kernel of a CG step being one launch over a table of
 * operator tiles and vector chunks of all systems (batch.hip).  FP32 CG with
 * FP64 scalars and per-system FP64 residual replacement, Jacobi from the
 * packed keys, stop rule delta <= tol^2 delta0 per system (FP64).
 *   chfemBatchAddSample: a .nf file (thermal or elastic, isotropic, <= 4
 *     materials; every sample of a batch has the same analysis type) and its
 *     uint8 image in file order; returns 1 on success
 *   chfemBatchSolve: returns 1 when every system converged, 2 when some
 *     did not (iteration cap or FP32 stagnation), 0 on failure
 *   chfemBatchGetC: effective matrix of a sample (row-major n x n)
 *   chfemBatchGetStats: per load case iterations, FP64 replacements and
 *     final relative residual
 *   chfemBatchSeconds: device time of the last solve (RHS to convergence) */
void *chfemBatchCreate(void);
void chfemBatchDestroy(void *batch);
int chfemBatchAddSample(void *batch, const char *nf_file, const uint8_t *img);
int chfemBatchSolve(void *batch, double tol, unsigned int maxit);
int chfemBatchGetC(void *batch, unsigned int sample, double *C);
int chfemBatchGetStats(void *batch, unsigned int sample, chfemSolveStats_t *stats, unsigned int *count);
double chfemBatchSeconds(void *batch);
/* Learned initial guess for every system of a batch (no reference
 * counterpart): each sample's load cases start from ramp + the network's
 * prediction (chfemGuessCreate, same physics as the batch; grids multiples of
 * 8), with the zero guess kept for a system whose guess leaves a larger
 * residual (chfemSolveStats_t x0_used / x0_residual).  NULL: zero guesses.
 * Returns 1 on success. */
int chfemBatchSetInitialGuessNetwork(void *batch, void *net);

/* BF16 matrix-core lane-layout check: D = A * B for 16x16 row-major BF16
 * A and B, FP32 D. */
int chfemGuessMfmaCheck(const uint16_t *A, const uint16_t *B, float *D);
/* FP32 / FP64 16x16x4 matrix-core lane-layout check used by the EBE
 * operator: D = A (16x4) * B (4x16), row-major. */
int chfemMfmaCheckF(const float *A, const float *B, float *D, const double *A2, const double *B2, double *D2);
/* FP32 4x4x1 16-block matrix-core lane-layout check used by the FP32 EBE
 * operator: lane l of one wavefront supplies A[l] and B[l]; its 4 result
 * items land at D[4 l .. 4 l + 3] (64 inputs each, 256 outputs). */
int chfemMfmaCheck4(const float *A, const float *B, float *D);

/* Operator-launch timing inside the Krylov loops: with timing on, every
 * operator launch of the inner solve is bracketed with HIP events on the
 * solver's stream and the durations are summed. */
cudapcgFlag_t cudapcgSetKernelTiming(cudapcgFlag_t on);
cudapcgFlag_t cudapcgGetKernelTiming(double *op_ms_total, unsigned int *op_launches);
/* Walsh-basis operator tile in use for a precision (rows, columns and node
 * layers per block), as fixed by CHFEM_WALSH_TILE or chosen by the
 * autotuner before the first solve. */
cudapcgFlag_t cudapcgGetOperatorTile(cudapcgFlag_t precision, int *tile);
/* Form of the Jacobi-PCG loop on the Walsh-basis operators for a precision:
 * 0 separate kernels, 1 direction and iterate updates inside the operator
 * plus a z update kernel, 2 single-kernel iterations; fixed by
 * CHFEM_FUSE_DIR or chosen by the autotuner (-1 without a solver). */
int cudapcgGetLoopForm(cudapcgFlag_t precision);
/* Same counters for the last hmgSolveResident / hmgSolveHomogenization. */
cudapcgFlag_t hmgGetKernelTiming(double *op_ms_total, unsigned int *op_launches);

/* Operator micro-benchmark on the current solver (after cudapcgAllocateArrays
 * or hmgPrepare): q = A v for reps launches alternating between two input
 * vectors, precision CHFEM_PREC_FP64 or CHFEM_PREC_MIXED (FP32 operator),
 * with the solver's fused v.Av reduction when fused_dot is 1; fused_dot 2
 * times the CG loop's form instead: the direction update d = M^-1 r + beta d
 * inside the operator (isotropic elastic NBN with the Jacobi preconditioner;
 * reads r and the previous direction, writes d and q).  Returns the mean
 * launch time in ms from HIP events on the solver's stream, or a negative
 * value on failure.  fused_dot 3 times the single-kernel CG iteration (the
 * z update z -= alpha M^-1 q inside as well; reduces d.q, z.q, q.M^-1 q). */
double cudapcgOperatorBench(cudapcgFlag_t precision, unsigned int reps, cudapcgFlag_t fused_dot);

/* Calibration kernel for hardware-counter traffic: a grid-stride FP64 copy of
 * n entries, run reps times on device buffers it allocates; returns the mean
 * duration in ms (HIP events), or a negative value on failure. */
double chfemCopyKernelBench(size_t n, unsigned int reps);

#ifdef __cplusplus
}
#endif

#endif /* CHFEM_AMD_EXT_H */
