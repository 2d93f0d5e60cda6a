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

/* Statistics of the last cudapcgSolve: total Krylov iterations (inner
 * iterations summed over refinement sweeps), refinement sweeps, device time in
 * seconds of the solve. */
typedef struct {
    unsigned int iterations;
    unsigned int sweeps;
    double solve_seconds;
    double residual;
} chfemSolveStats_t;
cudapcgFlag_t cudapcgGetSolveStats(chfemSolveStats_t *stats);

/* Per-direction statistics of the last hmgSolveHomogenization (up to 6). */
cudapcgFlag_t hmgGetSolveStats(chfemSolveStats_t *stats, unsigned int *count);

/* Test hooks on the current solver model: q = A v and z = M^-1 v on host
 * arrays of nvars entries, through the same device kernels the solver uses. */
cudapcgFlag_t cudapcgApplyOperator(const cudapcgVar_t *v, cudapcgVar_t *q);
cudapcgFlag_t cudapcgApplyPreconditioner(const cudapcgVar_t *v, cudapcgVar_t *z);

/* Right-hand side and effective-tensor column computed on the device for the
 * current hmg model: b for load case dir (host array of nvars entries), and
 * the column dir of C from solution x (writes C[row*dim + dir]). */
logical hmgAssembleRHS(unsigned int dir, cudapcgVar_t *b);
logical hmgUpdateC(unsigned int dir, const cudapcgVar_t *x, var *C);

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

#ifdef __cplusplus
}
#endif

#endif /* CHFEM_AMD_EXT_H */
