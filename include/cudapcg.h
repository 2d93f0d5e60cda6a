/*
 * cudapcg.h -- solver plugin C ABI of the MI355X chfem library.
 *
 * Drop-in for the reference's solver layer: every type, flag value and entry
 * point below has the name, signature and meaning of its counterpart in
 * chfem/src/femhmg/cudapcg/includes.h:11-131 and
 * chfem/src/femhmg/cudapcg/cudapcg.h:19-62.  The implementation behind it is
 * hand-written gfx950 HIP (libchfem.so); no CUDA is involved.
 *
 * Differences a caller can observe:
 *   - device errors are returned as CUDAPCG_FALSE with a message on stderr
 *     instead of the reference's cudaDeviceReset()+exit(0)
 *     (chfem/src/femhmg/cudapcg/error_handling.h:6-14);
 *   - cudapcgEnd() releases this library's device memory but does not reset
 *     the device (other users of the GPU in the process are left intact).
 */
#ifndef CHFEM_AMD_CUDAPCG_H
#define CHFEM_AMD_CUDAPCG_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* flag values: chfem/src/femhmg/cudapcg/includes.h:11-44 */
#define CUDAPCG_TRUE 1
#define CUDAPCG_FALSE 0

#define CUDAPCG_NBN 0
#define CUDAPCG_EBE 1

#define CUDAPCG_POREMAP_IMG 0
#define CUDAPCG_POREMAP_NUM 1

#define CUDAPCG_TOLERANCE 0.00001
#define CUDAPCG_MAX_ITERATIONS 100000

#define CUDAPCG_L2_NORM 0
#define CUDAPCG_INF_NORM 1
#define CUDAPCG_ERROR_NORM 2

#define CUDAPCG_THERMAL_2D 0
#define CUDAPCG_THERMAL_3D 1
#define CUDAPCG_ELASTIC_2D 2
#define CUDAPCG_ELASTIC_3D 3
#define CUDAPCG_FLUID_2D 4
#define CUDAPCG_FLUID_3D 5

#define CUDAPCG_CG_SOLVER 0
#define CUDAPCG_MINRES_SOLVER 1
#define CUDAPCG_CG3_SOLVER 2
#define CUDAPCG_MINRES3_SOLVER 3
#define CUDAPCG_CG2_SOLVER 4
#define CUDAPCG_MINRES2_SOLVER 5

#define CUDAPCG_XREDUCE_NONE 0
#define CUDAPCG_XREDUCE_ONLYDIR 1
#define CUDAPCG_XREDUCE_FULL 2

/* ABI types (includes.h:46-91): the FP64 / 16-bit-key default build */
typedef double cudapcgVar_t;
typedef double cudapcgTol_t;
typedef unsigned short int cudapcgMap_t;
typedef uint8_t cudapcgFlag_t;
typedef unsigned int cudapcgIdMap_t;
typedef uint16_t parametricScalarField_t;

#define MATKEY_BITSTEP_3D 2
#define MATKEY_BITSTEP_RANGE_3D 3

/* model description, field for field (includes.h:93-129).  Array members are
 * device pointers owned by the library. */
typedef struct _pcgmodel {
    char *name;
    cudapcgFlag_t freeAllowed_flag;
    cudapcgFlag_t parStrategy_flag;
    cudapcgFlag_t poremap_flag;
    cudapcgFlag_t parametric_density_field_flag;
    cudapcgFlag_t SBS_flag;
    unsigned int nrows;
    unsigned int ncols;
    unsigned int nlayers;
    unsigned int nelem;
    unsigned int nvars;
    unsigned int nvarspernode;
    unsigned int nkeys;
    unsigned int localmtxdim;
    unsigned int nporenodes;
    unsigned int nbordernodes;
    unsigned int nhmgvars;
    double limits_density_field[2];
    cudapcgMap_t *image;
    parametricScalarField_t *parametric_density_field;
    cudapcgFlag_t *pore_map;
    cudapcgFlag_t *border_pore_map;
    cudapcgIdMap_t *periodic2DOF_map;
    cudapcgIdMap_t *DOF2periodic_map;
} cudapcgModel_t;

/* entry points: chfem/src/femhmg/cudapcg/cudapcg.h:19-62 */
cudapcgFlag_t cudapcgInit(cudapcgFlag_t analysis_flag, cudapcgFlag_t parStrategy_flag);
cudapcgFlag_t cudapcgEnd(void);

cudapcgFlag_t cudapcgSetModelConstructorFcn(cudapcgFlag_t (*fcn)(cudapcgModel_t **, const void *));
cudapcgFlag_t cudapcgBuildModel(const void *data);

cudapcgModel_t *cudapcgNewModel(void);
cudapcgFlag_t cudapcgSetModel(cudapcgModel_t *model);

cudapcgFlag_t cudapcgSetNumTol(cudapcgTol_t t);
cudapcgFlag_t cudapcgSetMaxIterations(unsigned int n);

cudapcgFlag_t cudapcgSetXReduceMode(cudapcgFlag_t flag);
cudapcgFlag_t cudapcgSetXReduceShift(unsigned int shift);
cudapcgFlag_t cudapcgSetXReduceScale(double scl);
cudapcgFlag_t cudapcgSetReduceStabFactor(double scl);

cudapcgFlag_t cudapcgSetSolver(cudapcgFlag_t flag);
cudapcgFlag_t cudapcgSetPreconditioner(cudapcgFlag_t flag);
cudapcgFlag_t cudapcgSetResNorm(cudapcgFlag_t flag);

cudapcgFlag_t cudapcgAllocateArrays(void);
cudapcgFlag_t cudapcgFreeArrays(void);

cudapcgFlag_t cudapcgSetRHS(cudapcgVar_t *RHS);
cudapcgFlag_t cudapcgSetX0(cudapcgVar_t *x0, cudapcgFlag_t mustInterpolate);
cudapcgFlag_t cudapcgSetImage(cudapcgMap_t *img);
cudapcgFlag_t cudapcgSetParametricDensityField(parametricScalarField_t *field, double fmin, double fmax);
cudapcgFlag_t cudapcgSetPoreMap(cudapcgFlag_t *pores);
cudapcgFlag_t cudapcgSetPeriodic2DOFMap(cudapcgIdMap_t *vars);
cudapcgFlag_t cudapcgSetDOF2PeriodicMap(cudapcgIdMap_t *nodes);
cudapcgFlag_t cudapcgSetLclMtxs(cudapcgVar_t *LclMtxs);

cudapcgFlag_t cudapcgSolve(cudapcgVar_t *x);

cudapcgFlag_t cudapcgSetHeaderString(char *header);
cudapcgFlag_t cudapcgPrintSolverReport(void);
cudapcgFlag_t cudapcgPrintSolverReport2(char *dest);
cudapcgFlag_t cudapcgPrintSolverMetrics(void);
cudapcgFlag_t cudapcgPrintSolverMetrics2(char *dest);

unsigned int cudapcgGetNumIterations(void);
unsigned int cudapcgGetMaxNumIterations(void);
cudapcgVar_t cudapcgGetResidual(void);

#ifdef __cplusplus
}
#endif

#endif /* CHFEM_AMD_CUDAPCG_H */
