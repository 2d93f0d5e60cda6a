/*
 * femhmg.h -- homogenization-model C ABI of the MI355X chfem library.
 *
 * Same entry points, flag values and semantics as the reference's model layer
 * (chfem/src/femhmg/femhmg.h:6-36, flags chfem/src/femhmg/includes.h:24-48)
 * plus the CLI/Python driver pair runAnalysis/initDefaultInput
 * (chfem/src/main.c:72-152, :300-323; struct chfem/src/includes.h:10-47).
 */
#ifndef CHFEM_AMD_FEMHMG_H
#define CHFEM_AMD_FEMHMG_H

#include <stdint.h>
#include "cudapcg.h"

#ifdef __cplusplus
extern "C" {
#endif

#define HMG_THERMAL 0
#define HMG_ELASTIC 1
#define HMG_FLUID 2

#define HMG_2D 2
#define HMG_3D 3

#define HMG_TRUE 1
#define HMG_FALSE 0

#define HMG_DIR_X 0
#define HMG_DIR_Y 1
#define HMG_DIR_Z 2

#define HOMOGENIZE_X 0
#define HOMOGENIZE_Y 1
#define HOMOGENIZE_Z 2
#define HOMOGENIZE_YZ 3
#define HOMOGENIZE_XZ 4
#define HOMOGENIZE_XY 5
#define HOMOGENIZE_ALL 6
#define HOMOGENIZE_THERMAL_EXPANSION 7

#define HMG_FLOAT32 0
#define HMG_FLOAT64 1

typedef double var;
typedef uint8_t logical;
typedef uint16_t hmgFlag_t;
typedef uint8_t reportFlag_t;

logical hmgInit(char *data_filename, char *elem_filename, char *sdf_filename, uint8_t *data);
logical hmgEnd(void);

void hmgSetXReduceFlag(cudapcgFlag_t flag);
void hmgSetParallelStrategyFlag(cudapcgFlag_t flag);
void hmgSetSolverFlag(cudapcgFlag_t flag);
void hmgSetPreConditionerFlag(cudapcgFlag_t flag);
void hmgSetStoppingCriteria(cudapcgFlag_t flag);
logical hmgSetPoreMappingStrategy(cudapcgFlag_t flag);

void hmgFindInitialGuesses(unsigned int nlevels);
void hmgExportX(logical flag);
void hmgImportX(char *file);

void hmgSetHomogenizationFlag(hmgFlag_t flag);
void hmgSolveHomogenization(void);

void hmgPrintModelData(void);
void hmgPrintConstitutiveMtx(void);

var *hmgGetConstitutiveMtx(void);
unsigned int hmgGetConstitutiveMtxDim(void);
var *hmgGetThermalExpansion(void);
unsigned int hmgGetThermalExpansionDim(void);

logical hmgSaveConstitutiveMtx(const char *filename);

void hmgSaveFields(logical mustExport_flag, logical byElems_flag);

void hmgKeepTrackOfAnalysisReport(reportFlag_t flag);
reportFlag_t hmgPrintReport(const char *filename);

/* CLI / Python driver (chfem/src/includes.h:10-47, chfem/src/main.c:72-152) */
typedef struct _input {
    char *neutral_file;
    char *raw_file;
    char *sdf_bin_file;
    unsigned char exportFields_flag;
    unsigned char fieldsByElem_flag;
    unsigned char save2binary_flag;
    char *binary_file;
    unsigned char writeReport_flag;
    char *report_file;
    unsigned char exportX_flag;
    unsigned char importX_flag;
    char *x0_file;
    unsigned int stopcrit_flag;
    unsigned int poremap_flag;
    unsigned int hmg_direction_flag;
    unsigned int solver_flag;
    unsigned int parallel_flag;
    unsigned int num_of_recursions_initguess;
    unsigned int preconditioner_flag;
    unsigned int xreduce_flag;
    var eff_coeff[42];
    unsigned int num_of_thermal_expansion_coeffs;
} chfemgpuInput_t;

void initDefaultInput(chfemgpuInput_t *user_input);
int runAnalysis(chfemgpuInput_t *user_input);

#ifdef __cplusplus
}
#endif

#endif /* CHFEM_AMD_FEMHMG_H */
