// solver.h -- internal interface between the model layer (femhmg.cpp) and the
// solver layer (cudapcg.cpp).
#pragma once
#include "kernels.h"
#include "../../include/cudapcg.h"
#include "../../include/chfem_amd.h"

namespace chfem {

#define CHECK_HIP(call)                                                                          \
    do {                                                                                         \
        hipError_t err_ = (call);                                                                \
        if (err_ != hipSuccess) {                                                                \
            fprintf(stderr, "chfem: HIP error %s at %s:%d\n", hipGetErrorString(err_), __FILE__, \
                    __LINE__);                                                                   \
            return CUDAPCG_FALSE;                                                                \
        }                                                                                        \
    } while (0)

// device-resident model + local matrices of the active solver
struct DeviceModel {
    OpDesc op64, op32;          // same geometry, local matrices in FP64 / FP32
    double *K64 = nullptr;      // dense local matrices, FP64
    float *K32 = nullptr;
    size_t ksz = 0;             // dense entries per material
    unsigned nmat = 0;
};

// Accessors used by the model layer.
DeviceModel *active_model();
// the operator descriptors as last built, without rebuilding them (nullptr
// before the first build)
const DeviceModel *built_model();
hipStream_t active_stream();
// device FP64 vector holding the right-hand side of the next solve
double *rhs_buffer();
// device FP64 vector holding the last solution (valid after solve_device)
double *solution_buffer();
// run the configured solver on the RHS already in rhs_buffer(); x0 (device,
// FP64) may be null.  Returns 1 if converged.
int solve_device(const double *x0_dev);
// the values the last two-vector solve (CG2, MINRES2) returned, already
// scaled (count, 0 if the last solve was not one)
int reduced_solution(double *out);
// scalar reduction scratch of the active solver
Reduce make_reduce(int fin, int slot);
double *scalars_host();
int sync_scalars();

}  // namespace chfem
