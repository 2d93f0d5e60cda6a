// noalloc.h -- the static device scratch of the allocation-free entry point
// hmgUpdateCNoAlloc (noalloc.hip): one file-scope device array per GPU holds
// the reduction partials, the finished sums and the seam table of a load
// case, so the entry point calls no allocator.
#pragma once
#include <hip/hip_runtime.h>
#include "kernels.h"

namespace chfem {

// materials a seam table may hold (6 rows per material, passed to the device
// by value in one kernel argument of 3 KiB)
constexpr int NOALLOC_MAX_MAT = 64;
// layout of the scratch array, in doubles: block partials of a 6-value
// reduction over at most MAX_RED_GRID blocks, then the 6 finished sums, then
// the seam table
constexpr size_t NOALLOC_PARTIALS = (size_t)6 * MAX_RED_GRID;
constexpr size_t NOALLOC_SUMS_OFF = NOALLOC_PARTIALS;
constexpr size_t NOALLOC_SEAM_OFF = NOALLOC_SUMS_OFF + 64;
constexpr size_t NOALLOC_SCRATCH_DOUBLES = NOALLOC_SEAM_OFF + (size_t)6 * NOALLOC_MAX_MAT;

// Device address of the current GPU's scratch array, looked up once per
// device (hipGetSymbolAddress) and cached; nullptr (with a message) if the
// lookup fails.  The array's host-side name is not a device address and is
// never passed to a kernel or a copy.
double *noalloc_scratch();

// dst[0 .. n) = seam[0 .. n) (n <= 6 * NOALLOC_MAX_MAT), with the table
// passed by value as a kernel argument
void launch_noalloc_seam(double *dst, const double *seam, int n, hipStream_t s);
// out[k] = sums[k] / v (divide) or sums[k] * v, k < n
void launch_noalloc_column(const double *sums, double *out, int n, double v, int divide, hipStream_t s);

}  // namespace chfem
