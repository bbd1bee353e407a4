// pt_launch.h -- host entry points of the kernels in pt_kernels.hip.
#pragma once
#include <hip/hip_runtime.h>
#include "pt_plan.h"

namespace ptk {
size_t shared_static_bytes();
size_t plain_scratch_bytes();
hipError_t set_lds_limits(size_t bytes);
hipError_t launch_eval_batch(const DevPlan &P, const double *X, double *chi2, int ntrials, int grid, int block, int ch,
                             size_t dyn, hipStream_t s);
hipError_t launch_pattern(const DevPlan &P, const double *X, double *y, int block, int ch, size_t dyn, hipStream_t s);
hipError_t launch_workload(const DevPlan &P, const double *X, long long *stats, int block, int ch, size_t dyn,
                           hipStream_t s);
hipError_t launch_pt_sweep(const DevPlan &P, int steps, int bootstrap, int block, int ch, size_t dyn,
                           unsigned long long counter0, hipStream_t s);
hipError_t launch_exchange(const DevPlan &P, int parity, unsigned long long counter, hipStream_t s);
hipError_t launch_plain(const DevPlan &P, const double *X, void *scratch, LineRec *rec, int4 *span, double *partial,
                        double *out, int ch, hipStream_t s);
}  // namespace ptk
