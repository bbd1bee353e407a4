// pt_launch.h -- host entry points of the kernels in pt_kernels.hip.
#pragma once
#include <hip/hip_runtime.h>
#include "pt_plan.h"

namespace ptk {
size_t shared_static_bytes();
size_t plain_scratch_bytes();
size_t trial_image_bytes();
int accum_points_per_group();
hipError_t set_lds_limits(size_t bytes);
hipError_t launch_stage_setup(const DevPlan &P, const double *X, void *img, int ntrials, int pt_mode, int bootstrap,
                              hipStream_t s);
hipError_t launch_stage_lines(const DevPlan &P, const void *img, LineRec *rec, int4 *span, int ntrials, hipStream_t s);
hipError_t launch_stage_accum(const DevPlan &P, const void *img, const LineRec *rec, const int4 *span, double2 *partial,
                              int nsplit, int ntrials, hipStream_t s);
hipError_t launch_stage_finish(const DevPlan &P, const void *img, const double2 *partial, int nsplit, int ntrials,
                               double *chi2, int pt_mode, hipStream_t s);
hipError_t launch_stage_finish_setup(const DevPlan &P, void *img, const double2 *partial, int nsplit, hipStream_t s);
hipError_t launch_pattern(const DevPlan &P, const double *X, double *y, int block, int ch, size_t dyn, hipStream_t s);
hipError_t launch_workload(const DevPlan &P, const double *X, long long *stats, int block, int ch, size_t dyn,
                           hipStream_t s);
hipError_t launch_exchange(const DevPlan &P, hipStream_t s);
hipError_t launch_plain(const DevPlan &P, const double *X, void *scratch, LineRec *rec, int4 *span, double *partial,
                        double *out, int ch, hipStream_t s);
}  // namespace ptk
