// pt_launch.h -- host entry points of the kernels in pt_kernels.hip.
#pragma once
#include <hip/hip_runtime.h>
#include "pt_plan.h"

namespace ptk {
size_t shared_static_bytes();
size_t plain_scratch_bytes();
size_t trial_image_bytes();
// the allocation-free evaluation's fixed device scratch: its address on `device` (current)
#define PTK_NOALLOC_BYTES (32u << 20)
#define PTK_NOALLOC_DEVICES 64
hipError_t noalloc_scratch(int device, unsigned char **ptr, size_t *bytes);
int accum_points_per_group(int tpw);   // points per accumulate workgroup at tpw tiles per wavefront
hipError_t set_lds_limits(size_t bytes);
// Every stage launch covers trials [t_off, t_off + ntrials) (t_off a multiple of 8, so
// that trial t keeps XCD t mod 8).
// pick (PT mode): the move of each chain chosen by the surrogate screen, else the
// setup draws the plain proposal itself
hipError_t launch_stage_setup(const DevPlan &P, const double *X, void *img, int ntrials, int pt_mode, int bootstrap,
                              hipStream_t s, const double2 *pick = nullptr, int t_off = 0);
hipError_t launch_stage_nnls(const DevPlan &P, const void *img, const StagedLine *rec, double2 *partial, int nsplit, int ntrials,
                             hipStream_t s, int t_off = 0);
size_t nnls_shared_bytes(int m);           // dynamic LDS of the free-intensity solve for m reflections
size_t nnls_slot_bytes(int nf);             // global workspace per trial slot for a solve of up to nf reflections
hipError_t nnls_lds_limit(size_t bytes);
hipError_t launch_stage_follow(const DevPlan &P, const double *mstate, int mstride, const void *mimg, void *img, int ntrials,
                               hipStream_t s, int t_off = 0);
hipError_t launch_stage_lines(const DevPlan &P, const void *img, StagedLine *rec, int ntrials, hipStream_t s, int t_off = 0);
// partial chi^2 per 512-point group: ngroups = ceil(np / 512) per trial, whatever tpw
hipError_t launch_stage_accum(const DevPlan &P, const void *img, const StagedLine *rec, double2 *partial, int ngroups,
                              int ntrials, int tpw, hipStream_t s, int t_off = 0);
hipError_t launch_stage_finish(const DevPlan &P, const void *img, const double2 *partial, int nsplit, int ntrials,
                               double *chi2, int pt_mode, hipStream_t s, int t_off = 0);
hipError_t launch_stage_finish_setup(const DevPlan &P, void *img, const double2 *partial, int nsplit, hipStream_t s, int ntrials,
                                     int t_off = 0);
// finish, next setup and lines in one launch (plans with P.nr <= 512, no surrogate skip
// test)
hipError_t launch_stage_finish_lines(const DevPlan &P, void *img, const double2 *partial, int nsplit, StagedLine *rec, int ntrials,
                                     hipStream_t s, int t_off = 0);
hipError_t launch_matrix_columns(const DevPlan &P, const void *img, LineRec *rec, int4 *span, float *factor, hipStream_t s);
// bf16 != 0: the hi/lo pairs as BF16 instead of FP16
hipError_t launch_matrix_fill(const DevPlan &P, const LineRec *rec, const int4 *span, unsigned short *phi, unsigned short *plo,
                              int kpad, int bf16, hipStream_t s);
// per-reflection constants of the intensities (representative phases, f0 w0, s^2/4 with the
// height factor) and the form's cell validity, from the start vector's image
hipError_t launch_matrix_tables(const DevPlan &P, const void *img, const float *factor, double4 *hr, float *f0w, float2 *qk,
                               int *go, int kpad, int nrep_max, hipStream_t s);
// intensities of the n trial vectors X as hi/lo pairs (structure parameters materialised in the kernel)
hipError_t launch_matrix_intensity(const DevPlan &P, const double *X, const double4 *hr, const float *f0w, const float2 *qk,
                                   unsigned short *ihi, unsigned short *ilo, float *iscale, int kpad, int n, int bf16,
                                   hipStream_t s);
// trials per workgroup tile of the matrix product (64 or 128); a pass pads its trials to
// a multiple of it
constexpr int kMatrixTrialTile = 128;
// yout (when not null) receives Y of the first ny trials (ny x np doubles)
hipError_t launch_matrix_chi2(const DevPlan &P, const unsigned short *phi, const unsigned short *plo, const unsigned short *ihi,
                              const unsigned short *ilo, const float *iscale, const int2 *band, const float2 *yw,
                              float *partial, int kpad, int tpad, int row_tiles, double *yout, int ny, int bf16, hipStream_t s);
hipError_t launch_matrix_reduce(const float *partial, int row_tiles, int tpad, int ntrials, const int *go, double *chi2,
                                hipStream_t s);
hipError_t launch_pattern(const DevPlan &P, const double *X, double *y, int block, int ch, size_t dyn, hipStream_t s);
hipError_t launch_workload(const DevPlan &P, const double *X, long long *stats, int block, int ch, size_t dyn,
                           hipStream_t s);
hipError_t launch_exchange(const DevPlan &P, hipStream_t s);
hipError_t launch_group_setup(const DevPlan *plans, const DevMember *mem, const int2 *tab, int n, int bootstrap,
                              hipStream_t s);
hipError_t launch_group_lines(const DevPlan *plans, const DevMember *mem, const int4 *tab, int n, hipStream_t s);
hipError_t launch_group_accum(const DevPlan *plans, const DevMember *mem, const int4 *tab, int n, int tpw, hipStream_t s);
hipError_t launch_group_finish(const DevPlan *plans, const DevMember *mem, const int2 *tab, int nchains, int pt_mode,
                               hipStream_t s);
hipError_t launch_group_finish_setup(const DevPlan *plans, const DevMember *mem, const int2 *tab, int nchains, hipStream_t s);
hipError_t launch_group_exchange(const DevPlan *plans, const int2 *tab, int npairs, hipStream_t s);
hipError_t launch_copy_probe(const void *a, void *b, size_t bytes, hipStream_t s);
hipError_t launch_plain(const DevPlan &P, const double *X, void *scratch, LineRec *rec, int4 *span, double *partial,
                        double *out, int ch, hipStream_t s);
}  // namespace ptk
