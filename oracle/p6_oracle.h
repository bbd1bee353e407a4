/*
 * p6_oracle.h -- FP64 CPU restatement of the Powder 6 Rietveld cost and CPU parallel
 * tempering (js/rietveld_core.js, js/profile.js, js/crystal.js, js/constants.js).
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker the HIP path is compared against and
 * the CPU baseline that bench.py times; the product library never links, loads or
 * calls it.  Each function cites the reference file:line it restates.
 */
#ifndef P6_ORACLE_H
#define P6_ORACLE_H

#include <stdint.h>
#include "../include/powder6_pt.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_model or_model;

/* ---- building blocks (pinned by closed-form tests) ---- */
double or_soft_positive(double x, double floor);                 /* constants.js:138-143 */
double or_soft_floor(double x, double f);                        /* profile.js:182-189 */
double or_erfc(double x);                                        /* profile.js:791-797 */
double or_hypot2(double x, double y);                            /* Math.hypot, 2 args */
double or_lp(double tth, double k);                              /* crystal.js:487-497 */
double or_lorentz(double tth);                                   /* crystal.js:465-471 */
double or_polarization(double tth, double k);                    /* crystal.js:474-478 */
double or_tch_fwhm(double gg, double gl);                        /* profile.js:552-565 */
int or_inv_dsq(const double cell[6], double form[6]);            /* crystal.js:326-366 */
double or_cell_volume(const double cell[6]);                     /* crystal.js:245-252 */
double or_peak_shift(double tth, double disp, double transp);    /* profile.js:108-132 */

/* Prepared pseudo-Voigt of one line (profile.js:603-718), flattened.  kind = profile
 * type; p = the profile slots of the phase (PT_TCH_* / PT_SIM_* / PT_SPL_*). */
typedef struct {
    int type;
    double x0, asym, pedestal, max_d, window_fwhm, fwhm, eta;
    double inv_g[2], inv_l[2], cl[2], cg[2];   /* [0] left / centre, [1] right */
    double analytic_area, area;                /* area = analytic + asymmetry correction */
} or_voigt;
int or_prepare_voigt(int type, const double *prof, double min_fwhm, double tth_peak, double x0,
                     or_voigt *out);

/* A separated (GSAS) profile at one reflection (profile.js:384-422): the instrument
 * values (defaults applied), the phase's PT_GS_* sample slots, the broadening models,
 * the cell, the reflection's Miller indices and d, and the Stephens basis. */
typedef struct {
    double inst[PT_N_INST];
    const double *sample;
    int size_model, strain_model;
    double size_axis[3], strain_axis[3];
    double cell[6];
    double h[3], d;
    int n_basis;
    const int32_t *basis_key;
    const double *basis_coeff;
} or_sep;
/* prepareVoigt of a separated profile: `prof` holds the family's eta / SL / HL slots
 * (PT_SIM_ETA, PT_SPL_ETA, PT_TCH_SL/HL), the widths come from sep at wavelength lambda */
int or_prepare_voigt_sep(int type, const double *prof, const or_sep *sep, double lambda, double min_fwhm,
                         double tth_peak, double x0, or_voigt *out);
/* calculateProfileWidths of a separated profile (floored), side 0 left / 1 right */
void or_gsas_widths(int type, const or_sep *sep, double lambda, double min_fwhm, double tth, int side,
                    double *gamma_g, double *gamma_l);
/* min over x >= 0 of x'Hx - 2b'x (H: m x m row major, symmetric positive semidefinite),
 * Lawson-Hanson active set on equilibrated columns; returns outer iterations or -1 */
int or_nnls_normal(int m, const double *H, const double *b, double *x);
int or_stephens_basis(int n_ops, const int32_t *op_r, int32_t *keys, double *coeff);   /* profile.js:334-363 */
double or_axis_angle(const double cell[6], const double n[3], const double axis[3]);  /* profile.js:390 */
double or_stephens_sum(const double h[3], const double *s15, int n_basis, const int32_t *keys,
                       const double *coeff);                                          /* profile.js:375-378 */
double or_eval_voigt(double x, const or_voigt *p);               /* profile.js:733-759 */
double or_asym_area(const or_voigt *p);                          /* profile.js:799-831 */

/* Reflection list (rietveld_core.js:247-276).  cell = a,b,c,alpha,beta,gamma.  box NULL
 * for the metric cut.  Writes up to cap rows of (h,k,l,m); d0 optional.  Returns the
 * total count. */
int or_generate_reflections(const double cell[6], int n_ops, const int32_t *op_r,
                            const double *op_t, double dmin, const int32_t *box,
                            int32_t *hklm, double *d0, int cap);
/* Coset representatives modulo centring (rietveld_core.js:164-179); returns the count
 * and writes the representative indices; *n_cent = number of centring vectors. */
int or_cosets(int n_ops, const int32_t *op_r, const double *op_t, int32_t *rep_index,
              int *n_cent);
/* reflectionDmin (rietveld_core.js:238-245) at the given zero/disp/transp magnitudes */
double or_reflection_dmin(const pt_problem *prob, double zero, double disp, double transp);

/* ---- the model ---- */
or_model *or_create(const pt_problem *prob);
void or_destroy(or_model *m);
const char *or_error(void);
/* model parameters at the variables x: p = base + M x */
void or_params(const or_model *m, const double *x, double *p);
/* chi^2 at variables x: rietveld_core.js:1878-1889 (+inf when invalid) */
double or_cost(or_model *m, const double *x);
/* the spline background (rietveld_core.js:549-612): tangents mm[na] and interval widths
 * h[na - 1] of the anchors (xs, ys); values at the n sorted points tth */
void or_spline_slopes(const double *xs, const double *ys, int na, double *h, double *mm);
void or_spline_values(const double *tth, int n, const double *xs, const double *ys, int na, double *out);
/* calculated pattern (background + phases) at x; returns 0, or -1 for an invalid cell */
int or_pattern(or_model *m, const double *x, double *ycalc);
/* per-reflection diagnostics of phase 0 at x: F^2, height, Kalpha1 centre (count out) */
int or_reflection_table(or_model *m, const double *x, int phase, int32_t *hklm, double *f2,
                        double *height, double *centre, int cap);
/* per-line Kalpha1 / Kalpha2 contributions of phase q at x (present, centre, weight, area;
 * 8 values per reflection) and the Bragg angle of each wavelength (2 per reflection) */
int or_line_table(or_model *m, const double *x, int phase, double *line, double *bragg, int cap);
/* varianceViolation (rietveld_core.js:1489-1496) at x */
double or_violation(or_model *m, const double *x);
/* the profile-feasibility rule of pt_gpu.js propose (js/pt_gpu.js:361): 1 valid */
int or_cell_valid(or_model *m, const double *x);

/* The normal equations of one LM step at x (rietveld_core.js:1953-1983 normal with the
 * columns of Model.derivatives :1245-1298): A[nv*nv] = J'WJ, g[nv] = J'W(y - ycalc),
 * *f = sum_{w>0} w (y - ycalc)^2, and J[nv*n] (may be NULL).  Returns 0, -1 for an
 * invalid cell, -2 with free-intensity phases (their variable projection is not restated). */
int or_normal(or_model *m, const double *x, double *A, double *g, double *f, double *J);

/* CPU PT temper() (rietveld_core.js:2171-2240).  scales[nv]; best_x[nv] out;
 * stats[4] = accepted, proposals, swaps, swap attempts.  Returns best cost. */
double or_temper(or_model *m, const double *x0, const double *scales, int sweeps, int chains,
                 double max_temp, uint32_t seed, double *best_x, double *stats);

/* One independent temper() per thread on `threads` POSIX threads, each on its own
 * model and seed+i; returns proposals per second over all threads (wall clock). */
double or_temper_parallel(const pt_problem *prob, const double *scales, int sweeps, int chains,
                          double max_temp, uint32_t seed, int threads, double *seconds);

#ifdef __cplusplus
}
#endif
#endif
