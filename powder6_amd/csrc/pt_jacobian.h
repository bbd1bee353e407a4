// pt_jacobian.h -- device layout of the FP64 derivative columns and normal equations of
// one Levenberg-Marquardt step (js/rietveld_core.js:1245-1298 Model.derivatives and
// :1953-1983 normal), shared by the host (pt_host.cpp) and the kernels (pt_jacobian.hip).
//
// The columns are linear combinations of "evaluations": patterns of the phases (or of the
// background) at one parameter vector, with one kind of line height.  A finite-difference
// column is (y(p + h e_k) - y(p - h e_k)) / (2 h) over the phases the parameter moves; an
// analytic column (scale, atom, background) is one evaluation with its own heights.  The
// variables' columns are then J = C Y with C = (coefficients of the evaluations) M, where
// M maps variables to model parameters (p = base + M x), and A = J W J', g = J W r.
#pragma once
#include <hip/hip_runtime.h>
#include "pt_plan.h"

namespace ptk {

// an atom in FP64 (the Rietveld structure factor and its gradient)
struct JacAtom {
    double r0[3];
    double basis[3][3];
    double ff_a[5], ff_b[5], ff_c;
    double w0;                 // site multiplicity / number of coset representatives (full list)
    double b_neutron, fp, fpp;
    int free_mask, n_gauss;
};

// line heights of an evaluation
enum { JAC_NORMAL = 0,         // S m Lp PO |F|^2 / area
       JAC_UNIT = 1,           // m Lp PO |F|^2 / area (the scale column)
       JAC_GRAD = 2,           // S m Lp PO / area d|F|^2/d(slot): slot = 5 atom + (0 x, 1 y, 2 z, 3 occ, 4 B) of the phase
       JAC_PO = 3,             // height (PO(r + h) / PO(r) - 1) / h (the preferred-orientation column)
       JAC_BKG = 4,            // the background alone at the evaluation's parameters
       JAC_CHEB = 5,           // Chebyshev polynomial T_slot at every point
       JAC_SPLINE = 6 };       // (spline(ys + h e_slot) - spline(ys)) / h

struct JacEval {
    int pvec;                  // parameter vector (row of JacPlan::pvecs)
    int mask;                  // phases whose lines the evaluation sums (bit q)
    int mode;                  // JAC_*
    int slot;
    double h;                  // step of JAC_PO / JAC_SPLINE
};

// one FP64 line (a Kalpha line of one reflection in one evaluation)
struct JacLine {
    double x0, asym, ped, hgt;
    double ig[2], il[2], cl[2], cg[2];   // [0] left flank (delta < 0), [1] right
    int start, stop;           // points [start, stop) of the window; empty when stop <= start
    int pad[2];
};

struct JacPlan {
    const JacAtom *atoms;      // per phase atom (DevPhase::atom_off order)
    const double *yobs, *w;    // FP64 observations and weights
    const double *pvecs;       // [n_pvec][n_params] parameter vectors
    const JacEval *evals;      // [n_evals]
    int n_evals, n_params;
    JacLine *lines;            // [n_evals][2 nr]
    double *ye;                // [n_evals][np] the evaluations' patterns
    const double *C;           // [nv][n_evals] the variables' columns as combinations
    int nv, pad;
    double *jr;                // [nv + 1][np]: the columns J, then the residual r
    double *ext;               // [(nv + 1)(nv + 2) / 2] packed upper triangle of [J; r] W [J; r]'
};

hipError_t launch_jac_lines(const DevPlan &P, const JacPlan &J, hipStream_t s);
hipError_t launch_jac_patterns(const DevPlan &P, const JacPlan &J, hipStream_t s);
hipError_t launch_jac_normal(const DevPlan &P, const JacPlan &J, hipStream_t s);

}  // namespace ptk
