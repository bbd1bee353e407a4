// pt_plan.h -- device-side layout of a compiled PT plan (shared by host and kernels).
//
// HBM layout (one allocation per array, all read-only during a launch except state):
//   tth[np] f64             2-theta axis (window searches)
//   tff[np] float2          the same as a float-float pair (per-point delta = t - centre)
//   obs[np] float2          {y_obs, w} per point
//   hklm[nr] int4           envelope reflections of every phase, phase blocks contiguous
//   eq_off[nr+1], eq[] short3   Laue-orbit equivalents (March-Dollase average)
//   rep_r[nrep*9] int, rep_t[nrep*3] f64   coset representatives of every phase
//   atoms[natoms] DevAtom   Wyckoff sites, scattering factors, weights
//   base[n_params] f64, term_off[n_params+1], term_var[], term_coeff[] f64
//                           model parameters p = base + M x (CSR by parameter)
//   vlo/vhi/vscale[nv] f64  variable boxes and PT step scales
//   state[chains*stride] f64  PT chain state (js/pt_gpu.js state layout)
#pragma once
#include <stdint.h>

#define PTK_MAX_PHASES 4
#define PTK_MAX_PARAMS 512
#define PTK_MAX_ATOMS 160
#define PTK_MAX_REPS 48
#define PTK_MAX_REPS_TOTAL 64
#define PTK_MAX_VARS 256
#define PTK_MAX_BKG 24       // background coefficients or spline anchors
#define PTK_MAX_BKG_F 48     // a trial's background slots: values, then spline tangents
#define PTK_P_INST0 27       // the dataset's instrument slots (separated profiles)
#define PTK_N_INST 18
#define PTK_PH_NPROF 32      // profile slots of a phase block (from slot 8)
#define PTK_PH_ATOM0 40      // first atom slot of a phase block
#define PTK_MAX_FREE 2048    // free reflections per dataset (the envelope of the free phases)
#define PTK_MAX_FREE_LIVE 120  // free reflections with a window in one trial (the NNLS stage holds them in LDS)
#define PTK_MAX_FREE_WS_LIVE 256  // live free reflections one trial may solve in its slot's global workspace;
                                  // a trial with more is reported unsolved (the caller takes the CPU path)
#define PTK_PROF_LDS 48      // a trial's profile slots of a phase in LDS: the 32 model slots, then
                             // (separated profiles) the family's eta / SL / HL slots from the instrument
#define PTK_S_COUNT_K 12
#define PTK_MATRIX_MAX_REFL 4096
#define PTK_MAX_GROUPS 64          // 512-point groups per pattern (32768 points)
#define PTK_REP_CHUNK_ATOMS 8      // atoms per phase up to which the lines stage may chunk the representative sums

namespace ptk {

struct DevAtom {
    double r0[3];
    double basis[3][3];
    int free_mask;
    int n_gauss;
    float ff_a[5], ff_b[5], ff_c;
    float w0;            // site multiplicity / number of coset representatives
    float b_neutron, fp, fpp;
    float pad;
};

struct DevPhase {
    int system, unique_axis, profile_type;
    int n_reps, rep_off;       // coset representatives [rep_off, rep_off + n_reps)
    int n_atoms, atom_off;     // atoms [atom_off, atom_off + n_atoms)
    int refl_off, refl_count;  // envelope reflections [refl_off, refl_off + refl_count)
    int param_base;            // first model parameter of the phase block
    int fixed_angles;          // 1: the cell angles do not move; trig below is used
    int centro;                // 1: the representatives are one of each (R, t) / (-R, -t) pair
    int n_reps_full;           // representatives before that halving
    double po_dir[3];
    double ang[6];             // alpha, beta, gamma (deg), cos alpha, cos beta, cos gamma
    double ang_sin[3];
};

// A phase's separated (GSAS) profile: instrument function from the inst slots, sample
// broadening from the phase's 26 sample slots (profile_type is then the instrument
// family).  Kept in global memory, apart from the kernel arguments.
struct DevSepPhase {
    int sep;
    int size_model, strain_model;  // PT_SIZE_*, PT_STRAIN_*
    int stephens;                  // 1: the phase's rows of sbasis hold its Stephens basis
    int n_sdir, sdir_off;          // Stephens guard directions [sdir_off, sdir_off + n_sdir) of sguard
    float size_axis[3], strain_axis[3];  // uniaxial axes (Miller indices)
    int free_ph;                   // 1: a Le Bail / WPPF phase (free reflection intensities)
};
// another dataset of a multi-dataset plan, as the first dataset's finish reads it: the
// trial images and chi^2 partials of its own pipeline (js/pt_gpu.js:259-297 sums the
// datasets' chi^2 per chain)
struct DevFollow {
    const void *img;
    const double2 *partial;
    const int *unsolved;       // per trial slot: 1 when the dataset's intensity solve failed (free phases), else null
    int nsplit;
    int pad;
};

struct DevSepPlan {
    DevSepPhase ph[PTK_MAX_PHASES];
    const float *sbasis;           // Stephens basis per phase, [phase][key 15][coefficient 15]
    const double *sguard;          // Stephens guard: per direction the 15 key coefficients
    // free intensities (set when the plan has free phases; the per-slot arrays follow the
    // trial slots)
    const int *free_k;             // [n_free] envelope index of every free reflection
    int4 *fspan;                   // per trial slot: the spans of the free reflections' lines (2 nr)
    float *resid;                  // per trial slot: y_obs - background - Rietveld phases (np)
    double *fx;                    // per trial slot: the last solve's free intensities (its warm start)
    int *unsolved;                 // per trial slot: 1 when the last solve did not converge (the trial costs 3e38)
    unsigned char *nnls_ws;        // per trial slot: the solve's arrays when more intensities are live
    size_t nnls_stride;            //   than LDS holds (set when n_free exceeds that), bytes per slot
};

struct DevPlan {
    // sizes
    int np, nr, nv, n_params, n_phases, n_bkg;
    int chains, stride, steps_per_sweep;
    int uniform, radiation, doublet;
    int max_lines_lds;         // line-record chunk held in LDS
    short any_sep;             // some phase has a separated profile
    short has_stephens;        // some separated phase has Stephens strain
    int n_follow;              // further datasets whose chi^2 the finish adds (first dataset only)
    int has_free;              // some phase has free intensities (the NNLS stage runs)
    int n_free;                // free reflections (all reflections of the free phases)
    int nnls_cap;              // active-set iterations of one solve (0: 3 m + 10 for m live columns)
    int rep_chunks;             // the lines stage sums each reflection's representatives in this many
                               // chunks on separate threads (1: one thread per reflection)
    // dataset
    double t0, step;           // uniform grid: t_i = t0 + i*step
    double inv_step;           // 1 / step
    float step_hi, step_lo;    // float-float split of step
    double tlo, thi;           // Chebyshev range (first and last point)
    double lambda1, lambda2, ratio, pol_k, min_fwhm;
    double qlim;               // envelope cut 1/dmin^2 (1 + 1e-9) at the limits
    double guard_lo, guard_hi; // tan(theta) range of the variance guard
    double max_temp;
    uint32_t seed;
    int bkg_spline;            // 1: background is the spline through the anchors, 0: Chebyshev
    DevPhase phase[PTK_MAX_PHASES];
    // device arrays
    const double *tth;
    const float2 *tff;         // float-float split of tth
    const float *zcheb;        // Chebyshev abscissa of every point, 2 (tth - tlo) / (thi - tlo) - 1
    const int *bseg;           // spline: anchor interval of every point
    const float4 *bcoef;       // spline: Hermite weights of every point (y_s, m_s, y_s+1, m_s+1)
    const double *bknots;      // spline: anchor positions
    const float2 *obs;
    const int4 *hklm;
    const int *eq_off;
    const short *eq;
    const int *rep_r;
    const double *rep_t;
    const DevAtom *atoms;
    const double *base;
    const int *term_off;
    const int *term_var;
    const double *term_coeff;
    const double *vlo, *vhi, *vscale;
    const double *temp;        // temperature of chain c: max_temp^(c / (chains - 1))
    const DevSepPlan *sp;      // separated profiles and free intensities (always set)
    const DevFollow *follow;   // [n_follow]
    double *state;
    double *noeval;            // per chain: proposals rejected without an evaluation (invalid cell,
                               // infeasible move of a feasible chain, or skipped by the screen)
    const struct ScreenFused *sf;   // the surrogate's skip test inside the PT setup (null: off)
};

// The surrogate screen as the PT stages run it when each chain has one proposal per step:
// the setup predicts log Rwp at the proposal and at the chain's state with the network
// (BF16 weights, FP32 sums) and may mark the proposal not evaluated; the finish records
// every evaluated trial into the replay ring.  The network trains between sweeps only.
struct ScreenFused {
    const uint16_t *w;         // BF16 weights in the network's flat layout
    const float *p;            // FP32 parameters (the biases)
    const int *iter;           // Adam steps taken
    const double *center, *inv_half;   // network input u_i = (x_i - center_i) inv_half_i
    uint16_t *ring_x;          // replay ring: cap rows of 64 BF16 inputs
    float *ring_t;             //   and their targets (log Rwp, NaN = infeasible)
    unsigned *skipped;         // count of proposals not evaluated
    double log_wss;            // log sum w y_obs^2
    float skip_margin;         // log Rwp allowance taken off a prediction before a skip
    float skip_log;            // skip when the predicted acceptance < exp(-skip_log); 0 off
    int cold_chains;           // chains [0, cold_chains) are screened and recorded
    int cap, warmup_iters, pad;
};

// a separated-profile phase of one trial, formed once in the setup (FP32)
struct SepPhase {
    float rb[9];               // reciprocal basis rows (Cartesian, 1/A): h . rows = reciprocal vector
    float axs[3], axe[3];      // unit reciprocal vectors of the size and strain axes
    float rot[9];              // ellipsoid: inverse Euler rotation (rows)
    float e[15];               // Stephens: sum_k S_k basis[k][i]
    float inv_l2[3];           // ellipsoid: 1 / max(0.001, L)^2
    float inv_i2, inv_a2;      // uniaxial size: 1 / size^2, 1 / sizeAxial^2 (floored at 0.001)
    float inv_size;            // isotropic size: 1 / max(0.001, size)
    float si, sa;              // strain, strainAxial (>= 0)
    float sm, em;              // sizeMix, strainMix clamped to [0, 1]
    float pad;
};

// line record (one Kalpha line of one reflection), as held in LDS / global memory:
//   a = {centre_hi, centre_lo (float-float centre, deg), amplitude, pedestal}
//   r = {kg, kl, cl, cg} right flank (delta >= 0), l = the same for the left flank, where
//       g = exp2(kg delta^2), l = 1/(1 + kl delta^2), value = cl l + cg g - pedestal
//   span = {start, stop, asymmetry (float bits), 1 when the flanks differ}: points in [start, stop)
struct LineRec {
    float4 a, r, l;            // r first: a line with equal flanks needs only a and r
};

// The staged pipeline's line record (64 bytes, one scalar load): a line record with its
// window, written by the lines stage in the form the accumulate stage sums.  On a uniform
// grid the centre is taken relative to a grid point k near it: a = (k, c - t_k, height,
// pedestal), and a plain line (equal flanks, no asymmetry) gets its profile in scaled
// form, l = (step s, (c - t_k) s, -B, 0) with s = sqrt(-r.x) and B = r.y / s^2, so that
// u = (i - k) step s - (c - t_k) s is the Gaussian's argument and -u^2 serves both terms;
// r.xy = (step q, (c - t_k) q) with q = sqrt(r.y) scale the Lorentzian alone, whose
// denominator is 1 + w^2, w = (i - k) step q - (c - t_k) q; and span.z packs the points
// [span.x + lo, span.y - hi) outside which the Gaussian term is below 2^-40 of the line's
// peak (lo, hi: 16 bits each, rounded so that the range only grows).  Other lines, lines
// on other grids and lines without a window (among them the free-intensity lines, which
// the intensity solve reads) keep the LineRec form, span.z the asymmetry's bits.
struct StagedLine {
    float4 a, r, l;
    int4 span;
};

// staged-pipeline buffers of one member of a group (per chain: a trial image, 2 nr staged
// line records, one partial chi^2 per 512-point group, the chi^2)
struct DevMember {
    void *img;
    StagedLine *rec;
    double2 *partial;
    double *chi2;
    int nsplit[3];             // accumulate workgroups per trial for 1, 2, 4 tiles per wavefront;
                               // nsplit[0] is also the trial's 512-point group count
    int pad;
};

}  // namespace ptk
