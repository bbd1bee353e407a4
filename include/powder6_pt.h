/*
 * powder6_pt.h -- C ABI of the MI355X parallel-tempering trial evaluator.
 *
 * This is the drop-in boundary for the WebGPU backend `RVPTGPU = {compile, create, run}`
 * of Powder 6 (js/pt_gpu.js:424).  A host (the Node N-API addon under powder6_amd/node,
 * the Python binding powder6_amd/pt.py, or any C caller) flattens the PT context that
 * RVCore.refineAsync yields (js/rietveld_core.js:2271-2273) into a `pt_problem`, creates
 * a plan, and then drives it with the calls below.  Every call returns 0 on success and a
 * negative status on failure; pt_last_error() then holds a message whose wording follows
 * the reference's thrown errors (js/pt_gpu.js:119,121,373,380,402,415) so that callers can
 * keep matching on /disagrees/, /unavailable/, /did not converge/ and the budget phrases.
 *
 * Ownership: every input array is read during pt_plan_create and copied; outputs are
 * written into caller buffers; device memory belongs to the plan until pt_plan_destroy.
 * Threading: plans are independent (one HIP stream each); a plan is not re-entrant.
 */
#ifndef POWDER6_PT_H
#define POWDER6_PT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PT_ABI_VERSION 3

/* ---- status codes -------------------------------------------------------------- */
#define PT_OK 0
#define PT_ERR_ARG -1         /* invalid argument or problem description          */
#define PT_ERR_BUDGET -2      /* a setup budget is exceeded (pt_gpu.js:121,175,381)*/
#define PT_ERR_DEVICE -3      /* no usable GPU, or a HIP call failed (pt_gpu.js:373,380) */
#define PT_ERR_NUMERIC -4     /* a device result disagrees / is not finite (pt_gpu.js:415) */
#define PT_ERR_STATE -5       /* call out of order (e.g. pt_batch before pt_set_states) */

/* ---- enumerations ------------------------------------------------------------- */
/* crystal systems, js/crystal.js:76-119 (rhombohedral uses hexagonal axes) */
enum { PT_TRICLINIC = 0, PT_MONOCLINIC = 1, PT_ORTHORHOMBIC = 2, PT_TETRAGONAL = 3,
       PT_TRIGONAL = 4, PT_HEXAGONAL = 5, PT_CUBIC = 6 };
/* profile families, js/profile.js:603-718 (type 1 simple, 2 split, 3 TCH); 4 is the
 * reference's 'gsas_cw', a separated profile with a TCH instrument function
 * (js/rietveld_core.js:443-460) */
enum { PT_PROFILE_SIMPLE = 1, PT_PROFILE_SPLIT = 2, PT_PROFILE_TCH = 3, PT_PROFILE_GSAS_CW = 4 };
/* sample broadening models of a separated profile, js/rietveld_core.js:416-421
 * broadeningConfig and js/profile.js:384-422 gsasCwComponents */
enum { PT_SIZE_ISOTROPIC = 0, PT_SIZE_UNIAXIAL = 1, PT_SIZE_ELLIPSOID = 2 };
enum { PT_STRAIN_ISOTROPIC = 0, PT_STRAIN_UNIAXIAL = 1, PT_STRAIN_STEPHENS = 2 };
enum { PT_XRAY = 0, PT_NEUTRON = 1 };
/* background forms, js/rietveld_core.js:622-627 bkgType ('cheb' | 'spline') */
enum { PT_BKG_CHEBYSHEV = 0, PT_BKG_SPLINE = 1 };

/* ---- model-parameter layout ----------------------------------------------------
 * The model parameters of one dataset, the names js/rietveld_core.js reads from its
 * value map (`zero`, `disp`, `transp`, `bkg<k>`, `p1.a`, `p1.U`, `p1.C1.x`, ...), in a
 * fixed order.  Dataset-level parameters come first; each phase then has a block of
 * PT_PH_ATOM0 slots followed by PT_AT_STRIDE slots per atom.  pt_phase_base() gives
 * the first index of a phase block.                                                  */
enum { PT_P_ZERO = 0, PT_P_DISP = 1, PT_P_TRANSP = 2, PT_P_BKG0 = 3, PT_MAX_BKG = 24,
       PT_P_INST0 = 3 + 24, PT_N_INST = 18, PT_P_PHASE0 = 3 + 24 + 18 };
/* instrument slots (relative to PT_P_INST0): the dataset's `inst.<name>` values that a
 * separated profile reads (js/rietveld_core.js:413, 454-459 ALL_INSTRUMENT_PROFILE);
 * a value that is not finite reads as the listed default (:700-701) */
enum { PT_INST_U = 0, PT_INST_V, PT_INST_W, PT_INST_X, PT_INST_Y, PT_INST_Z, PT_INST_SL, PT_INST_HL,
       PT_INST_ETA, PT_INST_U_L, PT_INST_V_L, PT_INST_W_L, PT_INST_X_L, PT_INST_U_R, PT_INST_V_R,
       PT_INST_W_R, PT_INST_X_R, PT_INST_ETA_SPLIT };
enum { PT_PH_A = 0, PT_PH_B = 1, PT_PH_C = 2, PT_PH_ALPHA = 3, PT_PH_BETA = 4, PT_PH_GAMMA = 5,
       PT_PH_SCALE = 6, PT_PH_PO = 7, PT_PH_PROF0 = 8, PT_PH_NPROF = 32, PT_PH_ATOM0 = 8 + 32 };
/* profile slots (relative to PT_PH_PROF0) -- js/rietveld_core.js:446-451 names */
enum { PT_TCH_U = 0, PT_TCH_V, PT_TCH_W, PT_TCH_X, PT_TCH_Y, PT_TCH_SL, PT_TCH_HL };
enum { PT_SIM_GU = 0, PT_SIM_GV, PT_SIM_GW, PT_SIM_GP, PT_SIM_LX, PT_SIM_ETA };
enum { PT_SPL_GU_L = 0, PT_SPL_GV_L, PT_SPL_GW_L, PT_SPL_LX_L,
       PT_SPL_GU_R, PT_SPL_GV_R, PT_SPL_GW_R, PT_SPL_LX_R, PT_SPL_ETA };
/* sample slots of a separated profile, js/rietveld_core.js:414-415 SAMPLE_PROFILE order:
 * size, sizeAxial, sizeB, sizeC, rotA..rotC, strain, strainAxial, sizeMix, strainMix,
 * then the 15 Stephens coefficients S400 S040 S004 S220 S202 S022 S310 S301 S130 S103
 * S031 S013 S211 S121 S112 (js/profile.js:332 GSAS_QUARTIC) */
enum { PT_GS_SIZE = 0, PT_GS_SIZE_AXIAL, PT_GS_SIZE_B, PT_GS_SIZE_C, PT_GS_ROT_A, PT_GS_ROT_B, PT_GS_ROT_C,
       PT_GS_STRAIN, PT_GS_STRAIN_AXIAL, PT_GS_SIZE_MIX, PT_GS_STRAIN_MIX, PT_GS_S0, PT_GS_COUNT = 26 };
/* atom slots: free-coordinate values, occupancy, isotropic B */
enum { PT_AT_X = 0, PT_AT_Y = 1, PT_AT_Z = 2, PT_AT_OCC = 3, PT_AT_B = 4, PT_AT_STRIDE = 5 };

#define PT_MAX_GAUSS 5

/* One atom: its Wyckoff site r = r0 + sum_f u_f basis[f] over the free coordinates f
 * (js/rietveld_core.js:305-378 siteSymmetry/sitePosition), multiplicity, and scattering
 * factor f0(s) = c + sum_i a_i exp(-b_i (s/2)^2) (+f', f'') or a neutron length
 * (js/rietveld_core.js:383-390).                                                      */
typedef struct {
    double r0[3];
    double basis[3][3];   /* basis[f][i]: component i of the direction of free coordinate f */
    int32_t free_mask;    /* bit f set: coordinate f (0=x,1=y,2=z) is free */
    int32_t n_gauss;      /* 0..PT_MAX_GAUSS */
    double mult;          /* site multiplicity, site.mult */
    double ff_a[PT_MAX_GAUSS], ff_b[PT_MAX_GAUSS], ff_c;
    double b_neutron;     /* neutron scattering length (PT_NEUTRON datasets) */
    double fp, fpp;       /* anomalous f', f'' of the dataset for this atom */
} pt_atom;

/* One phase of the dataset (all Rietveld).  `op_r`/`op_t` hold EVERY operator of the
 * setting, centring translations included (js/rietveld_core.js:139-153 expandOps):
 * op_r[9*i + 3*row + col] integer rotation, op_t[3*i + row] translation.               */
typedef struct {
    int32_t system;        /* PT_CUBIC ... */
    int32_t unique_axis;   /* monoclinic unique axis: 0=a, 1=b, 2=c */
    int32_t profile_type;  /* PT_PROFILE_* */
    int32_t n_ops;
    const int32_t *op_r;
    const double *op_t;
    double po_dir[3];      /* March-Dollase direction, phase.poDir */
    int32_t n_atoms;
    const pt_atom *atoms;
    /* separated (GSAS) profile: instrument function from the dataset's inst.* slots,
     * sample broadening from the phase's PT_GS_* slots (js/rietveld_core.js:460
     * usesSeparatedProfile: profile_type PT_PROFILE_GSAS_CW, or separated = 1 with any
     * family).  size_axis / strain_axis: the uniaxial axes (Miller indices; an axis that
     * is not three finite numbers with one non-zero reads as 0 0 1, :416-421). */
    int32_t separated;
    int32_t size_model;    /* PT_SIZE_* */
    int32_t strain_model;  /* PT_STRAIN_* */
    double size_axis[3], strain_axis[3];
    /* 1: a Le Bail or WPPF phase (P.method 'lebail' / 'wppf'): every reflection carries a
     * free intensity >= 0 instead of scale |F|^2 Lp PO, and each trial's chi^2 is the
     * least one over those intensities (js/pt_gpu.js:174-186, 259-297 per-trial NNLS;
     * js/rietveld_core.js:1019-1090).  Atoms, scale and po of the phase are not used. */
    int32_t free_intensities;
} pt_phase;

/* The whole PT problem of one dataset.  Model parameters are an affine function of
 * the refined variables x:  p[k] = param_base[k] + sum_{t: term_param[t]==k}
 * term_coeff[t] * x[term_var[t]]  -- plain variables (coefficient 1, base 0) and the
 * exact affine links of js/atomic_constraints.js:76-155 use the same form.              */
typedef struct {
    /* dataset: js/rietveld_core.js:622-627 (tth ascending, w = 0 excludes a point) */
    int32_t n_points;
    const double *tth, *y_obs, *weight;
    double lambda1, lambda2, ratio, pol_k, min_fwhm;
    int32_t radiation;     /* PT_XRAY / PT_NEUTRON */
    int32_t n_bkg;         /* Chebyshev terms, or spline anchors (bkg_type), 0..PT_MAX_BKG */
    /* phases */
    int32_t n_phases;
    const pt_phase *phases;
    /* model parameters */
    int32_t n_params;
    const double *param_base;
    /* refined variables: start x0, box [lo, hi] (+-inf allowed), PT step scale */
    int32_t n_vars;
    const double *x0, *lo, *hi, *scale;
    int32_t n_terms;
    const int32_t *term_param, *term_var;
    const double *term_coeff;
    /* PT settings (js/rietveld_core.js:2147-2158); steps_per_sweep 0 means n_vars */
    int32_t chains, sweeps, steps_per_sweep;
    double max_temp;
    uint32_t seed;
    /* background form.  PT_BKG_SPLINE: a monotone cubic Hermite spline through n_bkg
     * anchors at bkg_x (ascending 2theta), anchor j at the value of the model parameter
     * in slot PT_P_BKG0 + j (the reference's bga<j>), flat outside the anchors
     * (js/rietveld_core.js:549-612 splineSlopes / splineValues).  bkg_x is unused for
     * PT_BKG_CHEBYSHEV. */
    int32_t bkg_type;
    const double *bkg_x;
    /* instrument function family of every separated phase (the dataset's
     * resolutionType, js/rietveld_core.js:700): PT_PROFILE_*, or 0 for each phase's own
     * profile_type.  PT_PROFILE_GSAS_CW means TCH. */
    int32_t resolution_type;
} pt_problem;

/* Plan summary: the numeric fields of pt_gpu.js compile() (js/pt_gpu.js:226).        */
typedef struct {
    int32_t nv, chains, stride, point_count, reflection_count, line_count;
    int32_t steps_per_sweep, n_params;
    int32_t uniform_grid;   /* 1 when tth is t0 + i*step to 1e-9 step */
    int32_t device_ok;      /* 1 when a gfx950 device is attached to the plan */
} pt_plan_info;

/* State layout per chain, js/pt_gpu.js state buffer (stride 3*nv + 12):
 *   [0,nv) current x, [nv,2nv) proposal, [2nv,3nv) best,
 *   3nv+0 cost, +1 best cost, +2 step counter (-1 = bootstrap), +3 accepted,
 *   +4 proposals, +5 swaps, +6 swap attempts, +7 trial valid, +8 log u,
 *   +9 unsolved flag (1 once an intensity solve of the chain's trial did not converge;
 *   such a trial costs 3e38; the flag stays until pt_set_states), +10 trial cost,
 *   +11 feasible.
 * Snapshots are returned in double precision.                                         */
enum { PT_S_COST = 0, PT_S_BEST = 1, PT_S_STEP = 2, PT_S_ACCEPTED = 3, PT_S_PROPOSALS = 4,
       PT_S_SWAPS = 5, PT_S_ATTEMPTS = 6, PT_S_VALID = 7, PT_S_LOGU = 8, PT_S_UNSOLVED = 9,
       PT_S_TRIAL = 10, PT_S_FEASIBLE = 11, PT_S_COUNT = 12 };

/* Evaluation forms for pt_eval */
enum { PT_FORM_PLAIN = 0,    /* one trial per launch sequence, whole-GPU kernels        */
       PT_FORM_BATCHED = 1,  /* many trials per launch, one workgroup per trial (VALU) */
       PT_FORM_MATRIX = 2,   /* shared profile matrix x intensities on MFMA (atom-only trials),
                                FP16 hi/lo pairs                                           */
       PT_FORM_MATRIX_BF16 = 3 }; /* the same with BF16 hi/lo pairs (precision comparison) */

typedef struct pt_plan pt_plan;

/* Library version and the last error of the calling thread (js/pt_gpu.js error strings). */
int pt_abi_version(void);
const char *pt_last_error(void);

/* First model-parameter index of phase `phase` (layout above). */
int pt_phase_base(const pt_problem *prob, int phase);

/* compile(ctx) + create(ctx): js/pt_gpu.js:20-227 and :371-406.  Enumerates the
 * reflection envelope on the host, packs the problem, allocates the device buffers. */
int pt_plan_create(const pt_problem *prob, pt_plan **out);
void pt_plan_destroy(pt_plan *plan);                      /* close(), js/pt_gpu.js:376 */
/* compile(ctx) alone (js/pt_gpu.js:20-227): the host part of pt_plan_create without a
 * device -- envelope, counts and validation.  Only pt_plan_info_get,
 * pt_plan_reflections and pt_plan_destroy accept such a plan. */
int pt_plan_compile(const pt_problem *prob, pt_plan **out);
int pt_plan_info_get(const pt_plan *plan, pt_plan_info *info);

/* compile(ctx) + create(ctx) for a context with several datasets (js/pt_gpu.js:20-227
 * loops problem.datasets; the per-chain cost is the sum of every dataset's chi^2,
 * :259-297).  Each pt_problem is one dataset in its own model-parameter layout; all
 * share the variables (n_vars, x0, lo, hi, scale identical; a variable a dataset does not
 * use has no term there) and the PT settings of datasets[0].  pt_set_states, pt_batch,
 * pt_costs and pt_eval (PT_FORM_BATCHED) then work on the sum; pt_pattern gives the
 * first dataset's pattern and pt_plan_dataset the plan of dataset d (owned by `plan`)
 * for the others'.  Surrogate screening, the plain and the shared-profile forms take one
 * dataset per plan. */
int pt_plan_create_multi(const pt_problem *datasets, int n_datasets, pt_plan **out);
int pt_plan_dataset(pt_plan *plan, int dataset, pt_plan **sub);

/* Reflection envelope of one phase as enumerated by the plan (h,k,l,m per row):
 * the list js/pt_gpu.js:110-131 builds.  Returns the row count, or <0 on error; writes
 * at most `cap` rows.                                                                */
int pt_plan_reflections(const pt_plan *plan, int phase, int32_t *hklm, int cap);

/* initial(vectors): js/pt_gpu.js:400.  x holds nvec vectors of nv doubles, cycled over
 * the chains. */
int pt_set_states(pt_plan *plan, const double *x, int nvec);

/* costs(vectors): js/pt_gpu.js:401.  chi^2 of every chain at the vectors (cycled),
 * written to chi2[chains].  PT_ERR_NUMERIC with "... intensity solve did not converge"
 * when a free-intensity solve of some chain failed (js/pt_gpu.js:402). */
int pt_costs(pt_plan *plan, const double *x, int nvec, double *chi2);

/* batch(count, bootstrap): js/pt_gpu.js:402.  `count` sweeps (or the bootstrap step),
 * then a snapshot of chains*stride doubles into state_out (may be NULL).  With free
 * intensities, a chain whose unsolved flag (+9) is set makes the call return
 * PT_ERR_NUMERIC "... intensity solve did not converge" after the snapshot is written
 * (the reference's run() throws on it, js/pt_gpu.js:418). */
int pt_batch(pt_plan *plan, int count, int bootstrap, double *state_out);

/* Evaluate ntrials arbitrary variable vectors (ntrials*nv doubles) with one of the
 * evaluation forms; chi2[ntrials] out.  PT_FORM_MATRIX requires that only structure
 * (atom) and scale variables differ between the trials and the plan's start vector. */
int pt_eval(pt_plan *plan, int form, const double *x, int ntrials, double *chi2);

/* No counterpart in the reference.  For callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc): the
 * chi^2 of ntrials vectors that are already in device memory (d_x: ntrials * nv doubles;
 * d_chi2: ntrials doubles, both device pointers), by the kernels of pt_eval's
 * PT_FORM_BATCHED, with every per-trial buffer taken from one fixed device scratch area
 * of 32 MiB per GPU (held until the process ends) instead of the plan's buffers.  It calls
 * no hipMalloc, hipFree or other allocating routine.  Runs on the plan's stream and
 * synchronises it.  PT_ERR_BUDGET when the trials do not fit the scratch (the message
 * states this plan's maximum, about 32 MiB / (10 KB + 128 B x reflections) trials);
 * PT_ERR_STATE for multi-dataset plans and plans with free intensities.  All calls on one
 * GPU share the scratch: calls that overlap in time on one GPU race. */
int pt_eval_noalloc(pt_plan *plan, const double *d_x, int ntrials, double *d_chi2);

/* The normal equations of one Levenberg-Marquardt step at the variables x, the step
 * after PT (js/rietveld_core.js:1953-1983 normal): A = J'WJ (nv x nv, row major),
 * g = J'W(y - ycalc) (nv), *chi2 = sum_{w>0} w (y - ycalc)^2, summed over the plan's
 * datasets; with J not NULL also the derivative columns, J[j * N + i] = d ycalc_i / d x_j
 * over the N points of all datasets in order.  The columns are the reference's
 * (Model.derivatives, js/rietveld_core.js:1245-1298): symmetric finite differences with
 * its fdStep and defaultLimits for cell, zero, displacement, transparency, profile and
 * instrument parameters, the Chebyshev polynomials (a one-sided spline difference for
 * spline anchors), the unit-scale pattern, the March-Dollase difference and the analytic
 * |F|^2 gradients for atoms, through the plan's links.  Everything runs in FP64 on the GPU;
 * the elapsed device time is reported by pt_last_timing.  PT_ERR_STATE for plans with
 * free intensities or separated profiles. */
int pt_normal_equations(pt_plan *plan, const double *x, double *A, double *g, double *chi2, double *J);

/* Pattern of one structure-only vector through the shared profile matrix (PT_FORM_MATRIX
 * path): background of the start vector + P I.  No counterpart in the reference. */
int pt_matrix_pattern(pt_plan *plan, const double *x, double *ycalc);
/* Milliseconds of the intensity and matrix-product kernels of the last PT_FORM_MATRIX
 * evaluation (HIP events).  No counterpart in the reference (measurement). */
int pt_matrix_timing(const pt_plan *plan, double *intensity_ms, double *product_ms);

/* Calculated pattern (background + phases) of one vector, FP32 kernels, n_points out. */
int pt_pattern(pt_plan *plan, const double *x, double *ycalc);

/* Work of one trial at vector x, for the roofline: out[0] lines with a window, out[1]
 * window points summed over the lines (W_tot), out[2] reflections with a Kalpha1 window
 * (N_hkl), out[3] structure-factor terms per reflection S = sum_atoms |coset reps|,
 * out[4] points, out[5] Chebyshev terms.  No counterpart in the reference (measurement). */
int pt_workload(pt_plan *plan, const double *x, double *out);

/* Kernel timing mode (0 off, default): when on, pt_batch launches every kernel directly
 * instead of replaying the captured sweep graph and brackets each launch of the dominant
 * kernel (the banded profile sum) with HIP events.  No counterpart in the reference. */
int pt_set_kernel_timing(pt_plan *plan, int on);

/* Active-set iterations one free-intensity solve may take per trial (0, the default:
 * 3 m + 10 for m live intensities; applies to every dataset of the plan).  A trial whose
 * solve needs more is unsolved: it costs 3e38, its chain's +9 flag is set and pt_costs /
 * pt_batch report "did not converge" -- the reference bounds its solve the same way, at
 * 8 solve / polish passes and a convergence flag (js/pt_gpu.js:398, 255, 363). */
int pt_set_solve_limit(pt_plan *plan, int max_iterations);

/* PT proposals since pt_set_states over all chains: out[0] evaluated exactly, out[1]
 * rejected without an evaluation (invalid cell, an infeasible move of a feasible chain,
 * or skipped by the surrogate screen).  No counterpart in the reference (measurement). */
int pt_trial_counts(pt_plan *plan, double *out);

/* Timing of the last pt_batch / pt_eval call on the plan's stream (HIP events):
 * elapsed milliseconds of the whole call, summed milliseconds of the dominant kernel's
 * timed launches (kernel timing mode, or batched pt_eval), and that launch count. */
int pt_last_timing(const pt_plan *plan, double *total_ms, double *kernel_ms, int *launches);

/* Tile shape of the banded profile sum: 128-point tiles per wavefront (1, 2 or 4; a
 * workgroup of four wavefronts covers 512 * tiles points).  pt_batch tunes it on its
 * first sweep call by timing each shape on the chains' current vectors.  request: 0
 * reports, -1 runs the tuner now, 1 / 2 / 4 fixes that shape (no tuning afterwards).
 * Reports the shape in use and the tuner's three timings (ms, 0 before tuning).
 * No counterpart in the reference (the WGSL path has fixed 64-wide dispatches). */
int pt_tiles(pt_plan *plan, int request, int *tiles_per_wave, double *candidate_ms);

/* ---- surrogate network and screened PT ----------------------------------------------
 * A 64 -> 256 -> 256 -> 256 -> 1 SiLU network (BF16 products with FP32 accumulation on
 * MFMA, FP32 master weights, Adam) that predicts log Rwp of a trial vector.  No
 * counterpart in the reference.  With pt_screen on, every PT step of pt_batch draws
 * `oversample` candidate moves per chain (candidate 0 is exactly the plain proposal) and
 * predicts their log Rwp and that of the chain's current vector; a screened chain (the
 * coldest cold_fraction of the ladder) evaluates the best candidate exactly when its
 * predicted gain exceeds `margin`, else the plain proposal, and runs the usual Metropolis
 * step on the exact chi^2.  Every screened chain's trial (vector, exact log Rwp) goes
 * into a replay ring on which the network trains `train_iters` Adam steps per PT step.
 * Until the network has trained `warmup_iters` steps the pick is the plain proposal, and
 * with the screen off (cfg NULL, the default) pt_batch is plain PT. */
typedef struct {
    int oversample;      /* candidates per chain and step, 2..16                       */
    int train_iters;     /* Adam steps per PT step                                      */
    int batch;           /* minibatch rows, a multiple of 64 in 64..4096                */
    int warmup_iters;    /* Adam steps before the network's picks are used             */
    int capacity;        /* replay ring entries (>= batch and >= chains)                */
    float lr;            /* Adam learning rate                                          */
    uint32_t seed;       /* network initialisation and minibatch sampling               */
    float margin;        /* predicted log Rwp gain over the current vector to take a pick */
    float cold_fraction; /* screened share of the ladder, coldest chains first (0, 1]   */
    float skip_margin;   /* log Rwp allowance taken off a prediction before a skip      */
    float skip_eps;      /* skip the exact evaluation of a screened chain's pick when its
                            predicted Metropolis acceptance is below skip_eps (0: never) */
    int train_every;     /* train (train_iters Adam steps) every train_every PT steps;
                            0 or 1: every step.  With oversample 1 and train_every at
                            least the steps of a sweep, the network trains once per sweep
                            (train_iters steps, before its PT steps) and the skip test runs
                            inside the PT step itself, whose chains keep running as
                            independent parts on their own streams                       */
} pt_screen_config;
int pt_screen(pt_plan *plan, const pt_screen_config *cfg);
/* out[0] Adam steps taken, out[1] ring entries filled, out[2] running mean of the
 * minibatch loss (squared error in log Rwp, masked rows counted as zero), out[3] picks of
 * a candidate other than the plain proposal, out[4] proposals skipped (not evaluated) */
int pt_screen_stats(pt_plan *plan, double *out);
/* log Rwp predicted by the plan's network for n vectors of nv doubles */
int pt_screen_predict(pt_plan *plan, const double *x, int n, float *pred);

/* The network alone: parameters are one flat FP32 vector (W1 [64][256] row major by input,
 * b1, W2 [256][256], b2, W3, b3, w4 [256], b4: pt_mlp_param_count() values).  Inputs are
 * n rows of n_in <= 64 values (zero padded to 64). */
typedef struct pt_mlp pt_mlp;
int pt_mlp_create(uint32_t seed, pt_mlp **out);
void pt_mlp_destroy(pt_mlp *net);
int pt_mlp_param_count(void);
int pt_mlp_get_params(pt_mlp *net, float *params);
int pt_mlp_set_params(pt_mlp *net, const float *params);     /* also resets Adam's state */
int pt_mlp_forward(pt_mlp *net, const float *x, int n, int n_in, float *y);
/* loss = (1/n) sum mask (y - t)^2 (mask NULL: ones) and its gradient (param layout) */
int pt_mlp_gradient(pt_mlp *net, const float *x, const float *t, const float *mask, int n, int n_in, float *loss,
                    float *grad);
/* `iters` Adam steps on minibatches of `batch` rows drawn from the n given pairs; *loss
 * receives the last minibatch loss */
int pt_mlp_train(pt_mlp *net, const float *x, const float *t, int n, int n_in, int iters, int batch, float lr,
                 uint32_t seed, float *loss);

/* HBM calibration: `repeats` launches of a plain streaming copy kernel (k_copy_probe) of
 * `bytes` bytes, after one untimed launch; *gbps = bytes read + written per second.  Used
 * to scale the FETCH_SIZE / WRITE_SIZE counters.  No counterpart in the reference. */
int pt_hbm_copy_probe(size_t bytes, int repeats, double *gbps);

/* ---- groups: many independent problems in the same launches ------------------------
 * The reference fits several problems side by side as independent workers
 * (js/rietveld/rv_commands.js:11-100, up to min(8, cores - 1) at once), each with its own
 * PT run.  A group runs the PT runs of `count` problems -- each with its own chains,
 * temperatures, seed, reflections and points -- through shared launches: every stage of
 * a PT step is one launch over all members, and one sweep of all members is one graph
 * launch.  All members take `steps_per_sweep` steps per sweep (0: the largest member
 * value; a member's own 0 means its nv), then exchange replicas.  Member m's chains
 * occupy state_out[offset_m, offset_m + chains_m * stride_m), offsets from
 * pt_group_member.  No counterpart in the reference. */
typedef struct pt_group pt_group;
int pt_group_create(const pt_problem *problems, int count, int steps_per_sweep, pt_group **out);
void pt_group_destroy(pt_group *group);
int pt_group_member(const pt_group *group, int member, pt_plan_info *info, int64_t *state_offset);
/* initial(vectors) of one member (pt_set_states) */
int pt_group_set_states(pt_group *group, int member, const double *x, int nvec);
/* chi^2 of every member's chains at their current vectors, concatenated member by member
 * (sum of chains doubles); the costs() step of js/pt_gpu.js:401 for all members at once */
int pt_group_costs(pt_group *group, double *chi2);
/* batch(count, bootstrap) for every member (pt_batch); state_out holds the concatenated
 * snapshots (may be NULL) */
int pt_group_batch(pt_group *group, int count, int bootstrap, double *state_out);
/* the group's tile shape and tuner timings (request as for pt_tiles) */
int pt_group_tiles(pt_group *group, int request, int *tiles_per_wave, double *candidate_ms);
/* milliseconds of the last pt_group_batch / pt_group_costs call (HIP events) */
int pt_group_last_timing(const pt_group *group, double *total_ms);

#ifdef __cplusplus
}
#endif

#endif /* POWDER6_PT_H */
