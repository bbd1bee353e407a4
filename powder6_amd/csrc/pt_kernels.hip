// pt_kernels.hip -- gfx950 kernels of the PT trial evaluator.
//
// One PT step of every chain (js/pt_gpu.js propose/geometry/columns/targets/energy/accept,
// :361-364) is a staged pipeline:
//   k_stage_setup        one workgroup per trial: the proposal, the model parameters, the
//                        cell metric, the atom positions, the variance guard -> trial image
//   k_stage_lines        one thread per (reflection, trial): Bragg angle, widths, prepared
//                        pseudo-Voigt, Lp, |F|^2, heights, the Ka2 companion -> staged line
//                        records (64 bytes each, in the form the sum evaluates)
//   k_stage_accum<T>     the banded profile sum over 128-point tiles (line records by
//                        scalar loads) and the compensated weighted chi^2 of every
//                        512-point group
//   k_stage_nnls         free intensities (Le Bail / WPPF): per-trial normal equations and
//                        a non-negative solve
//   k_stage_finish(_setup)  the ordered chi^2 sum and the FP64 Metropolis step, fused with
//                        the next step's setup
//   k_exchange           replica exchange between neighbouring temperatures (:365)
// k_group_*: the same stages over the members of a group of independent problems;
// k_plain_*: one trial spread over the whole GPU (the plain form); k_matrix_*: the shared
// profile matrix on MFMA; k_pattern / k_workload: one vector's pattern and work counts.
#include "pt_device.h"
#include "pt_launch.h"
#include "pt_screen_fused.h"
#include <algorithm>
#include <mutex>

namespace ptk {

// ---- per-workgroup shared state ----------------------------------------------------
// what a trial's setup forms (parameters, metrics, atoms, profile and background slots)
struct SetupShared {
    double p[PTK_MAX_PARAMS];                 // model parameters of the trial
    double metric[PTK_MAX_PHASES][6];         // reciprocal quadratic forms
    double apos[PTK_MAX_ATOMS][3];            // atom positions (fractional)
    float aocc[PTK_MAX_ATOMS], aB[PTK_MAX_ATOMS];
    float prof[PTK_MAX_PHASES][PTK_PROF_LDS]; // profile slots (FP32)
    float inst[PTK_N_INST];                   // instrument values of separated profiles (defaults applied)
    SepPhase sep[PTK_MAX_PHASES];             // separated profiles: per-trial sample broadening terms
    float bkg[PTK_MAX_BKG_F];                 // Chebyshev coefficients, or spline values [0, 24) and tangents [24, 48)
    float red_s[32], red_c[32];
    int phase_ok[PTK_MAX_PHASES];
    int cell_ok, prof_ok, go, accept;
    int skipped;                              // the surrogate screen marked the proposal not evaluated
    double chi2;
};

// ... and the symmetry and scattering constants the structure factors need
struct Shared : SetupShared {
    int rep_r[PTK_MAX_REPS_TOTAL][9];
    double rep_t[PTK_MAX_REPS_TOTAL][3];
    float aff[PTK_MAX_ATOMS][12];             // atom constants: a1..a5, b1..b5, c, w0
    float afp[PTK_MAX_ATOMS][2];              // f', f'' (or the neutron length in [0])
};

__device__ __forceinline__ int wave_id() { return threadIdx.x >> 6; }
__device__ __forceinline__ int lane_id() { return threadIdx.x & 63; }

// stage the coset representatives and the atoms' scattering constants of every phase in
// LDS (once per launch)
__device__ inline void stage_reps(const DevPlan &P, Shared &sh) {
    int total = 0, na = 0;
    for (int q = 0; q < P.n_phases; ++q) { total += P.phase[q].n_reps; na += P.phase[q].n_atoms; }
    for (int i = threadIdx.x; i < total * 9; i += blockDim.x) sh.rep_r[i / 9][i % 9] = P.rep_r[i];
    for (int i = threadIdx.x; i < total * 3; i += blockDim.x) sh.rep_t[i / 3][i % 3] = P.rep_t[i];
    for (int j = threadIdx.x; j < na; j += blockDim.x) {
        const DevAtom &A = P.atoms[j];
        for (int g = 0; g < 5; ++g) {
            sh.aff[j][g] = g < A.n_gauss ? A.ff_a[g] : 0.0f;
            sh.aff[j][5 + g] = g < A.n_gauss ? A.ff_b[g] : 0.0f;
        }
        sh.aff[j][10] = A.ff_c;
        sh.aff[j][11] = A.w0;
        sh.afp[j][0] = P.radiation == 1 ? A.b_neutron : A.fp;
        sh.afp[j][1] = P.radiation == 1 ? 0.0f : A.fpp;
    }
}

// Spline background of one trial: anchor values (NaN read as 0) into out[0, n) and their
// Fritsch-Carlson tangents into out[24, 24 + n), slopes in FP64 from the FP64 parameters
// (rietveld_core.js:549-561, 597-612).  One thread; n <= 24; out is LDS, so no private array.
__device__ inline double spline_y(const double *pv, int j) { const double v = pv[j]; return isnan(v) ? 0.0 : v; }
__device__ inline void spline_setup(const DevPlan &P, const double *pv, float *out) {
    const int na = P.n_bkg;
    const double *xs = P.bknots;
    float *m = out + PTK_MAX_BKG;
    for (int j = 0; j < na; ++j) { out[j] = (float)spline_y(pv, j); m[j] = 0.0f; }
    if (na < 2) return;
    double dprev = 0.0, hprev = 0.0;
    for (int j = 0; j + 1 < na; ++j) {
        const double h = xs[j + 1] - xs[j], d = (spline_y(pv, j + 1) - spline_y(pv, j)) / h;
        if (j == 0) m[0] = (float)d;
        else m[j] = dprev * d <= 0.0 ? 0.0f : (float)((h * dprev + hprev * d) / (hprev + h));
        dprev = d; hprev = h;
    }
    m[na - 1] = (float)dprev;
    for (int j = 0; j + 1 < na; ++j) {
        const double d = (spline_y(pv, j + 1) - spline_y(pv, j)) / (xs[j + 1] - xs[j]);
        if (d == 0.0) { m[j] = 0.0f; m[j + 1] = 0.0f; continue; }
        const double a = m[j] / d, b = m[j + 1] / d, tau = a * a + b * b;
        if (tau > 9.0) { const double f = 3.0 / sqrt(tau); m[j] = (float)(a * d * f); m[j + 1] = (float)(b * d * f); }
    }
}

// the instrument defaults of rietveld_core.js:413, 454-459 (PT_INST_* order)
__device__ __forceinline__ float inst_default(int i) {
    const float d[PTK_N_INST] = {2.f, -2.f, 5.f, 1.f, 0.f, 0.f, 0.f, 0.f, 0.5f, 2.f, -2.f, 5.f, 1.f, 2.f, -2.f, 5.f, 1.f, 0.5f};
    return d[i];
}
__device__ __forceinline__ double finite_or(double v, double d) { return isfinite(v) ? v : d; }

// The per-trial terms of a separated profile (profile.js:384-422 gsasCwComponents with
// crystal.js:263-275 cellBasis and profile.js:364-369 gsasReciprocalVector): the
// reciprocal basis of the trial cell, the unit reciprocal vectors of the uniaxial axes,
// the ellipsoid's inverse Euler rotation (about c by -rotC, then b by -rotB, then a by
// -rotA) and inverse squared lengths, and the Stephens form e = sum_k S_k basis_k.
// FP64 here, stored FP32.  One thread per phase.
__device__ inline void sep_phase_setup(const DevPlan &P, const DevPhase &ph, const double *pp, int q, SepPhase &o) {
    const DevSepPhase &sd = P.sp->ph[q];
    const double *S = pp + 8;
    double a = pp[0], b = pp[1] > 0.0 ? pp[1] : a, c = pp[2] > 0.0 ? pp[2] : a;
    double al = 90.0, be = 90.0, ga = 90.0;
    switch (ph.system) {
    case 6: b = c = a; break;
    case 3: b = a; break;
    case 4: case 5: b = a; ga = 120.0; break;
    case 2: break;
    case 1: {
        const int ax = ph.unique_axis;
        const double v = finite_or(pp[3 + (ax == 0 ? 0 : ax == 2 ? 2 : 1)], 90.0);
        if (ax == 0) al = v; else if (ax == 2) ga = v; else be = v;
        break;
    }
    default: al = finite_or(pp[3], 90.0); be = finite_or(pp[4], 90.0); ga = finite_or(pp[5], 90.0);
    }
    const double ca = cos(al * kD2R), cb = cos(be * kD2R), cg = cos(ga * kD2R), sg = sin(ga * kD2R);
    const double vf = sqrt(fmax(0.0, 1.0 - ca * ca - cb * cb - cg * cg + 2.0 * ca * cb * cg));
    const double av[3] = {a, 0.0, 0.0}, bv[3] = {b * cg, b * sg, 0.0};
    const double cv[3] = {c * cb, c * (ca - cb * cg) / sg, c * vf / sg};
    double rows[3][3] = {{bv[1] * cv[2] - bv[2] * cv[1], bv[2] * cv[0] - bv[0] * cv[2], bv[0] * cv[1] - bv[1] * cv[0]},
                         {cv[1] * av[2] - cv[2] * av[1], cv[2] * av[0] - cv[0] * av[2], cv[0] * av[1] - cv[1] * av[0]},
                         {av[1] * bv[2] - av[2] * bv[1], av[2] * bv[0] - av[0] * bv[2], av[0] * bv[1] - av[1] * bv[0]}};
    const double V = av[0] * rows[0][0] + av[1] * rows[0][1] + av[2] * rows[0][2];
    for (int r = 0; r < 3; ++r)
        for (int i = 0; i < 3; ++i) { rows[r][i] /= V; o.rb[3 * r + i] = (float)rows[r][i]; }
    for (int w = 0; w < 2; ++w) {
        const float *ax = w ? sd.strain_axis : sd.size_axis;
        double A[3];
        for (int i = 0; i < 3; ++i) A[i] = ax[0] * rows[0][i] + ax[1] * rows[1][i] + ax[2] * rows[2][i];
        double n = sqrt(A[0] * A[0] + A[1] * A[1] + A[2] * A[2]);
        if (n == 0.0) n = 1.0;
        float *out = w ? o.axe : o.axs;
        for (int i = 0; i < 3; ++i) out[i] = (float)(A[i] / n);
    }
    // rotation: the columns are the images of the unit vectors under the three plane
    // rotations in their order
    const double ang[3] = {finite_or(S[6], 0.0), finite_or(S[5], 0.0), finite_or(S[4], 0.0)};   // rotC, rotB, rotA
    const int axis_of[3] = {2, 1, 0};
    for (int col = 0; col < 3; ++col) {
        double v[3] = {col == 0 ? 1.0 : 0.0, col == 1 ? 1.0 : 0.0, col == 2 ? 1.0 : 0.0};
        for (int r = 0; r < 3; ++r) {
            const double t = -ang[r] * kPi / 180.0, cs = cos(t), sn = sin(t);
            const int i = (axis_of[r] + 1) % 3, j = (axis_of[r] + 2) % 3;
            const double x = v[i], y = v[j];
            v[i] = cs * x - sn * y; v[j] = sn * x + cs * y;
        }
        for (int i = 0; i < 3; ++i) o.rot[3 * i + col] = (float)v[i];
    }
    const double size = finite_or(S[0], 1.0), sizeAx = finite_or(S[1], 1.0);
    const double L[3] = {size, finite_or(S[2], 1.0), finite_or(S[3], 1.0)};
    for (int i = 0; i < 3; ++i) { const double m = fmax(0.001, L[i]); o.inv_l2[i] = (float)(1.0 / (m * m)); }
    const double i1 = fmax(0.001, size), a1 = fmax(0.001, sizeAx);
    o.inv_i2 = (float)(1.0 / (i1 * i1)); o.inv_a2 = (float)(1.0 / (a1 * a1));
    o.inv_size = (float)(1.0 / i1);
    o.si = (float)fmax(0.0, finite_or(S[7], 0.0));
    o.sa = (float)fmax(0.0, finite_or(S[8], 0.0));
    o.sm = (float)fmin(1.0, fmax(0.0, finite_or(S[9], 1.0)));
    o.em = (float)fmin(1.0, fmax(0.0, finite_or(S[10], 1.0)));
    for (int i = 0; i < 15; ++i) {
        double e = 0.0;
        if (sd.stephens)
            for (int k = 0; k < 15; ++k) e += finite_or(S[11 + k], 0.0) * (double)P.sp->sbasis[(q * 15 + k) * 15 + i];
        o.e[i] = (float)e;
    }
}

// the sample terms of one reflection (h k l, spacing d) of a separated phase:
// 1/size along it and its microstrain (profile.js:384-411)
__device__ __forceinline__ void sep_reflection(const DevSepPhase &ph, const SepPhase &sp, int h, int k, int l, float d,
                                               float &inv_size, float &strain) {
    const float fh = (float)h, fk = (float)k, fl = (float)l;
    float n[3] = {0.f, 0.f, 0.f};
    if (ph.size_model != 0 || ph.strain_model == 1) {
        float qv[3];
        for (int i = 0; i < 3; ++i) qv[i] = fh * sp.rb[i] + fk * sp.rb[3 + i] + fl * sp.rb[6 + i];
        float nn = sqrtf(qv[0] * qv[0] + qv[1] * qv[1] + qv[2] * qv[2]);
        if (nn == 0.0f) nn = 1.0f;
        for (int i = 0; i < 3; ++i) n[i] = qv[i] / nn;
    }
    inv_size = sp.inv_size;
    if (ph.size_model == 1) {
        float c = n[0] * sp.axs[0] + n[1] * sp.axs[1] + n[2] * sp.axs[2];
        const float c2 = fminf(1.0f, fmaxf(0.0f, c * c));
        inv_size = sqrtf((1.0f - c2) * sp.inv_i2 + c2 * sp.inv_a2);
    } else if (ph.size_model == 2) {
        float acc = 0.0f;
        for (int i = 0; i < 3; ++i) {
            const float v = sp.rot[3 * i] * n[0] + sp.rot[3 * i + 1] * n[1] + sp.rot[3 * i + 2] * n[2];
            acc += v * v * sp.inv_l2[i];
        }
        inv_size = sqrtf(acc);
    }
    strain = sp.si;
    if (ph.strain_model == 1) {
        float c = n[0] * sp.axe[0] + n[1] * sp.axe[1] + n[2] * sp.axe[2];
        const float c2 = fminf(1.0f, fmaxf(0.0f, c * c));
        const float i = sp.si, a = sp.sa;
        strain = c2 <= 1e-14f ? i : c2 >= 1.0f ? a : i * a / fmaxf(1e-30f, sqrtf(i * i * c2 + a * a * (1.0f - c2)));
    } else if (ph.strain_model == 2) {
        const float h2 = fh * fh, k2 = fk * fk, l2 = fl * fl;
        const float m[15] = {h2 * h2, k2 * k2, l2 * l2, h2 * k2, h2 * l2, k2 * l2, h2 * fh * fk, h2 * fh * fl, k2 * fk * fh,
                             l2 * fl * fh, k2 * fk * fl, l2 * fl * fk, h2 * fk * fl, k2 * fh * fl, l2 * fh * fk};
        float sum = 0.0f;
#pragma unroll
        for (int i = 0; i < 15; ++i) sum = fmaf(sp.e[i], m[i], sum);
        strain = d * d * sqrtf(fmaxf(0.0f, sum));
    }
}

// the Stephens guards (rietveld_core.js:1456-1478, js/pt_gpu.js:201-205): the strain
// quartic of every separated Stephens phase non-negative at every probe direction and
// envelope reflection.  Every thread takes some directions; true when this thread saw
// a negative value.
__device__ inline bool stephens_bad(const DevPlan &P, const double *p) {
    bool bad = false;
    for (int q = 0; q < P.n_phases; ++q) {
        const DevSepPhase &ph = P.sp->ph[q];
        if (!ph.sep || !ph.stephens) continue;
        const double *S = p + P.phase[q].param_base + 8 + 11;
        for (int dn = threadIdx.x; dn < ph.n_sdir; dn += blockDim.x) {
            const double *g = P.sp->sguard + (size_t)(ph.sdir_off + dn) * 15;
            double v = 0.0;
            for (int k = 0; k < 15; ++k) v += g[k] * finite_or(S[k], 0.0);
            if (v < 0.0) bad = true;
        }
    }
    return bad;
}

// p = base + M x  (rietveld_core.js:1855-1859 apply, affine links of atomic_constraints.js)
__device__ inline void materialize(const DevPlan &P, const double *x, SetupShared &sh) {
    for (int k = threadIdx.x; k < P.n_params; k += blockDim.x) {
        double v = P.base[k];
        for (int t = P.term_off[k]; t < P.term_off[k + 1]; ++t) v += P.term_coeff[t] * x[P.term_var[t]];
        sh.p[k] = v;
    }
}

// per-trial setup after materialize (+barrier): metrics, atoms, profile slots, guards.
// Ends with a barrier; sets sh.cell_ok and sh.prof_ok.
// SEP = false: a plan without separated profiles (P.any_sep == 0); the separated code is
// compiled out of that instance
template <bool SEP = true>
__device__ inline void trial_setup(const DevPlan &P, SetupShared &sh) {
    int tid = threadIdx.x, nt = blockDim.x;
    if (tid < P.n_phases) {
        const DevPhase &ph = P.phase[tid];
        sh.phase_ok[tid] = phase_metric(ph, sh.p + ph.param_base, sh.metric[tid]) ? 1 : 0;
    }
    if (tid == 63 || (nt < 64 && tid == 0)) sh.prof_ok = profile_valid<SEP>(P, sh.p) ? 1 : 0;
    // profile slots: the families read nine, a separated phase all 32
    if (SEP && P.any_sep) {
        for (int q = 0; q < P.n_phases; ++q)
            for (int s = tid; s < PTK_PH_NPROF; s += nt) sh.prof[q][s] = (float)sh.p[P.phase[q].param_base + 8 + s];
    } else {
        for (int i = tid; i < P.n_phases * 16; i += nt) {
            int q = i >> 4, s = i & 15;
            sh.prof[q][s] = (float)sh.p[P.phase[q].param_base + 8 + s];
        }
    }
    if (SEP && P.any_sep)
        for (int i = tid; i < PTK_N_INST; i += nt) sh.inst[i] = (float)finite_or(sh.p[PTK_P_INST0 + i], inst_default(i));
    for (int q = 0; q < P.n_phases; ++q)
        if (SEP && P.sp->ph[q].sep && tid == 32 + q) {           // every setup block has at least 64 threads
            sep_phase_setup(P, P.phase[q], sh.p + P.phase[q].param_base, q, sh.sep[q]);
            // the family's eta / SL / HL slots, read from the instrument values
            // (rietveld_core.js:699-701), after the model slots
            float *fs = sh.prof[q] + PTK_PH_NPROF;
            for (int i = 0; i < 9; ++i) fs[i] = 0.0f;
            const int type = P.phase[q].profile_type;
            const double *in = sh.p + PTK_P_INST0;
            if (type == 3) { fs[5] = (float)finite_or(in[6], 0.0); fs[6] = (float)finite_or(in[7], 0.0); }
            else if (type == 1) fs[5] = (float)finite_or(in[8], 0.5);
            else fs[8] = (float)finite_or(in[17], 0.5);
        }
    const bool strain_bad = SEP && P.has_stephens ? stephens_bad(P, sh.p) : false;
    if (!P.bkg_spline) {
        for (int i = tid; i < P.n_bkg; i += nt) sh.bkg[i] = (float)sh.p[3 + i];
    } else if (tid == nt - 1) {
        spline_setup(P, sh.p + 3, sh.bkg);
    }
    for (int q = 0; q < P.n_phases; ++q) {
        const DevPhase &ph = P.phase[q];
        for (int j = tid; j < ph.n_atoms; j += nt) {
            int ga = ph.atom_off + j;
            const DevAtom &A = P.atoms[ga];
            const double *ap = sh.p + ph.param_base + PTK_PH_ATOM0 + 5 * j;
            double r[3] = {A.r0[0], A.r0[1], A.r0[2]};      // sitePosition, rietveld_core.js:372-378
            for (int f = 0; f < 3; ++f)
                if (A.free_mask & (1 << f)) {
                    double v = ap[f];
                    r[0] += v * A.basis[f][0]; r[1] += v * A.basis[f][1]; r[2] += v * A.basis[f][2];
                }
            sh.apos[ga][0] = r[0]; sh.apos[ga][1] = r[1]; sh.apos[ga][2] = r[2];
            sh.aocc[ga] = (float)ap[3];
            sh.aB[ga] = (float)ap[4];
        }
    }
    int any_bad = 0;
    if (SEP && P.has_stephens) any_bad = __syncthreads_or(strain_bad ? 1 : 0);   // uniform over the block
    else __syncthreads();
    if (tid == 0) {
        int ok = 1;
        for (int q = 0; q < P.n_phases; ++q) ok &= sh.phase_ok[q];
        sh.cell_ok = ok;
        if (any_bad) sh.prof_ok = 0;
    }
    __syncthreads();
}

// |F|^2 of one reflection (rietveld_core.js:749-821): A^2 + B^2 + C^2 + D^2 over the atoms
// and coset representatives.  Phases h.(R r + t) are formed in FP64 and reduced to turns
// in [0, 1); sine and cosine of the turn come from the hardware (v_sin_f32 / v_cos_f32
// take revolutions).  Loop order: representatives outside, up to four atoms inside, so
// h R and h.t are formed once per representative.  CENTRO: the representatives are one
// of each (R, t) / (-R, -t) pair, so A = 2 sum cos, B = 0 (and C, D alike).
// Chunked representative sums: with few reflections per workgroup and many representatives
// (a cubic phase), each reflection's representatives are split into DevPlan::rep_chunks
// contiguous chunks, every chunk is summed from zero and the chunks are added in order.
// The single-plan lines stage sums the chunks on otherwise idle threads (kSfSlots floats
// per (reflection, chunk) in LDS: cos and sin sums of up to PTK_REP_CHUNK_ATOMS atoms);
// every other caller sums them in the reflection's own thread, with the same arithmetic.
constexpr int kSfSlots = 2 * PTK_REP_CHUNK_ATOMS;

// the representatives [r0, r1) of atoms j0 .. j0 + nj - 1 (nj <= 4): ca[a] += sum cos, sa[a] +=
// sum sin of the turns
template <bool CENTRO>
__device__ __forceinline__ void sf_rep_sums(const Shared &sh, const DevPhase &ph, int h, int k, int l, int j0, int nj, int r0,
                                            int r1, float *ca, float *sa) {
    double px[4], py[4], pz[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        const int ga = ph.atom_off + j0 + (a < nj ? a : 0);
        px[a] = sh.apos[ga][0]; py[a] = sh.apos[ga][1]; pz[a] = sh.apos[ga][2];
    }
    for (int r = r0; r < r1; ++r) {
        const int *R = sh.rep_r[ph.rep_off + r];
        const double *t = sh.rep_t[ph.rep_off + r];
        const double H0 = __mul24(h, R[0]) + __mul24(k, R[3]) + __mul24(l, R[6]);
        const double H1 = __mul24(h, R[1]) + __mul24(k, R[4]) + __mul24(l, R[7]);
        const double H2 = __mul24(h, R[2]) + __mul24(k, R[5]) + __mul24(l, R[8]);
        const double ht = (double)h * t[0] + (double)k * t[1] + (double)l * t[2];
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            if (a < nj) {
                const double phs = H0 * px[a] + H1 * py[a] + H2 * pz[a] + ht;
                const float fr = (float)(phs - floor(phs));
                ca[a] += __builtin_amdgcn_cosf(fr);
                if (!CENTRO) sa[a] += __builtin_amdgcn_sinf(fr);
            }
        }
    }
}

// nch > 1: the representative sums in nch chunks; parts (optional) holds them already
// formed by other threads, else they are formed here
template <bool CENTRO>
__device__ __forceinline__ float structure_f2_impl(const DevPlan &P, const Shared &sh, const DevPhase &ph, int h, int k,
                                                   int l, float s2, const float *parts, int nch) {
    float Ar = 0.f, Br = 0.f, Cr = 0.f, Dr = 0.f;
    const float q4 = 0.25f * s2;
    for (int j0 = 0; j0 < ph.n_atoms; j0 += 4) {
        const int nj = min(4, ph.n_atoms - j0);
        float ca[4] = {0.f, 0.f, 0.f, 0.f}, sa[4] = {0.f, 0.f, 0.f, 0.f};
        if (parts) {
            for (int c = 0; c < nch; ++c)
#pragma unroll
                for (int a = 0; a < 4; ++a)
                    if (a < nj) {
                        ca[a] += parts[c * kSfSlots + 2 * (j0 + a)];
                        if (!CENTRO) sa[a] += parts[c * kSfSlots + 2 * (j0 + a) + 1];
                    }
        } else if (nch > 1) {
            const int rpc = (ph.n_reps + nch - 1) / nch;
            for (int c = 0; c < nch; ++c) {
                const int r0 = min(ph.n_reps, c * rpc), r1 = min(ph.n_reps, r0 + rpc);
                float cc[4] = {0.f, 0.f, 0.f, 0.f}, sc[4] = {0.f, 0.f, 0.f, 0.f};
                sf_rep_sums<CENTRO>(sh, ph, h, k, l, j0, nj, r0, r1, cc, sc);
#pragma unroll
                for (int a = 0; a < 4; ++a) {
                    ca[a] += cc[a];
                    if (!CENTRO) sa[a] += sc[a];
                }
            }
        } else {
            sf_rep_sums<CENTRO>(sh, ph, h, k, l, j0, nj, 0, ph.n_reps, ca, sa);
        }
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            if (a < nj) {
                const int ga = ph.atom_off + j0 + a;
                const float *F = sh.aff[ga];
                float f0;
                if (P.radiation == 1) f0 = sh.afp[ga][0];
                else {
                    f0 = F[10];
                    for (int g = 0; g < 5; ++g) f0 += F[g] * __expf(-F[5 + g] * q4);
                    f0 += sh.afp[ga][0];
                }
                const float dw = __expf(-sh.aB[ga] * q4);
                const float wj = sh.aocc[ga] * f0 * dw * F[11];
                const float wi = (P.radiation == 1 || !(sh.afp[ga][1] != 0.f)) ? 0.f : sh.aocc[ga] * sh.afp[ga][1] * dw * F[11];
                Ar += wj * ca[a]; Cr += wi * ca[a];
                if (!CENTRO) { Br += wj * sa[a]; Dr += wi * sa[a]; }
            }
        }
    }
    if (CENTRO) return 4.0f * (Ar * Ar + Cr * Cr);
    return Ar * Ar + Br * Br + Cr * Cr + Dr * Dr;
}

__device__ __forceinline__ float structure_f2(const DevPlan &P, const Shared &sh, const DevPhase &ph, int h, int k, int l,
                                              float s2, const float *parts = nullptr, int nch = 1) {
    return ph.centro ? structure_f2_impl<true>(P, sh, ph, h, k, l, s2, parts, nch)
                     : structure_f2_impl<false>(P, sh, ph, h, k, l, s2, parts, nch);
}

// March-Dollase factor averaged over the equivalents (rietveld_core.js:824-843)
__device__ inline float march_dollase(const DevPlan &P, const DevPhase &ph, const double *f, double r, int kref) {
    if (!(r > 0.0) || fabs(r - 1.0) < 1e-12) return 1.0f;
    const double *H = ph.po_dir;
#define PTK_G(a0, a1, a2, b0, b1, b2) ((a0) * (b0) * f[0] + (a1) * (b1) * f[1] + (a2) * (b2) * f[2] + \
    ((a0) * (b1) + (a1) * (b0)) * f[3] * 0.5 + ((a1) * (b2) + (a2) * (b1)) * f[4] * 0.5 + ((a0) * (b2) + (a2) * (b0)) * f[5] * 0.5)
    double HH = PTK_G(H[0], H[1], H[2], H[0], H[1], H[2]);
    if (!(HH > 0.0)) return 1.0f;
    float r2 = (float)(r * r), ir = (float)(1.0 / r);
    int e0 = P.eq_off[kref], e1 = P.eq_off[kref + 1];
    float sum = 0.f;
    for (int e = e0; e < e1; ++e) {
        double a0 = P.eq[3 * e], a1 = P.eq[3 * e + 1], a2 = P.eq[3 * e + 2];
        double hh = PTK_G(a0, a1, a2, a0, a1, a2);
        double c = PTK_G(a0, a1, a2, H[0], H[1], H[2]) / sqrt(hh * HH);
        float c2 = (float)fmin(1.0, c * c);
        sum += powf(r2 * c2 + (1.0f - c2) * ir, -1.5f);
    }
#undef PTK_G
    return sum / (float)(e1 - e0);
}

// write one line record: window [start, stop) of the axis, float-float centre, flanks
__device__ inline void emit_line(const DevPlan &P, const Prep &pk, double x0, float amp, LineRec &rec, int4 &span) {
    span = make_int4(0, 0, 0, 0);
    double lo = x0 - (double)pk.max_d, hi = x0 + (double)pk.max_d;
    int start = lower_bound_t(P, lo);
    if (start >= P.np) return;
    int stop = upper_stop(P, start, hi);
    if (stop <= start) return;
    float xh = (float)x0;
    rec.a = make_float4(xh, (float)(x0 - (double)xh), amp, pk.pedestal);
    const float kG = -kCg * kLog2e;
    rec.l = make_float4(kG * pk.ig[0] * pk.ig[0], 4.0f * pk.il[0] * pk.il[0], pk.cl[0], pk.cg[0]);
    rec.r = make_float4(kG * pk.ig[1] * pk.ig[1], 4.0f * pk.il[1] * pk.il[1], pk.cl[1], pk.cg[1]);
    const int split = (pk.ig[0] != pk.ig[1] || pk.il[0] != pk.il[1] || pk.cl[0] != pk.cl[1] || pk.cg[0] != pk.cg[1]) ? 1 : 0;
    span = make_int4(start, stop, __float_as_int(pk.asym), split | (pk.asym != 0.0f ? 2 : 0));
}

// Both lines (Kalpha1, Kalpha2) of envelope reflection k: Model.geometry (:719-736),
// buildPeakContributions (profile.js:1013-1080), evalPhase heights (:855-897).
// With unit_factor != null the lines are written at unit height (the Kalpha2 line at its
// doublet weight) and the reflection's height factor m Lp PO / area goes to
// unit_factor[k] instead: the columns of the shared profile matrix.
// fspan != null: the trial's span slots for free-intensity lines.  A reflection of a free
// phase (Le Bail / WPPF) is written at unit height (its Kalpha2 line at the doublet
// weight), its spans go to fspan and its own spans stay empty, so the accumulate stage
// leaves it to the intensity solve.
template <bool SEP = true>
__device__ __forceinline__ void build_reflection(const DevPlan &P, const Shared &sh, int k, LineRec *rec, int4 *span,
                                        float *unit_factor = nullptr, const int4 *hk_given = nullptr, int4 *fspan = nullptr,
                                        const float *sf_parts = nullptr) {
    if (unit_factor) unit_factor[k] = 0.0f;
    span[0] = make_int4(0, 0, 0, 0);
    span[1] = make_int4(0, 0, 0, 0);
    if (fspan) { fspan[0] = make_int4(0, 0, 0, 0); fspan[1] = make_int4(0, 0, 0, 0); }
    int q = 0;
    while (q + 1 < P.n_phases && k >= P.phase[q + 1].refl_off) ++q;
    const DevPhase &ph = P.phase[q];
    if (!sh.phase_ok[q]) return;
    int4 hk = hk_given ? *hk_given : P.hklm[k];
    const double *f = sh.metric[q];
    double h = hk.x, kk = hk.y, l = hk.z;
    double qv = h * h * f[0] + kk * kk * f[1] + l * l * f[2] + h * kk * f[3] + kk * l * f[4] + h * l * f[5];
    if (!(qv > 0.0) || qv > P.qlim) return;
    double s = sqrt(qv), sinT = P.lambda1 * s * 0.5;
    if (!(sinT < 1.0)) return;
    double bragg = 2.0 * asin(sinT) * kR2D;
    float disp = (float)sh.p[1], transp = (float)sh.p[2];
    double zero = sh.p[0];
    double tth = bragg + (double)peak_shift(bragg, disp, transp);
    if (!(tth > 0.0 && tth < 180.0)) return;                 // hasRenderablePosition
    double base1 = tth + zero;
    Prep pk1;
    // separated profile: the sample terms of this reflection, and the family's eta / SL /
    // HL slots filled from the instrument values (rietveld_core.js:694-708)
    SepLine sl{nullptr, 0.f, 0.f, 0.f, 0.f, 0.f};
    const float *prof = sh.prof[q];
    if (SEP && P.sp->ph[q].sep) {
        sep_reflection(P.sp->ph[q], sh.sep[q], hk.x, hk.y, hk.z, (float)(1.0 / s), sl.inv_size, sl.strain);
        sl.inst = sh.inst; sl.sm = sh.sep[q].sm; sl.em = sh.sep[q].em; sl.lam = (float)P.lambda1;
        prof = sh.prof[q] + PTK_PH_NPROF;
    }
    if (SEP) prepare_line(ph.profile_type, prof, (float)P.min_fwhm, (float)base1, pk1, sl);
    else prepare_line(ph.profile_type, prof, (float)P.min_fwhm, (float)base1, pk1);
    if (!(pk1.area > 0.0f)) return;
    float lp1 = lp_factor((float)base1, (float)P.pol_k);
    const double *pp = sh.p + ph.param_base;
    float po = march_dollase(P, ph, f, pp[7], k);
    float height;
    const bool free_ph = SEP && fspan && P.has_free && P.sp->ph[q].free_ph;
    if (free_ph) {
        height = 1.0f;
    } else if (unit_factor) {
        unit_factor[k] = (float)hk.w * lp1 * po / pk1.area;
        height = 1.0f;
    } else {
        float f2 = structure_f2(P, sh, ph, hk.x, hk.y, hk.z, (float)qv, sf_parts, P.rep_chunks);
        double S = pp[6];
        if (!(S != 0.0) || isnan(S)) S = 0.0;
        height = (float)S * ((float)hk.w * lp1 * po / pk1.area * f2);
    }
    emit_line(P, pk1, base1, height, rec[0], span[0]);
    if (P.doublet) {
        double st2 = sinT * (P.lambda2 / P.lambda1);   // sin(theta1) lambda2/lambda1
        if (fabs(st2) < 1.0) {
            double b2 = 2.0 * asin(st2) * kR2D;
            double base2 = b2 + (double)peak_shift(b2, disp, transp) + zero;
            Prep pk2;
            sl.lam = (float)P.lambda2;
            if (SEP) prepare_line(ph.profile_type, prof, (float)P.min_fwhm, (float)base2, pk2, sl);
            else prepare_line(ph.profile_type, prof, (float)P.min_fwhm, (float)base2, pk2);
            float w2 = (float)P.ratio * pk1.area / pk2.area * (lp_factor((float)base2, (float)P.pol_k) / lp1);
            emit_line(P, pk2, base2, height * w2, rec[1], span[1]);
        }
    }
    if (free_ph) {
        fspan[0] = span[0]; fspan[1] = span[1];
        span[0] = make_int4(0, 0, 0, 0); span[1] = make_int4(0, 0, 0, 0);
    }
}

// value of one line at one point (delta = t - centre), profile.js:733-759 inside the window
__device__ __forceinline__ float line_value(const LineRec &L, float asym, float delta) {
    if (asym != 0.0f && fabsf(delta) >= 1e-9f) {
        float t = fmaxf_js(-0.95f, fminf_js(0.95f, asym * delta));
        delta = delta / (1.0f - t);
    }
    float4 fl = delta < 0.0f ? L.l : L.r;
    float d2 = delta * delta;
    float g = __builtin_amdgcn_exp2f(fl.x * d2);
    float lo = __builtin_amdgcn_rcpf(fmaf(fl.y, d2, 1.0f));
    float v = fmaf(fl.z, lo, fmaf(fl.w, g, -L.a.w));
    return v > 0.0f ? v : 0.0f;
}

// Banded profile sum of `nlines` LDS line records into y[] (LDS or global) for the 64-point
// tiles tile0, tile0 + tstep, ...  One wavefront per tile; every lane owns one point and
// walks the lines whose window meets the tile (ballot over the line spans), in line
// order, so the sum is deterministic.
__device__ inline void accumulate_tiles(const DevPlan &P, const LineRec *rec, const int4 *span, int nlines, float *y,
                                        int tile0, int tstep) {
    const int lane = lane_id();
    const int ntiles = (P.np + 63) >> 6;
    for (int tile = tile0; tile < ntiles; tile += tstep) {
        const int t_lo = tile << 6, t_hi = t_lo + 64;
        const int i = t_lo + lane;
        const bool live = i < P.np;
        float2 ti = live ? P.tff[i] : make_float2(0.f, 0.f);
        float acc = 0.0f;
        for (int lb = 0; lb < nlines; lb += 64) {
            int L = lb + lane;
            bool ov = false;
            if (L < nlines) {
                int4 sp = span[L];
                ov = sp.x < t_hi && sp.y > t_lo;
            }
            unsigned long long mask = __ballot(ov);
            while (mask) {
                int b = __ffsll((long long)mask) - 1;
                mask &= mask - 1ull;
                int line = lb + b;
                int4 sp = span[line];
                const LineRec &R = rec[line];
                if (i >= sp.x && i < sp.y) {
                    float delta = (ti.x - R.a.x) + (ti.y - R.a.y);
                    acc = fmaf(R.a.z, line_value(R, __int_as_float(sp.z), delta), acc);
                }
            }
        }
        if (live) y[i] += acc;
    }
}

// Neumaier two-sum accumulate
__device__ __forceinline__ void neumaier(float &s, float &c, float v) {
    float t = s + v;
    if (fabsf(s) >= fabsf(v)) c += (s - t) + v; else c += (v - t) + s;
    s = t;
}

// Background at point i from a trial's coefficients: the Chebyshev sum over the fitted range
// (rietveld_core.js:647-653, 920-931) as an FP32 recurrence, or the spline's cubic Hermite
// form on the point's anchor interval with weights computed once per plan
// (rietveld_core.js:597-612)
__device__ __forceinline__ float bkg_value(const DevPlan &P, const float *bkg, int i) {
    if (P.n_bkg <= 0) return 0.0f;
    if (P.bkg_spline) {
        const int s = P.bseg[i];
        const float4 c = P.bcoef[i];
        const float y1 = P.n_bkg > 1 ? bkg[s + 1] : 0.0f, m1 = P.n_bkg > 1 ? bkg[PTK_MAX_BKG + s + 1] : 0.0f;
        return fmaf(c.x, bkg[s], fmaf(c.y, bkg[PTK_MAX_BKG + s], fmaf(c.z, y1, c.w * m1)));
    }
    const float z = P.zcheb[i];
    float a = 1.0f, b = z, y = bkg[0];
    if (P.n_bkg > 1) y = fmaf(bkg[1], z, y);
    for (int k = 2; k < P.n_bkg; ++k) {
        float nx = 2.0f * z * b - a;
        a = b; b = nx;
        y = fmaf(bkg[k], nx, y);
    }
    return y;
}

__device__ __forceinline__ float background_at(const DevPlan &P, const Shared &sh, int i) { return bkg_value(P, sh.bkg, i); }

// chi^2 = sum_{w>0} w (y_obs - y_calc)^2 (rietveld_core.js:1885) over points [i0, i1)
// with stride, Neumaier per lane, two-sum tree over the wavefront and the block.
// Result valid in thread 0 (and written to sh.chi2).  ypk indexed by point.
__device__ inline double chi2_block(const DevPlan &P, Shared &sh, const float *ypk, int i0, int i1) {
    float s = 0.f, c = 0.f;
    for (int i = i0 + (int)threadIdx.x; i < i1; i += blockDim.x) {
        float2 ow = P.obs[i];
        if (!(ow.y > 0.0f)) continue;
        float yc = background_at(P, sh, i) + ypk[i];
        float r = ow.x - yc;
        neumaier(s, c, ow.y * r * r);
    }
    // wavefront: combine (s, c) pairs with two-sum
    for (int off = 32; off > 0; off >>= 1) {
        float so = __shfl_xor(s, off), co = __shfl_xor(c, off);
        float t = s + so;
        float e = fabsf(s) >= fabsf(so) ? (s - t) + so : (so - t) + s;
        s = t; c = c + co + e;
    }
    int w = wave_id(), nw = (blockDim.x + 63) >> 6;
    if (lane_id() == 0) { sh.red_s[w] = s; sh.red_c[w] = c; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double tot = 0.0;
        for (int k = 0; k < nw; ++k) tot += (double)sh.red_s[k] + (double)sh.red_c[k];
        sh.chi2 = tot;
    }
    __syncthreads();
    return sh.chi2;
}

// Full evaluation of the trial held in sh (after trial_setup): lines in chunks of `ch`
// records, banded sum into ylds[np], then chi^2.  Returns chi^2 (all threads).
__device__ inline double eval_trial(const DevPlan &P, Shared &sh, float *ylds, LineRec *rec, int4 *span, int ch) {
    for (int i = threadIdx.x; i < P.np; i += blockDim.x) ylds[i] = 0.0f;
    const int nwaves = (blockDim.x + 63) >> 6;
    const int per = ch >> 1;   // reflections per chunk
    for (int k0 = 0; k0 < P.nr; k0 += per) {
        int kn = min(per, P.nr - k0);
        __syncthreads();
        for (int kk = threadIdx.x; kk < kn; kk += blockDim.x) build_reflection(P, sh, k0 + kk, rec + 2 * kk, span + 2 * kk);
        __syncthreads();
        accumulate_tiles(P, rec, span, 2 * kn, ylds, wave_id(), nwaves);
    }
    __syncthreads();
    return chi2_block(P, sh, ylds, 0, P.np);
}

__device__ __forceinline__ void dyn_layout(int np, int ch, float *&ylds, LineRec *&rec, int4 *&span) {
    extern __shared__ __align__(16) unsigned char dyn[];
    rec = reinterpret_cast<LineRec *>(dyn);
    span = reinterpret_cast<int4 *>(dyn + sizeof(LineRec) * ch);
    ylds = reinterpret_cast<float *>(dyn + (sizeof(LineRec) + sizeof(int4)) * ch);
}

// ---- calculated pattern of one vector -----------------------------------------------
__global__ void __launch_bounds__(512) k_pattern(DevPlan P, const double *__restrict__ X, double *__restrict__ yout,
                                                 int ch) {
    __shared__ Shared sh;
    float *ylds; LineRec *rec; int4 *span;
    dyn_layout(P.np, ch, ylds, rec, span);
    stage_reps(P, sh);
    __syncthreads();
    materialize(P, X, sh);
    __syncthreads();
    trial_setup(P, sh);
    if (!sh.cell_ok) {
        for (int i = threadIdx.x; i < P.np; i += blockDim.x) yout[i] = NAN;
        return;
    }
    eval_trial(P, sh, ylds, rec, span, ch);
    for (int i = threadIdx.x; i < P.np; i += blockDim.x) yout[i] = (double)(background_at(P, sh, i) + ylds[i]);
}

// ---- workload of one vector: active lines, window points, active reflections --------
// stats[0] = lines with a window, stats[1] = sum of window lengths (points),
// stats[2] = reflections with a Kalpha1 window
__global__ void __launch_bounds__(512) k_workload(DevPlan P, const double *__restrict__ X, long long *stats, int ch) {
    __shared__ Shared sh;
    __shared__ unsigned long long cnt[3];
    float *ylds; LineRec *rec; int4 *span;
    dyn_layout(P.np, ch, ylds, rec, span);
    stage_reps(P, sh);
    if (threadIdx.x < 3) cnt[threadIdx.x] = 0ull;
    __syncthreads();
    materialize(P, X, sh);
    __syncthreads();
    trial_setup(P, sh);
    if (sh.cell_ok) {
        const int per = ch >> 1;
        for (int k0 = 0; k0 < P.nr; k0 += per) {
            int kn = min(per, P.nr - k0);
            __syncthreads();
            for (int kk = threadIdx.x; kk < kn; kk += blockDim.x) {
                build_reflection(P, sh, k0 + kk, rec + 2 * kk, span + 2 * kk);
                for (int s = 0; s < 2; ++s) {
                    int4 sp = span[2 * kk + s];
                    if (sp.y > sp.x) {
                        atomicAdd(&cnt[0], 1ull);
                        atomicAdd(&cnt[1], (unsigned long long)(sp.y - sp.x));
                        if (s == 0) atomicAdd(&cnt[2], 1ull);
                    }
                }
            }
        }
    }
    __syncthreads();
    if (threadIdx.x < 3) stats[threadIdx.x] = (long long)cnt[threadIdx.x];
}

// ---- staged pipeline: setup -> lines -> accumulate -> finish -------------------------
// One PT step (or one batch of independent trial vectors) runs as four launches over the
// whole GPU: per-trial setup (and the proposal), per-reflection line records written to
// HBM (they stay in the 256 MB Infinity Cache between launches), the banded profile sum
// split over several workgroups per trial, and the ordered reduction with the FP64
// Metropolis decision.  Every kernel is small in registers, so each CU holds many waves.

// per-trial image of the setup: what the line and accumulate kernels need
struct TrialImage {
    double p[PTK_MAX_PARAMS];
    double metric[PTK_MAX_PHASES][6];
    double apos[PTK_MAX_ATOMS][3];
    float aocc[PTK_MAX_ATOMS], aB[PTK_MAX_ATOMS];
    float prof[PTK_MAX_PHASES][PTK_PROF_LDS];
    float inst[PTK_N_INST];
    SepPhase sep[PTK_MAX_PHASES];
    float bkg[PTK_MAX_BKG_F];
    int phase_ok[PTK_MAX_PHASES];
    int cell_ok, prof_ok, go, boot, skipped;
    // per 512-point group: the first and last line index whose window meets the group,
    // reset by the setup and gathered by the lines stage (atomic min / max)
    int grange[PTK_MAX_GROUPS][2];
};

__device__ inline int total_atoms(const DevPlan &P) {
    const DevPhase &l = P.phase[P.n_phases - 1];
    return l.atom_off + l.n_atoms;
}

template <bool SEP = true>
__device__ inline void image_store(const DevPlan &P, const SetupShared &sh, TrialImage *img) {
    const int tid = threadIdx.x, nt = blockDim.x, na = total_atoms(P);
    for (int i = tid; i < P.n_params; i += nt) img->p[i] = sh.p[i];
    for (int i = tid; i < P.n_phases * 6; i += nt) img->metric[i / 6][i % 6] = sh.metric[i / 6][i % 6];
    for (int i = tid; i < na * 3; i += nt) img->apos[i / 3][i % 3] = sh.apos[i / 3][i % 3];
    for (int i = tid; i < na; i += nt) { img->aocc[i] = sh.aocc[i]; img->aB[i] = sh.aB[i]; }
    if (SEP && P.any_sep) {
        for (int q = 0; q < P.n_phases; ++q)
            for (int s = tid; s < PTK_PROF_LDS; s += nt) img->prof[q][s] = sh.prof[q][s];
    } else {
        for (int i = tid; i < P.n_phases * 16; i += nt) img->prof[i >> 4][i & 15] = sh.prof[i >> 4][i & 15];
    }
    if (SEP && P.any_sep) {
        for (int i = tid; i < PTK_N_INST; i += nt) img->inst[i] = sh.inst[i];
        const int nw = (int)(sizeof(SepPhase) / sizeof(float));
        for (int i = tid; i < P.n_phases * nw; i += nt)
            reinterpret_cast<float *>(img->sep)[i] = reinterpret_cast<const float *>(sh.sep)[i];
    }
    for (int i = tid; i < PTK_MAX_BKG_F; i += nt) img->bkg[i] = sh.bkg[i];
    if (tid < P.n_phases) img->phase_ok[tid] = sh.phase_ok[tid];
    if (tid == 0) {
        img->cell_ok = sh.cell_ok; img->prof_ok = sh.prof_ok; img->go = sh.go; img->boot = sh.accept;
        img->skipped = sh.skipped;
    }
    for (int g = tid; g < (P.np + 511) >> 9; g += nt) { img->grange[g][0] = 0x7fffffff; img->grange[g][1] = -1; }
}

template <bool SEP = true>
__device__ inline void image_load(const DevPlan &P, const TrialImage *img, SetupShared &sh) {
    const int tid = threadIdx.x, nt = blockDim.x, na = total_atoms(P);
    for (int i = tid; i < P.n_params; i += nt) sh.p[i] = img->p[i];
    for (int i = tid; i < P.n_phases * 6; i += nt) sh.metric[i / 6][i % 6] = img->metric[i / 6][i % 6];
    for (int i = tid; i < na * 3; i += nt) sh.apos[i / 3][i % 3] = img->apos[i / 3][i % 3];
    for (int i = tid; i < na; i += nt) { sh.aocc[i] = img->aocc[i]; sh.aB[i] = img->aB[i]; }
    if (SEP && P.any_sep) {
        for (int q = 0; q < P.n_phases; ++q)
            for (int s = tid; s < PTK_PROF_LDS; s += nt) sh.prof[q][s] = img->prof[q][s];
    } else {
        for (int i = tid; i < P.n_phases * 16; i += nt) sh.prof[i >> 4][i & 15] = img->prof[i >> 4][i & 15];
    }
    if (SEP && P.any_sep) {
        for (int i = tid; i < PTK_N_INST; i += nt) sh.inst[i] = img->inst[i];
        const int nw = (int)(sizeof(SepPhase) / sizeof(float));
        for (int i = tid; i < P.n_phases * nw; i += nt)
            reinterpret_cast<float *>(sh.sep)[i] = reinterpret_cast<const float *>(img->sep)[i];
    }
    for (int i = tid; i < PTK_MAX_BKG_F; i += nt) sh.bkg[i] = img->bkg[i];
    if (tid < P.n_phases) sh.phase_ok[tid] = img->phase_ok[tid];
    if (tid == 0) { sh.cell_ok = img->cell_ok; sh.prof_ok = img->prof_ok; sh.go = img->go; }
}

// Setup of trial t: a given vector (X != null), or the PT proposal of chain t
// (js/pt_gpu.js:361 propose; rietveld_core.js:2196-2212): one coordinate j drawn
// uniformly, x_j + z scale_j sqrt(T_c) reflected into its box, log u for the Metropolis
// test.  Random numbers are keyed on (seed, chain, the chain's step counter).
template <bool SEP = true>
__device__ __forceinline__ void setup_body(const DevPlan &P, const double *X, TrialImage *img, int t, int pt_mode, int bootstrap,
                                  SetupShared &sh, double *prop, const double2 *pick = nullptr) {
    const int nv = P.nv;
    bool boot = false, skip = false;
    double *st = nullptr;
    if (pt_mode) {
        st = P.state + (size_t)t * P.stride;
        const double step = st[3 * nv + 2];
        boot = bootstrap || step < 0.0;
        for (int j = threadIdx.x; j < nv; j += blockDim.x) prop[j] = st[j];
        __syncthreads();
        if (threadIdx.x == 0 && !boot) {
            const unsigned long long ctr = (unsigned long long)step;
            const double temp = P.temp[t];
            const double u3 = uniform01(P.seed, t, ctr, 3);
            if (pick) {
                // the move chosen by the surrogate screen among candidates drawn from
                // the same stream (candidate 0 is the move below)
                const double2 mv = pick[t];
                if (mv.x < 0.0) skip = true;     // predicted hopeless: not evaluated, rejected
                else prop[(int)mv.x] = mv.y;
            } else {
                double u0 = uniform01(P.seed, t, ctr, 0), u1 = uniform01(P.seed, t, ctr, 1);
                double u2 = uniform01(P.seed, t, ctr, 2);
                int j = (int)floor(u0 * nv);
                if (j > nv - 1) j = nv - 1;
                double z = sqrt(-2.0 * log(u1)) * cos(2.0 * kPi * u2);
                prop[j] = reflect_box(prop[j] + z * P.vscale[j] * sqrt(temp), P.vlo[j], P.vhi[j]);
            }
            st[3 * nv + 8] = log(u3);
        }
        __syncthreads();
        if (P.sf && !boot && !pick) {
            // the surrogate's skip test (one proposal per chain): the network's log Rwp at the
            // proposal and at the chain's state, and the predicted Metropolis acceptance
            const ScreenFused &F = *P.sf;
            __shared__ float sf_u[2][256], sf_h[2][2][256];
            const int i = threadIdx.x;                   // 64 threads: one input each
            sf_u[0][i] = i < nv ? sf_from_bf16(sf_input(F, i, prop[i])) : 0.0f;
            sf_u[1][i] = i < nv ? sf_from_bf16(sf_input(F, i, st[i])) : 0.0f;
            __syncthreads();
            float y0, y1;
            sf_predict2(F, sf_u, sf_h, y0, y1);
            if (threadIdx.x == 0 && F.skip_log > 0.0f && *F.iter >= F.warmup_iters && t < F.cold_chains) {
                const float d = y0 - y1 - F.skip_margin;
                const double cur = st[3 * nv];
                if (d > 0.0f && cur > 0.0 && cur < 3e38 && cur * expm1(2.0 * (double)d) > 2.0 * P.temp[t] * (double)F.skip_log) {
                    skip = true;
                    atomicAdd(F.skipped, 1u);
                }
            }
        }
        for (int j = threadIdx.x; j < nv; j += blockDim.x) st[nv + j] = prop[j];
        materialize(P, prop, sh);
    } else {
        materialize(P, X + (size_t)t * nv, sh);
    }
    __syncthreads();
    trial_setup<SEP>(P, sh);
    if (threadIdx.x == 0) {
        bool go = sh.cell_ok;
        if (pt_mode) {
            // a feasible replica may not step outside the variance region (pt_gpu.js:361)
            go = go && (boot || st[3 * nv + 11] < 0.5 || sh.prof_ok) && !skip;
            st[3 * nv + 7] = go ? 1.0 : 0.0;
        }
        sh.go = go ? 1 : 0;
        sh.accept = boot ? 1 : 0;
        sh.skipped = skip ? 1 : 0;
    }
    __syncthreads();
    image_store<SEP>(P, sh, img + t);
}

// (t_off: the first trial of a launch over a contiguous part of the trials)
__global__ void __launch_bounds__(64) k_stage_setup(DevPlan P, const double *__restrict__ X, TrialImage *img,
                                                    int pt_mode, int bootstrap, const double2 *__restrict__ pick, int t_off) {
    __shared__ SetupShared sh;
    __shared__ double prop[PTK_MAX_VARS];
    if (P.any_sep) setup_body<true>(P, X, img, t_off + blockIdx.x, pt_mode, bootstrap, sh, prop, pick);
    else setup_body<false>(P, X, img, t_off + blockIdx.x, pt_mode, bootstrap, sh, prop, pick);
}

// Setup of trial t in a further dataset of a multi-dataset plan: the first dataset's PT
// proposal of chain t (state slots [nv, 2 nv)) materialised in this dataset's parameter
// layout; the trial goes when it goes in the first dataset and its cell and, for a
// feasible chain, its profiles are valid here (js/pt_gpu.js:361 rule per dataset)
template <bool SEP>
__device__ __forceinline__ void follow_body(const DevPlan &P, const double *mstate, int mstride, const TrialImage *mimg,
                                            TrialImage *img, int t, SetupShared &sh) {
    const double *st = mstate + (size_t)t * mstride;
    materialize(P, st + P.nv, sh);
    __syncthreads();
    trial_setup<SEP>(P, sh);
    if (threadIdx.x == 0) {
        const bool boot = mimg[t].boot != 0;
        const bool go = mimg[t].go && sh.cell_ok && (boot || st[3 * P.nv + 11] < 0.5 || sh.prof_ok);
        sh.go = go ? 1 : 0;
        sh.accept = boot ? 1 : 0;
        sh.skipped = 0;
    }
    __syncthreads();
    image_store<SEP>(P, sh, img + t);
}

__global__ void __launch_bounds__(64) k_stage_follow(DevPlan P, const double *__restrict__ mstate, int mstride,
                                                     const TrialImage *__restrict__ mimg, TrialImage *img, int t_off) {
    __shared__ SetupShared sh;
    if (P.any_sep) follow_body<true>(P, mstate, mstride, mimg, img, t_off + blockIdx.x, sh);
    else follow_body<false>(P, mstate, mstride, mimg, img, t_off + blockIdx.x, sh);
}

// span.w bit of a staged line that must take the general path
constexpr int kGeneralLine = 4;
// |u| beyond which exp2(-u^2) < 2^-40: the Gaussian term of a plain line is dropped there
constexpr float kGaussCutF = 6.3245553f;

// a line record and its window in the staged form (StagedLine); uni: the grid is uniform
__device__ __forceinline__ StagedLine stage_line(const DevPlan &P, const LineRec &R, int4 v, bool uni) {
    if (!uni) return StagedLine{R.a, R.r, R.l, v};
    // k need not be the nearest grid point: any k within a few steps keeps c - t_k small
    const double c = (double)R.a.x + (double)R.a.y;
    const double k = (double)rintf((float)(c - P.t0) * (1.0f / P.step_hi));
    const double off = c - (P.t0 + k * P.step);
    StagedLine o{make_float4((float)k, (float)off, R.a.z, R.a.w), R.r, R.l, v};
    if (v.w == 0) {
        if (R.r.x < 0.0f) {
            const double sc = (double)sqrtf(-R.r.x);
            o.l = make_float4((float)(P.step * sc), (float)(off * sc), R.r.y / R.r.x, 0.0f);
            const double sq = (double)sqrtf(fmaxf(R.r.y, 0.0f));
            o.r.x = (float)(P.step * sq);
            o.r.y = (float)(off * sq);
            // the Gaussian's range in grid points, i - k in ((l.y - cut) / l.x, (l.y + cut) / l.x),
            // widened by a point on each side for the rounding of u and of this estimate
            const float il = __builtin_amdgcn_rcpf(o.l.x), kf = o.a.x;
            const float lo = floorf((o.l.y - kGaussCutF) * il) - 2.0f + kf, hi = ceilf((o.l.y + kGaussCutF) * il) + 3.0f + kf;
            const float sx = (float)v.x, sy = (float)v.y;
            const unsigned lo_off = lo > sx ? (unsigned)fminf(lo - sx, 65535.0f) : 0u;
            const unsigned hi_off = hi < sy ? (unsigned)fminf(sy - hi, 65535.0f) : 0u;
            o.span.z = (int)(lo_off | (hi_off << 16));
        } else {
            o.span.w = kGeneralLine;
        }
    }
    return o;
}

// staged line records of reflections [kb * blockDim.x, ...) of trial t; every line with a
// window widens the line range of the 512-point groups it meets (reduced in LDS, then one
// global atomic min / max per group and workgroup)
// set_up: the trial's setup values are already in sh (the same workgroup formed them), so
// its image is not read back
template <bool SEP = true>
__device__ __forceinline__ void lines_body(const DevPlan &P, TrialImage *img, StagedLine *rec, int t, int kb, Shared &sh,
                                           int (*gr)[2], float *sfpart, bool set_up = false) {
    const int k = kb * blockDim.x + threadIdx.x;
    const int4 hk = k < P.nr ? P.hklm[k] : make_int4(0, 0, 0, 0);     // issued before the staging waits
    stage_reps(P, sh);
    if (!set_up) image_load<SEP>(P, img + t, sh);
    const int ng = (P.np + 511) >> 9;
    for (int g = threadIdx.x; g < ng; g += blockDim.x) { gr[g][0] = 0x7fffffff; gr[g][1] = -1; }
    __syncthreads();
    // chunked representative sums (a plan with few reflections and many representatives):
    // with sfpart, thread p takes chunk p % nch of the workgroup's reflection p / nch
    const int nch = P.rep_chunks;
    if (nch > 1 && sfpart && sh.go) {
        const int k0 = kb * blockDim.x, nk = min((int)blockDim.x, P.nr - k0);
        for (int p = threadIdx.x; p < nk * nch; p += blockDim.x) {
            const int kc = k0 + p / nch, c = p % nch;
            int q = 0;
            while (q + 1 < P.n_phases && kc >= P.phase[q + 1].refl_off) ++q;
            const DevPhase &ph = P.phase[q];
            float *o = sfpart + (size_t)p * kSfSlots;
            const int rpc = (ph.n_reps + nch - 1) / nch, r0 = min(ph.n_reps, c * rpc), r1 = min(ph.n_reps, r0 + rpc);
            const int4 hc = P.hklm[kc];
            for (int j0 = 0; j0 < ph.n_atoms; j0 += 4) {
                const int nj = min(4, ph.n_atoms - j0);
                float ca[4] = {0.f, 0.f, 0.f, 0.f}, sa[4] = {0.f, 0.f, 0.f, 0.f};
                if (ph.centro) sf_rep_sums<true>(sh, ph, hc.x, hc.y, hc.z, j0, nj, r0, r1, ca, sa);
                else sf_rep_sums<false>(sh, ph, hc.x, hc.y, hc.z, j0, nj, r0, r1, ca, sa);
#pragma unroll
                for (int a = 0; a < 4; ++a)
                    if (a < nj) { o[2 * (j0 + a)] = ca[a]; o[2 * (j0 + a) + 1] = sa[a]; }
            }
        }
        __syncthreads();
    }
    if (k < P.nr) {
        const size_t base = (size_t)t * 2 * P.nr + 2 * (size_t)k;
        const float4 z4 = make_float4(0.f, 0.f, 0.f, 0.f);
        if (sh.go) {
            LineRec lr[2] = {{z4, z4, z4}, {z4, z4, z4}};
            int4 ls[2];
            build_reflection<SEP>(P, sh, k, lr, ls, nullptr, &hk, SEP && P.has_free ? P.sp->fspan + base : nullptr,
                                  nch > 1 && sfpart ? sfpart + (size_t)threadIdx.x * nch * kSfSlots : nullptr);
            const bool uni = P.uniform != 0;
            for (int q = 0; q < 2; ++q) {
                const int4 sp = ls[q];
                if (sp.y > sp.x) {
                    rec[base + q] = stage_line(P, lr[q], sp, uni);
                    for (int g = sp.x >> 9; g <= (sp.y - 1) >> 9; ++g) {
                        atomicMin(&gr[g][0], 2 * k + q);
                        atomicMax(&gr[g][1], 2 * k + q);
                    }
                } else {
                    rec[base + q] = StagedLine{lr[q].a, lr[q].r, lr[q].l, sp};   // no window here (a free line's is in fspan)
                }
            }
        } else {
            rec[base] = StagedLine{z4, z4, z4, make_int4(0, 0, 0, 0)};
            rec[base + 1] = StagedLine{z4, z4, z4, make_int4(0, 0, 0, 0)};
            if (SEP && P.has_free) { P.sp->fspan[base] = make_int4(0, 0, 0, 0); P.sp->fspan[base + 1] = make_int4(0, 0, 0, 0); }
        }
    }
    __syncthreads();
    for (int g = threadIdx.x; g < ng; g += blockDim.x)
        if (gr[g][1] >= 0) {
            atomicMin(&img[t].grange[g][0], gr[g][0]);
            atomicMax(&img[t].grange[g][1], gr[g][1]);
        }
}

// Workgroup b of a 1-D launch over ntrials x `per` units: unit u of trial t, with every
// unit of trial t on XCD t mod 8 (workgroups go to the XCDs round robin), so a trial's line
// records are written and read through one L2.  The grid is 8 ceil(ntrials / 8) per;
// returns false for the padding slots.
__device__ __forceinline__ bool xcd_unit(int b, int per, int ntrials, int &t, int &u) {
    const int x = b & 7, j = b >> 3;
    t = (j / per) * 8 + x;
    u = j % per;
    return t < ntrials;
}

// CHUNK: the plan chunks its representative sums (DevPlan::rep_chunks > 1); this instance
// holds the chunk sums in LDS and forms them on otherwise idle threads
template <bool CHUNK>
__global__ void __launch_bounds__(512) k_stage_lines(DevPlan P, TrialImage *img, StagedLine *rec, int per, int ntrials,
                                                     int t_off) {
    __shared__ Shared sh;
    __shared__ int gr[PTK_MAX_GROUPS][2];
    __shared__ float sfpart[CHUNK ? 256 * kSfSlots : 1];
    int t, u;
    if (!xcd_unit(blockIdx.x, per, ntrials, t, u)) return;
    float *parts = CHUNK ? sfpart : nullptr;
    // the lean instance for plans with neither separated profiles nor free intensities
    if (P.any_sep || P.has_free) lines_body<true>(P, img, rec, t_off + t, u, sh, gr, parts);
    else lines_body<false>(P, img, rec, t_off + t, u, sh, gr, parts);
}

constexpr int kAccWaves = 4;           // wavefronts per accumulate workgroup

// points per accumulate workgroup for TPW 128-point tiles per wavefront (two points per lane
// and tile); TPW is the tile shape the autotuner picks (1, 2 or 4)
__host__ __device__ constexpr int acc_points(int tpw) { return kAccWaves * tpw * 128; }

// one line at one point of a lane (delta = t - centre); `split` and `asym` are uniform
// over the wavefront (scalar branches), the window test is a select
__device__ __forceinline__ float line_point(const LineRec &R, bool split, float asym, float2 t, bool inwin) {
    float d = (t.x - R.a.x) + (t.y - R.a.y);
    if (asym != 0.0f && fabsf(d) >= 1e-9f) d = d / (1.0f - fmaxf_js(-0.95f, fminf_js(0.95f, asym * d)));
    float4 f = R.r;
    if (split && d < 0.0f) f = R.l;
    const float q = d * d;
    const float g = __builtin_amdgcn_exp2f(f.x * q);
    const float l = __builtin_amdgcn_rcpf(fmaf(f.y, q, 1.0f));
    const float v = fmaxf(fmaf(f.z, l, fmaf(f.w, g, -R.a.w)), 0.0f);
    return inwin ? R.a.z * v : 0.0f;
}

// One staged line of the general form (split flanks, asymmetry) at a lane's two points
// (indices fi0, fi1; float-float 2theta t0, t1).  No window test: every profile here
// decreases away from its centre on each side (the asymmetry map d / (1 - clamp(a d)) is
// increasing), and the pedestal is the larger of its two values at the window edges, so
// at a point outside the window max(0, profile - pedestal) is zero -- up to the rounding
// of the pedestal itself, as at the window edge.  The profile is at most 1 at the centre
// (its two weights sum to 1 on each flank), so the clip is a clamp to [0, 1].
template <bool UNI>
__device__ __forceinline__ void line_general(const float4 &ra, const float4 &rr, const StagedLine &L, bool split, bool has_asym,
                                             float asym, float2 t0, float2 t1, float fi0, float fi1, float step, float &a0,
                                             float &a1) {
    float d0, d1;
    if (UNI) {
        d0 = fmaf(fi0 - ra.x, step, -ra.y);
        d1 = fmaf(fi1 - ra.x, step, -ra.y);
    } else {
        d0 = (t0.x - ra.x) + (t0.y - ra.y);
        d1 = (t1.x - ra.x) + (t1.y - ra.y);
    }
    if (has_asym) {
        if (fabsf(d0) >= 1e-9f) d0 = d0 / (1.0f - fmaxf_js(-0.95f, fminf_js(0.95f, asym * d0)));
        if (fabsf(d1) >= 1e-9f) d1 = d1 / (1.0f - fmaxf_js(-0.95f, fminf_js(0.95f, asym * d1)));
    }
    float4 f0 = rr, f1 = rr;
    if (split) {
        const float4 fl = L.l;
        if (d0 < 0.0f) f0 = fl;
        if (d1 < 0.0f) f1 = fl;
    }
    const float q0 = d0 * d0, q1 = d1 * d1;
    const float g0 = __builtin_amdgcn_exp2f(f0.x * q0), g1 = __builtin_amdgcn_exp2f(f1.x * q1);
    const float l0 = __builtin_amdgcn_rcpf(fmaf(f0.y, q0, 1.0f)), l1 = __builtin_amdgcn_rcpf(fmaf(f1.y, q1, 1.0f));
    const float v0 = __builtin_amdgcn_fmed3f(fmaf(f0.z, l0, fmaf(f0.w, g0, -ra.w)), 0.0f, 1.0f);
    const float v1 = __builtin_amdgcn_fmed3f(fmaf(f1.z, l1, fmaf(f1.w, g1, -ra.w)), 0.0f, 1.0f);
    a0 = fmaf(ra.z, v0, a0);
    a1 = fmaf(ra.z, v1, a1);
}

// A plain staged line (equal flanks, no asymmetry) at one point of a lane (index fi,
// float-float 2theta t), before the line's height: clamp(cl L + cg G - pedestal, 0, 1).
template <bool UNI>
__device__ __forceinline__ float plain_value(const float4 &a, const float4 &r, const float4 &l, float2 t, float fi) {
    float n, c;
    if (UNI) {
        const float u = fmaf(fi - a.x, l.x, -l.y);
        n = u * -u;                     // exp2 argument of the Gaussian
        c = l.z;
    } else {
        const float d = (t.x - a.x) + (t.y - a.y);
        n = d * d;
        c = r.y;
    }
    const float g = __builtin_amdgcn_exp2f(UNI ? n : r.x * n);
    const float lz = __builtin_amdgcn_rcpf(fmaf(c, n, 1.0f));
    return __builtin_amdgcn_fmed3f(fmaf(r.z, lz, fmaf(r.w, g, -a.w)), 0.0f, 1.0f);
}

constexpr int kStage = 256;            // line records staged in LDS per pass (16 KB: 8 waves per SIMD)

// LDS of one accumulate workgroup: a chunk of the trial's staged line records, the chi^2
// of every wavefront tile
struct AccShared {
    StagedLine line[kStage];
    float red_s[4][kAccWaves], red_c[4][kAccWaves];
};

// A plain staged line on a uniform grid at a point where its Gaussian term is dropped
// (index fi): clamp(cl / (1 + w^2) - pedestal, 0, 1), w in the Lorentzian's scaled form.
__device__ __forceinline__ float lorentz_value(const float4 &a, const float4 &r, float fi) {
    const float w = fmaf(fi - a.x, r.x, -r.y);
    return __builtin_amdgcn_fmed3f(fmaf(r.z, __builtin_amdgcn_rcpf(fmaf(w, w, 1.0f)), -a.w), 0.0f, 1.0f);
}

// Per-point sums of one wavefront tile, one per class of (line, tile) pair: plain lines
// that meet both 64-point halves of the tile, with their Gaussian term (two) or without it
// (twl: the Gaussian is negligible over the whole tile); plain lines that meet one half,
// with (one) or without (onl) the Gaussian there; other lines (gen).  Each class is summed
// in line order and the classes are added in a fixed order at the end, so the order of
// the additions does not depend on how lines fall into LDS passes or tile shapes.
struct TileAcc {
    float two0, two1, twl0, twl1, one0, one1, onl0, onl1, gen0, gen1;
};

// The staged lines in LDS that meet one wavefront's 128-point tile [t_lo, t_lo + 128),
// summed into the lane's two points (t_lo + lane, t_lo + lane + 64).  Lines whose window
// misses the tile take no part (their mask bits are clear).
template <bool UNI>
__device__ __forceinline__ void tile_lines(const AccShared &S, int count, int t_lo, float2 t0, float2 t1, float step, TileAcc &A) {
    const int lane = threadIdx.x & 63, t_mid = t_lo + 64, t_hi = t_lo + 128;
    const float fi0 = (float)(t_lo + lane), fi1 = fi0 + 64.0f;
    const float4 zero4 = make_float4(0.f, 0.f, 0.f, 0.f);
    for (int lb = 0; lb < count; lb += 64) {
        // per staged line, as bit masks over the 64 lines of this block (scalar registers)
        const int q = lb + lane;
        bool oa = false, ob = false, pl = false, sf = false, af = false, ga = false, gb = false;
        if (q < count) {
            const int4 v = S.line[q].span;
            oa = v.x < t_mid && v.y > t_lo;
            ob = v.x < t_hi && v.y > t_mid;
            pl = v.w == 0;
            sf = (v.w & 1) != 0;
            af = (v.w & 2) != 0;
            // the Gaussian's range (the window itself off a uniform grid)
            const int glo = UNI && pl ? v.x + (v.z & 0xffff) : v.x;
            const int ghi = UNI && pl ? v.y - (int)((unsigned)v.z >> 16) : v.y;
            ga = glo < t_mid && ghi > t_lo;
            gb = glo < t_hi && ghi > t_mid;
        }
        const bool both = pl && oa && ob, only_a = pl && oa && !ob, only_b = pl && ob && !oa;
        unsigned long long m_two = __ballot(both && (ga || gb)), m_a = __ballot(only_a && ga), m_b = __ballot(only_b && gb);
        unsigned long long m_twl = UNI ? __ballot(both && !ga && !gb) : 0ull;
        const unsigned long long m_al = UNI ? __ballot(only_a && !ga) : 0ull, m_bl = UNI ? __ballot(only_b && !gb) : 0ull;
        unsigned long long m_gen = __ballot(!pl && (oa || ob));
        const unsigned long long m_split = __ballot(sf), m_asym = __ballot(af);
        // plain lines over the whole tile, two per iteration (all their LDS reads first)
        int n = __popcll(m_two);
        for (; n >= 2; n -= 2) {
            const int b1 = __ffsll((long long)m_two) - 1;
            m_two &= m_two - 1ull;
            const int b2 = __ffsll((long long)m_two) - 1;
            m_two &= m_two - 1ull;
            const StagedLine &L1 = S.line[lb + b1], &L2 = S.line[lb + b2];
            const float4 a1 = L1.a, r1 = L1.r, l1 = UNI ? L1.l : zero4, a2 = L2.a, r2 = L2.r, l2 = UNI ? L2.l : zero4;
            const float v10 = plain_value<UNI>(a1, r1, l1, t0, fi0), v11 = plain_value<UNI>(a1, r1, l1, t1, fi1);
            const float v20 = plain_value<UNI>(a2, r2, l2, t0, fi0), v21 = plain_value<UNI>(a2, r2, l2, t1, fi1);
            A.two0 = fmaf(a1.z, v10, A.two0);
            A.two1 = fmaf(a1.z, v11, A.two1);
            A.two0 = fmaf(a2.z, v20, A.two0);
            A.two1 = fmaf(a2.z, v21, A.two1);
        }
        if (n) {
            const StagedLine &L1 = S.line[lb + __ffsll((long long)m_two) - 1];
            const float4 a1 = L1.a, r1 = L1.r, l1 = UNI ? L1.l : zero4;
            const float v10 = plain_value<UNI>(a1, r1, l1, t0, fi0), v11 = plain_value<UNI>(a1, r1, l1, t1, fi1);
            A.two0 = fmaf(a1.z, v10, A.two0);
            A.two1 = fmaf(a1.z, v11, A.two1);
        }
        if (UNI) {
            // the same without the Gaussian term
            n = __popcll(m_twl);
            for (; n >= 2; n -= 2) {
                const int b1 = __ffsll((long long)m_twl) - 1;
                m_twl &= m_twl - 1ull;
                const int b2 = __ffsll((long long)m_twl) - 1;
                m_twl &= m_twl - 1ull;
                const StagedLine &L1 = S.line[lb + b1], &L2 = S.line[lb + b2];
                const float4 a1 = L1.a, r1 = L1.r, a2 = L2.a, r2 = L2.r;
                const float v10 = lorentz_value(a1, r1, fi0), v11 = lorentz_value(a1, r1, fi1);
                const float v20 = lorentz_value(a2, r2, fi0), v21 = lorentz_value(a2, r2, fi1);
                A.twl0 = fmaf(a1.z, v10, A.twl0);
                A.twl1 = fmaf(a1.z, v11, A.twl1);
                A.twl0 = fmaf(a2.z, v20, A.twl0);
                A.twl1 = fmaf(a2.z, v21, A.twl1);
            }
            if (n) {
                const StagedLine &L1 = S.line[lb + __ffsll((long long)m_twl) - 1];
                const float4 a1 = L1.a, r1 = L1.r;
                A.twl0 = fmaf(a1.z, lorentz_value(a1, r1, fi0), A.twl0);
                A.twl1 = fmaf(a1.z, lorentz_value(a1, r1, fi1), A.twl1);
            }
        }
        // plain lines over one half: two lines at the lane's point of that half
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            unsigned long long m = h ? m_b : m_a;
            const float2 t = h ? t1 : t0;
            const float fi = h ? fi1 : fi0;
            float acc = h ? A.one1 : A.one0;
            int k = __popcll(m);
            for (; k >= 2; k -= 2) {
                const int b1 = __ffsll((long long)m) - 1;
                m &= m - 1ull;
                const int b2 = __ffsll((long long)m) - 1;
                m &= m - 1ull;
                const StagedLine &L1 = S.line[lb + b1], &L2 = S.line[lb + b2];
                const float4 a1 = L1.a, r1 = L1.r, l1 = UNI ? L1.l : zero4, a2 = L2.a, r2 = L2.r, l2 = UNI ? L2.l : zero4;
                const float v1 = plain_value<UNI>(a1, r1, l1, t, fi), v2 = plain_value<UNI>(a2, r2, l2, t, fi);
                acc = fmaf(a1.z, v1, acc);
                acc = fmaf(a2.z, v2, acc);
            }
            if (k) {
                const StagedLine &L1 = S.line[lb + __ffsll((long long)m) - 1];
                const float4 a1 = L1.a;
                acc = fmaf(a1.z, plain_value<UNI>(a1, L1.r, UNI ? L1.l : zero4, t, fi), acc);
            }
            if (h) A.one1 = acc; else A.one0 = acc;
            if (UNI) {
                m = h ? m_bl : m_al;
                acc = h ? A.onl1 : A.onl0;
                k = __popcll(m);
                for (; k >= 2; k -= 2) {
                    const int b1 = __ffsll((long long)m) - 1;
                    m &= m - 1ull;
                    const int b2 = __ffsll((long long)m) - 1;
                    m &= m - 1ull;
                    const StagedLine &L1 = S.line[lb + b1], &L2 = S.line[lb + b2];
                    const float4 a1 = L1.a, r1 = L1.r, a2 = L2.a, r2 = L2.r;
                    const float v1 = lorentz_value(a1, r1, fi), v2 = lorentz_value(a2, r2, fi);
                    acc = fmaf(a1.z, v1, acc);
                    acc = fmaf(a2.z, v2, acc);
                }
                if (k) {
                    const StagedLine &L1 = S.line[lb + __ffsll((long long)m) - 1];
                    const float4 a1 = L1.a;
                    acc = fmaf(a1.z, lorentz_value(a1, L1.r, fi), acc);
                }
                if (h) A.onl1 = acc; else A.onl0 = acc;
            }
        }
        // the others (split flanks, asymmetry, or a profile that needs the general form)
        while (m_gen) {
            const int b = __ffsll((long long)m_gen) - 1;
            m_gen &= m_gen - 1ull;
            const StagedLine &L = S.line[lb + b];
            const bool has_asym = (m_asym >> b) & 1ull;
            const float asym = has_asym ? __int_as_float(__builtin_amdgcn_readfirstlane(L.span.z)) : 0.0f;
            line_general<UNI>(L.a, L.r, L, (m_split >> b) & 1ull, has_asym, asym, t0, t1, fi0, fi1, step, A.gen0, A.gen1);
        }
    }
}

// Banded profile sum and partial chi^2 of points [P s, P s + P) of trial t, P =
// acc_points(TPW): the staged line records of the workgroup's line range (the union of
// its 512-point groups' ranges) are copied into LDS a chunk at a time; every wavefront
// sums those that meet each of its TPW 128-point tiles (two points per lane), and the
// chi^2 of every 512-point group (the tiles j of the four wavefronts) is reduced with
// two-sum into partial[t * nsplit + group], nsplit = the trial's 512-point group count.
// FREE: the plan has free intensities, so the residual of every point is kept for their solve
template <int TPW, bool FREE = false>
__device__ inline void accum_body(const DevPlan &P, const TrialImage *__restrict__ img, const StagedLine *__restrict__ grec,
                                  double2 *partial, int nsplit, int t, int split, AccShared &S) {
    // the wavefront index is uniform: keep it (and every tile bound) in scalar registers
    const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const TrialImage *im = img + t;
    if (!im->go) {
        if (tid < TPW && split * TPW + tid < nsplit) partial[(size_t)t * nsplit + split * TPW + tid] = make_double2(0.0, 0.0);
        return;
    }
    const int lo_pt = split * acc_points(TPW), hi_pt = min(P.np, lo_pt + acc_points(TPW));
    const int nl = 2 * P.nr;
    const StagedLine *rec = grec + (size_t)t * nl;
    const bool uni = P.uniform != 0;
    // lines that can meet this workgroup's points: the union of its groups' ranges
    int first = 0x7fffffff, last = -1;
    for (int g = split * TPW; g < min(nsplit, split * TPW + TPW); ++g) {
        first = min(first, im->grange[g][0]);
        last = max(last, im->grange[g][1]);
    }
    const int lend = min(nl, last + 1);
    TileAcc ta[TPW];
    float2 tf[2 * TPW], ow[2 * TPW];
#pragma unroll
    for (int j = 0; j < TPW; ++j) {
        const int base = lo_pt + (wave + kAccWaves * j) * 128 + lane;
        ta[j] = TileAcc{0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        tf[2 * j] = !uni && base < hi_pt ? P.tff[base] : make_float2(0.f, 0.f);
        tf[2 * j + 1] = !uni && base + 64 < hi_pt ? P.tff[base + 64] : make_float2(0.f, 0.f);
        ow[2 * j] = base < hi_pt ? P.obs[base] : make_float2(0.f, 0.f);
        ow[2 * j + 1] = base + 64 < hi_pt ? P.obs[base + 64] : make_float2(0.f, 0.f);
    }
    for (int c0 = first; c0 < lend; c0 += kStage) {
        const int count = min(kStage, lend - c0);
        if (c0 > first) __syncthreads();                 // the last chunk's readers are done
        for (int i = tid; i < count; i += blockDim.x) S.line[i] = rec[c0 + i];
        __syncthreads();
#pragma unroll
        for (int j = 0; j < TPW; ++j) {
            const int t_lo = lo_pt + (wave + kAccWaves * j) * 128;
            if (t_lo < hi_pt) {
                if (uni) tile_lines<true>(S, count, t_lo, tf[2 * j], tf[2 * j + 1], P.step_hi, ta[j]);
                else tile_lines<false>(S, count, t_lo, tf[2 * j], tf[2 * j + 1], P.step_hi, ta[j]);
            }
        }
    }
    // chi^2 of this workgroup's points, one partial per 512-point group (the 128-point
    // tiles j of the four wavefronts), so that the partials -- and the ordered sum the
    // finish takes over them -- do not depend on the tile shape
#pragma unroll
    for (int j = 0; j < TPW; ++j) {
        float s = 0.f, c = 0.f;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int i = lo_pt + (wave + kAccWaves * j) * 128 + h * 64 + lane;
            if (i < hi_pt) {
                const float2 o = ow[2 * j + h];
                if (o.y > 0.0f) {
                    const float yb = bkg_value(P, im->bkg, i);
                    const float yp = h ? ((ta[j].two1 + ta[j].twl1) + (ta[j].one1 + ta[j].onl1)) + ta[j].gen1
                                       : ((ta[j].two0 + ta[j].twl0) + (ta[j].one0 + ta[j].onl0)) + ta[j].gen0;
                    const float r = o.x - (yb + yp);
                    neumaier(s, c, o.y * r * r);
                    if (FREE) P.sp->resid[(size_t)t * P.np + i] = r;      // the free intensities' target
                } else if (FREE) {
                    P.sp->resid[(size_t)t * P.np + i] = 0.0f;
                }
            }
        }
        for (int o = 32; o > 0; o >>= 1) {
            float so = __shfl_xor(s, o), co = __shfl_xor(c, o);
            float tt = s + so;
            float e = fabsf(s) >= fabsf(so) ? (s - tt) + so : (so - tt) + s;
            s = tt; c = c + co + e;
        }
        if (lane == 0) { S.red_s[j][wave] = s; S.red_c[j][wave] = c; }
    }
    __syncthreads();
    if (tid < TPW) {
        const int g = split * TPW + tid;
        if (g < nsplit) {
            double hi = 0.0, lo = 0.0;
            for (int w = 0; w < kAccWaves; ++w) { hi += (double)S.red_s[tid][w]; lo += (double)S.red_c[tid][w]; }
            partial[(size_t)t * nsplit + g] = make_double2(hi, lo);
        }
    }
}

template <int TPW, bool FREE>
__global__ void __launch_bounds__(256) k_stage_accum(DevPlan P, const TrialImage *__restrict__ img,
                                                     const StagedLine *__restrict__ grec, double2 *partial, int nsplit, int per,
                                                     int ntrials, int t_off) {
    __shared__ AccShared S;
    int t, u;
    if (!xcd_unit(blockIdx.x, per, ntrials, t, u)) return;
    accum_body<TPW, FREE>(P, img, grec, partial, nsplit, t_off + t, u, S);
}

// chi^2 of trial t: the ordered sum of its partials (3e38 for an invalid trial)
__device__ inline double trial_cost(const TrialImage *im, const double2 *partial, int nsplit, int t) {
    double cost = 3e38;
    if (im->go) {
        double hi = 0.0, lo = 0.0;
        for (int k = 0; k < nsplit; ++k) { double2 v = partial[(size_t)t * nsplit + k]; hi += v.x; lo += v.y; }
        cost = hi + lo;
        if (!isfinite(cost) || cost > 3e38) cost = 3e38;
    }
    return cost;
}

// ---- free intensities (Le Bail / WPPF phases) -------------------------------------
// One workgroup per trial: the weighted normal equations of the free reflections' columns
// against the residual the accumulate stage left (y_obs - background - Rietveld phases),
// then min over I >= 0 of I'HI - 2 b'I by Lawson and Hanson's active set (Solving Least
// Squares Problems, 1974, ch. 23) on columns equilibrated to unit diagonal
// (js/rietveld_core.js:1078-1082), all in FP64: the normal equations by the whole
// workgroup from columns cached in LDS, the active set by one wavefront (nnls_wave), in LDS
// up to 120 live columns and in the trial's global workspace up to 256.  chi^2 =
// c - 2 b'I + I'HI replaces the trial's partial sums (js/pt_gpu.js:259-297 solves the same
// problem per chain).
// the solve's arrays for up to c live columns: line records and spans, the packed normal
// matrix and its factor, three FP64 and six int vectors
__host__ __device__ inline size_t nnls_region_bytes(int c) {
    const size_t tri = (size_t)c * (c + 1) / 2;
    return 16 * (size_t)(2 * c) * 4 + ((tri * 2 + 3 * (size_t)c) * 8 + 15) / 16 * 16 + (size_t)c * 4 * 6;
}
// LDS for up to m live columns out of nf free reflections (the region, then nf flags and
// nf envelope indices)
__host__ __device__ inline size_t nnls_lds_bytes(int m, int nf) { return nnls_region_bytes(m) + (size_t)nf * 8 + 64; }
// a trial slot's global workspace for the solve when more columns are live than LDS holds
size_t nnls_slot_bytes(int nf) { return (nnls_region_bytes(std::min(nf, PTK_MAX_FREE_WS_LIVE)) + 255) / 256 * 256; }
// dynamic LDS of the solve: the region and flags, then a cache of the live columns' values
// over their windows in what is left of the CU's 160 KB (the static reductions aside)
constexpr int kNnlsThreads = 1024;     // 16 wavefronts per trial: the CU's latency is hidden by them
constexpr size_t kNnlsLdsMax = 160 * 1024 - 12 * 1024;
__host__ __device__ inline size_t nnls_base_bytes(int nf) {
    return (nnls_lds_bytes(nf < PTK_MAX_FREE_LIVE ? nf : PTK_MAX_FREE_LIVE, nf) + 15) / 16 * 16;
}
__host__ __device__ inline int nnls_cache_floats(int nf) {
    const size_t b = nnls_base_bytes(nf), f = b >= kNnlsLdsMax ? 0 : (kNnlsLdsMax - b) / 4;
    return (int)(f < 32768 ? f : 32768);
}
__host__ __device__ inline size_t nnls_total_bytes(int nf) { return nnls_base_bytes(nf) + 4 * (size_t)nnls_cache_floats(nf); }

struct NnlsArrays {
    LineRec *R;
    int4 *Sp;
    double *H, *Lf, *b, *x, *nrm;
    int *pas, *blk, *idx, *cs, *ce, *live;
};
// the arrays for capacity c carved from `base` (16-byte aligned), in nnls_region_bytes' order
__device__ __forceinline__ NnlsArrays nnls_carve(unsigned char *base, int c) {
    NnlsArrays A;
    A.R = reinterpret_cast<LineRec *>(base);
    A.Sp = reinterpret_cast<int4 *>(A.R + 2 * c);
    const int Tc = c * (c + 1) / 2;
    A.H = reinterpret_cast<double *>(A.Sp + 2 * c);
    A.Lf = A.H + Tc; A.b = A.Lf + Tc; A.x = A.b + c; A.nrm = A.x + c;
    A.pas = reinterpret_cast<int *>(A.nrm + c);
    A.blk = A.pas + c; A.idx = A.blk + c; A.cs = A.idx + c; A.ce = A.cs + c; A.live = A.ce + c;
    return A;
}
__device__ __forceinline__ int tri_at(int i, int j) { return i >= j ? i * (i + 1) / 2 + j : j * (j + 1) / 2 + i; }

// float-float 2theta of point i: computed on a uniform grid (t0 + i step, within 1e-9 of
// a step of the axis), loaded otherwise
__device__ __forceinline__ float2 point_tff(const DevPlan &P, int i) {
    if (!P.uniform) return P.tff[i];
    const double t = P.t0 + (double)i * P.step;
    const float hi = (float)t;
    return make_float2(hi, (float)(t - (double)hi));
}

// column j (the free reflection's lines at their weights) at point i
__device__ __forceinline__ float free_column(const DevPlan &P, const LineRec *R, const int4 *Sp, int j, int i) {
    float v = 0.0f;
    const float2 ti = point_tff(P, i);
    for (int q = 0; q < 2; ++q) {
        const int4 sp = Sp[2 * j + q];
        if (i >= sp.x && i < sp.y) {
            const LineRec &L = R[2 * j + q];
            v += L.a.z * line_value(L, __int_as_float(sp.z), (ti.x - L.a.x) + (ti.y - L.a.y));
        }
    }
    return v;
}

// ---- the active set of a small solve in one wavefront ----------------------------------
// FP64 helpers: lane k's value (k uniform), lane src's value (src per lane), the value of
// the first active lane (for a value every lane holds alike), the wavefront's sum
__device__ __forceinline__ double lane_value(double v, int k) {
    const long long u = __double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)u, k), hi = (unsigned)__builtin_amdgcn_readlane((int)(u >> 32), k);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
__device__ __forceinline__ double lane_fetch(double v, int src) {
    const long long u = __double_as_longlong(v);
    const unsigned lo = (unsigned)__shfl((int)u, src), hi = (unsigned)__shfl((int)(u >> 32), src);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
__device__ __forceinline__ double uniform_value(double v) {
    const long long u = __double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)u), hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(u >> 32));
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
// Wavefront reductions without LDS round trips: four DPP row rotations (row_ror 8, 4, 2, 1)
// leave every lane with a combination of all 16 lanes of its row, then lanes 0, 16, 32 and
// 48 are combined in that order.  The result is the same in every lane and depends only on
// the inputs.  All 64 lanes must be active.
template <int N>
__device__ __forceinline__ double row_ror(double v) {
    const long long u = __double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_update_dpp(0, (int)u, 0x120 + N, 0xf, 0xf, false),
                   hi = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(u >> 32), 0x120 + N, 0xf, 0xf, false);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
template <int N>
__device__ __forceinline__ int row_ror(int v) { return __builtin_amdgcn_update_dpp(0, v, 0x120 + N, 0xf, 0xf, false); }
__device__ __forceinline__ double wave_sum(double v) {
    v += row_ror<8>(v);
    v += row_ror<4>(v);
    v += row_ror<2>(v);
    v += row_ror<1>(v);
    return (lane_value(v, 0) + lane_value(v, 16)) + (lane_value(v, 32) + lane_value(v, 48));
}
__device__ __forceinline__ double wave_min(double v) {
    v = fmin(v, row_ror<8>(v));
    v = fmin(v, row_ror<4>(v));
    v = fmin(v, row_ror<2>(v));
    v = fmin(v, row_ror<1>(v));
    return fmin(fmin(lane_value(v, 0), lane_value(v, 16)), fmin(lane_value(v, 32), lane_value(v, 48)));
}
// the index of the largest value, the smallest index on ties (an exact total order, so the
// combination order does not matter)
__device__ __forceinline__ int wave_argmax(double bv, int bi) {
    auto take = [&](double ov, int oi) { if (ov > bv || (ov == bv && oi < bi)) { bv = ov; bi = oi; } };
    take(row_ror<8>(bv), row_ror<8>(bi));
    take(row_ror<4>(bv), row_ror<4>(bi));
    take(row_ror<2>(bv), row_ror<2>(bi));
    take(row_ror<1>(bv), row_ror<1>(bi));
    double rv = lane_value(bv, 0);
    int ri = __builtin_amdgcn_readlane(bi, 0);
#pragma unroll
    for (int r = 16; r < 64; r += 16) {
        const double ov = lane_value(bv, r);
        const int oi = __builtin_amdgcn_readlane(bi, r);
        if (ov > rv || (ov == rv && oi < ri)) { rv = ov; ri = oi; }
    }
    return ri;
}
// LDS written by some lanes of the wavefront, then read by others: the writes complete
// (s_waitcnt lgkmcnt(0)) before the wavefront goes on, and no access moves across
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_s_waitcnt(0xc07f);
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
__device__ __forceinline__ int tri_row(int i) { return i * (i + 1) / 2; }

// Lawson and Hanson's active set for m <= 64 S equilibrated columns, run by the 64 lanes
// of one wavefront (lane l, slot s: variable and passive position l + 64 s; x, b, the
// passive and blocked flags and the factor's inverse diagonal in registers) with no
// workgroup barrier.  The passive set keeps its order of entry; its Cholesky factor (packed
// rows in L, plist[position] = variable, pos[variable] = position) grows by one row when a
// column enters and is formed again from the first changed row when columns leave.  The
// triangular solves take a batch of columns per round of loads, so only the readlane /
// FMA chain is sequential.  The steps are those of the workgroup form: the largest
// gradient entry above tol enters (the first on ties), the step back to the boundary when
// the passive solution is not positive, a column that cannot stay is blocked; a warm start
// more than a quarter of whose passive set comes out not positive, or, with more than 64
// passive variables, that needs a third step back, gives way to x = 0 (the optimum is the
// same, reached in fewer steps; small passive sets step back faster than they start again).
// In: x (warm start), pas, blk; out: x, pas, s_int[2] = a block was not positive definite,
// s_int[3] = converged.  The arrays are in LDS, or for more columns than it holds in the
// trial's global workspace; either way only this wavefront touches them here, and its
// memory accesses are in order.
template <int S>
__device__ __forceinline__ void nnls_wave(const double *H, const double *b, double *x, int *pas, int *blk, int m, double tol,
                                          int outer_cap, double *L, int *plist, int *pos, int *s_int) {
    constexpr int B = S == 1 ? 8 : S == 2 ? 2 : 1;      // columns per batch of loads
    const int ln = threadIdx.x & 63;
    const unsigned long long below = (1ull << ln) - 1ull;
    double xj[S], bj[S], dinv[S];
    bool pj[S], kj[S], act[S];
#pragma unroll
    for (int s = 0; s < S; ++s) {
        const int j = ln + 64 * s;
        act[s] = j < m;
        xj[s] = act[s] ? x[j] : 0.0;
        bj[s] = act[s] ? b[j] : 0.0;
        pj[s] = act[s] && pas[j] != 0;
        kj[s] = act[s] && blk[j] != 0;
        dinv[s] = 0.0;
    }
    int fail = 0, conv = 0;
    // the value a register array holds for index c (uniform): lane c & 63, slot c >> 6
    auto at = [&](const double (&v)[S], int c) -> double {
        double r = lane_value(v[0], c & 63);
#pragma unroll
        for (int s = 1; s < S; ++s)
            if ((c >> 6) == s) r = lane_value(v[s], c & 63);
        return r;
    };
    // the value for index idx (per lane); every lane takes part in every shuffle
    auto fetch = [&](const double (&v)[S], int idx) -> double {
        double r = 0.0;
#pragma unroll
        for (int s = 0; s < S; ++s) {
            const double t = lane_fetch(v[s], idx & 63);
            if ((idx >> 6) == s) r = t;
        }
        return r;
    };
    auto fetch_flag = [&](const bool (&f)[S], int idx) -> bool {
        bool r = false;
#pragma unroll
        for (int s = 0; s < S; ++s) {
            const int t = __shfl(f[s] ? 1 : 0, idx & 63);
            if ((idx >> 6) == s) r = t != 0;
        }
        return r;
    };
    // L y = v over positions < p
    auto forward = [&](double (&v)[S], int p) {
        for (int c0 = 0; c0 < p; c0 += B) {
            double lr[B][S];
#pragma unroll
            for (int u = 0; u < B; ++u)
#pragma unroll
                for (int s = 0; s < S; ++s) {
                    const int c = c0 + u, ps = ln + 64 * s;
                    const bool use = c < p && ps > c && ps < p;
                    lr[u][s] = L[use ? tri_row(ps) + c : 0];
                    if (!use) lr[u][s] = 0.0;
                }
#pragma unroll
            for (int u = 0; u < B; ++u) {
                const int c = c0 + u;
                if (c < p) {
                    const double yc = at(v, c) * at(dinv, c);
#pragma unroll
                    for (int s = 0; s < S; ++s) v[s] = ln + 64 * s == c ? yc : v[s] - lr[u][s] * yc;
                }
            }
        }
    };
    // L' z = v over positions < p
    auto backward = [&](double (&v)[S], int p) {
        for (int c1 = p - 1; c1 >= 0; c1 -= B) {
            double lr[B][S];
#pragma unroll
            for (int u = 0; u < B; ++u)
#pragma unroll
                for (int s = 0; s < S; ++s) {
                    const int c = c1 - u, ps = ln + 64 * s;
                    const bool use = c >= 0 && ps < c;
                    lr[u][s] = L[use ? tri_row(c) + ps : 0];
                    if (!use) lr[u][s] = 0.0;
                }
#pragma unroll
            for (int u = 0; u < B; ++u) {
                const int c = c1 - u;
                if (c >= 0) {
                    const double zc = at(v, c) * at(dinv, c);
#pragma unroll
                    for (int s = 0; s < S; ++s) v[s] = ln + 64 * s == c ? zc : v[s] - lr[u][s] * zc;
                }
            }
        }
    };
    // variable q enters at position p: the new row solves L y = H[passive, q]
    auto append = [&](int q, int p) -> bool {
        double h[S];
#pragma unroll
        for (int s = 0; s < S; ++s) {
            const int ps = ln + 64 * s;
            h[s] = ps < p ? H[tri_at(plist[ps], q)] : 0.0;
        }
        forward(h, p);
        double ss = 0.0;
#pragma unroll
        for (int s = 0; s < S; ++s) if (ln + 64 * s < p) ss += h[s] * h[s];
        const double d = uniform_value(H[tri_at(q, q)] - wave_sum(ss));
        if (!(d > 0.0)) return false;
        const double sd = sqrt(d);
#pragma unroll
        for (int s = 0; s < S; ++s) {
            const int ps = ln + 64 * s;
            if (ps < p) L[tri_row(p) + ps] = h[s];
            if (ps == p) dinv[s] = 1.0 / sd;
        }
        if (ln == 0) { L[tri_row(p) + p] = sd; plist[p] = q; pos[q] = p; }
        wave_lds_sync();
        return true;
    };
    // the factor of the passive list, formed again from row r0 (the rows above r0 depend only
    // on the passive columns above r0).  Column by column (Crout), every row below advancing
    // at once: entry (i, k) = (A_ik - L_i0 L_k0 - L_i1 L_k1 - ... - L_i,k-1 L_k,k-1) / L_kk,
    // subtracted in that order and multiplied by 1 / L_kk, and L_kk^2 = A_kk minus the
    // wavefront sum of row k's squares: the operations of append's forward substitution for
    // row i, in the same order, so the factor is the one appending row after row gives.  The
    // first column whose pivot is not positive is the row at which append would stop.
    auto factor = [&](int p, int r0) -> bool {
        if (r0 >= p) return true;
        int qv[S];
#pragma unroll
        for (int s = 0; s < S; ++s) { const int ps = ln + 64 * s; qv[s] = ps < p ? plist[ps] : 0; }
        for (int k = 0; k < p; ++k) {
            int qk = __builtin_amdgcn_readlane(qv[0], k & 63);
#pragma unroll
            for (int s = 1; s < S; ++s) if ((k >> 6) == s) qk = __builtin_amdgcn_readlane(qv[s], k & 63);
            double acc[S];
            bool row[S];
#pragma unroll
            for (int s = 0; s < S; ++s) {
                const int ps = ln + 64 * s;
                row[s] = ps > k && ps >= r0 && ps < p;
                acc[s] = row[s] ? H[tri_at(qv[s], qk)] : 0.0;
            }
            double dk;
            if (k >= r0) {                               // the pivot: row k is complete
                double ss = 0.0;
#pragma unroll
                for (int s = 0; s < S; ++s) {
                    const int ps = ln + 64 * s;
                    if (ps < k) { const double l = L[tri_row(k) + ps]; ss += l * l; }
                }
                const double d = uniform_value(H[tri_at(qk, qk)] - wave_sum(ss));
                if (!(d > 0.0)) return false;
                const double sd = sqrt(d);
                dk = 1.0 / sd;
#pragma unroll
                for (int s = 0; s < S; ++s) if (ln + 64 * s == k) dinv[s] = dk;
                if (ln == 0) L[tri_row(k) + k] = sd;
            } else {
                dk = at(dinv, k);
            }
            for (int c0 = 0; c0 < k; c0 += B) {
                double lk[B], li[B][S];
#pragma unroll
                for (int u = 0; u < B; ++u) {
                    const int c = c0 + u;
                    lk[u] = c < k ? L[tri_row(k) + c] : 0.0;
#pragma unroll
                    for (int s = 0; s < S; ++s) li[u][s] = row[s] && c < k ? L[tri_row(ln + 64 * s) + c] : 0.0;
                }
#pragma unroll
                for (int u = 0; u < B; ++u)
                    if (c0 + u < k) {
#pragma unroll
                        for (int s = 0; s < S; ++s) if (row[s]) acc[s] = acc[s] - li[u][s] * lk[u];
                    }
            }
#pragma unroll
            for (int s = 0; s < S; ++s) if (row[s]) L[tri_row(ln + 64 * s) + k] = acc[s] * dk;
            wave_lds_sync();
        }
        return true;
    };
    // the passive solution L L' z = b_P, in variable space (z of each slot's variable)
    auto solve = [&](int p, double (&z)[S]) {
        double y[S];
#pragma unroll
        for (int s = 0; s < S; ++s) {
            const int ps = ln + 64 * s;
            y[s] = fetch(bj, ps < p ? plist[ps] : 0);
        }
        forward(y, p);
        backward(y, p);
#pragma unroll
        for (int s = 0; s < S; ++s) z[s] = fetch(y, pj[s] ? pos[ln + 64 * s] : 0);
    };
    // drop the variables that left the passive set from the list, keeping the order; r0:
    // the first position that changed
    auto compact = [&](int p, int &r0) -> int {
        int vi[S];
        bool keep[S];
        unsigned long long km[S];
        int nkeep = 0;
        r0 = p;
#pragma unroll
        for (int s = 0; s < S; ++s) {
            const int ps = ln + 64 * s;
            vi[s] = ps < p ? plist[ps] : 0;
            const bool still = fetch_flag(pj, vi[s]);   // every lane takes part in its shuffles
            keep[s] = ps < p && still;
            km[s] = __ballot(keep[s]);
            const unsigned long long gone = __ballot(ps < p && !keep[s]);
            if (gone && r0 == p) r0 = 64 * s + __ffsll((long long)gone) - 1;
        }
        wave_lds_sync();
#pragma unroll
        for (int s = 0; s < S; ++s) {
            const int newpos = nkeep + __popcll(km[s] & below);
            if (keep[s]) { plist[newpos] = vi[s]; pos[vi[s]] = newpos; }
            nkeep += __popcll(km[s]);
        }
        wave_lds_sync();
        return nkeep;
    };
    // the warm passive set in index order
    int p = 0;
#pragma unroll
    for (int s = 0; s < S; ++s) {
        const unsigned long long pm = __ballot(pj[s]);
        if (pj[s]) { const int ps = p + __popcll(pm & below); plist[ps] = ln + 64 * s; pos[ln + 64 * s] = ps; }
        p += __popcll(pm);
    }
    wave_lds_sync();
    bool warm = p > 0;
    if (warm && !factor(p, 0)) fail = 1;
    for (int it = 0; !fail && it < outer_cap; ++it) {
        int jm = -2;
        if (!warm) {
            double gj[S];
#pragma unroll
            for (int s = 0; s < S; ++s) gj[s] = bj[s];
            for (int k0 = 0; k0 < m; k0 += B) {
                double hr[B][S];
#pragma unroll
                for (int u = 0; u < B; ++u)
#pragma unroll
                    for (int s = 0; s < S; ++s) {
                        const bool use = act[s] && k0 + u < m;
                        hr[u][s] = H[use ? tri_at(ln + 64 * s, k0 + u) : 0];
                    }
#pragma unroll
                for (int u = 0; u < B; ++u) {
                    const double xk = at(xj, k0 + u < m ? k0 + u : 0);
#pragma unroll
                    for (int s = 0; s < S; ++s)
                        if (act[s] && k0 + u < m) gj[s] -= hr[u][s] * xk;
                }
            }
            // the largest candidate, the first on ties: each lane's slots in index order,
            // then the wavefront
            double bv = -1.0;
            int bi = 1 << 20;
#pragma unroll
            for (int s = 0; s < S; ++s)
                if (act[s] && !pj[s] && !kj[s] && gj[s] > tol && gj[s] > bv) { bv = gj[s]; bi = ln + 64 * s; }
            jm = __builtin_amdgcn_readfirstlane(wave_argmax(bv, bi));
            if (jm >= m) { conv = 1; break; }
#pragma unroll
            for (int s = 0; s < S; ++s) if (ln + 64 * s == jm) pj[s] = true;
            if (!append(jm, p)) { fail = 1; break; }
            ++p;
        }
        warm = false;
        bool settled = false;
        for (int inner = 0; inner < 3 * m + 10; ++inner) {
            double z[S];
            solve(p, z);
            bool bad[S];
            int nbad = 0;
#pragma unroll
            for (int s = 0; s < S; ++s) {
                bad[s] = pj[s] && !(z[s] > 0.0);
                nbad += __popcll(__ballot(bad[s]));
            }
            if (!nbad) {
#pragma unroll
                for (int s = 0; s < S; ++s) xj[s] = pj[s] ? z[s] : 0.0;
                settled = true;
                break;
            }
            if (jm == -2 && ((inner >= 2 && p > 64) || 4 * nbad > p)) {
#pragma unroll
                for (int s = 0; s < S; ++s) { pj[s] = false; xj[s] = 0.0; }
                p = 0;
                settled = true;
                break;
            }
            double al = 2.0;
#pragma unroll
            for (int s = 0; s < S; ++s) if (bad[s]) al = fmin(al, xj[s] / (xj[s] - z[s]));
            const double alpha = uniform_value(wave_min(al));
#pragma unroll
            for (int s = 0; s < S; ++s) {
                if (pj[s]) xj[s] += alpha * (z[s] - xj[s]);
                if (pj[s] && !(xj[s] > 0.0)) { pj[s] = false; xj[s] = 0.0; }
            }
            int r0 = 0;
            p = compact(p, r0);
            if (!factor(p, r0)) { fail = 1; break; }
        }
        if (fail || !settled) break;
#pragma unroll
        for (int s = 0; s < S; ++s)
            if (jm >= 0 && ln + 64 * s == jm && !pj[s]) kj[s] = true;   // could not stay positive: keep it out
    }
#pragma unroll
    for (int s = 0; s < S; ++s)
        if (act[s]) { const int j = ln + 64 * s; x[j] = xj[s]; pas[j] = pj[s] ? 1 : 0; blk[j] = kj[s] ? 1 : 0; }
    if (ln == 0) { s_int[2] = fail; s_int[3] = conv; }
}

// the solve of trial t over its m live columns (flag[gi]: 1 + the column of free
// reflection gi when it is live, else 0; kk[gi]: its envelope index)
__device__ __forceinline__ void nnls_core(const DevPlan &P, const StagedLine *__restrict__ rec, double2 *partial, int nsplit,
                                          int t, const NnlsArrays &A, int m, const int *flag, const int *kk, double *red,
                                          int *s_int, double &s_tol, float *colc, int ncache, bool in_lds) {
    const int nf = P.n_free, tid = threadIdx.x, nt = blockDim.x, T = m * (m + 1) / 2;
    const size_t base = (size_t)t * 2 * P.nr;
    LineRec *R = A.R;
    int4 *Sp = A.Sp;
    double *H = A.H, *Lf = A.Lf, *b = A.b, *x = A.x, *nrm = A.nrm;
    int *pas = A.pas, *blk = A.blk, *idx = A.idx, *cs = A.cs, *ce = A.ce, *live = A.live;
    for (int gi = tid; gi < nf; gi += nt)               // in envelope order
        if (flag[gi]) live[flag[gi] - 1] = gi;
    const float *r = P.sp->resid + (size_t)t * P.np;
    // the points' weights and residuals in LDS when they fit the cache (the columns' values
    // after them), and c = sum w r^2 in the order of a strided loop (four loads in flight)
    const int np = P.np;
    const bool pts = 2 * np <= ncache;
    float *wl = colc, *rl = colc + np;
    double cw = 0.0;
    for (int i0 = tid; i0 < np; i0 += 4 * nt) {
        float2 o[4];
        float rv[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int i = i0 + u * nt;
            o[u] = i < np ? P.obs[i] : make_float2(0.f, 0.f);
            rv[u] = i < np ? r[i] : 0.0f;
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int i = i0 + u * nt;
            if (i < np) {
                if (pts) { wl[i] = o[u].y; rl[i] = rv[u]; }
                if (o[u].y > 0.0f) cw += (double)o[u].y * (double)rv[u] * (double)rv[u];
            }
        }
    }
    if (pts) { colc += 2 * np; ncache -= 2 * np; }
    __syncthreads();
    for (int j = tid; j < 2 * m; j += nt) {
        const int k = kk[live[j >> 1]];
        const StagedLine &q = rec[base + 2 * k + (j & 1)];      // a line without a window: the LineRec form
        R[j] = LineRec{q.a, q.r, q.l};
        Sp[j] = P.sp->fspan[base + 2 * k + (j & 1)];
    }
    __syncthreads();
    for (int j = tid; j < m; j += nt) {
        int lo = P.np, hi = 0;                           // a column without a window: [np, 0), empty
        for (int q = 0; q < 2; ++q) {
            const int4 sp = Sp[2 * j + q];
            if (sp.y > sp.x) { lo = min(lo, sp.x); hi = max(hi, sp.y); }
        }
        cs[j] = lo; ce[j] = hi;
    }
    __syncthreads();
    // the columns' values over their windows, once, when they fit the LDS cache (offsets in idx)
    if (tid == 0) {
        int tot = 0;
        for (int j = 0; j < m; ++j) { idx[j] = tot; tot += ce[j] > cs[j] ? ce[j] - cs[j] : 0; }
        s_int[1] = tot;
    }
    __syncthreads();
    const bool cached = s_int[1] <= ncache;
    const int wv = tid >> 6, ln = tid & 63, nw = nt >> 6;
    if (cached) {                                        // one wavefront per column
        for (int j = wv; j < m; j += nw)
            for (int i = cs[j] + ln; i < ce[j]; i += 64) colc[idx[j] + i - cs[j]] = free_column(P, R, Sp, j, i);
        __syncthreads();
    }
    auto column = [&](int j, int i) -> float { return cached ? colc[idx[j] + i - cs[j]] : free_column(P, R, Sp, j, i); };
    // normal matrix and right-hand side (FP64 sums of FP32 terms): one wavefront per row j
    // (k <= j, then b_j), its lanes over the overlap of the two windows (coalesced point
    // loads), then a wavefront sum; entries whose windows do not meet are zero
    for (int j = wv; j < m; j += nw) {
        const int cj0 = cs[j], cj1 = ce[j];
        for (int k = 0; k <= j + 1; ++k) {
            const bool rhs = k == j + 1;                 // b_j
            const int lo = rhs ? cj0 : max(cj0, cs[min(k, j)]), hi = rhs ? cj1 : min(cj1, ce[min(k, j)]);
            double a = 0.0;
            if (lo < hi) {                               // 0 <= lo, hi <= np; the same in every lane
                for (int i = lo + ln; i < hi; i += 64) {
                    const float w = pts ? wl[i] : P.obs[i].y;
                    if (w > 0.0f) a += (double)w * (double)column(j, i) * (double)(rhs ? (pts ? rl[i] : r[i]) : column(k, i));
                }
                a = wave_sum(a);
            }
            if (ln == 0) {
                if (rhs) b[j] = a;
                else H[tri_at(j, k)] = a;
            }
        }
    }
    red[tid] = cw;
    __syncthreads();
    for (int o = nt >> 1; o > 0; o >>= 1) {
        if (tid < o) red[tid] += red[tid + o];
        __syncthreads();
    }
    const double csum = red[0];
    // equilibrate; the slot's last solution is the warm start (any x >= 0 with its positive
    // set passive is a valid start for the active set, and the optimum is the same)
    double *fx = P.sp->fx + (size_t)t * nf;
    for (int j = tid; j < m; j += nt) {
        const double d = H[tri_at(j, j)];
        nrm[j] = d > 0.0 ? sqrt(d) : 0.0;
        blk[j] = d > 0.0 ? 0 : 1;
        const double w0 = blk[j] ? 0.0 : fx[live[j]] * nrm[j];
        x[j] = w0 > 0.0 ? w0 : 0.0;
        pas[j] = x[j] > 0.0 ? 1 : 0;
    }
    __syncthreads();
    for (int e = tid; e < T; e += nt) {
        int j = (int)((sqrt(8.0 * e + 1.0) - 1.0) * 0.5);
        while ((j + 1) * (j + 2) / 2 <= e) ++j;
        while (j * (j + 1) / 2 > e) --j;
        const int k = e - j * (j + 1) / 2;
        H[e] = (blk[j] || blk[k]) ? (j == k ? 1.0 : 0.0) : H[e] / (nrm[j] * nrm[k]);
    }
    for (int j = tid; j < m; j += nt) b[j] = blk[j] ? 0.0 : b[j] / nrm[j];
    __syncthreads();
    if (tid == 0) {
        double bm = 0.0;
        for (int j = 0; j < m; ++j) bm = fmax(bm, fabs(b[j]));
        s_tol = 1e-12 * fmax(bm, 1.0);
        s_int[2] = 0;                                    // solve failure
    }
    __syncthreads();
    const double tol = s_tol;
    // converged: the gradient has no positive entry off the passive set.  The iteration
    // caps, or a passive block that is not positive definite, leave the trial unsolved
    // (js/pt_gpu.js:363 solved(): its chi^2 is not used and the state's unsolved flag is set)
    const int outer_cap = P.nnls_cap > 0 ? P.nnls_cap : 3 * m + 10;
    // one wavefront, no workgroup barrier inside (plist in idx, pos in cs); in LDS at most
    // PTK_MAX_FREE_LIVE = 120 columns, beyond (up to PTK_MAX_FREE_WS_LIVE = 256) in the workspace
    if (tid < 64) {
        if (in_lds) {
            if (m <= 64) nnls_wave<1>(H, b, x, pas, blk, m, tol, outer_cap, Lf, idx, cs, s_int);
            else nnls_wave<2>(H, b, x, pas, blk, m, tol, outer_cap, Lf, idx, cs, s_int);
        } else {
            if (m <= 128) nnls_wave<2>(H, b, x, pas, blk, m, tol, outer_cap, Lf, idx, cs, s_int);
            else nnls_wave<4>(H, b, x, pas, blk, m, tol, outer_cap, Lf, idx, cs, s_int);
        }
    }
    __syncthreads();
    const bool converged = s_int[3] != 0;
    for (int gi = tid; gi < nf; gi += nt) if (!flag[gi]) fx[gi] = 0.0;
    for (int j = tid; j < m; j += nt) fx[live[j]] = pas[j] && nrm[j] > 0.0 ? x[j] / nrm[j] : 0.0;
    // chi^2 at the solution: c - sum_j x_j (2 b_j - (H x)_j), in the equilibrated frame
    double q = 0.0;
    for (int j = tid; j < m; j += nt) {
        const double xj = pas[j] ? x[j] : 0.0;
        if (xj == 0.0) continue;
        double hx = 0.0;
        for (int k = 0; k < m; ++k) hx += H[tri_at(j, k)] * (pas[k] ? x[k] : 0.0);
        q += xj * (2.0 * b[j] - hx);
    }
    __syncthreads();
    red[tid] = q;
    __syncthreads();
    for (int o = nt >> 1; o > 0; o >>= 1) {
        if (tid < o) red[tid] += red[tid + o];
        __syncthreads();
    }
    const bool ok = converged && !s_int[2];
    if (tid == 0) P.sp->unsolved[t] = ok ? 0 : 1;
    for (int k = tid; k < nsplit; k += nt)
        partial[(size_t)t * nsplit + k] = make_double2(k == 0 ? (ok ? csum - red[0] : 3e38) : 0.0, 0.0);
}

__global__ void __launch_bounds__(kNnlsThreads) k_stage_nnls(DevPlan P, const TrialImage *__restrict__ img, const StagedLine *__restrict__ rec,
                                                    double2 *partial, int nsplit, int t_off) {
    const int t = t_off + blockIdx.x;
    if (!img[t].go) return;                              // not evaluated: the trial costs 3e38
    extern __shared__ __align__(16) unsigned char dyn[];
    __shared__ double red[kNnlsThreads];
    __shared__ int s_int[4], s_wc[kNnlsThreads / 64];
    __shared__ double s_tol;
    const int nf = P.n_free, cap = min(nf, PTK_MAX_FREE_LIVE), tid = threadIdx.x, nt = blockDim.x;
    const size_t base = (size_t)t * 2 * P.nr;
    const int *free_k = P.sp->free_k;
    const int4 *fspan = P.sp->fspan;
    // the live columns of this trial: free reflections with a line window in the pattern
    // (the others have no data and stay at zero), numbered in envelope order by a prefix
    // count over the workgroup
    int *flag = reinterpret_cast<int *>(dyn + nnls_region_bytes(cap)), *kk = flag + nf;
    if (tid == 0) s_int[0] = 0;
    for (int c0 = 0; c0 < nf; c0 += nt) {
        const int gi = c0 + tid;
        bool f = false;
        if (gi < nf) {
            const int k = free_k[gi];
            const int4 s0 = fspan[base + 2 * k], s1 = fspan[base + 2 * k + 1];
            kk[gi] = k;
            f = s0.y > s0.x || s1.y > s1.x;
        }
        const unsigned long long bm = __ballot(f);
        const int w = tid >> 6, l = tid & 63;
        if (l == 0) s_wc[w] = __popcll(bm);
        __syncthreads();
        int off = s_int[0];
        for (int q = 0; q < w; ++q) off += s_wc[q];
        if (gi < nf) flag[gi] = f ? off + __popcll(bm & ((1ull << l) - 1ull)) + 1 : 0;
        __syncthreads();
        if (tid == 0) { int n = s_int[0]; for (int q = 0; q < (nt >> 6); ++q) n += s_wc[q]; s_int[0] = n; }
        __syncthreads();
    }
    if (nf <= 0) __syncthreads();                        // (no chunk: the count's reset)
    const int m = s_int[0];
    float *colc = reinterpret_cast<float *>(dyn + nnls_base_bytes(nf));
    const int ncache = nnls_cache_floats(nf);
    if (m <= cap) {
        nnls_core(P, rec, partial, nsplit, t, nnls_carve(dyn, cap), m, flag, kk, red, s_int, s_tol, colc, ncache, true);
    } else if (P.sp->nnls_ws && m <= PTK_MAX_FREE_WS_LIVE) {   // more than LDS holds: the slot's workspace
        nnls_core(P, rec, partial, nsplit, t, nnls_carve(P.sp->nnls_ws + (size_t)t * P.sp->nnls_stride, PTK_MAX_FREE_WS_LIVE), m,
                  flag, kk, red, s_int, s_tol, colc, ncache, false);
    } else {
        // more live intensities than one trial's solve takes: unsolved
        if (tid == 0) P.sp->unsolved[t] = 1;
        for (int k = tid; k < nsplit; k += nt) partial[(size_t)t * nsplit + k] = make_double2(k == 0 ? 3e38 : 0.0, 0.0);
    }
}

// the chi^2 of trial t over every dataset of the plan (js/pt_gpu.js:259-297: the targets
// of all datasets add into one cost per chain), and whether every dataset's profile is
// feasible
// feasible, and (uns) whether an intensity solve of some dataset did not converge
__device__ inline double trial_cost_all(const DevPlan &P, const TrialImage *im, const double2 *partial, int nsplit, int t,
                                        int &pok, int &uns) {
    double c = trial_cost(im, partial, nsplit, t);
    pok = im->prof_ok;
    uns = (P.has_free && im->go) ? P.sp->unsolved[t] : 0;
    for (int f = 0; f < P.n_follow; ++f) {
        const DevFollow &F = P.follow[f];
        const TrialImage *fi = (const TrialImage *)F.img + t;
        const double cf = trial_cost(fi, F.partial, F.nsplit, t);
        c = (c >= 3e38 || cf >= 3e38) ? 3e38 : c + cf;
        pok = pok && fi->prof_ok;
        if (F.unsolved && fi->go && F.unsolved[t]) uns = 1;
    }
    return c;
}

// the Metropolis step of chain t in FP64 (js/pt_gpu.js:363-364, rietveld_core.js:2213-2222),
// run by one thread; returns 0 (rejected), 1 (accepted) or 2 (accepted, new best).
// prof_ok: the trial's profiles are feasible in every dataset; uns: an intensity solve of
// the trial did not converge, which sets the chain's unsolved flag (js/pt_gpu.js:363,
// state +9; it stays set until the states are initialised again) and costs 3e38.
__device__ inline int accept_chain(const DevPlan &P, const TrialImage *im, double cost, int prof_ok, int uns, int t) {
    const int nv = P.nv;
    double *sc = P.state + (size_t)t * P.stride + 3 * nv;
    const bool boot = im->boot != 0;
    const double step = sc[2];
    const double temp = P.temp[t];
    if (uns) { sc[9] = 1.0; cost = 3e38; }
    sc[10] = cost;
    const double old = sc[0];
    const bool acc = boot || (cost < 3e38 && (cost <= old || sc[8] < (old - cost) / (2.0 * temp)));
    int r = 0;
    if (acc) {
        const bool take_best = prof_ok && (boot || cost < sc[1]);
        sc[0] = cost;
        sc[11] = prof_ok ? 1.0 : 0.0;
        if (!boot) sc[3] += 1.0;
        if (take_best) sc[1] = cost;
        r = take_best ? 2 : 1;
    }
    if (!boot) sc[4] += 1.0;
    if (!boot && !im->go && P.noeval) P.noeval[t] += 1.0;
    sc[2] = step + 1.0;
    return r;
}

// after the Metropolis step of chain t (surrogate screen on): the trial just evaluated (the
// proposal, its exact chi^2) goes into ring slot ((step - 1) chains + t) mod cap with target
// log Rwp = (log chi^2 - log sum w y^2) / 2, NaN (left out of training) for an infeasible
// trial or a chain outside the screened ones; a proposal not evaluated is not recorded.
// Inputs i = i0, i0 + di, ... of the row are written by the caller's thread.
__device__ inline void screen_record_chain(const DevPlan &P, const TrialImage *im, int t, int i0, int di) {
    if (!P.sf || im->boot || im->skipped) return;
    const ScreenFused &F = *P.sf;
    const int nv = P.nv;
    const double *st = P.state + (size_t)t * P.stride;
    const long long step = (long long)st[3 * nv + 2];
    if (step < 1) return;
    const size_t pos = (size_t)(((step - 1) * P.chains + t) % F.cap);
    for (int i = i0; i < MLP_IN; i += di) F.ring_x[pos * MLP_IN + i] = i < nv ? sf_input(F, i, st[nv + i]) : (uint16_t)0;
    if (i0 == 0) {
        const double cost = st[3 * nv + 10];
        const bool ok = t < F.cold_chains && st[3 * nv + 7] > 0.5 && cost > 0.0 && cost < 3e38;
        F.ring_t[pos] = ok ? (float)(0.5 * (log(cost) - F.log_wss)) : __int_as_float(0x7fc00000);
    }
}

// the end of one PT step for chain t: Metropolis accept, and on acceptance the proposal
// becomes the chain's state (and its best when it is a new feasible best)
__device__ inline void finish_chain(const DevPlan &P, const TrialImage *im, double cost, int prof_ok, int uns, int t) {
    const int r = accept_chain(P, im, cost, prof_ok, uns, t);
    if (r) {
        double *st = P.state + (size_t)t * P.stride;
        const int nv = P.nv;
        for (int j = 0; j < nv; ++j) {
            st[j] = st[nv + j];
            if (r == 2) st[2 * nv + j] = st[nv + j];
        }
    }
}

// chi^2 of every trial (batch mode) or the Metropolis step of every chain (PT mode)
__global__ void __launch_bounds__(64) k_stage_finish(DevPlan P, const TrialImage *__restrict__ img,
                                                     const double2 *__restrict__ partial, int nsplit, int ntrials,
                                                     double *chi2, int pt_mode, int t_off) {
    const int tl = blockIdx.x * blockDim.x + threadIdx.x;
    if (tl >= ntrials) return;
    const int t = t_off + tl;
    int pok, uns;
    const double cost = trial_cost_all(P, img + t, partial, nsplit, t, pok, uns);
    // batch mode: NaN marks an unsolved trial for the host (a cost is otherwise finite)
    if (!pt_mode) { chi2[t] = uns ? __builtin_nan("") : cost; return; }
    finish_chain(P, img + t, cost, pok, uns, t);
    screen_record_chain(P, img + t, t, 0, 1);
}

// the Metropolis step of chain t, then its next proposal and setup: the finish of one PT
// step fused with the setup of the next (one launch less per step)
template <bool SEP = true>
__device__ __forceinline__ void finish_setup_body(const DevPlan &P, TrialImage *img, const double2 *__restrict__ partial,
                                         int nsplit, int t, SetupShared &sh, double *prop, int &s_acc) {
    const int nv = P.nv;
    if (threadIdx.x == 0) {
        int pok, uns;
        const double cost = trial_cost_all(P, img + t, partial, nsplit, t, pok, uns);
        s_acc = accept_chain(P, img + t, cost, pok, uns, t);
    }
    __syncthreads();
    const int r = s_acc;
    if (r) {
        double *st = P.state + (size_t)t * P.stride;
        for (int j = threadIdx.x; j < nv; j += blockDim.x) {
            const double v = st[nv + j];
            st[j] = v;
            if (r == 2) st[2 * nv + j] = v;
        }
    }
    screen_record_chain(P, img + t, t, threadIdx.x, blockDim.x);
    __syncthreads();
    setup_body<SEP>(P, nullptr, img, t, 1, 0, sh, prop);
}

// one instance per plan kind: a kernel holding both runs the plain one slower
template <bool SEP>
__global__ void __launch_bounds__(64) k_stage_finish_setup(DevPlan P, TrialImage *img, const double2 *__restrict__ partial,
                                                           int nsplit, int t_off) {
    __shared__ SetupShared sh;
    __shared__ double prop[PTK_MAX_VARS];
    __shared__ int s_acc;
    finish_setup_body<SEP>(P, img, partial, nsplit, t_off + blockIdx.x, sh, prop, s_acc);
}

// The finish of one PT step, the setup of the next and its lines stage in one workgroup
// per chain, for a plan whose reflections fit one lines workgroup (P.nr <= 512), without
// the surrogate's skip test: one launch and one
// round trip of the trial image less per step.  The lines see the setup's values in LDS;
// the image (which the accumulate stage reads) is still written.  CHUNK as in
// k_stage_lines.
template <bool SEP, bool CHUNK>
__global__ void __launch_bounds__(512) k_stage_finish_lines(DevPlan P, TrialImage *img, const double2 *__restrict__ partial,
                                                            int nsplit, StagedLine *rec, int ntrials, int t_off) {
    __shared__ Shared sh;
    __shared__ double prop[PTK_MAX_VARS];
    __shared__ int s_acc;
    __shared__ int gr[PTK_MAX_GROUPS][2];
    __shared__ float sfpart[CHUNK ? 256 * kSfSlots : 1];
    int t, u;
    if (!xcd_unit(blockIdx.x, 1, ntrials, t, u)) return;
    finish_setup_body<SEP>(P, img, partial, nsplit, t_off + t, sh, prop, s_acc);
    __syncthreads();
    lines_body<SEP>(P, img, rec, t_off + t, 0, sh, gr, CHUNK ? sfpart : nullptr, true);
}

// ---- replica exchange (js/pt_gpu.js:365, rietveld_core.js:2229-2236) ------------------
// pair `pair` of the sweep's parity: chains c = 2 pair + parity and c + 1
__device__ inline void exchange_body(const DevPlan &P, int pair) {
    const int nv = P.nv;
    // sweeps completed, from chain 0's step counter (js/pt_gpu.js:365 parity rule)
    const double done = P.state[3 * nv + 2] / (double)P.steps_per_sweep;
    const long long sw = (long long)done;
    const int parity = (int)(((sw - 1) % 2 + 2) % 2);
    const int c = 2 * pair + parity;
    if (c + 1 >= P.chains) return;
    double *a = P.state + (size_t)c * P.stride, *b = P.state + (size_t)(c + 1) * P.stride;
    const double t0 = P.temp[c], t1 = P.temp[c + 1];
    const double la = 0.5 * (1.0 / t0 - 1.0 / t1) * (a[3 * nv] - b[3 * nv]);
    a[3 * nv + 6] += 1.0;
    bool swp = la >= 0.0;
    if (!swp) swp = log(uniform01(P.seed ^ 0x5bd1e995u, c, (unsigned long long)sw, 7)) < la;
    if (swp) {
        for (int j = 0; j < nv; ++j) { double v = a[j]; a[j] = b[j]; b[j] = v; }
        double v = a[3 * nv]; a[3 * nv] = b[3 * nv]; b[3 * nv] = v;
        v = a[3 * nv + 11]; a[3 * nv + 11] = b[3 * nv + 11]; b[3 * nv + 11] = v;
        a[3 * nv + 5] += 1.0;
    }
}

__global__ void k_exchange(DevPlan P) { exchange_body(P, blockIdx.x * blockDim.x + threadIdx.x); }

// ---- groups: independent problems sharing every launch --------------------------------
// Each stage of a group runs as one launch over all members.  Workgroup b takes its
// (member, chain, unit) from a table built on the host: member-major for the per-chain
// stages; for the lines and accumulate stages grouped so that every unit of one trial
// lands on the same XCD (workgroup b runs on XCD b mod 8, so the trial's line records are
// written and read through one L2), heaviest trials first.  A member entry of -1 is an
// empty slot.
__global__ void __launch_bounds__(64) k_group_setup(const DevPlan *__restrict__ plans, const DevMember *__restrict__ mem,
                                                    const int2 *__restrict__ tab, int bootstrap) {
    __shared__ Shared sh;
    __shared__ double prop[PTK_MAX_VARS];
    const int2 w = tab[blockIdx.x];
    const DevPlan &P = plans[w.x];
    if (P.any_sep) setup_body<true>(P, nullptr, (TrialImage *)mem[w.x].img, w.y, 1, bootstrap, sh, prop);
    else setup_body<false>(P, nullptr, (TrialImage *)mem[w.x].img, w.y, 1, bootstrap, sh, prop);
}

__global__ void __launch_bounds__(256) k_group_lines(const DevPlan *__restrict__ plans, const DevMember *__restrict__ mem,
                                                     const int4 *__restrict__ tab) {
    __shared__ Shared sh;
    __shared__ int gr[PTK_MAX_GROUPS][2];
    const int4 w = tab[blockIdx.x];
    if (w.x < 0) return;
    const DevPlan &P = plans[w.x];
    const DevMember &M = mem[w.x];
    // a group launch is throughput-bound: chunked sums stay in the reflection's thread
    if (P.any_sep) lines_body<true>(P, (TrialImage *)M.img, M.rec, w.y, w.z, sh, gr, nullptr);
    else lines_body<false>(P, (TrialImage *)M.img, M.rec, w.y, w.z, sh, gr, nullptr);
}

template <int TPW>
__global__ void __launch_bounds__(256) k_group_accum(const DevPlan *__restrict__ plans, const DevMember *__restrict__ mem,
                                                     const int4 *__restrict__ tab) {
    __shared__ AccShared S;
    const int4 w = tab[blockIdx.x];
    if (w.x < 0) return;
    const DevPlan &P = plans[w.x];
    const DevMember &M = mem[w.x];
    accum_body<TPW>(P, (const TrialImage *)M.img, M.rec, M.partial, M.nsplit[0], w.y, w.z, S);
}

// one thread per chain: the Metropolis step (pt_mode) or the chi^2 of the chain's vector
__global__ void __launch_bounds__(64) k_group_finish(const DevPlan *__restrict__ plans, const DevMember *__restrict__ mem,
                                                     const int2 *__restrict__ tab, int nchains, int pt_mode) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nchains) return;
    const int2 w = tab[b];
    const DevPlan &P = plans[w.x];
    const DevMember &M = mem[w.x];
    const TrialImage *img = (const TrialImage *)M.img;
    const double cost = trial_cost(img + w.y, M.partial, M.nsplit[0], w.y);
    if (pt_mode) finish_chain(P, img + w.y, cost, img[w.y].prof_ok, 0, w.y);
    else M.chi2[w.y] = cost;
}

__global__ void __launch_bounds__(64) k_group_finish_setup(const DevPlan *__restrict__ plans,
                                                           const DevMember *__restrict__ mem,
                                                           const int2 *__restrict__ tab) {
    __shared__ Shared sh;
    __shared__ double prop[PTK_MAX_VARS];
    __shared__ int s_acc;
    const int2 w = tab[blockIdx.x];
    const DevPlan &P = plans[w.x];
    const DevMember &M = mem[w.x];
    if (P.any_sep) finish_setup_body<true>(P, (TrialImage *)M.img, M.partial, M.nsplit[0], w.y, sh, prop, s_acc);
    else finish_setup_body<false>(P, (TrialImage *)M.img, M.partial, M.nsplit[0], w.y, sh, prop, s_acc);
}

// one thread per (member, pair)
__global__ void __launch_bounds__(64) k_group_exchange(const DevPlan *__restrict__ plans, const int2 *__restrict__ tab,
                                                       int npairs) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= npairs) return;
    const int2 w = tab[b];
    exchange_body(plans[w.x], w.y);
}

// ---- plain single-trial path: setup, lines, points, reduce ---------------------------
// scratch layout (global): [0] Shared image of the trial (params, metrics, atoms, profile)
struct PlainScratch {
    Shared sh;
};

template <bool SEP>
__global__ void __launch_bounds__(256) k_plain_setup(DevPlan P, const double *__restrict__ X, PlainScratch *scr) {
    __shared__ Shared sh;
    stage_reps(P, sh);
    __syncthreads();
    materialize(P, X, sh);
    __syncthreads();
    trial_setup<SEP>(P, sh);
    // copy the shared image out (byte-wise, all threads)
    const int n = (int)(sizeof(Shared) / 4);
    const int *src = reinterpret_cast<const int *>(&sh);
    int *dst = reinterpret_cast<int *>(&scr->sh);
    for (int i = threadIdx.x; i < n; i += blockDim.x) dst[i] = src[i];
}

template <bool SEP>
__global__ void __launch_bounds__(256) k_plain_lines(DevPlan P, const PlainScratch *__restrict__ scr, LineRec *rec, int4 *span) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= P.nr) return;
    if (!scr->sh.cell_ok) { span[2 * k] = make_int4(0, 0, 0, 0); span[2 * k + 1] = make_int4(0, 0, 0, 0); return; }
    build_reflection<SEP>(P, scr->sh, k, rec + 2 * k, span + 2 * k);
}

// each workgroup: 4 wavefronts = 4 tiles of 64 points; line records streamed through LDS
__global__ void __launch_bounds__(256) k_plain_points(DevPlan P, const PlainScratch *__restrict__ scr,
                                                     const LineRec *__restrict__ grec, const int4 *__restrict__ gspan,
                                                     double *partial, int ch) {
    __shared__ Shared sh;
    extern __shared__ __align__(16) unsigned char dyn[];
    LineRec *rec = reinterpret_cast<LineRec *>(dyn);
    int4 *span = reinterpret_cast<int4 *>(dyn + sizeof(LineRec) * ch);
    float *ytile = reinterpret_cast<float *>(dyn + (sizeof(LineRec) + sizeof(int4)) * ch);
    // bring the trial image (bkg coefficients) in
    for (int i = threadIdx.x; i < PTK_MAX_BKG_F; i += blockDim.x) sh.bkg[i] = scr->sh.bkg[i];
    const int tile = blockIdx.x * 4 + wave_id();
    const int p0 = blockIdx.x * 256;
    ytile[threadIdx.x] = 0.0f;
    const int nl = 2 * P.nr;
    for (int l0 = 0; l0 < nl; l0 += ch) {
        int n = min(ch, nl - l0);
        __syncthreads();
        for (int i = threadIdx.x; i < n; i += blockDim.x) { rec[i] = grec[l0 + i]; span[i] = gspan[l0 + i]; }
        __syncthreads();
        // accumulate into ytile, indexed relative to p0
        accumulate_tiles(P, rec, span, n, ytile - p0, tile, 1 << 30);
    }
    __syncthreads();
    double c = 3e38;
    if (scr->sh.cell_ok) c = chi2_block(P, sh, ytile - p0, p0, min(P.np, p0 + 256));
    if (threadIdx.x == 0) partial[blockIdx.x] = c;
}

__global__ void k_plain_reduce(const double *partial, int n, double *out) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        double s = 0.0;
        for (int i = 0; i < n; ++i) s += partial[i];
        *out = (isfinite(s) && s < 3e38) ? s : 3e38;
    }
}


// ---- shared-profile matrix form: Y = P I on MFMA ---------------------------------------
// When only structure parameters (coordinates, occupancies, B, the phase scale) differ
// between trials, the profile of every reflection is fixed and the pattern is
//   y(t_i) = bkg(t_i) + sum_k P[i, k] I[k, trial],
// P[i, k] the unit-height column of reflection k (Kalpha1 + weighted Kalpha2) and
// I[k, trial] = S m Lp PO |F|^2 / area.  P and I are held as FP16 hi/lo pairs (P scaled
// by 2^12, each trial's I by a power of two), the products hi*hi + hi*lo + lo*hi are
// accumulated in FP32 by gfx950's v_mfma_f32_16x16x32_f16 (K = 32: twice the work per
// instruction of the 16x16x16 shape), and the weighted chi^2 is reduced in the epilogue
// -- Y is never stored.  Only the band of reflections whose windows meet a 128-point row
// tile is multiplied.  The same kernels also run the pairs as BF16
// (v_mfma_f32_16x16x32_bf16), the precision comparison of PT_FORM_MATRIX_BF16.

typedef _Float16 half8_t __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float floatx4_t __attribute__((ext_vector_type(4)));
constexpr float kPScale = 4096.0f;   // 2^12

// v as a hi/lo pair of 16-bit floats, FP16 (22 significant bits together) or BF16 (16):
// hi = round(v), lo = round(v - hi), round to nearest even
__device__ __forceinline__ unsigned short bf16_bits(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
__device__ __forceinline__ void split_pair(float v, bool bf16, unsigned short &hi, unsigned short &lo) {
    if (bf16) {
        hi = bf16_bits(v);
        lo = bf16_bits(v - __uint_as_float((uint32_t)hi << 16));
    } else {
        const _Float16 h = (_Float16)v;
        hi = __builtin_bit_cast(unsigned short, h);
        lo = __builtin_bit_cast(unsigned short, (_Float16)(v - (float)h));
    }
}

// unit-height lines of the start vector and the per-reflection height factors
__global__ void __launch_bounds__(128) k_matrix_columns(DevPlan P, const TrialImage *__restrict__ img, LineRec *rec,
                                                        int4 *span, float *factor) {
    __shared__ Shared sh;
    stage_reps(P, sh);
    image_load(P, img, sh);
    __syncthreads();
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= P.nr) return;
    build_reflection(P, sh, k, rec + 2 * k, span + 2 * k, factor);
}

// column k of P: Kalpha1 then Kalpha2 line values over the union of their windows
__global__ void __launch_bounds__(256) k_matrix_fill(DevPlan P, const LineRec *__restrict__ rec,
                                                     const int4 *__restrict__ span, unsigned short *phi,
                                                     unsigned short *plo, int kpad, int bf16) {
    const int k = blockIdx.x;
    const int4 s1 = span[2 * k], s2 = span[2 * k + 1];
    int lo = P.np, hi = 0;
    if (s1.y > s1.x) { lo = min(lo, s1.x); hi = max(hi, s1.y); }
    if (s2.y > s2.x) { lo = min(lo, s2.x); hi = max(hi, s2.y); }
    for (int i = lo + threadIdx.x; i < hi; i += blockDim.x) {
        const float2 t = P.tff[i];
        float v = 0.0f;
        if (i >= s1.x && i < s1.y)
            v += line_point(rec[2 * k], s1.w & 1, __int_as_float(s1.z), t, true);
        if (i >= s2.x && i < s2.y)
            v += line_point(rec[2 * k + 1], s2.w & 1, __int_as_float(s2.z), t, true);
        unsigned short h, l;
        split_pair(v * kPScale, bf16 != 0, h, l);
        phi[(size_t)i * kpad + k] = h;
        plo[(size_t)i * kpad + k] = l;
    }
}

// Per-reflection constants of the shared-profile intensities, formed once from the start
// vector's image (every trial of the form has the start vector's cell, so 1/d^2 and the
// form factors are shared):
//   hr[r kpad + k]  = (h R_r, h . t_r) in FP64 for representative r of k's phase (zero past
//                     the phase's n_reps),
//   f0w[ga kpad + k] = (f0 + f') w0 of atom ga at k (the neutron length times w0), zero for
//                     the atoms of other phases,
//   qk[k]           = (log2(e) s^2 / 4, the Debye-Waller exponent per unit B in exp2 form;
//                     the height factor m Lp PO / area, or 0 when k's phase cell is invalid),
//   *go             = the start vector's cell validity (a trial of the form goes when it does).
__global__ void __launch_bounds__(256) k_matrix_tables(DevPlan P, const TrialImage *__restrict__ img,
                                                       const float *__restrict__ factor, double4 *hr, float *f0w, float2 *qk,
                                                       int *go, int kpad, int nrep_max) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k == 0) *go = img->go;
    if (k >= kpad) return;
    const int na = total_atoms(P);
    if (k >= P.nr) {
        qk[k] = make_float2(0.f, 0.f);
        for (int r = 0; r < nrep_max; ++r) hr[(size_t)r * kpad + k] = make_double4(0.0, 0.0, 0.0, 0.0);
        for (int ga = 0; ga < na; ++ga) f0w[(size_t)ga * kpad + k] = 0.f;
        return;
    }
    int q = 0;
    while (q + 1 < P.n_phases && k >= P.phase[q + 1].refl_off) ++q;
    const DevPhase &ph = P.phase[q];
    const int4 hk = P.hklm[k];
    const double *f = img->metric[q];
    const double h = hk.x, kx = hk.y, l = hk.z;
    const double qv = h * h * f[0] + kx * kx * f[1] + l * l * f[2] + h * kx * f[3] + kx * l * f[4] + h * l * f[5];
    const float q4 = 0.25f * (float)qv;
    qk[k] = make_float2(q4 * 1.44269504f, img->phase_ok[q] ? factor[k] : 0.0f);
    for (int r = 0; r < nrep_max; ++r) {
        double4 e = make_double4(0.0, 0.0, 0.0, 0.0);
        if (r < ph.n_reps) {
            const int *R = P.rep_r + 9 * (ph.rep_off + r);
            const double *t = P.rep_t + 3 * (ph.rep_off + r);
            e.x = hk.x * R[0] + hk.y * R[3] + hk.z * R[6];
            e.y = hk.x * R[1] + hk.y * R[4] + hk.z * R[7];
            e.z = hk.x * R[2] + hk.y * R[5] + hk.z * R[8];
            e.w = (double)hk.x * t[0] + (double)hk.y * t[1] + (double)hk.z * t[2];
        }
        hr[(size_t)r * kpad + k] = e;
    }
    for (int ga = 0; ga < na; ++ga) {
        float v = 0.f;
        if (ga >= ph.atom_off && ga < ph.atom_off + ph.n_atoms) {
            const DevAtom &A = P.atoms[ga];
            float f0;
            if (P.radiation == 1) f0 = A.b_neutron;
            else {
                f0 = A.ff_c;
                for (int g = 0; g < A.n_gauss; ++g) f0 += A.ff_a[g] * __expf(-A.ff_b[g] * q4);
                f0 += A.fp;
            }
            v = f0 * A.w0;
        }
        f0w[(size_t)ga * kpad + k] = v;
    }
}

// model parameter pi of the trial vector x: base + M x (the materialize sum)
__device__ __forceinline__ double matrix_param(const DevPlan &P, const double *x, int pi) {
    double v = P.base[pi];
    for (int t = P.term_off[pi]; t < P.term_off[pi + 1]; ++t) v += P.term_coeff[t] * x[P.term_var[t]];
    return v;
}

// I[k, t] = S m Lp PO |F|^2 / area of every reflection k of kMxT trials per workgroup,
// formed straight from the trial vectors (only the structure parameters are materialised:
// positions, occupancies, B, the phase scales), as FP16 (or BF16) hi/lo pairs scaled by
// 2^-e_t with max_k |I| 2^-e_t <= 2^14; iscale[t] = 2^e_t.  A thread takes reflections k,
// k + blockDim, ... for the workgroup's trials (blockDim is sized so every thread has the
// same count and the whole grid is resident at once); per atom it sums the
// representatives' cos (and sin) of the turn h R r + h.t, FP64 phases reduced to [0, 1),
// with h R and h.t read from the table four representatives at a time (coalesced 32-byte
// loads, in flight together, shared by the trials); the atoms' weights occ dw f0 w0 then
// multiply the sums (structure_f2's order: atoms outer, representatives inner).  The
// intensities wait in LDS for the trial's maximum, then go out as pairs.
//
// Dynamic LDS, na atoms, nv variables: xs [T][nv] f64 (the trial vectors), apos
// [T][na][3] f64, aocc [T][na], aB [T][na], afi [na] (f'' w0),
// S [T][PTK_MAX_PHASES], wmax [T][8], ival [T][kpad] floats.
constexpr int kMxT = 4;                  // trials per workgroup
__host__ __device__ constexpr size_t matrix_intensity_lds(int T, int na, int nv, int kpad) {
    return sizeof(double) * ((size_t)T * nv + 3 * (size_t)T * na) +
           sizeof(float) * ((size_t)2 * T * na + na + (size_t)T * PTK_MAX_PHASES + 8 * T + (size_t)T * kpad);
}

// dwk: the Debye-Waller exponent per unit B in exp2 form (log2(e) s^2 / 4)
template <int T, bool CENTRO>
__device__ __forceinline__ void matrix_f2(const DevPhase &ph, const double *apos, const float *aocc, const float *aB,
                                          const float *afi, int na, const double4 *__restrict__ hr, const float *__restrict__ f0w,
                                          int k, int kpad, float dwk, float *f2) {
    float Ar[T], Br[T], Cr[T], Dr[T];
#pragma unroll
    for (int tt = 0; tt < T; ++tt) Ar[tt] = Br[tt] = Cr[tt] = Dr[tt] = 0.f;
    const int nrep = ph.n_reps;
    // with four trials, the first four representatives' h R, h.t are loaded once for all
    // atoms (with more trials the registers go to the trials)
    constexpr bool kHoist = T <= 4;
    double4 e4[4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
        e4[i] = kHoist && 4 <= nrep ? hr[(size_t)i * kpad + k] : make_double4(0.0, 0.0, 0.0, 0.0);
    for (int j = 0; j < ph.n_atoms; ++j) {
        const int ga = ph.atom_off + j;
        const float fw = f0w[(size_t)ga * kpad + k], fi = afi[ga];
        double px[T], py[T], pz[T];
        float ca[T], sa[T];
#pragma unroll
        for (int tt = 0; tt < T; ++tt) {
            const double *a = apos + 3 * (tt * na + ga);
            px[tt] = a[0]; py[tt] = a[1]; pz[tt] = a[2];
            ca[tt] = sa[tt] = 0.f;
        }
        // whole groups of four representatives: sixteen independent phase chains, summed
        // afterwards in representative order
        int r = 0;
        for (; r + 4 <= nrep; r += 4) {
            double4 e[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) e[i] = kHoist && r == 0 ? e4[i] : hr[(size_t)(r + i) * kpad + k];
            float cv[4][T], sv[4][T];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int tt = 0; tt < T; ++tt) {
                    const double phs = fma(e[i].x, px[tt], fma(e[i].y, py[tt], fma(e[i].z, pz[tt], e[i].w)));
                    const float fr = (float)__builtin_amdgcn_fract(phs);
                    cv[i][tt] = __builtin_amdgcn_cosf(fr);
                    sv[i][tt] = CENTRO ? 0.f : __builtin_amdgcn_sinf(fr);
                }
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int tt = 0; tt < T; ++tt) {
                    ca[tt] = r == 0 && i == 0 ? cv[i][tt] : ca[tt] + cv[i][tt];
                    if (!CENTRO) sa[tt] = r == 0 && i == 0 ? sv[i][tt] : sa[tt] + sv[i][tt];
                }
        }
        for (; r < nrep; ++r) {
            const double4 e = hr[(size_t)r * kpad + k];
#pragma unroll
            for (int tt = 0; tt < T; ++tt) {
                const double phs = fma(e.x, px[tt], fma(e.y, py[tt], fma(e.z, pz[tt], e.w)));
                const float fr = (float)__builtin_amdgcn_fract(phs);
                ca[tt] += __builtin_amdgcn_cosf(fr);
                if (!CENTRO) sa[tt] += __builtin_amdgcn_sinf(fr);
            }
        }
        float od[T];
#pragma unroll
        for (int tt = 0; tt < T; ++tt) {
            od[tt] = aocc[tt * na + ga] * __builtin_amdgcn_exp2f(-aB[tt * na + ga] * dwk);
            const float wj = od[tt] * fw;
            Ar[tt] += wj * ca[tt];
            if (!CENTRO) Br[tt] += wj * sa[tt];
        }
        if (fi != 0.0f) {               // the f'' terms (anomalous scattering)
#pragma unroll
            for (int tt = 0; tt < T; ++tt) {
                const float wi = od[tt] * fi;
                Cr[tt] += wi * ca[tt];
                if (!CENTRO) Dr[tt] += wi * sa[tt];
            }
        }
    }
#pragma unroll
    for (int tt = 0; tt < T; ++tt)
        f2[tt] = CENTRO ? 4.0f * (Ar[tt] * Ar[tt] + Cr[tt] * Cr[tt])
                        : Ar[tt] * Ar[tt] + Br[tt] * Br[tt] + Cr[tt] * Cr[tt] + Dr[tt] * Dr[tt];
}

// X: the chunk's n trial vectors (nv each)
template <int T>
__global__ void __launch_bounds__(256) k_matrix_intensity(DevPlan P, const double *__restrict__ X,
                                                          const double4 *__restrict__ hr, const float *__restrict__ f0w,
                                                          const float2 *__restrict__ qk, unsigned short *ihi, unsigned short *ilo,
                                                          float *iscale, int kpad, int n, int bf16) {
    extern __shared__ __align__(16) unsigned char mx_raw[];
    const int tid = threadIdx.x, nt = blockDim.x, tb = blockIdx.x * T;
    const int na = total_atoms(P);
    double *xs = reinterpret_cast<double *>(mx_raw);
    double *apos = xs + (size_t)T * P.nv;
    float *aocc = reinterpret_cast<float *>(apos + 3 * (size_t)T * na);
    float *aB = aocc + T * na, *afi = aB + T * na, *Sx = afi + na, *wmax = Sx + T * PTK_MAX_PHASES;
    float *ival = wmax + 8 * T;
    // the workgroup's trial vectors into LDS first (their loads do not wait on the
    // parameter map's)
    for (int i = tid; i < T * P.nv; i += nt) {
        const int tt = i / P.nv;
        xs[i] = tb + tt < n ? X[(size_t)tb * P.nv + i] : 0.0;
    }
    __syncthreads();
    // the trials' structure parameters, one (trial, atom, slot) per thread: the coordinate
    // slots wait in apos, occupancy and B go to aocc / aB
    for (int i = tid; i < T * na * 5; i += nt) {
        const int tt = i / (na * 5), ga = (i / 5) % na, f = i % 5;
        double v = 0.0;
        if (tb + tt < n) {
            int q = 0;
            while (q + 1 < P.n_phases && ga >= P.phase[q + 1].atom_off) ++q;
            const DevPhase &ph = P.phase[q];
            if (f >= 3 || (P.atoms[ga].free_mask & (1 << f)))
                v = matrix_param(P, xs + (size_t)tt * P.nv, ph.param_base + PTK_PH_ATOM0 + 5 * (ga - ph.atom_off) + f);
        }
        if (f < 3) apos[3 * ((size_t)tt * na + ga) + f] = v;
        else if (f == 3) aocc[tt * na + ga] = (float)v;
        else aB[tt * na + ga] = (float)v;
    }
    for (int i = tid; i < T * P.n_phases; i += nt) {
        const int tt = i / P.n_phases, q = i % P.n_phases;
        float S = 0.f;
        if (tb + tt < n) {
            S = (float)matrix_param(P, xs + (size_t)tt * P.nv, P.phase[q].param_base + 6);
            if (!(S != 0.0f) || isnan(S)) S = 0.0f;
        }
        Sx[tt * PTK_MAX_PHASES + q] = S;
    }
    for (int ga = tid; ga < na; ga += nt) {
        const DevAtom &A = P.atoms[ga];
        afi[ga] = (P.radiation == 1 || !(A.fpp != 0.f)) ? 0.f : A.fpp * A.w0;
    }
    __syncthreads();
    // positions r0 + sum over the free slots of v basis (trial_setup's arithmetic)
    for (int i = tid; i < T * na; i += nt) {
        const int ga = i % na;
        double *a = apos + 3 * (size_t)i;
        const DevAtom &A = P.atoms[ga];
        double r[3] = {A.r0[0], A.r0[1], A.r0[2]};
        for (int f = 0; f < 3; ++f)
            if (A.free_mask & (1 << f)) {
                const double v = a[f];
                r[0] += v * A.basis[f][0]; r[1] += v * A.basis[f][1]; r[2] += v * A.basis[f][2];
            }
        a[0] = r[0]; a[1] = r[1]; a[2] = r[2];
    }
    __syncthreads();
    float mx[T];
#pragma unroll
    for (int tt = 0; tt < T; ++tt) mx[tt] = 0.f;
    for (int k = tid; k < P.nr; k += nt) {
        const float2 qf = qk[k];
        float v[T];
#pragma unroll
        for (int tt = 0; tt < T; ++tt) v[tt] = 0.f;
        if (qf.y != 0.0f) {
            int q = 0;
            while (q + 1 < P.n_phases && k >= P.phase[q + 1].refl_off) ++q;
            const DevPhase &ph = P.phase[q];
            float f2[T];
            if (ph.centro) matrix_f2<T, true>(ph, apos, aocc, aB, afi, na, hr, f0w, k, kpad, qf.x, f2);
            else matrix_f2<T, false>(ph, apos, aocc, aB, afi, na, hr, f0w, k, kpad, qf.x, f2);
#pragma unroll
            for (int tt = 0; tt < T; ++tt) {
                v[tt] = Sx[tt * PTK_MAX_PHASES + q] * (qf.y * f2[tt]);
                mx[tt] = fmaxf(mx[tt], fabsf(v[tt]));
            }
        }
#pragma unroll
        for (int tt = 0; tt < T; ++tt) ival[tt * kpad + k] = v[tt];
    }
    // per trial: a power-of-two scale that keeps the largest intensity near 2^14
#pragma unroll
    for (int tt = 0; tt < T; ++tt) {
        float m = mx[tt];
        for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
        if ((tid & 63) == 0) wmax[tt * 8 + (tid >> 6)] = m;
    }
    __syncthreads();
    if (tid < T) {
        float m = 0.0f;
        for (int w = 0; w < (nt >> 6); ++w) m = fmaxf(m, wmax[tid * 8 + w]);
        int e = 0;
        if (m > 0.0f && isfinite(m)) { frexpf(m, &e); e -= 14; }
        wmax[tid * 8] = ldexpf(1.0f, e);
    }
    __syncthreads();
    for (int tt = 0; tt < T; ++tt) {
        const int t = tb + tt;
        if (t >= n) break;
        const float sc = wmax[tt * 8], inv = 1.0f / sc;
        for (int k = tid; k < kpad; k += nt) {
            unsigned short hv, lv;
            split_pair(k < P.nr ? ival[tt * kpad + k] * inv : 0.0f, bf16 != 0, hv, lv);
            ihi[(size_t)t * kpad + k] = hv;
            ilo[(size_t)t * kpad + k] = lv;
        }
        if (tid == 0) iscale[t] = sc;
    }
}

// 16-bit operand quad as the MFMA input type
template <bool BF16>
struct Quad;
template <>
struct Quad<false> {
    typedef half8_t T;
    static __device__ __forceinline__ T from(uint4 u) { return __builtin_bit_cast(half8_t, u); }
    static __device__ __forceinline__ floatx4_t mfma(T a, T b, floatx4_t c) {
        return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
    }
};
template <>
struct Quad<true> {
    typedef bf16x8_t T;
    static __device__ __forceinline__ T from(uint4 u) { return __builtin_bit_cast(bf16x8_t, u); }
    static __device__ __forceinline__ floatx4_t mfma(T a, T b, floatx4_t c) {
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
    }
};

constexpr int kMxK = 32;                 // reflections per staged slice: one MFMA K step
constexpr int kMxLd = kMxK + 8;          // LDS row stride (16-bit elements), padded; rows 16-byte aligned

// chi^2 partials of a 128-point x 64-trial tile.  The trial operand (I hi/lo, 64 trials x
// 32 reflections per slice) is staged through LDS, double buffered, and shared by the four
// wavefronts; wavefront w multiplies its 32 rows [i0 + 32 w, i0 + 32 w + 32) (two 16-row
// blocks, P read from global) by the four 16-trial column blocks over the tile's band.
// MFMA operand layout (16x16x32): lane l supplies A[l % 16][8 (l / 16) + e] and
// B[8 (l / 16) + e][l % 16], e < 8 (16 bytes of consecutive reflections each), and holds
// D[4 (l / 16) + e][l % 16], e < 4.
template <bool BF16, int TT>
__global__ void __launch_bounds__(256) k_matrix_chi2(DevPlan P, const unsigned short *__restrict__ phi,
                                                     const unsigned short *__restrict__ plo,
                                                     const unsigned short *__restrict__ ihi,
                                                     const unsigned short *__restrict__ ilo,
                                                     const float *__restrict__ iscale, const int2 *__restrict__ band,
                                                     const float2 *__restrict__ yw, float *partial, int kpad, int tpad,
                                                     int row_tiles, double *yout, int ny) {
    typedef Quad<BF16> Q;
    constexpr int NT = 64 * TT;          // trials per workgroup tile
    __shared__ __align__(16) unsigned short sB[2][2][NT][kMxLd];   // [buffer][hi, lo][trial][k]
    __shared__ float red[4][NT];
    // XCD-aware tiles: workgroups go to the eight XCDs round robin, and XCD x takes the
    // trial tiles x, x + 8, ..., each with all its row tiles in a row: a trial tile's
    // operand is read from HBM once and stays in that XCD's L2 while its row tiles pass,
    // and P's band (the only part of P any tile reads) stays resident in every XCD's L2
    const int xcd = blockIdx.x & 7, kx = blockIdx.x >> 3, ntt = tpad / NT;
    const int ntx = (ntt - xcd + 7) / 8;
    if (kx >= ntx * row_tiles) return;
    const int rt = kx % row_tiles, j0 = (xcd + 8 * (kx / row_tiles)) * NT;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int i0 = rt * 128 + wave * 32;
    const int2 b = band[rt];
    const int li = lane & 15, kq = 8 * (lane >> 4), dq = 4 * (lane >> 4);
    float2 ow[2][4];
#pragma unroll
    for (int rb = 0; rb < 2; ++rb)
#pragma unroll
        for (int e = 0; e < 4; ++e) ow[rb][e] = yw[i0 + 16 * rb + dq + e];
    floatx4_t acc[2][4 * TT];
#pragma unroll
    for (int rb = 0; rb < 2; ++rb)
#pragma unroll
        for (int nb = 0; nb < 4 * TT; ++nb) acc[rb][nb] = (floatx4_t){0.f, 0.f, 0.f, 0.f};
    // staging: thread t moves trials j0 + t / 4 (+ 64 u), reflections [8 (t % 4), 8 (t % 4)
    // + 8) of the slice, hi and lo (16 bytes each)
    const int sc_col = tid >> 2, sc_k = 8 * (tid & 3);
    const size_t src = (size_t)(j0 + sc_col) * kpad + sc_k;
    uint4 nh[TT], nl[TT];
#pragma unroll
    for (int u = 0; u < TT; ++u) {
        nh[u] = make_uint4(0u, 0u, 0u, 0u); nl[u] = nh[u];
        if (b.x < b.y) {
            nh[u] = *reinterpret_cast<const uint4 *>(ihi + src + (size_t)64 * u * kpad + b.x);
            nl[u] = *reinterpret_cast<const uint4 *>(ilo + src + (size_t)64 * u * kpad + b.x);
        }
    }
    int buf = 0;
    for (int k0 = b.x; k0 < b.y; k0 += kMxK) {
        // the slice k0 into LDS, then the next slice's loads in flight while computing
#pragma unroll
        for (int u = 0; u < TT; ++u) {
            *reinterpret_cast<uint4 *>(&sB[buf][0][sc_col + 64 * u][sc_k]) = nh[u];
            *reinterpret_cast<uint4 *>(&sB[buf][1][sc_col + 64 * u][sc_k]) = nl[u];
        }
        __syncthreads();
        if (k0 + kMxK < b.y) {
#pragma unroll
            for (int u = 0; u < TT; ++u) {
                nh[u] = *reinterpret_cast<const uint4 *>(ihi + src + (size_t)64 * u * kpad + k0 + kMxK);
                nl[u] = *reinterpret_cast<const uint4 *>(ilo + src + (size_t)64 * u * kpad + k0 + kMxK);
            }
        }
        {
            // the band ends on a slice boundary (bands are whole 32-reflection slices)
            typename Q::T ah[2], al[2];
#pragma unroll
            for (int rb = 0; rb < 2; ++rb) {
                const size_t a = (size_t)(i0 + 16 * rb + li) * kpad + k0 + kq;
                ah[rb] = Q::from(*reinterpret_cast<const uint4 *>(phi + a));
                al[rb] = Q::from(*reinterpret_cast<const uint4 *>(plo + a));
            }
#pragma unroll
            for (int nb = 0; nb < 4 * TT; ++nb) {
                const typename Q::T bh = Q::from(*reinterpret_cast<const uint4 *>(&sB[buf][0][16 * nb + li][kq]));
                const typename Q::T bl = Q::from(*reinterpret_cast<const uint4 *>(&sB[buf][1][16 * nb + li][kq]));
#pragma unroll
                for (int rb = 0; rb < 2; ++rb) {
                    acc[rb][nb] = Q::mfma(ah[rb], bh, acc[rb][nb]);
                    acc[rb][nb] = Q::mfma(ah[rb], bl, acc[rb][nb]);
                    acc[rb][nb] = Q::mfma(al[rb], bh, acc[rb][nb]);
                }
            }
        }
        buf ^= 1;
    }
    // epilogue: weighted residuals, summed over the 32 rows of this wavefront
#pragma unroll
    for (int nb = 0; nb < 4 * TT; ++nb) {
        const int col = j0 + 16 * nb + li;
        const float sc = iscale[col] * (1.0f / kPScale);
        float sum = 0.0f;
#pragma unroll
        for (int rb = 0; rb < 2; ++rb)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const float r = ow[rb][e].x - acc[rb][nb][e] * sc;
                sum = fmaf(ow[rb][e].y * r, r, sum);
            }
        if (col < ny) {
#pragma unroll
            for (int rb = 0; rb < 2; ++rb)
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int row = i0 + 16 * rb + dq + e;
                    if (row < P.np) yout[(size_t)col * P.np + row] = (double)(acc[rb][nb][e] * sc);
                }
        }
        sum += __shfl_xor(sum, 16);
        sum += __shfl_xor(sum, 32);
        if (lane < 16) red[wave][16 * nb + lane] = sum;
    }
    __syncthreads();
    for (int c = tid; c < NT; c += blockDim.x) {
        const float s = red[0][c] + red[1][c] + red[2][c] + red[3][c];
        partial[(size_t)rt * tpad + j0 + c] = s;
    }
}

// chi^2 of each trial: FP64 sum of its row-tile partials, one wavefront per trial (lane l
// takes the row tiles l, l + 64, ... in order, then a fixed shuffle tree), so every load
// of a trial is in flight at once and the sum is deterministic; go: the form's cell
// validity (3e38 when the cell is invalid or the sum is not finite)
__global__ void __launch_bounds__(256) k_matrix_reduce(const float *__restrict__ partial, int row_tiles, int tpad, int ntrials,
                                                       const int *__restrict__ go, double *chi2) {
    const int t = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (t >= ntrials) return;                   // uniform over the wavefront
    double s = 0.0;
    for (int r = lane; r < row_tiles; r += 64) s += (double)partial[(size_t)r * tpad + t];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if (lane == 0) {
        if (!*go || !isfinite(s) || s > 3e38) s = 3e38;
        chi2[t] = s;
    }
}

// ---- HBM calibration: a plain streaming copy of known size ------------------------------
__global__ void __launch_bounds__(256) k_copy_probe(const float4 *__restrict__ a, float4 *__restrict__ b, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) b[i] = a[i];
}

}  // namespace ptk

// ---- host-side launchers (used by pt_host.cpp) ----------------------------------------
namespace ptk {

size_t shared_static_bytes() { return sizeof(Shared); }
size_t plain_scratch_bytes() { return sizeof(PlainScratch); }
size_t trial_image_bytes() { return sizeof(TrialImage); }

// ---- the fixed scratch of the allocation-free evaluation ------------------------------
// pt_eval_noalloc stages its trials (trial images, line records and spans, chi^2
// partials) here instead of in the plan's buffers.  Every GPU the library runs on holds
// one copy for the life of the process.
static __device__ __align__(256) unsigned char scratch[PTK_NOALLOC_BYTES];

// The device address of `scratch` on `device` (which must be current), looked up once per
// device: on the host the array's name is not a device address.
hipError_t noalloc_scratch(int device, unsigned char **ptr, size_t *bytes) {
    static std::mutex mu;
    static unsigned char *addr[PTK_NOALLOC_DEVICES] = {};
    if (device < 0 || device >= PTK_NOALLOC_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(mu);
    if (!addr[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        addr[device] = static_cast<unsigned char *>(p);
    }
    *ptr = addr[device];
    *bytes = PTK_NOALLOC_BYTES;
    return hipSuccess;
}
int accum_points_per_group(int tpw) { return acc_points(tpw); }

// The dynamic-LDS limit of a kernel is one setting per process; every plan raises it to
// what it launches with and none lowers it, so a plan created later with smaller needs
// leaves the limit of an earlier one in place.
static hipError_t raise_lds_limit(const void *fn, size_t bytes, size_t &current) {
    static std::mutex mu;
    std::lock_guard<std::mutex> lock(mu);
    if (bytes <= current) return hipSuccess;
    hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
    if (e == hipSuccess) current = bytes;
    return e;
}
hipError_t set_lds_limits(size_t bytes) {
    static size_t cur[3] = {0, 0, 0};
    hipError_t e;
    if ((e = raise_lds_limit((const void *)k_workload, bytes, cur[0])) != hipSuccess) return e;
    if ((e = raise_lds_limit((const void *)k_pattern, bytes, cur[1])) != hipSuccess) return e;
    return raise_lds_limit((const void *)k_plain_points, bytes, cur[2]);
}

hipError_t launch_stage_setup(const DevPlan &P, const double *X, void *img, int ntrials, int pt_mode, int bootstrap,
                              hipStream_t s, const double2 *pick, int t_off) {
    hipLaunchKernelGGL(k_stage_setup, dim3(ntrials), dim3(64), 0, s, P, X, (TrialImage *)img, pt_mode, bootstrap, pick, t_off);
    return hipGetLastError();
}
hipError_t launch_stage_nnls(const DevPlan &P, const void *img, const StagedLine *rec, double2 *partial, int nsplit, int ntrials,
                             hipStream_t s, int t_off) {
    hipLaunchKernelGGL(k_stage_nnls, dim3(ntrials), dim3(kNnlsThreads), nnls_total_bytes(P.n_free), s, P, (const TrialImage *)img, rec,
                       partial, nsplit, t_off);
    return hipGetLastError();
}
size_t nnls_shared_bytes(int nf) { return nnls_total_bytes(nf); }
hipError_t nnls_lds_limit(size_t bytes) {
    static size_t cur = 0;
    return raise_lds_limit((const void *)k_stage_nnls, bytes, cur);
}

hipError_t launch_stage_follow(const DevPlan &P, const double *mstate, int mstride, const void *mimg, void *img, int ntrials,
                               hipStream_t s, int t_off) {
    hipLaunchKernelGGL(k_stage_follow, dim3(ntrials), dim3(64), 0, s, P, mstate, mstride, (const TrialImage *)mimg,
                       (TrialImage *)img, t_off);
    return hipGetLastError();
}
static inline unsigned xcd_grid(int per, int ntrials) { return 8u * (unsigned)((ntrials + 7) / 8) * (unsigned)per; }

// threads of a single-plan lines workgroup: 512 for 257-512 reflections (one workgroup per
// trial, so the finish can join the lines stage), else 256 (larger patterns ran 10% slower
// in 512-thread workgroups)
static int lines_threads(const DevPlan &P) { return P.nr > 256 && P.nr <= 512 ? 512 : 256; }
hipError_t launch_stage_lines(const DevPlan &P, const void *img, StagedLine *rec, int ntrials, hipStream_t s, int t_off) {
    // one workgroup per trial up to 512 reflections, 256-reflection workgroups beyond
    const int threads = lines_threads(P), per = (P.nr + threads - 1) / threads;
    if (P.rep_chunks > 1)
        hipLaunchKernelGGL(k_stage_lines<true>, dim3(xcd_grid(per, ntrials)), dim3(threads), 0, s, P, (TrialImage *)img, rec, per,
                           ntrials, t_off);
    else
        hipLaunchKernelGGL(k_stage_lines<false>, dim3(xcd_grid(per, ntrials)), dim3(threads), 0, s, P, (TrialImage *)img, rec, per,
                           ntrials, t_off);
    return hipGetLastError();
}
hipError_t launch_stage_accum(const DevPlan &P, const void *img, const StagedLine *rec, double2 *partial, int ngroups,
                              int ntrials, int tpw, hipStream_t s, int t_off) {
    const int per = (ngroups + tpw - 1) / tpw;
    const dim3 grid(xcd_grid(per, ntrials));
    const TrialImage *im = (const TrialImage *)img;
#define PTK_ACCUM(T, F) hipLaunchKernelGGL((k_stage_accum<T, F>), grid, dim3(256), 0, s, P, im, rec, partial, ngroups, per, ntrials, t_off)
    if (P.has_free) {
        if (tpw == 1) PTK_ACCUM(1, true); else if (tpw == 4) PTK_ACCUM(4, true); else PTK_ACCUM(2, true);
    } else {
        if (tpw == 1) PTK_ACCUM(1, false); else if (tpw == 4) PTK_ACCUM(4, false); else PTK_ACCUM(2, false);
    }
#undef PTK_ACCUM
    return hipGetLastError();
}
hipError_t launch_stage_finish(const DevPlan &P, const void *img, const double2 *partial, int nsplit, int ntrials,
                               double *chi2, int pt_mode, hipStream_t s, int t_off) {
    hipLaunchKernelGGL(k_stage_finish, dim3((ntrials + 63) / 64), dim3(64), 0, s, P, (const TrialImage *)img, partial, nsplit,
                       ntrials, chi2, pt_mode, t_off);
    return hipGetLastError();
}

hipError_t launch_stage_finish_setup(const DevPlan &P, void *img, const double2 *partial, int nsplit, hipStream_t s, int ntrials,
                                     int t_off) {
    if (P.any_sep) hipLaunchKernelGGL(k_stage_finish_setup<true>, dim3(ntrials), dim3(64), 0, s, P, (TrialImage *)img, partial, nsplit, t_off);
    else hipLaunchKernelGGL(k_stage_finish_setup<false>, dim3(ntrials), dim3(64), 0, s, P, (TrialImage *)img, partial, nsplit, t_off);
    return hipGetLastError();
}
hipError_t launch_stage_finish_lines(const DevPlan &P, void *img, const double2 *partial, int nsplit, StagedLine *rec, int ntrials,
                                     hipStream_t s, int t_off) {
    const dim3 grid(xcd_grid(1, ntrials));
    TrialImage *im = (TrialImage *)img;
    const int threads = lines_threads(P);
#define PTK_FL(SEP, CH) hipLaunchKernelGGL((k_stage_finish_lines<SEP, CH>), grid, dim3(threads), 0, s, P, im, partial, nsplit, rec, ntrials, t_off)
    // the general instance for separated profiles or free intensities (its lines stage
    // writes the free lines' spans), as k_stage_lines takes it
    if (P.any_sep || P.has_free) {
        if (P.rep_chunks > 1) PTK_FL(true, true); else PTK_FL(true, false);
    } else {
        if (P.rep_chunks > 1) PTK_FL(false, true); else PTK_FL(false, false);
    }
#undef PTK_FL
    return hipGetLastError();
}

hipError_t launch_matrix_columns(const DevPlan &P, const void *img, LineRec *rec, int4 *span, float *factor, hipStream_t s) {
    hipLaunchKernelGGL(k_matrix_columns, dim3((P.nr + 127) / 128), dim3(128), 0, s, P, (const TrialImage *)img, rec, span, factor);
    return hipGetLastError();
}
hipError_t launch_matrix_fill(const DevPlan &P, const LineRec *rec, const int4 *span, unsigned short *phi, unsigned short *plo,
                              int kpad, int bf16, hipStream_t s) {
    hipLaunchKernelGGL(k_matrix_fill, dim3(P.nr), dim3(256), 0, s, P, rec, span, phi, plo, kpad, bf16);
    return hipGetLastError();
}
hipError_t launch_matrix_tables(const DevPlan &P, const void *img, const float *factor, double4 *hr, float *f0w, float2 *qk,
                               int *go, int kpad, int nrep_max, hipStream_t s) {
    hipLaunchKernelGGL(k_matrix_tables, dim3((kpad + 255) / 256), dim3(256), 0, s, P, (const TrialImage *)img, factor, hr, f0w, qk,
                       go, kpad, nrep_max);
    return hipGetLastError();
}
hipError_t launch_matrix_intensity(const DevPlan &P, const double *X, const double4 *hr, const float *f0w, const float2 *qk,
                                   unsigned short *ihi, unsigned short *ilo, float *iscale, int kpad, int n, int bf16,
                                   hipStream_t s) {
    // at most 256 threads, as few passes over the reflections as that allows and the same
    // count for every thread: four-wavefront-or-fewer workgroups keep the chunk's grid
    // resident at once at this kernel's register count
    int na = 0;
    for (int q = 0; q < P.n_phases; ++q) na = std::max(na, P.phase[q].atom_off + P.phase[q].n_atoms);
    const int passes = std::max(1, (P.nr + 255) / 256);
    const int threads = std::max(64, (P.nr + 64 * passes - 1) / (64 * passes) * 64);
    static std::once_flag once;
    static hipError_t attr = hipSuccess;
    std::call_once(once, [] {
        attr = hipFuncSetAttribute((const void *)k_matrix_intensity<kMxT>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)matrix_intensity_lds(kMxT, PTK_MAX_ATOMS, PTK_MAX_VARS, PTK_MATRIX_MAX_REFL));
    });
    if (attr != hipSuccess) return attr;
    hipLaunchKernelGGL(k_matrix_intensity<kMxT>, dim3((n + kMxT - 1) / kMxT), dim3(threads), matrix_intensity_lds(kMxT, na, P.nv, kpad),
                       s, P, X, hr, f0w, qk, ihi, ilo, iscale, kpad, n, bf16);
    return hipGetLastError();
}
hipError_t launch_matrix_chi2(const DevPlan &P, const unsigned short *phi, const unsigned short *plo, const unsigned short *ihi,
                              const unsigned short *ilo, const float *iscale, const int2 *band, const float2 *yw,
                              float *partial, int kpad, int tpad, int row_tiles, double *yout, int ny, int bf16, hipStream_t s) {
    // row_tiles: 128-point tiles; the grid is eight XCD slots of ceil(trial tiles / 8) trial
    // tiles x the row tiles; tpad is a multiple of kMatrixTrialTile
    constexpr int TT = kMatrixTrialTile / 64;
    const dim3 grid(8 * ((tpad / kMatrixTrialTile + 7) / 8) * row_tiles);
    if (bf16)
        hipLaunchKernelGGL((k_matrix_chi2<true, TT>), grid, dim3(256), 0, s, P, phi, plo, ihi, ilo, iscale, band, yw, partial, kpad,
                           tpad, row_tiles, yout, yout ? ny : 0);
    else
        hipLaunchKernelGGL((k_matrix_chi2<false, TT>), grid, dim3(256), 0, s, P, phi, plo, ihi, ilo, iscale, band, yw, partial, kpad,
                           tpad, row_tiles, yout, yout ? ny : 0);
    return hipGetLastError();
}
hipError_t launch_matrix_reduce(const float *partial, int row_tiles, int tpad, int ntrials, const int *go, double *chi2,
                                hipStream_t s) {
    hipLaunchKernelGGL(k_matrix_reduce, dim3((ntrials + 3) / 4), dim3(256), 0, s, partial, row_tiles, tpad, ntrials, go, chi2);
    return hipGetLastError();
}

hipError_t launch_pattern(const DevPlan &P, const double *X, double *y, int block, int ch, size_t dyn, hipStream_t s) {
    hipLaunchKernelGGL(k_pattern, dim3(1), dim3(block), dyn, s, P, X, y, ch);
    return hipGetLastError();
}

hipError_t launch_workload(const DevPlan &P, const double *X, long long *stats, int block, int ch, size_t dyn,
                           hipStream_t s) {
    hipLaunchKernelGGL(k_workload, dim3(1), dim3(block), dyn, s, P, X, stats, ch);
    return hipGetLastError();
}

hipError_t launch_exchange(const DevPlan &P, hipStream_t s) {
    int pairs = (P.chains + 1) / 2;
    hipLaunchKernelGGL(k_exchange, dim3((pairs + 63) / 64), dim3(64), 0, s, P);
    return hipGetLastError();
}

hipError_t launch_group_setup(const DevPlan *plans, const DevMember *mem, const int2 *tab, int n, int bootstrap,
                              hipStream_t s) {
    hipLaunchKernelGGL(k_group_setup, dim3(n), dim3(64), 0, s, plans, mem, tab, bootstrap);
    return hipGetLastError();
}
hipError_t launch_group_lines(const DevPlan *plans, const DevMember *mem, const int4 *tab, int n, hipStream_t s) {
    hipLaunchKernelGGL(k_group_lines, dim3(n), dim3(256), 0, s, plans, mem, tab);
    return hipGetLastError();
}
hipError_t launch_group_accum(const DevPlan *plans, const DevMember *mem, const int4 *tab, int n, int tpw, hipStream_t s) {
    if (tpw == 1) hipLaunchKernelGGL(k_group_accum<1>, dim3(n), dim3(256), 0, s, plans, mem, tab);
    else if (tpw == 4) hipLaunchKernelGGL(k_group_accum<4>, dim3(n), dim3(256), 0, s, plans, mem, tab);
    else hipLaunchKernelGGL(k_group_accum<2>, dim3(n), dim3(256), 0, s, plans, mem, tab);
    return hipGetLastError();
}
hipError_t launch_group_finish(const DevPlan *plans, const DevMember *mem, const int2 *tab, int nchains, int pt_mode,
                               hipStream_t s) {
    hipLaunchKernelGGL(k_group_finish, dim3((nchains + 63) / 64), dim3(64), 0, s, plans, mem, tab, nchains, pt_mode);
    return hipGetLastError();
}
hipError_t launch_group_finish_setup(const DevPlan *plans, const DevMember *mem, const int2 *tab, int nchains, hipStream_t s) {
    hipLaunchKernelGGL(k_group_finish_setup, dim3(nchains), dim3(64), 0, s, plans, mem, tab);
    return hipGetLastError();
}
hipError_t launch_group_exchange(const DevPlan *plans, const int2 *tab, int npairs, hipStream_t s) {
    if (npairs > 0) hipLaunchKernelGGL(k_group_exchange, dim3((npairs + 63) / 64), dim3(64), 0, s, plans, tab, npairs);
    return hipGetLastError();
}

hipError_t launch_copy_probe(const void *a, void *b, size_t bytes, hipStream_t s) {
    hipLaunchKernelGGL(k_copy_probe, dim3(4096), dim3(256), 0, s, (const float4 *)a, (float4 *)b, bytes / sizeof(float4));
    return hipGetLastError();
}

hipError_t launch_plain(const DevPlan &P, const double *X, void *scratch, LineRec *rec, int4 *span, double *partial,
                        double *out, int ch, hipStream_t s) {
    PlainScratch *scr = reinterpret_cast<PlainScratch *>(scratch);
    if (P.any_sep) {
        hipLaunchKernelGGL(k_plain_setup<true>, dim3(1), dim3(256), 0, s, P, X, scr);
        hipLaunchKernelGGL(k_plain_lines<true>, dim3((P.nr + 255) / 256), dim3(256), 0, s, P, scr, rec, span);
    } else {
        hipLaunchKernelGGL(k_plain_setup<false>, dim3(1), dim3(256), 0, s, P, X, scr);
        hipLaunchKernelGGL(k_plain_lines<false>, dim3((P.nr + 255) / 256), dim3(256), 0, s, P, scr, rec, span);
    }
    int nblk = (P.np + 255) / 256;
    size_t dyn = (sizeof(LineRec) + sizeof(int4)) * (size_t)ch + 256 * sizeof(float);
    hipLaunchKernelGGL(k_plain_points, dim3(nblk), dim3(256), dyn, s, P, scr, rec, span, partial, ch);
    hipLaunchKernelGGL(k_plain_reduce, dim3(1), dim3(64), 0, s, partial, nblk, out);
    return hipGetLastError();
}

}  // namespace ptk
