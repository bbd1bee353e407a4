/*
 * p6_oracle.c -- FP64 CPU restatement of the Powder 6 Rietveld trial cost and the CPU
 * parallel-tempering loop.
 *
 * TEST INFRASTRUCTURE.  This file is the oracle: the parity checker for the HIP path and
 * the CPU baseline that bench.py times.  Nothing in the product library uses it.
 *
 * Restated from the reference's JavaScript (file:line cited per function), keeping its
 * floating-point operation order: compile with -ffp-contract=off.  Transcendentals are
 * glibc's, where the reference's are V8's (fdlibm port); the two can differ in the last
 * ulp, so cost() agrees with the reference to ~1e-12 relative, not bitwise.  Math.hypot
 * is restated with V8's Kahan-compensated algorithm.
 */
#define _GNU_SOURCE
#include "p6_oracle.h"

#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

/* ------------------------------------------------------------------------------ */
/*  constants (js/constants.js, js/crystal.js:38, js/rietveld_core.js:108)          */
/* ------------------------------------------------------------------------------ */
#define JS_PI 3.141592653589793
#define JS_LN2 0.6931471805599453
static double D2R, R2D, TWO_PI, PV_G0, PV_L0, PV_GAUSS_AREA, PV_LORENTZ_AREA, GCONV, LCONV;
static const double WINDOW_G = 8.0, WINDOW_L = 25.0, WINDOW_MAX = 10.0; /* constants.js:59,70,84 */
static const double MIN_FWHM_DEFAULT = 1e-3;                              /* constants.js:108 */
static const double CG = 2.772588722239781;                               /* profile.js:606 */
static const double GL8_X[4] = {0.1834346424956498, 0.5255324099163290, 0.7966664774136267, 0.9602898564975363};
static const double GL8_W[4] = {0.3626837833783620, 0.3137066458778873, 0.2223810344533745, 0.1012285362903763};

static pthread_once_t const_once = PTHREAD_ONCE_INIT;
static void init_constants(void) {
    D2R = JS_PI / 180;                  /* crystal.js:38, rietveld_core.js:108 */
    R2D = 180 / JS_PI;
    TWO_PI = 2 * JS_PI;
    PV_G0 = 2 * sqrt(JS_LN2 / JS_PI);   /* constants.js:155 */
    PV_L0 = 2 / JS_PI;                  /* constants.js:162 */
    PV_GAUSS_AREA = 1 / PV_G0;
    PV_LORENTZ_AREA = 1 / PV_L0;
    GCONV = sqrt(8 * JS_LN2) / 100.0;   /* constants.js:248 */
    LCONV = 1.0 / 100.0;                /* constants.js:251 */
}
static void consts(void) { pthread_once(&const_once, init_constants); }

static __thread char err_buf[256];
const char *or_error(void) { return err_buf; }
static void set_err(const char *m) { snprintf(err_buf, sizeof err_buf, "%s", m); }

static int js_finite(double v) { return isfinite(v); }
static double dmax(double a, double b) { return a > b ? a : b; }  /* Math.max for finite args */
static double dmin(double a, double b) { return a < b ? a : b; }
static double js_round(double v) { return floor(v + 0.5); }       /* Math.round */

/* Math.hypot(x, y): V8 scales by the largest magnitude and sums squares with Kahan
 * compensation, then returns sqrt(sum) * max. */
double or_hypot2(double x, double y) {
    double ax = fabs(x), ay = fabs(y);
    if (isinf(ax) || isinf(ay)) return INFINITY;
    if (isnan(ax) || isnan(ay)) return NAN;
    double mx = ax > ay ? ax : ay;
    if (mx == 0) return 0;
    double v[2] = {ax, ay}, sum = 0, comp = 0;
    for (int i = 0; i < 2; i++) {
        double n = v[i] / mx;
        double summand = n * n - comp;
        double prelim = sum + summand;
        comp = (prelim - sum) - summand;
        sum = prelim;
    }
    return sqrt(sum) * mx;
}

/* Math.hypot with n arguments, the same compensated sum */
static double hypot_n(const double *v, int n) {
    double mx = 0;
    for (int i = 0; i < n; i++) {
        double a = fabs(v[i]);
        if (isinf(a)) return INFINITY;
        if (isnan(a)) return NAN;
        if (a > mx) mx = a;
    }
    if (mx == 0) return 0;
    double sum = 0, comp = 0;
    for (int i = 0; i < n; i++) {
        double q = fabs(v[i]) / mx, summand = q * q - comp, prelim = sum + summand;
        comp = (prelim - sum) - summand;
        sum = prelim;
    }
    return sqrt(sum) * mx;
}

/* constants.js:138-143 softPositive */
double or_soft_positive(double x, double f) {
    consts();
    if (!js_finite(x)) return f;
    double h = or_hypot2(x, 2 * f);
    return x < 0 ? f * (f / (0.5 * h - 0.5 * x)) : 0.5 * x + 0.5 * h;
}

/* profile.js:182-189 softFloor */
double or_soft_floor(double x, double f) {
    if (!js_finite(x) || !(x > 0)) return f;
    if (!(f > 0)) return x;
    double r = x / f;
    if (r > 20) return x;
    return f * pow(pow(r, 4) + 1, 0.25);
}

/* profile.js:791-797 sharkoErfc (Numerical Recipes erfcc) */
double or_erfc(double x) {
    double z = fabs(x), t = 1 / (1 + 0.5 * z);
    double ans = t * exp(-z * z - 1.26551223 + t * (1.00002368 + t * (0.37409196 +
                 t * (0.09678418 + t * (-0.18628806 + t * (0.27886807 + t * (-1.13520398 +
                 t * (1.48851587 + t * (-0.82215223 + t * 0.17087277)))))))));
    return x >= 0 ? ans : 2 - ans;
}

/* crystal.js:465-471 sharkoLorentzFactor */
double or_lorentz(double tth) {
    consts();
    if (!js_finite(tth) || tth <= 0 || tth >= 180) return NAN;
    double th = 0.5 * tth * D2R, s = sin(th), c = cos(th);
    if (s < 1e-6 || c < 1e-6) return NAN;
    return 1 / (s * s * c);
}
/* crystal.js:474-478 sharkoPolarizationFactor */
double or_polarization(double tth, double K) {
    consts();
    if (!js_finite(tth)) return NAN;
    double k = (js_finite(K) && K >= 0) ? K : 1;
    double c = cos(tth * D2R);
    return (1 + k * (c * c)) / (1 + k);
}
/* crystal.js:487-497 sharkoLorentzPolarization / sharkoLp({K}) */
double or_lp(double tth, double K) {
    double L = or_lorentz(tth);
    if (!js_finite(L)) return 1;
    double P = or_polarization(tth, (js_finite(K) && K >= 0) ? K : 1);
    return js_finite(P) ? L * P : 1;
}

/* profile.js:552-565 getPeakFWHM (Thompson, Cox & Hastings 1987) */
double or_tch_fwhm(double gamma_g, double gamma_l) {
    double gG = dmax(1e-9, (gamma_g != 0 && !isnan(gamma_g)) ? gamma_g : 1e-9);
    double gL = dmax(1e-9, (gamma_l != 0 && !isnan(gamma_l)) ? gamma_l : 1e-9);
    double g5 = pow(gG, 5), l5 = pow(gL, 5);
    double g4l = 2.69269 * pow(gG, 4) * gL;
    double g3l2 = 2.42843 * pow(gG, 3) * pow(gL, 2);
    double g2l3 = 4.47163 * pow(gG, 2) * pow(gL, 3);
    double gl4 = 0.07842 * gG * pow(gL, 4);
    double s = g5 + g4l + g3l2 + g2l3 + gl4 + l5;
    if (s < 0 || !js_finite(s)) return dmax(dmax(gG, gL), 1e-6);
    return dmax(1e-6, pow(s, 0.2));
}

/* crystal.js:245-252 cellVolume */
double or_cell_volume(const double c[6]) {
    consts();
    double ca = cos(c[3] * D2R), cb = cos(c[4] * D2R), cg = cos(c[5] * D2R);
    double t = 1 - ca * ca - cb * cb - cg * cg + 2 * ca * cb * cg;
    for (int i = 0; i < 3; i++) if (!(js_finite(c[i]) && c[i] > 0)) return NAN;
    for (int i = 3; i < 6; i++) if (!(js_finite(c[i]) && c[i] > 0 && c[i] < 180)) return NAN;
    if (!(t > 0)) return NAN;
    return c[0] * c[1] * c[2] * sqrt(t);
}

/* crystal.js:326-366 buildInvDsq -> form {aa,bb,cc,ab,bc,ac}; returns 0 for a
 * degenerate cell */
int or_inv_dsq(const double cell[6], double f[6]) {
    consts();
    double a = cell[0], b = cell[1] > 0 ? cell[1] : a, c = cell[2] > 0 ? cell[2] : a;
    if (!(a > 0) || !(b > 0) || !(c > 0)) return 0;
    double al = cell[3] * D2R, be = cell[4] * D2R, ga = cell[5] * D2R;
    double cosA = cos(al), cosB = cos(be), cosG = cos(ga);
    double sinA = sin(al), sinB = sin(be), sinG = sin(ga);
    double vf = 1 - cosA * cosA - cosB * cosB - cosG * cosG + 2 * cosA * cosB * cosG;
    if (vf <= 1e-9 || fabs(sinA) < 1e-9 || fabs(sinB) < 1e-9 || fabs(sinG) < 1e-9) return 0;
    double V = a * b * c * sqrt(vf);
    double as = b * c * sinA / V, bs = a * c * sinB / V, cs = a * b * sinG / V;
    double cA = (cosB * cosG - cosA) / (sinB * sinG);
    double cB = (cosA * cosG - cosB) / (sinA * sinG);
    double cG = (cosA * cosB - cosG) / (sinA * sinB);
    f[0] = as * as; f[1] = bs * bs; f[2] = cs * cs;
    f[3] = 2 * as * bs * cG; f[4] = 2 * bs * cs * cA; f[5] = 2 * as * cs * cB;
    return 1;
}
/* crystal.js:374-377 invDsq */
static double inv_dsq(const double f[6], double h, double k, double l) {
    return h * h * f[0] + k * k * f[1] + l * l * f[2] + h * k * f[3] + k * l * f[4] + h * l * f[5];
}

/* profile.js:108-132 calculatePeakShift with profileType 'simple_pvoigt' (the sample
 * displacement and transparency shift applied in Model.geometry and in
 * buildPeakContributions' shifted()). */
double or_peak_shift(double tth, double shft, double trns) {
    consts();
    double th = tth * (JS_PI / 180) / 2;
    if (fabs(th - JS_PI / 2.0) < 1e-6) return 0;
    double cosT = cos(th), sin2T = sin(2 * th);
    double d = -(shft / 1000) * cosT * (180 / JS_PI);
    double t = trns * sin2T * (180 / JS_PI);
    double tot = d + t;
    return js_finite(tot) ? tot : 0;
}

/* ------------------------------------------------------------------------------ */
/*  symmetry and reflections (rietveld_core.js:155-276)                              */
/* ------------------------------------------------------------------------------ */
static int is_identity_r(const int32_t *r) {
    return r[0] == 1 && r[4] == 1 && r[8] == 1 && !r[1] && !r[2] && !r[3] && !r[5] && !r[6] && !r[7];
}

/* rietveld_core.js:164-179 cosets */
int or_cosets(int n_ops, const int32_t *op_r, const double *op_t, int32_t *rep_index, int *n_cent) {
    int nC = 0;
    double *C = (double *)malloc(sizeof(double) * 3 * (n_ops + 1));
    for (int i = 0; i < n_ops; i++)
        if (is_identity_r(op_r + 9 * i)) { memcpy(C + 3 * nC, op_t + 3 * i, 3 * sizeof(double)); nC++; }
    if (!nC) { C[0] = C[1] = C[2] = 0; nC = 1; }
    int nrep = 0;
    for (int i = 0; i < n_ops; i++) {
        int dup = 0;
        for (int q = 0; q < nrep && !dup; q++) {
            int j = rep_index[q];
            if (memcmp(op_r + 9 * i, op_r + 9 * j, 9 * sizeof(int32_t)) != 0) continue;
            for (int c = 0; c < nC && !dup; c++) {
                int ok = 1;
                for (int x = 0; x < 3; x++) {
                    double d = op_t[3 * i + x] - op_t[3 * j + x] - C[3 * c + x];
                    if (fabs(d - js_round(d)) > 1e-6) { ok = 0; break; }
                }
                if (ok) dup = 1;
            }
        }
        if (!dup) rep_index[nrep++] = i;
    }
    free(C);
    if (n_cent) *n_cent = nC;
    return nrep;
}

/* rietveld_core.js:182-188 laueRotations: +R and -R of every op, first occurrence kept */
static int laue_rotations(int n_ops, const int32_t *op_r, int32_t *rots) {
    int n = 0;
    for (int i = 0; i < n_ops; i++) {
        for (int s = 0; s < 2; s++) {
            int32_t r[9];
            for (int x = 0; x < 9; x++) r[x] = s ? -op_r[9 * i + x] : op_r[9 * i + x];
            int seen = 0;
            for (int q = 0; q < n && !seen; q++) if (!memcmp(rots + 9 * q, r, sizeof r)) seen = 1;
            if (!seen) { memcpy(rots + 9 * n, r, sizeof r); n++; }
        }
    }
    return n;
}

/* rietveld_core.js:192-200 isAllowed */
static int is_allowed(int h, int k, int l, int n_ops, const int32_t *op_r, const double *op_t) {
    for (int i = 0; i < n_ops; i++) {
        const int32_t *r = op_r + 9 * i;
        const double *t = op_t + 3 * i;
        if (h * r[0] + k * r[3] + l * r[6] != h || h * r[1] + k * r[4] + l * r[7] != k ||
            h * r[2] + k * r[5] + l * r[8] != l) continue;
        double p = h * t[0] + k * t[1] + l * t[2];
        if (fabs(p - js_round(p)) > 1e-6) return 0;
    }
    return 1;
}

typedef struct { int h, k, l, m; double d0; int eq_off; } refl_t;

static int refl_cmp(const void *pa, const void *pb) {
    const refl_t *a = (const refl_t *)pa, *b = (const refl_t *)pb;
    double d = b->d0 - a->d0;                      /* b.d0 - a.d0 || b.h - a.h || ... */
    if (d != 0 && !isnan(d)) return d < 0 ? -1 : 1;
    if (b->h != a->h) return b->h - a->h < 0 ? -1 : 1;
    if (b->k != a->k) return b->k - a->k < 0 ? -1 : 1;
    if (b->l != a->l) return b->l - a->l < 0 ? -1 : 1;
    return 0;
}

typedef struct {
    int n, cap;
    refl_t *r;
    int n_eq, cap_eq;
    int32_t *eq;   /* 3 ints per equivalent */
} refl_list;

static void rl_free(refl_list *L) { free(L->r); free(L->eq); memset(L, 0, sizeof *L); }

/* rietveld_core.js:247-276 generateReflections */
static int gen_reflections(const double cell[6], int n_ops, const int32_t *op_r, const double *op_t,
                           double dmin_, const int32_t *box, refl_list *out) {
    double f[6];
    out->n = 0; out->n_eq = 0;
    if (!or_inv_dsq(cell, f) || !(dmin_ > 0)) return 0;
    double lim = 1 / (dmin_ * dmin_) * (1 + 1e-9);
    int32_t *rots = (int32_t *)malloc(sizeof(int32_t) * 18 * (n_ops > 0 ? n_ops : 1));
    int nr = laue_rotations(n_ops, op_r, rots);
    double cb = cell[1] > 0 ? cell[1] : cell[0], cc = cell[2] > 0 ? cell[2] : cell[0];
    int hmax = box ? box[0] : (int)floor(cell[0] / dmin_);
    int kmax = box ? box[1] : (int)floor(cb / dmin_);
    int lmax = box ? box[2] : (int)floor(cc / dmin_);
    int32_t *orbit = (int32_t *)malloc(sizeof(int32_t) * 3 * (nr + 1));
    for (int h = -hmax; h <= hmax; h++)
        for (int k = -kmax; k <= kmax; k++)
            for (int l = -lmax; l <= lmax; l++) {
                if (!h && !k && !l) continue;
                double q = inv_dsq(f, h, k, l);
                if (!(q > 0) || (!box && q > lim)) continue;
                int ch = -1000000000, ck = 0, cl = 0, cp = -1, no = 0;
                for (int i = 0; i < nr; i++) {
                    const int32_t *r = rots + 9 * i;
                    int v0 = h * r[0] + k * r[3] + l * r[6], v1 = h * r[1] + k * r[4] + l * r[7],
                        v2 = h * r[2] + k * r[5] + l * r[8];
                    int np = (v0 >= 0) + (v1 >= 0) + (v2 >= 0);
                    int seen = 0;
                    for (int o = 0; o < no && !seen; o++)
                        if (orbit[3 * o] == v0 && orbit[3 * o + 1] == v1 && orbit[3 * o + 2] == v2) seen = 1;
                    if (!seen) { orbit[3 * no] = v0; orbit[3 * no + 1] = v1; orbit[3 * no + 2] = v2; no++; }
                    if (np > cp || (np == cp && (v0 > ch || (v0 == ch && (v1 > ck || (v1 == ck && v2 > cl)))))) {
                        cp = np; ch = v0; ck = v1; cl = v2;
                    }
                }
                if (ch != h || ck != k || cl != l) continue;
                if (!is_allowed(h, k, l, n_ops, op_r, op_t)) continue;
                if (out->n == out->cap) {
                    out->cap = out->cap ? 2 * out->cap : 256;
                    out->r = (refl_t *)realloc(out->r, sizeof(refl_t) * out->cap);
                }
                if (out->n_eq + no > out->cap_eq) {
                    out->cap_eq = 2 * (out->n_eq + no) + 1024;
                    out->eq = (int32_t *)realloc(out->eq, sizeof(int32_t) * 3 * out->cap_eq);
                }
                refl_t *R = &out->r[out->n++];
                R->h = h; R->k = k; R->l = l; R->m = no; R->d0 = 1 / sqrt(q); R->eq_off = out->n_eq;
                memcpy(out->eq + 3 * out->n_eq, orbit, sizeof(int32_t) * 3 * no);
                out->n_eq += no;
            }
    free(orbit); free(rots);
    qsort(out->r, out->n, sizeof(refl_t), refl_cmp);
    return out->n;
}

int or_generate_reflections(const double cell[6], int n_ops, const int32_t *op_r, const double *op_t,
                            double dmin_, const int32_t *box, int32_t *hklm, double *d0, int cap) {
    refl_list L = {0};
    int n = gen_reflections(cell, n_ops, op_r, op_t, dmin_, box, &L);
    for (int i = 0; i < n && i < cap; i++) {
        if (hklm) { hklm[4 * i] = L.r[i].h; hklm[4 * i + 1] = L.r[i].k; hklm[4 * i + 2] = L.r[i].l; hklm[4 * i + 3] = L.r[i].m; }
        if (d0) d0[i] = L.r[i].d0;
    }
    rl_free(&L);
    return n;
}

/* crystal.js:76-119 normaliseCell + rietveld_core.js:207-221 cellOf: cell of phase at p */
static int cell_of(const pt_phase *ph, const double *pp, double cell[6]) {
    double a = pp[PT_PH_A];
    if (!(a > 0)) return 0;
    double b = pp[PT_PH_B] > 0 ? pp[PT_PH_B] : a, c = pp[PT_PH_C] > 0 ? pp[PT_PH_C] : a;
    cell[0] = a; cell[1] = b; cell[2] = c; cell[3] = cell[4] = cell[5] = 90;
    switch (ph->system) {
    case PT_CUBIC: cell[1] = cell[2] = a; break;
    case PT_TETRAGONAL: cell[1] = a; break;
    case PT_HEXAGONAL: case PT_TRIGONAL: cell[1] = a; cell[5] = 120; break;
    case PT_ORTHORHOMBIC: break;
    case PT_MONOCLINIC: {
        int ax = ph->unique_axis;
        double v = pp[PT_PH_ALPHA + (ax == 0 ? 0 : ax == 2 ? 2 : 1)];
        cell[3 + (ax == 0 ? 0 : ax == 2 ? 2 : 1)] = js_finite(v) ? v : 90;
        break;
    }
    default:
        cell[3] = js_finite(pp[PT_PH_ALPHA]) ? pp[PT_PH_ALPHA] : 90;
        cell[4] = js_finite(pp[PT_PH_BETA]) ? pp[PT_PH_BETA] : 90;
        cell[5] = js_finite(pp[PT_PH_GAMMA]) ? pp[PT_PH_GAMMA] : 90;
    }
    return 1;
}

/* rietveld_core.js:238-245 reflectionDmin */
double or_reflection_dmin(const pt_problem *prob, double zero, double disp, double transp) {
    consts();
    double lambda = prob->lambda1;
    if (prob->ratio > 1e-6 && prob->lambda2 > 0) lambda = dmin(lambda, prob->lambda2);
    double margin = WINDOW_MAX + fabs(zero) + (fabs(disp) * .001 + fabs(transp)) * R2D;
    double upper = dmin(179.999999, prob->tth[prob->n_points - 1] + margin + .001);
    return lambda / (2 * sin(upper * D2R / 2));
}

/* ------------------------------------------------------------------------------ */
/*  profile (profile.js:234-759, constants.js:194-201)                              */
/* ------------------------------------------------------------------------------ */
/* constants.js:194-201 pvMixCoeffs */
static void pv_mix(double eta, double hg, double hl, double *cL, double *cG) {
    double e = dmax(0, dmin(1, eta));
    double wL = e * PV_L0 / dmax(1e-12, hl);
    double wG = (1 - e) * PV_G0 / dmax(1e-12, hg);
    double s = wL + wG;
    if (!(s > 0) || !js_finite(s)) { *cL = e; *cG = 1 - e; return; }
    *cL = wL / s; *cG = wG / s;
}
/* profile.js:149-153 gaussianFromSigmaSq */
static double gaussian_from_sigma_sq(double s2, double minF) {
    double fwhmMin = minF > 0 ? minF : MIN_FWHM_DEFAULT;
    double floorSq = pow(0.25 * fwhmMin / GCONV, 2);
    return sqrt(or_soft_positive(s2, floorSq)) * GCONV;
}
/* profile.js:160-164 lorentzianFromGamma */
static double lorentzian_from_gamma(double g, double minF) {
    double fwhmMin = minF > 0 ? minF : MIN_FWHM_DEFAULT;
    double fl = 0.25 * fwhmMin / LCONV;
    return or_soft_positive(g, fl) * LCONV;
}
static double orz(double v) { return (v != 0 && !isnan(v)) ? v : 0; }  /* `x || 0` */

/* ---- separated (GSAS) profile ------------------------------------------------ */
/* profile.js:332 GSAS_QUARTIC exponents (h, k, l) */
static const int QUARTIC[15][3] = {{4,0,0},{0,4,0},{0,0,4},{2,2,0},{2,0,2},{0,2,2},{3,1,0},{3,0,1},
                                   {1,3,0},{1,0,3},{0,3,1},{0,1,3},{2,1,1},{1,2,1},{1,1,2}};
/* rietveld_core.js:413, 454-459: instrument defaults, in PT_INST_* order */
static const double INST_DEFAULT[PT_N_INST] = {2, -2, 5, 1, 0, 0, 0, 0, 0.5, 2, -2, 5, 1, 2, -2, 5, 1, 0.5};

/* profile.js:334-363 gsasStephensBasis: the quartic forms in h invariant under the
 * setting's rotations.  For every GSAS_QUARTIC key j (weight 1, 2, 3 or 4 as its
 * exponents hold a 4, a 3, a 0 or none), the weighted monomial of h' = R^T h averaged
 * over the distinct rotations; kept when independent of the terms before it (Gram-Schmidt,
 * norm >= 1e-8), scaled so that its leading coefficient is the key's weight.  Writes the
 * key index and the 15 coefficients of every term; returns the term count. */
int or_stephens_basis(int n_ops, const int32_t *op_r, int32_t *keys, double *coeff) {
    int nrot = 0;
    int32_t *rots = (int32_t *)malloc(sizeof(int32_t) * 9 * (n_ops > 0 ? n_ops : 1));
    for (int o = 0; o < n_ops; o++) {
        int dup = 0;
        for (int r = 0; r < nrot && !dup; r++) dup = !memcmp(rots + 9 * r, op_r + 9 * o, sizeof(int32_t) * 9);
        if (!dup) memcpy(rots + 9 * nrot++, op_r + 9 * o, sizeof(int32_t) * 9);
    }
    if (!nrot) { free(rots); return 0; }
    double orth[15][15];
    int north = 0, nb = 0;
    for (int j = 0; j < 15; j++) {
        const int *ex = QUARTIC[j];
        int has4 = ex[0] == 4 || ex[1] == 4 || ex[2] == 4, has3 = ex[0] == 3 || ex[1] == 3 || ex[2] == 3;
        int has0 = ex[0] == 0 || ex[1] == 0 || ex[2] == 0;
        double wj = has4 ? 1 : has3 ? 2 : has0 ? 3 : 4;
        double c[15] = {0};
        for (int r = 0; r < nrot; r++) {
            const int32_t *R = rots + 9 * r;
            /* polynomial in h as a dense 5x5x5 table of exponents, built factor by factor:
             * h'_axis = sum_k R[3k + axis] h_k */
            double poly[5][5][5], next[5][5][5];
            memset(poly, 0, sizeof poly);
            poly[0][0][0] = wj;
            for (int axis = 0; axis < 3; axis++)
                for (int rep = 0; rep < ex[axis]; rep++) {
                    memset(next, 0, sizeof next);
                    for (int a = 0; a < 5; a++) for (int b = 0; b < 5; b++) for (int d = 0; d < 5; d++) {
                        double v = poly[a][b][d];
                        if (v == 0) continue;
                        for (int k = 0; k < 3; k++) {
                            int cc = R[3 * k + axis];
                            if (!cc) continue;
                            int e0 = a + (k == 0), e1 = b + (k == 1), e2 = d + (k == 2);
                            next[e0][e1][e2] += v * cc;
                        }
                    }
                    memcpy(poly, next, sizeof poly);
                }
            for (int i = 0; i < 15; i++) c[i] += poly[QUARTIC[i][0]][QUARTIC[i][1]][QUARTIC[i][2]] / nrot;
        }
        double v[15];
        memcpy(v, c, sizeof v);
        for (int q = 0; q < north; q++) {
            double dot = 0;
            for (int i = 0; i < 15; i++) dot += v[i] * orth[q][i];
            for (int i = 0; i < 15; i++) v[i] -= dot * orth[q][i];
        }
        double norm = hypot_n(v, 15);
        if (norm < 1e-8) continue;
        for (int i = 0; i < 15; i++) orth[north][i] = v[i] / norm;
        north++;
        int lead = -1;
        for (int i = 0; i < 15 && lead < 0; i++) if (fabs(c[i]) > 1e-8) lead = i;
        const int *el = QUARTIC[lead];
        int l4 = el[0] == 4 || el[1] == 4 || el[2] == 4, l3 = el[0] == 3 || el[1] == 3 || el[2] == 3;
        int l0 = el[0] == 0 || el[1] == 0 || el[2] == 0;
        double factor = (l4 ? 1 : l3 ? 2 : l0 ? 3 : 4) / c[lead];
        keys[nb] = j;
        for (int i = 0; i < 15; i++) coeff[15 * nb + i] = fabs(c[i]) < 1e-9 ? 0 : c[i] * factor;
        nb++;
    }
    free(rots);
    return nb;
}

/* crystal.js:263-275 cellBasis and profile.js:364-369 gsasReciprocalVector */
static void reciprocal_vector(const double cell[6], const double h[3], double out[3]) {
    double a = cell[0], b = cell[1], c = cell[2];
    double al = cell[3] * D2R, be = cell[4] * D2R, ga = cell[5] * D2R;
    double ca = cos(al), cb = cos(be), cg = cos(ga), sg = sin(ga);
    double vol = or_cell_volume(cell) / (a * b * c);
    double av[3] = {a, 0, 0}, bv[3] = {b * cg, b * sg, 0}, cv[3] = {c * cb, c * (ca - cb * cg) / sg, c * vol / sg};
    double bc[3] = {bv[1] * cv[2] - bv[2] * cv[1], bv[2] * cv[0] - bv[0] * cv[2], bv[0] * cv[1] - bv[1] * cv[0]};
    double cav[3] = {cv[1] * av[2] - cv[2] * av[1], cv[2] * av[0] - cv[0] * av[2], cv[0] * av[1] - cv[1] * av[0]};
    double ab[3] = {av[1] * bv[2] - av[2] * bv[1], av[2] * bv[0] - av[0] * bv[2], av[0] * bv[1] - av[1] * bv[0]};
    double v = av[0] * bc[0] + av[1] * bc[1] + av[2] * bc[2];
    for (int i = 0; i < 3; i++) out[i] = (h[0] * bc[i] + h[1] * cav[i] + h[2] * ab[i]) / v;
}

static double sval(const double *sample, int k, double d) { return js_finite(sample[k]) ? sample[k] : d; }

/* profile.js:390 angle(): squared cosine between the unit reciprocal direction n of a
 * reflection and the reciprocal vector of `axis`, clamped to [0, 1] */
double or_axis_angle(const double cell[6], const double n[3], const double axis[3]) {
    double a[3];
    reciprocal_vector(cell, axis, a);
    double na = hypot_n(a, 3), dot = 0;
    if (na == 0) na = 1;
    for (int i = 0; i < 3; i++) dot = dot + n[i] * a[i];
    return dmin(1, dmax(0, pow(dot / na, 2)));
}

/* profile.js:375-378 gsasStephensVariance with a basis: sum over the terms of
 * S_term (coefficients . monomials of h); S holds the 15 coefficients in GSAS_QUARTIC
 * order (a value that is not finite reads as 0) */
double or_stephens_sum(const double h[3], const double *S, int nb, const int32_t *keys, const double *coeff) {
    double mono[15], total = 0;
    for (int i = 0; i < 15; i++) mono[i] = pow(h[0], QUARTIC[i][0]) * pow(h[1], QUARTIC[i][1]) * pow(h[2], QUARTIC[i][2]);
    for (int t = 0; t < nb; t++) {
        double inner = 0;
        for (int i = 0; i < 15; i++) inner = inner + coeff[15 * t + i] * mono[i];
        double v = js_finite(S[keys[t]]) ? S[keys[t]] : 0;
        total = total + v * inner;
    }
    return total;
}

/* profile.js:384-422 gsasCwComponents, returned as (Gaussian, Lorentzian) FWHM before the
 * floor.  flank: -1 the family's own U..Z; 0 / 1 the split flank (Y = Z = 0). */
static void gsas_components(double tth, const or_sep *sp, double lambda, int flank, double minF, double *gG, double *gL) {
    double theta = tth * JS_PI / 360, tn = tan(theta), cs = dmax(1e-9, cos(theta));
    const double *S = sp->sample;
    int directional = sp->size_model == PT_SIZE_UNIAXIAL || sp->size_model == PT_SIZE_ELLIPSOID ||
                      sp->strain_model == PT_STRAIN_UNIAXIAL;
    double n[3] = {0, 0, 0};
    if (directional) {
        double q[3];
        reciprocal_vector(sp->cell, sp->h, q);
        double norm = hypot_n(q, 3), dv = norm != 0 ? norm : 1;
        for (int i = 0; i < 3; i++) n[i] = q[i] / dv;
    }
    double invSize = 1 / dmax(0.001, sval(S, PT_GS_SIZE, 1)), strain = dmax(0, sval(S, PT_GS_STRAIN, 0));
    if (sp->size_model == PT_SIZE_UNIAXIAL) {
        double c2 = or_axis_angle(sp->cell, n, sp->size_axis);
        double i = dmax(0.001, sval(S, PT_GS_SIZE, 1)), a = dmax(0.001, sval(S, PT_GS_SIZE_AXIAL, 1));
        invSize = sqrt((1 - c2) / (i * i) + c2 / (a * a));
    } else if (sp->size_model == PT_SIZE_ELLIPSOID) {
        double v[3] = {n[0], n[1], n[2]};
        const int axes[3] = {2, 1, 0}, keys[3] = {PT_GS_ROT_C, PT_GS_ROT_B, PT_GS_ROT_A};
        for (int r = 0; r < 3; r++) {
            double t = -sval(S, keys[r], 0) * JS_PI / 180, c = cos(t), sn = sin(t);
            int i = (axes[r] + 1) % 3, j = (axes[r] + 2) % 3;
            double x = v[i], y = v[j];
            v[i] = c * x - sn * y; v[j] = sn * x + c * y;
        }
        double L[3] = {sval(S, PT_GS_SIZE, 1), sval(S, PT_GS_SIZE_B, 1), sval(S, PT_GS_SIZE_C, 1)}, acc = 0;
        for (int i = 0; i < 3; i++) acc = acc + v[i] * v[i] / pow(dmax(0.001, L[i]), 2);
        invSize = sqrt(acc);
    }
    if (sp->strain_model == PT_STRAIN_UNIAXIAL) {
        double c2 = or_axis_angle(sp->cell, n, sp->strain_axis);
        double i = dmax(0, sval(S, PT_GS_STRAIN, 0)), a = dmax(0, sval(S, PT_GS_STRAIN_AXIAL, 0));
        strain = c2 <= 1e-14 ? i : c2 >= 1 - 1e-14 ? a : i * a / dmax(1e-30, sqrt(i * i * c2 + a * a * (1 - c2)));
    } else if (sp->strain_model == PT_STRAIN_STEPHENS) {
        double sum = or_stephens_sum(sp->h, S + PT_GS_S0, sp->n_basis, sp->basis_key, sp->basis_coeff);
        strain = (sp->d * sp->d) * sqrt(dmax(0, sum));
    }
    double sizeGamma = 1.8 * (js_finite(lambda) ? lambda : 1.540598) * invSize / (JS_PI * cs);
    double strainGamma = 0.018 * strain * tn / JS_PI;
    double sm = dmax(0, dmin(1, sval(S, PT_GS_SIZE_MIX, 1))), em = dmax(0, dmin(1, sval(S, PT_GS_STRAIN_MIX, 1)));
    const double *I = sp->inst;
    double U, V, W, X, Y, Z;
    if (flank < 0) { U = I[PT_INST_U]; V = I[PT_INST_V]; W = I[PT_INST_W]; X = I[PT_INST_X]; Y = I[PT_INST_Y]; Z = I[PT_INST_Z]; }
    else {
        int o = flank == 0 ? 0 : PT_INST_U_R - PT_INST_U_L;
        U = I[PT_INST_U_L + o]; V = I[PT_INST_V_L + o]; W = I[PT_INST_W_L + o]; X = I[PT_INST_X_L + o]; Y = 0; Z = 0;
    }
    double instrumentSigma = U * tn * tn + V * tn + W;
    double instrumentGamma = X / cs + Y * tn + Z;
    double sampleSigma = (pow(sizeGamma * (1 - sm), 2) + pow(strainGamma * (1 - em), 2)) / (8 * log(2));
    double sampleGamma = sizeGamma * sm + strainGamma * em;
    double fwhmMin = minF > 0 ? minF : MIN_FWHM_DEFAULT;
    *gG = gaussian_from_sigma_sq(instrumentSigma + sampleSigma, fwhmMin);
    *gL = lorentzian_from_gamma(instrumentGamma + sampleGamma, fwhmMin);
}

/* profile.js:501-537 calculateProfileWidths (non-separated TCH / simple / split);
 * side: 0 left, 1 right/centre.  sep != NULL: the separated profile (profile.js:235, 271,
 * 427), the family's instrument function plus the sample broadening. */
static void profile_widths(int type, const double *p, const or_sep *sep, double lambda, double minF, double tth, int side,
                           double *gG, double *gL) {
    double fwhmMin = minF > 0 ? minF : MIN_FWHM_DEFAULT;
    double th = tth * (JS_PI / 180) / 2, MAXA = JS_PI / 2.0 - 1e-6, safe = dmin(th, MAXA);
    if (isnan(th)) safe = NAN;
    if (safe < 1e-6) { *gG = fwhmMin; *gL = fwhmMin; return; }
    double tn = tan(safe), cs = cos(safe), cs_safe = dmax(cs, 1e-9), cs2_safe = dmax(cs * cs, 1e-9);
    double g, l;
    if (sep) {
        gsas_components(tth, sep, lambda, type == PT_PROFILE_SPLIT ? side : -1, minF, &g, &l);
    } else if (type == PT_PROFILE_SIMPLE) {     /* profile.js:234-254 */
        double GU = orz(p[PT_SIM_GU]), GV = orz(p[PT_SIM_GV]), GW = orz(p[PT_SIM_GW]), GP = orz(p[PT_SIM_GP]), LX = orz(p[PT_SIM_LX]);
        double s2 = GU * tn * tn + GV * tn + GW + GP / cs2_safe;
        g = gaussian_from_sigma_sq(s2, fwhmMin);
        l = lorentzian_from_gamma(LX / cs_safe, fwhmMin);
    } else if (type == PT_PROFILE_SPLIT) {      /* profile.js:270-298 */
        int o = side == 0 ? 0 : 4;
        double GU = orz(p[PT_SPL_GU_L + o]), GV = orz(p[PT_SPL_GV_L + o]), GW = orz(p[PT_SPL_GW_L + o]), LX = orz(p[PT_SPL_LX_L + o]);
        double s2 = GU * tn * tn + GV * tn + GW;
        g = gaussian_from_sigma_sq(s2, fwhmMin);
        l = lorentzian_from_gamma(LX / cs_safe, fwhmMin);
    } else {                                    /* profile.js:426-482 (no Stephens terms) */
        double U = orz(p[PT_TCH_U]), V = orz(p[PT_TCH_V]), W = orz(p[PT_TCH_W]), X = orz(p[PT_TCH_X]), Y = orz(p[PT_TCH_Y]);
        double s2 = U * tn * tn + V * tn + W;
        g = gaussian_from_sigma_sq(s2, fwhmMin);
        l = lorentzian_from_gamma(X / cs_safe + Y * tn, fwhmMin);
    }
    *gG = or_soft_floor(g, fwhmMin);            /* profile.js:213-219 applyFwhmFloor */
    *gL = or_soft_floor(l, fwhmMin);
}

/* profile.js:733-759 evalVoigt */
double or_eval_voigt(double x, const or_voigt *p) {
    double delta = x - p->x0;
    if (delta > p->max_d || delta < -p->max_d) return 0.0;
    if (p->asym != 0 && fabs(delta) >= 1e-9) {
        double t = dmax(-0.95, dmin(0.95, p->asym * delta));
        delta /= (1.0 - t);
    }
    int s = (p->type == PT_PROFILE_SPLIT && !(delta < 0)) ? 1 : 0;
    double ig = p->inv_g[s], il = p->inv_l[s], cL = p->cl[s], cG = p->cg[s];
    double dg = (delta * ig) * (delta * ig);
    double dl = (delta * il) * (delta * il);
    double g = exp(-CG * dg);
    double l = 1.0 / (1.0 + 4.0 * dl);
    double v = cL * l + cG * g - p->pedestal;
    return v > 0 ? v : 0.0;
}

/* profile.js:799-831 asymmetryAreaCorrection */
double or_asym_area(const or_voigt *prep) {
    consts();
    if (prep->asym == 0 || !(prep->max_d > 0) || prep->type != PT_PROFILE_TCH) return 0;
    or_voigt as = *prep, sy;
    as.x0 = 0; as.max_d = INFINITY; as.pedestal = 0;
    sy = as; sy.asym = 0;
    double a = prep->asym, H = prep->fwhm, w = dmax(1e-6, prep->window_fwhm);
    double clampAt = 0.95 / fabs(a), D = dmax(prep->max_d, dmin(clampAt, 4096 * w));
    double total = 0;
    for (int si = 0; si < 2; si++) {
        double side = si ? 1 : -1, lo = 0, e = 0.25 * w;
        for (;;) {
            double hi = dmin(e, D);
            double bounds[2][2]; int np = 0;
            if (clampAt > lo && clampAt < hi) { bounds[np][0] = lo; bounds[np][1] = clampAt; np++; lo = clampAt; }
            bounds[np][0] = lo; bounds[np][1] = hi; np++;
            for (int b = 0; b < np; b++) {
                double m = 0.5 * (bounds[b][0] + bounds[b][1]), h = 0.5 * (bounds[b][1] - bounds[b][0]), s = 0;
                for (int i = 0; i < 4; i++)
                    for (int sg = -1; sg <= 1; sg += 2) {
                        double d = side * (m + sg * h * GL8_X[i]);
                        s += GL8_W[i] * (or_eval_voigt(d, &as) - or_eval_voigt(d, &sy));
                    }
                total += s * h;
            }
            lo = hi;
            if (hi >= D) break;
            e *= 2;
        }
        double k = 1 - dmax(-0.95, dmin(0.95, a * side * D));
        double x1 = D / k, x2 = D;
        double t1 = prep->cl[0] * 0.5 * H * (0.5 * JS_PI - atan(2 * x1 / H)) +
                    prep->cg[0] * 0.5 * H * sqrt(JS_PI / CG) * or_erfc(sqrt(CG) * x1 / H);
        double t2 = prep->cl[0] * 0.5 * H * (0.5 * JS_PI - atan(2 * x2 / H)) +
                    prep->cg[0] * 0.5 * H * sqrt(JS_PI / CG) * or_erfc(sqrt(CG) * x2 / H);
        total += k * t1 - t2;
    }
    return total;
}

/* profile.js:603-718 prepareVoigt (+ preparedProfileArea, profile.js:854-858) */
int or_prepare_voigt_sep(int type, const double *p, const or_sep *sep, double lambda, double minF, double tth_peak,
                         double x0, or_voigt *o) {
    consts();
    memset(o, 0, sizeof *o);
    o->type = type; o->x0 = x0;
    double asym = 0;
    if (type == PT_PROFILE_TCH && (orz(p[PT_TCH_SL]) != 0 || orz(p[PT_TCH_HL]) != 0) && tth_peak >= 0.1 && tth_peak < 180) {
        double th = tth_peak * (JS_PI / 180) / 2.0;
        double safe = dmax(1e-6, dmin(JS_PI / 2.0 - 1e-6, th));
        double tn = tan(safe);
        if (fabs(tn) >= 1e-9) asym = orz(p[PT_TCH_SL]) / tn + orz(p[PT_TCH_HL]);
    }
    o->asym = asym;
    double fwhm_total, eta;
    if (type == PT_PROFILE_SIMPLE) {
        double gG, gL;
        profile_widths(type, p, sep, lambda, minF, tth_peak, 1, &gG, &gL);
        double HG = dmax(1e-9, gG), HL = dmax(1e-9, gL);
        double e = p[PT_SIM_ETA];
        eta = dmax(0, dmin(1, js_finite(e) ? e : 0.5));   /* readEta */
        fwhm_total = or_tch_fwhm(HG, HL);
        double cL, cG; pv_mix(eta, HG, HL, &cL, &cG);
        o->inv_g[0] = o->inv_g[1] = 1 / HG; o->inv_l[0] = o->inv_l[1] = 1 / HL;
        o->cl[0] = o->cl[1] = cL; o->cg[0] = o->cg[1] = cG;
        o->analytic_area = cL * PV_LORENTZ_AREA * HL + cG * PV_GAUSS_AREA * HG;
        o->fwhm = fwhm_total;
    } else if (type == PT_PROFILE_SPLIT) {
        double gGL, gLL, gGR, gLR;
        profile_widths(type, p, sep, lambda, minF, tth_peak, 0, &gGL, &gLL);
        profile_widths(type, p, sep, lambda, minF, tth_peak, 1, &gGR, &gLR);
        double HGL = dmax(1e-9, gGL), HLL = dmax(1e-9, gLL), HGR = dmax(1e-9, gGR), HLR = dmax(1e-9, gLR);
        double e = p[PT_SPL_ETA];
        eta = dmax(0, dmin(1, js_finite(e) ? e : 0.5));
        fwhm_total = dmax(or_tch_fwhm(HGL, HLL), or_tch_fwhm(HGR, HLR));
        double cLL, cGL, cLR, cGR;
        pv_mix(eta, HGL, HLL, &cLL, &cGL);
        pv_mix(eta, HGR, HLR, &cLR, &cGR);
        o->inv_g[0] = 1 / HGL; o->inv_l[0] = 1 / HLL; o->inv_g[1] = 1 / HGR; o->inv_l[1] = 1 / HLR;
        o->cl[0] = cLL; o->cg[0] = cGL; o->cl[1] = cLR; o->cg[1] = cGR;
        o->analytic_area = 0.5 * (cLL * PV_LORENTZ_AREA * HLL + cGL * PV_GAUSS_AREA * HGL) +
                           0.5 * (cLR * PV_LORENTZ_AREA * HLR + cGR * PV_GAUSS_AREA * HGR);
        o->fwhm = fwhm_total;
    } else {
        double gG, gL;
        profile_widths(PT_PROFILE_TCH, p, sep, lambda, minF, tth_peak, 1, &gG, &gL);
        double HG = dmax(1e-9, gG), HL = dmax(1e-9, gL);
        double fwhm = dmax(1e-9, or_tch_fwhm(HG, HL));
        eta = 0;
        if (fwhm > 1e-9) {
            double r = HL / fwhm;
            eta = dmax(0, dmin(1, 1.36603 * r - 0.47719 * (r * r) + 0.11116 * pow(r, 3)));
        }
        fwhm_total = fwhm;
        double cL, cG; pv_mix(eta, fwhm, fwhm, &cL, &cG);
        o->inv_g[0] = o->inv_g[1] = o->inv_l[0] = o->inv_l[1] = 1 / fwhm;
        o->cl[0] = o->cl[1] = cL; o->cg[0] = o->cg[1] = cG;
        o->analytic_area = (cL * PV_LORENTZ_AREA + cG * PV_GAUSS_AREA) * fwhm;
        o->fwhm = fwhm;
    }
    o->eta = eta;
    o->window_fwhm = dmax(1e-6, fwhm_total);
    double etaWin = dmax(0, dmin(1, eta));
    double mult = WINDOW_G + (WINDOW_L - WINDOW_G) * etaWin;
    double R = dmin(WINDOW_MAX, mult * dmax(0.01, o->window_fwhm) * (asym != 0 ? 2.0 : 1.0));
    o->pedestal = 0;
    o->max_d = INFINITY;
    o->pedestal = dmax(or_eval_voigt(x0 + R, o), or_eval_voigt(x0 - R, o));
    o->max_d = R;
    o->area = o->analytic_area + or_asym_area(o);
    return 1;
}

/* ------------------------------------------------------------------------------ */
/*  the model (rietveld_core.js:629-931, 1152-1164, 1878-1889)                      */
/* ------------------------------------------------------------------------------ */
typedef struct {
    int nrep, ncent;
    int32_t *rep;                    /* indices into the phase's ops */
    refl_list refl;                  /* current reflection list */
    int sep, fam;                    /* separated profile; line family (PT_PROFILE_SIMPLE..TCH) */
    double size_axis[3], strain_axis[3];
    int nb;                          /* Stephens basis (separated, PT_STRAIN_STEPHENS) */
    int32_t bkey[15];
    double bcoef[15 * 15];
} or_phase;

/* rietveld_core.js:416-421 broadeningConfig axis(): three finite numbers, one non-zero,
 * else 0 0 1 */
static void broadening_axis(const double *a, double out[3]) {
    int ok = 1, nz = 0;
    for (int i = 0; i < 3; i++) { if (!js_finite(a[i])) ok = 0; if (a[i] != 0) nz = 1; }
    if (ok && nz) { out[0] = a[0]; out[1] = a[1]; out[2] = a[2]; }
    else { out[0] = 0; out[1] = 0; out[2] = 1; }
}

struct or_model {
    pt_problem P;                    /* shallow copy; arrays owned below */
    double *tth, *y, *w;
    pt_phase *phases;
    pt_atom **atoms;
    int32_t **op_r;
    double **op_t;
    double *base, *lo, *hi, *scale, *x0, *coeff;
    int32_t *tparam, *tvar;
    or_phase *ph;
    double **T;                      /* Chebyshev T_k(x_i) */
    double *bx;                      /* spline anchors (PT_BKG_SPLINE) */
    double guard_lo, guard_hi;
    /* columns of the free-intensity reflections of the current evaluation */
    int nf, fcap;
    int *fstart, *flen;
    double *fvals;
    size_t fvals_n, fvals_cap;
    int has_key;
    double *key;                     /* cache key: zero, disp, transp, cells */
    /* scratch */
    double *params, *ybuf, *col;
};

static double *dup_d(const double *s, int n) {
    double *d = (double *)malloc(sizeof(double) * (n > 0 ? n : 1));
    if (n > 0) memcpy(d, s, sizeof(double) * n);
    return d;
}

or_model *or_create(const pt_problem *prob) {
    consts();
    if (!prob || prob->n_points < 2 || prob->n_phases < 1 || prob->n_bkg < 0 || prob->n_bkg > PT_MAX_BKG) { set_err("invalid problem"); return NULL; }
    or_model *m = (or_model *)calloc(1, sizeof *m);
    m->P = *prob;
    int n = prob->n_points;
    m->tth = dup_d(prob->tth, n); m->y = dup_d(prob->y_obs, n); m->w = dup_d(prob->weight, n);
    m->P.tth = m->tth; m->P.y_obs = m->y; m->P.weight = m->w;
    m->phases = (pt_phase *)malloc(sizeof(pt_phase) * prob->n_phases);
    m->atoms = (pt_atom **)calloc(prob->n_phases, sizeof(pt_atom *));
    m->op_r = (int32_t **)calloc(prob->n_phases, sizeof(int32_t *));
    m->op_t = (double **)calloc(prob->n_phases, sizeof(double *));
    m->ph = (or_phase *)calloc(prob->n_phases, sizeof(or_phase));
    for (int p = 0; p < prob->n_phases; p++) {
        const pt_phase *s = &prob->phases[p];
        m->phases[p] = *s;
        m->atoms[p] = (pt_atom *)malloc(sizeof(pt_atom) * (s->n_atoms > 0 ? s->n_atoms : 1));
        memcpy(m->atoms[p], s->atoms, sizeof(pt_atom) * s->n_atoms);
        m->op_r[p] = (int32_t *)malloc(sizeof(int32_t) * 9 * s->n_ops);
        memcpy(m->op_r[p], s->op_r, sizeof(int32_t) * 9 * s->n_ops);
        m->op_t[p] = dup_d(s->op_t, 3 * s->n_ops);
        m->phases[p].atoms = m->atoms[p]; m->phases[p].op_r = m->op_r[p]; m->phases[p].op_t = m->op_t[p];
        m->ph[p].rep = (int32_t *)malloc(sizeof(int32_t) * (s->n_ops + 1));
        m->ph[p].nrep = or_cosets(s->n_ops, s->op_r, s->op_t, m->ph[p].rep, &m->ph[p].ncent);
        if (!s->n_ops) { set_err("phase without operators"); }
        /* rietveld_core.js:460 usesSeparatedProfile; :700-702 the instrument family */
        or_phase *op = &m->ph[p];
        op->sep = s->separated || s->profile_type == PT_PROFILE_GSAS_CW;
        op->fam = op->sep && prob->resolution_type ? prob->resolution_type : s->profile_type;
        if (op->fam == PT_PROFILE_GSAS_CW) op->fam = PT_PROFILE_TCH;
        broadening_axis(s->size_axis, op->size_axis);
        broadening_axis(s->strain_axis, op->strain_axis);
        op->nb = op->sep && s->strain_model == PT_STRAIN_STEPHENS ? or_stephens_basis(s->n_ops, s->op_r, op->bkey, op->bcoef) : 0;
    }
    m->P.phases = m->phases;
    m->base = dup_d(prob->param_base, prob->n_params);
    m->lo = dup_d(prob->lo, prob->n_vars); m->hi = dup_d(prob->hi, prob->n_vars);
    m->scale = dup_d(prob->scale, prob->n_vars); m->x0 = dup_d(prob->x0, prob->n_vars);
    m->coeff = dup_d(prob->term_coeff, prob->n_terms);
    m->tparam = (int32_t *)malloc(sizeof(int32_t) * (prob->n_terms + 1));
    m->tvar = (int32_t *)malloc(sizeof(int32_t) * (prob->n_terms + 1));
    if (prob->n_terms) {
        memcpy(m->tparam, prob->term_param, sizeof(int32_t) * prob->n_terms);
        memcpy(m->tvar, prob->term_var, sizeof(int32_t) * prob->n_terms);
    }
    if (prob->bkg_type == PT_BKG_SPLINE) {
        m->bx = dup_d(prob->bkg_x, prob->n_bkg);
        m->P.bkg_x = m->bx;
        for (int j = 1; j < prob->n_bkg; j++)
            if (!(m->bx[j] >= m->bx[j - 1])) { set_err("spline anchors must ascend"); }
    }
    /* rietveld_core.js:647-653 Chebyshev polynomials over the fitted range (none for a
     * spline background, :639-640) */
    double lo = m->tth[0], hi = m->tth[n - 1];
    const int n_cheb = prob->bkg_type == PT_BKG_SPLINE ? 0 : prob->n_bkg;
    m->P.n_bkg = prob->n_bkg;
    m->T = (double **)calloc(prob->n_bkg > 0 ? prob->n_bkg : 1, sizeof(double *));
    for (int k = 0; k < n_cheb; k++) m->T[k] = (double *)malloc(sizeof(double) * n);
    for (int i = 0; i < n; i++) {
        double x = hi > lo ? 2 * (m->tth[i] - lo) / (hi - lo) - 1 : 0;
        if (n_cheb > 0) m->T[0][i] = 1;
        if (n_cheb > 1) m->T[1][i] = x;
        for (int k = 2; k < n_cheb; k++) m->T[k][i] = 2 * x * m->T[k - 1][i] - m->T[k - 2][i];
    }
    /* rietveld_core.js:1422-1424 guard range in tan(theta) */
    m->guard_lo = tan(dmax(1e-6, dmin(179.999999, m->tth[0])) * JS_PI / 360);
    m->guard_hi = tan(dmax(1e-6, dmin(179.999999, m->tth[n - 1])) * JS_PI / 360);
    m->key = (double *)malloc(sizeof(double) * (3 + 6 * prob->n_phases));
    m->params = (double *)malloc(sizeof(double) * (prob->n_params > 0 ? prob->n_params : 1));
    m->ybuf = (double *)malloc(sizeof(double) * n);
    m->col = (double *)malloc(sizeof(double) * n);
    return m;
}

void or_destroy(or_model *m) {
    if (!m) return;
    for (int p = 0; p < m->P.n_phases; p++) {
        free(m->atoms[p]); free(m->op_r[p]); free(m->op_t[p]); free(m->ph[p].rep); rl_free(&m->ph[p].refl);
    }
    if (m->P.bkg_type != PT_BKG_SPLINE)
        for (int k = 0; k < m->P.n_bkg; k++) free(m->T[k]);
    free(m->T); free(m->bx); free(m->atoms); free(m->op_r); free(m->op_t); free(m->ph); free(m->phases);
    free(m->tth); free(m->y); free(m->w); free(m->base); free(m->lo); free(m->hi); free(m->scale);
    free(m->x0); free(m->coeff); free(m->tparam); free(m->tvar); free(m->key); free(m->params);
    free(m->fstart); free(m->flen); free(m->fvals);
    free(m->ybuf); free(m->col); free(m);
}

/* apply(x), rietveld_core.js:1855-1859: every model parameter at the variables */
void or_params(const or_model *m, const double *x, double *p) {
    memcpy(p, m->base, sizeof(double) * m->P.n_params);
    for (int t = 0; t < m->P.n_terms; t++) p[m->tparam[t]] += m->coeff[t] * x[m->tvar[t]];
}

static int phase_base(const pt_problem *P, int phase) {
    int b = PT_P_PHASE0;
    for (int q = 0; q < phase; q++) b += PT_PH_ATOM0 + PT_AT_STRIDE * P->phases[q].n_atoms;
    return b;
}

/* Model.refreshReflections, rietveld_core.js:680-689: re-enumerate when the cell, zero,
 * disp or transp change */
static void refresh_reflections(or_model *m, const double *p) {
    int np = m->P.n_phases, changed = !m->has_key;
    double key[3 + 6 * 8];
    double *kk = np <= 8 ? key : (double *)malloc(sizeof(double) * (3 + 6 * np));
    kk[0] = p[PT_P_ZERO]; kk[1] = p[PT_P_DISP]; kk[2] = p[PT_P_TRANSP];
    for (int q = 0; q < np; q++) {
        const double *pp = p + phase_base(&m->P, q);
        for (int c = 0; c < 6; c++) kk[3 + 6 * q + c] = pp[c];
    }
    if (!changed) for (int i = 0; i < 3 + 6 * np; i++) if (memcmp(&kk[i], &m->key[i], sizeof(double))) { changed = 1; break; }
    if (changed) {
        double dm = or_reflection_dmin(&m->P, p[PT_P_ZERO], p[PT_P_DISP], p[PT_P_TRANSP]);
        for (int q = 0; q < np; q++) {
            double cell[6];
            const double *pp = p + phase_base(&m->P, q);
            if (cell_of(&m->phases[q], pp, cell))
                gen_reflections(cell, m->phases[q].n_ops, m->phases[q].op_r, m->phases[q].op_t, dm, NULL, &m->ph[q].refl);
            else m->ph[q].refl.n = 0;
        }
        memcpy(m->key, kk, sizeof(double) * (3 + 6 * np));
        m->has_key = 1;
    }
    if (kk != key) free(kk);
}

/* rietveld_core.js:383-390 formFactor */
static double form_factor(const pt_atom *a, int radiation, double s) {
    if (radiation == PT_NEUTRON) return js_finite(a->b_neutron) ? a->b_neutron : 0;
    double q = (s / 2) * (s / 2), f = a->ff_c;
    for (int i = 0; i < a->n_gauss; i++) f += a->ff_a[i] * exp(-a->ff_b[i] * q);
    return f;
}

/* per-reflection working data of one phase */
typedef struct {
    double tth, bragg, s, d, f2, po, lp, K, height;
    int has_col, start, stop;
    or_voigt v[2];
    double weight[2];
    int present[2], cstart[2], cstop[2];
} refl_work;

/* Model.structure, rietveld_core.js:749-821 (no gradients) */
static void structure(or_model *m, int q, const double *pp, refl_work *rw) {
    const pt_phase *ph = &m->phases[q];
    or_phase *op = &m->ph[q];
    int nA = ph->n_atoms, nR = op->nrep, rad = m->P.radiation;
    double pos[64][3];
    double *posd = nA <= 64 ? &pos[0][0] : (double *)malloc(sizeof(double) * 3 * nA);
    for (int j = 0; j < nA; j++) {
        const pt_atom *a = &ph->atoms[j];
        const double *ap = pp + PT_PH_ATOM0 + PT_AT_STRIDE * j;
        double r[3] = {a->r0[0], a->r0[1], a->r0[2]};     /* sitePosition, :372-378 */
        for (int f = 0; f < 3; f++) if (a->free_mask & (1 << f)) {
            double v = ap[f];
            r[0] += v * a->basis[f][0]; r[1] += v * a->basis[f][1]; r[2] += v * a->basis[f][2];
        }
        posd[3 * j] = r[0]; posd[3 * j + 1] = r[1]; posd[3 * j + 2] = r[2];
    }
    refl_list *L = &op->refl;
    for (int k = 0; k < L->n; k++) {
        refl_work *R = &rw[k];
        R->f2 = 0;
        if (!js_finite(R->tth)) continue;
        double s = R->s, s2 = s * s, Ar = 0, Br = 0, Cr = 0, Dr = 0;
        int h = L->r[k].h, kk = L->r[k].k, l = L->r[k].l;
        for (int j = 0; j < nA; j++) {
            const pt_atom *a = &ph->atoms[j];
            const double *ap = pp + PT_PH_ATOM0 + PT_AT_STRIDE * j;
            double occ = ap[PT_AT_OCC], B = ap[PT_AT_B];
            double w0 = (a->mult != 0 ? a->mult : 1) / nR;
            double dw = exp(-B * s2 / 4), f0 = form_factor(a, rad, s);
            int neutron = rad == PT_NEUTRON;
            double wj = occ * (neutron ? f0 : f0 + a->fp) * dw * w0;
            double wi = (neutron || !(a->fpp != 0)) ? 0 : occ * a->fpp * dw * w0;
            double c = 0, sn = 0;
            const double *p = posd + 3 * j;
            for (int qq = 0; qq < nR; qq++) {
                const int32_t *r = ph->op_r + 9 * op->rep[qq];
                const double *t = ph->op_t + 3 * op->rep[qq];
                double H0 = h * r[0] + kk * r[3] + l * r[6], H1 = h * r[1] + kk * r[4] + l * r[7], H2 = h * r[2] + kk * r[5] + l * r[8];
                double phs = TWO_PI * (H0 * p[0] + H1 * p[1] + H2 * p[2] + h * t[0] + kk * t[1] + l * t[2]);
                c += cos(phs); sn += sin(phs);
            }
            Ar += wj * c; Br += wj * sn; Cr += wi * c; Dr += wi * sn;
        }
        R->f2 = Ar * Ar + Br * Br + Cr * Cr + Dr * Dr;
    }
    if (posd != &pos[0][0]) free(posd);
}

/* Model.preferredOrientation, rietveld_core.js:824-843 */
static void preferred_orientation(or_model *m, int q, const double *pp, const double f[6], int have_f, refl_work *rw) {
    refl_list *L = &m->ph[q].refl;
    double r = pp[PT_PH_PO];
    for (int k = 0; k < L->n; k++) rw[k].po = 1;
    if (!(r > 0) || fabs(r - 1) < 1e-12 || !have_f) return;
    const double *H = m->phases[q].po_dir;
#define GMET(a, b) ((a)[0] * (b)[0] * f[0] + (a)[1] * (b)[1] * f[1] + (a)[2] * (b)[2] * f[2] + \
                    ((a)[0] * (b)[1] + (a)[1] * (b)[0]) * f[3] / 2 + ((a)[1] * (b)[2] + (a)[2] * (b)[1]) * f[4] / 2 + \
                    ((a)[0] * (b)[2] + (a)[2] * (b)[0]) * f[5] / 2)
    double HH = GMET(H, H), r2 = r * r, ir = 1 / r;
    if (!(HH > 0)) return;
    for (int k = 0; k < L->n; k++) {
        int mm = L->r[k].m;
        const int32_t *eq = L->eq + 3 * L->r[k].eq_off;
        double sum = 0;
        for (int e = 0; e < mm; e++) {
            double hv[3] = {eq[3 * e], eq[3 * e + 1], eq[3 * e + 2]};
            double hh = GMET(hv, hv), c = GMET(hv, H) / sqrt(hh * HH), c2 = dmin(1, c * c);
            sum += pow(r2 * c2 + (1 - c2) * ir, -1.5);
        }
        rw[k].po = sum / mm;
    }
#undef GMET
}

/* constants.js:506-513 lowerBound */
static int lower_bound(const double *a, int n, double target) {
    int lo = 0, hi = n;
    while (lo < hi) { int mid = (lo + hi) >> 1; if (a[mid] < target) lo = mid + 1; else hi = mid; }
    return lo;
}

int or_prepare_voigt(int type, const double *p, double minF, double tth_peak, double x0, or_voigt *o) {
    return or_prepare_voigt_sep(type, p, NULL, 0, minF, tth_peak, x0, o);
}

void or_gsas_widths(int type, const or_sep *sep, double lambda, double minF, double tth, int side, double *gG, double *gL) {
    consts();
    profile_widths(type, NULL, sep, lambda, minF, tth, side, gG, gL);
}

/* one phase: Model.geometry (:719-736), buildPeakContributions (profile.js:1013-1080),
 * evalPhase (:855-897), render (:908-917).  Adds the phase pattern into y. */
static int eval_phase(or_model *m, int q, const double *p, double *y, refl_work **rw_out, int *nr_out) {
    const pt_phase *ph = &m->phases[q];
    const double *pp = p + phase_base(&m->P, q);
    refl_list *L = &m->ph[q].refl;
    int n = m->P.n_points, nr = L->n;
    double cell[6], f[6];
    int have_f = cell_of(ph, pp, cell) && or_inv_dsq(cell, f);
    double lam = m->P.lambda1, disp = p[PT_P_DISP], transp = p[PT_P_TRANSP], zero = p[PT_P_ZERO];
    refl_work *rw = (refl_work *)calloc(nr > 0 ? nr : 1, sizeof(refl_work));
    /* geometry */
    for (int k = 0; k < nr; k++) {
        double qv = have_f ? inv_dsq(f, L->r[k].h, L->r[k].k, L->r[k].l) : NAN, s = sqrt(qv), sinT = lam * s / 2, tth = NAN;
        if (qv > 0 && sinT < 1) tth = 2 * asin(sinT) * R2D;
        double bragg = tth;
        if (js_finite(tth)) tth += or_peak_shift(tth, disp, transp);
        rw[k].tth = tth; rw[k].bragg = bragg; rw[k].s = s; rw[k].d = 1 / s;
    }
    /* contributions: all Kalpha1, then all Kalpha2 */
    const double *prof = pp + PT_PH_PROF0;
    /* separated profile (Model.profileParams, rietveld_core.js:694-708): instrument values
     * with their defaults, eta / SL / HL of the family from them, sample slots as given */
    const or_phase *oph = &m->ph[q];
    int fam = oph->fam;
    or_sep sp;
    double sslots[PT_PH_NPROF];
    if (oph->sep) {
        memset(&sp, 0, sizeof sp);
        for (int i = 0; i < PT_N_INST; i++) {
            double v = p[PT_P_INST0 + i];
            sp.inst[i] = js_finite(v) ? v : INST_DEFAULT[i];
        }
        sp.sample = prof;
        sp.size_model = ph->size_model; sp.strain_model = ph->strain_model;
        memcpy(sp.size_axis, oph->size_axis, sizeof sp.size_axis);
        memcpy(sp.strain_axis, oph->strain_axis, sizeof sp.strain_axis);
        memcpy(sp.cell, cell, sizeof cell);
        sp.n_basis = oph->nb; sp.basis_key = oph->bkey; sp.basis_coeff = oph->bcoef;
        for (int i = 0; i < PT_PH_NPROF; i++) sslots[i] = 0;
        sslots[PT_TCH_SL] = sp.inst[PT_INST_SL]; sslots[PT_TCH_HL] = sp.inst[PT_INST_HL];
        if (fam == PT_PROFILE_SIMPLE) sslots[PT_SIM_ETA] = sp.inst[PT_INST_ETA];
        if (fam == PT_PROFILE_SPLIT) sslots[PT_SPL_ETA] = sp.inst[PT_INST_ETA_SPLIT];
    }
    double l1 = m->P.lambda1 != 0 ? m->P.lambda1 : 1.54056, l2 = m->P.lambda2, ratio = m->P.ratio;
    int doublet = ratio > 1e-6 && l2 > 1e-6 && fabs(l1 - l2) > 1e-6;
    double lratio = doublet ? l2 / l1 : 1.0;
    for (int pass = 0; pass < (doublet ? 2 : 1); pass++) {
        for (int k = 0; k < nr; k++) {
            refl_work *R = &rw[k];
            if (!(js_finite(R->tth) && R->tth > 0 && R->tth < 180)) continue;   /* hasRenderablePosition */
            double bragg = js_finite(R->bragg) ? R->bragg : R->tth, tl = bragg;
            if (pass == 1) {
                double st2 = sin(bragg * (JS_PI / 180) / 2.0) * lratio;
                if (fabs(st2) >= 1) continue;
                tl = 2 * asin(st2) / (JS_PI / 180);
            }
            double base = tl + or_peak_shift(tl, disp, transp) + zero;
            double centre = base;   /* calculatePeakShift(base, params) is 0: no shft/trns in params */
            or_voigt *v = &R->v[pass];
            if (oph->sep) {
                sp.h[0] = L->r[k].h; sp.h[1] = L->r[k].k; sp.h[2] = L->r[k].l; sp.d = R->d;
                or_prepare_voigt_sep(fam, sslots, &sp, pass ? l2 : m->P.lambda1, m->P.min_fwhm, base, centre, v);
            } else {
                or_prepare_voigt(fam, prof, m->P.min_fwhm, base, centre, v);
            }
            R->present[pass] = 1;
            if (pass == 0) R->weight[0] = 1.0;
            else {
                double lpr = js_finite(m->P.pol_k) ? or_lp(base, m->P.pol_k) / or_lp(R->v[0].x0, m->P.pol_k) : 1;
                R->weight[1] = ratio * R->v[0].area / v->area * lpr;
            }
            /* push(): window in the axis */
            R->cstart[pass] = R->cstop[pass] = 0;
            if (!js_finite(v->max_d)) { R->present[pass] = 0; continue; }
            double lo = centre - v->max_d, hi = centre + v->max_d;
            int start = lower_bound(m->tth, n, lo);
            if (start >= n) { R->present[pass] = 0; continue; }
            int stop = start;
            while (stop < n && m->tth[stop] <= hi) stop++;
            if (stop <= start) { R->present[pass] = 0; continue; }
            R->cstart[pass] = start; R->cstop[pass] = stop;
        }
    }
    /* evalPhase: K = area1, Lp, structure, PO, heights */
    for (int k = 0; k < nr; k++) {
        refl_work *R = &rw[k];
        R->has_col = 0;
        if (!(R->present[0] || R->present[1]) || !js_finite(R->tth)) continue;
        double tc = R->tth + zero, a = R->v[0].area;
        if (!(a > 0)) continue;
        R->has_col = 1; R->K = a; R->lp = or_lp(tc, m->P.pol_k);
        R->start = R->present[0] ? R->cstart[0] : R->cstart[1];
        R->stop = R->present[0] ? R->cstop[0] : R->cstop[1];
        if (R->present[1]) { if (R->cstart[1] < R->start) R->start = R->cstart[1]; if (R->cstop[1] > R->stop) R->stop = R->cstop[1]; }
    }
    structure(m, q, pp, rw);
    preferred_orientation(m, q, pp, f, have_f, rw);
    double S = pp[PT_PH_SCALE];
    if (!(S != 0) || isnan(S)) S = 0;
    double *col = m->col;
    for (int k = 0; k < nr; k++) {
        refl_work *R = &rw[k];
        R->height = 0;
        if (!R->has_col) continue;
        double factor = L->r[k].m * R->lp * R->po / R->K;
        R->height = S * (factor * R->f2);
        /* intensityDerivativeColumn (profile.js:1131-1150), then render */
        int len = R->stop - R->start;
        for (int i = 0; i < len; i++) col[i] = 0;
        for (int pass = 0; pass < 2; pass++) {
            if (!R->present[pass]) continue;
            double wgt = R->weight[pass];
            for (int i = R->cstart[pass]; i < R->cstop[pass]; i++)
                col[i - R->start] += wgt * or_eval_voigt(m->tth[i], &R->v[pass]);
        }
        if (ph->free_intensities) {
            /* a free intensity: keep the column for the per-evaluation NNLS */
            if (m->nf == m->fcap) {
                m->fcap = m->fcap ? 2 * m->fcap : 64;
                m->fstart = (int *)realloc(m->fstart, sizeof(int) * m->fcap);
                m->flen = (int *)realloc(m->flen, sizeof(int) * m->fcap);
            }
            if (m->fvals_n + len > m->fvals_cap) {
                m->fvals_cap = 2 * (m->fvals_n + len);
                m->fvals = (double *)realloc(m->fvals, sizeof(double) * m->fvals_cap);
            }
            m->fstart[m->nf] = R->start; m->flen[m->nf] = len;
            memcpy(m->fvals + m->fvals_n, col, sizeof(double) * len);
            m->fvals_n += len; m->nf++;
            continue;
        }
        double hk = R->height;
        if (!(hk != 0)) continue;
        for (int i = 0; i < len; i++) y[R->start + i] += hk * col[i];
    }
    if (rw_out) { *rw_out = rw; *nr_out = nr; } else free(rw);
    return 0;
}

static int cell_ok(or_model *m, const double *p) {
    for (int q = 0; q < m->P.n_phases; q++) {
        double cell[6], f[6];
        const double *pp = p + phase_base(&m->P, q);
        if (!cell_of(&m->phases[q], pp, cell)) return 0;
        if (!(or_cell_volume(cell) > 0) || !or_inv_dsq(cell, f)) return 0;
    }
    return 1;
}
int or_cell_valid(or_model *m, const double *x) {
    or_params(m, x, m->params);
    return cell_ok(m, m->params);
}

/* rietveld_core.js:549-561 splineSlopes: Fritsch-Carlson tangents of the monotone cubic
 * Hermite spline through (xs, ys), limited to tau = a^2 + b^2 <= 9 per interval */
void or_spline_slopes(const double *xs, const double *ys, int na, double *h, double *mm) {
    if (na < 2) { if (na == 1) mm[0] = 0; return; }
    double *d = (double *)malloc(sizeof(double) * (na - 1));
    for (int i = 0; i < na - 1; i++) { h[i] = xs[i + 1] - xs[i]; d[i] = (ys[i + 1] - ys[i]) / h[i]; }
    mm[0] = d[0]; mm[na - 1] = d[na - 2];
    for (int i = 1; i < na - 1; i++) mm[i] = d[i - 1] * d[i] <= 0 ? 0 : (h[i] * d[i - 1] + h[i - 1] * d[i]) / (h[i - 1] + h[i]);
    for (int i = 0; i < na - 1; i++) {
        if (d[i] == 0) { mm[i] = 0; mm[i + 1] = 0; continue; }
        double a = mm[i] / d[i], b = mm[i + 1] / d[i], tau = a * a + b * b;
        if (tau > 9) { double f = 3 / sqrt(tau); mm[i] = a * d[i] * f; mm[i + 1] = b * d[i] * f; }
    }
    free(d);
}

/* rietveld_core.js:597-612 splineValues: the spline at every point of the sorted axis,
 * flat outside the anchors; one anchor gives a constant, none zero */
void or_spline_values(const double *tth, int n, const double *xs, const double *ys, int na, double *out) {
    for (int i = 0; i < n; i++) out[i] = 0;
    if (na < 1) return;
    if (na == 1) { for (int i = 0; i < n; i++) out[i] = ys[0]; return; }
    double *h = (double *)malloc(sizeof(double) * na), *mm = (double *)malloc(sizeof(double) * na);
    or_spline_slopes(xs, ys, na, h, mm);
    int seg = 0;
    for (int i = 0; i < n; i++) {
        double t = tth[i];
        if (t <= xs[0]) { out[i] = ys[0]; continue; }
        if (t >= xs[na - 1]) { out[i] = ys[na - 1]; continue; }
        while (seg < na - 2 && xs[seg + 1] <= t) seg++;
        while (seg > 0 && xs[seg] > t) seg--;
        double hi = h[seg], u = hi > 1e-12 ? (t - xs[seg]) / hi : 0, u2 = u * u, u3 = u2 * u;
        out[i] = (2 * u3 - 3 * u2 + 1) * ys[seg] + (u3 - 2 * u2 + u) * hi * mm[seg] + (-2 * u3 + 3 * u2) * ys[seg + 1] +
                 (u3 - u2) * hi * mm[seg + 1];
    }
    free(h); free(mm);
}

/* ---- free intensities (Le Bail / WPPF phases) ----------------------------------- */
/* Cholesky solve of the passive block of a symmetric positive definite H (row major,
 * m x m): H_PP z_P = b_P, z zero elsewhere.  Returns 0, or -1 when the block is not
 * positive definite. */
static int passive_solve(int m, const double *H, const double *b, const int *P, double *z, double *L, int *idx) {
    int p = 0;
    for (int j = 0; j < m; j++) { z[j] = 0; if (P[j]) idx[p++] = j; }
    for (int i = 0; i < p; i++)
        for (int j = 0; j <= i; j++) {
            double a = H[idx[i] * m + idx[j]];
            for (int k = 0; k < j; k++) a -= L[i * p + k] * L[j * p + k];
            if (i == j) { if (!(a > 0)) return -1; L[i * p + i] = sqrt(a); }
            else L[i * p + j] = a / L[j * p + j];
        }
    double *t = (double *)malloc(sizeof(double) * (p ? p : 1));
    for (int i = 0; i < p; i++) {
        double a = b[idx[i]];
        for (int k = 0; k < i; k++) a -= L[i * p + k] * t[k];
        t[i] = a / L[i * p + i];
    }
    for (int i = p - 1; i >= 0; i--) {
        double a = t[i];
        for (int k = i + 1; k < p; k++) a -= L[k * p + i] * t[k];
        t[i] = a / L[i * p + i];
    }
    for (int i = 0; i < p; i++) z[idx[i]] = t[i];
    free(t);
    return 0;
}

/* min over x >= 0 of x'Hx - 2 b'x for a symmetric positive semidefinite H (the normal
 * equations of a weighted least-squares fit): Lawson and Hanson's active-set algorithm
 * (Solving Least Squares Problems, 1974, ch. 23), with the columns equilibrated to unit
 * diagonal first as js/rietveld_core.js:1078-1082 does.  Columns with no weight
 * (H_jj <= 0) stay at zero.  Returns the outer iterations, or -1. */
int or_nnls_normal(int m, const double *H0, const double *b0, double *x) {
    if (m <= 0) return 0;
    double *H = (double *)malloc(sizeof(double) * m * m), *b = (double *)malloc(sizeof(double) * m);
    double *nrm = (double *)malloc(sizeof(double) * m), *w = (double *)malloc(sizeof(double) * m);
    double *z = (double *)malloc(sizeof(double) * m), *L = (double *)malloc(sizeof(double) * m * m);
    int *P = (int *)calloc(m, sizeof(int)), *idx = (int *)malloc(sizeof(int) * m), *blocked = (int *)calloc(m, sizeof(int));
    for (int j = 0; j < m; j++) { nrm[j] = H0[j * m + j] > 0 ? sqrt(H0[j * m + j]) : 0; if (!(nrm[j] > 0)) blocked[j] = 1; x[j] = 0; }
    double bmax = 0;
    for (int j = 0; j < m; j++) {
        b[j] = blocked[j] ? 0 : b0[j] / nrm[j];
        if (fabs(b[j]) > bmax) bmax = fabs(b[j]);
        for (int k = 0; k < m; k++) H[j * m + k] = blocked[j] || blocked[k] ? (j == k ? 1 : 0) : H0[j * m + k] / (nrm[j] * nrm[k]);
    }
    const double tol = 1e-12 * (bmax > 1 ? bmax : 1);
    int it = 0, rc = 0;
    for (; it < 3 * m + 10; it++) {
        for (int j = 0; j < m; j++) { double g = b[j]; for (int k = 0; k < m; k++) g -= H[j * m + k] * x[k]; w[j] = g; }
        int jmax = -1;
        for (int j = 0; j < m; j++) if (!P[j] && !blocked[j] && w[j] > tol && (jmax < 0 || w[j] > w[jmax])) jmax = j;
        if (jmax < 0) break;
        P[jmax] = 1;
        int inner = 0;
        for (; inner < 3 * m + 10; inner++) {
            if (passive_solve(m, H, b, P, z, L, idx)) { rc = -1; break; }
            int allpos = 1;
            for (int j = 0; j < m; j++) if (P[j] && !(z[j] > 0)) { allpos = 0; break; }
            if (allpos) { memcpy(x, z, sizeof(double) * m); break; }
            double alpha = 2;
            for (int j = 0; j < m; j++)
                if (P[j] && !(z[j] > 0)) { double a = x[j] / (x[j] - z[j]); if (a < alpha) alpha = a; }
            for (int j = 0; j < m; j++) if (P[j]) x[j] += alpha * (z[j] - x[j]);
            for (int j = 0; j < m; j++) if (P[j] && !(x[j] > 0)) { P[j] = 0; x[j] = 0; }
        }
        if (rc) break;
        /* a column that could not stay positive would be picked again: keep it out */
        if (!P[jmax]) blocked[jmax] = 1;
    }
    for (int j = 0; j < m; j++) x[j] = P[j] && nrm[j] > 0 ? x[j] / nrm[j] : 0;
    free(H); free(b); free(nrm); free(w); free(z); free(L); free(P); free(idx); free(blocked);
    return rc ? -1 : it;
}

/* the free intensities of the evaluation (js/pt_gpu.js:259-297; rietveld_core.js:1057-1090
 * wppf): the target y_obs - (background + Rietveld phases), the weighted normal
 * equations of the free columns, their non-negative least-squares solution, and the
 * columns rendered at it into y */
static void free_intensities(or_model *m, double *y) {
    const int nf = m->nf;
    double *H = (double *)calloc((size_t)nf * nf, sizeof(double)), *b = (double *)calloc(nf, sizeof(double));
    double *x = (double *)calloc(nf, sizeof(double));
    size_t *off = (size_t *)malloc(sizeof(size_t) * nf);
    size_t o = 0;
    for (int j = 0; j < nf; j++) { off[j] = o; o += m->flen[j]; }
    for (int j = 0; j < nf; j++) {
        const double *cj = m->fvals + off[j];
        const int sj = m->fstart[j], ej = sj + m->flen[j];
        double r = 0;
        for (int i = sj; i < ej; i++) r += cj[i - sj] * m->w[i] * (m->y[i] - y[i]);
        b[j] = r;
        for (int k = 0; k <= j; k++) {
            const double *ck = m->fvals + off[k];
            const int sk = m->fstart[k], ek = sk + m->flen[k];
            const int lo = sj > sk ? sj : sk, hi = ej < ek ? ej : ek;
            double a = 0;
            for (int i = lo; i < hi; i++) a += cj[i - sj] * ck[i - sk] * m->w[i];
            H[j * nf + k] = H[k * nf + j] = a;
        }
    }
    or_nnls_normal(nf, H, b, x);
    for (int j = 0; j < nf; j++) {
        if (!(x[j] > 0)) continue;
        const double *cj = m->fvals + off[j];
        for (int i = 0; i < m->flen[j]; i++) y[m->fstart[j] + i] += x[j] * cj[i];
    }
    free(H); free(b); free(x); free(off);
}

/* Model.background (:920-931) + evalDataset (:1152-1164) */
static int pattern_at_params(or_model *m, const double *p, double *y) {
    int n = m->P.n_points;
    refresh_reflections(m, p);
    for (int i = 0; i < n; i++) y[i] = 0;
    if (m->P.bkg_type == PT_BKG_SPLINE) {
        /* Model.splineY (:932-936): anchor j takes bga<j> (0 when unset) */
        double ys[PT_MAX_BKG];
        for (int j = 0; j < m->P.n_bkg; j++) ys[j] = isnan(p[PT_P_BKG0 + j]) ? 0 : p[PT_P_BKG0 + j];
        or_spline_values(m->tth, n, m->bx, ys, m->P.n_bkg, y);
    } else {
        for (int k = 0; k < m->P.n_bkg; k++) {
            double b = p[PT_P_BKG0 + k];
            if (!(b != 0) || isnan(b)) continue;
            for (int i = 0; i < n; i++) y[i] += b * m->T[k][i];
        }
    }
    /* each phase renders into its own array, then addTo(y, ph.y) */
    double *yp = (double *)malloc(sizeof(double) * n);
    m->nf = 0; m->fvals_n = 0;
    for (int q = 0; q < m->P.n_phases; q++) {
        for (int i = 0; i < n; i++) yp[i] = 0;
        eval_phase(m, q, p, yp, NULL, NULL);
        for (int i = 0; i < n; i++) y[i] += yp[i];
    }
    free(yp);
    if (m->nf) free_intensities(m, y);
    return 0;
}

int or_pattern(or_model *m, const double *x, double *ycalc) {
    or_params(m, x, m->params);
    if (!cell_ok(m, m->params)) return -1;
    return pattern_at_params(m, m->params, ycalc);
}

/* cost(), rietveld_core.js:1878-1889 */
double or_cost(or_model *m, const double *x) {
    or_params(m, x, m->params);
    if (!cell_ok(m, m->params)) return INFINITY;
    double *y = m->ybuf;
    pattern_at_params(m, m->params, y);
    double s = 0;
    for (int i = 0; i < m->P.n_points; i++) {
        if (!(m->w[i] > 0)) continue;
        double r = m->y[i] - y[i];
        s += m->w[i] * r * r;
    }
    if (!(s < INFINITY)) return INFINITY;
    return s;
}

int or_reflection_table(or_model *m, const double *x, int q, int32_t *hklm, double *f2, double *height,
                        double *centre, int cap) {
    or_params(m, x, m->params);
    if (!cell_ok(m, m->params)) return -1;
    refresh_reflections(m, m->params);
    double *yp = (double *)calloc(m->P.n_points, sizeof(double));
    refl_work *rw = NULL; int nr = 0;
    eval_phase(m, q, m->params, yp, &rw, &nr);
    refl_list *L = &m->ph[q].refl;
    for (int k = 0; k < nr && k < cap; k++) {
        if (hklm) { hklm[4 * k] = L->r[k].h; hklm[4 * k + 1] = L->r[k].k; hklm[4 * k + 2] = L->r[k].l; hklm[4 * k + 3] = L->r[k].m; }
        if (f2) f2[k] = rw[k].f2;
        if (height) height[k] = rw[k].height;
        if (centre) centre[k] = rw[k].v[0].x0;
    }
    free(rw); free(yp);
    return nr;
}

/* The contributions buildPeakContributions (profile.js:1013-1080) makes for phase q at x:
 * per reflection k and pass (0 Kalpha1, 1 Kalpha2), line[8 * k + 4 * pass + 0..3] =
 * present (0/1), centre x0, doublet weight, prepared area (preparedProfileArea); and
 * bragg[2 * k + pass] = the Bragg angle of the pass's wavelength (no shift, no zero). */
int or_line_table(or_model *m, const double *x, int q, double *line, double *bragg, int cap) {
    or_params(m, x, m->params);
    if (!cell_ok(m, m->params)) return -1;
    refresh_reflections(m, m->params);
    double *yp = (double *)calloc(m->P.n_points, sizeof(double));
    refl_work *rw = NULL; int nr = 0;
    eval_phase(m, q, m->params, yp, &rw, &nr);
    double l1 = m->P.lambda1 != 0 ? m->P.lambda1 : 1.54056, lr = m->P.lambda2 / l1;
    for (int k = 0; k < nr && k < cap; k++) {
        for (int pass = 0; pass < 2; pass++) {
            double *o = line + 8 * k + 4 * pass;
            o[0] = rw[k].present[pass]; o[1] = rw[k].v[pass].x0; o[2] = rw[k].weight[pass]; o[3] = rw[k].v[pass].area;
        }
        bragg[2 * k] = rw[k].bragg;
        double st2 = sin(rw[k].bragg * (JS_PI / 180) / 2.0) * lr;
        bragg[2 * k + 1] = fabs(st2) < 1 ? 2 * asin(st2) / (JS_PI / 180) : NAN;
    }
    free(rw); free(yp);
    return nr;
}

/* gaussianVarianceMinimum (:1393-1398) */
static double variance_min(double u, double v, double w, double lo, double hi) {
    double c[3] = {lo, hi, 0}; int nc = 2;
    if (u > 0) { double vx = -v / (2 * u); if (vx > lo && vx < hi) c[nc++] = vx; }
    double best = (u * c[0] + v) * c[0] + w;
    for (int i = 1; i < nc; i++) { double val = (u * c[i] + v) * c[i] + w; if (val < best) best = val; }
    return best;
}

/* The Stephens guard of a separated phase (profileVarianceGuards + guardVarianceMinimum,
 * rietveld_core.js:1456-1478): the strain quartic at the unit direction of every probe
 * h k l in [-2, 2] (one of each +-pair) and of every reflection of the current list; the
 * least value.  Every direction's coefficient of term n is the quartic with S_n = 1. */
static double stephens_guard_min(or_model *m, int q, const double *S) {
    const or_phase *oph = &m->ph[q];
    const refl_list *L = &oph->refl;
    int cap = 62 + L->n, nd = 0;
    int32_t *dirs = (int32_t *)malloc(sizeof(int32_t) * 3 * cap);
    for (int h = -2; h <= 2; h++) for (int k = -2; k <= 2; k++) for (int l = -2; l <= 2; l++)
        if (h > 0 || (h == 0 && (k > 0 || (k == 0 && l > 0)))) { dirs[3 * nd] = h; dirs[3 * nd + 1] = k; dirs[3 * nd + 2] = l; nd++; }
    for (int r = 0; r < L->n; r++) {
        int h = L->r[r].h, k = L->r[r].k, l = L->r[r].l, dup = 0;
        if (!h && !k && !l) continue;
        for (int e = 0; e < nd && !dup; e++) dup = dirs[3 * e] == h && dirs[3 * e + 1] == k && dirs[3 * e + 2] == l;
        if (!dup) { dirs[3 * nd] = h; dirs[3 * nd + 1] = k; dirs[3 * nd + 2] = l; nd++; }
    }
    double worst = INFINITY, one[15];
    for (int e = 0; e < nd; e++) {
        double hv[3] = {dirs[3 * e], dirs[3 * e + 1], dirs[3 * e + 2]}, len = hypot_n(hv, 3), unit[3], value = 0;
        for (int i = 0; i < 3; i++) unit[i] = hv[i] / len;
        for (int n = 0; n < oph->nb; n++) {
            for (int i = 0; i < 15; i++) one[i] = 0;
            one[oph->bkey[n]] = 1;
            double c = or_stephens_sum(unit, one, oph->nb, oph->bkey, oph->bcoef);
            double sv = js_finite(S[oph->bkey[n]]) ? S[oph->bkey[n]] : 0;
            value = value + c * sv;
        }
        if (value < worst) worst = value;
    }
    free(dirs);
    return worst;
}

/* gaussianVarianceGuards + varianceViolation (:1413-1496): every Gaussian polynomial
 * (a phase's own, or the dataset's instrument one for a separated phase, its values
 * defaulted), and the Stephens guards */
static double violation_at_params(or_model *m, const double *p) {
    double worst = 0;
    int stephens = 0;
    for (int q = 0; q < m->P.n_phases; q++) stephens |= m->ph[q].nb > 0;
    if (stephens) refresh_reflections(m, p);
    for (int q = 0; q < m->P.n_phases; q++) {
        const double *pr = p + phase_base(&m->P, q) + PT_PH_PROF0;
        const or_phase *oph = &m->ph[q];
        int type = oph->fam;
        int sets = type == PT_PROFILE_SPLIT ? 2 : 1;
        for (int s = 0; s < sets; s++) {
            double u, v, w;
            if (oph->sep) {
                int o = type == PT_PROFILE_SPLIT ? (s ? PT_INST_U_R : PT_INST_U_L) : PT_INST_U;
                double c[3];
                for (int i = 0; i < 3; i++) { double x = p[PT_P_INST0 + o + i]; c[i] = js_finite(x) ? x : INST_DEFAULT[o + i]; }
                u = c[0]; v = c[1]; w = c[2];
            }
            else if (type == PT_PROFILE_TCH) { u = pr[PT_TCH_U]; v = pr[PT_TCH_V]; w = pr[PT_TCH_W]; }
            else if (type == PT_PROFILE_SIMPLE) { double gp = pr[PT_SIM_GP]; u = pr[PT_SIM_GU] + gp; v = pr[PT_SIM_GV]; w = pr[PT_SIM_GW] + gp; }
            else { int o = s ? 4 : 0; u = pr[PT_SPL_GU_L + o]; v = pr[PT_SPL_GV_L + o]; w = pr[PT_SPL_GW_L + o]; }
            double mn = variance_min(u, v, w, m->guard_lo, m->guard_hi);
            if (mn < 0 && -mn > worst) worst = -mn;
        }
        if (oph->nb > 0) {
            double mn = stephens_guard_min(m, q, pr + PT_GS_S0);
            if (mn < 0 && -mn > worst) worst = -mn;
        }
    }
    return worst;
}
double or_violation(or_model *m, const double *x) {
    or_params(m, x, m->params);
    return violation_at_params(m, m->params);
}

/* temper(), rietveld_core.js:2171-2240 */
typedef struct { uint32_t s; } xs32;
static double xs_random(xs32 *r) {
    uint32_t v = r->s;
    v ^= v << 13; v ^= v >> 17; v ^= v << 5;
    r->s = v;
    return ((double)v + 0.5) / 4294967296.0;
}
static double xs_gauss(xs32 *r) {
    double u1 = xs_random(r);
    double a = sqrt(-2 * log(u1));
    double u2 = xs_random(r);
    return a * cos(2 * JS_PI * u2);
}
static double reflect_js(double value, double low, double high) {
    if (!js_finite(value)) return NAN;
    if (js_finite(low) && js_finite(high)) {
        double span = high - low;
        if (!(span > 0)) return low;
        double folded = fmod(fmod(value - low, 2 * span) + 2 * span, 2 * span);
        return low + (folded <= span ? folded : 2 * span - folded);
    }
    if (value < low) return low + (low - value);
    if (value > high) return high - (value - high);
    return value;
}

double or_temper(or_model *m, const double *x0, const double *scales, int sweeps, int chains, double max_temp,
                 uint32_t seed, double *best_x, double *stats) {
    int nv = m->P.n_vars;
    xs32 rng = {seed};
    double *xs = (double *)malloc(sizeof(double) * nv * chains), *costs = (double *)malloc(sizeof(double) * chains);
    int *feas = (int *)malloc(sizeof(int) * chains);
    double *temps = (double *)malloc(sizeof(double) * chains), *prop = (double *)malloc(sizeof(double) * nv);
    double initial = or_cost(m, x0);
    double startViolation = or_violation(m, x0);
    double bestCost = initial;
    memcpy(best_x, x0, sizeof(double) * nv);
    for (int r = 0; r < chains; r++) {
        temps[r] = pow(max_temp, (double)r / (chains - 1));
        memcpy(xs + r * nv, x0, sizeof(double) * nv);
        costs[r] = initial;
        feas[r] = startViolation == 0;
    }
    double accepted = 0, proposals = 0, swaps = 0, attempts = 0;
    for (int sweep = 0; sweep < sweeps; sweep++) {
        for (int r = 0; r < chains; r++) {
            double *x = xs + r * nv;
            for (int move = 0; move < nv; move++) {
                int c = (int)floor(xs_random(&rng) * nv);
                memcpy(prop, x, sizeof(double) * nv);
                double g = xs_gauss(&rng);
                prop[c] = reflect_js(prop[c] + g * scales[c] * sqrt(temps[r]), m->lo[c], m->hi[c]);
                if (!js_finite(prop[c])) continue;
                double cc = or_cost(m, prop);
                proposals++;
                double viol = or_violation(m, prop);
                if (viol > 0 && feas[r]) cc = INFINITY;
                if (js_finite(cc) && (cc <= costs[r] || log(xs_random(&rng)) < (costs[r] - cc) / (2 * temps[r]))) {
                    memcpy(x, prop, sizeof(double) * nv);
                    costs[r] = cc; feas[r] = !(viol > 0); accepted++;
                    if (cc < bestCost) { bestCost = cc; memcpy(best_x, prop, sizeof(double) * nv); }
                }
            }
        }
        for (int r = sweep % 2; r < chains - 1; r += 2) {
            double la = 0.5 * (1 / temps[r] - 1 / temps[r + 1]) * (costs[r] - costs[r + 1]);
            attempts++;
            if (la >= 0 || log(xs_random(&rng)) < la) {
                double *a = xs + r * nv, *b = xs + (r + 1) * nv;
                for (int j = 0; j < nv; j++) { double t = a[j]; a[j] = b[j]; b[j] = t; }
                double tc = costs[r]; costs[r] = costs[r + 1]; costs[r + 1] = tc;
                int tf = feas[r]; feas[r] = feas[r + 1]; feas[r + 1] = tf;
                swaps++;
            }
        }
    }
    if (stats) { stats[0] = accepted; stats[1] = proposals; stats[2] = swaps; stats[3] = attempts; }
    free(xs); free(costs); free(feas); free(temps); free(prop);
    return bestCost;
}

typedef struct {
    const pt_problem *prob;
    const double *scales;
    int sweeps, chains;
    double max_temp;
    uint32_t seed;
    double proposals;
} temper_job;

static void *temper_thread(void *arg) {
    temper_job *j = (temper_job *)arg;
    or_model *m = or_create(j->prob);
    double *best = (double *)malloc(sizeof(double) * j->prob->n_vars), st[4] = {0};
    or_temper(m, j->prob->x0, j->scales, j->sweeps, j->chains, j->max_temp, j->seed, best, st);
    j->proposals = st[1];
    free(best);
    or_destroy(m);
    return NULL;
}

double or_temper_parallel(const pt_problem *prob, const double *scales, int sweeps, int chains, double max_temp,
                          uint32_t seed, int threads, double *seconds) {
    if (threads < 1) threads = 1;
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * threads);
    temper_job *jobs = (temper_job *)calloc(threads, sizeof(temper_job));
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    for (int i = 0; i < threads; i++) {
        jobs[i].prob = prob; jobs[i].scales = scales; jobs[i].sweeps = sweeps; jobs[i].chains = chains;
        jobs[i].max_temp = max_temp; jobs[i].seed = seed + (uint32_t)i;
        pthread_create(&th[i], NULL, temper_thread, &jobs[i]);
    }
    double total = 0;
    for (int i = 0; i < threads; i++) { pthread_join(th[i], NULL); total += jobs[i].proposals; }
    clock_gettime(CLOCK_MONOTONIC, &t1);
    double sec = (t1.tv_sec - t0.tv_sec) + 1e-9 * (t1.tv_nsec - t0.tv_nsec);
    if (seconds) *seconds = sec;
    free(th); free(jobs);
    return total / sec;
}

/* ------------------------------------------------------------------------------ */
/*  the LM normal equations: Model.derivatives (rietveld_core.js:1245-1298) and      */
/*  normal() (:1953-1983), FP64                                                      */
/* ------------------------------------------------------------------------------ */
static const char *const OR_INST_NAMES[PT_N_INST] = {"U", "V", "W", "X", "Y", "Z", "SL", "HL", "eta", "U_L", "V_L",
                                                     "W_L", "X_L", "U_R", "V_R", "W_R", "X_R", "eta_split"};
static const char *const OR_TCH_NAMES[7] = {"U", "V", "W", "X", "Y", "SL", "HL"};
static const char *const OR_SIM_NAMES[6] = {"GU", "GV", "GW", "GP", "LX", "eta"};
static const char *const OR_SPL_NAMES[9] = {"GU_L", "GV_L", "GW_L", "LX_L", "GU_R", "GV_R", "GW_R", "LX_R", "eta_split"};
static const char *const OR_CELL_NAMES[6] = {"a", "b", "c", "alpha", "beta", "gamma"};
static const char *const OR_ATOM_COMP[5] = {"x", "y", "z", "occ", "B"};

/* The reference's name of model parameter k (the value-map names Model reads: zero,
 * bkg<k> / bga<k>, inst.<n>, p<q>.<cell|scale|po|profile slot>, p<q>.A<j>.<comp>). */
static void or_param_name(const or_model *m, int k, char *out, size_t cap) {
    if (k == PT_P_ZERO) { snprintf(out, cap, "zero"); return; }
    if (k == PT_P_DISP) { snprintf(out, cap, "disp"); return; }
    if (k == PT_P_TRANSP) { snprintf(out, cap, "transp"); return; }
    if (k < PT_P_INST0) { snprintf(out, cap, "%s%d", m->P.bkg_type == PT_BKG_SPLINE ? "bga" : "bkg", k - PT_P_BKG0); return; }
    if (k < PT_P_PHASE0) { snprintf(out, cap, "inst.%s", OR_INST_NAMES[k - PT_P_INST0]); return; }
    for (int q = 0; q < m->P.n_phases; q++) {
        int rel = k - phase_base(&m->P, q), na = m->phases[q].n_atoms;
        if (rel < 0 || rel >= PT_PH_ATOM0 + PT_AT_STRIDE * na) continue;
        if (rel < 6) { snprintf(out, cap, "p%d.%s", q + 1, OR_CELL_NAMES[rel]); return; }
        if (rel == PT_PH_SCALE) { snprintf(out, cap, "p%d.scale", q + 1); return; }
        if (rel == PT_PH_PO) { snprintf(out, cap, "p%d.po", q + 1); return; }
        if (rel < PT_PH_ATOM0) {
            int s = rel - PT_PH_PROF0, fam = m->phases[q].profile_type;
            const char *n = NULL;
            if (fam == PT_PROFILE_SIMPLE && s < 6) n = OR_SIM_NAMES[s];
            else if (fam == PT_PROFILE_SPLIT && s < 9) n = OR_SPL_NAMES[s];
            else if ((fam == PT_PROFILE_TCH || fam == PT_PROFILE_GSAS_CW) && s < 7) n = OR_TCH_NAMES[s];
            if (n) snprintf(out, cap, "p%d.%s", q + 1, n); else snprintf(out, cap, "p%d.slot%d", q + 1, s);
            return;
        }
        rel -= PT_PH_ATOM0;
        snprintf(out, cap, "p%d.A%d.%s", q + 1, rel / PT_AT_STRIDE, OR_ATOM_COMP[rel % PT_AT_STRIDE]);
        return;
    }
    snprintf(out, cap, "?");
}

/* classify (:470-484) */
enum { OK_UNKNOWN, OK_INSTRUMENT, OK_ZERO, OK_DISP, OK_TRANSP, OK_BKG, OK_BGA, OK_SCALE, OK_PO, OK_CELL, OK_PROFILE, OK_ATOM };
typedef struct { int kind, phase, index, atom; char comp[64]; } or_class;
static or_class or_classify(const char *name) {
    or_class c;
    memset(&c, 0, sizeof c);
    c.kind = OK_UNKNOWN; c.phase = -1;
    if (!strncmp(name, "inst.", 5)) { c.kind = OK_INSTRUMENT; snprintf(c.comp, sizeof c.comp, "%s", name + 5); return c; }
    if (!strcmp(name, "zero")) { c.kind = OK_ZERO; return c; }
    if (!strcmp(name, "disp")) { c.kind = OK_DISP; return c; }
    if (!strcmp(name, "transp")) { c.kind = OK_TRANSP; return c; }
    if (!strncmp(name, "bkg", 3)) { c.kind = OK_BKG; c.index = atoi(name + 3); return c; }
    if (!strncmp(name, "bga", 3)) { c.kind = OK_BGA; c.index = atoi(name + 3); return c; }
    if (name[0] == 'p') {
        int q = atoi(name + 1) - 1;
        const char *d1 = strchr(name, '.');
        if (!d1) return c;
        const char *d2 = strchr(d1 + 1, '.');
        c.phase = q;
        if (!d2) {
            const char *r = d1 + 1;
            if (!strcmp(r, "scale")) c.kind = OK_SCALE;
            else if (!strcmp(r, "po")) c.kind = OK_PO;
            else {
                c.kind = OK_PROFILE;
                for (int i = 0; i < 6; i++) if (!strcmp(r, OR_CELL_NAMES[i])) c.kind = OK_CELL;
            }
            snprintf(c.comp, sizeof c.comp, "%s", r);
            return c;
        }
        c.kind = OK_ATOM; c.atom = atoi(d1 + 2);
        snprintf(c.comp, sizeof c.comp, "%s", d2 + 1);
    }
    return c;
}
/* fdStep (:526-537) */
static double or_fd_step(const or_class *c, double v) {
    double a = fabs(v);
    switch (c->kind) {
    case OK_CELL: return (!strcmp(c->comp, "a") || !strcmp(c->comp, "b") || !strcmp(c->comp, "c")) ? dmax(1e-6 * a, 1e-6) : 1e-4;
    case OK_ZERO: return 1e-4;
    case OK_DISP: return 1e-2;
    case OK_TRANSP: return 1e-6;
    case OK_PO: return 1e-4;
    case OK_BGA: return dmax(1e-4 * a, 1e-3);
    default: return dmax(1e-4 * a, 1e-5);
    }
}
static int or_is(const char *s, const char *const *list, int n) {
    for (int i = 0; i < n; i++) if (!strcmp(s, list[i])) return 1;
    return 0;
}
/* defaultLimits (:486-508) */
static void or_default_limits(const or_class *c, double v, double *lo, double *hi) {
    *lo = -INFINITY; *hi = INFINITY;
    const char *s = c->comp;
    switch (c->kind) {
    case OK_SCALE: *lo = 0; break;
    case OK_CELL:
        if (!strcmp(s, "a") || !strcmp(s, "b") || !strcmp(s, "c")) { *lo = 0.5 * v; *hi = 2 * v; } else { *lo = 30; *hi = 150; }
        break;
    case OK_ATOM:
        if (!strcmp(s, "occ")) { *lo = 0; *hi = 2; } else if (!strcmp(s, "B")) { *lo = -2; *hi = 30; }
        break;
    case OK_PO: *lo = 0.1; *hi = 5; break;
    case OK_ZERO: *lo = -2; *hi = 2; break;
    case OK_INSTRUMENT: {
        static const char *const pos[] = {"X", "X_L", "X_R", "Y", "Z", "SL"};
        if (!strncmp(s, "eta", 3)) { *lo = 0; *hi = 1; } else if (or_is(s, pos, 6)) *lo = 0;
        break;
    }
    case OK_PROFILE: {
        static const char *const sizes[] = {"size", "sizeAxial", "sizeB", "sizeC"}, *mixes[] = {"sizeMix", "strainMix"},
                                 *quart[] = {"S400", "S040", "S004"}, *pos[] = {"LX", "LX_L", "LX_R", "X", "Y", "GP"};
        if (or_is(s, sizes, 4)) { *lo = 0.001; *hi = 10; }
        else if (or_is(s, mixes, 2)) { *lo = 0; *hi = 1; }
        else if (!strncmp(s, "strain", 6)) { *lo = 0; *hi = 1e6; }
        else if (or_is(s, quart, 3)) *lo = 0;
        else if (!strncmp(s, "eta", 3)) { *lo = 0; *hi = 1; }
        else if (or_is(s, pos, 6)) *lo = 0;
        break;
    }
    default: break;
    }
}

/* combine(ph.cols, dh, n) (:1192-1200): sum over the reflections of dh_k times the
 * reflection's unit-height column (intensityDerivativeColumn, profile.js:1131-1150) */
static void or_combine(or_model *m, const refl_work *rw, int nr, const double *dh, double *out) {
    for (int k = 0; k < nr; k++) {
        const refl_work *R = &rw[k];
        if (!R->has_col || !(dh[k] != 0)) continue;
        int len = R->stop - R->start;
        double *col = m->col;
        for (int i = 0; i < len; i++) col[i] = 0;
        for (int pass = 0; pass < 2; pass++) {
            if (!R->present[pass]) continue;
            for (int i = R->cstart[pass]; i < R->cstop[pass]; i++) col[i - R->start] += R->weight[pass] * or_eval_voigt(m->tth[i], &R->v[pass]);
        }
        for (int i = 0; i < len; i++) out[R->start + i] += dh[k] * col[i];
    }
}

/* the gradient slot of Model.structure with grad (:749-821) for atom j, component comp
 * (0-2 a free coordinate, 3 occ, 4 B): G[k] = d|F_k|^2 / d(slot) */
static void or_structure_grad(or_model *m, int q, const double *pp, const refl_work *rw, int ja, int comp, double *G) {
    const pt_phase *ph = &m->phases[q];
    or_phase *op = &m->ph[q];
    int nA = ph->n_atoms, nR = op->nrep, rad = m->P.radiation, neutron = rad == PT_NEUTRON;
    double *pos = (double *)malloc(sizeof(double) * 3 * (nA > 0 ? nA : 1));
    for (int j = 0; j < nA; j++) {
        const pt_atom *a = &ph->atoms[j];
        const double *ap = pp + PT_PH_ATOM0 + PT_AT_STRIDE * j;
        double r[3] = {a->r0[0], a->r0[1], a->r0[2]};
        for (int f = 0; f < 3; f++) if (a->free_mask & (1 << f)) for (int i = 0; i < 3; i++) r[i] += ap[f] * a->basis[f][i];
        for (int i = 0; i < 3; i++) pos[3 * j + i] = r[i];
    }
    refl_list *L = &op->refl;
    for (int k = 0; k < L->n; k++) {
        G[k] = 0;
        if (!js_finite(rw[k].tth)) continue;
        double s = rw[k].s, s2 = s * s, Ar = 0, Br = 0, Cr = 0, Dr = 0, Aj = 0, Bj = 0, Cj = 0, Dj = 0;
        double dA[3] = {0, 0, 0}, dB[3] = {0, 0, 0}, dC[3] = {0, 0, 0}, dD[3] = {0, 0, 0};
        int h = L->r[k].h, kk = L->r[k].k, l = L->r[k].l;
        for (int j = 0; j < nA; j++) {
            const pt_atom *a = &ph->atoms[j];
            const double *ap = pp + PT_PH_ATOM0 + PT_AT_STRIDE * j;
            double occ = ap[PT_AT_OCC], B = ap[PT_AT_B], w0 = (a->mult != 0 ? a->mult : 1) / nR;
            double dw = exp(-B * s2 / 4), f0 = form_factor(a, rad, s);
            double wj = occ * (neutron ? f0 : f0 + a->fp) * dw * w0, wi = (neutron || !(a->fpp != 0)) ? 0 : occ * a->fpp * dw * w0;
            double c = 0, sn = 0, dc[3] = {0, 0, 0}, ds[3] = {0, 0, 0};
            for (int qq = 0; qq < nR; qq++) {
                const int32_t *r = ph->op_r + 9 * op->rep[qq];
                const double *t = ph->op_t + 3 * op->rep[qq];
                double H[3] = {h * r[0] + kk * r[3] + l * r[6], h * r[1] + kk * r[4] + l * r[7], h * r[2] + kk * r[5] + l * r[8]};
                double phs = TWO_PI * (H[0] * pos[3 * j] + H[1] * pos[3 * j + 1] + H[2] * pos[3 * j + 2] + h * t[0] + kk * t[1] + l * t[2]);
                double cs = cos(phs), si = sin(phs);
                c += cs; sn += si;
                for (int i = 0; i < 3; i++) { dc[i] -= H[i] * si; ds[i] += H[i] * cs; }
            }
            Ar += wj * c; Br += wj * sn; Cr += wi * c; Dr += wi * sn;
            if (j == ja) {
                Aj = wj * c; Bj = wj * sn; Cj = wi * c; Dj = wi * sn;
                for (int i = 0; i < 3; i++) {
                    dA[i] = TWO_PI * wj * dc[i]; dB[i] = TWO_PI * wj * ds[i]; dC[i] = TWO_PI * wi * dc[i]; dD[i] = TWO_PI * wi * ds[i];
                }
            }
        }
        const pt_atom *a = &ph->atoms[ja];
        const double *ap = pp + PT_PH_ATOM0 + PT_AT_STRIDE * ja;
        if (comp < 3) {
            const double *b = a->basis[comp];
            double dAr = dA[0] * b[0] + dA[1] * b[1] + dA[2] * b[2], dBr = dB[0] * b[0] + dB[1] * b[1] + dB[2] * b[2];
            double dCr = dC[0] * b[0] + dC[1] * b[1] + dC[2] * b[2], dDr = dD[0] * b[0] + dD[1] * b[1] + dD[2] * b[2];
            G[k] = 2 * (Ar * dAr + Br * dBr + Cr * dCr + Dr * dDr);
        } else if (comp == 3) {
            double oc = ap[PT_AT_OCC];
            if (fabs(oc) > 1e-12) G[k] = 2 * (Ar * Aj + Br * Bj + Cr * Cj + Dr * Dj) / oc;
            else {                                   /* occ = 0: the contribution per unit occupancy */
                double f1 = form_factor(a, rad, s), dw1 = exp(-ap[PT_AT_B] * s2 / 4), w0 = (a->mult != 0 ? a->mult : 1) / nR;
                double w1 = (neutron ? f1 : f1 + a->fp) * dw1 * w0, w1i = neutron ? 0 : a->fpp * dw1 * w0, c1 = 0, s1 = 0;
                for (int qq = 0; qq < nR; qq++) {
                    const int32_t *r = ph->op_r + 9 * op->rep[qq];
                    const double *t = ph->op_t + 3 * op->rep[qq];
                    double p1 = TWO_PI * ((h * r[0] + kk * r[3] + l * r[6]) * pos[3 * ja] + (h * r[1] + kk * r[4] + l * r[7]) * pos[3 * ja + 1] +
                                          (h * r[2] + kk * r[5] + l * r[8]) * pos[3 * ja + 2] + h * t[0] + kk * t[1] + l * t[2]);
                    c1 += cos(p1); s1 += sin(p1);
                }
                G[k] = 2 * (Ar * w1 * c1 + Br * w1 * s1 + Cr * w1i * c1 + Dr * w1i * s1);
            }
        } else {
            G[k] = 2 * (Ar * Aj + Br * Bj + Cr * Cj + Dr * Dj) * (-s2 / 4);
        }
    }
    free(pos);
}

/* the background at parameters p (Model.background, :920-936) */
static void or_background_at(or_model *m, const double *p, double *y) {
    int n = m->P.n_points;
    for (int i = 0; i < n; i++) y[i] = 0;
    if (m->P.bkg_type == PT_BKG_SPLINE) {
        double ys[PT_MAX_BKG];
        for (int j = 0; j < m->P.n_bkg; j++) ys[j] = isnan(p[PT_P_BKG0 + j]) ? 0 : p[PT_P_BKG0 + j];
        or_spline_values(m->tth, n, m->bx, ys, m->P.n_bkg, y);
    } else {
        for (int k = 0; k < m->P.n_bkg; k++) {
            double b = p[PT_P_BKG0 + k];
            if (!(b != 0) || isnan(b)) continue;
            for (int i = 0; i < n; i++) y[i] += b * m->T[k][i];
        }
    }
}

int or_normal(or_model *m, const double *x, double *A, double *g, double *f, double *J) {
    consts();
    const int n = m->P.n_points, nv = m->P.n_vars, npar = m->P.n_params, nph = m->P.n_phases;
    for (int q = 0; q < nph; q++) if (m->phases[q].free_intensities) return -2;   /* variable projection: not restated */
    or_params(m, x, m->params);
    if (!cell_ok(m, m->params)) return -1;
    double *p = dup_d(m->params, npar), *pm = dup_d(m->params, npar);
    refresh_reflections(m, p);                          /* evalDataset; derivatives keep this list */
    /* the evaluation at p: every phase's pattern and working data, the background */
    double **yph = (double **)calloc(nph, sizeof(double *));
    refl_work **rw = (refl_work **)calloc(nph, sizeof(refl_work *));
    int *nr = (int *)calloc(nph, sizeof(int));
    double *bkg = (double *)calloc(n, sizeof(double)), *yb = (double *)calloc(n, sizeof(double));
    double *tmp = (double *)calloc(n, sizeof(double)), *tmp2 = (double *)calloc(n, sizeof(double));
    for (int q = 0; q < nph; q++) {
        yph[q] = (double *)calloc(n, sizeof(double));
        m->nf = 0; m->fvals_n = 0;
        eval_phase(m, q, p, yph[q], &rw[q], &nr[q]);
    }
    or_background_at(m, p, bkg);
    /* the columns of the model parameters the variables act on */
    double **col = (double **)calloc(npar, sizeof(double *));
    for (int t = 0; t < m->P.n_terms; t++) {
        int k = m->tparam[t];
        if (col[k]) continue;
        col[k] = (double *)calloc(n, sizeof(double));
        char name[64];
        or_param_name(m, k, name, sizeof name);
        or_class c = or_classify(name);
        double v = p[k];
        int q = c.phase;
        if (c.kind == OK_BKG) {
            if (m->P.bkg_type != PT_BKG_SPLINE && c.index < m->P.n_bkg) memcpy(col[k], m->T[c.index], sizeof(double) * n);
        } else if (c.kind == OK_BGA) {
            if (m->P.bkg_type == PT_BKG_SPLINE && c.index < m->P.n_bkg) {
                double ys[PT_MAX_BKG];
                for (int j = 0; j < m->P.n_bkg; j++) ys[j] = isnan(p[PT_P_BKG0 + j]) ? 0 : p[PT_P_BKG0 + j];
                double hb = or_fd_step(&c, ys[c.index]);
                ys[c.index] += hb;
                or_spline_values(m->tth, n, m->bx, ys, m->P.n_bkg, tmp);
                for (int i = 0; i < n; i++) col[k][i] = (tmp[i] - bkg[i]) / hb;
            }
        } else if (c.kind == OK_SCALE || c.kind == OK_ATOM || c.kind == OK_PO) {
            const double *pp = p + phase_base(&m->P, q);
            double S = pp[PT_PH_SCALE];
            if (!(S != 0) || isnan(S)) S = 0;
            double *dh = (double *)calloc(nr[q] > 0 ? nr[q] : 1, sizeof(double));
            const refl_list *L = &m->ph[q].refl;
            int ok = 1;
            if (c.kind == OK_SCALE) {
                for (int kk = 0; kk < nr[q]; kk++) if (rw[q][kk].has_col) dh[kk] = L->r[kk].m * rw[q][kk].lp * rw[q][kk].po / rw[q][kk].K * rw[q][kk].f2;
            } else if (c.kind == OK_ATOM) {
                int comp = !strcmp(c.comp, "x") ? 0 : !strcmp(c.comp, "y") ? 1 : !strcmp(c.comp, "z") ? 2 : !strcmp(c.comp, "occ") ? 3 : 4;
                if (comp < 3 && !(m->phases[q].atoms[c.atom].free_mask & (1 << comp))) ok = 0;   /* no gradient slot */
                if (ok) {
                    double *G = (double *)calloc(nr[q] > 0 ? nr[q] : 1, sizeof(double));
                    or_structure_grad(m, q, pp, rw[q], c.atom, comp, G);
                    for (int kk = 0; kk < nr[q]; kk++) if (rw[q][kk].has_col) dh[kk] = S * (L->r[kk].m * rw[q][kk].lp * rw[q][kk].po / rw[q][kk].K) * G[kk];
                    free(G);
                }
            } else {
                double h = or_fd_step(&c, v), cell[6], ff[6];
                memcpy(pm, p, sizeof(double) * npar);
                pm[k] = v + h;
                const double *ppm = pm + phase_base(&m->P, q);
                int have_f = cell_of(&m->phases[q], ppm, cell) && or_inv_dsq(cell, ff);
                refl_work *rw2 = (refl_work *)malloc(sizeof(refl_work) * (nr[q] > 0 ? nr[q] : 1));
                memcpy(rw2, rw[q], sizeof(refl_work) * nr[q]);
                preferred_orientation(m, q, ppm, ff, have_f, rw2);
                for (int kk = 0; kk < nr[q]; kk++)
                    dh[kk] = rw[q][kk].po > 0 ? rw[q][kk].height * (rw2[kk].po / rw[q][kk].po - 1) / h : 0;
                free(rw2);
            }
            if (ok) or_combine(m, rw[q], nr[q], dh, col[k]);
            free(dh);
        } else if (c.kind != OK_UNKNOWN) {
            /* symmetric finite difference of the phases the parameter moves (:1275-1296) */
            double h0 = or_fd_step(&c, v), lo, hi;
            or_default_limits(&c, v, &lo, &hi);
            int plus = v + h0 <= hi, minus = v - h0 >= lo;
            double vp = plus ? v + h0 : v, vm = minus ? v - h0 : v, d = vp - vm;
            if (d != 0) {
                for (int qq = 0; qq < nph; qq++) {
                    if ((c.kind == OK_CELL || c.kind == OK_PROFILE) && qq != q) continue;
                    const double *ep = yph[qq], *em = yph[qq];
                    if (plus) {
                        memcpy(pm, p, sizeof(double) * npar); pm[k] = vp;
                        for (int i = 0; i < n; i++) tmp[i] = 0;
                        eval_phase(m, qq, pm, tmp, NULL, NULL);
                        ep = tmp;
                    }
                    if (minus) {
                        memcpy(pm, p, sizeof(double) * npar); pm[k] = vm;
                        for (int i = 0; i < n; i++) tmp2[i] = 0;
                        eval_phase(m, qq, pm, tmp2, NULL, NULL);
                        em = tmp2;
                    }
                    for (int i = 0; i < n; i++) col[k][i] += (ep[i] - em[i]) / d;
                }
            }
        }
    }
    /* the variables' columns through the links, and normal() */
    for (int i = 0; i < n; i++) { yb[i] = bkg[i]; for (int q = 0; q < nph; q++) yb[i] += yph[q][i]; }
    double *Jl = J ? J : (double *)calloc((size_t)nv * n, sizeof(double));
    for (size_t i = 0; i < (size_t)nv * n; i++) Jl[i] = 0;
    for (int t = 0; t < m->P.n_terms; t++) {
        int k = m->tparam[t], j = m->tvar[t];
        for (int i = 0; i < n; i++) Jl[(size_t)j * n + i] += m->coeff[t] * col[k][i];
    }
    double fs = 0;
    for (int i = 0; i < n; i++) { if (!(m->w[i] > 0)) { tmp[i] = 0; continue; } tmp[i] = m->y[i] - yb[i]; fs += m->w[i] * tmp[i] * tmp[i]; }
    for (int a = 0; a < nv; a++) {
        const double *Ja = Jl + (size_t)a * n;
        double ga = 0;
        for (int i = 0; i < n; i++) ga += m->w[i] * Ja[i] * tmp[i];
        g[a] = ga;
        for (int b = 0; b <= a; b++) {
            const double *Jb = Jl + (size_t)b * n;
            double s = 0;
            for (int i = 0; i < n; i++) s += m->w[i] * Ja[i] * Jb[i];
            A[(size_t)a * nv + b] = s;
            A[(size_t)b * nv + a] = s;
        }
    }
    *f = fs;
    if (!J) free(Jl);
    for (int k = 0; k < npar; k++) free(col[k]);
    for (int q = 0; q < nph; q++) { free(yph[q]); free(rw[q]); }
    free(col); free(yph); free(rw); free(nr); free(bkg); free(yb); free(tmp); free(tmp2); free(p); free(pm);
    return 0;
}
