// pt_device.h -- device math of one PT trial: parameters, cell metric, per-line profile
// preparation, structure factors, the banded profile sum and the weighted chi^2.
//
// Precision plan (the FP64 oracle is the reference):
//   * reciprocal metric, 1/d^2, Bragg angles and line centres in FP64 (a centre error of
//     1e-7 deg is 1e-5 of a 0.01 deg FWHM);
//   * structure-factor phases h.(R r + t) in FP64 reduced to [0,1) turns, then FP32
//     sincospi; widths, areas, Lp, heights in FP32;
//   * per-point profile values in FP32 with delta = t_i - centre formed from a local
//     grid origin (no cancellation), chi^2 with Neumaier-compensated FP32 sums.
#pragma once
#include <hip/hip_runtime.h>
#include "pt_plan.h"

namespace ptk {

constexpr double kPi = 3.141592653589793;
constexpr double kD2R = kPi / 180.0;
constexpr double kR2D = 180.0 / kPi;
constexpr float kPiF = 3.14159265358979f;
constexpr float kCg = 2.772588722239781f;               // 4 ln 2, profile.js:606
constexpr float kLog2e = 1.4426950408889634f;
constexpr float kG0 = 0.9394372786996513f;               // 2 sqrt(ln2/pi), constants.js:155
constexpr float kL0 = 0.6366197723675814f;               // 2/pi, constants.js:162
constexpr float kGArea = 1.0644670194312262f;            // 1/G0
constexpr float kLArea = 1.5707963267948966f;            // 1/L0
constexpr float kGConv = 0.023548200450309493f;          // sqrt(8 ln2)/100, constants.js:248
constexpr float kLConv = 0.01f;                          // constants.js:251
constexpr float kWinG = 8.0f, kWinL = 25.0f, kWinMax = 10.0f;   // constants.js:59,70,84
constexpr double BIG = 3e38;

// ---- scalar helpers ---------------------------------------------------------------
__device__ __forceinline__ float fmaxf_js(float a, float b) { return a > b ? a : b; }
__device__ __forceinline__ float fminf_js(float a, float b) { return a < b ? a : b; }

// constants.js:138-143 softPositive
__device__ __forceinline__ float soft_positive(float x, float f) {
    if (!isfinite(x)) return f;
    float h = hypotf(x, 2.0f * f);
    return x < 0.0f ? f * (f / (0.5f * h - 0.5f * x)) : 0.5f * x + 0.5f * h;
}
// profile.js:182-189 softFloor
__device__ __forceinline__ float soft_floor(float x, float f) {
    if (!isfinite(x) || !(x > 0.0f)) return f;
    if (!(f > 0.0f)) return x;
    float r = x / f;
    if (r > 20.0f) return x;
    float r2 = r * r;
    return f * sqrtf(sqrtf(r2 * r2 + 1.0f));
}
// profile.js:552-565 getPeakFWHM
__device__ __forceinline__ float tch_fwhm(float g, float l) {
    float gG = fmaxf_js(1e-9f, (g != 0.0f && !isnan(g)) ? g : 1e-9f);
    float gL = fmaxf_js(1e-9f, (l != 0.0f && !isnan(l)) ? l : 1e-9f);
    float g2 = gG * gG, l2 = gL * gL;
    float s = g2 * g2 * gG + 2.69269f * g2 * g2 * gL + 2.42843f * g2 * gG * l2 + 4.47163f * g2 * l2 * gL +
              0.07842f * gG * l2 * l2 + l2 * l2 * gL;
    if (s < 0.0f || !isfinite(s)) return fmaxf_js(fmaxf_js(gG, gL), 1e-6f);
    return fmaxf_js(1e-6f, powf(s, 0.2f));
}
// constants.js:194-201 pvMixCoeffs
__device__ __forceinline__ void pv_mix(float eta, float hg, float hl, float &cl, float &cg) {
    float e = fmaxf_js(0.0f, fminf_js(1.0f, eta));
    float wl = e * kL0 / fmaxf_js(1e-12f, hl);
    float wg = (1.0f - e) * kG0 / fmaxf_js(1e-12f, hg);
    float s = wl + wg;
    if (!(s > 0.0f) || !isfinite(s)) { cl = e; cg = 1.0f - e; return; }
    cl = wl / s; cg = wg / s;
}
// profile.js:149-164 gaussianFromSigmaSq / lorentzianFromGamma
__device__ __forceinline__ float gauss_width(float s2, float fmin) {
    float fl = 0.25f * fmin / kGConv;
    return sqrtf(soft_positive(s2, fl * fl)) * kGConv;
}
__device__ __forceinline__ float lorentz_width(float g, float fmin) {
    return soft_positive(g, 0.25f * fmin / kLConv) * kLConv;
}
__device__ __forceinline__ float orz(float v) { return (v != 0.0f && !isnan(v)) ? v : 0.0f; }

// One line of a separated (GSAS) profile: the dataset's instrument values (defaults
// applied, PT_INST_* order), the reflection's sample terms, the line's wavelength.
struct SepLine {
    const float *inst;         // nullptr: not a separated profile
    float inv_size, strain;    // 1/size along the reflection, microstrain of the reflection
    float sm, em;              // sizeMix, strainMix
    float lam;
};

// profile.js:501-537 calculateProfileWidths (side 0 = left, 1 = right / centre); with a
// SepLine the separated profile's gsasCwComponents (profile.js:384-422): the family's
// instrument Caglioti and Lorentzian terms plus size 1.8 lam / (pi D cos) and strain
// 0.018 e tan / pi, split between the Gaussian and Lorentzian parts by the mixes
__device__ inline void profile_widths(int type, const float *p, float fmin_in, float tth, int side,
                                      float &gG, float &gL, const SepLine sl = SepLine{nullptr, 0.f, 0.f, 0.f, 0.f, 0.f}) {
    float fmin = fmin_in > 0.0f ? fmin_in : 1e-3f;
    float th = tth * (kPiF / 180.0f) * 0.5f;
    float safe = fminf_js(th, 1.5707953267948966f);
    if (safe < 1e-6f) { gG = fmin; gL = fmin; return; }
    float sn, cs;
    sincosf(safe, &sn, &cs);
    float tn = sn / cs;
    float cs_safe = fmaxf_js(cs, 1e-9f), cs2_safe = fmaxf_js(cs * cs, 1e-9f);
    float g, l;
    if (sl.inst) {
        const float *I = sl.inst;
        float U, V, W, X, Y, Z;
        if (type == 2) {
            const int o = side == 0 ? 9 : 13;          // U_L.. or U_R..
            U = I[o]; V = I[o + 1]; W = I[o + 2]; X = I[o + 3]; Y = 0.0f; Z = 0.0f;
        } else {
            U = I[0]; V = I[1]; W = I[2]; X = I[3]; Y = I[4]; Z = I[5];
        }
        const float sg = 1.8f * sl.lam * sl.inv_size / (kPiF * cs_safe);
        const float eg = 0.018f * sl.strain * tn / kPiF;
        const float a = sg * (1.0f - sl.sm), b = eg * (1.0f - sl.em);
        const float s2 = U * tn * tn + V * tn + W + (a * a + b * b) * 0.18033688011112042f;   // 1 / (8 ln 2)
        g = gauss_width(s2, fmin);
        l = lorentz_width(X / cs_safe + Y * tn + Z + sg * sl.sm + eg * sl.em, fmin);
    } else if (type == 1) {
        float s2 = orz(p[0]) * tn * tn + orz(p[1]) * tn + orz(p[2]) + orz(p[3]) / cs2_safe;
        g = gauss_width(s2, fmin);
        l = lorentz_width(orz(p[4]) / cs_safe, fmin);
    } else if (type == 2) {
        int o = side == 0 ? 0 : 4;
        float s2 = orz(p[o]) * tn * tn + orz(p[o + 1]) * tn + orz(p[o + 2]);
        g = gauss_width(s2, fmin);
        l = lorentz_width(orz(p[o + 3]) / cs_safe, fmin);
    } else {
        float s2 = orz(p[0]) * tn * tn + orz(p[1]) * tn + orz(p[2]);
        g = gauss_width(s2, fmin);
        l = lorentz_width(orz(p[3]) / cs_safe + orz(p[4]) * tn, fmin);
    }
    gG = soft_floor(g, fmin);
    gL = soft_floor(l, fmin);
}

// A prepared line in FP32 (profile.js:603-718), flank 0 = left (delta < 0), 1 = right.
struct Prep {
    float ig[2], il[2], cl[2], cg[2];
    float asym, pedestal, max_d, window, fwhm, area;
};

// profile.js:733-759 evalVoigt without the window test (delta relative to the centre)
__device__ __forceinline__ float raw_profile(float delta, const Prep &p) {
    if (p.asym != 0.0f && fabsf(delta) >= 1e-9f) {
        float t = fmaxf_js(-0.95f, fminf_js(0.95f, p.asym * delta));
        delta /= (1.0f - t);
    }
    const bool left = delta < 0.0f;           // selects, not an indexed load (no scratch)
    const float ig = left ? p.ig[0] : p.ig[1], il = left ? p.il[0] : p.il[1];
    const float cl = left ? p.cl[0] : p.cl[1], cg = left ? p.cg[0] : p.cg[1];
    float u = delta * ig, v = delta * il;
    float g = __expf(-kCg * u * u);
    float l = 1.0f / (1.0f + 4.0f * v * v);
    return cl * l + cg * g;
}

// NR erfcc, profile.js:791-797
__device__ __forceinline__ float erfc_nr(float x) {
    float z = fabsf(x), t = 1.0f / (1.0f + 0.5f * z);
    float ans = t * expf(-z * z - 1.26551223f + t * (1.00002368f + t * (0.37409196f + t * (0.09678418f +
                t * (-0.18628806f + t * (0.27886807f + t * (-1.13520398f + t * (1.48851587f +
                t * (-0.82215223f + t * 0.17087277f)))))))));
    return x >= 0.0f ? ans : 2.0f - ans;
}

// profile.js:799-831 asymmetryAreaCorrection: doubling Gauss-Legendre panels out to D,
// closed-form tails beyond.
__device__ inline float asym_area(const Prep &pk) {
    const float X[4] = {0.1834346424956498f, 0.5255324099163290f, 0.7966664774136267f, 0.9602898564975363f};
    const float W[4] = {0.3626837833783620f, 0.3137066458778873f, 0.2223810344533745f, 0.1012285362903763f};
    Prep sym = pk;
    sym.asym = 0.0f;
    float a = pk.asym, H = pk.fwhm, w = fmaxf_js(1e-6f, pk.window);
    float clampAt = 0.95f / fabsf(a), D = fmaxf_js(pk.max_d, fminf_js(clampAt, 4096.0f * w));
    float total = 0.0f;
    for (int si = 0; si < 2; ++si) {
        float side = si ? 1.0f : -1.0f, lo = 0.0f, e = 0.25f * w;
        for (int guard = 0; guard < 64; ++guard) {
            float hi = fminf_js(e, D);
            // a panel that straddles the clamp point is split there (profile.js:821)
            for (int part = 0; part < 2; ++part) {
                float pa, pb;
                if (part == 0) {
                    if (!(clampAt > lo && clampAt < hi)) continue;
                    pa = lo; pb = clampAt; lo = clampAt;
                } else {
                    pa = lo; pb = hi;
                }
                float m = 0.5f * (pa + pb), h = 0.5f * (pb - pa), s = 0.0f;
                for (int i = 0; i < 4; ++i) {
                    float d1 = side * (m - h * X[i]), d2 = side * (m + h * X[i]);
                    s += W[i] * (raw_profile(d1, pk) - raw_profile(d1, sym) + raw_profile(d2, pk) - raw_profile(d2, sym));
                }
                total += s * h;
            }
            lo = hi;
            if (hi >= D) break;
            e *= 2.0f;
        }
        float k = 1.0f - fmaxf_js(-0.95f, fminf_js(0.95f, a * side * D));
        float sq = sqrtf(kPiF / kCg), sc = sqrtf(kCg);
        float x1 = D / k, x2 = D;
        float t1 = pk.cl[1] * 0.5f * H * (0.5f * kPiF - atanf(2.0f * x1 / H)) + pk.cg[1] * 0.5f * H * sq * erfc_nr(sc * x1 / H);
        float t2 = pk.cl[1] * 0.5f * H * (0.5f * kPiF - atanf(2.0f * x2 / H)) + pk.cg[1] * 0.5f * H * sq * erfc_nr(sc * x2 / H);
        total += k * t1 - t2;
    }
    return total;
}

// profile.js:603-718 prepareVoigt (+ preparedProfileArea): widths evaluated at tth_peak.
__device__ inline void prepare_line(int type, const float *prof, float fmin, float tth_peak, Prep &o,
                                    const SepLine sl = SepLine{nullptr, 0.f, 0.f, 0.f, 0.f, 0.f}) {
    float asym = 0.0f;
    if (type == 3 && (orz(prof[5]) != 0.0f || orz(prof[6]) != 0.0f) && tth_peak >= 0.1f && tth_peak < 180.0f) {
        float th = tth_peak * (kPiF / 180.0f) * 0.5f;
        float safe = fmaxf_js(1e-6f, fminf_js(1.5707953267948966f, th));
        float tn = tanf(safe);
        if (fabsf(tn) >= 1e-9f) asym = orz(prof[5]) / tn + orz(prof[6]);
    }
    o.asym = asym;
    float eta, fwhm_total, analytic;
    if (type == 1) {
        float gG, gL;
        profile_widths(1, prof, fmin, tth_peak, 1, gG, gL, sl);
        float HG = fmaxf_js(1e-9f, gG), HL = fmaxf_js(1e-9f, gL);
        eta = fmaxf_js(0.0f, fminf_js(1.0f, isfinite(prof[5]) ? prof[5] : 0.5f));
        fwhm_total = tch_fwhm(HG, HL);
        float cl, cg;
        pv_mix(eta, HG, HL, cl, cg);
        o.ig[0] = o.ig[1] = 1.0f / HG; o.il[0] = o.il[1] = 1.0f / HL;
        o.cl[0] = o.cl[1] = cl; o.cg[0] = o.cg[1] = cg;
        analytic = cl * kLArea * HL + cg * kGArea * HG;
    } else if (type == 2) {
        float gGL, gLL, gGR, gLR;
        profile_widths(2, prof, fmin, tth_peak, 0, gGL, gLL, sl);
        profile_widths(2, prof, fmin, tth_peak, 1, gGR, gLR, sl);
        float HGL = fmaxf_js(1e-9f, gGL), HLL = fmaxf_js(1e-9f, gLL), HGR = fmaxf_js(1e-9f, gGR), HLR = fmaxf_js(1e-9f, gLR);
        eta = fmaxf_js(0.0f, fminf_js(1.0f, isfinite(prof[8]) ? prof[8] : 0.5f));
        fwhm_total = fmaxf_js(tch_fwhm(HGL, HLL), tch_fwhm(HGR, HLR));
        float clL, cgL, clR, cgR;
        pv_mix(eta, HGL, HLL, clL, cgL);
        pv_mix(eta, HGR, HLR, clR, cgR);
        o.ig[0] = 1.0f / HGL; o.il[0] = 1.0f / HLL; o.ig[1] = 1.0f / HGR; o.il[1] = 1.0f / HLR;
        o.cl[0] = clL; o.cg[0] = cgL; o.cl[1] = clR; o.cg[1] = cgR;
        analytic = 0.5f * (clL * kLArea * HLL + cgL * kGArea * HGL) + 0.5f * (clR * kLArea * HLR + cgR * kGArea * HGR);
    } else {
        float gG, gL;
        profile_widths(3, prof, fmin, tth_peak, 1, gG, gL, sl);
        float HG = fmaxf_js(1e-9f, gG), HL = fmaxf_js(1e-9f, gL);
        float fw = fmaxf_js(1e-9f, tch_fwhm(HG, HL));
        eta = 0.0f;
        if (fw > 1e-9f) {
            float r = HL / fw;
            eta = fmaxf_js(0.0f, fminf_js(1.0f, 1.36603f * r - 0.47719f * (r * r) + 0.11116f * (r * r * r)));
        }
        fwhm_total = fw;
        float cl, cg;
        pv_mix(eta, fw, fw, cl, cg);
        o.ig[0] = o.ig[1] = o.il[0] = o.il[1] = 1.0f / fw;
        o.cl[0] = o.cl[1] = cl; o.cg[0] = o.cg[1] = cg;
        analytic = (cl * kLArea + cg * kGArea) * fw;
    }
    o.fwhm = fwhm_total;
    o.window = fmaxf_js(1e-6f, fwhm_total);
    float mult = kWinG + (kWinL - kWinG) * fmaxf_js(0.0f, fminf_js(1.0f, eta));
    float R = fminf_js(kWinMax, mult * fmaxf_js(0.01f, o.window) * (asym != 0.0f ? 2.0f : 1.0f));
    o.max_d = R;
    o.pedestal = 0.0f;
    o.pedestal = fmaxf_js(raw_profile(R, o), raw_profile(-R, o));
    o.area = analytic;
    if (type == 3 && asym != 0.0f) o.area += asym_area(o);
}

// crystal.js:465-497 sharkoLp at 2theta (degrees), K
__device__ __forceinline__ float lp_factor(float tth, float K) {
    if (!(tth > 0.0f) || !(tth < 180.0f)) return 1.0f;
    float th = 0.5f * tth * (kPiF / 180.0f);
    float s, c;
    sincosf(th, &s, &c);
    if (s < 1e-6f || c < 1e-6f) return 1.0f;
    float L = 1.0f / (s * s * c);
    float k = (isfinite(K) && K >= 0.0f) ? K : 1.0f;
    float c2 = cosf(tth * (kPiF / 180.0f));
    return L * ((1.0f + k * c2 * c2) / (1.0f + k));
}

// profile.js:108-132 sample displacement / transparency shift (deg), FP32 (|shift| < 1 deg)
__device__ __forceinline__ float peak_shift(double tth, float disp, float transp) {
    double th = tth * kD2R * 0.5;
    if (fabs(th - kPi * 0.5) < 1e-6) return 0.0f;
    float s, c;
    sincosf((float)th, &s, &c);
    float t = (-(disp / 1000.0f) * c + transp * (2.0f * s * c)) * (float)kR2D;
    return isfinite(t) ? t : 0.0f;
}

// constants.js:506-513 lowerBound over the 2theta axis
__device__ __forceinline__ int lower_bound_t(const DevPlan &P, double target) {
    int lo = 0, hi = P.np;
    if (P.uniform) {
        // on a uniform grid the positions are computed, not loaded: t_i = t0 + i step (the
        // plan accepts a grid as uniform within 1e-9 step of that)
        double g = ceil((target - P.t0) * P.inv_step);
        int i = g < 0.0 ? 0 : g > (double)P.np ? P.np : (int)g;
        while (i > 0 && P.t0 + (i - 1) * P.step >= target) --i;
        while (i < P.np && P.t0 + i * P.step < target) ++i;
        return i;
    }
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (P.tth[mid] < target) lo = mid + 1; else hi = mid;
    }
    return lo;
}
// first index >= start with t > target
__device__ __forceinline__ int upper_stop(const DevPlan &P, int start, double target) {
    if (P.uniform) {
        double g = floor((target - P.t0) * P.inv_step) + 1.0;
        int i = g < (double)start ? start : g > (double)P.np ? P.np : (int)g;
        while (i > start && P.t0 + (i - 1) * P.step > target) --i;
        while (i < P.np && P.t0 + i * P.step <= target) ++i;
        return i;
    }
    int lo = start, hi = P.np;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (P.tth[mid] <= target) lo = mid + 1; else hi = mid;
    }
    return lo;
}

// crystal.js:326-366 buildInvDsq for the phase cell at parameters pp (phase block)
// (cellOf / normaliseCell of crystal.js:76-119).  Returns false for a degenerate cell.
__device__ inline bool phase_metric(const DevPhase &ph, const double *pp, double f[6]) {
    double a = pp[0];
    if (!(a > 0.0)) return false;
    double b = pp[1] > 0.0 ? pp[1] : a, c = pp[2] > 0.0 ? pp[2] : a;
    double al = 90.0, be = 90.0, ga = 90.0;
    switch (ph.system) {
    case 6: b = c = a; break;
    case 3: b = a; break;
    case 4: case 5: b = a; ga = 120.0; break;
    case 2: break;
    case 1: {
        int ax = ph.unique_axis;
        double v = pp[3 + (ax == 0 ? 0 : ax == 2 ? 2 : 1)];
        v = isfinite(v) ? v : 90.0;
        if (ax == 0) al = v; else if (ax == 2) ga = v; else be = v;
        break;
    }
    default:
        al = isfinite(pp[3]) ? pp[3] : 90.0;
        be = isfinite(pp[4]) ? pp[4] : 90.0;
        ga = isfinite(pp[5]) ? pp[5] : 90.0;
    }
    double ca, cb, cg, sa, sb, sg;
    if (ph.fixed_angles) {
        // angles that do not move: their cosines and sines are computed once on the host
        al = ph.ang[0]; be = ph.ang[1]; ga = ph.ang[2];
        ca = ph.ang[3]; cb = ph.ang[4]; cg = ph.ang[5];
        sa = ph.ang_sin[0]; sb = ph.ang_sin[1]; sg = ph.ang_sin[2];
    } else {
        ca = cos(al * kD2R); cb = cos(be * kD2R); cg = cos(ga * kD2R);
        sa = sin(al * kD2R); sb = sin(be * kD2R); sg = sin(ga * kD2R);
    }
    double vf = 1.0 - ca * ca - cb * cb - cg * cg + 2.0 * ca * cb * cg;
    // cellVolume (crystal.js:245-252) validity + buildInvDsq degeneracy
    if (!(vf > 0.0) || !(al > 0.0 && al < 180.0 && be > 0.0 && be < 180.0 && ga > 0.0 && ga < 180.0)) return false;
    if (!isfinite(a) || !isfinite(b) || !isfinite(c) || !(b > 0.0) || !(c > 0.0)) return false;
    if (vf <= 1e-9 || fabs(sa) < 1e-9 || fabs(sb) < 1e-9 || fabs(sg) < 1e-9) return false;
    double V = a * b * c * sqrt(vf);
    double as = b * c * sa / V, bs = a * c * sb / V, cs = a * b * sg / V;
    double cA = (cb * cg - ca) / (sb * sg), cB = (ca * cg - cb) / (sa * sg), cG = (ca * cb - cg) / (sa * sb);
    f[0] = as * as; f[1] = bs * bs; f[2] = cs * cs;
    f[3] = 2.0 * as * bs * cG; f[4] = 2.0 * bs * cs * cA; f[5] = 2.0 * as * cs * cB;
    return true;
}

// rietveld_core.js:1393-1398 gaussianVarianceMinimum; returns the minimum value
__device__ __forceinline__ double variance_min(double u, double v, double w, double lo, double hi) {
    double best = (u * lo + v) * lo + w, t = (u * hi + v) * hi + w;
    if (t < best) best = t;
    if (u > 0.0) {
        double vx = -v / (2.0 * u);
        if (vx > lo && vx < hi) { double m = (u * vx + v) * vx + w; if (m < best) best = m; }
    }
    return best;
}

// the profile-feasibility rule (rietveld_core.js:1413-1496): every Gaussian polynomial
// of every phase non-negative over the fitted tan(theta) range
template <bool SEP = true>
__device__ inline bool profile_valid(const DevPlan &P, const double *p) {
    for (int q = 0; q < P.n_phases; ++q) {
        const double *pr = p + P.phase[q].param_base + 8;
        int type = P.phase[q].profile_type;
        for (int s = 0; s < (type == 2 ? 2 : 1); ++s) {
            double u, v, w;
            if (SEP && P.sp->ph[q].sep) {
                // the dataset's instrument polynomial (U, V, W or a split flank's), a value
                // that is not finite read as its default 2, -2, 5 (rietveld_core.js:1483-1488)
                const double *in = p + PTK_P_INST0 + (type == 2 ? (s ? 13 : 9) : 0);
                u = isfinite(in[0]) ? in[0] : 2.0; v = isfinite(in[1]) ? in[1] : -2.0; w = isfinite(in[2]) ? in[2] : 5.0;
            }
            else if (type == 3) { u = pr[0]; v = pr[1]; w = pr[2]; }
            else if (type == 1) { u = pr[0] + pr[3]; v = pr[1]; w = pr[2] + pr[3]; }
            else { int o = s ? 4 : 0; u = pr[o]; v = pr[o + 1]; w = pr[o + 2]; }
            if (variance_min(u, v, w, P.guard_lo, P.guard_hi) < 0.0) return false;
        }
    }
    return true;
}

// ---- counter-based random numbers (SplitMix64 finaliser, Steele et al. 2014) -------
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}
// uniform double in (0,1) for (seed, chain, counter, draw)
__device__ __forceinline__ double uniform01(uint32_t seed, uint32_t chain, uint64_t counter, uint32_t draw) {
    uint64_t k = mix64(((uint64_t)seed << 32) ^ 0x9e3779b97f4a7c15ull) ^ mix64(((uint64_t)chain << 40) ^ ((uint64_t)draw << 56) ^ counter);
    uint64_t r = mix64(k);
    return ((double)(r >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

// js/pt_gpu.js reflect / rietveld_core.js:2180-2190 reflect
__device__ __forceinline__ double reflect_box(double v, double lo, double hi) {
    if (!isfinite(v)) return NAN;
    if (isfinite(lo) && isfinite(hi)) {
        double span = hi - lo;
        if (!(span > 0.0)) return lo;
        double folded = fmod(fmod(v - lo, 2.0 * span) + 2.0 * span, 2.0 * span);
        return lo + (folded <= span ? folded : 2.0 * span - folded);
    }
    if (v < lo) return lo + (lo - v);
    if (v > hi) return hi - (v - hi);
    return v;
}

}  // namespace ptk
