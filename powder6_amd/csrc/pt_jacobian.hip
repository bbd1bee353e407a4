// pt_jacobian.hip -- FP64 derivative columns of the calculated pattern and the normal
// equations of one Levenberg-Marquardt step, on the GPU (js/rietveld_core.js:1245-1298
// Model.derivatives, :1953-1983 normal).  The LM finish stays in double precision: the
// reference's finite-difference steps are as small as 1e-6 of a cell edge, so the line
// geometry, the profiles and every sum here are FP64 (the PT trial kernels are FP32).
//
//   k_jac_lines     one thread per (evaluation, reflection): FP64 Bragg angle and sample
//                   shift, widths, prepared pseudo-Voigt (pedestal, area, asymmetry), Lp,
//                   March-Dollase, |F|^2 (or its gradient), the line height of the
//                   evaluation's kind, the Kalpha2 companion
//   k_jac_patterns  one workgroup per (256-point tile, evaluation): the evaluation's
//                   pattern -- its lines' banded sum, or a background column
//   k_jac_combine   the variables' columns J = C Y and the residual r = y_obs - y
//   k_jac_ext       [J; r] W [J; r]' (packed), one workgroup per entry: A, g and chi^2
#include "pt_device.h"
#include "pt_jacobian.h"

namespace ptk {
namespace {

constexpr double kCg64 = 2.772588722239781;         // 4 ln 2, profile.js:606
constexpr double kG064 = 0.9394372786996513;        // 2 sqrt(ln2/pi), constants.js:155
constexpr double kL064 = 0.6366197723675814;        // 2/pi, constants.js:162
constexpr double kGArea64 = 1.0644670194312262;     // 1/PV_G0
constexpr double kLArea64 = 1.5707963267948966;     // 1/PV_L0
constexpr double kGConv64 = 0.023548200450309493;   // sqrt(8 ln 2)/100, constants.js:248
constexpr double kLConv64 = 0.01;                   // constants.js:251
constexpr double kMaxTheta = 1.5707953267948966;    // pi/2 - 1e-6, profile.js:510-514

__device__ __forceinline__ double jmax(double a, double b) { return a > b ? a : b; }
__device__ __forceinline__ double jmin(double a, double b) { return a < b ? a : b; }
__device__ __forceinline__ double jorz(double v) { return (v != 0.0 && !isnan(v)) ? v : 0.0; }

// constants.js:138-143 softPositive
__device__ double soft_positive64(double x, double f) {
    if (!isfinite(x)) return f;
    const double h = hypot(x, 2.0 * f);
    return x < 0.0 ? f * (f / (0.5 * h - 0.5 * x)) : 0.5 * x + 0.5 * h;
}
// profile.js:182-189 softFloor
__device__ double soft_floor64(double x, double f) {
    if (!isfinite(x) || !(x > 0.0)) return f;
    if (!(f > 0.0)) return x;
    const double r = x / f;
    if (r > 20.0) return x;
    return f * pow(pow(r, 4.0) + 1.0, 0.25);
}
// profile.js:552-565 getPeakFWHM (Thompson, Cox & Hastings 1987)
__device__ double tch_fwhm64(double g, double l) {
    const double gG = jmax(1e-9, (g != 0.0 && !isnan(g)) ? g : 1e-9), gL = jmax(1e-9, (l != 0.0 && !isnan(l)) ? l : 1e-9);
    const double g2 = gG * gG, l2 = gL * gL;
    const double s = g2 * g2 * gG + 2.69269 * g2 * g2 * gL + 2.42843 * g2 * gG * l2 + 4.47163 * g2 * l2 * gL +
                     0.07842 * gG * l2 * l2 + l2 * l2 * gL;
    if (s < 0.0 || !isfinite(s)) return jmax(jmax(gG, gL), 1e-6);
    return jmax(1e-6, pow(s, 0.2));
}
// constants.js:194-201 pvMixCoeffs
__device__ void pv_mix64(double eta, double hg, double hl, double &cl, double &cg) {
    const double e = jmax(0.0, jmin(1.0, eta));
    const double wl = e * kL064 / jmax(1e-12, hl), wg = (1.0 - e) * kG064 / jmax(1e-12, hg), s = wl + wg;
    if (!(s > 0.0) || !isfinite(s)) { cl = e; cg = 1.0 - e; return; }
    cl = wl / s; cg = wg / s;
}
// profile.js:149-164 gaussianFromSigmaSq / lorentzianFromGamma
__device__ double gauss_width64(double s2, double fmin) {
    const double fl = 0.25 * fmin / kGConv64;
    return sqrt(soft_positive64(s2, fl * fl)) * kGConv64;
}
__device__ double lorentz_width64(double g, double fmin) { return soft_positive64(g, 0.25 * fmin / kLConv64) * kLConv64; }

// profile.js:501-537 calculateProfileWidths (TCH, simple, split; side 0 left, 1 right)
__device__ void profile_widths64(int type, const double *p, double fmin_in, double tth, int side, double &gG, double &gL) {
    const double fmin = fmin_in > 0.0 ? fmin_in : 1e-3;
    const double th = tth * (kPi / 180.0) * 0.5, safe = jmin(th, kMaxTheta);
    if (!(safe >= 1e-6)) { gG = fmin; gL = fmin; return; }
    const double tn = tan(safe), cs = cos(safe), cs_safe = jmax(cs, 1e-9), cs2_safe = jmax(cs * cs, 1e-9);
    double g, l;
    if (type == 1) {                         // profile.js:234-254
        const double s2 = jorz(p[0]) * tn * tn + jorz(p[1]) * tn + jorz(p[2]) + jorz(p[3]) / cs2_safe;
        g = gauss_width64(s2, fmin);
        l = lorentz_width64(jorz(p[4]) / cs_safe, fmin);
    } else if (type == 2) {                  // profile.js:270-298
        const int o = side == 0 ? 0 : 4;
        const double s2 = jorz(p[o]) * tn * tn + jorz(p[o + 1]) * tn + jorz(p[o + 2]);
        g = gauss_width64(s2, fmin);
        l = lorentz_width64(jorz(p[o + 3]) / cs_safe, fmin);
    } else {                                 // profile.js:426-482
        const double s2 = jorz(p[0]) * tn * tn + jorz(p[1]) * tn + jorz(p[2]);
        g = gauss_width64(s2, fmin);
        l = lorentz_width64(jorz(p[3]) / cs_safe + jorz(p[4]) * tn, fmin);
    }
    gG = soft_floor64(g, fmin);
    gL = soft_floor64(l, fmin);
}

struct Prep64 {
    double ig[2], il[2], cl[2], cg[2];
    double asym, ped, max_d, window, fwhm, area;
};

// profile.js:733-759 evalVoigt without the window test and the pedestal
__device__ double raw_profile64(double delta, const Prep64 &p) {
    if (p.asym != 0.0 && fabs(delta) >= 1e-9) delta /= (1.0 - jmax(-0.95, jmin(0.95, p.asym * delta)));
    const int s = delta < 0.0 ? 0 : 1;
    const double u = delta * p.ig[s], v = delta * p.il[s];
    return p.cl[s] / (1.0 + 4.0 * v * v) + p.cg[s] * exp(-kCg64 * u * u);
}

// profile.js:791-797 sharkoErfc (Numerical Recipes erfcc)
__device__ double erfc_nr64(double x) {
    const double z = fabs(x), t = 1.0 / (1.0 + 0.5 * z);
    const double a = t * exp(-z * z - 1.26551223 + t * (1.00002368 + t * (0.37409196 + t * (0.09678418 +
                     t * (-0.18628806 + t * (0.27886807 + t * (-1.13520398 + t * (1.48851587 +
                     t * (-0.82215223 + t * 0.17087277)))))))));
    return x >= 0.0 ? a : 2.0 - a;
}

// profile.js:799-831 asymmetryAreaCorrection: doubling Gauss-Legendre panels (8-point
// rule, profile.js:762-763) out to D, a panel that straddles the clamp point split there,
// closed-form tails beyond
__device__ double asym_area64(const Prep64 &pk) {
    const double X[4] = {0.1834346424956498, 0.5255324099163290, 0.7966664774136267, 0.9602898564975363};
    const double W[4] = {0.3626837833783620, 0.3137066458778873, 0.2223810344533745, 0.1012285362903763};
    Prep64 sym = pk;
    sym.asym = 0.0;
    const double a = pk.asym, H = pk.fwhm, w = jmax(1e-6, pk.window);
    const double clampAt = 0.95 / fabs(a), D = jmax(pk.max_d, jmin(clampAt, 4096.0 * w));
    double total = 0.0;
    for (int si = 0; si < 2; ++si) {
        const double side = si ? 1.0 : -1.0;
        double lo = 0.0, e = 0.25 * w;
        for (int guard = 0; guard < 64; ++guard) {
            const double hi = jmin(e, D);
            for (int part = 0; part < 2; ++part) {
                double pa, pb;
                if (part == 0) {
                    if (!(clampAt > lo && clampAt < hi)) continue;
                    pa = lo; pb = clampAt; lo = clampAt;
                } else {
                    pa = lo; pb = hi;
                }
                const double m = 0.5 * (pa + pb), h = 0.5 * (pb - pa);
                double s = 0.0;
                for (int i = 0; i < 4; ++i) {
                    const double d1 = side * (m - h * X[i]), d2 = side * (m + h * X[i]);
                    s += W[i] * (raw_profile64(d1, pk) - raw_profile64(d1, sym) + raw_profile64(d2, pk) - raw_profile64(d2, sym));
                }
                total += s * h;
            }
            lo = hi;
            if (hi >= D) break;
            e *= 2.0;
        }
        const double k = 1.0 - jmax(-0.95, jmin(0.95, a * side * D)), sq = sqrt(kPi / kCg64), sc = sqrt(kCg64);
        const double x1 = D / k, x2 = D;
        const double t1 = pk.cl[1] * 0.5 * H * (0.5 * kPi - atan(2.0 * x1 / H)) + pk.cg[1] * 0.5 * H * sq * erfc_nr64(sc * x1 / H);
        const double t2 = pk.cl[1] * 0.5 * H * (0.5 * kPi - atan(2.0 * x2 / H)) + pk.cg[1] * 0.5 * H * sq * erfc_nr64(sc * x2 / H);
        total += k * t1 - t2;
    }
    return total;
}

// profile.js:603-718 prepareVoigt (+ preparedProfileArea), widths at tth_peak
__device__ void prepare64(int type, const double *prof, double fmin, double tth_peak, Prep64 &o) {
    double asym = 0.0;
    if (type == 3 && (jorz(prof[5]) != 0.0 || jorz(prof[6]) != 0.0) && tth_peak >= 0.1 && tth_peak < 180.0) {
        const double safe = jmax(1e-6, jmin(kMaxTheta, tth_peak * (kPi / 180.0) * 0.5)), tn = tan(safe);
        if (fabs(tn) >= 1e-9) asym = jorz(prof[5]) / tn + jorz(prof[6]);
    }
    o.asym = asym;
    double eta, fwhm, analytic;
    if (type == 1) {
        double gG, gL;
        profile_widths64(1, prof, fmin, tth_peak, 1, gG, gL);
        const double HG = jmax(1e-9, gG), HL = jmax(1e-9, gL);
        eta = jmax(0.0, jmin(1.0, isfinite(prof[5]) ? prof[5] : 0.5));
        fwhm = tch_fwhm64(HG, HL);
        double cl, cg;
        pv_mix64(eta, HG, HL, cl, cg);
        o.ig[0] = o.ig[1] = 1.0 / HG; o.il[0] = o.il[1] = 1.0 / HL;
        o.cl[0] = o.cl[1] = cl; o.cg[0] = o.cg[1] = cg;
        analytic = cl * kLArea64 * HL + cg * kGArea64 * HG;
    } else if (type == 2) {
        double gGL, gLL, gGR, gLR;
        profile_widths64(2, prof, fmin, tth_peak, 0, gGL, gLL);
        profile_widths64(2, prof, fmin, tth_peak, 1, gGR, gLR);
        const double HGL = jmax(1e-9, gGL), HLL = jmax(1e-9, gLL), HGR = jmax(1e-9, gGR), HLR = jmax(1e-9, gLR);
        eta = jmax(0.0, jmin(1.0, isfinite(prof[8]) ? prof[8] : 0.5));
        fwhm = jmax(tch_fwhm64(HGL, HLL), tch_fwhm64(HGR, HLR));
        double clL, cgL, clR, cgR;
        pv_mix64(eta, HGL, HLL, clL, cgL);
        pv_mix64(eta, HGR, HLR, clR, cgR);
        o.ig[0] = 1.0 / HGL; o.il[0] = 1.0 / HLL; o.ig[1] = 1.0 / HGR; o.il[1] = 1.0 / HLR;
        o.cl[0] = clL; o.cg[0] = cgL; o.cl[1] = clR; o.cg[1] = cgR;
        analytic = 0.5 * (clL * kLArea64 * HLL + cgL * kGArea64 * HGL) + 0.5 * (clR * kLArea64 * HLR + cgR * kGArea64 * HGR);
    } else {
        double gG, gL;
        profile_widths64(3, prof, fmin, tth_peak, 1, gG, gL);
        const double HG = jmax(1e-9, gG), HL = jmax(1e-9, gL), fw = jmax(1e-9, tch_fwhm64(HG, HL));
        eta = 0.0;
        if (fw > 1e-9) {
            const double r = HL / fw;
            eta = jmax(0.0, jmin(1.0, 1.36603 * r - 0.47719 * (r * r) + 0.11116 * (r * r * r)));
        }
        fwhm = fw;
        double cl, cg;
        pv_mix64(eta, fw, fw, cl, cg);
        o.ig[0] = o.ig[1] = o.il[0] = o.il[1] = 1.0 / fw;
        o.cl[0] = o.cl[1] = cl; o.cg[0] = o.cg[1] = cg;
        analytic = (cl * kLArea64 + cg * kGArea64) * fw;
    }
    o.fwhm = fwhm;
    o.window = jmax(1e-6, fwhm);
    const double mult = 8.0 + (25.0 - 8.0) * jmax(0.0, jmin(1.0, eta));       // constants.js:59, 70
    const double R = jmin(10.0, mult * jmax(0.01, o.window) * (asym != 0.0 ? 2.0 : 1.0));   // constants.js:84
    o.max_d = R;
    o.ped = 0.0;
    o.ped = jmax(raw_profile64(R, o), raw_profile64(-R, o));
    o.area = analytic;
    if (type == 3 && asym != 0.0) o.area += asym_area64(o);
}

// crystal.js:465-497 sharkoLp
__device__ double lp64(double tth, double K) {
    if (!isfinite(tth) || tth <= 0.0 || tth >= 180.0) return 1.0;
    const double th = 0.5 * tth * kD2R, s = sin(th), c = cos(th);
    if (s < 1e-6 || c < 1e-6) return 1.0;
    const double k = (isfinite(K) && K >= 0.0) ? K : 1.0, c2 = cos(tth * kD2R);
    return 1.0 / (s * s * c) * ((1.0 + k * (c2 * c2)) / (1.0 + k));
}

// profile.js:108-132 calculatePeakShift (displacement and transparency), degrees
__device__ double shift64(double tth, double disp, double transp) {
    const double th = tth * (kPi / 180.0) / 2.0;
    if (fabs(th - kPi / 2.0) < 1e-6) return 0.0;
    const double t = -(disp / 1000.0) * cos(th) * (180.0 / kPi) + transp * sin(2.0 * th) * (180.0 / kPi);
    return isfinite(t) ? t : 0.0;
}

// rietveld_core.js:824-843 preferredOrientation: March-Dollase over the Laue equivalents
__device__ double march_dollase64(const DevPlan &P, const DevPhase &ph, const double *f, double r, int kref) {
    if (!(r > 0.0) || fabs(r - 1.0) < 1e-12) return 1.0;
    const double *H = ph.po_dir;
    auto g = [&](double a0, double a1, double a2, double b0, double b1, double b2) {
        return a0 * b0 * f[0] + a1 * b1 * f[1] + a2 * b2 * f[2] + (a0 * b1 + a1 * b0) * f[3] / 2 + (a1 * b2 + a2 * b1) * f[4] / 2 +
               (a0 * b2 + a2 * b0) * f[5] / 2;
    };
    const double HH = g(H[0], H[1], H[2], H[0], H[1], H[2]);
    if (!(HH > 0.0)) return 1.0;
    const double r2 = r * r, ir = 1.0 / r;
    const int e0 = P.eq_off[kref], e1 = P.eq_off[kref + 1];
    double sum = 0.0;
    for (int e = e0; e < e1; ++e) {
        const double a0 = P.eq[3 * e], a1 = P.eq[3 * e + 1], a2 = P.eq[3 * e + 2];
        const double c = g(a0, a1, a2, H[0], H[1], H[2]) / sqrt(g(a0, a1, a2, a0, a1, a2) * HH), c2 = jmin(1.0, c * c);
        sum += pow(r2 * c2 + (1.0 - c2) * ir, -1.5);
    }
    return sum / (double)(e1 - e0);
}

// rietveld_core.js:749-821 structure: |F|^2 = A^2 + B^2 + C^2 + D^2 over the atoms and
// coset representatives, and with slot >= 0 its derivative with respect to atom slot / 5's
// component slot % 5 (a free coordinate's value, occ, B).  Centrosymmetric phases keep
// one of each (R, t) / (-R, -t) pair: the cosine sums double, the sine sums vanish.
__device__ double structure64(const DevPlan &P, const JacPlan &J, const DevPhase &ph, const double *pph, int4 hk, double s2,
                              int slot, double *grad) {
    const double q4 = 0.25 * s2, c2x = ph.centro ? 2.0 : 1.0;
    const bool neutron = P.radiation == 1;
    double Ar = 0.0, Br = 0.0, Cr = 0.0, Dr = 0.0;
    // the slot atom's own sums (per unit occupancy) and coordinate derivatives
    double As = 0.0, Bs = 0.0, Cs = 0.0, Ds = 0.0, dA = 0.0, dB = 0.0, dC = 0.0, dD = 0.0;
    const int sa = slot >= 0 ? slot / 5 : -1, sc = slot >= 0 ? slot % 5 : -1;
    for (int j = 0; j < ph.n_atoms; ++j) {
        const JacAtom &a = J.atoms[ph.atom_off + j];
        const double *ap = pph + PTK_PH_ATOM0 + 5 * j;
        double r[3] = {a.r0[0], a.r0[1], a.r0[2]};             // sitePosition, rietveld_core.js:372-378
        for (int f = 0; f < 3; ++f)
            if (a.free_mask & (1 << f)) for (int i = 0; i < 3; ++i) r[i] += ap[f] * a.basis[f][i];
        double f0;
        if (neutron) f0 = isfinite(a.b_neutron) ? a.b_neutron : 0.0;
        else {
            f0 = a.ff_c;
            for (int gI = 0; gI < a.n_gauss; ++gI) f0 += a.ff_a[gI] * exp(-a.ff_b[gI] * q4);
        }
        const double dw = exp(-ap[4] * q4);
        const double w1 = (neutron ? f0 : f0 + a.fp) * dw * a.w0;              // per unit occupancy
        const double w1i = (neutron || !(a.fpp != 0.0)) ? 0.0 : a.fpp * dw * a.w0;
        const bool mine = j == sa;
        double bh = 0.0, b3[3] = {0.0, 0.0, 0.0};
        if (mine && sc < 3) for (int i = 0; i < 3; ++i) b3[i] = a.basis[sc][i];
        double c = 0.0, sn = 0.0, dc = 0.0, ds = 0.0;
        for (int q = 0; q < ph.n_reps; ++q) {
            const int *R = P.rep_r + 9 * (ph.rep_off + q);
            const double *t = P.rep_t + 3 * (ph.rep_off + q);
            const double H0 = hk.x * R[0] + hk.y * R[3] + hk.z * R[6], H1 = hk.x * R[1] + hk.y * R[4] + hk.z * R[7],
                         H2 = hk.x * R[2] + hk.y * R[5] + hk.z * R[8];
            const double phs = H0 * r[0] + H1 * r[1] + H2 * r[2] + (double)hk.x * t[0] + (double)hk.y * t[1] + (double)hk.z * t[2];
            double sv, cv;
            sincospi(2.0 * (phs - floor(phs)), &sv, &cv);
            c += cv; sn += sv;
            if (mine && sc < 3) {
                bh = H0 * b3[0] + H1 * b3[1] + H2 * b3[2];
                dc -= bh * sv; ds += bh * cv;
            }
        }
        c *= c2x; dc *= c2x;
        if (ph.centro) { sn = 0.0; ds = 0.0; }
        const double occ = ap[3];
        Ar += occ * w1 * c; Br += occ * w1 * sn; Cr += occ * w1i * c; Dr += occ * w1i * sn;
        if (mine) {
            As = w1 * c; Bs = w1 * sn; Cs = w1i * c; Ds = w1i * sn;
            const double tp = 2.0 * kPi * occ;
            dA = tp * w1 * dc; dB = tp * w1 * ds; dC = tp * w1i * dc; dD = tp * w1i * ds;
            if (sc == 4) { As *= occ; Bs *= occ; Cs *= occ; Ds *= occ; }
        }
    }
    if (grad) {
        if (sc < 3) *grad = 2.0 * (Ar * dA + Br * dB + Cr * dC + Dr * dD);
        else if (sc == 3) *grad = 2.0 * (Ar * As + Br * Bs + Cr * Cs + Dr * Ds);
        else *grad = 2.0 * (Ar * As + Br * Bs + Cr * Cs + Dr * Ds) * (-s2 / 4.0);
    }
    return Ar * Ar + Br * Br + Cr * Cr + Dr * Dr;
}

__device__ void write_line(const DevPlan &P, const Prep64 &pk, double x0, double hgt, JacLine &L) {
    L.start = L.stop = 0;
    if (!isfinite(x0) || !isfinite(pk.max_d)) return;
    const int start = lower_bound_t(P, x0 - pk.max_d);
    if (start >= P.np) return;
    const int stop = upper_stop(P, start, x0 + pk.max_d);
    if (stop <= start) return;
    L.x0 = x0; L.asym = pk.asym; L.ped = pk.ped; L.hgt = hgt;
    for (int s = 0; s < 2; ++s) { L.ig[s] = pk.ig[s]; L.il[s] = pk.il[s]; L.cl[s] = pk.cl[s]; L.cg[s] = pk.cg[s]; }
    L.start = start; L.stop = stop;
}

// the value of a line at a point inside its window (profile.js:733-759)
__device__ __forceinline__ double line64(const JacLine &L, double t) {
    double delta = t - L.x0;
    if (L.asym != 0.0 && fabs(delta) >= 1e-9) delta /= (1.0 - jmax(-0.95, jmin(0.95, L.asym * delta)));
    const int s = delta < 0.0 ? 0 : 1;
    const double u = delta * L.ig[s], v = delta * L.il[s];
    const double val = L.cl[s] / (1.0 + 4.0 * v * v) + L.cg[s] * exp(-kCg64 * u * u) - L.ped;
    return val > 0.0 ? val : 0.0;
}

// rietveld_core.js:549-561 splineSlopes (Fritsch-Carlson tangents, tau <= 9 per interval)
__device__ void spline_slopes64(const double *xs, const double *ys, int na, double *m) {
    if (na < 2) { if (na == 1) m[0] = 0.0; return; }
    double dprev = 0.0, hprev = 0.0;
    for (int j = 0; j + 1 < na; ++j) {
        const double h = xs[j + 1] - xs[j], d = (ys[j + 1] - ys[j]) / h;
        if (j == 0) m[0] = d;
        else m[j] = dprev * d <= 0.0 ? 0.0 : (h * dprev + hprev * d) / (hprev + h);
        dprev = d; hprev = h;
    }
    m[na - 1] = dprev;
    for (int j = 0; j + 1 < na; ++j) {
        const double d = (ys[j + 1] - ys[j]) / (xs[j + 1] - xs[j]);
        if (d == 0.0) { m[j] = 0.0; m[j + 1] = 0.0; continue; }
        const double a = m[j] / d, b = m[j + 1] / d, tau = a * a + b * b;
        if (tau > 9.0) { const double f = 3.0 / sqrt(tau); m[j] = a * d * f; m[j + 1] = b * d * f; }
    }
}
// rietveld_core.js:597-612 splineValues at one point
__device__ double spline_at64(const double *xs, const double *ys, const double *m, int na, double t) {
    if (na < 1) return 0.0;
    if (na == 1 || t <= xs[0]) return ys[0];
    if (t >= xs[na - 1]) return ys[na - 1];
    int s = 0;
    while (s < na - 2 && xs[s + 1] <= t) ++s;
    const double hi = xs[s + 1] - xs[s], u = hi > 1e-12 ? (t - xs[s]) / hi : 0.0, u2 = u * u, u3 = u2 * u;
    return (2 * u3 - 3 * u2 + 1) * ys[s] + (u3 - 2 * u2 + u) * hi * m[s] + (-2 * u3 + 3 * u2) * ys[s + 1] + (u3 - u2) * hi * m[s + 1];
}

__device__ __forceinline__ double cheb_z(const DevPlan &P, int i) {
    return P.thi > P.tlo ? 2.0 * (P.tth[i] - P.tlo) / (P.thi - P.tlo) - 1.0 : 0.0;
}

}  // namespace

// one thread per (evaluation, reflection): the reflection's Kalpha1 / Kalpha2 lines at the
// evaluation's parameters (Model.geometry rietveld_core.js:719-736, buildPeakContributions
// profile.js:1013-1080, evalPhase heights :855-897)
__global__ void __launch_bounds__(128) k_jac_lines(DevPlan P, JacPlan J) {
    const size_t gid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= (size_t)J.n_evals * P.nr) return;
    const int e = (int)(gid / P.nr), k = (int)(gid % P.nr);
    JacLine *L = J.lines + ((size_t)e * P.nr + k) * 2;
    L[0].start = L[0].stop = 0;
    L[1].start = L[1].stop = 0;
    const JacEval E = J.evals[e];
    if (E.mode >= JAC_BKG) return;
    int q = 0;
    while (q + 1 < P.n_phases && k >= P.phase[q + 1].refl_off) ++q;
    if (!((E.mask >> q) & 1)) return;
    const DevPhase &ph = P.phase[q];
    const double *p = J.pvecs + (size_t)E.pvec * J.n_params, *pph = p + ph.param_base;
    double f[6];
    if (!phase_metric(ph, pph, f)) return;
    const int4 hk = P.hklm[k];
    const double h = hk.x, kk = hk.y, l = hk.z;
    const double qv = h * h * f[0] + kk * kk * f[1] + l * l * f[2] + h * kk * f[3] + kk * l * f[4] + h * l * f[5];
    if (!(qv > 0.0) || qv > P.qlim) return;
    const double s = sqrt(qv), sinT = P.lambda1 * s * 0.5;
    if (!(sinT < 1.0)) return;
    const double bragg = 2.0 * asin(sinT) * kR2D, disp = p[1], transp = p[2], zero = p[0];
    const double tth = bragg + shift64(bragg, disp, transp);
    if (!(tth > 0.0 && tth < 180.0)) return;                 // hasRenderablePosition
    const double base1 = tth + zero;
    const double *prof = pph + 8;
    Prep64 pk1;
    prepare64(ph.profile_type, prof, P.min_fwhm, base1, pk1);
    if (!(pk1.area > 0.0)) return;
    const double lp1 = lp64(base1, P.pol_k);
    const double r = pph[7], po = march_dollase64(P, ph, f, r, k);
    double gradv = 0.0;
    const double f2 = structure64(P, J, ph, pph, hk, qv, E.mode == JAC_GRAD ? E.slot : -1, E.mode == JAC_GRAD ? &gradv : nullptr);
    double S = pph[6];
    if (!(S != 0.0) || isnan(S)) S = 0.0;
    const double factor = hk.w * lp1 * po / pk1.area;       // height per unit S |F|^2
    double hgt;
    if (E.mode == JAC_UNIT) hgt = factor * f2;
    else if (E.mode == JAC_GRAD) hgt = S * factor * gradv;
    else if (E.mode == JAC_PO) {
        const double po2 = march_dollase64(P, ph, f, r + E.h, k);
        hgt = po > 0.0 ? S * (factor * f2) * (po2 / po - 1.0) / E.h : 0.0;
    } else hgt = S * (factor * f2);
    write_line(P, pk1, base1, hgt, L[0]);
    if (P.doublet) {
        const double st2 = sin(bragg * (kPi / 180.0) / 2.0) * (P.lambda2 / P.lambda1);
        if (fabs(st2) < 1.0) {
            const double b2 = 2.0 * asin(st2) / (kPi / 180.0);
            const double base2 = b2 + shift64(b2, disp, transp) + zero;
            Prep64 pk2;
            prepare64(ph.profile_type, prof, P.min_fwhm, base2, pk2);
            const double lpr = isfinite(P.pol_k) ? lp64(base2, P.pol_k) / lp64(base1, P.pol_k) : 1.0;
            const double w2 = P.ratio * pk1.area / pk2.area * lpr;
            if (isfinite(pk2.max_d)) write_line(P, pk2, base2, hgt * w2, L[1]);
        }
    }
}

// one workgroup per (256-point tile, evaluation): the evaluation's pattern
__global__ void __launch_bounds__(256) k_jac_patterns(DevPlan P, JacPlan J) {
    __shared__ int2 win[4096];
    __shared__ double bx[PTK_MAX_BKG], by[2][PTK_MAX_BKG], bm[2][PTK_MAX_BKG];
    const int e = blockIdx.y, i = blockIdx.x * 256 + threadIdx.x;
    const JacEval E = J.evals[e];
    const double *p = J.pvecs + (size_t)E.pvec * J.n_params;
    double v = 0.0;
    if (E.mode == JAC_CHEB) {
        if (i < P.np) {
            const double z = cheb_z(P, i);
            double a = 1.0, b = z;
            v = E.slot == 0 ? 1.0 : z;
            for (int kx = 2; kx <= E.slot; ++kx) { const double nx = 2.0 * z * b - a; a = b; b = nx; v = nx; }
        }
    } else if (E.mode == JAC_BKG || E.mode == JAC_SPLINE) {
        const int na = P.n_bkg;
        if (P.bkg_spline) {
            if (threadIdx.x == 0) {
                for (int jx = 0; jx < na; ++jx) {
                    bx[jx] = P.bknots[jx];
                    by[0][jx] = p[3 + jx];
                    by[1][jx] = p[3 + jx] + (E.mode == JAC_SPLINE && jx == E.slot ? E.h : 0.0);
                }
                spline_slopes64(bx, by[0], na, bm[0]);
                if (E.mode == JAC_SPLINE) spline_slopes64(bx, by[1], na, bm[1]);
            }
            __syncthreads();
            if (i < P.np) {
                const double t = P.tth[i], y0 = spline_at64(bx, by[0], bm[0], na, t);
                v = E.mode == JAC_SPLINE ? (spline_at64(bx, by[1], bm[1], na, t) - y0) / E.h : y0;
            }
        } else if (E.mode == JAC_BKG && i < P.np) {
            // rietveld_core.js:920-931: sum_k bkg_k T_k(z)
            const double z = cheb_z(P, i);
            double a = 1.0, b = z;
            if (na > 0) v = p[3];
            if (na > 1) v += p[4] * z;
            for (int kx = 2; kx < na; ++kx) { const double nx = 2.0 * z * b - a; a = b; b = nx; v += p[3 + kx] * nx; }
        }
    } else {
        // the lines of the evaluation, in envelope order (Kalpha1 then Kalpha2 of each
        // reflection), windows staged in LDS in chunks
        const JacLine *L = J.lines + (size_t)e * 2 * P.nr;
        const int nl = 2 * P.nr, t_lo = blockIdx.x * 256, t_hi = min(P.np, t_lo + 256);
        const double t = i < P.np ? P.tth[i] : 0.0;
        for (int c0 = 0; c0 < nl; c0 += 4096) {
            const int cn = min(4096, nl - c0);
            __syncthreads();
            for (int x = threadIdx.x; x < cn; x += 256) win[x] = make_int2(L[c0 + x].start, L[c0 + x].stop);
            __syncthreads();
            for (int x = 0; x < cn; ++x) {
                const int2 w = win[x];
                if (w.y <= w.x || w.x >= t_hi || w.y <= t_lo) continue;     // uniform over the workgroup
                if (i >= w.x && i < w.y) v += L[c0 + x].hgt * line64(L[c0 + x], t);
            }
        }
    }
    if (i < P.np) J.ye[(size_t)e * P.np + i] = v;
}

// the variables' columns J = C Y (rows 0..nv-1) and the residual r = y_obs - y (row nv,
// zero where w = 0); evaluations 0 and 1 are the base phases and the base background
__global__ void __launch_bounds__(256) k_jac_combine(DevPlan P, JacPlan J) {
    const int row = blockIdx.y, i = blockIdx.x * 256 + threadIdx.x;
    if (i >= P.np) return;
    double v = 0.0;
    if (row < J.nv) {
        const double *c = J.C + (size_t)row * J.n_evals;
        for (int e = 0; e < J.n_evals; ++e)
            if (c[e] != 0.0) v += c[e] * J.ye[(size_t)e * P.np + i];
    } else if (J.w[i] > 0.0) {
        v = J.yobs[i] - (J.ye[i] + J.ye[(size_t)P.np + i]);
    }
    J.jr[(size_t)row * P.np + i] = v;
}

// entry (a, b), a <= b, of [J; r] W [J; r]': one workgroup, an ordered reduction
__global__ void __launch_bounds__(256) k_jac_ext(DevPlan P, JacPlan J) {
    __shared__ double red[256];
    const int n = J.nv + 1;
    int e = blockIdx.x, a = 0;
    while (e >= n - a) { e -= n - a; ++a; }
    const int b = a + e;
    const double *ja = J.jr + (size_t)a * P.np, *jb = J.jr + (size_t)b * P.np;
    double s = 0.0;
    for (int i = threadIdx.x; i < P.np; i += 256) s += J.w[i] * ja[i] * jb[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) J.ext[blockIdx.x] = red[0];
}

hipError_t launch_jac_lines(const DevPlan &P, const JacPlan &J, hipStream_t s) {
    const size_t n = (size_t)J.n_evals * P.nr;
    hipLaunchKernelGGL(k_jac_lines, dim3((unsigned)((n + 127) / 128)), dim3(128), 0, s, P, J);
    return hipGetLastError();
}
hipError_t launch_jac_patterns(const DevPlan &P, const JacPlan &J, hipStream_t s) {
    hipLaunchKernelGGL(k_jac_patterns, dim3((P.np + 255) / 256, J.n_evals), dim3(256), 0, s, P, J);
    return hipGetLastError();
}
hipError_t launch_jac_normal(const DevPlan &P, const JacPlan &J, hipStream_t s) {
    hipLaunchKernelGGL(k_jac_combine, dim3((P.np + 255) / 256, J.nv + 1), dim3(256), 0, s, P, J);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const int n = J.nv + 1;
    hipLaunchKernelGGL(k_jac_ext, dim3(n * (n + 1) / 2), dim3(256), 0, s, P, J);
    return hipGetLastError();
}

}  // namespace ptk
