// pt_host.cpp -- C ABI of the MI355X PT trial evaluator (include/powder6_pt.h).
//
// Host side of js/pt_gpu.js: compile() (reflection envelope, packing), create()
// (device buffers, stream), costs()/batch()/run() plumbing.  All arithmetic of a trial
// runs in the kernels of pt_kernels.hip; this file only prepares and launches.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/powder6_pt.h"
#include "pt_common.h"
#include "pt_jacobian.h"
#include "pt_launch.h"
#include "pt_mlp_host.h"

namespace {

thread_local std::string g_err;
int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}

const double kPi = 3.141592653589793;
const double D2R = kPi / 180, R2D = 180 / kPi;

}  // namespace

int ptk_host::fail(int code, const std::string &msg) { return ::fail(code, msg); }

namespace {

// ---------------------------------------------------------------------------------
// reflection enumeration: the envelope of js/pt_gpu.js:110-131 through
// generateReflections (js/rietveld_core.js:182-276)
// ---------------------------------------------------------------------------------
struct Refl {
    int h, k, l, m;
    double d0;
    std::vector<short> eq;   // 3 per equivalent
};

bool inv_dsq_form(const double cell[6], double f[6]) {   // crystal.js:326-366
    double a = cell[0], b = cell[1] > 0 ? cell[1] : a, c = cell[2] > 0 ? cell[2] : a;
    if (!(a > 0) || !(b > 0) || !(c > 0)) return false;
    double al = cell[3] * D2R, be = cell[4] * D2R, ga = cell[5] * D2R;
    double ca = std::cos(al), cb = std::cos(be), cg = std::cos(ga);
    double sa = std::sin(al), sb = std::sin(be), sg = std::sin(ga);
    double vf = 1 - ca * ca - cb * cb - cg * cg + 2 * ca * cb * cg;
    if (vf <= 1e-9 || std::fabs(sa) < 1e-9 || std::fabs(sb) < 1e-9 || std::fabs(sg) < 1e-9) return false;
    double V = a * b * c * std::sqrt(vf);
    double as = b * c * sa / V, bs = a * c * sb / V, cs = a * b * sg / V;
    double cA = (cb * cg - ca) / (sb * sg), cB = (ca * cg - cb) / (sa * sg), cG = (ca * cb - cg) / (sa * sb);
    f[0] = as * as; f[1] = bs * bs; f[2] = cs * cs;
    f[3] = 2 * as * bs * cG; f[4] = 2 * bs * cs * cA; f[5] = 2 * as * cs * cB;
    return true;
}

inline double rnd(double v) { return std::floor(v + 0.5); }

std::vector<Refl> enumerate_reflections(const double cell[6], int n_ops, const int32_t *R, const double *T, double dmin,
                                        const int *box) {
    std::vector<Refl> out;
    double f[6];
    if (!inv_dsq_form(cell, f) || !(dmin > 0)) return out;
    double lim = 1 / (dmin * dmin) * (1 + 1e-9);
    // Laue group: +R and -R of every operator, first occurrence kept (row-vector action h R)
    std::vector<std::array<int, 9>> rots;
    for (int i = 0; i < n_ops; ++i)
        for (int s = 0; s < 2; ++s) {
            std::array<int, 9> r;
            for (int x = 0; x < 9; ++x) r[x] = s ? -R[9 * i + x] : R[9 * i + x];
            if (std::find(rots.begin(), rots.end(), r) == rots.end()) rots.push_back(r);
        }
    double cb = cell[1] > 0 ? cell[1] : cell[0], cc = cell[2] > 0 ? cell[2] : cell[0];
    int hmax = box ? box[0] : (int)std::floor(cell[0] / dmin);
    int kmax = box ? box[1] : (int)std::floor(cb / dmin);
    int lmax = box ? box[2] : (int)std::floor(cc / dmin);
    std::vector<std::array<int, 3>> orbit;
    for (int h = -hmax; h <= hmax; ++h)
        for (int k = -kmax; k <= kmax; ++k)
            for (int l = -lmax; l <= lmax; ++l) {
                if (!h && !k && !l) continue;
                double q = h * h * f[0] + k * k * f[1] + l * l * f[2] + h * k * f[3] + k * l * f[4] + h * l * f[5];
                if (!(q > 0) || (!box && q > lim)) continue;
                // canonical member: most non-negative indices, then the largest (h, k, l)
                int bh = -1000000000, bk = 0, bl = 0, bp = -1;
                orbit.clear();
                for (const auto &r : rots) {
                    int v0 = h * r[0] + k * r[3] + l * r[6], v1 = h * r[1] + k * r[4] + l * r[7], v2 = h * r[2] + k * r[5] + l * r[8];
                    int np = (v0 >= 0) + (v1 >= 0) + (v2 >= 0);
                    std::array<int, 3> v{v0, v1, v2};
                    if (std::find(orbit.begin(), orbit.end(), v) == orbit.end()) orbit.push_back(v);
                    if (np > bp || (np == bp && (v0 > bh || (v0 == bh && (v1 > bk || (v1 == bk && v2 > bl)))))) {
                        bp = np; bh = v0; bk = v1; bl = v2;
                    }
                }
                if (bh != h || bk != k || bl != l) continue;
                // systematic absence: some (R, t) with h R = h and h.t not an integer
                bool allowed = true;
                for (int i = 0; i < n_ops && allowed; ++i) {
                    const int32_t *r = R + 9 * i;
                    if (h * r[0] + k * r[3] + l * r[6] != h || h * r[1] + k * r[4] + l * r[7] != k || h * r[2] + k * r[5] + l * r[8] != l) continue;
                    double p = h * T[3 * i] + k * T[3 * i + 1] + l * T[3 * i + 2];
                    if (std::fabs(p - rnd(p)) > 1e-6) allowed = false;
                }
                if (!allowed) continue;
                Refl rf;
                rf.h = h; rf.k = k; rf.l = l; rf.m = (int)orbit.size(); rf.d0 = 1 / std::sqrt(q);
                for (auto &v : orbit) { rf.eq.push_back((short)v[0]); rf.eq.push_back((short)v[1]); rf.eq.push_back((short)v[2]); }
                out.push_back(std::move(rf));
            }
    std::sort(out.begin(), out.end(), [](const Refl &a, const Refl &b) {
        double d = b.d0 - a.d0;
        if (d != 0 && !std::isnan(d)) return d < 0;
        if (a.h != b.h) return a.h > b.h;
        if (a.k != b.k) return a.k > b.k;
        return a.l > b.l;
    });
    return out;
}

// coset representatives modulo the centring translations (rietveld_core.js:164-179)
// centring translations of the operator list (identity rotation)
std::vector<std::array<double, 3>> centrings(int n_ops, const int32_t *R, const double *T) {
    std::vector<std::array<double, 3>> C;
    for (int i = 0; i < n_ops; ++i) {
        const int32_t *r = R + 9 * i;
        if (r[0] == 1 && r[4] == 1 && r[8] == 1 && !r[1] && !r[2] && !r[3] && !r[5] && !r[6] && !r[7])
            C.push_back({T[3 * i], T[3 * i + 1], T[3 * i + 2]});
    }
    if (C.empty()) C.push_back({0, 0, 0});
    return C;
}

// One of each (R, t) / (-R, -t) pair of representatives, when every representative has
// its partner (-R, -t + centring) in the list: the inversion lies at the origin, so
// cos h.(R r + t) is the same for both members and the sines cancel -- the structure
// factor sums cosines over half the list and doubles them.  Empty when some
// representative has no partner.
std::vector<int> centro_half(const std::vector<int> &reps, int n_ops, const int32_t *R, const double *T) {
    const auto C = centrings(n_ops, R, T);
    std::vector<int> partner(reps.size(), -1), half;
    for (size_t a = 0; a < reps.size(); ++a) {
        if (partner[a] >= 0) continue;
        const int i = reps[a];
        for (size_t b = 0; b < reps.size() && partner[a] < 0; ++b) {
            if (b == a || partner[b] >= 0) continue;
            const int j = reps[b];
            bool neg = true;
            for (int x = 0; x < 9; ++x) neg = neg && R[9 * j + x] == -R[9 * i + x];
            if (!neg) continue;
            for (auto &c : C) {
                bool ok = true;
                for (int x = 0; x < 3; ++x) {
                    const double d = T[3 * i + x] + T[3 * j + x] - c[x];
                    if (std::fabs(d - rnd(d)) > 1e-6) { ok = false; break; }
                }
                if (ok) { partner[a] = (int)b; partner[b] = (int)a; break; }
            }
        }
        if (partner[a] < 0) return {};
        half.push_back(i);
    }
    return half;
}

std::vector<int> coset_reps(int n_ops, const int32_t *R, const double *T) {
    std::vector<std::array<double, 3>> C;
    for (int i = 0; i < n_ops; ++i) {
        const int32_t *r = R + 9 * i;
        if (r[0] == 1 && r[4] == 1 && r[8] == 1 && !r[1] && !r[2] && !r[3] && !r[5] && !r[6] && !r[7])
            C.push_back({T[3 * i], T[3 * i + 1], T[3 * i + 2]});
    }
    if (C.empty()) C.push_back({0, 0, 0});
    std::vector<int> reps;
    for (int i = 0; i < n_ops; ++i) {
        bool dup = false;
        for (int j : reps) {
            if (std::memcmp(R + 9 * i, R + 9 * j, 9 * sizeof(int32_t))) continue;
            for (auto &c : C) {
                bool same = true;
                for (int x = 0; x < 3; ++x) {
                    double d = T[3 * i + x] - T[3 * j + x] - c[x];
                    if (std::fabs(d - rnd(d)) > 1e-6) { same = false; break; }
                }
                if (same) { dup = true; break; }
            }
            if (dup) break;
        }
        if (!dup) reps.push_back(i);
    }
    return reps;
}

}  // namespace

// ---------------------------------------------------------------------------------
// the plan
// ---------------------------------------------------------------------------------
struct pt_plan {
    int nv = 0, chains = 0, stride = 0, np = 0, nr = 0, n_params = 0, steps_per_sweep = 0, uniform = 0;
    int device = -1;
    int block = 512, ch = 256;
    size_t dyn = 0, plain_dyn = 0;
    std::vector<std::vector<int32_t>> phase_hklm;   // per phase (h,k,l,m)
    ptk::DevPlan P{};
    hipStream_t stream = nullptr;
    // a sweep's chains run as `parts` contiguous parts, part p > 0 on side[p - 1], so that
    // one part's latency-bound stages overlap another part's accumulation
    static constexpr int kMaxParts = 4;
    int parts = 1;
    hipStream_t side[kMaxParts - 1] = {};
    hipEvent_t fork = nullptr, join[kMaxParts - 1] = {};
    std::vector<void *> allocs;
    // device scratch
    double *d_x = nullptr, *d_chi2 = nullptr, *d_partial = nullptr, *d_y = nullptr;
    size_t cap_trials = 0;
    void *d_plain = nullptr;
    ptk::LineRec *d_rec = nullptr;
    int4 *d_span = nullptr;
    bool states_set = false;
    // staged pipeline buffers (per trial slot)
    int slots = 0, ngroups = 1;           // 512-point groups per trial (one partial chi^2 each)
    int tpw = 2;                          // accumulate tile shape: 128-point tiles per wavefront
    bool tuned = false;
    double tune_ms[3] = {0, 0, 0};        // accumulate time per step at 1, 2, 4 tiles per wavefront
    void *d_img = nullptr;
    ptk::StagedLine *d_lrec = nullptr;   // staged line records, 2 nr per slot
    ptk::LineRec *d_mrec = nullptr;      // unit-height lines of the shared profile matrix, 2 nr
    int4 *d_mspan = nullptr;
    double2 *d_part = nullptr;
    double *d_schi2 = nullptr;
    // one PT sweep captured as a graph
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    bool kernel_timing = false;
    std::vector<hipEvent_t> ev;
    // shared-profile matrix form
    bool matrix_ready = false;
    int matrix_bf16 = 0;                 // element type P is filled in (0 FP16, 1 BF16)
    int kpad = 0, np_pad = 0, row_tiles = 0, cap_mtrials = 0;
    unsigned short *d_phi = nullptr, *d_plo = nullptr, *d_ihi = nullptr, *d_ilo = nullptr;
    float *d_factor = nullptr, *d_iscale = nullptr, *d_mpart = nullptr;
    double4 *d_mhr = nullptr;            // per representative and reflection: h R, h.t
    float *d_mf0w = nullptr;             // per atom and reflection: (f0 + f') w0
    float2 *d_mqk = nullptr;             // per reflection: s^2 / 4, height factor
    int *d_mgo = nullptr;                // the start vector's cell validity
    // the matrix form's products (its intensities run on `stream`): the first side stream
    // when the plan has one (streams beyond the hardware queues would share one with
    // `stream` and run after it), else a stream of its own
    hipStream_t mstream = nullptr;
    bool mstream_own = false;
    hipEvent_t mjoin = nullptr;
    int2 *d_band = nullptr;
    float2 *d_yw = nullptr;
    double *d_myout = nullptr;
    std::vector<char> struct_var;        // variables that may differ from x0 in the matrix form
    std::vector<double> x0, base_params; // start vector and its model parameters
    std::vector<double> vlo_h, vhi_h, vscale_h;
    double log_wss = 0;                  // log sum w y_obs^2 (the Rwp denominator)
    // surrogate-screened PT
    pt_mlp *mlp = nullptr;
    bool screen_on = false;
    bool screen_fused = false;           // the skip test runs inside the PT setup (one proposal per chain)
    pt_screen_config scfg{};
    ptk::ScreenDev S{};
    ptk::ScreenFused *d_sf = nullptr;
    int screen_cap = 0;
    std::vector<double> tth_host;
    int n_bkg = 0;
    int bkg_type = PT_BKG_CHEBYSHEV;
    std::vector<double> bkg_x;         // spline anchors
    // multi-dataset plans: the further datasets (owned), each with its own pipeline buffers;
    // the first dataset's finish adds their chi^2 through d_follow
    std::vector<pt_plan *> follow;
    ptk::DevFollow *d_follow = nullptr;
    // the plan's separated-profile / free-intensity description (device copy and host image)
    ptk::DevSepPlan sp_host{};
    ptk::DevSepPlan *d_sp = nullptr;
    hipEvent_t ev_total[2] = {nullptr, nullptr};
    double last_total_ms = 0, last_kernel_ms = 0, last_matrix_ms[2] = {0, 0};
    int last_launches = 0;
    double qlim = 0;
    // the LM normal equations (FP64): what the column specification needs on the host, and
    // the device buffers (allocated at the first pt_normal_equations call)
    std::vector<ptk::JacAtom> jac_atoms_h;
    std::vector<double> yobs_h, w_h, param_base_h, term_coeff_h;
    std::vector<int> term_off_h, term_var_h;
    ptk::JacAtom *d_jatoms = nullptr;
    double *d_jy = nullptr, *d_jw = nullptr, *d_jp = nullptr, *d_jye = nullptr, *d_jC = nullptr, *d_jr = nullptr,
           *d_jext = nullptr;
    ptk::JacEval *d_jev = nullptr;
    ptk::JacLine *d_jlines = nullptr;
    int jac_cap_evals = 0;
    double last_normal_ms = 0;
};

namespace {

template <class Owner, class T>
int dalloc(Owner *pl, T **ptr, size_t count) {
    void *p = nullptr;
    if (hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T)) != hipSuccess) return fail(PT_ERR_DEVICE, "PT buffers exceed this GPU's memory (hipMalloc failed).");
    pl->allocs.push_back(p);
    *ptr = reinterpret_cast<T *>(p);
    return 0;
}
// frees a buffer dalloc gave out (a null pointer is a no-op)
template <class Owner, class T>
void dfree(Owner *pl, T **ptr) {
    void *p = reinterpret_cast<void *>(*ptr);
    if (!p) return;
    for (size_t i = 0; i < pl->allocs.size(); ++i)
        if (pl->allocs[i] == p) { (void)hipFree(p); pl->allocs.erase(pl->allocs.begin() + i); break; }
    *ptr = nullptr;
}
template <class Owner, class T>
int upload(Owner *pl, T **ptr, const std::vector<T> &v) {
    int rc = dalloc(pl, ptr, v.size());
    if (rc) return rc;
    if (!v.empty() && hipMemcpy(*ptr, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess)
        return fail(PT_ERR_DEVICE, "PT device upload failed.");
    return 0;
}

int device_check(int *dev) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n < 1)
        return fail(PT_ERR_DEVICE, "MI355X HIP device is unavailable in this process or context.");
    int d = 0;
    (void)hipGetDevice(&d);
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, d) != hipSuccess) return fail(PT_ERR_DEVICE, "MI355X HIP device is unavailable (device query failed).");
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(PT_ERR_DEVICE, std::string("MI355X HIP device is unavailable: found ") + prop.gcnArchName + ", need gfx950.");
    *dev = d;
    return 0;
}

#define HIPCHECK(expr, what)                                                                    \
    do {                                                                                        \
        hipError_t e_ = (expr);                                                                 \
        if (e_ != hipSuccess) return fail(PT_ERR_DEVICE, std::string("PT device lost during ") + what + ": " + hipGetErrorString(e_)); \
    } while (0)

int phase_base_of(const pt_problem *pb, int q) {
    int b = PT_P_PHASE0;
    for (int i = 0; i < q; ++i) b += PT_PH_ATOM0 + PT_AT_STRIDE * pb->phases[i].n_atoms;
    return b;
}

// range of a model parameter over the variable boxes: base + sum c [lo, hi]
void param_range(const pt_problem *pb, int k, double &lo, double &hi) {
    lo = hi = pb->param_base[k];
    for (int t = 0; t < pb->n_terms; ++t) {
        if (pb->term_param[t] != k) continue;
        double c = pb->term_coeff[t];
        int j = pb->term_var[t];
        lo += c * (c > 0 ? pb->lo[j] : pb->hi[j]);
        hi += c * (c > 0 ? pb->hi[j] : pb->lo[j]);
    }
}
double param_value(const pt_problem *pb, int k) {
    double v = pb->param_base[k];
    for (int t = 0; t < pb->n_terms; ++t)
        if (pb->term_param[t] == k) v += pb->term_coeff[t] * pb->x0[pb->term_var[t]];
    return v;
}

// cellOf (rietveld_core.js:207-221 with crystal.js:76-119) from per-slot values
void cell_from(const pt_phase &ph, const double v[6], double cell[6]) {
    double a = v[0], b = v[1] > 0 ? v[1] : a, c = v[2] > 0 ? v[2] : a;
    cell[0] = a; cell[1] = b; cell[2] = c; cell[3] = cell[4] = cell[5] = 90;
    switch (ph.system) {
    case PT_CUBIC: cell[1] = cell[2] = a; break;
    case PT_TETRAGONAL: cell[1] = a; break;
    case PT_HEXAGONAL: case PT_TRIGONAL: cell[1] = a; cell[5] = 120; break;
    case PT_ORTHORHOMBIC: break;
    case PT_MONOCLINIC: {
        int s = ph.unique_axis == 0 ? 3 : ph.unique_axis == 2 ? 5 : 4;
        cell[s] = std::isfinite(v[s]) ? v[s] : 90;
        break;
    }
    default:
        for (int s = 3; s < 6; ++s) cell[s] = std::isfinite(v[s]) ? v[s] : 90;
    }
}
// free cell slots of a system (rietveld_core.js:207-215 cellParams)
std::vector<int> cell_slots(const pt_phase &ph) {
    switch (ph.system) {
    case PT_CUBIC: return {0};
    case PT_TETRAGONAL: case PT_HEXAGONAL: case PT_TRIGONAL: return {0, 2};
    case PT_ORTHORHOMBIC: return {0, 1, 2};
    case PT_MONOCLINIC: return {0, 1, 2, ph.unique_axis == 0 ? 3 : ph.unique_axis == 2 ? 5 : 4};
    default: return {0, 1, 2, 3, 4, 5};
    }
}

int validate(const pt_problem *pb) {
    if (!pb) return fail(PT_ERR_ARG, "PT problem is missing.");
    if (pb->n_points < 2 || !pb->tth || !pb->y_obs || !pb->weight) return fail(PT_ERR_ARG, "a dataset has fewer than two points in its fitted range");
    for (int i = 0; i < pb->n_points; ++i) {
        if (!std::isfinite(pb->tth[i]) || (i && !(pb->tth[i] > pb->tth[i - 1]))) return fail(PT_ERR_ARG, "PT 2-theta axis must be finite and strictly ascending.");
        if (!std::isfinite(pb->y_obs[i]) || !std::isfinite(pb->weight[i])) return fail(PT_ERR_ARG, "PT observations must be finite.");
    }
    if (pb->n_phases < 1 || pb->n_phases > PTK_MAX_PHASES) return fail(PT_ERR_ARG, "PT supports 1 to 4 phases per dataset.");
    if (pb->n_bkg < 0 || pb->n_bkg > PT_MAX_BKG) return fail(PT_ERR_ARG, "Choose 0 to 24 Chebyshev terms.");
    if (pb->bkg_type != PT_BKG_CHEBYSHEV && pb->bkg_type != PT_BKG_SPLINE) return fail(PT_ERR_ARG, "unknown background form");
    if (pb->resolution_type < 0 || pb->resolution_type > 4) return fail(PT_ERR_ARG, "unknown instrument profile type");
    if (pb->bkg_type == PT_BKG_SPLINE) {
        if (pb->n_bkg > 0 && !pb->bkg_x) return fail(PT_ERR_ARG, "spline background anchors are missing");
        for (int j = 0; j < pb->n_bkg; ++j)
            if (!std::isfinite(pb->bkg_x[j]) || (j && !(pb->bkg_x[j] > pb->bkg_x[j - 1])))
                return fail(PT_ERR_ARG, "spline background anchors must be finite, distinct and ascending");
    }
    int atoms = 0;
    for (int q = 0; q < pb->n_phases; ++q) {
        const pt_phase &ph = pb->phases[q];
        if (ph.n_ops < 1 || !ph.op_r || !ph.op_t) return fail(PT_ERR_ARG, "a phase has no symmetry operators");
        if (ph.profile_type < 1 || ph.profile_type > 4) return fail(PT_ERR_ARG, "unknown profile type");
        if (ph.free_intensities != 0 && ph.free_intensities != 1) return fail(PT_ERR_ARG, "free_intensities must be 0 or 1");
        if (ph.separated || ph.profile_type == PT_PROFILE_GSAS_CW) {
            if (ph.size_model < 0 || ph.size_model > 2 || ph.strain_model < 0 || ph.strain_model > 2)
                return fail(PT_ERR_ARG, "unknown size or strain broadening model");
        }
        if (ph.n_atoms < 0 || (ph.n_atoms && !ph.atoms)) return fail(PT_ERR_ARG, "a phase has an invalid atom list");
        for (int j = 0; j < ph.n_atoms; ++j)
            if (ph.atoms[j].n_gauss < 0 || ph.atoms[j].n_gauss > PT_MAX_GAUSS) return fail(PT_ERR_ARG, "a scattering factor has more than 5 Gaussians");
        atoms += ph.n_atoms;
    }
    if (atoms > PTK_MAX_ATOMS) return fail(PT_ERR_BUDGET, "PT has more than 160 atoms; using the CPU path.");
    int expect = phase_base_of(pb, pb->n_phases);
    if (pb->n_params != expect || !pb->param_base) return fail(PT_ERR_ARG, "PT model-parameter vector has the wrong length.");
    if (expect > PTK_MAX_PARAMS) return fail(PT_ERR_BUDGET, "PT has more than 512 model parameters.");
    if (pb->n_vars < 1 || pb->n_vars > PTK_MAX_VARS) return fail(PT_ERR_ARG, "PT needs 1 to 256 refined variables.");
    if (!pb->x0 || !pb->lo || !pb->hi || !pb->scale) return fail(PT_ERR_ARG, "PT variables are missing.");
    for (int t = 0; t < pb->n_terms; ++t)
        if (pb->term_param[t] < 0 || pb->term_param[t] >= pb->n_params || pb->term_var[t] < 0 || pb->term_var[t] >= pb->n_vars || !std::isfinite(pb->term_coeff[t]))
            return fail(PT_ERR_ARG, "PT parameter link out of range.");
    if (pb->chains < 1 || pb->chains > 65536) return fail(PT_ERR_ARG, "invalid parallel-tempering settings");
    if (!(pb->max_temp >= 1) || !std::isfinite(pb->max_temp)) return fail(PT_ERR_ARG, "invalid parallel-tempering settings");
    return 0;
}

}  // namespace

extern "C" {

int pt_abi_version(void) { return PT_ABI_VERSION; }
const char *pt_last_error(void) { return g_err.c_str(); }
int pt_phase_base(const pt_problem *prob, int phase) {
    if (!prob || phase < 0 || phase > prob->n_phases) return fail(PT_ERR_ARG, "phase index out of range");
    return phase_base_of(prob, phase);
}

void pt_plan_destroy(pt_plan *pl) {
    if (!pl) return;
    for (pt_plan *f : pl->follow) pt_plan_destroy(f);
    pl->follow.clear();
    if (pl->device >= 0) {
        (void)hipSetDevice(pl->device);
        if (pl->stream) (void)hipStreamSynchronize(pl->stream);
        if (pl->mstream) (void)hipStreamSynchronize(pl->mstream);
        for (void *p : pl->allocs) (void)hipFree(p);
        for (auto e : pl->ev) (void)hipEventDestroy(e);
        if (pl->gexec) (void)hipGraphExecDestroy(pl->gexec);
        if (pl->graph) (void)hipGraphDestroy(pl->graph);
        if (pl->ev_total[0]) (void)hipEventDestroy(pl->ev_total[0]);
        if (pl->ev_total[1]) (void)hipEventDestroy(pl->ev_total[1]);
        for (int p = 0; p + 1 < pt_plan::kMaxParts; ++p) {
            if (pl->side[p]) { (void)hipStreamSynchronize(pl->side[p]); (void)hipStreamDestroy(pl->side[p]); }
            if (pl->join[p]) (void)hipEventDestroy(pl->join[p]);
        }
        if (pl->fork) (void)hipEventDestroy(pl->fork);
        if (pl->mjoin) (void)hipEventDestroy(pl->mjoin);
        if (pl->mstream && pl->mstream_own) (void)hipStreamDestroy(pl->mstream);
        if (pl->stream) (void)hipStreamDestroy(pl->stream);
        mlpx::destroy(pl->mlp);
    }
    delete pl;
}

// Spline background point table: for every point the anchor interval s and the four
// Hermite weights, so that the device forms bkg(t) = c0 y_s + c1 m_s + c2 y_s+1 + c3 m_s+1
// from a trial's anchor values y and tangents m (js/rietveld_core.js:597-612: flat
// outside the anchors, one anchor a constant, none zero; c1, c3 carry the interval width).
static void spline_point_table(const double *tth, int n, const double *xs, int na, std::vector<int> &seg,
                               std::vector<float4> &coef) {
    int s = 0;
    for (int i = 0; i < n; ++i) {
        const double t = tth[i];
        if (na < 1) { seg[i] = 0; coef[i] = make_float4(0.f, 0.f, 0.f, 0.f); continue; }
        if (na == 1 || t <= xs[0]) { seg[i] = 0; coef[i] = make_float4(1.f, 0.f, 0.f, 0.f); continue; }
        if (t >= xs[na - 1]) { seg[i] = na - 2; coef[i] = make_float4(0.f, 0.f, 1.f, 0.f); continue; }
        while (s < na - 2 && xs[s + 1] <= t) ++s;
        while (s > 0 && xs[s] > t) --s;
        const double h = xs[s + 1] - xs[s], u = h > 1e-12 ? (t - xs[s]) / h : 0.0, u2 = u * u, u3 = u2 * u;
        seg[i] = s;
        coef[i] = make_float4((float)(2 * u3 - 3 * u2 + 1), (float)((u3 - 2 * u2 + u) * h), (float)(-2 * u3 + 3 * u2),
                              (float)((u3 - u2) * h));
    }
}

// Fritsch-Carlson tangents of the anchors (js/rietveld_core.js:549-561)
static void spline_tangents(const double *xs, const double *ys, int na, double *m) {
    if (na < 2) { if (na == 1) m[0] = 0; return; }
    std::vector<double> h(na - 1), d(na - 1);
    for (int i = 0; i < na - 1; ++i) { h[i] = xs[i + 1] - xs[i]; d[i] = (ys[i + 1] - ys[i]) / h[i]; }
    m[0] = d[0]; m[na - 1] = d[na - 2];
    for (int i = 1; i < na - 1; ++i) m[i] = d[i - 1] * d[i] <= 0 ? 0 : (h[i] * d[i - 1] + h[i - 1] * d[i]) / (h[i - 1] + h[i]);
    for (int i = 0; i < na - 1; ++i) {
        if (d[i] == 0) { m[i] = 0; m[i + 1] = 0; continue; }
        const double a = m[i] / d[i], b = m[i + 1] / d[i], tau = a * a + b * b;
        if (tau > 9) { const double f = 3 / std::sqrt(tau); m[i] = a * d[i] * f; m[i + 1] = b * d[i] * f; }
    }
}

// the background of the start vector at every point, FP64 (Chebyshev terms over the fitted
// range, js/rietveld_core.js:647-653 and 920-931, or the spline)
static std::vector<double> start_background(const pt_plan *pl) {
    const int n = pl->np;
    std::vector<double> bk(n, 0.0);
    if (pl->bkg_type == PT_BKG_SPLINE) {
        const int na = pl->n_bkg;
        if (na < 1) return bk;
        std::vector<double> ys(na), m(na);
        for (int j = 0; j < na; ++j) { const double v = pl->base_params[PT_P_BKG0 + j]; ys[j] = std::isnan(v) ? 0.0 : v; }
        spline_tangents(pl->bkg_x.data(), ys.data(), na, m.data());
        std::vector<int> seg(n);
        std::vector<float4> c(n);
        const double *xs = pl->bkg_x.data();
        int s = 0;
        for (int i = 0; i < n; ++i) {
            const double t = pl->tth_host[i];
            if (na == 1 || t <= xs[0]) { bk[i] = ys[0]; continue; }
            if (t >= xs[na - 1]) { bk[i] = ys[na - 1]; continue; }
            while (s < na - 2 && xs[s + 1] <= t) ++s;
            while (s > 0 && xs[s] > t) --s;
            const double h = xs[s + 1] - xs[s], u = h > 1e-12 ? (t - xs[s]) / h : 0.0, u2 = u * u, u3 = u2 * u;
            bk[i] = (2 * u3 - 3 * u2 + 1) * ys[s] + (u3 - 2 * u2 + u) * h * m[s] + (-2 * u3 + 3 * u2) * ys[s + 1] +
                    (u3 - u2) * h * m[s + 1];
        }
        return bk;
    }
    const double lo = pl->tth_host[0], hi = pl->tth_host[n - 1];
    for (int i = 0; i < n; ++i) {
        double z = hi > lo ? 2 * (pl->tth_host[i] - lo) / (hi - lo) - 1 : 0, ta = 1, tb = z, b = 0;
        for (int k = 0; k < pl->n_bkg; ++k) {
            double T = k == 0 ? 1 : k == 1 ? z : 2 * z * tb - ta;
            if (k >= 2) { ta = tb; tb = T; }
            b += pl->base_params[PT_P_BKG0 + k] * T;
        }
        bk[i] = b;
    }
    return bk;
}

// ---- separated profiles ------------------------------------------------------------
// GSAS quartic monomials h^a k^b l^c, a + b + c = 4 (js/profile.js:332 order)
const int kQuartic[15][3] = {{4, 0, 0}, {0, 4, 0}, {0, 0, 4}, {2, 2, 0}, {2, 0, 2}, {0, 2, 2}, {3, 1, 0}, {3, 0, 1},
                             {1, 3, 0}, {1, 0, 3}, {0, 3, 1}, {0, 1, 3}, {2, 1, 1}, {1, 2, 1}, {1, 1, 2}};

int quartic_index(int a, int b, int c) {
    for (int i = 0; i < 15; ++i) if (kQuartic[i][0] == a && kQuartic[i][1] == b && kQuartic[i][2] == c) return i;
    return -1;
}
double quartic_weight(int j) {     // 1, 2, 3 or 4 as the exponents hold a 4, a 3, a 0 or none
    const int *e = kQuartic[j];
    auto has = [&](int v) { return e[0] == v || e[1] == v || e[2] == v; };
    return has(4) ? 1.0 : has(3) ? 2.0 : has(0) ? 3.0 : 4.0;
}

// The Stephens strain basis of a setting (js/profile.js:334-363 gsasStephensBasis):
// each weighted monomial of h' = R^T h averaged over the distinct rotations, kept when
// linearly independent of the kept ones, scaled to its key's weight at its leading
// coefficient.  Row k of `basis` (15 x 15) holds the term of key k, or zeros.
bool stephens_basis(const pt_phase &ph, std::vector<float> &basis) {
    std::vector<std::array<int, 9>> rots;
    for (int o = 0; o < ph.n_ops; ++o) {
        std::array<int, 9> r;
        for (int x = 0; x < 9; ++x) r[x] = ph.op_r[9 * o + x];
        if (std::find(rots.begin(), rots.end(), r) == rots.end()) rots.push_back(r);
    }
    basis.assign(15 * 15, 0.0f);
    std::vector<std::array<double, 15>> orth;
    bool any = false;
    for (int j = 0; j < 15; ++j) {
        std::array<double, 15> c{};
        for (const auto &R : rots) {
            // expand prod_axis (sum_k R[3k + axis] h_k)^e_axis as a table over exponents
            std::map<std::array<int, 3>, double> poly{{{0, 0, 0}, quartic_weight(j)}};
            for (int axis = 0; axis < 3; ++axis)
                for (int rep = 0; rep < kQuartic[j][axis]; ++rep) {
                    std::map<std::array<int, 3>, double> next;
                    for (const auto &t : poly)
                        for (int k = 0; k < 3; ++k) {
                            const int cc = R[3 * k + axis];
                            if (!cc) continue;
                            std::array<int, 3> e = t.first;
                            ++e[k];
                            next[e] += t.second * cc;
                        }
                    poly.swap(next);
                }
            for (const auto &t : poly) {
                const int i = quartic_index(t.first[0], t.first[1], t.first[2]);
                if (i >= 0) c[i] += t.second / (double)rots.size();
            }
        }
        std::array<double, 15> v = c;
        for (const auto &q : orth) {
            double dot = 0;
            for (int i = 0; i < 15; ++i) dot += v[i] * q[i];
            for (int i = 0; i < 15; ++i) v[i] -= dot * q[i];
        }
        double norm = 0;
        for (int i = 0; i < 15; ++i) norm += v[i] * v[i];
        norm = std::sqrt(norm);
        if (norm < 1e-8) continue;
        for (int i = 0; i < 15; ++i) v[i] /= norm;
        orth.push_back(v);
        int lead = 0;
        while (lead < 15 && !(std::fabs(c[lead]) > 1e-8)) ++lead;
        const double factor = quartic_weight(lead) / c[lead];
        for (int i = 0; i < 15; ++i) basis[15 * j + i] = std::fabs(c[i]) < 1e-9 ? 0.0f : (float)(c[i] * factor);
        any = true;
    }
    return any;
}

// Stephens guard directions (rietveld_core.js:1456-1478, js/pt_gpu.js:201-205): every
// h k l in [-2, 2] with the first non-zero index positive, then every envelope reflection
// not yet listed; per direction the 15 key coefficients of the quartic at h / |h|.
void stephens_guard(const std::vector<float> &basis, int q, const std::vector<int32_t> &hklm_rows,
                    std::vector<double> &out, int &count) {
    std::vector<std::array<int, 3>> dirs;
    for (int h = -2; h <= 2; ++h)
        for (int k = -2; k <= 2; ++k)
            for (int l = -2; l <= 2; ++l)
                if (h > 0 || (h == 0 && (k > 0 || (k == 0 && l > 0)))) dirs.push_back({h, k, l});
    for (size_t r = 0; r + 4 <= hklm_rows.size(); r += 4) {
        std::array<int, 3> d{hklm_rows[r], hklm_rows[r + 1], hklm_rows[r + 2]};
        if (!d[0] && !d[1] && !d[2]) continue;
        if (std::find(dirs.begin(), dirs.end(), d) == dirs.end()) dirs.push_back(d);
    }
    for (const auto &d : dirs) {
        const double len = std::sqrt((double)d[0] * d[0] + (double)d[1] * d[1] + (double)d[2] * d[2]);
        const double u[3] = {d[0] / len, d[1] / len, d[2] / len};
        double mono[15];
        for (int i = 0; i < 15; ++i)
            mono[i] = std::pow(u[0], kQuartic[i][0]) * std::pow(u[1], kQuartic[i][1]) * std::pow(u[2], kQuartic[i][2]);
        for (int k = 0; k < 15; ++k) {
            double g = 0;
            for (int i = 0; i < 15; ++i) g += (double)basis[15 * k + i] * mono[i];
            out.push_back(g);
        }
    }
    count = (int)dirs.size();
    (void)q;
}

static int create_impl(const pt_problem *pb, pt_plan **out, bool with_device) {
    if (!out) return fail(PT_ERR_ARG, "output handle is NULL");
    *out = nullptr;
    int rc = validate(pb);
    if (rc) return rc;
    pt_plan *pl = new pt_plan();
    pl->nv = pb->n_vars;
    pl->chains = pb->chains;
    pl->stride = 3 * pb->n_vars + PT_S_COUNT;
    pl->np = pb->n_points;
    pl->n_params = pb->n_params;
    pl->steps_per_sweep = pb->steps_per_sweep > 0 ? pb->steps_per_sweep : pb->n_vars;
    const int n = pb->n_points;

    // limits-based dmin (js/pt_gpu.js:104-106, rietveld_core.js:238-245)
    double mag[3];
    for (int s = 0; s < 3; ++s) {
        double lo, hi;
        param_range(pb, s, lo, hi);
        mag[s] = std::max(std::fabs(lo), std::fabs(hi));
    }
    double lambda = pb->lambda1;
    if (pb->ratio > 1e-6 && pb->lambda2 > 0) lambda = std::min(lambda, pb->lambda2);
    double margin = 10.0 + mag[0] + (mag[1] * .001 + mag[2]) * R2D;
    double upper = std::min(179.999999, pb->tth[n - 1] + margin + .001);
    double dmin = lambda / (2 * std::sin(upper * D2R / 2));
    pl->qlim = 1 / (dmin * dmin) * (1 + 1e-9);

    // envelope per phase (js/pt_gpu.js:107-131)
    std::vector<int4> hklm;
    std::vector<int> eq_off{0};
    std::vector<short> eq;
    std::vector<int> rep_r;
    std::vector<double> rep_t;
    std::vector<ptk::DevAtom> atoms;
    std::vector<float> sbasis((size_t)PTK_MAX_PHASES * 225, 0.0f);
    std::vector<double> sguard;
    ptk::DevSepPlan sep_plan{};
    std::vector<int> free_k;             // envelope index of every free-intensity reflection
    ptk::DevPlan &P = pl->P;
    P.any_sep = 0; P.has_stephens = 0;
    for (int q = 0; q < pb->n_phases; ++q) {
        const pt_phase &ph = pb->phases[q];
        int pbase = phase_base_of(pb, q);
        double v[6], lo6[6], hi6[6];
        for (int s = 0; s < 6; ++s) {
            v[s] = param_value(pb, pbase + s);
            param_range(pb, pbase + s, lo6[s], hi6[s]);
        }
        double cell[6];
        cell_from(ph, v, cell);
        double maxes[3] = {hi6[0], hi6[1], hi6[2]};
        if (ph.system == PT_CUBIC) maxes[1] = maxes[2] = maxes[0];
        else if (ph.system == PT_TETRAGONAL || ph.system == PT_HEXAGONAL || ph.system == PT_TRIGONAL) maxes[1] = maxes[0];
        // constrained systems: an unrefined b or c follows a
        for (int s = 0; s < 3; ++s) if (!(maxes[s] > 0)) maxes[s] = cell[s];
        for (int s = 0; s < 3; ++s)
            if (!(maxes[s] > 0) || !std::isfinite(maxes[s])) { delete pl; return fail(PT_ERR_ARG, "PT cell limits need finite positive upper bounds."); }
        int box[3];
        double prod = 1;
        for (int s = 0; s < 3; ++s) { box[s] = (int)std::floor(maxes[s] / dmin); prod *= 2.0 * box[s] + 1; }
        if (prod > 1e6) { delete pl; return fail(PT_ERR_BUDGET, "PT reflection envelope exceeds the setup budget."); }
        bool cellMoves = false, anglesMove = false;
        for (int s : cell_slots(ph)) {
            if (lo6[s] != hi6[s]) { cellMoves = true; if (s >= 3) anglesMove = true; }
        }
        bool orthogonal = cell[3] == 90 && cell[4] == 90 && cell[5] == 90;
        bool monotone = !anglesMove && (orthogonal || ph.system == PT_HEXAGONAL || ph.system == PT_TRIGONAL);
        double env[6] = {maxes[0], maxes[1], maxes[2], cell[3], cell[4], cell[5]};
        std::vector<Refl> list = !cellMoves ? enumerate_reflections(cell, ph.n_ops, ph.op_r, ph.op_t, dmin, nullptr)
                                 : monotone ? enumerate_reflections(env, ph.n_ops, ph.op_r, ph.op_t, dmin, nullptr)
                                            : enumerate_reflections(cell, ph.n_ops, ph.op_r, ph.op_t, dmin, box);
        const std::vector<int> all_reps = coset_reps(ph.n_ops, ph.op_r, ph.op_t);
        const std::vector<int> half = centro_half(all_reps, ph.n_ops, ph.op_r, ph.op_t);
        const std::vector<int> &reps = half.empty() ? all_reps : half;
        ptk::DevPhase &dp = P.phase[q];
        dp.centro = half.empty() ? 0 : 1;
        dp.n_reps_full = (int)all_reps.size();
        dp.system = ph.system;
        dp.unique_axis = ph.unique_axis;
        // a separated profile takes the dataset's instrument family (rietveld_core.js:700-702)
        ptk::DevSepPhase &sd = sep_plan.ph[q];
        sd.sep = (ph.separated || ph.profile_type == PT_PROFILE_GSAS_CW) ? 1 : 0;
        dp.profile_type = sd.sep && pb->resolution_type ? pb->resolution_type : ph.profile_type;
        if (dp.profile_type == PT_PROFILE_GSAS_CW) dp.profile_type = PT_PROFILE_TCH;
        sd.size_model = sd.sep ? ph.size_model : 0;
        sd.strain_model = sd.sep ? ph.strain_model : 0;
        for (int w = 0; w < 2; ++w) {
            // broadeningConfig axis(): three finite numbers, one non-zero, else 0 0 1
            const double *a = w ? ph.strain_axis : ph.size_axis;
            bool ok = std::isfinite(a[0]) && std::isfinite(a[1]) && std::isfinite(a[2]) && (a[0] != 0 || a[1] != 0 || a[2] != 0);
            float *o = w ? sd.strain_axis : sd.size_axis;
            for (int i = 0; i < 3; ++i) o[i] = ok ? (float)a[i] : (i == 2 ? 1.0f : 0.0f);
        }
        sd.stephens = 0; sd.n_sdir = 0; sd.sdir_off = 0;
        sd.free_ph = ph.free_intensities ? 1 : 0;
        dp.n_reps = (int)reps.size();
        dp.rep_off = (int)(rep_r.size() / 9);
        dp.n_atoms = ph.n_atoms;
        dp.atom_off = (int)atoms.size();
        dp.refl_off = (int)hklm.size();
        dp.refl_count = (int)list.size();
        dp.param_base = pbase;
        for (int s = 0; s < 3; ++s) dp.po_dir[s] = ph.po_dir[s];
        dp.fixed_angles = anglesMove ? 0 : 1;
        for (int s = 0; s < 3; ++s) {
            dp.ang[s] = cell[3 + s];
            dp.ang[3 + s] = std::cos(cell[3 + s] * D2R);
            dp.ang_sin[s] = std::sin(cell[3 + s] * D2R);
        }
        if (dp.rep_off + dp.n_reps > PTK_MAX_REPS_TOTAL) { delete pl; return fail(PT_ERR_BUDGET, "PT has more than 64 coset representatives."); }
        for (int r : reps) {
            for (int x = 0; x < 9; ++x) rep_r.push_back(ph.op_r[9 * r + x]);
            for (int x = 0; x < 3; ++x) rep_t.push_back(ph.op_t[3 * r + x]);
        }
        for (int j = 0; j < ph.n_atoms; ++j) {
            const pt_atom &a = ph.atoms[j];
            ptk::DevAtom d{};
            for (int x = 0; x < 3; ++x) { d.r0[x] = a.r0[x]; for (int y = 0; y < 3; ++y) d.basis[x][y] = a.basis[x][y]; }
            d.free_mask = a.free_mask;
            d.n_gauss = a.n_gauss;
            for (int g = 0; g < PT_MAX_GAUSS; ++g) { d.ff_a[g] = (float)a.ff_a[g]; d.ff_b[g] = (float)a.ff_b[g]; }
            d.ff_c = (float)a.ff_c;
            d.w0 = (float)((a.mult != 0 ? a.mult : 1) / (double)all_reps.size());
            d.b_neutron = (float)a.b_neutron;
            d.fp = (float)a.fp;
            d.fpp = (float)a.fpp;
            atoms.push_back(d);
            ptk::JacAtom ja{};
            for (int x = 0; x < 3; ++x) { ja.r0[x] = a.r0[x]; for (int y = 0; y < 3; ++y) ja.basis[x][y] = a.basis[x][y]; }
            for (int g = 0; g < PT_MAX_GAUSS; ++g) { ja.ff_a[g] = a.ff_a[g]; ja.ff_b[g] = a.ff_b[g]; }
            ja.ff_c = a.ff_c;
            ja.w0 = (a.mult != 0 ? a.mult : 1) / (double)all_reps.size();
            ja.b_neutron = a.b_neutron; ja.fp = a.fp; ja.fpp = a.fpp;
            ja.free_mask = a.free_mask; ja.n_gauss = a.n_gauss;
            pl->jac_atoms_h.push_back(ja);
        }
        std::vector<int32_t> rows;
        for (auto &r : list) {
            hklm.push_back(make_int4(r.h, r.k, r.l, r.m));
            for (short s : r.eq) eq.push_back(s);
            eq_off.push_back((int)(eq.size() / 3));
            rows.push_back(r.h); rows.push_back(r.k); rows.push_back(r.l); rows.push_back(r.m);
        }
        if (sd.sep && sd.strain_model == PT_STRAIN_STEPHENS) {
            std::vector<float> basis;
            if (stephens_basis(ph, basis)) {
                sd.stephens = 1;
                std::copy(basis.begin(), basis.end(), sbasis.begin() + (size_t)q * 225);
                sd.sdir_off = (int)(sguard.size() / 15);
                stephens_guard(basis, q, rows, sguard, sd.n_sdir);
            }
        }
        if (sd.sep) P.any_sep = 1;
        if (sd.stephens) P.has_stephens = 1;
        if (sd.free_ph)
            for (int j = 0; j < dp.refl_count; ++j) free_k.push_back(dp.refl_off + j);
        pl->phase_hklm.push_back(std::move(rows));
    }
    pl->nr = (int)hklm.size();
    if (!pl->nr) { delete pl; return fail(PT_ERR_ARG, "No reflections for PT."); }
    if ((int)free_k.size() > PTK_MAX_FREE) {
        delete pl;
        return fail(PT_ERR_BUDGET, "PT has more than 2048 potential free reflection intensities in one dataset; using the CPU solver.");
    }
    P.has_free = free_k.empty() ? 0 : 1;
    P.n_free = (int)free_k.size();

    // grid
    double t0 = pb->tth[0], step = (pb->tth[n - 1] - pb->tth[0]) / (n - 1);
    int uniform = 1;
    for (int i = 0; i < n && uniform; ++i)
        if (std::fabs(pb->tth[i] - (t0 + i * step)) > 1e-9 * step) uniform = 0;
    pl->uniform = uniform;
    std::vector<float2> tff(n), obs(n);
    std::vector<float> zcheb(n);
    std::vector<int> bseg(n, 0);
    std::vector<float4> bcoef(n, make_float4(0.f, 0.f, 0.f, 0.f));
    if (pb->bkg_type == PT_BKG_SPLINE) spline_point_table(pb->tth, n, pb->bkg_x, pb->n_bkg, bseg, bcoef);
    const double tlo = pb->tth[0], thi = pb->tth[n - 1];
    for (int i = 0; i < n; ++i) {
        float hi = (float)pb->tth[i];
        tff[i] = make_float2(hi, (float)(pb->tth[i] - (double)hi));
        obs[i] = make_float2((float)pb->y_obs[i], (float)pb->weight[i]);
        zcheb[i] = (float)(thi > tlo ? 2.0 * (pb->tth[i] - tlo) / (thi - tlo) - 1.0 : 0.0);
    }
    // CSR of the links by parameter
    std::vector<int> term_off(pb->n_params + 1, 0), term_var;
    std::vector<double> term_coeff;
    for (int k = 0; k < pb->n_params; ++k) {
        for (int t = 0; t < pb->n_terms; ++t)
            if (pb->term_param[t] == k) { term_var.push_back(pb->term_var[t]); term_coeff.push_back(pb->term_coeff[t]); }
        term_off[k + 1] = (int)term_var.size();
    }

    pl->yobs_h.assign(pb->y_obs, pb->y_obs + n);
    pl->w_h.assign(pb->weight, pb->weight + n);
    pl->param_base_h.assign(pb->param_base, pb->param_base + pb->n_params);
    pl->term_off_h = term_off; pl->term_var_h = term_var; pl->term_coeff_h = term_coeff;
    if (!with_device) { *out = pl; return 0; }
    rc = device_check(&pl->device);
    if (rc) { delete pl; return rc; }
    (void)hipSetDevice(pl->device);
    if (hipStreamCreateWithFlags(&pl->stream, hipStreamNonBlocking) != hipSuccess) { delete pl; return fail(PT_ERR_DEVICE, "PT stream creation failed."); }
    (void)hipEventCreate(&pl->ev_total[0]);
    (void)hipEventCreate(&pl->ev_total[1]);
    // sweep parts: as many as keep at least 64 chains each, up to four
    pl->parts = std::max(1, std::min(pt_plan::kMaxParts, pb->chains / 64));
    if (pl->parts > 1) {
        bool ok = hipEventCreateWithFlags(&pl->fork, hipEventDisableTiming) == hipSuccess;
        for (int p = 0; ok && p + 1 < pl->parts; ++p)
            ok = hipStreamCreateWithFlags(&pl->side[p], hipStreamNonBlocking) == hipSuccess &&
                 hipEventCreateWithFlags(&pl->join[p], hipEventDisableTiming) == hipSuccess;
        if (!ok) { pt_plan_destroy(pl); return fail(PT_ERR_DEVICE, "PT side stream creation failed."); }
    }

    P.np = n; P.nr = pl->nr; P.nv = pb->n_vars; P.n_params = pb->n_params; P.n_phases = pb->n_phases; P.n_bkg = pb->n_bkg;
    P.chains = pb->chains; P.stride = pl->stride; P.steps_per_sweep = pl->steps_per_sweep;
    P.uniform = uniform; P.radiation = pb->radiation;
    P.doublet = (pb->ratio > 1e-6 && pb->lambda2 > 1e-6 && std::fabs(pb->lambda1 - pb->lambda2) > 1e-6) ? 1 : 0;
    P.t0 = t0; P.step = step; P.inv_step = 1.0 / step;
    P.step_hi = (float)step; P.step_lo = (float)(step - (double)P.step_hi);
    P.tlo = pb->tth[0]; P.thi = pb->tth[n - 1];
    P.lambda1 = pb->lambda1 != 0 ? pb->lambda1 : 1.54056; P.lambda2 = pb->lambda2; P.ratio = pb->ratio;
    P.pol_k = pb->pol_k; P.min_fwhm = pb->min_fwhm > 0 ? pb->min_fwhm : 1e-3;
    P.qlim = pl->qlim;
    P.guard_lo = std::tan(std::max(1e-6, std::min(179.999999, pb->tth[0])) * kPi / 360);
    P.guard_hi = std::tan(std::max(1e-6, std::min(179.999999, pb->tth[n - 1])) * kPi / 360);
    P.max_temp = pb->max_temp;
    P.seed = pb->seed;
    {
        // chunked representative sums in the lines stage: a workgroup of 256 threads holds at
        // most 256 reflections; when it holds fewer and a phase has many representatives
        // (a cubic phase of a small pattern), each reflection's representatives are split
        // into chunks summed on otherwise idle threads (phases of at most
        // PTK_REP_CHUNK_ATOMS atoms)
        int maxrep = 1, maxat = 0;
        for (int q = 0; q < P.n_phases; ++q) {
            maxrep = std::max(maxrep, P.phase[q].n_reps);
            maxat = std::max(maxat, P.phase[q].n_atoms);
        }
        const int nk = std::max(1, std::min(P.nr, 256));
        int ch = 1;
        if (maxat <= PTK_REP_CHUNK_ATOMS && maxrep >= 4) {
            ch = std::max(1, std::min(maxrep, 256 / nk));
            const int rpc = (maxrep + ch - 1) / ch;
            ch = (maxrep + rpc - 1) / rpc;
        }
        P.rep_chunks = ch;
    }

    std::vector<double> tth(pb->tth, pb->tth + n), base(pb->param_base, pb->param_base + pb->n_params);
    pl->x0.assign(pb->x0, pb->x0 + pb->n_vars);
    pl->vlo_h.assign(pb->lo, pb->lo + pb->n_vars);
    pl->vhi_h.assign(pb->hi, pb->hi + pb->n_vars);
    pl->vscale_h.assign(pb->scale, pb->scale + pb->n_vars);
    {
        double wss = 0;
        for (int i = 0; i < pb->n_points; ++i)
            if (pb->weight[i] > 0) wss += pb->weight[i] * pb->y_obs[i] * pb->y_obs[i];
        pl->log_wss = std::log(std::max(wss, 1e-300));
    }
    pl->base_params.resize(pb->n_params);
    for (int k = 0; k < pb->n_params; ++k) pl->base_params[k] = param_value(pb, k);
    pl->tth_host = tth;
    pl->n_bkg = pb->n_bkg;
    pl->bkg_type = pb->bkg_type;
    if (pb->bkg_type == PT_BKG_SPLINE) pl->bkg_x.assign(pb->bkg_x, pb->bkg_x + pb->n_bkg);
    pl->struct_var.assign(pb->n_vars, 1);
    for (int t = 0; t < pb->n_terms; ++t) {
        // structure parameters: atom slots and the phase scale of some phase
        int k = pb->term_param[t], ok = 0;
        for (int q = 0; q < pb->n_phases; ++q) {
            int b = phase_base_of(pb, q);
            if (k == b + PT_PH_SCALE || (k >= b + PT_PH_ATOM0 && k < b + PT_PH_ATOM0 + PT_AT_STRIDE * pb->phases[q].n_atoms)) ok = 1;
        }
        if (!ok) pl->struct_var[pb->term_var[t]] = 0;
    }
    std::vector<double> vlo(pb->lo, pb->lo + pb->n_vars), vhi(pb->hi, pb->hi + pb->n_vars), vsc(pb->scale, pb->scale + pb->n_vars);
    if (eq.empty()) eq.push_back(0);
    if (sguard.empty()) sguard.assign(15, 0.0);
    if (free_k.empty()) free_k.push_back(0);
    int *d_free_k = nullptr;
    float *d_sbasis = nullptr; double *d_sguard = nullptr;
    double *d_tth, *d_bknots = nullptr; float2 *d_tff, *d_obs; float *d_zcheb; int4 *d_hklm; int *d_bseg = nullptr;
    float4 *d_bcoef = nullptr; int *d_eq_off, *d_rep_r, *d_term_off, *d_term_var;
    short *d_eq; double *d_rep_t, *d_base, *d_coeff, *d_lo, *d_hi, *d_sc, *d_state; ptk::DevAtom *d_atoms;
    if ((rc = upload(pl, &d_tth, tth)) || (rc = upload(pl, &d_tff, tff)) || (rc = upload(pl, &d_obs, obs)) ||
        (rc = upload(pl, &d_zcheb, zcheb)) ||
        (pb->bkg_type == PT_BKG_SPLINE &&
         ((rc = upload(pl, &d_bseg, bseg)) || (rc = upload(pl, &d_bcoef, bcoef)) ||
          (rc = upload(pl, &d_bknots, std::vector<double>(pb->bkg_x, pb->bkg_x + std::max(1, pb->n_bkg)))))) ||
        (rc = upload(pl, &d_hklm, hklm)) || (rc = upload(pl, &d_eq_off, eq_off)) || (rc = upload(pl, &d_eq, eq)) ||
        (rc = upload(pl, &d_rep_r, rep_r)) || (rc = upload(pl, &d_rep_t, rep_t)) || (rc = upload(pl, &d_atoms, atoms)) ||
        (rc = upload(pl, &d_base, base)) || (rc = upload(pl, &d_term_off, term_off)) || (rc = upload(pl, &d_term_var, term_var)) ||
        (rc = upload(pl, &d_coeff, term_coeff)) || (rc = upload(pl, &d_lo, vlo)) || (rc = upload(pl, &d_hi, vhi)) ||
        (rc = upload(pl, &d_sc, vsc)) || (rc = upload(pl, &d_sbasis, sbasis)) || (rc = upload(pl, &d_sguard, sguard)) ||
        (rc = upload(pl, &d_free_k, free_k)) ||
        (rc = dalloc(pl, &d_state, (size_t)pl->chains * pl->stride)) || (rc = dalloc(pl, &P.noeval, (size_t)pl->chains))) {
        pt_plan_destroy(pl);
        return rc;
    }
    P.bkg_spline = pb->bkg_type == PT_BKG_SPLINE ? 1 : 0; P.bseg = d_bseg; P.bcoef = d_bcoef; P.bknots = d_bknots;
    {
        if (P.has_free) HIPCHECK(ptk::nnls_lds_limit(ptk::nnls_shared_bytes(P.n_free)), "kernel setup");
        sep_plan.sbasis = d_sbasis; sep_plan.sguard = d_sguard; sep_plan.free_k = d_free_k;
        ptk::DevSepPlan *d_sp = nullptr;
        if ((rc = upload(pl, &d_sp, std::vector<ptk::DevSepPlan>{sep_plan}))) { pt_plan_destroy(pl); return rc; }
        P.sp = d_sp;
        pl->sp_host = sep_plan;
        pl->d_sp = d_sp;
    }
    P.tth = d_tth; P.tff = d_tff; P.zcheb = d_zcheb; P.obs = d_obs; P.hklm = d_hklm; P.eq_off = d_eq_off; P.eq = d_eq;
    P.rep_r = d_rep_r; P.rep_t = d_rep_t; P.atoms = d_atoms; P.base = d_base; P.term_off = d_term_off;
    P.term_var = d_term_var; P.term_coeff = d_coeff; P.vlo = d_lo; P.vhi = d_hi; P.vscale = d_sc; P.state = d_state;
    {
        // the temperature ladder once, as the reference forms it (rietveld_core.js:2195)
        std::vector<double> temps(pl->chains);
        for (int c = 0; c < pl->chains; ++c) temps[c] = std::pow(pb->max_temp, (double)c / (double)(pl->chains > 1 ? pl->chains - 1 : 1));
        double *d_temp = nullptr;
        if ((rc = upload(pl, &d_temp, temps))) { pt_plan_destroy(pl); return rc; }
        P.temp = d_temp;
    }

    // LDS: line-record chunk + y[np]; keep two workgroups per CU when the pattern allows
    size_t stat = ptk::shared_static_bytes() + (3 * PTK_MAX_VARS + PTK_S_COUNT_K) * sizeof(double);
    size_t ybytes = (size_t)n * sizeof(float);
    const size_t rec_bytes = sizeof(ptk::LineRec) + sizeof(int4);
    const size_t lds_max = 160 * 1024;
    if (stat + ybytes + 64 * rec_bytes > lds_max || n > PTK_MAX_GROUPS * 512) { pt_plan_destroy(pl); return fail(PT_ERR_BUDGET, "PT pattern has too many points for one workgroup (LDS)."); }
    size_t target = 80 * 1024;
    long want = 2L * pl->nr;
    long room = (long)((target > stat + ybytes ? target - stat - ybytes : 0) / rec_bytes);
    if (room < 128) room = (long)((lds_max - stat - ybytes) / rec_bytes);
    long chl = std::min(want, room);
    chl = std::max(64L, (chl / 64) * 64);
    pl->ch = (int)chl;
    pl->dyn = rec_bytes * pl->ch + ybytes;
    pl->plain_dyn = rec_bytes * pl->ch + 256 * sizeof(float);
    pl->block = 512;
    pl->ngroups = (n + 511) / 512;
    HIPCHECK(ptk::set_lds_limits(std::max(pl->dyn, pl->plain_dyn)), "kernel setup");
    *out = pl;
    return 0;
}

int pt_plan_create(const pt_problem *pb, pt_plan **out) { return create_impl(pb, out, true); }

int pt_plan_create_multi(const pt_problem *datasets, int n_datasets, pt_plan **out) {
    if (!datasets || !out || n_datasets < 1 || n_datasets > 8) return fail(PT_ERR_ARG, "PT takes 1 to 8 datasets per plan.");
    *out = nullptr;
    for (int d = 1; d < n_datasets; ++d)
        if (datasets[d].n_vars != datasets[0].n_vars) return fail(PT_ERR_ARG, "the datasets of a PT plan must share one variable list.");
    pt_plan *pl = nullptr;
    int rc = create_impl(&datasets[0], &pl, true);
    if (rc) return rc;
    for (int d = 1; d < n_datasets; ++d) {
        pt_problem q = datasets[d];
        // the PT settings are the first dataset's; the others only need room for its chains
        q.chains = datasets[0].chains; q.sweeps = datasets[0].sweeps; q.steps_per_sweep = datasets[0].steps_per_sweep;
        q.max_temp = datasets[0].max_temp; q.seed = datasets[0].seed;
        pt_plan *f = nullptr;
        if ((rc = create_impl(&q, &f, true))) { pt_plan_destroy(pl); return rc; }
        pl->follow.push_back(f);
    }
    *out = pl;
    return 0;
}

int pt_plan_dataset(pt_plan *pl, int dataset, pt_plan **sub) {
    if (!pl || !sub || dataset < 0 || dataset > (int)pl->follow.size()) return fail(PT_ERR_ARG, "dataset index out of range");
    *sub = dataset ? pl->follow[dataset - 1] : pl;
    return 0;
}
int pt_plan_compile(const pt_problem *pb, pt_plan **out) { return create_impl(pb, out, false); }

#define NEED_DEVICE(pl) \
    do { if ((pl)->device < 0) return fail(PT_ERR_STATE, "PT plan was compiled without a device."); } while (0)

int pt_plan_info_get(const pt_plan *pl, pt_plan_info *info) {
    if (!pl || !info) return fail(PT_ERR_ARG, "NULL argument");
    info->nv = pl->nv; info->chains = pl->chains; info->stride = pl->stride; info->point_count = pl->np;
    info->reflection_count = pl->nr; info->line_count = 2 * pl->nr; info->steps_per_sweep = pl->steps_per_sweep;
    info->n_params = pl->n_params; info->uniform_grid = pl->uniform; info->device_ok = pl->device >= 0;
    return 0;
}

int pt_plan_reflections(const pt_plan *pl, int phase, int32_t *hklm, int cap) {
    if (!pl || phase < 0 || phase >= (int)pl->phase_hklm.size()) return fail(PT_ERR_ARG, "phase index out of range");
    const auto &rows = pl->phase_hklm[phase];
    int n = (int)(rows.size() / 4);
    if (hklm) std::memcpy(hklm, rows.data(), sizeof(int32_t) * 4 * std::min(n, std::max(cap, 0)));
    return n;
}

static int ensure_trials(pt_plan *pl, size_t ntrials) {
    if (ntrials <= pl->cap_trials) return 0;
    size_t cap = std::max<size_t>(ntrials, 64);
    int rc;
    if ((rc = dalloc(pl, &pl->d_x, cap * pl->nv)) || (rc = dalloc(pl, &pl->d_chi2, cap))) return rc;
    pl->cap_trials = cap;
    return 0;
}

int pt_set_states(pt_plan *pl, const double *x, int nvec) {
    if (!pl || !x || nvec < 1) return fail(PT_ERR_ARG, "PT initial vectors are missing.");
    NEED_DEVICE(pl);
    (void)hipSetDevice(pl->device);
    std::vector<double> st((size_t)pl->chains * pl->stride, 0.0);
    const int nv = pl->nv;
    for (int c = 0; c < pl->chains; ++c) {
        double *b = st.data() + (size_t)c * pl->stride;
        const double *v = x + (size_t)(c % nvec) * nv;
        for (int j = 0; j < nv; ++j) b[j] = b[nv + j] = b[2 * nv + j] = v[j];
        b[3 * nv + PT_S_COST] = 3e38;
        b[3 * nv + PT_S_BEST] = 3e38;
        b[3 * nv + PT_S_STEP] = -1;
        b[3 * nv + PT_S_VALID] = 1;
    }
    HIPCHECK(hipMemcpyAsync(pl->P.state, st.data(), st.size() * sizeof(double), hipMemcpyHostToDevice, pl->stream), "state upload");
    HIPCHECK(hipMemsetAsync(pl->P.noeval, 0, sizeof(double) * pl->chains, pl->stream), "state upload");
    HIPCHECK(hipStreamSynchronize(pl->stream), "state upload");
    pl->states_set = true;
    return 0;
}

static void drop_graph(hipGraphExec_t &ge, hipGraph_t &g);

// staged-pipeline buffers for `n` trial slots
static int ensure_slots(pt_plan *pl, int n) {
    if (n <= pl->slots) return 0;
    int rc;
    size_t nl = 2 * (size_t)pl->nr;
    // the smaller buffers go (nothing outlives a call in them; the captured sweep graph
    // that points at them is dropped below)
    drop_graph(pl->gexec, pl->graph);
    (void)hipStreamSynchronize(pl->stream);
    dfree(pl, (char **)&pl->d_img); dfree(pl, &pl->d_lrec); dfree(pl, &pl->d_part);
    dfree(pl, &pl->d_schi2);
    if (pl->P.has_free) {
        ptk::DevSepPlan &h = pl->sp_host;
        dfree(pl, &h.fspan); dfree(pl, &h.resid); dfree(pl, &h.fx); dfree(pl, &h.unsolved); dfree(pl, &h.nnls_ws);
    }
    if ((rc = dalloc(pl, (char **)&pl->d_img, (size_t)n * ptk::trial_image_bytes())) ||
        (rc = dalloc(pl, &pl->d_lrec, (size_t)n * nl)) ||
        (rc = dalloc(pl, &pl->d_part, (size_t)n * pl->ngroups)) || (rc = dalloc(pl, &pl->d_schi2, (size_t)n)))
        return rc;
    if (pl->P.has_free) {
        // free intensities: their line spans and the residual they are fitted to, per slot
        ptk::DevSepPlan &h = pl->sp_host;
        if ((rc = dalloc(pl, &h.fspan, (size_t)n * nl)) || (rc = dalloc(pl, &h.resid, (size_t)n * pl->np)) ||
            (rc = dalloc(pl, &h.fx, (size_t)n * pl->P.n_free)) || (rc = dalloc(pl, &h.unsolved, (size_t)n)))
            return rc;
        if (pl->P.n_free > PTK_MAX_FREE_LIVE) {
            // a trial with more live intensities than LDS holds (up to PTK_MAX_FREE_WS_LIVE)
            // solves in its slot's workspace
            h.nnls_stride = ptk::nnls_slot_bytes(pl->P.n_free);
            if ((rc = dalloc(pl, &h.nnls_ws, (size_t)n * h.nnls_stride))) return rc;
        }
        HIPCHECK(hipMemset(h.fx, 0, sizeof(double) * (size_t)n * pl->P.n_free), "intensity start");
        HIPCHECK(hipMemset(h.unsolved, 0, sizeof(int) * (size_t)n), "intensity flags");
        HIPCHECK(hipMemcpy(pl->d_sp, &h, sizeof h, hipMemcpyHostToDevice), "intensity buffers");
        drop_graph(pl->gexec, pl->graph);
    }
    pl->slots = n;
    if (!pl->follow.empty()) {
        // the further datasets get as many slots, and the first dataset's finish their buffers
        std::vector<ptk::DevFollow> fv;
        for (pt_plan *f : pl->follow) {
            if ((rc = ensure_slots(f, n))) return rc;
            fv.push_back(ptk::DevFollow{f->d_img, f->d_part, f->P.has_free ? f->sp_host.unsolved : nullptr, f->ngroups, 0});
        }
        if (!pl->d_follow && (rc = dalloc(pl, &pl->d_follow, fv.size()))) return rc;
        HIPCHECK(hipMemcpy(pl->d_follow, fv.data(), sizeof(ptk::DevFollow) * fv.size(), hipMemcpyHostToDevice), "dataset links");
        pl->P.n_follow = (int)fv.size();
        pl->P.follow = pl->d_follow;
        drop_graph(pl->gexec, pl->graph);
    }
    return 0;
}

// one pass of the staged pipeline over `ntrials` slots; accum timed when ev_pair != null
static int run_stages(pt_plan *pl, const double *dX, int ntrials, int pt_mode, int bootstrap, double *d_chi2,
                      hipEvent_t *ev_pair) {
    HIPCHECK(ptk::launch_stage_setup(pl->P, dX, pl->d_img, ntrials, pt_mode, bootstrap, pl->stream), "setup");
    for (pt_plan *f : pl->follow) {
        if (pt_mode) HIPCHECK(ptk::launch_stage_follow(f->P, pl->P.state, pl->stride, pl->d_img, f->d_img, ntrials, pl->stream), "setup");
        else HIPCHECK(ptk::launch_stage_setup(f->P, dX, f->d_img, ntrials, 0, 0, pl->stream), "setup");
    }
    HIPCHECK(ptk::launch_stage_lines(pl->P, pl->d_img, pl->d_lrec, ntrials, pl->stream), "lines");
    if (ev_pair) HIPCHECK(hipEventRecord(ev_pair[0], pl->stream), "timing");
    HIPCHECK(ptk::launch_stage_accum(pl->P, pl->d_img, pl->d_lrec, pl->d_part, pl->ngroups, ntrials, pl->tpw, pl->stream), "accumulate");
    if (ev_pair) HIPCHECK(hipEventRecord(ev_pair[1], pl->stream), "timing");
    if (pl->P.has_free) HIPCHECK(ptk::launch_stage_nnls(pl->P, pl->d_img, pl->d_lrec, pl->d_part, pl->ngroups, ntrials, pl->stream), "intensities");
    for (pt_plan *f : pl->follow) {
        HIPCHECK(ptk::launch_stage_lines(f->P, f->d_img, f->d_lrec, ntrials, pl->stream), "lines");
        HIPCHECK(ptk::launch_stage_accum(f->P, f->d_img, f->d_lrec, f->d_part, f->ngroups, ntrials, f->tpw, pl->stream), "accumulate");
        if (f->P.has_free) HIPCHECK(ptk::launch_stage_nnls(f->P, f->d_img, f->d_lrec, f->d_part, f->ngroups, ntrials, pl->stream), "intensities");
    }
    HIPCHECK(ptk::launch_stage_finish(pl->P, pl->d_img, pl->d_part, pl->ngroups, ntrials, d_chi2, pt_mode, pl->stream), "finish");
    return 0;
}

static int ensure_events(pt_plan *pl, int pairs) {
    if ((int)pl->ev.size() >= 2 * pairs) return 0;
    for (auto e : pl->ev) (void)hipEventDestroy(e);
    pl->ev.assign(2 * pairs, nullptr);
    for (auto &e : pl->ev) HIPCHECK(hipEventCreate(&e), "event");
    return 0;
}

static double sum_pairs(pt_plan *pl, int pairs) {
    double k = 0;
    for (int j = 0; j < pairs; ++j) {
        float m = 0;
        (void)hipEventElapsedTime(&m, pl->ev[2 * j], pl->ev[2 * j + 1]);
        k += m;
    }
    return k;
}

// trial slots per pass of pt_eval: up to 16384 (larger passes amortise each stage
// launch's ramp and tail: 2048-trial passes ran 10% slower per trial on the Pnma bench
// plan), halved while the slots would take more than 1 GiB, down to 2048
static const int kEvalChunkMax = 16384, kEvalChunkMin = 2048;
static const size_t kEvalSlotBytes = (size_t)1 << 30;
static const int kEvalChunkFree = 512; // the same for plans with free intensities (per-slot solve buffers)

// a plan or one of its further datasets has free intensities
static bool any_free(const pt_plan *pl) {
    if (pl->P.has_free) return true;
    for (const pt_plan *f : pl->follow) if (f->P.has_free) return true;
    return false;
}
static const char *kUnsolved = "MI355X PT intensity solve did not converge.";

static int eval_matrix(pt_plan *pl, const double *x, int ntrials, double *chi2, double *yout, int bf16 = 0);

static int eval_chunk(const pt_plan *pl) {
    if (any_free(pl)) return kEvalChunkFree;
    const size_t per = ptk::trial_image_bytes() + 2 * (size_t)pl->nr * sizeof(ptk::StagedLine) +
                       (size_t)pl->ngroups * sizeof(double2) + sizeof(double);
    int c = kEvalChunkMax;
    while (c > kEvalChunkMin && (size_t)c * per > kEvalSlotBytes) c /= 2;
    return c;
}

static int eval_batched(pt_plan *pl, const double *x, int ntrials, double *chi2, bool time_it) {
    int rc = ensure_trials(pl, ntrials);
    if (rc) return rc;
    int chunk = std::min(ntrials, eval_chunk(pl));
    if ((rc = ensure_slots(pl, std::max(chunk, pl->chains)))) return rc;
    int passes = (ntrials + chunk - 1) / chunk;
    if (time_it && (rc = ensure_events(pl, passes))) return rc;
    HIPCHECK(hipMemcpyAsync(pl->d_x, x, sizeof(double) * ntrials * pl->nv, hipMemcpyHostToDevice, pl->stream), "trial upload");
    HIPCHECK(hipEventRecord(pl->ev_total[0], pl->stream), "timing");
    for (int p = 0; p < passes; ++p) {
        int t0 = p * chunk, n = std::min(chunk, ntrials - t0);
        if ((rc = run_stages(pl, pl->d_x + (size_t)t0 * pl->nv, n, 0, 0, pl->d_chi2 + t0, time_it ? &pl->ev[2 * p] : nullptr)))
            return rc;
    }
    HIPCHECK(hipEventRecord(pl->ev_total[1], pl->stream), "timing");
    HIPCHECK(hipMemcpyAsync(chi2, pl->d_chi2, sizeof(double) * ntrials, hipMemcpyDeviceToHost, pl->stream), "readback");
    HIPCHECK(hipStreamSynchronize(pl->stream), "batched evaluation");
    float ms = 0;
    (void)hipEventElapsedTime(&ms, pl->ev_total[0], pl->ev_total[1]);
    pl->last_total_ms = ms;
    pl->last_kernel_ms = time_it ? sum_pairs(pl, passes) : 0;
    pl->last_launches = time_it ? passes : 0;
    // an intensity solve that did not converge comes back as NaN (js/pt_gpu.js:402 throws)
    for (int t = 0; t < ntrials; ++t)
        if (std::isnan(chi2[t])) return fail(PT_ERR_NUMERIC, kUnsolved);
    return 0;
}

int pt_costs(pt_plan *pl, const double *x, int nvec, double *chi2) {
    if (!pl || !x || !chi2 || nvec < 1) return fail(PT_ERR_ARG, "NULL argument");
    NEED_DEVICE(pl);
    int rc = pt_set_states(pl, x, nvec);
    if (rc) return rc;
    std::vector<double> X((size_t)pl->chains * pl->nv);
    for (int c = 0; c < pl->chains; ++c) std::memcpy(&X[(size_t)c * pl->nv], x + (size_t)(c % nvec) * pl->nv, sizeof(double) * pl->nv);
    return eval_batched(pl, X.data(), pl->chains, chi2, false);
}

// ---- tile-shape autotuner ------------------------------------------------------------
// Times `launch(tpw)` (one accumulate stage over the live line records) at 1, 2 and 4
// 128-point tiles per wavefront: one untimed launch, then the best of three timed ones.
// Writes the times to ms[3] and the fastest shape to *best.
extern "C++" {
template <class F>
static int time_tiles(hipStream_t s, F launch, double ms[3], int *best) {
    static const int cand[3] = {1, 2, 4};
    hipEvent_t e[2];
    HIPCHECK(hipEventCreate(&e[0]), "event");
    HIPCHECK(hipEventCreate(&e[1]), "event");
    int rc = 0;
    for (int i = 0; i < 3 && !rc; ++i) {
        double m = 1e30;
        for (int r = 0; r < 4 && !rc; ++r) {
            if (hipEventRecord(e[0], s) != hipSuccess || launch(cand[i]) != hipSuccess || hipEventRecord(e[1], s) != hipSuccess ||
                hipEventSynchronize(e[1]) != hipSuccess) {
                rc = fail(PT_ERR_DEVICE, "PT device lost during tile tuning.");
                break;
            }
            float t = 0;
            (void)hipEventElapsedTime(&t, e[0], e[1]);
            if (r > 0) m = std::min(m, (double)t);
        }
        ms[i] = m;
    }
    (void)hipEventDestroy(e[0]);
    (void)hipEventDestroy(e[1]);
    if (rc) return rc;
    // the smallest shape unless a larger one is clearly faster: smaller tiles give more
    // workgroups per launch, which the sweep's parts overlap better than the isolated
    // timing shows, and a near tie is timing noise
    int b = 0;
    for (int i = 1; i < 3; ++i)
        if (ms[i] < 0.95 * ms[b]) b = i;
    *best = cand[b];
    return 0;
}
}  // extern "C++"

static void drop_graph(hipGraphExec_t &ge, hipGraph_t &g) {
    if (ge) (void)hipGraphExecDestroy(ge);
    if (g) (void)hipGraphDestroy(g);
    ge = nullptr;
    g = nullptr;
}

// tile shape of the accumulate stage for this plan: the chains' current vectors go
// through setup and lines once (batch mode, the PT state is not touched), then the
// accumulate stage is timed at each shape
static int tune_plan(pt_plan *pl) {
    int rc = ensure_trials(pl, pl->chains);
    if (rc || (rc = ensure_slots(pl, pl->chains))) return rc;
    const size_t nvb = sizeof(double) * pl->nv;
    HIPCHECK(hipMemcpy2DAsync(pl->d_x, nvb, pl->P.state, sizeof(double) * pl->stride, nvb, pl->chains, hipMemcpyDeviceToDevice,
                              pl->stream), "tile tuning");
    HIPCHECK(ptk::launch_stage_setup(pl->P, pl->d_x, pl->d_img, pl->chains, 0, 0, pl->stream), "setup");
    HIPCHECK(ptk::launch_stage_lines(pl->P, pl->d_img, pl->d_lrec, pl->chains, pl->stream), "lines");
    int best = 2;
    rc = time_tiles(pl->stream, [&](int tpw) {
        return ptk::launch_stage_accum(pl->P, pl->d_img, pl->d_lrec, pl->d_part, pl->ngroups, pl->chains, tpw, pl->stream);
    }, pl->tune_ms, &best);
    if (rc) return rc;
    if (best != pl->tpw) drop_graph(pl->gexec, pl->graph);
    pl->tpw = best;
    for (pt_plan *f : pl->follow) f->tpw = best;   // datasets of one plan share the tile shape
    pl->tuned = true;
    return 0;
}

int pt_tiles(pt_plan *pl, int request, int *tiles_per_wave, double *candidate_ms) {
    if (!pl) return fail(PT_ERR_ARG, "NULL plan");
    NEED_DEVICE(pl);
    if (request == -1) {
        if (!pl->states_set) return fail(PT_ERR_STATE, "PT tile tuning before initial state.");
        (void)hipSetDevice(pl->device);
        int rc = tune_plan(pl);
        if (rc) return rc;
    } else if (request == 1 || request == 2 || request == 4) {
        if (request != pl->tpw) drop_graph(pl->gexec, pl->graph);
        pl->tpw = request;
        pl->tuned = true;
    } else if (request != 0) {
        return fail(PT_ERR_ARG, "tile request must be 0, -1, 1, 2 or 4");
    }
    if (tiles_per_wave) *tiles_per_wave = pl->tpw;
    if (candidate_ms) for (int i = 0; i < 3; ++i) candidate_ms[i] = pl->tune_ms[i];
    return 0;
}

// one screened sweep: every step trains the surrogate on the replay ring, draws K
// candidate moves per chain, predicts their log Rwp, takes the lowest, then evaluates it
// exactly (setup, lines, accumulate) and runs the Metropolis step (finish); the trial is
// recorded into the ring.  Then the replica exchange.
static int enqueue_screened_sweep(pt_plan *pl, hipEvent_t *ev) {
    const pt_screen_config &c = pl->scfg;
    const double *step_ptr = pl->P.state + 3 * pl->nv + PT_S_STEP;
    for (int s = 0; s < pl->steps_per_sweep; ++s) {
        const bool train = c.train_every <= 1 || s % c.train_every == 0;
        for (int it = 0; train && it < c.train_iters; ++it) {
            int rc = mlpx::train_iter(pl->mlp, pl->S.ring_x, pl->S.ring_t, pl->screen_cap, step_ptr, pl->chains, 0, c.batch,
                                      c.seed, c.batch, pl->stream);
            if (rc) return rc;
        }
        HIPCHECK(ptk::screen_candidates(pl->P, pl->S, pl->stream), "screen");
        int rc = mlpx::infer(pl->mlp, pl->S.xs, pl->chains * (c.oversample + 1), pl->S.pred, pl->stream);
        if (rc) return rc;
        HIPCHECK(ptk::screen_pick(pl->P, pl->S, pl->mlp->iter, pl->stream), "screen");
        HIPCHECK(ptk::launch_stage_setup(pl->P, nullptr, pl->d_img, pl->chains, 1, 0, pl->stream, pl->S.pick), "setup");
        HIPCHECK(ptk::launch_stage_lines(pl->P, pl->d_img, pl->d_lrec, pl->chains, pl->stream), "lines");
        if (ev) HIPCHECK(hipEventRecord(ev[2 * s], pl->stream), "timing");
        HIPCHECK(ptk::launch_stage_accum(pl->P, pl->d_img, pl->d_lrec, pl->d_part, pl->ngroups, pl->chains, pl->tpw, pl->stream), "accumulate");
        if (ev) HIPCHECK(hipEventRecord(ev[2 * s + 1], pl->stream), "timing");
        HIPCHECK(ptk::launch_stage_finish(pl->P, pl->d_img, pl->d_part, pl->ngroups, pl->chains, pl->d_schi2, 1, pl->stream), "finish");
        HIPCHECK(ptk::screen_record(pl->P, pl->S, pl->stream), "screen");
    }
    HIPCHECK(ptk::launch_exchange(pl->P, pl->stream), "exchange");
    return 0;
}

// one sweep: steps_per_sweep x (setup/propose, lines, accumulate, finish/accept), then the
// replica exchange; the finish of each step but the last is fused with the next setup.
// Accumulate launches are bracketed by events when ev != null.
// the further datasets' setup of chains [t0, t0 + n) from the first dataset's proposals
static int enqueue_follow(pt_plan *pl, hipStream_t st, int t0, int n) {
    for (pt_plan *f : pl->follow)
        HIPCHECK(ptk::launch_stage_follow(f->P, pl->P.state, pl->stride, pl->d_img, f->d_img, n, st, t0), "setup");
    return 0;
}

static int enqueue_part(pt_plan *pl, hipStream_t st, int t0, int n, hipEvent_t *ev) {
    HIPCHECK(ptk::launch_stage_setup(pl->P, nullptr, pl->d_img, n, 1, 0, st, nullptr, t0), "setup");
    int rc = enqueue_follow(pl, st, t0, n);
    if (rc) return rc;
    // a plan whose reflections fit one lines workgroup runs each step's finish, the next
    // setup and its lines as one launch (not with the surrogate's skip test)
    const bool fuse = pl->P.nr <= 512 && !pl->P.sf;
    for (int s = 0; s < pl->steps_per_sweep; ++s) {
        if (!fuse || s == 0) HIPCHECK(ptk::launch_stage_lines(pl->P, pl->d_img, pl->d_lrec, n, st, t0), "lines");
        if (ev) HIPCHECK(hipEventRecord(ev[2 * s], st), "timing");
        HIPCHECK(ptk::launch_stage_accum(pl->P, pl->d_img, pl->d_lrec, pl->d_part, pl->ngroups, n, pl->tpw, st, t0),
                 "accumulate");
        if (ev) HIPCHECK(hipEventRecord(ev[2 * s + 1], st), "timing");
        if (pl->P.has_free) HIPCHECK(ptk::launch_stage_nnls(pl->P, pl->d_img, pl->d_lrec, pl->d_part, pl->ngroups, n, st, t0), "intensities");
        for (pt_plan *f : pl->follow) {
            HIPCHECK(ptk::launch_stage_lines(f->P, f->d_img, f->d_lrec, n, st, t0), "lines");
            HIPCHECK(ptk::launch_stage_accum(f->P, f->d_img, f->d_lrec, f->d_part, f->ngroups, n, f->tpw, st, t0),
                     "accumulate");
            if (f->P.has_free) HIPCHECK(ptk::launch_stage_nnls(f->P, f->d_img, f->d_lrec, f->d_part, f->ngroups, n, st, t0), "intensities");
        }
        if (s + 1 < pl->steps_per_sweep) {
            if (fuse) {
                HIPCHECK(ptk::launch_stage_finish_lines(pl->P, pl->d_img, pl->d_part, pl->ngroups, pl->d_lrec, n, st, t0), "finish");
                if ((rc = enqueue_follow(pl, st, t0, n))) return rc;
                continue;
            }
            HIPCHECK(ptk::launch_stage_finish_setup(pl->P, pl->d_img, pl->d_part, pl->ngroups, st, n, t0), "finish");
            if ((rc = enqueue_follow(pl, st, t0, n))) return rc;
        } else
            HIPCHECK(ptk::launch_stage_finish(pl->P, pl->d_img, pl->d_part, pl->ngroups, n, pl->d_schi2, 1, st, t0), "finish");
    }
    return 0;
}

static int enqueue_sweep(pt_plan *pl, hipEvent_t *ev) {
    if (pl->screen_on && !pl->screen_fused) return enqueue_screened_sweep(pl, ev);
    if (pl->screen_on) {
        // the network's training steps of this sweep, before its PT steps read the weights
        const pt_screen_config &c = pl->scfg;
        const double *step_ptr = pl->P.state + 3 * pl->nv + PT_S_STEP;
        for (int it = 0; it < c.train_iters; ++it) {
            int rc = mlpx::train_iter(pl->mlp, pl->S.ring_x, pl->S.ring_t, pl->screen_cap, step_ptr, pl->chains, 0, c.batch,
                                      c.seed, c.batch, pl->stream);
            if (rc) return rc;
        }
    }
    // chains take no part in each other's steps until the exchange, so the parts run
    // independently between a fork and a join (the timed form runs as one part)
    const int parts = ev ? 1 : pl->parts;
    if (parts <= 1) {
        int rc = enqueue_part(pl, pl->stream, 0, pl->chains, ev);
        if (rc) return rc;
    } else {
        const int chunk = ((pl->chains + parts - 1) / parts + 7) / 8 * 8;
        HIPCHECK(hipEventRecord(pl->fork, pl->stream), "fork");
        for (int p = 1; p < parts; ++p) HIPCHECK(hipStreamWaitEvent(pl->side[p - 1], pl->fork, 0), "fork");
        for (int p = 0; p < parts; ++p) {
            const int t0 = p * chunk, n = std::min(chunk, pl->chains - t0);
            if (n <= 0) continue;
            int rc = enqueue_part(pl, p ? pl->side[p - 1] : pl->stream, t0, n, nullptr);
            if (rc) return rc;
        }
        for (int p = 1; p < parts; ++p) {
            HIPCHECK(hipEventRecord(pl->join[p - 1], pl->side[p - 1]), "join");
            HIPCHECK(hipStreamWaitEvent(pl->stream, pl->join[p - 1], 0), "join");
        }
    }
    HIPCHECK(ptk::launch_exchange(pl->P, pl->stream), "exchange");
    return 0;
}

int pt_screen(pt_plan *pl, const pt_screen_config *cfg) {
    if (!pl) return fail(PT_ERR_ARG, "NULL plan");
    if (!pl->follow.empty() && cfg) return fail(PT_ERR_STATE, "PT surrogate screening takes one dataset per plan.");
    if (pl->P.has_free && cfg) return fail(PT_ERR_STATE, "PT surrogate screening takes Rietveld phases only.");
    NEED_DEVICE(pl);
    (void)hipSetDevice(pl->device);
    drop_graph(pl->gexec, pl->graph);
    if (!cfg) {
        pl->screen_on = false;
        pl->screen_fused = false;
        pl->P.sf = nullptr;
        return 0;
    }
    const pt_screen_config c = *cfg;
    if (pl->nv > ptk::MLP_IN) return fail(PT_ERR_ARG, "PT surrogate screen supports at most 64 variables.");
    if (c.oversample < 1 || c.oversample > 16) return fail(PT_ERR_ARG, "PT surrogate oversampling must be 1..16.");
    if (c.batch < 64 || c.batch > 4096 || c.batch % 64) return fail(PT_ERR_ARG, "PT surrogate batch must be a multiple of 64 in 64..4096.");
    if (c.train_iters < 0 || c.warmup_iters < 0 || !(c.lr > 0.0f) || !(c.margin >= 0.0f) || !(c.cold_fraction > 0.0f) ||
        c.cold_fraction > 1.0f || !(c.skip_margin >= 0.0f) || c.skip_eps < 0.0f || c.skip_eps >= 1.0f)
        return fail(PT_ERR_ARG, "invalid surrogate training settings.");
    if (c.capacity < c.batch || c.capacity < pl->chains) return fail(PT_ERR_ARG, "PT surrogate ring must hold a batch and one step of every chain.");
    int rc = 0;
    if (!pl->mlp && (rc = mlpx::create(c.seed ^ 0x6a09e667u, &pl->mlp))) return rc;
    const int K = c.oversample;
    if ((rc = mlpx::ensure_rows(pl->mlp, std::max(c.batch, pl->chains * (K + 1))))) return rc;
    pl->mlp->lr = c.lr;
    if (c.capacity != pl->screen_cap || K != pl->S.K) {
        ptk::ScreenDev &S = pl->S;
        if ((rc = dalloc(pl, &S.ring_x, (size_t)c.capacity * ptk::MLP_IN)) || (rc = dalloc(pl, &S.ring_t, (size_t)c.capacity)) ||
            (rc = dalloc(pl, &S.cand, (size_t)pl->chains * K)) || (rc = dalloc(pl, &S.xs, (size_t)pl->chains * (K + 1) * ptk::MLP_IN)) ||
            (rc = dalloc(pl, &S.pred, (size_t)pl->chains * (K + 1))) || (rc = dalloc(pl, &S.pick, (size_t)pl->chains)))
            return rc;
        if (!S.picked && ((rc = dalloc(pl, &S.picked, 1)) || (rc = dalloc(pl, &S.skipped, 1)))) return rc;
        if (!S.center) {
            std::vector<double> ctr(pl->nv), ih(pl->nv);
            for (int j = 0; j < pl->nv; ++j) {
                const double lo = pl->vlo_h[j], hi = pl->vhi_h[j];
                if (std::isfinite(lo) && std::isfinite(hi) && hi > lo) {
                    ctr[j] = 0.5 * (lo + hi);
                    ih[j] = 2.0 / (hi - lo);
                } else {
                    const double sc = pl->vscale_h[j] > 0 ? pl->vscale_h[j] : 1.0;
                    ctr[j] = pl->x0[j];
                    ih[j] = 1.0 / (8.0 * sc);
                }
            }
            double *dc = nullptr, *dh = nullptr;
            if ((rc = upload(pl, &dc, ctr)) || (rc = upload(pl, &dh, ih))) return rc;
            S.center = dc;
            S.inv_half = dh;
        }
        HIPCHECK(hipMemset(S.ring_x, 0, sizeof(uint16_t) * (size_t)c.capacity * ptk::MLP_IN), "screen setup");
        HIPCHECK(hipMemset(S.ring_t, 0xff, sizeof(float) * (size_t)c.capacity), "screen setup");   // NaN: empty
        pl->screen_cap = c.capacity;
        S.K = K;
        S.cap = c.capacity;
    }
    HIPCHECK(hipMemset(pl->S.picked, 0, sizeof(unsigned)), "screen setup");
    HIPCHECK(hipMemset(pl->S.skipped, 0, sizeof(unsigned)), "screen setup");
    pl->S.skip_margin = c.skip_margin;
    pl->S.skip_log = c.skip_eps > 0.0f ? (float)-std::log((double)c.skip_eps) : 0.0f;
    pl->S.warmup_iters = c.warmup_iters;
    pl->S.log_wss = pl->log_wss;
    pl->S.margin = c.margin;
    pl->S.cold_chains = std::max(1, (int)std::lround(c.cold_fraction * pl->chains));
    pl->scfg = c;
    pl->screen_on = true;
    // one proposal per chain and training between sweeps: the skip test runs inside the PT
    // setup and the sweep keeps its parts on their streams
    pl->screen_fused = c.oversample == 1 && c.train_every >= pl->steps_per_sweep;
    pl->P.sf = nullptr;
    if (pl->screen_fused) {
        ptk::ScreenFused F{};
        F.w = pl->mlp->wb; F.p = pl->mlp->p; F.iter = pl->mlp->iter;
        F.center = pl->S.center; F.inv_half = pl->S.inv_half;
        F.ring_x = pl->S.ring_x; F.ring_t = pl->S.ring_t; F.skipped = pl->S.skipped;
        F.log_wss = pl->S.log_wss; F.skip_margin = pl->S.skip_margin; F.skip_log = pl->S.skip_log;
        F.cold_chains = pl->S.cold_chains; F.cap = pl->S.cap; F.warmup_iters = pl->S.warmup_iters;
        if (!pl->d_sf && (rc = dalloc(pl, &pl->d_sf, 1))) return rc;
        HIPCHECK(hipMemcpy(pl->d_sf, &F, sizeof F, hipMemcpyHostToDevice), "screen setup");
        pl->P.sf = pl->d_sf;
    }
    return 0;
}

int pt_screen_stats(pt_plan *pl, double *out) {
    if (!pl || !out) return fail(PT_ERR_ARG, "NULL argument");
    NEED_DEVICE(pl);
    for (int i = 0; i < 5; ++i) out[i] = 0;
    if (!pl->mlp) return 0;
    (void)hipSetDevice(pl->device);
    int it = 0;
    float loss[2] = {0, 0};
    unsigned picked = 0, skipped = 0;
    double step = 0;
    HIPCHECK(hipStreamSynchronize(pl->stream), "screen statistics");
    HIPCHECK(hipMemcpy(&it, pl->mlp->iter, sizeof(int), hipMemcpyDeviceToHost), "screen statistics");
    HIPCHECK(hipMemcpy(loss, pl->mlp->loss, 2 * sizeof(float), hipMemcpyDeviceToHost), "screen statistics");
    if (pl->S.picked) HIPCHECK(hipMemcpy(&picked, pl->S.picked, sizeof(unsigned), hipMemcpyDeviceToHost), "screen statistics");
    if (pl->S.skipped) HIPCHECK(hipMemcpy(&skipped, pl->S.skipped, sizeof(unsigned), hipMemcpyDeviceToHost), "screen statistics");
    HIPCHECK(hipMemcpy(&step, pl->P.state + 3 * pl->nv + PT_S_STEP, sizeof(double), hipMemcpyDeviceToHost), "screen statistics");
    out[0] = it;
    out[1] = std::min((double)pl->screen_cap, std::max(0.0, step) * pl->chains);
    out[2] = loss[1];
    out[3] = picked;
    out[4] = skipped;
    return 0;
}

int pt_screen_predict(pt_plan *pl, const double *x, int n, float *pred) {
    if (!pl || !x || !pred || n < 1) return fail(PT_ERR_ARG, "NULL argument");
    NEED_DEVICE(pl);
    if (!pl->mlp || !pl->S.center) return fail(PT_ERR_STATE, "PT surrogate screen is not set up on this plan.");
    (void)hipSetDevice(pl->device);
    int rc = ensure_trials(pl, n);
    if (rc || (rc = mlpx::ensure_rows(pl->mlp, n))) return rc;
    HIPCHECK(hipMemcpyAsync(pl->d_x, x, sizeof(double) * n * pl->nv, hipMemcpyHostToDevice, pl->stream), "upload");
    HIPCHECK(ptk::screen_inputs(pl->d_x, n, pl->nv, pl->S, pl->mlp->x16, pl->stream), "screen");
    if ((rc = mlpx::infer(pl->mlp, pl->mlp->x16, n, pl->mlp->y, pl->stream))) return rc;
    HIPCHECK(hipMemcpyAsync(pred, pl->mlp->y, sizeof(float) * n, hipMemcpyDeviceToHost, pl->stream), "readback");
    HIPCHECK(hipStreamSynchronize(pl->stream), "screen prediction");
    return 0;
}

int pt_set_kernel_timing(pt_plan *pl, int on) {
    if (!pl) return fail(PT_ERR_ARG, "NULL plan");
    pl->kernel_timing = on != 0;
    return 0;
}

// The chi^2 of ntrials device-resident vectors through the staged pipeline of the batched
// form, staged in the library's fixed device scratch: no allocation of any kind, on the
// plan's stream, synchronised before returning.  Calls that overlap on one GPU share the
// scratch and race.
int pt_eval_noalloc(pt_plan *pl, const double *d_x, int ntrials, double *d_chi2) {
    if (!pl || !d_x || !d_chi2 || ntrials < 1) return fail(PT_ERR_ARG, "NULL argument");
    NEED_DEVICE(pl);
    if (!pl->follow.empty() || pl->P.has_free)
        return fail(PT_ERR_STATE, "PT allocation-free evaluation takes one dataset without free intensities (use pt_eval).");
    (void)hipSetDevice(pl->device);
    unsigned char *base = nullptr;
    size_t cap = 0;
    HIPCHECK(ptk::noalloc_scratch(pl->device, &base, &cap), "scratch lookup");
    auto up = [](size_t b) { return (b + 255) / 256 * 256; };
    const size_t n = (size_t)ntrials, nl = 2 * (size_t)pl->nr;
    const size_t img_b = up(n * ptk::trial_image_bytes()), rec_b = up(n * nl * sizeof(ptk::StagedLine)),
                 part_b = up(n * pl->ngroups * sizeof(double2));
    if (img_b + rec_b + part_b > cap) {
        const size_t per = ptk::trial_image_bytes() + nl * sizeof(ptk::StagedLine) + pl->ngroups * sizeof(double2);
        return fail(PT_ERR_BUDGET, "PT allocation-free evaluation takes at most " + std::to_string((cap - 1024) / per) +
                                       " trials of this plan per call (" + std::to_string(cap >> 20) + " MiB scratch).");
    }
    void *img = base;
    ptk::StagedLine *rec = reinterpret_cast<ptk::StagedLine *>(base + img_b);
    double2 *part = reinterpret_cast<double2 *>(base + img_b + rec_b);
    HIPCHECK(ptk::launch_stage_setup(pl->P, d_x, img, ntrials, 0, 0, pl->stream), "setup");
    HIPCHECK(ptk::launch_stage_lines(pl->P, img, rec, ntrials, pl->stream), "lines");
    HIPCHECK(ptk::launch_stage_accum(pl->P, img, rec, part, pl->ngroups, ntrials, pl->tpw, pl->stream), "accumulate");
    HIPCHECK(ptk::launch_stage_finish(pl->P, img, part, pl->ngroups, ntrials, d_chi2, 0, pl->stream), "finish");
    HIPCHECK(hipStreamSynchronize(pl->stream), "allocation-free evaluation");
    return 0;
}

int pt_set_solve_limit(pt_plan *pl, int max_iterations) {
    if (!pl) return fail(PT_ERR_ARG, "NULL plan");
    if (max_iterations < 0) return fail(PT_ERR_ARG, "negative solve limit");
    pl->P.nnls_cap = max_iterations;
    for (pt_plan *f : pl->follow) f->P.nnls_cap = max_iterations;
    drop_graph(pl->gexec, pl->graph);          // the sweep graph holds the plan by value
    return 0;
}

int pt_batch(pt_plan *pl, int count, int bootstrap, double *state_out) {
    if (!pl) return fail(PT_ERR_ARG, "NULL plan");
    NEED_DEVICE(pl);
    if (!pl->states_set) return fail(PT_ERR_STATE, "PT batch before initial state.");
    if (count < 0) return fail(PT_ERR_ARG, "negative sweep count");
    (void)hipSetDevice(pl->device);
    int rc = ensure_slots(pl, pl->chains);
    if (rc) return rc;
    HIPCHECK(hipEventRecord(pl->ev_total[0], pl->stream), "timing");
    int timed = 0;
    if (bootstrap) {
        if ((rc = run_stages(pl, nullptr, pl->chains, 1, 1, pl->d_schi2, nullptr))) return rc;
    } else if (pl->kernel_timing) {
        if (!pl->tuned && (rc = tune_plan(pl))) return rc;
        timed = count * pl->steps_per_sweep;
        if ((rc = ensure_events(pl, timed))) return rc;
        for (int j = 0; j < count; ++j)
            if ((rc = enqueue_sweep(pl, &pl->ev[2 * (size_t)j * pl->steps_per_sweep]))) return rc;
    } else if (count > 0) {
        if (!pl->tuned && (rc = tune_plan(pl))) return rc;
        if (!pl->gexec) {
            HIPCHECK(hipStreamBeginCapture(pl->stream, hipStreamCaptureModeThreadLocal), "graph capture");
            rc = enqueue_sweep(pl, nullptr);
            hipGraph_t g = nullptr;
            hipError_t e = hipStreamEndCapture(pl->stream, &g);
            if (rc) return rc;
            if (e != hipSuccess) return fail(PT_ERR_DEVICE, std::string("PT graph capture failed: ") + hipGetErrorString(e));
            pl->graph = g;
            HIPCHECK(hipGraphInstantiate(&pl->gexec, pl->graph, nullptr, nullptr, 0), "graph instantiation");
        }
        for (int j = 0; j < count; ++j) HIPCHECK(hipGraphLaunch(pl->gexec, pl->stream), "sweep");
    }
    HIPCHECK(hipEventRecord(pl->ev_total[1], pl->stream), "timing");
    // with free intensities the snapshot is read anyway: a chain whose solve did not
    // converge has its unsolved flag set (state +9), and the batch reports it as the
    // reference's run() does (js/pt_gpu.js:418), after filling state_out
    std::vector<double> own;
    double *snap = state_out;
    if (!snap && any_free(pl)) { own.resize((size_t)pl->chains * pl->stride); snap = own.data(); }
    if (snap)
        HIPCHECK(hipMemcpyAsync(snap, pl->P.state, sizeof(double) * pl->chains * pl->stride, hipMemcpyDeviceToHost, pl->stream), "readback");
    HIPCHECK(hipStreamSynchronize(pl->stream), "sweeps");
    float ms = 0;
    (void)hipEventElapsedTime(&ms, pl->ev_total[0], pl->ev_total[1]);
    pl->last_total_ms = ms;
    pl->last_kernel_ms = timed ? sum_pairs(pl, timed) : 0;
    pl->last_launches = timed;
    if (snap && any_free(pl))
        for (int c = 0; c < pl->chains; ++c)
            if (snap[(size_t)c * pl->stride + 3 * pl->nv + PT_S_UNSOLVED] != 0.0) return fail(PT_ERR_NUMERIC, kUnsolved);
    return 0;
}

int pt_eval(pt_plan *pl, int form, const double *x, int ntrials, double *chi2) {
    if (!pl || !x || !chi2 || ntrials < 1) return fail(PT_ERR_ARG, "NULL argument");
    NEED_DEVICE(pl);
    (void)hipSetDevice(pl->device);
    if (form == PT_FORM_BATCHED) return eval_batched(pl, x, ntrials, chi2, true);
    if (!pl->follow.empty()) return fail(PT_ERR_STATE, "PT evaluation form not available for a multi-dataset plan (use the batched form).");
    if (pl->P.has_free) return fail(PT_ERR_STATE, "PT evaluation form not available with free intensities (use the batched form).");
    if (form == PT_FORM_PLAIN) {
        int rc = ensure_trials(pl, ntrials);
        if (rc) return rc;
        if (!pl->d_plain) {
            int nblk = (pl->np + 255) / 256;
            if ((rc = dalloc(pl, (char **)&pl->d_plain, ptk::plain_scratch_bytes())) || (rc = dalloc(pl, &pl->d_rec, 2 * (size_t)pl->nr)) ||
                (rc = dalloc(pl, &pl->d_span, 2 * (size_t)pl->nr)) || (rc = dalloc(pl, &pl->d_partial, (size_t)nblk)))
                return rc;
        }
        HIPCHECK(hipMemcpyAsync(pl->d_x, x, sizeof(double) * ntrials * pl->nv, hipMemcpyHostToDevice, pl->stream), "trial upload");
        HIPCHECK(hipEventRecord(pl->ev_total[0], pl->stream), "timing");
        for (int t = 0; t < ntrials; ++t)
            HIPCHECK(ptk::launch_plain(pl->P, pl->d_x + (size_t)t * pl->nv, pl->d_plain, pl->d_rec, pl->d_span, pl->d_partial,
                                       pl->d_chi2 + t, pl->ch, pl->stream), "plain evaluation");
        HIPCHECK(hipEventRecord(pl->ev_total[1], pl->stream), "timing");
        HIPCHECK(hipMemcpyAsync(chi2, pl->d_chi2, sizeof(double) * ntrials, hipMemcpyDeviceToHost, pl->stream), "readback");
        HIPCHECK(hipStreamSynchronize(pl->stream), "plain evaluation");
        float ms = 0;
        (void)hipEventElapsedTime(&ms, pl->ev_total[0], pl->ev_total[1]);
        pl->last_total_ms = ms;
        pl->last_kernel_ms = 0;
        pl->last_launches = 0;
        return 0;
    }
    if (form == PT_FORM_MATRIX) return eval_matrix(pl, x, ntrials, chi2, nullptr, 0);
    if (form == PT_FORM_MATRIX_BF16) return eval_matrix(pl, x, ntrials, chi2, nullptr, 1);
    return fail(PT_ERR_STATE, "PT evaluation form not available for this plan.");
}

// ---- shared-profile matrix form ----------------------------------------------------
static int matrix_fill(pt_plan *pl, int bf16) {
    HIPCHECK(hipMemsetAsync(pl->d_phi, 0, sizeof(unsigned short) * (size_t)pl->np_pad * pl->kpad, pl->stream), "clear");
    HIPCHECK(hipMemsetAsync(pl->d_plo, 0, sizeof(unsigned short) * (size_t)pl->np_pad * pl->kpad, pl->stream), "clear");
    HIPCHECK(ptk::launch_matrix_fill(pl->P, pl->d_mrec, pl->d_mspan, pl->d_phi, pl->d_plo, pl->kpad, bf16, pl->stream), "matrix fill");
    pl->matrix_bf16 = bf16;
    return 0;
}

static int matrix_prepare(pt_plan *pl, int bf16) {
    if (pl->matrix_ready && pl->matrix_bf16 == bf16) return 0;
    if (pl->matrix_ready) {
        // same columns, other element type: rebuild the unit-height lines and refill P
        HIPCHECK(hipMemcpyAsync(pl->d_x, pl->x0.data(), sizeof(double) * pl->nv, hipMemcpyHostToDevice, pl->stream), "upload");
        HIPCHECK(ptk::launch_stage_setup(pl->P, pl->d_x, pl->d_img, 1, 0, 0, pl->stream), "setup");
        HIPCHECK(ptk::launch_matrix_columns(pl->P, pl->d_img, pl->d_mrec, pl->d_mspan, pl->d_factor, pl->stream), "matrix columns");
        int rc = matrix_fill(pl, bf16);
        if (rc) return rc;
        HIPCHECK(hipStreamSynchronize(pl->stream), "matrix build");
        return 0;
    }
    if (pl->nr > PTK_MATRIX_MAX_REFL) return fail(PT_ERR_BUDGET, "PT shared-profile form supports at most 4096 reflections.");
    if (!pl->mstream) {
        if (pl->side[0]) pl->mstream = pl->side[0];
        else if (hipStreamCreateWithFlags(&pl->mstream, hipStreamNonBlocking) == hipSuccess) pl->mstream_own = true;
        else return fail(PT_ERR_DEVICE, "PT stream creation failed.");
        if (hipEventCreateWithFlags(&pl->mjoin, hipEventDisableTiming) != hipSuccess) return fail(PT_ERR_DEVICE, "PT event creation failed.");
    }
    int rc = ensure_trials(pl, 1);
    if (rc || (rc = ensure_slots(pl, std::max(1, pl->chains)))) return rc;
    pl->kpad = (pl->nr + 31) / 32 * 32;           // 32-reflection slices
    pl->np_pad = (pl->np + 127) / 128 * 128;      // 128-point row tiles
    pl->row_tiles = pl->np_pad / 128;
    if ((rc = dalloc(pl, &pl->d_factor, (size_t)pl->nr)) || (rc = dalloc(pl, &pl->d_mrec, 2 * (size_t)pl->nr)) ||
        (rc = dalloc(pl, &pl->d_mspan, 2 * (size_t)pl->nr)) || (rc = dalloc(pl, &pl->d_phi, (size_t)pl->np_pad * pl->kpad)) ||
        (rc = dalloc(pl, &pl->d_plo, (size_t)pl->np_pad * pl->kpad)) || (rc = dalloc(pl, &pl->d_band, (size_t)pl->row_tiles)) ||
        (rc = dalloc(pl, &pl->d_yw, (size_t)pl->np_pad)))
        return rc;
    // unit-height lines of the start vector
    HIPCHECK(hipMemcpyAsync(pl->d_x, pl->x0.data(), sizeof(double) * pl->nv, hipMemcpyHostToDevice, pl->stream), "upload");
    HIPCHECK(ptk::launch_stage_setup(pl->P, pl->d_x, pl->d_img, 1, 0, 0, pl->stream), "setup");
    HIPCHECK(ptk::launch_matrix_columns(pl->P, pl->d_img, pl->d_mrec, pl->d_mspan, pl->d_factor, pl->stream), "matrix columns");
    {
        int nrep_max = 1, na = 0;
        for (int q = 0; q < pl->P.n_phases; ++q) {
            nrep_max = std::max(nrep_max, pl->P.phase[q].n_reps);
            na = std::max(na, pl->P.phase[q].atom_off + pl->P.phase[q].n_atoms);
        }
        if ((rc = dalloc(pl, &pl->d_mhr, (size_t)nrep_max * pl->kpad)) ||
            (rc = dalloc(pl, &pl->d_mf0w, (size_t)std::max(1, na) * pl->kpad)) || (rc = dalloc(pl, &pl->d_mqk, (size_t)pl->kpad)) ||
            (rc = dalloc(pl, &pl->d_mgo, 1)))
            return rc;
        HIPCHECK(ptk::launch_matrix_tables(pl->P, pl->d_img, pl->d_factor, pl->d_mhr, pl->d_mf0w, pl->d_mqk, pl->d_mgo, pl->kpad,
                                           nrep_max, pl->stream),
                 "matrix tables");
    }
    std::vector<int4> spans(2 * (size_t)pl->nr);
    HIPCHECK(hipMemcpyAsync(spans.data(), pl->d_mspan, sizeof(int4) * spans.size(), hipMemcpyDeviceToHost, pl->stream), "readback");
    if ((rc = matrix_fill(pl, bf16))) return rc;
    HIPCHECK(hipStreamSynchronize(pl->stream), "matrix build");
    // band of reflections meeting each 128-point row tile, in whole 32-reflection slices
    std::vector<int2> band(pl->row_tiles);
    for (int rt = 0; rt < pl->row_tiles; ++rt) {
        int lo_pt = 128 * rt, hi_pt = lo_pt + 128, kmin = -1, kmax = -1;
        for (int k = 0; k < pl->nr; ++k) {
            const int4 &a = spans[2 * k], &b = spans[2 * k + 1];
            bool hit = (a.y > a.x && a.x < hi_pt && a.y > lo_pt) || (b.y > b.x && b.x < hi_pt && b.y > lo_pt);
            if (hit) { if (kmin < 0) kmin = k; kmax = k; }
        }
        band[rt] = kmin < 0 ? make_int2(0, 0) : make_int2(kmin / 32 * 32, std::min(pl->kpad, (kmax + 32) / 32 * 32));
    }
    // y_obs - background of the start vector, FP64 on the host
    std::vector<float2> yw(pl->np_pad, make_float2(0.f, 0.f));
    {
        std::vector<float2> obs(pl->np);
        HIPCHECK(hipMemcpy(obs.data(), pl->P.obs, sizeof(float2) * pl->np, hipMemcpyDeviceToHost), "readback");
        const std::vector<double> bk = start_background(pl);
        for (int i = 0; i < pl->np; ++i) yw[i] = make_float2((float)((double)obs[i].x - bk[i]), obs[i].y);
    }
    HIPCHECK(hipMemcpy(pl->d_band, band.data(), sizeof(int2) * band.size(), hipMemcpyHostToDevice), "upload");
    HIPCHECK(hipMemcpy(pl->d_yw, yw.data(), sizeof(float2) * yw.size(), hipMemcpyHostToDevice), "upload");
    pl->matrix_ready = true;
    return 0;
}

static int ensure_matrix_trials(pt_plan *pl, int tpad) {
    if (tpad <= pl->cap_mtrials) return 0;
    int rc;
    // the pairs and scales twice: chunk p + 1's intensities are formed while chunk p's
    // product reads the other buffer
    if ((rc = dalloc(pl, &pl->d_ihi, 2 * (size_t)tpad * pl->kpad)) || (rc = dalloc(pl, &pl->d_ilo, 2 * (size_t)tpad * pl->kpad)) ||
        (rc = dalloc(pl, &pl->d_iscale, 2 * (size_t)tpad)) || (rc = dalloc(pl, &pl->d_mpart, (size_t)tpad * pl->row_tiles)))
        return rc;
    pl->cap_mtrials = tpad;
    return 0;
}

// trials per pass of the matrix form: large grids amortise each launch's ramp and tail
// (4096-trial passes ran 36% slower per trial); passes beyond one overlap the next pass's
// intensities with the current product
static const int kMatrixChunk = 32768;
static const int kMatrixYout = 64;      // patterns one eval_matrix call may return (d_myout rows)

// chi^2 of ntrials structure-only vectors through Y = P I; yout (optional, ntrials <= 64)
// receives Y (without background) per trial and point
static int eval_matrix(pt_plan *pl, const double *x, int ntrials, double *chi2, double *yout, int bf16) {
    if (yout && ntrials > kMatrixYout) return fail(PT_ERR_ARG, "PT shared-profile pattern: at most 64 vectors.");
    int rc = matrix_prepare(pl, bf16);
    if (rc) return rc;
    for (int t = 0; t < ntrials; ++t)
        for (int j = 0; j < pl->nv; ++j)
            if (!pl->struct_var[j] && x[(size_t)t * pl->nv + j] != pl->x0[j])
                return fail(PT_ERR_ARG, "PT shared-profile form: only structure variables (coordinates, occupancies, B, scale) "
                                        "may differ from the start vector.");
    if ((rc = ensure_trials(pl, ntrials))) return rc;
    int chunk = std::min(ntrials, kMatrixChunk);
    const int tile = ptk::kMatrixTrialTile;
    int tpad_max = (chunk + tile - 1) / tile * tile;       // whole product tiles
    if ((rc = ensure_matrix_trials(pl, tpad_max))) return rc;
    if (yout && !pl->d_myout && (rc = dalloc(pl, &pl->d_myout, (size_t)kMatrixYout * pl->np))) return rc;
    int passes = (ntrials + chunk - 1) / chunk;
    if ((rc = ensure_events(pl, 2 * passes))) return rc;
    // two streams: the intensities of chunk p (on `stream`) overlap the product and chi^2
    // of chunk p - 1 (on mstream); the pair buffers alternate between chunks
    HIPCHECK(hipMemcpyAsync(pl->d_x, x, sizeof(double) * ntrials * pl->nv, hipMemcpyHostToDevice, pl->stream), "trial upload");
    HIPCHECK(hipEventRecord(pl->ev_total[0], pl->stream), "timing");
    HIPCHECK(hipStreamWaitEvent(pl->mstream, pl->ev_total[0], 0), "fork");
    const size_t bstride = (size_t)pl->cap_mtrials * pl->kpad;
    for (int p = 0; p < passes; ++p) {
        int t0 = p * chunk, n = std::min(chunk, ntrials - t0), tpad = (n + tile - 1) / tile * tile;
        const int bsel = p & 1;
        unsigned short *ihi = pl->d_ihi + bsel * bstride, *ilo = pl->d_ilo + bsel * bstride;
        float *isc = pl->d_iscale + bsel * (size_t)pl->cap_mtrials;
        if (p >= 2) HIPCHECK(hipStreamWaitEvent(pl->stream, pl->ev[4 * (p - 2) + 3], 0), "buffer reuse");
        if (tpad > n) {
            // the intensity kernel writes every reflection slot of the n trials; the padding
            // trials of the last product tile stay zero
            const size_t off = (size_t)n * pl->kpad, cnt = sizeof(unsigned short) * (size_t)(tpad - n) * pl->kpad;
            HIPCHECK(hipMemsetAsync(ihi + off, 0, cnt, pl->stream), "clear");
            HIPCHECK(hipMemsetAsync(ilo + off, 0, cnt, pl->stream), "clear");
        }
        HIPCHECK(hipEventRecord(pl->ev[4 * p], pl->stream), "timing");
        HIPCHECK(ptk::launch_matrix_intensity(pl->P, pl->d_x + (size_t)t0 * pl->nv, pl->d_mhr, pl->d_mf0w, pl->d_mqk, ihi, ilo, isc,
                                              pl->kpad, n, bf16, pl->stream),
                 "intensities");
        HIPCHECK(hipEventRecord(pl->ev[4 * p + 1], pl->stream), "timing");
        HIPCHECK(hipStreamWaitEvent(pl->mstream, pl->ev[4 * p + 1], 0), "intensities ready");
        HIPCHECK(hipEventRecord(pl->ev[4 * p + 2], pl->mstream), "timing");
        HIPCHECK(ptk::launch_matrix_chi2(pl->P, pl->d_phi, pl->d_plo, ihi, ilo, isc, pl->d_band, pl->d_yw, pl->d_mpart, pl->kpad,
                                         tpad, pl->row_tiles, yout ? pl->d_myout : nullptr, n, bf16, pl->mstream),
                 "matrix product");
        HIPCHECK(hipEventRecord(pl->ev[4 * p + 3], pl->mstream), "timing");
        HIPCHECK(ptk::launch_matrix_reduce(pl->d_mpart, pl->row_tiles, tpad, n, pl->d_mgo, pl->d_chi2 + t0, pl->mstream), "reduce");
    }
    HIPCHECK(hipEventRecord(pl->mjoin, pl->mstream), "join");
    HIPCHECK(hipStreamWaitEvent(pl->stream, pl->mjoin, 0), "join");
    HIPCHECK(hipEventRecord(pl->ev_total[1], pl->stream), "timing");
    HIPCHECK(hipMemcpyAsync(chi2, pl->d_chi2, sizeof(double) * ntrials, hipMemcpyDeviceToHost, pl->stream), "readback");
    if (yout)
        HIPCHECK(hipMemcpyAsync(yout, pl->d_myout, sizeof(double) * (size_t)ntrials * pl->np, hipMemcpyDeviceToHost, pl->stream), "readback");
    HIPCHECK(hipStreamSynchronize(pl->stream), "matrix evaluation");
    float ms = 0;
    (void)hipEventElapsedTime(&ms, pl->ev_total[0], pl->ev_total[1]);
    pl->last_total_ms = ms;
    double ki = 0, kg = 0;
    for (int p = 0; p < passes; ++p) {
        float a = 0, b = 0;
        (void)hipEventElapsedTime(&a, pl->ev[4 * p], pl->ev[4 * p + 1]);
        (void)hipEventElapsedTime(&b, pl->ev[4 * p + 2], pl->ev[4 * p + 3]);
        ki += a; kg += b;
    }
    pl->last_kernel_ms = ki + kg;
    pl->last_launches = passes;
    pl->last_matrix_ms[0] = ki;
    pl->last_matrix_ms[1] = kg;
    return 0;
}

int pt_matrix_pattern(pt_plan *pl, const double *x, double *ycalc) {
    if (!pl || !x || !ycalc) return fail(PT_ERR_ARG, "NULL argument");
    NEED_DEVICE(pl);
    (void)hipSetDevice(pl->device);
    double chi2 = 0;
    int rc = eval_matrix(pl, x, 1, &chi2, ycalc);
    if (rc) return rc;
    // add the background of the start vector back
    const std::vector<double> bk = start_background(pl);
    for (int i = 0; i < pl->np; ++i) ycalc[i] += bk[i];
    return 0;
}

int pt_matrix_timing(const pt_plan *pl, double *intensity_ms, double *product_ms) {
    if (!pl) return fail(PT_ERR_ARG, "NULL plan");
    if (intensity_ms) *intensity_ms = pl->last_matrix_ms[0];
    if (product_ms) *product_ms = pl->last_matrix_ms[1];
    return 0;
}

// ---- the LM normal equations (FP64 derivative columns) ------------------------------
namespace {

// what the derivative column of one model parameter is (js/rietveld_core.js:1245-1298):
// a symmetric finite difference of the phase patterns (all phases, or the parameter's
// own), a Chebyshev polynomial, a one-sided spline difference, the unit-scale pattern,
// the March-Dollase difference, an atom's analytic |F|^2 gradient, or nothing
enum { COL_NONE, COL_FD_ALL, COL_FD_PHASE, COL_CHEB, COL_SPLINE, COL_SCALE, COL_PO, COL_ATOM };
struct ColSpec {
    int kind, phase, slot;
    double h, lo, hi;          // step (fdStep, :526-537) and box (defaultLimits, :486-508)
};

const char *const kInstNames[PTK_N_INST] = {"U", "V", "W", "X", "Y", "Z", "SL", "HL", "eta", "U_L", "V_L", "W_L",
                                            "X_L", "U_R", "V_R", "W_R", "X_R", "eta_split"};
const char *const kTchNames[] = {"U", "V", "W", "X", "Y", "SL", "HL"};
const char *const kSimNames[] = {"GU", "GV", "GW", "GP", "LX", "eta"};
const char *const kSplNames[] = {"GU_L", "GV_L", "GW_L", "LX_L", "GU_R", "GV_R", "GW_R", "LX_R", "eta_split"};

std::string profile_slot_name(int family, int s) {
    if (family == PT_PROFILE_SIMPLE) return s < 6 ? kSimNames[s] : "";
    if (family == PT_PROFILE_SPLIT) return s < 9 ? kSplNames[s] : "";
    return s < 7 ? kTchNames[s] : "";
}

ColSpec column_spec(const pt_plan *pl, int k, double v) {
    const double inf = INFINITY;
    ColSpec c{COL_NONE, -1, 0, 0.0, -inf, inf};
    const double h_default = std::max(1e-4 * std::fabs(v), 1e-5);      // FD_BASE_FRACTION, constants.js:411
    if (k == PT_P_ZERO) { c.kind = COL_FD_ALL; c.h = 1e-4; c.lo = -2; c.hi = 2; return c; }
    if (k == PT_P_DISP) { c.kind = COL_FD_ALL; c.h = 1e-2; return c; }
    if (k == PT_P_TRANSP) { c.kind = COL_FD_ALL; c.h = 1e-6; return c; }
    if (k >= PT_P_BKG0 && k < PT_P_BKG0 + PT_MAX_BKG) {
        const int b = k - PT_P_BKG0;
        if (b >= pl->n_bkg) return c;
        if (pl->bkg_type == PT_BKG_SPLINE) { c.kind = COL_SPLINE; c.slot = b; c.h = std::max(1e-4 * std::fabs(v), 1e-3); }
        else { c.kind = COL_CHEB; c.slot = b; }
        return c;
    }
    if (k >= PT_P_INST0 && k < PT_P_INST0 + PT_N_INST) {
        const std::string n = kInstNames[k - PT_P_INST0];
        c.kind = COL_FD_ALL; c.h = h_default;
        if (n.compare(0, 3, "eta") == 0) { c.lo = 0; c.hi = 1; }
        else if (n == "X" || n == "X_L" || n == "X_R" || n == "Y" || n == "Z" || n == "SL") c.lo = 0;
        return c;
    }
    for (int q = 0; q < pl->P.n_phases; ++q) {
        const ptk::DevPhase &ph = pl->P.phase[q];
        const int rel = k - ph.param_base;
        if (rel < 0 || rel >= PTK_PH_ATOM0 + PT_AT_STRIDE * ph.n_atoms) continue;
        c.phase = q;
        if (rel < 6) {                                   // cell
            c.kind = COL_FD_PHASE;
            if (rel < 3) { c.h = std::max(1e-6 * std::fabs(v), 1e-6); c.lo = 0.5 * v; c.hi = 2 * v; }
            else { c.h = 1e-4; c.lo = 30; c.hi = 150; }
            return c;
        }
        if (rel == PT_PH_SCALE) { c.kind = COL_SCALE; return c; }
        if (rel == PT_PH_PO) { c.kind = COL_PO; c.h = 1e-4; return c; }
        if (rel < PTK_PH_ATOM0) {                        // profile slot
            const std::string n = profile_slot_name(ph.profile_type, rel - PT_PH_PROF0);
            c.kind = COL_FD_PHASE; c.h = h_default;
            if (n.compare(0, 3, "eta") == 0) { c.lo = 0; c.hi = 1; }
            else if (n == "LX" || n == "LX_L" || n == "LX_R" || n == "X" || n == "Y" || n == "GP") c.lo = 0;
            return c;
        }
        const int a = (rel - PTK_PH_ATOM0) / PT_AT_STRIDE, comp = (rel - PTK_PH_ATOM0) % PT_AT_STRIDE;
        // a coordinate the site fixes has no gradient slot, so no column (:1264-1266)
        if (comp < 3 && !(pl->jac_atoms_h[ph.atom_off + a].free_mask & (1 << comp))) return c;
        c.kind = COL_ATOM; c.slot = PT_AT_STRIDE * a + comp;
        return c;
    }
    return c;
}

// one dataset: the packed upper triangle of [J; r] W [J; r]' into ext, and [J; r] into jr
// when it is not null ((nv + 1) rows of np)
int normal_dataset(pt_plan *pl, const double *x, std::vector<double> &ext, std::vector<double> *jr) {
    const int nv = pl->nv, np = pl->np, npar = pl->n_params;
    std::vector<double> p(pl->param_base_h);
    for (int k = 0; k < npar; ++k)
        for (int t = pl->term_off_h[k]; t < pl->term_off_h[k + 1]; ++t) p[k] += pl->term_coeff_h[t] * x[pl->term_var_h[t]];
    const int all = (1 << pl->P.n_phases) - 1;
    std::vector<double> pv(p);
    std::vector<ptk::JacEval> ev;
    ev.push_back(ptk::JacEval{0, all, ptk::JAC_NORMAL, 0, 0.0});   // 0: the phases at p
    ev.push_back(ptk::JacEval{0, 0, ptk::JAC_BKG, 0, 0.0});        // 1: the background at p
    std::vector<std::vector<std::pair<int, double>>> coef(npar);
    for (int k = 0; k < npar; ++k) {
        if (pl->term_off_h[k + 1] == pl->term_off_h[k]) continue;
        const double v = p[k];
        const ColSpec c = column_spec(pl, k, v);
        auto add_eval = [&](int row, int mask, int mode, int slot, double h) {
            ev.push_back(ptk::JacEval{row, mask, mode, slot, h});
            return (int)ev.size() - 1;
        };
        auto add_row = [&](double value) {
            const int row = (int)(pv.size() / npar);
            pv.insert(pv.end(), p.begin(), p.end());
            pv[(size_t)row * npar + k] = value;
            return row;
        };
        switch (c.kind) {
        case COL_FD_ALL:
        case COL_FD_PHASE: {
            // symmetric where the box allows (:1277-1295)
            const bool plus = v + c.h <= c.hi, minus = v - c.h >= c.lo;
            const double vp = plus ? v + c.h : v, vm = minus ? v - c.h : v, d = vp - vm;
            if (!(d != 0.0)) break;
            const int mask = c.kind == COL_FD_ALL ? all : (1 << c.phase);
            const int ep = add_eval(add_row(vp), mask, ptk::JAC_NORMAL, 0, 0.0);
            const int em = add_eval(add_row(vm), mask, ptk::JAC_NORMAL, 0, 0.0);
            coef[k] = {{ep, 1.0 / d}, {em, -1.0 / d}};
            break;
        }
        case COL_CHEB: coef[k] = {{add_eval(0, 0, ptk::JAC_CHEB, c.slot, 0.0), 1.0}}; break;
        case COL_SPLINE: coef[k] = {{add_eval(0, 0, ptk::JAC_SPLINE, c.slot, c.h), 1.0}}; break;
        case COL_SCALE: coef[k] = {{add_eval(0, 1 << c.phase, ptk::JAC_UNIT, 0, 0.0), 1.0}}; break;
        case COL_PO: coef[k] = {{add_eval(0, 1 << c.phase, ptk::JAC_PO, 0, c.h), 1.0}}; break;
        case COL_ATOM: coef[k] = {{add_eval(0, 1 << c.phase, ptk::JAC_GRAD, c.slot, 0.0), 1.0}}; break;
        default: break;
        }
    }
    const int ne = (int)ev.size();
    // the variables' columns: J_j = sum over the terms (k, j, c) of c times column k
    std::vector<double> C((size_t)nv * ne, 0.0);
    for (int k = 0; k < npar; ++k)
        for (int t = pl->term_off_h[k]; t < pl->term_off_h[k + 1]; ++t)
            for (const auto &ec : coef[k]) C[(size_t)pl->term_var_h[t] * ne + ec.first] += pl->term_coeff_h[t] * ec.second;
    int rc;
    if (!pl->d_jatoms) {
        if ((rc = upload(pl, &pl->d_jatoms, pl->jac_atoms_h.empty() ? std::vector<ptk::JacAtom>(1) : pl->jac_atoms_h)) ||
            (rc = upload(pl, &pl->d_jy, pl->yobs_h)) || (rc = upload(pl, &pl->d_jw, pl->w_h)) ||
            (rc = dalloc(pl, &pl->d_jr, (size_t)(nv + 1) * np)) || (rc = dalloc(pl, &pl->d_jext, (size_t)(nv + 1) * (nv + 2) / 2)))
            return rc;
    }
    if (ne > pl->jac_cap_evals) {
        dfree(pl, &pl->d_jp); dfree(pl, &pl->d_jev); dfree(pl, &pl->d_jlines); dfree(pl, &pl->d_jye); dfree(pl, &pl->d_jC);
        const size_t cap = (size_t)ne;
        if ((rc = dalloc(pl, &pl->d_jp, cap * npar)) || (rc = dalloc(pl, &pl->d_jev, cap)) ||
            (rc = dalloc(pl, &pl->d_jlines, cap * 2 * pl->nr)) || (rc = dalloc(pl, &pl->d_jye, cap * np)) ||
            (rc = dalloc(pl, &pl->d_jC, cap * nv)))
            return rc;
        pl->jac_cap_evals = ne;
    }
    HIPCHECK(hipMemcpyAsync(pl->d_jp, pv.data(), sizeof(double) * pv.size(), hipMemcpyHostToDevice, pl->stream), "normal equations");
    HIPCHECK(hipMemcpyAsync(pl->d_jev, ev.data(), sizeof(ptk::JacEval) * ne, hipMemcpyHostToDevice, pl->stream), "normal equations");
    HIPCHECK(hipMemcpyAsync(pl->d_jC, C.data(), sizeof(double) * C.size(), hipMemcpyHostToDevice, pl->stream), "normal equations");
    ptk::JacPlan J{};
    J.atoms = pl->d_jatoms; J.yobs = pl->d_jy; J.w = pl->d_jw; J.pvecs = pl->d_jp; J.evals = pl->d_jev;
    J.n_evals = ne; J.n_params = npar; J.lines = pl->d_jlines; J.ye = pl->d_jye; J.C = pl->d_jC; J.nv = nv;
    J.jr = pl->d_jr; J.ext = pl->d_jext;
    HIPCHECK(ptk::launch_jac_lines(pl->P, J, pl->stream), "normal equations (lines)");
    HIPCHECK(ptk::launch_jac_patterns(pl->P, J, pl->stream), "normal equations (patterns)");
    HIPCHECK(ptk::launch_jac_normal(pl->P, J, pl->stream), "normal equations (products)");
    ext.assign((size_t)(nv + 1) * (nv + 2) / 2, 0.0);
    HIPCHECK(hipMemcpyAsync(ext.data(), pl->d_jext, sizeof(double) * ext.size(), hipMemcpyDeviceToHost, pl->stream), "readback");
    if (jr) {
        jr->assign((size_t)(nv + 1) * np, 0.0);
        HIPCHECK(hipMemcpyAsync(jr->data(), pl->d_jr, sizeof(double) * jr->size(), hipMemcpyDeviceToHost, pl->stream), "readback");
    }
    HIPCHECK(hipStreamSynchronize(pl->stream), "normal equations");
    return 0;
}

}  // namespace

int pt_normal_equations(pt_plan *pl, const double *x, double *A, double *g, double *chi2, double *J) {
    if (!pl || !x || !A || !g || !chi2) return fail(PT_ERR_ARG, "NULL argument");
    NEED_DEVICE(pl);
    std::vector<pt_plan *> ds{pl};
    for (pt_plan *f : pl->follow) ds.push_back(f);
    for (pt_plan *d : ds)
        if (d->P.has_free || d->P.any_sep)
            return fail(PT_ERR_STATE, "MI355X LM normal equations take Rietveld phases without separated profiles; using the CPU path.");
    (void)hipSetDevice(pl->device);
    const int nv = pl->nv;
    std::fill(A, A + (size_t)nv * nv, 0.0);
    std::fill(g, g + nv, 0.0);
    *chi2 = 0.0;
    size_t ntot = 0;
    for (pt_plan *d : ds) ntot += d->np;
    size_t off = 0;
    float ms_sum = 0;
    for (pt_plan *d : ds) {
        std::vector<double> ext, jr;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        HIPCHECK(hipEventCreate(&e0), "timing");
        HIPCHECK(hipEventCreate(&e1), "timing");
        (void)hipEventRecord(e0, d->stream);
        int rc = normal_dataset(d, x, ext, J ? &jr : nullptr);
        (void)hipEventRecord(e1, d->stream);
        (void)hipEventSynchronize(e1);
        float ms = 0;
        (void)hipEventElapsedTime(&ms, e0, e1);
        ms_sum += ms;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        if (rc) return rc;
        // unpack: row a of the packed upper triangle starts at a (nv + 1) - a (a - 1) / 2
        const int n = nv + 1;
        auto at = [&](int a, int b) { return ext[(size_t)a * n - (size_t)a * (a - 1) / 2 + (b - a)]; };
        for (int a = 0; a < nv; ++a) {
            for (int b = a; b < nv; ++b) {
                const double v = at(a, b);
                A[(size_t)a * nv + b] += v;
                if (b != a) A[(size_t)b * nv + a] += v;
            }
            g[a] += at(a, nv);
        }
        *chi2 += at(nv, nv);
        if (J)
            for (int j = 0; j < nv; ++j)
                std::memcpy(J + (size_t)j * ntot + off, jr.data() + (size_t)j * d->np, sizeof(double) * d->np);
        off += d->np;
    }
    pl->last_normal_ms = ms_sum;
    pl->last_total_ms = ms_sum;
    return 0;
}

int pt_pattern(pt_plan *pl, const double *x, double *ycalc) {
    if (!pl || !x || !ycalc) return fail(PT_ERR_ARG, "NULL argument");
    NEED_DEVICE(pl);
    if (pl->P.has_free) return fail(PT_ERR_STATE, "PT pattern not available with free intensities.");
    (void)hipSetDevice(pl->device);
    int rc = ensure_trials(pl, 1);
    if (rc) return rc;
    if (!pl->d_y && (rc = dalloc(pl, &pl->d_y, (size_t)pl->np))) return rc;
    HIPCHECK(hipMemcpyAsync(pl->d_x, x, sizeof(double) * pl->nv, hipMemcpyHostToDevice, pl->stream), "upload");
    HIPCHECK(ptk::launch_pattern(pl->P, pl->d_x, pl->d_y, pl->block, pl->ch, pl->dyn, pl->stream), "pattern");
    HIPCHECK(hipMemcpyAsync(ycalc, pl->d_y, sizeof(double) * pl->np, hipMemcpyDeviceToHost, pl->stream), "readback");
    HIPCHECK(hipStreamSynchronize(pl->stream), "pattern");
    return 0;
}

int pt_workload(pt_plan *pl, const double *x, double *out) {
    if (!pl || !x || !out) return fail(PT_ERR_ARG, "NULL argument");
    NEED_DEVICE(pl);
    (void)hipSetDevice(pl->device);
    int rc = ensure_trials(pl, 1);
    if (rc) return rc;
    long long *d_stats = nullptr;
    if ((rc = dalloc(pl, &d_stats, 3))) return rc;
    HIPCHECK(hipMemcpyAsync(pl->d_x, x, sizeof(double) * pl->nv, hipMemcpyHostToDevice, pl->stream), "upload");
    HIPCHECK(ptk::launch_workload(pl->P, pl->d_x, d_stats, pl->block, pl->ch, pl->dyn, pl->stream), "workload");
    long long h[3] = {0, 0, 0};
    HIPCHECK(hipMemcpyAsync(h, d_stats, sizeof(h), hipMemcpyDeviceToHost, pl->stream), "readback");
    HIPCHECK(hipStreamSynchronize(pl->stream), "workload");
    int S = 0;
    for (int q = 0; q < pl->P.n_phases; ++q) S += pl->P.phase[q].n_atoms * pl->P.phase[q].n_reps_full;
    out[0] = (double)h[0]; out[1] = (double)h[1]; out[2] = (double)h[2]; out[3] = (double)S;
    out[4] = (double)pl->np; out[5] = (double)pl->P.n_bkg;
    return 0;
}

int pt_trial_counts(pt_plan *pl, double *out) {
    if (!pl || !out) return fail(PT_ERR_ARG, "NULL argument");
    NEED_DEVICE(pl);
    (void)hipSetDevice(pl->device);
    std::vector<double> st((size_t)pl->chains * pl->stride), ne(pl->chains);
    HIPCHECK(hipMemcpyAsync(st.data(), pl->P.state, sizeof(double) * st.size(), hipMemcpyDeviceToHost, pl->stream), "readback");
    HIPCHECK(hipMemcpyAsync(ne.data(), pl->P.noeval, sizeof(double) * ne.size(), hipMemcpyDeviceToHost, pl->stream), "readback");
    HIPCHECK(hipStreamSynchronize(pl->stream), "trial counts");
    double props = 0, skipped = 0;
    for (int c = 0; c < pl->chains; ++c) {
        props += st[(size_t)c * pl->stride + 3 * pl->nv + PT_S_PROPOSALS];
        skipped += ne[c];
    }
    out[0] = props - skipped;
    out[1] = skipped;
    return 0;
}

int pt_last_timing(const pt_plan *pl, double *total_ms, double *kernel_ms, int *launches) {
    if (!pl) return fail(PT_ERR_ARG, "NULL plan");
    if (total_ms) *total_ms = pl->last_total_ms;
    if (kernel_ms) *kernel_ms = pl->last_kernel_ms;
    if (launches) *launches = pl->last_launches;
    return 0;
}

// ---------------------------------------------------------------------------------
// groups: independent problems sharing every launch
// ---------------------------------------------------------------------------------
struct pt_group {
    std::vector<pt_plan *> members;
    std::vector<int64_t> offsets;          // state offsets of the members (doubles)
    std::vector<int> chain_base;           // first entry of each member in the chain table
    int device = -1, steps = 0, tpw = 2, total_chains = 0, total_pairs = 0, n_lines = 0;
    int n_accum[3] = {0, 0, 0};
    bool tuned = false;
    double tune_ms[3] = {0, 0, 0};
    double last_ms = 0;
    hipStream_t stream = nullptr;
    std::vector<void *> allocs;
    std::vector<ptk::DevMember> hmem;
    ptk::DevPlan *d_plans = nullptr;
    ptk::DevMember *d_mem = nullptr;
    int2 *d_chain_tab = nullptr, *d_pair_tab = nullptr;
    int4 *d_lines_tab = nullptr;
    int4 *d_accum_tab[3] = {nullptr, nullptr, nullptr};
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    hipEvent_t ev[2] = {nullptr, nullptr};
    // sweep parts: disjoint sets of members, each with its own work tables, part p > 0 on
    // side[p - 1] between a fork and a join (members are independent PT runs)
    struct Part {
        int2 *chain_tab = nullptr, *pair_tab = nullptr;
        int4 *lines_tab = nullptr, *accum_tab[3] = {nullptr, nullptr, nullptr};
        int n_chains = 0, n_pairs = 0, n_lines = 0, n_accum[3] = {0, 0, 0};
    };
    static constexpr int kMaxParts = 4;
    std::vector<Part> parts;
    hipStream_t side[kMaxParts - 1] = {};
    hipEvent_t fork = nullptr, join[kMaxParts - 1] = {};
};

void pt_group_destroy(pt_group *g) {
    if (!g) return;
    if (g->device >= 0) {
        (void)hipSetDevice(g->device);
        if (g->stream) (void)hipStreamSynchronize(g->stream);
        drop_graph(g->gexec, g->graph);
        for (void *p : g->allocs) (void)hipFree(p);
        for (auto e : g->ev) if (e) (void)hipEventDestroy(e);
        for (int p = 0; p + 1 < pt_group::kMaxParts; ++p) {
            if (g->side[p]) { (void)hipStreamSynchronize(g->side[p]); (void)hipStreamDestroy(g->side[p]); }
            if (g->join[p]) (void)hipEventDestroy(g->join[p]);
        }
        if (g->fork) (void)hipEventDestroy(g->fork);
        if (g->stream) (void)hipStreamDestroy(g->stream);
    }
    for (pt_plan *m : g->members) pt_plan_destroy(m);
    delete g;
}

static inline int tiles_of(int np, int tpw) {
    return (np + ptk::accum_points_per_group(tpw) - 1) / ptk::accum_points_per_group(tpw);
}

// estimated cost of one trial of member m (lines x points of the banded sum)
static double trial_cost_estimate(const pt_plan *pl) {
    const double nl = 2.0 * pl->nr;
    return nl * tiles_of(pl->np, 2) + nl * (pl->np + 100.0 * tiles_of(pl->np, 2)) / std::max(1, pl->np) * 8.0;
}

// XCD placement of the trials of the members flagged in `use` (all when empty): heaviest
// first onto the least-loaded XCD.  Returns the trials in placement order with their XCD.
struct GroupTrial { double cost; int m, c, x; };
static std::vector<GroupTrial> place_trials(const pt_group *g, const std::vector<char> &use = {}) {
    std::vector<GroupTrial> trials;
    for (int m = 0; m < (int)g->members.size(); ++m) {
        if (!use.empty() && !use[m]) continue;
        const pt_plan *pl = g->members[m];
        const double cost = trial_cost_estimate(pl);
        for (int c = 0; c < pl->chains; ++c) trials.push_back({cost, m, c, 0});
    }
    std::stable_sort(trials.begin(), trials.end(), [](const GroupTrial &a, const GroupTrial &b) { return a.cost > b.cost; });
    double load[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (GroupTrial &t : trials) {
        int x = 0;
        for (int k = 1; k < 8; ++k)
            if (load[k] < load[x]) x = k;
        load[x] += t.cost;
        t.x = x;
    }
    return trials;
}

// work table with `units(m)` units per trial: every unit of a trial in slots of its XCD
// (table slot i runs on XCD i mod 8), each XCD's list in placement order (heaviest trials
// first); unequal lists are padded with empty slots (member -1)
extern "C++" {
template <class F>
static std::vector<int4> xcd_table(const std::vector<GroupTrial> &trials, F units) {
    std::vector<std::vector<int4>> lists(8);
    for (const GroupTrial &t : trials)
        for (int u = 0; u < units(t.m); ++u) lists[t.x].push_back(make_int4(t.m, t.c, u, 0));
    size_t len = 0;
    for (auto &l : lists) len = std::max(len, l.size());
    std::vector<int4> tab(len * 8, make_int4(-1, 0, 0, 0));
    for (int x = 0; x < 8; ++x)
        for (size_t j = 0; j < lists[x].size(); ++j) tab[j * 8 + x] = lists[x][j];
    return tab;
}
}  // extern "C++"

int pt_group_create(const pt_problem *problems, int count, int steps_per_sweep, pt_group **out) {
    if (!out) return fail(PT_ERR_ARG, "output handle is NULL");
    *out = nullptr;
    if (!problems || count < 1) return fail(PT_ERR_ARG, "PT group needs at least one problem.");
    pt_group *g = new pt_group();
    int rc = 0;
    for (int m = 0; m < count && !rc; ++m) {
        pt_plan *pl = nullptr;
        rc = create_impl(&problems[m], &pl, true);
        if (!rc && pl->P.has_free) {
            pt_plan_destroy(pl);
            rc = fail(PT_ERR_ARG, "free intensities are not available in a group");
        }
        if (rc) {
            g_err = "PT group member " + std::to_string(m) + ": " + g_err;
            break;
        }
        g->members.push_back(pl);
    }
    if (rc) { pt_group_destroy(g); return rc; }
    g->device = g->members[0]->device;
    (void)hipSetDevice(g->device);
    int steps = steps_per_sweep;
    if (steps <= 0)
        for (pt_plan *pl : g->members) steps = std::max(steps, pl->steps_per_sweep);
    g->steps = steps;
    auto bail = [&](int code) { pt_group_destroy(g); return code; };
    if (hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&g->ev[0]) != hipSuccess || hipEventCreate(&g->ev[1]) != hipSuccess)
        return bail(fail(PT_ERR_DEVICE, "PT device lost during group setup."));

    // per-member device copies of the plans (common sweep length) and staged buffers
    std::vector<ptk::DevPlan> hp;
    std::vector<int2> chain_tab, pair_tab;
    std::vector<int4> lines_tab;
    int64_t off = 0;
    for (int m = 0; m < count; ++m) {
        pt_plan *pl = g->members[m];
        ptk::DevPlan P = pl->P;
        P.steps_per_sweep = steps;
        pl->steps_per_sweep = steps;
        hp.push_back(P);
        ptk::DevMember M{};
        const size_t C = (size_t)pl->chains, nl = 2 * (size_t)pl->nr;
        for (int v = 0; v < 3; ++v) M.nsplit[v] = tiles_of(pl->np, 1 << v);
        if ((rc = dalloc(g, (char **)&M.img, C * ptk::trial_image_bytes())) || (rc = dalloc(g, &M.rec, C * nl)) ||
            (rc = dalloc(g, &M.partial, C * M.nsplit[0])) || (rc = dalloc(g, &M.chi2, C)))
            return bail(rc);
        g->hmem.push_back(M);
        g->offsets.push_back(off);
        off += (int64_t)pl->chains * pl->stride;
        g->chain_base.push_back((int)chain_tab.size());
        for (int c = 0; c < pl->chains; ++c) chain_tab.push_back(make_int2(m, c));
        for (int q = 0; q < (pl->chains + 1) / 2; ++q) pair_tab.push_back(make_int2(m, q));
    }
    const std::vector<GroupTrial> placed = place_trials(g);
    lines_tab = xcd_table(placed, [&](int m) { return (g->members[m]->nr + 255) / 256; });
    g->total_chains = (int)chain_tab.size();
    g->total_pairs = (int)pair_tab.size();
    g->n_lines = (int)lines_tab.size();
    if ((rc = upload(g, &g->d_plans, hp)) || (rc = upload(g, &g->d_mem, g->hmem)) || (rc = upload(g, &g->d_chain_tab, chain_tab)) ||
        (rc = upload(g, &g->d_pair_tab, pair_tab)) || (rc = upload(g, &g->d_lines_tab, lines_tab)))
        return bail(rc);
    for (int v = 0; v < 3; ++v) {
        std::vector<int4> tab = xcd_table(placed, [&](int m) { return tiles_of(g->members[m]->np, 1 << v); });
        g->n_accum[v] = (int)tab.size();
        if ((rc = upload(g, &g->d_accum_tab[v], tab))) return bail(rc);
    }
    // sweep parts: members heaviest first onto the lightest part, as many parts as keep at
    // least 64 chains each (up to four, never more than the members)
    const int nparts = std::max(1, std::min({pt_group::kMaxParts, count, g->total_chains / 64}));
    if (nparts > 1) {
        std::vector<int> order(count);
        for (int m = 0; m < count; ++m) order[m] = m;
        auto mcost = [&](int m) { return trial_cost_estimate(g->members[m]) * g->members[m]->chains; };
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return mcost(a) > mcost(b); });
        std::vector<double> load(nparts, 0.0);
        std::vector<int> part_of(count, 0);
        for (int m : order) {
            int p = 0;
            for (int k = 1; k < nparts; ++k)
                if (load[k] < load[p]) p = k;
            load[p] += mcost(m);
            part_of[m] = p;
        }
        bool ok = hipEventCreateWithFlags(&g->fork, hipEventDisableTiming) == hipSuccess;
        for (int p = 0; ok && p + 1 < nparts; ++p)
            ok = hipStreamCreateWithFlags(&g->side[p], hipStreamNonBlocking) == hipSuccess &&
                 hipEventCreateWithFlags(&g->join[p], hipEventDisableTiming) == hipSuccess;
        if (!ok) return bail(fail(PT_ERR_DEVICE, "PT group side stream creation failed."));
        g->parts.resize(nparts);
        for (int p = 0; p < nparts; ++p) {
            pt_group::Part &P = g->parts[p];
            std::vector<char> use(count, 0);
            std::vector<int2> ct, pt;
            for (int m = 0; m < count; ++m) {
                if (part_of[m] != p) continue;
                use[m] = 1;
                for (int c = 0; c < g->members[m]->chains; ++c) ct.push_back(make_int2(m, c));
                for (int q = 0; q < (g->members[m]->chains + 1) / 2; ++q) pt.push_back(make_int2(m, q));
            }
            const std::vector<GroupTrial> pl_p = place_trials(g, use);
            std::vector<int4> lt = xcd_table(pl_p, [&](int m) { return (g->members[m]->nr + 255) / 256; });
            P.n_chains = (int)ct.size();
            P.n_pairs = (int)pt.size();
            P.n_lines = (int)lt.size();
            if ((rc = upload(g, &P.chain_tab, ct)) || (rc = upload(g, &P.pair_tab, pt)) || (rc = upload(g, &P.lines_tab, lt)))
                return bail(rc);
            for (int v = 0; v < 3; ++v) {
                std::vector<int4> tab = xcd_table(pl_p, [&](int m) { return tiles_of(g->members[m]->np, 1 << v); });
                P.n_accum[v] = (int)tab.size();
                if ((rc = upload(g, &P.accum_tab[v], tab))) return bail(rc);
            }
        }
    }
    *out = g;
    return 0;
}

#define NEED_GROUP(g) \
    do { if (!(g)) return fail(PT_ERR_ARG, "NULL group"); } while (0)

int pt_group_member(const pt_group *g, int member, pt_plan_info *info, int64_t *state_offset) {
    NEED_GROUP(g);
    if (member < 0 || member >= (int)g->members.size()) return fail(PT_ERR_ARG, "group member out of range");
    if (info) { int rc = pt_plan_info_get(g->members[member], info); if (rc) return rc; }
    if (state_offset) *state_offset = g->offsets[member];
    return 0;
}

int pt_group_set_states(pt_group *g, int member, const double *x, int nvec) {
    NEED_GROUP(g);
    if (member < 0 || member >= (int)g->members.size()) return fail(PT_ERR_ARG, "group member out of range");
    return pt_set_states(g->members[member], x, nvec);
}

static int group_ready(pt_group *g) {
    for (size_t m = 0; m < g->members.size(); ++m)
        if (!g->members[m]->states_set)
            return fail(PT_ERR_STATE, "PT group batch before initial state of member " + std::to_string(m) + ".");
    (void)hipSetDevice(g->device);
    return 0;
}

static inline int vidx(int tpw) { return tpw == 1 ? 0 : tpw == 4 ? 2 : 1; }

// one bootstrap-style pass over every chain's current vector: setup (no move), lines,
// accumulate, then the Metropolis bootstrap (pt_mode) or the chi^2 into the members' chi2
static int group_pass(pt_group *g, int pt_mode) {
    hipStream_t s = g->stream;
    HIPCHECK(ptk::launch_group_setup(g->d_plans, g->d_mem, g->d_chain_tab, g->total_chains, 1, s), "setup");
    HIPCHECK(ptk::launch_group_lines(g->d_plans, g->d_mem, g->d_lines_tab, g->n_lines, s), "lines");
    HIPCHECK(ptk::launch_group_accum(g->d_plans, g->d_mem, g->d_accum_tab[vidx(g->tpw)], g->n_accum[vidx(g->tpw)], g->tpw, s),
             "accumulate");
    HIPCHECK(ptk::launch_group_finish(g->d_plans, g->d_mem, g->d_chain_tab, g->total_chains, pt_mode, s), "finish");
    return 0;
}

// one sweep of the members behind the given work tables, on stream s
static int group_sweep_tables(pt_group *g, hipStream_t s, const int2 *chain_tab, int n_chains, const int2 *pair_tab, int n_pairs,
                              const int4 *lines_tab, int n_lines, const int4 *accum_tab, int n_accum) {
    HIPCHECK(ptk::launch_group_setup(g->d_plans, g->d_mem, chain_tab, n_chains, 0, s), "setup");
    for (int k = 0; k < g->steps; ++k) {
        HIPCHECK(ptk::launch_group_lines(g->d_plans, g->d_mem, lines_tab, n_lines, s), "lines");
        HIPCHECK(ptk::launch_group_accum(g->d_plans, g->d_mem, accum_tab, n_accum, g->tpw, s), "accumulate");
        if (k + 1 < g->steps)
            HIPCHECK(ptk::launch_group_finish_setup(g->d_plans, g->d_mem, chain_tab, n_chains, s), "finish");
        else
            HIPCHECK(ptk::launch_group_finish(g->d_plans, g->d_mem, chain_tab, n_chains, 1, s), "finish");
    }
    HIPCHECK(ptk::launch_group_exchange(g->d_plans, pair_tab, n_pairs, s), "exchange");
    return 0;
}

static int group_sweep(pt_group *g) {
    const int v = vidx(g->tpw);
    if (g->parts.size() < 2)
        return group_sweep_tables(g, g->stream, g->d_chain_tab, g->total_chains, g->d_pair_tab, g->total_pairs, g->d_lines_tab,
                                  g->n_lines, g->d_accum_tab[v], g->n_accum[v]);
    const int np = (int)g->parts.size();
    HIPCHECK(hipEventRecord(g->fork, g->stream), "fork");
    for (int p = 1; p < np; ++p) HIPCHECK(hipStreamWaitEvent(g->side[p - 1], g->fork, 0), "fork");
    for (int p = 0; p < np; ++p) {
        const pt_group::Part &P = g->parts[p];
        int rc = group_sweep_tables(g, p ? g->side[p - 1] : g->stream, P.chain_tab, P.n_chains, P.pair_tab, P.n_pairs, P.lines_tab,
                                    P.n_lines, P.accum_tab[v], P.n_accum[v]);
        if (rc) return rc;
    }
    for (int p = 1; p < np; ++p) {
        HIPCHECK(hipEventRecord(g->join[p - 1], g->side[p - 1]), "join");
        HIPCHECK(hipStreamWaitEvent(g->stream, g->join[p - 1], 0), "join");
    }
    return 0;
}

static int tune_group(pt_group *g) {
    hipStream_t s = g->stream;
    HIPCHECK(ptk::launch_group_setup(g->d_plans, g->d_mem, g->d_chain_tab, g->total_chains, 1, s), "setup");
    HIPCHECK(ptk::launch_group_lines(g->d_plans, g->d_mem, g->d_lines_tab, g->n_lines, s), "lines");
    int best = 2;
    int rc = time_tiles(s, [&](int tpw) {
        return ptk::launch_group_accum(g->d_plans, g->d_mem, g->d_accum_tab[vidx(tpw)], g->n_accum[vidx(tpw)], tpw, s);
    }, g->tune_ms, &best);
    if (rc) return rc;
    if (best != g->tpw) drop_graph(g->gexec, g->graph);
    g->tpw = best;
    g->tuned = true;
    return 0;
}

int pt_group_tiles(pt_group *g, int request, int *tiles_per_wave, double *candidate_ms) {
    NEED_GROUP(g);
    if (request == -1) {
        int rc = group_ready(g);
        if (rc || (rc = tune_group(g))) return rc;
    } else if (request == 1 || request == 2 || request == 4) {
        if (request != g->tpw) drop_graph(g->gexec, g->graph);
        g->tpw = request;
        g->tuned = true;
    } else if (request != 0) {
        return fail(PT_ERR_ARG, "tile request must be 0, -1, 1, 2 or 4");
    }
    if (tiles_per_wave) *tiles_per_wave = g->tpw;
    if (candidate_ms) for (int i = 0; i < 3; ++i) candidate_ms[i] = g->tune_ms[i];
    return 0;
}

int pt_group_costs(pt_group *g, double *chi2) {
    NEED_GROUP(g);
    if (!chi2) return fail(PT_ERR_ARG, "NULL argument");
    int rc = group_ready(g);
    if (rc) return rc;
    HIPCHECK(hipEventRecord(g->ev[0], g->stream), "timing");
    if ((rc = group_pass(g, 0))) return rc;
    HIPCHECK(hipEventRecord(g->ev[1], g->stream), "timing");
    for (size_t m = 0; m < g->members.size(); ++m)
        HIPCHECK(hipMemcpyAsync(chi2 + g->chain_base[m], g->hmem[m].chi2, sizeof(double) * g->members[m]->chains,
                                hipMemcpyDeviceToHost, g->stream), "readback");
    HIPCHECK(hipStreamSynchronize(g->stream), "group costs");
    float ms = 0;
    (void)hipEventElapsedTime(&ms, g->ev[0], g->ev[1]);
    g->last_ms = ms;
    return 0;
}

int pt_group_batch(pt_group *g, int count, int bootstrap, double *state_out) {
    NEED_GROUP(g);
    if (count < 0) return fail(PT_ERR_ARG, "negative sweep count");
    int rc = group_ready(g);
    if (rc) return rc;
    if (!bootstrap && count > 0 && !g->tuned && (rc = tune_group(g))) return rc;
    HIPCHECK(hipEventRecord(g->ev[0], g->stream), "timing");
    if (bootstrap) {
        if ((rc = group_pass(g, 1))) return rc;
    } else if (count > 0) {
        if (!g->gexec) {
            HIPCHECK(hipStreamBeginCapture(g->stream, hipStreamCaptureModeThreadLocal), "graph capture");
            rc = group_sweep(g);
            hipGraph_t gr = nullptr;
            hipError_t e = hipStreamEndCapture(g->stream, &gr);
            if (rc) return rc;
            if (e != hipSuccess) return fail(PT_ERR_DEVICE, std::string("PT graph capture failed: ") + hipGetErrorString(e));
            g->graph = gr;
            HIPCHECK(hipGraphInstantiate(&g->gexec, g->graph, nullptr, nullptr, 0), "graph instantiation");
        }
        for (int j = 0; j < count; ++j) HIPCHECK(hipGraphLaunch(g->gexec, g->stream), "sweep");
    }
    HIPCHECK(hipEventRecord(g->ev[1], g->stream), "timing");
    if (state_out)
        for (size_t m = 0; m < g->members.size(); ++m) {
            const pt_plan *pl = g->members[m];
            HIPCHECK(hipMemcpyAsync(state_out + g->offsets[m], pl->P.state, sizeof(double) * pl->chains * pl->stride,
                                    hipMemcpyDeviceToHost, g->stream), "readback");
        }
    HIPCHECK(hipStreamSynchronize(g->stream), "sweeps");
    float ms = 0;
    (void)hipEventElapsedTime(&ms, g->ev[0], g->ev[1]);
    g->last_ms = ms;
    return 0;
}

int pt_group_last_timing(const pt_group *g, double *total_ms) {
    NEED_GROUP(g);
    if (total_ms) *total_ms = g->last_ms;
    return 0;
}

int pt_hbm_copy_probe(size_t bytes, int repeats, double *gbps) {
    if (bytes < 16 || repeats < 1) return fail(PT_ERR_ARG, "copy probe needs bytes >= 16 and repeats >= 1");
    int dev = -1, rc = device_check(&dev);
    if (rc) return rc;
    bytes &= ~(size_t)15;
    void *a = nullptr, *b = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e[2] = {nullptr, nullptr};
    rc = 0;
    if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess || hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&e[0]) != hipSuccess || hipEventCreate(&e[1]) != hipSuccess || hipMemsetAsync(a, 0, bytes, st) != hipSuccess)
        rc = fail(PT_ERR_DEVICE, "copy probe buffers could not be set up.");
    float ms = 0;
    if (!rc && (ptk::launch_copy_probe(a, b, bytes, st) != hipSuccess || hipEventRecord(e[0], st) != hipSuccess))
        rc = fail(PT_ERR_DEVICE, "copy probe launch failed.");
    for (int r = 0; r < repeats && !rc; ++r)
        if (ptk::launch_copy_probe(a, b, bytes, st) != hipSuccess) rc = fail(PT_ERR_DEVICE, "copy probe launch failed.");
    if (!rc && (hipEventRecord(e[1], st) != hipSuccess || hipEventSynchronize(e[1]) != hipSuccess ||
                hipEventElapsedTime(&ms, e[0], e[1]) != hipSuccess))
        rc = fail(PT_ERR_DEVICE, "copy probe timing failed.");
    if (!rc && gbps) *gbps = 2.0 * (double)bytes * repeats / (ms * 1e-3) / 1e9;
    if (st) (void)hipStreamSynchronize(st);
    for (auto ev : e) if (ev) (void)hipEventDestroy(ev);
    if (st) (void)hipStreamDestroy(st);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    return rc;
}

}  // extern "C"
