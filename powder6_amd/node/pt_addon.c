/*
 * pt_addon.c -- Node N-API addon over the C ABI of include/powder6_pt.h.
 *
 * The thin native layer of powder6_amd/node/pt_hip.js, the RVPTGPU-shaped module that
 * rietveld_worker.js loads in place of js/pt_gpu.js.  N-API version 8 (Node >= 12).
 * Device work (costs, batch, eval) runs on the libuv thread pool through
 * napi_create_async_work and resolves a Promise, so the JS thread never blocks -- the
 * reference awaits the same calls (js/pt_gpu.js:401-403).  Errors reject with the
 * library's message (pt_last_error), which keeps the reference's wording.
 */
#define NAPI_VERSION 8
#include <node_api.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/powder6_pt.h"

#define CHECK(env, call)                                                    \
    do {                                                                    \
        if ((call) != napi_ok) {                                            \
            napi_throw_error((env), NULL, "N-API call failed: " #call);    \
            return NULL;                                                    \
        }                                                                   \
    } while (0)

/* ---- problem conversion ------------------------------------------------------- */
typedef struct {
    pt_problem prob;
    pt_phase *phases;
    pt_atom *atoms;
    void **owned;
    int n_owned, cap_owned;
} holder;

static void own(holder *h, void *p) {
    if (h->n_owned == h->cap_owned) {
        h->cap_owned = h->cap_owned ? 2 * h->cap_owned : 32;
        h->owned = (void **)realloc(h->owned, sizeof(void *) * h->cap_owned);
    }
    h->owned[h->n_owned++] = p;
}
static void holder_free(holder *h) {
    for (int i = 0; i < h->n_owned; ++i) free(h->owned[i]);
    free(h->owned);
    free(h->phases);
    free(h->atoms);
}

static int get_prop(napi_env env, napi_value obj, const char *name, napi_value *out) {
    bool has = false;
    if (napi_has_named_property(env, obj, name, &has) != napi_ok || !has) return 0;
    return napi_get_named_property(env, obj, name, out) == napi_ok;
}
static double num(napi_env env, napi_value obj, const char *name, double dflt) {
    napi_value v;
    double d = dflt;
    if (get_prop(env, obj, name, &v)) napi_get_value_double(env, v, &d);
    return d;
}

/* typed array or plain array -> freshly allocated doubles / ints; returns the length */
static int to_doubles(napi_env env, napi_value v, double **out, holder *h) {
    bool is_ta = false, is_arr = false;
    napi_is_typedarray(env, v, &is_ta);
    napi_is_array(env, v, &is_arr);
    uint32_t n = 0;
    if (is_ta) {
        napi_typedarray_type t;
        size_t len;
        void *data;
        napi_get_typedarray_info(env, v, &t, &len, &data, NULL, NULL);
        n = (uint32_t)len;
        double *d = (double *)malloc(sizeof(double) * (n ? n : 1));
        for (uint32_t i = 0; i < n; ++i) {
            switch (t) {
            case napi_float64_array: d[i] = ((double *)data)[i]; break;
            case napi_float32_array: d[i] = ((float *)data)[i]; break;
            case napi_int32_array: d[i] = ((int32_t *)data)[i]; break;
            case napi_int16_array: d[i] = ((int16_t *)data)[i]; break;
            default: d[i] = 0;
            }
        }
        own(h, d);
        *out = d;
        return (int)n;
    }
    if (is_arr) {
        napi_get_array_length(env, v, &n);
        double *d = (double *)malloc(sizeof(double) * (n ? n : 1));
        for (uint32_t i = 0; i < n; ++i) {
            napi_value e;
            napi_get_element(env, v, i, &e);
            d[i] = 0;
            napi_get_value_double(env, e, &d[i]);
        }
        own(h, d);
        *out = d;
        return (int)n;
    }
    *out = NULL;
    return -1;
}
static int to_ints(napi_env env, napi_value v, int32_t **out, holder *h) {
    double *d;
    int n = to_doubles(env, v, &d, h);
    if (n < 0) { *out = NULL; return n; }
    int32_t *r = (int32_t *)malloc(sizeof(int32_t) * (n ? n : 1));
    for (int i = 0; i < n; ++i) r[i] = (int32_t)d[i];
    own(h, r);
    *out = r;
    return n;
}
static int field_doubles(napi_env env, napi_value obj, const char *name, double **out, holder *h) {
    napi_value v;
    if (!get_prop(env, obj, name, &v)) { *out = NULL; return -1; }
    return to_doubles(env, v, out, h);
}
static int field_ints(napi_env env, napi_value obj, const char *name, int32_t **out, holder *h) {
    napi_value v;
    if (!get_prop(env, obj, name, &v)) { *out = NULL; return -1; }
    return to_ints(env, v, out, h);
}

/* the flat problem of pt_hip.js (field names mirror pt_problem) -> pt_problem */
static const char *convert(napi_env env, napi_value o, holder *h) {
    memset(h, 0, sizeof *h);
    pt_problem *p = &h->prob;
    double *d;
    p->n_points = field_doubles(env, o, "tth", &d, h); p->tth = d;
    if (field_doubles(env, o, "y", &d, h) != p->n_points) return "problem.y must match problem.tth";
    p->y_obs = d;
    if (field_doubles(env, o, "w", &d, h) != p->n_points) return "problem.w must match problem.tth";
    p->weight = d;
    p->lambda1 = num(env, o, "lambda1", 1.54056);
    p->lambda2 = num(env, o, "lambda2", 0);
    p->ratio = num(env, o, "ratio", 0);
    p->pol_k = num(env, o, "polK", 1);
    p->min_fwhm = num(env, o, "minFwhm", 1e-3);
    p->radiation = (int32_t)num(env, o, "radiation", 0);
    p->n_bkg = (int32_t)num(env, o, "nBkg", 0);
    p->bkg_type = (int32_t)num(env, o, "bkgType", PT_BKG_CHEBYSHEV);
    p->resolution_type = (int32_t)num(env, o, "resolutionType", 0);
    if (p->bkg_type == PT_BKG_SPLINE) {
        if (field_doubles(env, o, "bkgX", &d, h) != p->n_bkg) return "problem.bkgX must hold nBkg anchors";
        p->bkg_x = d;
    }
    napi_value phases;
    if (!get_prop(env, o, "phases", &phases)) return "problem.phases is missing";
    uint32_t nph = 0;
    napi_get_array_length(env, phases, &nph);
    p->n_phases = (int32_t)nph;
    h->phases = (pt_phase *)calloc(nph ? nph : 1, sizeof(pt_phase));
    int total_atoms = 0;
    for (uint32_t q = 0; q < nph; ++q) {
        napi_value ph, atoms;
        napi_get_element(env, phases, q, &ph);
        if (get_prop(env, ph, "atoms", &atoms)) { uint32_t na = 0; napi_get_array_length(env, atoms, &na); total_atoms += (int)na; }
    }
    h->atoms = (pt_atom *)calloc(total_atoms ? total_atoms : 1, sizeof(pt_atom));
    int ao = 0;
    for (uint32_t q = 0; q < nph; ++q) {
        napi_value ph, atoms;
        napi_get_element(env, phases, q, &ph);
        pt_phase *P = &h->phases[q];
        P->system = (int32_t)num(env, ph, "system", 0);
        P->unique_axis = (int32_t)num(env, ph, "uniqueAxis", 1);
        P->profile_type = (int32_t)num(env, ph, "profileType", PT_PROFILE_TCH);
        int32_t *r;
        int nr9 = field_ints(env, ph, "opR", &r, h);
        double *t;
        int nt3 = field_doubles(env, ph, "opT", &t, h);
        if (nr9 < 9 || nr9 % 9 || nt3 != nr9 / 3) return "phase.opR / phase.opT are inconsistent";
        P->n_ops = nr9 / 9; P->op_r = r; P->op_t = t;
        double *pd;
        if (field_doubles(env, ph, "poDir", &pd, h) == 3) { P->po_dir[0] = pd[0]; P->po_dir[1] = pd[1]; P->po_dir[2] = pd[2]; }
        else { P->po_dir[2] = 1; }
        P->separated = (int32_t)num(env, ph, "separated", 0);
        P->free_intensities = (int32_t)num(env, ph, "freeIntensities", 0);
        P->size_model = (int32_t)num(env, ph, "sizeModel", PT_SIZE_ISOTROPIC);
        P->strain_model = (int32_t)num(env, ph, "strainModel", PT_STRAIN_ISOTROPIC);
        double *ax;
        if (field_doubles(env, ph, "sizeAxis", &ax, h) == 3) { P->size_axis[0] = ax[0]; P->size_axis[1] = ax[1]; P->size_axis[2] = ax[2]; }
        else { P->size_axis[2] = 1; }
        if (field_doubles(env, ph, "strainAxis", &ax, h) == 3) { P->strain_axis[0] = ax[0]; P->strain_axis[1] = ax[1]; P->strain_axis[2] = ax[2]; }
        else { P->strain_axis[2] = 1; }
        P->atoms = h->atoms + ao;
        if (get_prop(env, ph, "atoms", &atoms)) {
            uint32_t na = 0;
            napi_get_array_length(env, atoms, &na);
            P->n_atoms = (int32_t)na;
            for (uint32_t j = 0; j < na; ++j) {
                napi_value at;
                napi_get_element(env, atoms, j, &at);
                pt_atom *A = &h->atoms[ao + j];
                double *v;
                if (field_doubles(env, at, "r0", &v, h) == 3) memcpy(A->r0, v, sizeof A->r0);
                if (field_doubles(env, at, "basis", &v, h) == 9) memcpy(A->basis, v, sizeof A->basis);
                A->free_mask = (int32_t)num(env, at, "freeMask", 0);
                A->mult = num(env, at, "mult", 1);
                int ng = field_doubles(env, at, "ffA", &v, h);
                if (ng > PT_MAX_GAUSS) return "a scattering factor has more than 5 Gaussians";
                A->n_gauss = ng > 0 ? ng : 0;
                for (int g = 0; g < A->n_gauss; ++g) A->ff_a[g] = v[g];
                if (field_doubles(env, at, "ffB", &v, h) != A->n_gauss && A->n_gauss) return "atom.ffB must match atom.ffA";
                for (int g = 0; g < A->n_gauss; ++g) A->ff_b[g] = v[g];
                A->ff_c = num(env, at, "ffC", 0);
                A->b_neutron = num(env, at, "bNeutron", 0);
                A->fp = num(env, at, "fp", 0);
                A->fpp = num(env, at, "fpp", 0);
            }
            ao += (int)na;
        }
    }
    p->phases = h->phases;
    p->n_params = field_doubles(env, o, "paramBase", &d, h); p->param_base = d;
    p->n_vars = field_doubles(env, o, "x0", &d, h); p->x0 = d;
    if (field_doubles(env, o, "lo", &d, h) != p->n_vars) return "problem.lo must match problem.x0";
    p->lo = d;
    if (field_doubles(env, o, "hi", &d, h) != p->n_vars) return "problem.hi must match problem.x0";
    p->hi = d;
    if (field_doubles(env, o, "scale", &d, h) != p->n_vars) return "problem.scale must match problem.x0";
    p->scale = d;
    int32_t *ti;
    p->n_terms = field_ints(env, o, "termParam", &ti, h); p->term_param = ti;
    if (p->n_terms < 0) p->n_terms = 0;
    if (field_ints(env, o, "termVar", &ti, h) != p->n_terms) return "problem.termVar must match termParam";
    p->term_var = ti;
    if (field_doubles(env, o, "termCoeff", &d, h) != p->n_terms) return "problem.termCoeff must match termParam";
    p->term_coeff = d;
    p->chains = (int32_t)num(env, o, "chains", 6);
    p->sweeps = (int32_t)num(env, o, "sweeps", 50);
    p->steps_per_sweep = (int32_t)num(env, o, "stepsPerSweep", 0);
    p->max_temp = num(env, o, "maxTemp", 100);
    p->seed = (uint32_t)num(env, o, "seed", 12345);
    return NULL;
}

/* ---- plan handles ----------------------------------------------------------------- */
typedef struct {
    pt_plan *plan;
    pt_plan_info info;
    int busy;
} handle;

static void finalize_handle(napi_env env, void *data, void *hint) {
    (void)env; (void)hint;
    handle *h = (handle *)data;
    if (h->plan) pt_plan_destroy(h->plan);
    free(h);
}

static handle *get_handle(napi_env env, napi_value v) {
    handle *h = NULL;
    if (napi_get_value_external(env, v, (void **)&h) != napi_ok) return NULL;
    return h;
}

static napi_value info_object(napi_env env, const pt_plan_info *in) {
    napi_value o, v;
    napi_create_object(env, &o);
#define SET(name, val) napi_create_int32(env, (val), &v); napi_set_named_property(env, o, name, v)
    SET("nv", in->nv); SET("chains", in->chains); SET("stride", in->stride); SET("pointCount", in->point_count);
    SET("reflectionCount", in->reflection_count); SET("lineCount", in->line_count);
    SET("stepsPerSweep", in->steps_per_sweep); SET("paramCount", in->n_params);
    napi_get_boolean(env, in->uniform_grid != 0, &v); napi_set_named_property(env, o, "uniformGrid", v);
    napi_get_boolean(env, in->device_ok != 0, &v); napi_set_named_property(env, o, "device", v);
#undef SET
    return o;
}

/* planCreate(problem, withDevice) -> {handle, info}; throws with the library message.
 * An array of problems (one per dataset, one variable list) makes a multi-dataset plan
 * (pt_plan_create_multi; device only). */
static napi_value plan_create(napi_env env, napi_callback_info cbi) {
    size_t argc = 2;
    napi_value argv[2];
    CHECK(env, napi_get_cb_info(env, cbi, &argc, argv, NULL, NULL));
    if (argc < 1) { napi_throw_type_error(env, NULL, "planCreate(problem, withDevice)"); return NULL; }
    bool dev = true;
    if (argc > 1) napi_get_value_bool(env, argv[1], &dev);
    bool is_array = false;
    napi_is_array(env, argv[0], &is_array);
    pt_plan *plan = NULL;
    int rc;
    if (is_array) {
        uint32_t n = 0;
        napi_get_array_length(env, argv[0], &n);
        if (n < 1 || n > 8) { napi_throw_error(env, NULL, "PT takes 1 to 8 datasets per plan."); return NULL; }
        if (!dev) { napi_throw_error(env, NULL, "a multi-dataset plan needs the device"); return NULL; }
        holder hs[8];
        pt_problem probs[8];
        const char *err = NULL;
        uint32_t made = 0;
        for (; made < n && !err; ++made) {
            napi_value e;
            napi_get_element(env, argv[0], made, &e);
            err = convert(env, e, &hs[made]);
            probs[made] = hs[made].prob;
        }
        rc = err ? 0 : pt_plan_create_multi(probs, (int)n, &plan);
        for (uint32_t i = 0; i < made; ++i) holder_free(&hs[i]);
        if (err) { napi_throw_error(env, NULL, err); return NULL; }
    } else {
        holder h;
        const char *err = convert(env, argv[0], &h);
        if (err) { holder_free(&h); napi_throw_error(env, NULL, err); return NULL; }
        rc = dev ? pt_plan_create(&h.prob, &plan) : pt_plan_compile(&h.prob, &plan);
        holder_free(&h);
    }
    if (rc) { napi_throw_error(env, NULL, pt_last_error()); return NULL; }
    handle *hd = (handle *)calloc(1, sizeof(handle));
    hd->plan = plan;
    pt_plan_info_get(plan, &hd->info);
    napi_value ext, out;
    CHECK(env, napi_create_external(env, hd, finalize_handle, NULL, &ext));
    CHECK(env, napi_create_object(env, &out));
    CHECK(env, napi_set_named_property(env, out, "handle", ext));
    CHECK(env, napi_set_named_property(env, out, "info", info_object(env, &hd->info)));
    return out;
}

/* planDestroy(handle): releases the device memory now (the finalizer also would) */
static napi_value plan_destroy(napi_env env, napi_callback_info cbi) {
    size_t argc = 1;
    napi_value argv[1];
    CHECK(env, napi_get_cb_info(env, cbi, &argc, argv, NULL, NULL));
    handle *h = argc ? get_handle(env, argv[0]) : NULL;
    if (h && h->plan && !h->busy) { pt_plan_destroy(h->plan); h->plan = NULL; }
    return NULL;
}

/* reflections(handle, phase) -> Int32Array of h,k,l,m rows */
static napi_value reflections(napi_env env, napi_callback_info cbi) {
    size_t argc = 2;
    napi_value argv[2];
    CHECK(env, napi_get_cb_info(env, cbi, &argc, argv, NULL, NULL));
    handle *h = get_handle(env, argv[0]);
    int32_t phase = 0;
    if (argc > 1) napi_get_value_int32(env, argv[1], &phase);
    if (!h || !h->plan) { napi_throw_error(env, NULL, "PT plan is closed."); return NULL; }
    int n = pt_plan_reflections(h->plan, phase, NULL, 0);
    if (n < 0) { napi_throw_error(env, NULL, pt_last_error()); return NULL; }
    napi_value ab, ta;
    void *data;
    CHECK(env, napi_create_arraybuffer(env, sizeof(int32_t) * 4 * (size_t)n, &data, &ab));
    pt_plan_reflections(h->plan, phase, (int32_t *)data, n);
    CHECK(env, napi_create_typedarray(env, napi_int32_array, 4 * (size_t)n, ab, 0, &ta));
    return ta;
}

/* ---- asynchronous device calls ----------------------------------------------------- */
enum { OP_COSTS = 1, OP_BATCH = 2, OP_EVAL = 3, OP_INITIAL = 4, OP_PATTERN = 5, OP_NORMAL = 6 };
typedef struct {
    napi_async_work work;
    napi_deferred deferred;
    handle *h;
    napi_ref href;
    int op, count, bootstrap, form, nvec;
    double *in;
    double *out;
    size_t n_out;
    int rc;
    char err[512];
} job;

static void execute(napi_env env, void *data) {
    (void)env;
    job *j = (job *)data;
    pt_plan *p = j->h->plan;
    switch (j->op) {
    case OP_COSTS: j->rc = pt_costs(p, j->in, j->nvec, j->out); break;
    case OP_BATCH: j->rc = pt_batch(p, j->count, j->bootstrap, j->out); break;
    case OP_EVAL: j->rc = pt_eval(p, j->form, j->in, j->nvec, j->out); break;
    case OP_INITIAL: j->rc = pt_set_states(p, j->in, j->nvec); break;
    case OP_PATTERN: j->rc = pt_pattern(p, j->in, j->out); break;
    case OP_NORMAL: {
        const int nv = j->h->info.nv;
        j->rc = pt_normal_equations(p, j->in, j->out, j->out + (size_t)nv * nv, j->out + (size_t)nv * nv + nv, NULL);
        break;
    }
    }
    if (j->rc) {
        strncpy(j->err, pt_last_error(), sizeof j->err - 1);
        j->err[sizeof j->err - 1] = 0;
    }
}

static void complete(napi_env env, napi_status status, void *data) {
    job *j = (job *)data;
    j->h->busy = 0;
    if (status != napi_ok || j->rc) {
        napi_value msg, errv;
        napi_create_string_utf8(env, j->rc ? j->err : "PT device call was cancelled.", NAPI_AUTO_LENGTH, &msg);
        napi_create_error(env, NULL, msg, &errv);
        napi_reject_deferred(env, j->deferred, errv);
    } else if (j->out) {
        napi_value ab, ta;
        void *buf;
        napi_create_arraybuffer(env, sizeof(double) * j->n_out, &buf, &ab);
        memcpy(buf, j->out, sizeof(double) * j->n_out);
        napi_create_typedarray(env, napi_float64_array, j->n_out, ab, 0, &ta);
        napi_resolve_deferred(env, j->deferred, ta);
    } else {
        napi_value u;
        napi_get_undefined(env, &u);
        napi_resolve_deferred(env, j->deferred, u);
    }
    napi_delete_async_work(env, j->work);
    napi_delete_reference(env, j->href);
    free(j->in);
    free(j->out);
    free(j);
}

static napi_value start(napi_env env, napi_value hval, job *j) {
    napi_value promise, name;
    handle *h = j->h;
    if (!h || !h->plan) { free(j->in); free(j->out); free(j); napi_throw_error(env, NULL, "PT plan is closed."); return NULL; }
    if (h->busy) { free(j->in); free(j->out); free(j); napi_throw_error(env, NULL, "PT plan is busy with another call."); return NULL; }
    h->busy = 1;
    napi_create_reference(env, hval, 1, &j->href);
    CHECK(env, napi_create_promise(env, &j->deferred, &promise));
    CHECK(env, napi_create_string_utf8(env, "powder6_pt", NAPI_AUTO_LENGTH, &name));
    CHECK(env, napi_create_async_work(env, NULL, name, execute, complete, j, &j->work));
    CHECK(env, napi_queue_async_work(env, j->work));
    return promise;
}

static double *copy_vectors(napi_env env, napi_value v, int nv, int *nvec) {
    holder tmp;
    memset(&tmp, 0, sizeof tmp);
    double *d;
    int n = to_doubles(env, v, &d, &tmp);
    double *out = NULL;
    if (n > 0 && nv > 0 && n % nv == 0) {
        out = (double *)malloc(sizeof(double) * n);
        memcpy(out, d, sizeof(double) * n);
        *nvec = n / nv;
    }
    holder_free(&tmp);
    return out;
}

/* costs(handle, vectors) -> Promise<Float64Array(chains)>; vectors: nvec*nv doubles */
static napi_value costs(napi_env env, napi_callback_info cbi) {
    size_t argc = 2;
    napi_value argv[2];
    CHECK(env, napi_get_cb_info(env, cbi, &argc, argv, NULL, NULL));
    job *j = (job *)calloc(1, sizeof(job));
    j->h = get_handle(env, argv[0]);
    j->op = OP_COSTS;
    if (j->h) {
        j->in = copy_vectors(env, argv[1], j->h->info.nv, &j->nvec);
        j->n_out = (size_t)j->h->info.chains;
        j->out = (double *)calloc(j->n_out, sizeof(double));
    }
    if (!j->h || !j->in) { free(j->out); free(j); napi_throw_error(env, NULL, "costs(handle, vectors): vectors of nv doubles"); return NULL; }
    return start(env, argv[0], j);
}

/* initial(handle, vectors) -> Promise<undefined> */
static napi_value initial(napi_env env, napi_callback_info cbi) {
    size_t argc = 2;
    napi_value argv[2];
    CHECK(env, napi_get_cb_info(env, cbi, &argc, argv, NULL, NULL));
    job *j = (job *)calloc(1, sizeof(job));
    j->h = get_handle(env, argv[0]);
    j->op = OP_INITIAL;
    if (j->h) j->in = copy_vectors(env, argv[1], j->h->info.nv, &j->nvec);
    if (!j->h || !j->in) { free(j); napi_throw_error(env, NULL, "initial(handle, vectors): vectors of nv doubles"); return NULL; }
    return start(env, argv[0], j);
}

/* batch(handle, count, bootstrap) -> Promise<Float64Array(chains*stride)> */
static napi_value batch(napi_env env, napi_callback_info cbi) {
    size_t argc = 3;
    napi_value argv[3];
    CHECK(env, napi_get_cb_info(env, cbi, &argc, argv, NULL, NULL));
    job *j = (job *)calloc(1, sizeof(job));
    j->h = get_handle(env, argv[0]);
    j->op = OP_BATCH;
    int32_t count = 1;
    bool boot = false;
    if (argc > 1) napi_get_value_int32(env, argv[1], &count);
    if (argc > 2) napi_get_value_bool(env, argv[2], &boot);
    j->count = count;
    j->bootstrap = boot ? 1 : 0;
    if (!j->h) { free(j); napi_throw_error(env, NULL, "batch(handle, count, bootstrap)"); return NULL; }
    j->n_out = (size_t)j->h->info.chains * j->h->info.stride;
    j->out = (double *)calloc(j->n_out, sizeof(double));
    return start(env, argv[0], j);
}

/* evalTrials(handle, form, vectors) -> Promise<Float64Array(ntrials)> */
static napi_value eval_trials(napi_env env, napi_callback_info cbi) {
    size_t argc = 3;
    napi_value argv[3];
    CHECK(env, napi_get_cb_info(env, cbi, &argc, argv, NULL, NULL));
    job *j = (job *)calloc(1, sizeof(job));
    j->h = get_handle(env, argv[0]);
    j->op = OP_EVAL;
    int32_t form = PT_FORM_BATCHED;
    napi_get_value_int32(env, argv[1], &form);
    j->form = form;
    if (j->h) j->in = copy_vectors(env, argv[2], j->h->info.nv, &j->nvec);
    if (!j->h || !j->in) { free(j); napi_throw_error(env, NULL, "evalTrials(handle, form, vectors)"); return NULL; }
    j->n_out = (size_t)j->nvec;
    j->out = (double *)calloc(j->n_out, sizeof(double));
    return start(env, argv[0], j);
}

/* pattern(handle, vector) -> Promise<Float64Array(points)> */
static napi_value pattern(napi_env env, napi_callback_info cbi) {
    size_t argc = 2;
    napi_value argv[2];
    CHECK(env, napi_get_cb_info(env, cbi, &argc, argv, NULL, NULL));
    job *j = (job *)calloc(1, sizeof(job));
    j->h = get_handle(env, argv[0]);
    j->op = OP_PATTERN;
    if (j->h) j->in = copy_vectors(env, argv[1], j->h->info.nv, &j->nvec);
    if (!j->h || !j->in || j->nvec != 1) { free(j->in); free(j); napi_throw_error(env, NULL, "pattern(handle, vector)"); return NULL; }
    j->n_out = (size_t)j->h->info.point_count;
    j->out = (double *)calloc(j->n_out, sizeof(double));
    return start(env, argv[0], j);
}

/* normal(handle, vector) -> Promise<Float64Array(nv*nv + nv + 1)>: A (row major), g, chi^2
 * of one LM step (pt_normal_equations) */
static napi_value normal(napi_env env, napi_callback_info cbi) {
    size_t argc = 2;
    napi_value argv[2];
    CHECK(env, napi_get_cb_info(env, cbi, &argc, argv, NULL, NULL));
    job *j = (job *)calloc(1, sizeof(job));
    j->h = get_handle(env, argv[0]);
    j->op = OP_NORMAL;
    if (j->h) j->in = copy_vectors(env, argv[1], j->h->info.nv, &j->nvec);
    if (!j->h || !j->in || j->nvec != 1) { free(j->in); free(j); napi_throw_error(env, NULL, "normal(handle, vector)"); return NULL; }
    j->n_out = (size_t)j->h->info.nv * j->h->info.nv + j->h->info.nv + 1;
    j->out = (double *)calloc(j->n_out, sizeof(double));
    return start(env, argv[0], j);
}

static napi_value version(napi_env env, napi_callback_info cbi) {
    (void)cbi;
    napi_value v;
    napi_create_int32(env, pt_abi_version(), &v);
    return v;
}

static napi_value init(napi_env env, napi_value exports) {
    napi_property_descriptor d[] = {
        {"planCreate", NULL, plan_create, NULL, NULL, NULL, napi_enumerable, NULL},
        {"planDestroy", NULL, plan_destroy, NULL, NULL, NULL, napi_enumerable, NULL},
        {"reflections", NULL, reflections, NULL, NULL, NULL, napi_enumerable, NULL},
        {"initial", NULL, initial, NULL, NULL, NULL, napi_enumerable, NULL},
        {"costs", NULL, costs, NULL, NULL, NULL, napi_enumerable, NULL},
        {"batch", NULL, batch, NULL, NULL, NULL, napi_enumerable, NULL},
        {"evalTrials", NULL, eval_trials, NULL, NULL, NULL, napi_enumerable, NULL},
        {"pattern", NULL, pattern, NULL, NULL, NULL, napi_enumerable, NULL},
        {"normal", NULL, normal, NULL, NULL, NULL, napi_enumerable, NULL},
        {"abiVersion", NULL, version, NULL, NULL, NULL, napi_enumerable, NULL},
    };
    napi_define_properties(env, exports, sizeof d / sizeof d[0], d);
    return exports;
}

NAPI_MODULE(NODE_GYP_MODULE_NAME, init)
