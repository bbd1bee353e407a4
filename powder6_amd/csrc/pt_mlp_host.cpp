// pt_mlp_host.cpp -- host side of the surrogate network: buffers, the forward / backward
// / Adam launch sequences, and the pt_mlp_* calls of include/powder6_pt.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <string>

#include "../../include/powder6_pt.h"
#include "pt_common.h"
#include "pt_mlp_host.h"

using ptk_host::fail;
using namespace ptk;

#define MLPCHECK(expr, what)                                                                        \
    do {                                                                                            \
        hipError_t e_ = (expr);                                                                     \
        if (e_ != hipSuccess) return fail(PT_ERR_DEVICE, std::string("PT surrogate device error during ") + what + ": " + hipGetErrorString(e_)); \
    } while (0)

namespace {

template <class T>
int nalloc(pt_mlp *m, T **ptr, size_t count) {
    void *p = nullptr;
    if (hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T)) != hipSuccess)
        return fail(PT_ERR_DEVICE, "PT surrogate buffers exceed this GPU's memory (hipMalloc failed).");
    m->allocs.push_back(p);
    *ptr = reinterpret_cast<T *>(p);
    return 0;
}

// the fused kernel's launch for `rows` rows: given rows (x16, t, mask) or, with x16 null,
// a minibatch drawn from a replay ring (filled in by the caller)
MlpRun run_of(pt_mlp *m, int rows) {
    MlpRun R{};
    R.rows = rows;
    R.inv_n = 1.0f / (float)std::max(1, rows);
    R.wb = m->wb;
    R.wbt = m->wbt;
    R.p = m->p;
    R.z = m->zf;
    R.part = m->partf;
    R.iter = m->iter;
    R.active = m->active;
    return R;
}

// loss and gradient (into m->g, m->loss) of `rows` given rows in m->x16, m->t, m->mask
int gradient_rows(pt_mlp *m, int rows, hipStream_t s) {
    MlpRun R = run_of(m, rows);
    R.x16 = m->x16;
    R.t = m->t;
    R.mask = m->mask;
    MLPCHECK(mlp_fused(R, true, s), "gradient");
    MLPCHECK(mlp_reduce_adam(m->partf, rows, R.inv_n, m->g, m->loss, 0, m->p, m->m, m->v, m->wb, m->wbt, m->lr, m->beta1, m->beta2,
                             m->eps, m->iter, m->active, s), "gradient sum");
    return 0;
}

int stage(pt_mlp *m, const float *host, size_t count) {
    if (count > m->staging_cap) {
        size_t cap = std::max(count, (size_t)1 << 16);
        int rc = nalloc(m, &m->staging, cap);
        if (rc) return rc;
        m->staging_cap = cap;
    }
    MLPCHECK(hipMemcpyAsync(m->staging, host, sizeof(float) * count, hipMemcpyHostToDevice, m->stream), "upload");
    return 0;
}

}  // namespace

namespace mlpx {

int create(uint32_t seed, pt_mlp **out) {
    *out = nullptr;
    pt_mlp *m = new pt_mlp();
    (void)hipGetDevice(&m->device);
    int rc = 0;
    if (hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess) {
        delete m;
        return fail(PT_ERR_DEVICE, "PT surrogate stream creation failed.");
    }
    if ((rc = nalloc(m, &m->p, MLP_PARAMS)) || (rc = nalloc(m, &m->g, MLP_PARAMS)) || (rc = nalloc(m, &m->m, MLP_PARAMS)) ||
        (rc = nalloc(m, &m->v, MLP_PARAMS)) || (rc = nalloc(m, &m->wb, MLP_PARAMS)) || (rc = nalloc(m, &m->wbt, MLP_PARAMS)) ||
        (rc = nalloc(m, &m->iter, 1)) ||
        (rc = nalloc(m, &m->active, 1)) || (rc = nalloc(m, &m->loss, 2))) {
        destroy(m);
        return rc;
    }
    std::vector<float> p(MLP_PARAMS, 0.0f);
    std::mt19937_64 gen(seed);
    auto fill = [&](int off, int fan_in, int count, double gain) {
        std::normal_distribution<double> nd(0.0, gain / std::sqrt((double)fan_in));
        for (int i = 0; i < count; ++i) p[off + i] = (float)nd(gen);
    };
    fill(MLP_OFF_W1, MLP_IN, MLP_IN * MLP_H, std::sqrt(2.0));
    fill(MLP_OFF_W2, MLP_H, MLP_H * MLP_H, std::sqrt(2.0));
    fill(MLP_OFF_W3, MLP_H, MLP_H * MLP_H, std::sqrt(2.0));
    fill(MLP_OFF_W4, MLP_H, MLP_H, 1.0);
    int zero = 1;
    if (hipMemcpy(m->p, p.data(), sizeof(float) * MLP_PARAMS, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(m->m, 0, sizeof(float) * MLP_PARAMS) != hipSuccess || hipMemset(m->v, 0, sizeof(float) * MLP_PARAMS) != hipSuccess ||
        hipMemset(m->g, 0, sizeof(float) * MLP_PARAMS) != hipSuccess || hipMemset(m->iter, 0, sizeof(int)) != hipSuccess ||
        hipMemcpy(m->active, &zero, sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(m->loss, 0, 2 * sizeof(float)) != hipSuccess || mlp_to_bf16(m->p, m->wb, MLP_PARAMS, m->stream) != hipSuccess ||
        mlp_to_bf16_t(m->p, m->wbt, MLP_PARAMS, m->stream) != hipSuccess ||
        hipStreamSynchronize(m->stream) != hipSuccess) {
        destroy(m);
        return fail(PT_ERR_DEVICE, "PT surrogate initialisation failed.");
    }
    *out = m;
    return 0;
}

void destroy(pt_mlp *m) {
    if (!m) return;
    if (m->stream) (void)hipStreamSynchronize(m->stream);
    for (void *p : m->allocs) (void)hipFree(p);
    if (m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
}

int ensure_rows(pt_mlp *m, int rows) {
    if (rows <= m->rows_cap) return 0;
    const int cap = (std::max(rows, 256) + 63) / 64 * 64;
    const size_t blocks = (size_t)cap / 32;                  // 32-row blocks of the fused kernel
    int rc;
    if ((rc = nalloc(m, &m->x16, (size_t)cap * MLP_IN)) || (rc = nalloc(m, &m->y, cap)) || (rc = nalloc(m, &m->t, cap)) ||
        (rc = nalloc(m, &m->mask, cap)) || (rc = nalloc(m, &m->zf, blocks * 3 * 32 * MLP_H)) ||
        (rc = nalloc(m, &m->partf, blocks * kMlpPart)))
        return rc;
    m->rows_cap = cap;
    return 0;
}

int infer(pt_mlp *m, const uint16_t *x16, int rows, float *y, hipStream_t s) {
    MlpRun R = run_of(m, rows);
    R.x16 = x16;
    R.y = y;
    MLPCHECK(mlp_fused(R, false, s), "inference");
    return 0;
}

int train_iter(pt_mlp *m, const uint16_t *ring_x, const float *ring_t, int cap, const double *step_ptr, int chains,
               int fixed_count, int min_fill, uint32_t seed, int batch, hipStream_t s) {
    // one fused launch (minibatch drawn from the ring, forward, backward, per-block
    // gradients), one for the ordered gradient sum and Adam, one for the step count
    MlpRun R = run_of(m, batch);
    R.ring_x = ring_x;
    R.ring_t = ring_t;
    R.cap = cap;
    R.step_ptr = step_ptr;
    R.chains = chains;
    R.fixed_count = fixed_count;
    R.min_fill = min_fill;
    R.seed = seed;
    MLPCHECK(mlp_fused(R, true, s), "training");
    MLPCHECK(mlp_reduce_adam(m->partf, batch, R.inv_n, m->g, m->loss, 1, m->p, m->m, m->v, m->wb, m->wbt, m->lr, m->beta1, m->beta2,
                             m->eps, m->iter, m->active, s), "Adam");
    return 0;
}

}  // namespace mlpx

// ---------------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------------
extern "C" {

int pt_mlp_create(uint32_t seed, pt_mlp **out) {
    if (!out) return fail(PT_ERR_ARG, "output handle is NULL");
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n < 1) return fail(PT_ERR_DEVICE, "MI355X HIP device is unavailable in this process or context.");
    hipDeviceProp_t prop;
    int d = 0;
    (void)hipGetDevice(&d);
    if (hipGetDeviceProperties(&prop, d) != hipSuccess || std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(PT_ERR_DEVICE, "MI355X HIP device is unavailable (need gfx950).");
    return mlpx::create(seed, out);
}

void pt_mlp_destroy(pt_mlp *net) { mlpx::destroy(net); }

int pt_mlp_param_count(void) { return MLP_PARAMS; }

int pt_mlp_get_params(pt_mlp *net, float *params) {
    if (!net || !params) return fail(PT_ERR_ARG, "NULL argument");
    (void)hipSetDevice(net->device);
    MLPCHECK(hipMemcpyAsync(params, net->p, sizeof(float) * MLP_PARAMS, hipMemcpyDeviceToHost, net->stream), "readback");
    MLPCHECK(hipStreamSynchronize(net->stream), "readback");
    return 0;
}

int pt_mlp_set_params(pt_mlp *net, const float *params) {
    if (!net || !params) return fail(PT_ERR_ARG, "NULL argument");
    (void)hipSetDevice(net->device);
    MLPCHECK(hipMemcpyAsync(net->p, params, sizeof(float) * MLP_PARAMS, hipMemcpyHostToDevice, net->stream), "upload");
    MLPCHECK(hipMemsetAsync(net->m, 0, sizeof(float) * MLP_PARAMS, net->stream), "reset");
    MLPCHECK(hipMemsetAsync(net->v, 0, sizeof(float) * MLP_PARAMS, net->stream), "reset");
    MLPCHECK(hipMemsetAsync(net->iter, 0, sizeof(int), net->stream), "reset");
    MLPCHECK(mlp_to_bf16(net->p, net->wb, MLP_PARAMS, net->stream), "BF16 copy");
    MLPCHECK(mlp_to_bf16_t(net->p, net->wbt, MLP_PARAMS, net->stream), "BF16 copy");
    MLPCHECK(hipStreamSynchronize(net->stream), "upload");
    return 0;
}

static int check_rows(const pt_mlp *net, const float *x, int n, int n_in) {
    if (!net || !x || n < 1) return fail(PT_ERR_ARG, "NULL argument or empty batch");
    if (n_in < 1 || n_in > MLP_IN) return fail(PT_ERR_ARG, "surrogate input width must be 1..64");
    return 0;
}

int pt_mlp_forward(pt_mlp *net, const float *x, int n, int n_in, float *y) {
    int rc = check_rows(net, x, n, n_in);
    if (rc) return rc;
    if (!y) return fail(PT_ERR_ARG, "NULL argument");
    (void)hipSetDevice(net->device);
    if ((rc = mlpx::ensure_rows(net, n)) || (rc = stage(net, x, (size_t)n * n_in))) return rc;
    MLPCHECK(mlp_pack(net->staging, n, n_in, net->x16, net->stream), "pack");
    if ((rc = mlpx::infer(net, net->x16, n, net->y, net->stream))) return rc;
    MLPCHECK(hipMemcpyAsync(y, net->y, sizeof(float) * n, hipMemcpyDeviceToHost, net->stream), "readback");
    MLPCHECK(hipStreamSynchronize(net->stream), "forward");
    return 0;
}

int pt_mlp_gradient(pt_mlp *net, const float *x, const float *t, const float *mask, int n, int n_in, float *loss,
                    float *grad) {
    int rc = check_rows(net, x, n, n_in);
    if (rc) return rc;
    if (!t || !grad) return fail(PT_ERR_ARG, "NULL argument");
    (void)hipSetDevice(net->device);
    if ((rc = mlpx::ensure_rows(net, n)) || (rc = stage(net, x, (size_t)n * n_in))) return rc;
    MLPCHECK(mlp_pack(net->staging, n, n_in, net->x16, net->stream), "pack");
    MLPCHECK(hipStreamSynchronize(net->stream), "upload");
    std::vector<float> ones;
    if (!mask) { ones.assign(n, 1.0f); mask = ones.data(); }
    MLPCHECK(hipMemcpyAsync(net->t, t, sizeof(float) * n, hipMemcpyHostToDevice, net->stream), "upload");
    MLPCHECK(hipMemcpyAsync(net->mask, mask, sizeof(float) * n, hipMemcpyHostToDevice, net->stream), "upload");
    if ((rc = gradient_rows(net, n, net->stream))) return rc;
    MLPCHECK(hipMemcpyAsync(grad, net->g, sizeof(float) * MLP_PARAMS, hipMemcpyDeviceToHost, net->stream), "readback");
    float l = 0;
    MLPCHECK(hipMemcpyAsync(&l, net->loss, sizeof(float), hipMemcpyDeviceToHost, net->stream), "readback");
    MLPCHECK(hipStreamSynchronize(net->stream), "gradient");
    if (loss) *loss = l;
    return 0;
}

int pt_mlp_train(pt_mlp *net, const float *x, const float *t, int n, int n_in, int iters, int batch, float lr, uint32_t seed,
                 float *loss) {
    int rc = check_rows(net, x, n, n_in);
    if (rc) return rc;
    if (!t || iters < 0 || batch < 1 || batch > 4096 || !(lr > 0.0f)) return fail(PT_ERR_ARG, "invalid training arguments");
    (void)hipSetDevice(net->device);
    if ((rc = mlpx::ensure_rows(net, std::max(n, batch)))) return rc;
    if (n > net->data_cap) {
        if ((rc = nalloc(net, &net->data_x, (size_t)n * MLP_IN)) || (rc = nalloc(net, &net->data_t, (size_t)n))) return rc;
        net->data_cap = n;
    }
    if ((rc = stage(net, x, (size_t)n * n_in))) return rc;
    MLPCHECK(mlp_pack(net->staging, n, n_in, net->data_x, net->stream), "pack");
    MLPCHECK(hipStreamSynchronize(net->stream), "upload");
    MLPCHECK(hipMemcpyAsync(net->data_t, t, sizeof(float) * n, hipMemcpyHostToDevice, net->stream), "upload");
    net->lr = lr;
    for (int i = 0; i < iters; ++i)
        if ((rc = mlpx::train_iter(net, net->data_x, net->data_t, n, nullptr, 0, n, 1, seed, batch, net->stream))) return rc;
    float l = 0;
    MLPCHECK(hipMemcpyAsync(&l, net->loss, sizeof(float), hipMemcpyDeviceToHost, net->stream), "readback");
    MLPCHECK(hipStreamSynchronize(net->stream), "training");
    if (loss) *loss = l;
    return 0;
}

}  // extern "C"
