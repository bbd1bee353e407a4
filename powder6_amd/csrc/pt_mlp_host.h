// pt_mlp_host.h -- the surrogate network object (host side) shared by the library's
// translation units: device buffers, and the training / inference sequences that the
// screened PT sweep records into its graph.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "pt_mlp.h"

struct pt_mlp {
    int device = -1;
    hipStream_t stream = nullptr;      // stream of the standalone calls
    std::vector<void *> allocs;
    // parameters, gradients, Adam moments (FP32, MLP_PARAMS each), BF16 copy
    float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
    uint16_t *wb = nullptr, *wbt = nullptr;   // BF16 copy; the same with the weight blocks transposed
    int *iter = nullptr, *active = nullptr;   // Adam step count; minibatch active flag
    float *loss = nullptr;                    // [0] last minibatch loss, [1] its running mean
    float lr = 1e-3f, beta1 = 0.9f, beta2 = 0.999f, eps = 1e-8f;
    // activations for up to rows_cap rows
    int rows_cap = 0;
    uint16_t *x16 = nullptr;
    float *y = nullptr, *t = nullptr, *mask = nullptr;
    float *zf = nullptr, *partf = nullptr;    // fused kernel: pre-activations and gradient partials per 32-row block
    // host-given training set of pt_mlp_train (rows of MLP_IN BF16, targets)
    int data_cap = 0;
    uint16_t *data_x = nullptr;
    float *data_t = nullptr;
    float *staging = nullptr;                 // FP32 upload area
    size_t staging_cap = 0;
};

namespace mlpx {
// a network on the current device, parameters drawn from `seed` (He-scaled normals)
int create(uint32_t seed, pt_mlp **out);
void destroy(pt_mlp *m);
int ensure_rows(pt_mlp *m, int rows);
// predictions of `rows` BF16 input rows into y (device)
int infer(pt_mlp *m, const uint16_t *x16, int rows, float *y, hipStream_t s);
// one Adam step on a minibatch of `batch` rows drawn from a replay ring (see k_mlp_gather)
int train_iter(pt_mlp *m, const uint16_t *ring_x, const float *ring_t, int cap, const double *step_ptr, int chains,
               int fixed_count, int min_fill, uint32_t seed, int batch, hipStream_t s);
}  // namespace mlpx
