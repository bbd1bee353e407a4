// pt_mlp.h -- the surrogate network's device layout and kernel entry points.
//
// Network 64 -> 256 -> 256 -> 256 -> 1 with SiLU, BF16 operands and FP32 accumulation on
// MFMA; FP32 master parameters in one flat vector (W1 row major by input, b1, W2, b2,
// W3, b3, w4, b4) with a BF16 copy of the same layout for the products.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pt_plan.h"

namespace ptk {

constexpr int MLP_IN = 64;
constexpr int MLP_H = 256;
constexpr int MLP_OFF_W1 = 0;
constexpr int MLP_OFF_B1 = MLP_OFF_W1 + MLP_IN * MLP_H;
constexpr int MLP_OFF_W2 = MLP_OFF_B1 + MLP_H;
constexpr int MLP_OFF_B2 = MLP_OFF_W2 + MLP_H * MLP_H;
constexpr int MLP_OFF_W3 = MLP_OFF_B2 + MLP_H;
constexpr int MLP_OFF_B3 = MLP_OFF_W3 + MLP_H * MLP_H;
constexpr int MLP_OFF_W4 = MLP_OFF_B3 + MLP_H;
constexpr int MLP_OFF_B4 = MLP_OFF_W4 + MLP_H;
constexpr int MLP_PARAMS = MLP_OFF_B4 + 1;

// per-block partial row of the fused training kernel: every parameter's gradient, then
// the block's sum of m e^2 (at MLP_PARAMS), padded to a multiple of four floats
constexpr int kMlpPart = (MLP_PARAMS + 1 + 3) / 4 * 4;

// one launch of k_mlp_fused: the rows (given in x16 / t / mask, or drawn from a replay
// ring when x16 is null), the weights and the per-block areas
struct MlpRun {
    const uint16_t *x16;                 // rows x MLP_IN BF16 inputs (null: gather from the ring)
    const float *t, *mask;               // targets and row weights of given rows (mask may be null: 0)
    const uint16_t *ring_x;              // replay ring (rows of MLP_IN BF16) and its targets
    const float *ring_t;
    int cap;
    const double *step_ptr;              // ring fill = step * chains (PT), or fixed_count when null
    int chains, fixed_count, min_fill;
    uint32_t seed;
    const int *iter;                     // Adam step count (the minibatch draw's counter)
    int *active;                         // out: whether the minibatch trains (gathered rows)
    int rows;                            // rows of the minibatch: ceil(rows / 32) workgroups
    float inv_n;                         // 1 / rows
    const uint16_t *wb, *wbt;            // BF16 parameters, and with the weight blocks transposed
    const float *p;                      // FP32 parameters (biases)
    float *z;                            // 3 x 32 x MLP_H pre-activations per workgroup
    float *part;                         // kMlpPart per workgroup (training)
    float *y;                            // predictions (inference)
};

// device view of a plan's screening state
struct ScreenDev {
    const double *center, *inv_half;   // network input u_i = (x_i - center_i) inv_half_i
    uint16_t *ring_x;                  // replay ring: cap rows of MLP_IN BF16 inputs
    float *ring_t;                     //   and their targets (log Rwp, NaN = infeasible)
    double2 *cand;                     // chains * K candidate moves (j, value)
    uint16_t *xs;                      // chains * (K + 1) network inputs (row K: current x)
    float *pred;                       // chains * (K + 1) predictions
    double2 *pick;                     // chosen move per chain
    unsigned *picked;                  // count of picks other than the plain proposal
    unsigned *skipped;                 // count of proposals not evaluated (predicted hopeless)
    double log_wss;                    // log sum w y_obs^2
    float margin;                      // predicted log Rwp gain a pick must show
    float skip_margin;                 // log Rwp allowance taken off a prediction before a skip
    float skip_log;                    // skip when the predicted acceptance < exp(-skip_log); 0 off
    int cold_chains;                   // chains [0, cold_chains) are screened and recorded
    int cap, K, warmup_iters, pad;
};

hipError_t mlp_to_bf16(const float *p, uint16_t *wb, int n, hipStream_t s);
hipError_t mlp_pack(const float *x, int n, int n_in, uint16_t *x16, hipStream_t s);
// training (forward, backward, per-block gradients) or inference (forward, head) of
// R.rows rows in ceil(rows / 32) workgroups
hipError_t mlp_fused(const MlpRun &R, bool train, hipStream_t s);
// the blocks' gradients summed into g (and the loss); with adam, one Adam step and the step count
hipError_t mlp_reduce_adam(const float *part, int rows, float inv_n, float *g, float *loss, int adam, float *p, float *m, float *v,
                           uint16_t *wb, uint16_t *wbt, float lr, float b1, float b2, float eps, const int *iter, const int *active,
                           hipStream_t s);
// BF16 copy of the parameters with the weight blocks transposed (out x in)
hipError_t mlp_to_bf16_t(const float *p, uint16_t *wbt, int n, hipStream_t s);
size_t mlp_fused_lds_bytes();
hipError_t screen_candidates(const DevPlan &P, const ScreenDev &S, hipStream_t s);
hipError_t screen_inputs(const double *x, int n, int nv, const ScreenDev &S, uint16_t *xs, hipStream_t s);
hipError_t screen_pick(const DevPlan &P, const ScreenDev &S, const int *iter, hipStream_t s);
hipError_t screen_record(const DevPlan &P, const ScreenDev &S, hipStream_t s);

}  // namespace ptk
