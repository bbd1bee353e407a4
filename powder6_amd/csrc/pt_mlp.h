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

// GEMM epilogue operands (row stride ldc for every output)
struct GemmEpi {
    float *c32;          // FP32 output (store, or dz)
    uint16_t *c16;       // BF16 output (silu(z), or dz)
    const float *bias;   // bias per column (forward)
    const float *z32;    // pre-activations (backward)
    float *zout;         // pre-activations out (forward)
    float *colpart;      // per row-block column sums of dz (backward), [gridDim.y][N]
    int ldc;
    int pad;
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

hipError_t mlp_gemm_forward(const uint16_t *A, const uint16_t *W, int M, int K, const float *bias, float *z, uint16_t *a,
                            hipStream_t s);
hipError_t mlp_gemm_dact(const uint16_t *dz, const uint16_t *W, int M, const float *z, float *d32, uint16_t *d16, float *colpart,
                         hipStream_t s);
hipError_t mlp_gemm_wgrad(const uint16_t *a, int a_cols, const uint16_t *dz, int M, float *gW, hipStream_t s);
hipError_t mlp_head(const uint16_t *a3, const float *z3, const uint16_t *w4b, const float *b4, int M, float *y, const float *t,
                    const float *mask, float inv_n, float *dz32, uint16_t *dz16, float *part, hipStream_t s);
hipError_t mlp_bias_grads(const float *pb1, const float *pb2, const float *ph, int nblk, float inv_n, float *g, float *loss,
                          hipStream_t s);
hipError_t mlp_adam(float *p, const float *g, float *m, float *v, uint16_t *wb, int n, float lr, float b1, float b2, float eps,
                    const int *iter, const int *active, hipStream_t s);
hipError_t mlp_to_bf16(const float *p, uint16_t *wb, int n, hipStream_t s);
hipError_t mlp_pack(const float *x, int n, int n_in, uint16_t *x16, hipStream_t s);
hipError_t mlp_gather(const uint16_t *ring_x, const float *ring_t, int cap, const double *step_ptr, int chains, int fixed_count,
                      int min_fill, uint32_t seed, const int *iter, int B, uint16_t *xb, float *tb, float *mb, int *active,
                      hipStream_t s);
hipError_t screen_candidates(const DevPlan &P, const ScreenDev &S, hipStream_t s);
hipError_t screen_inputs(const double *x, int n, int nv, const ScreenDev &S, uint16_t *xs, hipStream_t s);
hipError_t screen_pick(const DevPlan &P, const ScreenDev &S, const int *iter, hipStream_t s);
hipError_t screen_record(const DevPlan &P, const ScreenDev &S, hipStream_t s);

}  // namespace ptk
