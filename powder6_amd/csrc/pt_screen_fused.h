// pt_screen_fused.h -- the surrogate network evaluated inside the PT setup stage: one
// wavefront predicts log Rwp for two inputs (a chain's proposal and its state) with the
// flat BF16 weights of the trained network (64 -> 256 -> 256 -> 256 -> 1, SiLU), FP32
// sums, activations rounded to BF16 between layers as the MFMA inference stores them.
#pragma once
#include "pt_device.h"
#include "pt_mlp.h"

namespace ptk {

__device__ __forceinline__ float sf_from_bf16(uint32_t h16) { return __uint_as_float(h16 << 16); }
// round to nearest even (finite inputs)
__device__ __forceinline__ uint16_t sf_to_bf16(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float sf_silu(float z) { return z * (1.0f / (1.0f + __expf(-z))); }

// network input i of parameter value x: BF16 of (x - center_i) inv_half_i
__device__ __forceinline__ uint16_t sf_input(const ScreenFused &F, int i, double x) {
    return sf_to_bf16((float)((x - F.center[i]) * F.inv_half[i]));
}

// one layer for two rows: lane l forms units 4l .. 4l+3 of out[r][.] = BF16(SiLU(in[r] W + b)),
// W row major by input (K rows of 256), eight input rows per round of loads
template <int K>
__device__ __forceinline__ void sf_layer(const uint16_t *__restrict__ W, const float *__restrict__ b, const float (*in)[256],
                                         float (*out)[256], int lane) {
    float a0[4] = {0.f, 0.f, 0.f, 0.f}, a1[4] = {0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < K; k0 += 8) {
        uint2 wv[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) wv[u] = *reinterpret_cast<const uint2 *>(W + (size_t)(k0 + u) * MLP_H + 4 * lane);
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const float w[4] = {sf_from_bf16(wv[u].x & 0xffffu), sf_from_bf16(wv[u].x >> 16), sf_from_bf16(wv[u].y & 0xffffu),
                                sf_from_bf16(wv[u].y >> 16)};
            const float x0 = in[0][k0 + u], x1 = in[1][k0 + u];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                a0[q] = fmaf(x0, w[q], a0[q]);
                a1[q] = fmaf(x1, w[q], a1[q]);
            }
        }
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const float bb = b[4 * lane + q];
        out[0][4 * lane + q] = sf_from_bf16(sf_to_bf16(sf_silu(a0[q] + bb)));
        out[1][4 * lane + q] = sf_from_bf16(sf_to_bf16(sf_silu(a1[q] + bb)));
    }
}

// predicted log Rwp of the two input rows u[0], u[1] (64 values each, BF16 as float, in
// LDS; h: LDS for two layers of activations), by the 64 lanes of one wavefront: y0, y1 in
// every lane
__device__ inline void sf_predict2(const ScreenFused &F, const float (*u)[256], float (*h)[2][256], float &y0, float &y1) {
    const int lane = threadIdx.x & 63;
    sf_layer<MLP_IN>(F.w + MLP_OFF_W1, F.p + MLP_OFF_B1, u, h[0], lane);
    __syncthreads();
    sf_layer<MLP_H>(F.w + MLP_OFF_W2, F.p + MLP_OFF_B2, h[0], h[1], lane);
    __syncthreads();
    sf_layer<MLP_H>(F.w + MLP_OFF_W3, F.p + MLP_OFF_B3, h[1], h[0], lane);
    __syncthreads();
    float s0 = 0.f, s1 = 0.f;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const float w4 = sf_from_bf16(F.w[MLP_OFF_W4 + lane + 64 * q]);
        s0 = fmaf(h[0][0][lane + 64 * q], w4, s0);
        s1 = fmaf(h[0][1][lane + 64 * q], w4, s1);
    }
    for (int o = 32; o > 0; o >>= 1) {
        s0 += __shfl_xor(s0, o);
        s1 += __shfl_xor(s1, o);
    }
    y0 = s0 + F.p[MLP_OFF_B4];
    y1 = s1 + F.p[MLP_OFF_B4];
}

}  // namespace ptk
