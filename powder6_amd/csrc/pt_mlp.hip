// pt_mlp.hip -- gfx950 kernels of the surrogate network and of surrogate-screened PT.
//
//   k_mlp_fused<TRAIN>       32 rows per workgroup through the network with every activation
//                            in LDS (v_mfma_f32_16x16x32_bf16, FP32 accumulation): the
//                            forward pass and head (inference), and with TRAIN the loss,
//                            the backward pass and the block's gradient partials; training
//                            rows are drawn from the replay ring inside the kernel.
//   k_mlp_reduce_adam        ordered sums of the blocks' partials (deterministic), Adam on the
//                            FP32 master weights, the BF16 copies (row-major and transposed).
//   k_mlp_tick               the Adam step count.
//   k_mlp_to_bf16(_t) / k_mlp_pack  BF16 copies of the parameters / host rows to BF16.
//   k_screen_*               the screened PT step: K candidate moves per chain, their
//                            predicted log Rwp, the pick, and the record of every exactly
//                            evaluated trial into the replay ring.
#include <mutex>

#include "pt_device.h"
#include "pt_mlp.h"

namespace ptk {

typedef float f32x4_t __attribute__((ext_vector_type(4)));

// round to nearest even (finite inputs)
__device__ __forceinline__ uint16_t to_bf16(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float from_bf16(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
__device__ __forceinline__ float sigm(float z) { return 1.0f / (1.0f + __expf(-z)); }
__device__ __forceinline__ float silu(float z) { return z * sigm(z); }
__device__ __forceinline__ float dsilu(float z) {
    const float s = sigm(z);
    return s * (1.0f + z * (1.0f - s));
}

__global__ void k_mlp_tick(int *iter, const int *active) {
    if (threadIdx.x == 0 && blockIdx.x == 0 && *active) *iter += 1;
}

__global__ void __launch_bounds__(256) k_mlp_to_bf16(const float *__restrict__ p, uint16_t *wb, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) wb[i] = to_bf16(p[i]);
}

// host rows (n x n_in FP32) -> BF16 rows of MLP_IN (zero padded)
__global__ void __launch_bounds__(64) k_mlp_pack(const float *__restrict__ x, int n_in, uint16_t *x16) {
    const int r = blockIdx.x, i = threadIdx.x;
    x16[(size_t)r * MLP_IN + i] = i < n_in ? to_bf16(x[(size_t)r * n_in + i]) : (uint16_t)0;
}

// ---- one training step or inference in a few launches: 32 rows per workgroup ----------
// k_mlp_fused<TRAIN> takes a block of 32 rows (given, or drawn from the replay ring: row r
// takes entry floor(u count) of the ring's count filled entries, u uniform from (seed,
// Adam step, r)) through the three hidden layers and the head with every
// activation in LDS; with TRAIN it goes on through the backward pass and writes this
// block's gradient of every parameter (the weight gradients as the 32-row products
// A^T dZ, the bias gradients as column sums) and its loss into its own partial row.
// k_mlp_reduce_adam sums the blocks' partials in block order (deterministic, no atomics)
// and runs Adam; k_mlp_tick counts the step.  The layers run on v_mfma_f32_16x16x32_bf16
// with FP32 accumulation: a wavefront owns 64 output columns of the block's 32 rows (2 x 4
// blocks of 16 x 16).  Operands are K-contiguous per lane, so each activation and
// gradient is kept in LDS row-major (the forward's and backward's A operand) and
// transposed (the weight gradients' operands), and the forward reads the transposed BF16
// weights wbt (out x in).  Pre-activations z go to a per-block global area; each lane
// reads back only the elements it wrote.
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));

constexpr int kTR = 32;                  // rows per workgroup
constexpr int kLdR = MLP_H + 8;          // row-major LDS stride (BF16), 16-byte aligned rows
constexpr int kLdX = MLP_IN + 8;
constexpr int kLdT = kTR + 8;            // transposed LDS stride (BF16)

struct FusedLds {
    uint16_t x[kTR][kLdX];               // input rows
    uint16_t xt[MLP_IN][kLdT];           // the same, transposed
    uint16_t r1[kTR][kLdR];              // a1, then dz2
    uint16_t t1[MLP_H][kLdT];            // a1 transposed
    uint16_t r2[kTR][kLdR];              // a2, then dz3
    uint16_t t2[MLP_H][kLdT];            // a2 transposed, then dz1 transposed
    uint16_t r3[kTR][kLdR];              // a3
    uint16_t t3[MLP_H][kLdT];            // dz3 transposed
    uint16_t t4[MLP_H][kLdT];            // dz2 transposed
    float tgt[kTR], msk[kTR], dy[kTR];
    float red[4][MLP_H + 2];             // head: per-wavefront w4 gradient, b4 gradient, loss
};
size_t mlp_fused_lds_bytes() { return sizeof(FusedLds); }

__device__ __forceinline__ bf16x8_t frag(const uint16_t *p) { return __builtin_bit_cast(bf16x8_t, *reinterpret_cast<const uint4 *>(p)); }
__device__ __forceinline__ f32x4_t mfma32(bf16x8_t a, bf16x8_t b, f32x4_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// z = a W + b for the block's rows: A row-major in LDS (K = KIN), W^T rows of the
// wavefront's 64 outputs from global; z to zg, BF16 SiLU(z) to ar (and at, transposed)
template <int KIN>
__device__ __forceinline__ void fused_forward(const uint16_t *A, int lda, const uint16_t *__restrict__ wt,
                                              const float *__restrict__ bias, float *zg, uint16_t *ar, uint16_t (*at)[kLdT],
                                              bool store_z) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, li = lane & 15, kq = 8 * (lane >> 4), dq = 4 * (lane >> 4);
    f32x4_t acc[2][4];
#pragma unroll
    for (int rb = 0; rb < 2; ++rb)
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[rb][c] = f32x4_t{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int k0 = 0; k0 < KIN; k0 += 32) {
        bf16x8_t b[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) b[c] = frag(wt + (size_t)(64 * wave + 16 * c + li) * KIN + k0 + kq);
        const bf16x8_t a0 = frag(A + li * lda + k0 + kq), a1 = frag(A + (16 + li) * lda + k0 + kq);
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            acc[0][c] = mfma32(a0, b[c], acc[0][c]);
            acc[1][c] = mfma32(a1, b[c], acc[1][c]);
        }
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const int o = 64 * wave + 16 * c + li;
        const float bo = bias[o];
#pragma unroll
        for (int rb = 0; rb < 2; ++rb)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int row = 16 * rb + dq + e;
                const float z = acc[rb][c][e] + bo;
                if (store_z) zg[row * MLP_H + o] = z;
                const uint16_t v = to_bf16(silu(z));
                ar[row * kLdR + o] = v;
                if (at) at[o][row] = v;
            }
    }
}

// dz_prev = (dZ W^T) SiLU'(z_prev): A = dZ row-major in LDS, W rows of the wavefront's 64
// outputs from global (K = the layer's outputs); BF16 dz_prev to dr (if given) and dt
// (transposed), its column sums (the bias gradient) to bgrad
__device__ __forceinline__ void fused_backward(const uint16_t *D, const uint16_t *__restrict__ w, const float *zg, uint16_t *dr,
                                               uint16_t (*dt)[kLdT], float *bgrad) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, li = lane & 15, kq = 8 * (lane >> 4), dq = 4 * (lane >> 4);
    f32x4_t acc[2][4];
#pragma unroll
    for (int rb = 0; rb < 2; ++rb)
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[rb][c] = f32x4_t{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int k0 = 0; k0 < MLP_H; k0 += 32) {
        bf16x8_t b[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) b[c] = frag(w + (size_t)(64 * wave + 16 * c + li) * MLP_H + k0 + kq);
        const bf16x8_t a0 = frag(D + li * kLdR + k0 + kq), a1 = frag(D + (16 + li) * kLdR + k0 + kq);
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            acc[0][c] = mfma32(a0, b[c], acc[0][c]);
            acc[1][c] = mfma32(a1, b[c], acc[1][c]);
        }
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const int i = 64 * wave + 16 * c + li;
        float cs = 0.f;
#pragma unroll
        for (int rb = 0; rb < 2; ++rb)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int row = 16 * rb + dq + e;
                const float d = acc[rb][c][e] * dsilu(zg[row * MLP_H + i]);
                const uint16_t v = to_bf16(d);
                if (dr) dr[row * kLdR + i] = v;
                dt[i][row] = v;
                cs += d;
            }
        cs += __shfl_xor(cs, 16);
        cs += __shfl_xor(cs, 32);
        if (lane < 16) bgrad[i] = cs;
    }
}

// gW[i][o] = sum over the block's rows of A[row][i] dZ[row][o], from the transposed
// copies (K = the 32 rows: one MFMA per 16 x 16 block); MI rows of gW, the wavefront's 64
// columns
template <int MI>
__device__ __forceinline__ void fused_wgrad(const uint16_t (*at)[kLdT], const uint16_t (*dt)[kLdT], float *gw) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, li = lane & 15, kq = 8 * (lane >> 4), dq = 4 * (lane >> 4);
    bf16x8_t b[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) b[c] = frag(&dt[64 * wave + 16 * c + li][kq]);
#pragma unroll 4
    for (int ib = 0; ib < MI / 16; ++ib) {
        const bf16x8_t a = frag(&at[16 * ib + li][kq]);
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const f32x4_t r = mfma32(a, b[c], f32x4_t{0.f, 0.f, 0.f, 0.f});
            const int o = 64 * wave + 16 * c + li;
#pragma unroll
            for (int e = 0; e < 4; ++e) gw[(size_t)(16 * ib + dq + e) * MLP_H + o] = r[e];
        }
    }
}

template <bool TRAIN>
__global__ void __launch_bounds__(256) k_mlp_fused(MlpRun R) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    FusedLds &S = *reinterpret_cast<FusedLds *>(lds_raw);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, blk = blockIdx.x, r0 = blk * kTR;
    // the block's rows: given, or drawn from the replay ring (inactive, all rows weightless,
    // until the ring holds min_fill entries; entries with a NaN target weigh nothing)
    bool act = true;
    if (!R.x16) {
        long long count = R.fixed_count;
        if (R.step_ptr) count = (long long)fmax(0.0, *R.step_ptr) * R.chains;
        if (count > R.cap) count = R.cap;
        act = count >= R.min_fill && count > 0;
        if (blk == 0 && tid == 0) *R.active = act ? 1 : 0;
        for (int q = tid; q < kTR * (MLP_IN / 8); q += blockDim.x) {
            const int rr = q / (MLP_IN / 8), c8 = 8 * (q % (MLP_IN / 8)), row = r0 + rr;
            uint4 v = make_uint4(0u, 0u, 0u, 0u);
            float tv = 0.f, mv = 0.f;
            if (act && row < R.rows) {
                const double u = uniform01(R.seed, row, (unsigned long long)(*R.iter), 11);
                long long idx = (long long)(u * (double)count);
                if (idx >= count) idx = count - 1;
                v = *reinterpret_cast<const uint4 *>(R.ring_x + (size_t)idx * MLP_IN + c8);
                const float tr = R.ring_t[idx];
                const bool ok = tr == tr;
                tv = ok ? tr : 0.f;
                mv = ok ? 1.f : 0.f;
            }
            *reinterpret_cast<uint4 *>(&S.x[rr][c8]) = v;
            if (c8 == 0) { S.tgt[rr] = tv; S.msk[rr] = mv; }
        }
    } else {
        for (int q = tid; q < kTR * (MLP_IN / 8); q += blockDim.x) {
            const int rr = q / (MLP_IN / 8), c8 = 8 * (q % (MLP_IN / 8)), row = r0 + rr;
            const bool in = row < R.rows;
            *reinterpret_cast<uint4 *>(&S.x[rr][c8]) =
                in ? *reinterpret_cast<const uint4 *>(R.x16 + (size_t)row * MLP_IN + c8) : make_uint4(0u, 0u, 0u, 0u);
            if (c8 == 0) { S.tgt[rr] = in && R.t ? R.t[row] : 0.f; S.msk[rr] = in && R.mask ? R.mask[row] : 0.f; }
        }
    }
    __syncthreads();
    if (TRAIN)
        for (int q = tid; q < kTR * MLP_IN; q += blockDim.x) S.xt[q % MLP_IN][q / MLP_IN] = S.x[q / MLP_IN][q % MLP_IN];
    float *z = R.z + (size_t)blk * 3 * kTR * MLP_H;
    fused_forward<MLP_IN>(&S.x[0][0], kLdX, R.wbt + MLP_OFF_W1, R.p + MLP_OFF_B1, z, &S.r1[0][0], TRAIN ? S.t1 : nullptr, TRAIN);
    __syncthreads();
    fused_forward<MLP_H>(&S.r1[0][0], kLdR, R.wbt + MLP_OFF_W2, R.p + MLP_OFF_B2, z + kTR * MLP_H, &S.r2[0][0],
                         TRAIN ? S.t2 : nullptr, TRAIN);
    __syncthreads();
    fused_forward<MLP_H>(&S.r2[0][0], kLdR, R.wbt + MLP_OFF_W3, R.p + MLP_OFF_B3, z + 2 * kTR * MLP_H, &S.r3[0][0], nullptr, TRAIN);
    __syncthreads();
    // head (256 -> 1, loss, and its backward): wavefront w takes rows 8w..8w+7, lanes the hidden units
    float w4[MLP_H / 64], gw4[MLP_H / 64] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int q = 0; q < MLP_H / 64; ++q) w4[q] = from_bf16(R.wb[MLP_OFF_W4 + lane + 64 * q]);
    float gb4 = 0.f, gl = 0.f;
    for (int rr = 8 * wave; rr < 8 * wave + 8; ++rr) {
        float av[MLP_H / 64], s = 0.f;
#pragma unroll
        for (int q = 0; q < MLP_H / 64; ++q) {
            av[q] = from_bf16(S.r3[rr][lane + 64 * q]);
            s = fmaf(av[q], w4[q], s);
        }
        for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
        const float yv = s + R.p[MLP_OFF_B4];
        if (!TRAIN) {
            if (lane == 0 && r0 + rr < R.rows) R.y[r0 + rr] = yv;
            continue;
        }
        const float m = S.msk[rr], e = yv - S.tgt[rr];
        const float d = 2.0f * m * e * R.inv_n;
        gb4 += d;
        gl += m * e * e;
#pragma unroll
        for (int q = 0; q < MLP_H / 64; ++q) gw4[q] = fmaf(d, av[q], gw4[q]);
        if (lane == 0) S.dy[rr] = d;
    }
    if (!TRAIN) return;
#pragma unroll
    for (int q = 0; q < MLP_H / 64; ++q) S.red[wave][lane + 64 * q] = gw4[q];
    if (lane == 0) { S.red[wave][MLP_H] = gb4; S.red[wave][MLP_H + 1] = gl; }
    __syncthreads();
    float *part = R.part + (size_t)blk * kMlpPart;
    for (int i = tid; i < MLP_H + 2; i += blockDim.x) {
        const float v = S.red[0][i] + S.red[1][i] + S.red[2][i] + S.red[3][i];
        if (i < MLP_H) part[MLP_OFF_W4 + i] = v;
        else if (i == MLP_H) part[MLP_OFF_B4] = v;
        else part[MLP_PARAMS] = v;                           // sum of m e^2
    }
    // dz3 = dy w4 SiLU'(z3), in the forward's element mapping (each lane's own z3)
    {
        const int li = lane & 15, dq = 4 * (lane >> 4);
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const int o = 64 * wave + 16 * c + li;
            const float w4o = from_bf16(R.wb[MLP_OFF_W4 + o]);
            float cs = 0.f;
#pragma unroll
            for (int rb = 0; rb < 2; ++rb)
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int row = 16 * rb + dq + e;
                    const float g = S.dy[row] * w4o * dsilu(z[2 * kTR * MLP_H + row * MLP_H + o]);
                    const uint16_t v = to_bf16(g);
                    S.r2[row][o] = v;                        // a2 (row-major) is no longer needed
                    S.t3[o][row] = v;
                    cs += g;
                }
            cs += __shfl_xor(cs, 16);
            cs += __shfl_xor(cs, 32);
            if (lane < 16) part[MLP_OFF_B3 + o] = cs;
        }
    }
    __syncthreads();
    fused_wgrad<MLP_H>(S.t2, S.t3, part + MLP_OFF_W3);                                   // dW3 = a2^T dz3
    fused_backward(&S.r2[0][0], R.wb + MLP_OFF_W3, z + kTR * MLP_H, &S.r1[0][0], S.t4, part + MLP_OFF_B2);   // dz2
    __syncthreads();
    fused_wgrad<MLP_H>(S.t1, S.t4, part + MLP_OFF_W2);                                   // dW2 = a1^T dz2
    fused_backward(&S.r1[0][0], R.wb + MLP_OFF_W2, z, nullptr, S.t2, part + MLP_OFF_B1); // dz1 (a2^T is used up)
    __syncthreads();
    fused_wgrad<MLP_IN>(S.xt, S.t2, part + MLP_OFF_W1);                                  // dW1 = x^T dz1
}

// g = sum of the blocks' partials in block order; loss[0] = the partials' loss entries summed
// and scaled by inv_n, loss[1] its running average; with adam, Adam on every parameter and the BF16 copies (wb row-major, wbt the
// weight blocks transposed) refreshed
__device__ __forceinline__ int mlp_transposed_index(int i) {
    if (i < MLP_OFF_B1) return (i % MLP_H) * MLP_IN + i / MLP_H;
    if (i >= MLP_OFF_W2 && i < MLP_OFF_B2) { const int j = i - MLP_OFF_W2; return MLP_OFF_W2 + (j % MLP_H) * MLP_H + j / MLP_H; }
    if (i >= MLP_OFF_W3 && i < MLP_OFF_B3) { const int j = i - MLP_OFF_W3; return MLP_OFF_W3 + (j % MLP_H) * MLP_H + j / MLP_H; }
    return i;
}
__global__ void __launch_bounds__(256) k_mlp_reduce_adam(const float *__restrict__ part, int nblk, float inv_n, float *g,
                                                         float *loss, int adam, float *p, float *m, float *v, uint16_t *wb,
                                                         uint16_t *wbt, float lr, float b1, float b2, float eps,
                                                         const int *iter, const int *active) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == MLP_PARAMS) {
        float l = 0.f;
        for (int k = 0; k < nblk; ++k) l += part[(size_t)k * kMlpPart + MLP_PARAMS];
        loss[0] = l * inv_n;
        loss[1] = loss[1] > 0.f ? loss[1] + (loss[0] - loss[1]) * (1.0f / 64.0f) : loss[0];
        return;
    }
    if (i > MLP_PARAMS) return;
    float s = 0.f;
    for (int k = 0; k < nblk; ++k) s += part[(size_t)k * kMlpPart + i];
    g[i] = s;
    if (!adam || !*active) return;
    const float t = (float)(*iter + 1);
    const float mi = b1 * m[i] + (1.0f - b1) * s;
    const float vi = b2 * v[i] + (1.0f - b2) * s * s;
    m[i] = mi;
    v[i] = vi;
    const float mh = mi / (1.0f - powf(b1, t)), vh = vi / (1.0f - powf(b2, t));
    const float pi = p[i] - lr * mh / (sqrtf(vh) + eps);
    p[i] = pi;
    const uint16_t h = to_bf16(pi);
    wb[i] = h;
    wbt[mlp_transposed_index(i)] = h;
}

__global__ void __launch_bounds__(256) k_mlp_to_bf16_t(const float *__restrict__ p, uint16_t *wbt, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) wbt[mlp_transposed_index(i)] = to_bf16(p[i]);
}

// ---- surrogate-screened PT step ---------------------------------------------------------
// network input of a parameter vector: u_i = (x_i - c_i) h_i for the nv variables
__device__ __forceinline__ uint16_t screen_input(const ScreenDev &S, int i, double x) {
    return to_bf16((float)((x - S.center[i]) * S.inv_half[i]));
}

// candidate k < K of chain c: one coordinate move drawn like the plain PT proposal
// (candidate 0 uses exactly the plain proposal's draws, rietveld_core.js:2196-2212), and
// the network input of the moved vector; row K is the chain's current vector.  One
// workgroup of 64 per (chain, row).
__global__ void __launch_bounds__(64) k_screen_candidates(DevPlan P, ScreenDev S) {
    const int R = S.K + 1;
    const int c = blockIdx.x / R, k = blockIdx.x % R;
    const int nv = P.nv;
    const double *st = P.state + (size_t)c * P.stride;
    __shared__ int sj;
    __shared__ double sv;
    if (threadIdx.x == 0) {
        int j = -1;
        double v = 0.0;
        if (k < S.K) {
            const unsigned long long ctr = (unsigned long long)fmax(0.0, st[3 * nv + 2]);
            const double temp = P.temp[c];
            const int base = k == 0 ? 0 : 16 + 3 * k;
            const double u0 = uniform01(P.seed, c, ctr, base), u1 = uniform01(P.seed, c, ctr, base + 1);
            const double u2 = uniform01(P.seed, c, ctr, base + 2);
            j = (int)floor(u0 * nv);
            if (j > nv - 1) j = nv - 1;
            const double z = sqrt(-2.0 * log(u1)) * cos(2.0 * kPi * u2);
            v = reflect_box(st[j] + z * P.vscale[j] * sqrt(temp), P.vlo[j], P.vhi[j]);
            S.cand[(size_t)c * S.K + k] = make_double2((double)j, v);
        }
        sj = j;
        sv = v;
    }
    __syncthreads();
    const int i = threadIdx.x;
    uint16_t u = 0;
    if (i < nv) u = screen_input(S, i, i == sj ? sv : st[i]);
    S.xs[(size_t)blockIdx.x * MLP_IN + i] = u;
}

// network inputs of n given parameter vectors (n x nv doubles), one workgroup per row
__global__ void __launch_bounds__(64) k_screen_inputs(const double *__restrict__ x, int nv, ScreenDev S, uint16_t *xs) {
    const int r = blockIdx.x, i = threadIdx.x;
    xs[(size_t)r * MLP_IN + i] = i < nv ? screen_input(S, i, x[(size_t)r * nv + i]) : (uint16_t)0;
}

// per chain: once the network has trained warmup_iters iterations, a screened (cold)
// chain takes the candidate with the lowest predicted log Rwp if that beats the
// prediction at its current vector by the margin; otherwise, and for every other chain,
// candidate 0 (the plain PT proposal).  With skipping on, a screened chain whose pick is
// predicted hopeless -- even after skip_margin is taken off its predicted log Rwp rise d,
// the Metropolis acceptance exp(-(chi2 (e^{2d} - 1)) / 2T) (rietveld_core.js:2213-2222)
// stays below exp(-skip_log) -- is marked (move -1) and not evaluated.
__global__ void __launch_bounds__(64) k_screen_pick(DevPlan P, ScreenDev S, const int *iter) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= P.chains) return;
    const int R = S.K + 1;
    int best = 0;
    bool skip = false;
    if (*iter >= S.warmup_iters && c < S.cold_chains) {
        const float *pr = S.pred + (size_t)c * R;
        int b = 0;
        for (int k = 1; k < S.K; ++k)
            if (pr[k] < pr[b]) b = k;
        if (pr[b] < pr[S.K] - S.margin) best = b;
        const float d = pr[best] - pr[S.K] - S.skip_margin;
        if (S.skip_log > 0.0f && d > 0.0f) {
            const double *st = P.state + (size_t)c * P.stride;
            const double cur = st[3 * P.nv];
            const double temp = P.temp[c];
            if (cur > 0.0 && cur < 3e38 && cur * expm1(2.0 * (double)d) > 2.0 * temp * (double)S.skip_log) skip = true;
        }
    }
    S.pick[c] = skip ? make_double2(-1.0, 0.0) : S.cand[(size_t)c * S.K + best];
    if (best != 0 && !skip) atomicAdd(S.picked, 1u);
    if (skip) atomicAdd(S.skipped, 1u);
}

// after the Metropolis step: the trial just evaluated (the proposal, its exact chi^2) goes
// into ring slot ((step - 1) chains + c) mod cap with target log Rwp =
// (log chi^2 - log sum w y^2) / 2; NaN (left out of training) for an infeasible trial or
// a chain outside the screened ones
__global__ void __launch_bounds__(64) k_screen_record(DevPlan P, ScreenDev S) {
    const int c = blockIdx.x, i = threadIdx.x, nv = P.nv;
    const double *st = P.state + (size_t)c * P.stride;
    const long long step = (long long)st[3 * nv + 2];
    if (step < 1 || S.pick[c].x < 0.0) return;     // a skipped proposal has no evaluation
    const size_t pos = (size_t)(((step - 1) * P.chains + c) % S.cap);
    uint16_t u = 0;
    if (i < nv) u = screen_input(S, i, st[nv + i]);
    S.ring_x[pos * MLP_IN + i] = u;
    if (i == 0) {
        const double cost = st[3 * nv + 10];
        const bool ok = c < S.cold_chains && st[3 * nv + 7] > 0.5 && cost > 0.0 && cost < 3e38;
        S.ring_t[pos] = ok ? (float)(0.5 * (log(cost) - S.log_wss)) : __int_as_float(0x7fc00000);
    }
}

// ---- launchers ------------------------------------------------------------------------
hipError_t mlp_to_bf16(const float *p, uint16_t *wb, int n, hipStream_t s) {
    hipLaunchKernelGGL(k_mlp_to_bf16, dim3((n + 255) / 256), dim3(256), 0, s, p, wb, n);
    return hipGetLastError();
}
hipError_t mlp_pack(const float *x, int n, int n_in, uint16_t *x16, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_mlp_pack, dim3(n), dim3(64), 0, s, x, n_in, x16);
    return hipGetLastError();
}
hipError_t mlp_fused(const MlpRun &R, bool train, hipStream_t s) {
    static std::mutex mu;
    static bool raised = false;
    const size_t bytes = sizeof(FusedLds);
    {
        std::lock_guard<std::mutex> lock(mu);
        if (!raised) {
            hipError_t e = hipFuncSetAttribute((const void *)k_mlp_fused<true>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
            if (e == hipSuccess)
                e = hipFuncSetAttribute((const void *)k_mlp_fused<false>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
            if (e != hipSuccess) return e;
            raised = true;
        }
    }
    const dim3 grid((R.rows + kTR - 1) / kTR);
    if (train) hipLaunchKernelGGL(k_mlp_fused<true>, grid, dim3(256), bytes, s, R);
    else hipLaunchKernelGGL(k_mlp_fused<false>, grid, dim3(256), bytes, s, R);
    return hipGetLastError();
}
hipError_t mlp_reduce_adam(const float *part, int rows, float inv_n, float *g, float *loss, int adam, float *p, float *m, float *v,
                           uint16_t *wb, uint16_t *wbt, float lr, float b1, float b2, float eps, const int *iter, const int *active,
                           hipStream_t s) {
    const int nblk = (rows + kTR - 1) / kTR;
    hipLaunchKernelGGL(k_mlp_reduce_adam, dim3((MLP_PARAMS + 1 + 255) / 256), dim3(256), 0, s, part, nblk, inv_n, g, loss, adam, p, m,
                       v, wb, wbt, lr, b1, b2, eps, iter, active);
    if (adam) hipLaunchKernelGGL(k_mlp_tick, dim3(1), dim3(64), 0, s, (int *)iter, active);
    return hipGetLastError();
}
hipError_t mlp_to_bf16_t(const float *p, uint16_t *wbt, int n, hipStream_t s) {
    hipLaunchKernelGGL(k_mlp_to_bf16_t, dim3((n + 255) / 256), dim3(256), 0, s, p, wbt, n);
    return hipGetLastError();
}
hipError_t screen_candidates(const DevPlan &P, const ScreenDev &S, hipStream_t s) {
    hipLaunchKernelGGL(k_screen_candidates, dim3(P.chains * (S.K + 1)), dim3(64), 0, s, P, S);
    return hipGetLastError();
}
hipError_t screen_inputs(const double *x, int n, int nv, const ScreenDev &S, uint16_t *xs, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_screen_inputs, dim3(n), dim3(64), 0, s, x, nv, S, xs);
    return hipGetLastError();
}
hipError_t screen_pick(const DevPlan &P, const ScreenDev &S, const int *iter, hipStream_t s) {
    hipLaunchKernelGGL(k_screen_pick, dim3((P.chains + 63) / 64), dim3(64), 0, s, P, S, iter);
    return hipGetLastError();
}
hipError_t screen_record(const DevPlan &P, const ScreenDev &S, hipStream_t s) {
    hipLaunchKernelGGL(k_screen_record, dim3(P.chains), dim3(64), 0, s, P, S);
    return hipGetLastError();
}

}  // namespace ptk
