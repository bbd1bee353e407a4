// pt_mlp.hip -- gfx950 kernels of the surrogate network and of surrogate-screened PT.
//
//   k_mlp_gemm<TA, TB, EPI>  C = A B on BF16 MFMA (16x16x16, FP32 accumulation), 64x64
//                            output tile per workgroup, K staged through LDS in 32-wide
//                            slices; epilogues: plain FP32 store (weight gradients), bias
//                            + SiLU (forward layers), times SiLU'(z) with per-tile column
//                            sums for the bias gradients (backward layers).
//   k_mlp_head               y = a3 . w4 + b4 (64 rows per workgroup); in training also
//                            dz3 = dy w4 SiLU'(z3), dy = 2 m (y - t) / N, and partial
//                            sums of the b3 / w4 / b4 gradients and the loss.
//   k_mlp_bias_grads         ordered sums of the per-row-block partials (deterministic).
//   k_mlp_adam / k_mlp_tick  Adam on the FP32 master weights, BF16 copies refreshed.
//   k_mlp_gather / k_mlp_pack  minibatch rows from the replay ring / host rows to BF16.
//   k_screen_*               the screened PT step: K candidate moves per chain, their
//                            predicted log Rwp, the pick, and the record of every exactly
//                            evaluated trial into the replay ring.
#include "pt_device.h"
#include "pt_mlp.h"

namespace ptk {

typedef short bf16x4_t __attribute__((ext_vector_type(4)));
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

enum { EPI_STORE = 0, EPI_BIAS_SILU = 1, EPI_DSILU = 2 };

// four consecutive BF16 values of a row (8-byte load when in range and aligned)
__device__ __forceinline__ uint2 load4(const uint16_t *__restrict__ p, int avail) {
    if (avail >= 4) return *reinterpret_cast<const uint2 *>(p);
    uint16_t v[4] = {0, 0, 0, 0};
    for (int u = 0; u < avail; ++u) v[u] = p[u];
    return make_uint2((uint32_t)v[0] | ((uint32_t)v[1] << 16), (uint32_t)v[2] | ((uint32_t)v[3] << 16));
}

// C[M x N] = A[M x K] B[K x N].  A(m, k) = TA ? A[k lda + m] : A[m lda + k];
// B(k, n) = TB ? B[n ldb + k] : B[k ldb + n]; leading dimensions are multiples of 4.
// Four wavefronts, each a 32 x 32 quarter of the 64 x 64 tile as 2 x 2 MFMA blocks; the
// operands are staged in LDS as sA[m][k] and sB[n][k] with 8-byte global loads.
// Operand layout of v_mfma_f32_16x16x16_bf16 for lane l (i = l % 16, q = l / 16):
// a[e] = A(i, 4q + e), b[e] = B(4q + e, i), and the result d[e] = C(4q + e, i).
// With ep.colpart set (backward epilogue) the workgroup also writes the column sums of
// its 64 rows of dz to colpart[blockIdx.y][n] (the bias gradient, reduced in order later).
template <bool TA, bool TB, int EPI>
__global__ void __launch_bounds__(256) k_mlp_gemm(const uint16_t *__restrict__ A, int lda, const uint16_t *__restrict__ B,
                                                  int ldb, int M, int N, int K, GemmEpi ep) {
    __shared__ __align__(16) uint16_t sA[64][40];   // [m][k]
    __shared__ __align__(16) uint16_t sB[64][40];   // [n][k]
    __shared__ float csum[2][64];
    const int m0 = blockIdx.y * 64, n0 = blockIdx.x * 64;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    const int li = lane & 15, lq = lane >> 4;
    f32x4_t acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = f32x4_t{0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < K; k0 += 32) {
#pragma unroll
        for (int r = 0; r < 2; ++r) {
            const int e = tid + 256 * r;
            if (!TA) {          // rows of A along k: 64 rows x 8 quads
                const int mm = e >> 3, kq = (e & 7) * 4, gm = m0 + mm, gk = k0 + kq;
                uint2 v = make_uint2(0u, 0u);
                if (gm < M && gk < K) v = load4(A + (size_t)gm * lda + gk, K - gk);
                *reinterpret_cast<uint2 *>(&sA[mm][kq]) = v;
            } else {            // A stored k-major: 32 k x 16 quads of m, transposed into LDS
                const int kk = e >> 4, mq = (e & 15) * 4, gm = m0 + mq, gk = k0 + kk;
                uint2 v = make_uint2(0u, 0u);
                if (gm < M && gk < K) v = load4(A + (size_t)gk * lda + gm, M - gm);
                sA[mq][kk] = (uint16_t)(v.x & 0xffffu);
                sA[mq + 1][kk] = (uint16_t)(v.x >> 16);
                sA[mq + 2][kk] = (uint16_t)(v.y & 0xffffu);
                sA[mq + 3][kk] = (uint16_t)(v.y >> 16);
            }
            if (TB) {           // B stored n-major (k contiguous)
                const int nn = e >> 3, kq = (e & 7) * 4, gn = n0 + nn, gk = k0 + kq;
                uint2 v = make_uint2(0u, 0u);
                if (gn < N && gk < K) v = load4(B + (size_t)gn * ldb + gk, K - gk);
                *reinterpret_cast<uint2 *>(&sB[nn][kq]) = v;
            } else {            // B stored k-major (n contiguous), transposed into LDS
                const int kk = e >> 4, nq = (e & 15) * 4, gn = n0 + nq, gk = k0 + kk;
                uint2 v = make_uint2(0u, 0u);
                if (gn < N && gk < K) v = load4(B + (size_t)gk * ldb + gn, N - gn);
                sB[nq][kk] = (uint16_t)(v.x & 0xffffu);
                sB[nq + 1][kk] = (uint16_t)(v.x >> 16);
                sB[nq + 2][kk] = (uint16_t)(v.y & 0xffffu);
                sB[nq + 3][kk] = (uint16_t)(v.y >> 16);
            }
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < 32; kk += 16) {
            bf16x4_t a[2], b[2];
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                a[i] = *reinterpret_cast<const bf16x4_t *>(&sA[wm + 16 * i + li][kk + 4 * lq]);
                b[i] = *reinterpret_cast<const bf16x4_t *>(&sB[wn + 16 * i + li][kk + 4 * lq]);
            }
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[i], b[j], acc[i][j], 0, 0, 0);
        }
        __syncthreads();
    }
    float cs[2] = {0.f, 0.f};
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int m = m0 + wm + 16 * i + 4 * lq + e, n = n0 + wn + 16 * j + li;
                if (m >= M || n >= N) continue;
                const size_t o = (size_t)m * ep.ldc + n;
                const float v = acc[i][j][e];
                if (EPI == EPI_STORE) {
                    ep.c32[o] = v;
                } else if (EPI == EPI_BIAS_SILU) {
                    const float z = v + ep.bias[n];
                    ep.zout[o] = z;
                    ep.c16[o] = to_bf16(silu(z));
                } else {
                    const float d = v * dsilu(ep.z32[o]);
                    ep.c32[o] = d;
                    ep.c16[o] = to_bf16(d);
                    cs[j] += d;
                }
            }
    if (EPI == EPI_DSILU && ep.colpart) {
        // column sums: over e and i in the lane, over the four lane quads, then the two
        // wavefronts that share the columns
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            cs[j] += __shfl_xor(cs[j], 16);
            cs[j] += __shfl_xor(cs[j], 32);
        }
        if (lq == 0) {
            csum[wave >> 1][wn + li] = cs[0];
            csum[wave >> 1][wn + 16 + li] = cs[1];
        }
        __syncthreads();
        if (tid < 64 && n0 + tid < N) ep.colpart[(size_t)blockIdx.y * N + n0 + tid] = csum[0][tid] + csum[1][tid];
    }
}

// y = a3 . w4 + b4 for 64 rows per workgroup (16 per wavefront, lanes over the hidden
// units); with t != null also the head's backward pass -- dy = 2 m (y - t) / N,
// dz3 = dy w4 SiLU'(z3) (FP32 and BF16) -- and this workgroup's partial sums of the b3
// and w4 gradients (part[blk][0..255], part[blk][256..511]) and of b4 and the loss
// (part[blk][512], part[blk][513]), reduced in order by k_mlp_bias_grads
__global__ void __launch_bounds__(256) k_mlp_head(const uint16_t *__restrict__ a3, const float *__restrict__ z3,
                                                  const uint16_t *__restrict__ w4b, const float *__restrict__ b4, int M,
                                                  float *y, const float *__restrict__ t, const float *__restrict__ mask,
                                                  float inv_n, float *dz32, uint16_t *dz16, float *part) {
    __shared__ float red[4][2 * MLP_H + 2];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    float w4[MLP_H / 64];
#pragma unroll
    for (int q = 0; q < MLP_H / 64; ++q) w4[q] = from_bf16(w4b[lane + 64 * q]);
    float gb3[MLP_H / 64] = {0.f, 0.f, 0.f, 0.f}, gw4[MLP_H / 64] = {0.f, 0.f, 0.f, 0.f};
    float gb4 = 0.f, gl = 0.f;
    for (int rr = 0; rr < 16; ++rr) {
        const int row = blockIdx.x * 64 + wave * 16 + rr;
        if (row >= M) break;
        float av[MLP_H / 64];
        float s = 0.f;
#pragma unroll
        for (int q = 0; q < MLP_H / 64; ++q) {
            av[q] = from_bf16(a3[(size_t)row * MLP_H + lane + 64 * q]);
            s = fmaf(av[q], w4[q], s);
        }
        for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
        const float yv = s + b4[0];
        if (lane == 0) y[row] = yv;
        if (!t) continue;
        const float m = mask[row], e = yv - t[row];
        const float d = 2.0f * m * e * inv_n;
        gb4 += d;
        gl += m * e * e;
#pragma unroll
        for (int q = 0; q < MLP_H / 64; ++q) {
            const size_t o = (size_t)row * MLP_H + lane + 64 * q;
            const float g = d * w4[q] * dsilu(z3[o]);
            dz32[o] = g;
            dz16[o] = to_bf16(g);
            gb3[q] += g;
            gw4[q] = fmaf(d, av[q], gw4[q]);
        }
    }
    if (!t) return;
#pragma unroll
    for (int q = 0; q < MLP_H / 64; ++q) {
        red[wave][lane + 64 * q] = gb3[q];
        red[wave][MLP_H + lane + 64 * q] = gw4[q];
    }
    if (lane == 0) { red[wave][2 * MLP_H] = gb4; red[wave][2 * MLP_H + 1] = gl; }
    __syncthreads();
    for (int i = threadIdx.x; i < 2 * MLP_H + 2; i += blockDim.x)
        part[(size_t)blockIdx.x * (2 * MLP_H + 2) + i] = red[0][i] + red[1][i] + red[2][i] + red[3][i];
}

// gradients of b1, b2 (column partials of the backward GEMMs), b3, w4, b4 and the loss
// (head partials): ordered sums over the nblk row blocks, so the result is deterministic
__global__ void __launch_bounds__(256) k_mlp_bias_grads(const float *__restrict__ pb1, const float *__restrict__ pb2,
                                                        const float *__restrict__ ph, int nblk, float inv_n, float *g,
                                                        float *loss) {
    const int i = threadIdx.x;
    float s1 = 0.f, s2 = 0.f, s3 = 0.f, s4 = 0.f;
    for (int b = 0; b < nblk; ++b) {
        s1 += pb1[(size_t)b * MLP_H + i];
        s2 += pb2[(size_t)b * MLP_H + i];
        s3 += ph[(size_t)b * (2 * MLP_H + 2) + i];
        s4 += ph[(size_t)b * (2 * MLP_H + 2) + MLP_H + i];
    }
    g[MLP_OFF_B1 + i] = s1;
    g[MLP_OFF_B2 + i] = s2;
    g[MLP_OFF_B3 + i] = s3;
    g[MLP_OFF_W4 + i] = s4;
    if (i == 0) {
        float b = 0.f, l = 0.f;
        for (int k = 0; k < nblk; ++k) {
            b += ph[(size_t)k * (2 * MLP_H + 2) + 2 * MLP_H];
            l += ph[(size_t)k * (2 * MLP_H + 2) + 2 * MLP_H + 1];
        }
        g[MLP_OFF_B4] = b;
        loss[0] = l * inv_n;
        // loss[1]: running mean over recent minibatches (weight 1/64), seeded by the first
        loss[1] = loss[1] > 0.f ? loss[1] + (loss[0] - loss[1]) * (1.0f / 64.0f) : loss[0];
    }
}

// Adam on the FP32 master parameters (step t = iter + 1), BF16 copy refreshed; nothing
// when the minibatch was inactive (ring not yet filled)
__global__ void __launch_bounds__(256) k_mlp_adam(float *p, const float *__restrict__ g, float *m, float *v, uint16_t *wb, int n,
                                                  float lr, float b1, float b2, float eps, const int *iter, const int *active) {
    if (!*active) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float t = (float)(*iter + 1);
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float mh = mi / (1.0f - powf(b1, t)), vh = vi / (1.0f - powf(b2, t));
    const float pi = p[i] - lr * mh / (sqrtf(vh) + eps);
    p[i] = pi;
    wb[i] = to_bf16(pi);
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

// minibatch row r of B from the replay ring: index hash(seed, iteration, r) mod count,
// where count = min(cap, step * chains) (ring fed by the PT steps) or fixed_count;
// inactive (all rows masked) while count < min_fill.  Entries with a NaN target
// (infeasible trials) are masked out.
__global__ void __launch_bounds__(64) k_mlp_gather(const uint16_t *__restrict__ ring_x, const float *__restrict__ ring_t,
                                                   int cap, const double *step_ptr, int chains, int fixed_count,
                                                   int min_fill, uint32_t seed, const int *iter, uint16_t *xb, float *tb,
                                                   float *mb, int *active) {
    long long count = fixed_count;
    if (step_ptr) count = (long long)fmax(0.0, *step_ptr) * chains;
    if (count > cap) count = cap;
    const bool act = count >= min_fill && count > 0;
    const int r = blockIdx.x, i = threadIdx.x;
    if (r == 0 && i == 0) *active = act ? 1 : 0;
    if (!act) {
        xb[(size_t)r * MLP_IN + i] = 0;
        if (i == 0) { tb[r] = 0.f; mb[r] = 0.f; }
        return;
    }
    const double u = uniform01(seed, r, (unsigned long long)(*iter), 11);
    long long idx = (long long)(u * (double)count);
    if (idx >= count) idx = count - 1;
    xb[(size_t)r * MLP_IN + i] = ring_x[(size_t)idx * MLP_IN + i];
    if (i == 0) {
        const float tv = ring_t[idx];
        const bool ok = tv == tv;
        tb[r] = ok ? tv : 0.f;
        mb[r] = ok ? 1.f : 0.f;
    }
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
template <bool TA, bool TB, int EPI>
static hipError_t gemm(const uint16_t *A, int lda, const uint16_t *B, int ldb, int M, int N, int K, const GemmEpi &ep,
                       hipStream_t s) {
    hipLaunchKernelGGL((k_mlp_gemm<TA, TB, EPI>), dim3((N + 63) / 64, (M + 63) / 64), dim3(256), 0, s, A, lda, B, ldb, M, N, K, ep);
    return hipGetLastError();
}

hipError_t mlp_gemm_forward(const uint16_t *A, const uint16_t *W, int M, int K, const float *bias, float *z, uint16_t *a,
                            hipStream_t s) {
    GemmEpi ep{};
    ep.bias = bias; ep.zout = z; ep.c16 = a; ep.ldc = MLP_H;
    return gemm<false, false, EPI_BIAS_SILU>(A, K, W, MLP_H, M, MLP_H, K, ep, s);
}
hipError_t mlp_gemm_dact(const uint16_t *dz, const uint16_t *W, int M, const float *z, float *d32, uint16_t *d16, float *colpart,
                         hipStream_t s) {
    GemmEpi ep{};
    ep.z32 = z; ep.c32 = d32; ep.c16 = d16; ep.ldc = MLP_H; ep.colpart = colpart;
    return gemm<false, true, EPI_DSILU>(dz, MLP_H, W, MLP_H, M, MLP_H, MLP_H, ep, s);
}
hipError_t mlp_gemm_wgrad(const uint16_t *a, int a_cols, const uint16_t *dz, int M, float *gW, hipStream_t s) {
    GemmEpi ep{};
    ep.c32 = gW; ep.ldc = MLP_H;
    return gemm<true, false, EPI_STORE>(a, a_cols, dz, MLP_H, a_cols, MLP_H, M, ep, s);
}
hipError_t mlp_head(const uint16_t *a3, const float *z3, const uint16_t *w4b, const float *b4, int M, float *y, const float *t,
                    const float *mask, float inv_n, float *dz32, uint16_t *dz16, float *part, hipStream_t s) {
    hipLaunchKernelGGL(k_mlp_head, dim3((M + 63) / 64), dim3(256), 0, s, a3, z3, w4b, b4, M, y, t, mask, inv_n, dz32, dz16, part);
    return hipGetLastError();
}
hipError_t mlp_bias_grads(const float *pb1, const float *pb2, const float *ph, int nblk, float inv_n, float *g, float *loss,
                          hipStream_t s) {
    hipLaunchKernelGGL(k_mlp_bias_grads, dim3(1), dim3(MLP_H), 0, s, pb1, pb2, ph, nblk, inv_n, g, loss);
    return hipGetLastError();
}
hipError_t mlp_adam(float *p, const float *g, float *m, float *v, uint16_t *wb, int n, float lr, float b1, float b2, float eps,
                    const int *iter, const int *active, hipStream_t s) {
    hipLaunchKernelGGL(k_mlp_adam, dim3((n + 255) / 256), dim3(256), 0, s, p, g, m, v, wb, n, lr, b1, b2, eps, iter, active);
    hipLaunchKernelGGL(k_mlp_tick, dim3(1), dim3(64), 0, s, (int *)iter, active);
    return hipGetLastError();
}
hipError_t mlp_to_bf16(const float *p, uint16_t *wb, int n, hipStream_t s) {
    hipLaunchKernelGGL(k_mlp_to_bf16, dim3((n + 255) / 256), dim3(256), 0, s, p, wb, n);
    return hipGetLastError();
}
hipError_t mlp_pack(const float *x, int n, int n_in, uint16_t *x16, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_mlp_pack, dim3(n), dim3(64), 0, s, x, n_in, x16);
    return hipGetLastError();
}
hipError_t mlp_gather(const uint16_t *ring_x, const float *ring_t, int cap, const double *step_ptr, int chains, int fixed_count,
                      int min_fill, uint32_t seed, const int *iter, int B, uint16_t *xb, float *tb, float *mb, int *active,
                      hipStream_t s) {
    hipLaunchKernelGGL(k_mlp_gather, dim3(B), dim3(64), 0, s, ring_x, ring_t, cap, step_ptr, chains, fixed_count, min_fill, seed,
                       iter, xb, tb, mb, active);
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
