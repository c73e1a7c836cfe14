// mlp.h -- the per-atom MLP of the learned CG initial guess: BF16 inference on the matrix cores
// (v_mfma_f32_16x16x16_bf16, FP32 accumulation) and FP32 training kernels (forward with saved
// activations, backward, Adam).
//
// Network: h_0 = x (D0 = 16 S0 inputs, zero padded), h_l = silu(h_{l-1} W_l + b_l) for the
// hidden layers of width 128, y = h_L W_out + b_out (DOUT <= 16 outputs, padded to 16).
//
// Inference layout.  Atoms are the N dimension of every product: with A = W^T (16 x 16 tiles of
// the weights) and B = h^T, lane i of a wave holds, for atom i % 16, the four consecutive features
// 4 (i / 16) + j of each 16-wide feature tile -- which is both the B-operand layout of the
// 16x16x16 MFMA and the layout of its result, so layers chain in registers with no transpose.
// The weights are packed once per update into that A-operand order (pack index
// ((t * S + s) * 64 + lane) * 4 + j for output tile t, input tile s) and staged in LDS.
#pragma once

#include "common.h"

namespace pgm {

constexpr int MLP_W = 128;      // hidden width
constexpr int MLP_T = MLP_W / 16;

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

PGM_DEV short to_bf16(float f) {  // round to nearest even (finite inputs)
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (short)(u >> 16);
}
PGM_DEV float silu_f(float z) { return z / (1.0f + __expf(-z)); }

// Pack W (in x out, FP32 row-major, in <= 16 S, out <= 16 T) into the A-operand order (BF16),
// zero padded: Wp[((t * S + s) * 64 + lane) * 4 + j] = W[16 s + 4 (lane / 16) + j][16 t + lane % 16].
__global__ void k_mlp_pack(int in, int out, int S, int T, const float *__restrict__ W, short *__restrict__ Wp) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;  // over T * S * 64 * 4
    if (idx >= T * S * 256) return;
    const int j = idx & 3, lane = (idx >> 2) & 63, ts = idx >> 8;
    const int s = ts % S, t = ts / S;
    const int k = 16 * s + 4 * (lane / 16) + j, m = 16 * t + lane % 16;
    Wp[idx] = (k < in && m < out) ? to_bf16(W[(size_t)k * out + m]) : (short)0;
}

// Inference: y (n x dout, FP32) = MLP(x (n x d0, FP32)).  Layers: [d0 -> 128] silu, NH - 1 times
// [128 -> 128] silu, [128 -> dout].  S0 = ceil(d0 / 16) <= 8 input tiles; wp: packed weights of
// all layers back to back, bias: FP32 biases back to back (128 per hidden layer, then dout).
// 256 threads = 4 waves; each wave takes 16 atoms at a time.  LDS: every packed weight
// (mlp_lds_bytes<S0, NH>()).
template <int S0, int NH> constexpr int mlp_packed_shorts() {
    return MLP_T * S0 * 256 + (NH - 1) * MLP_T * MLP_T * 256 + MLP_T * 256;
}

template <int S0, int NH>
__global__ __launch_bounds__(256) void k_mlp_infer(int n, int d0, int dout, const short *__restrict__ wp,
                                                   const float *__restrict__ bias, const float *__restrict__ x,
                                                   float *__restrict__ y) {
    extern __shared__ __align__(16) unsigned char mlp_lds[];
    constexpr int total = mlp_packed_shorts<S0, NH>();
    short *w = (short *)mlp_lds;
    for (int k = threadIdx.x * 4; k < total; k += blockDim.x * 4) *(bf16x4 *)(w + k) = *(const bf16x4 *)(wp + k);
    __syncthreads();
    const int lane = threadIdx.x % WAVE, wave = threadIdx.x / WAVE;
    const int r = lane % 16, q = lane / 16;
    const int ntile = (n + 15) / 16;
    for (int tile = blockIdx.x * 4 + wave; tile < ntile; tile += gridDim.x * 4) {
        const int atom = tile * 16 + r;
        bf16x4 hin[MLP_T];  // features 16 s + 4 q + j of this lane's atom
#pragma unroll
        for (int s = 0; s < MLP_T; ++s) {
            bf16x4 v;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int f = 16 * s + 4 * q + j;
                v[j] = to_bf16(s < S0 && atom < n && f < d0 ? x[(size_t)atom * d0 + f] : 0.0f);
            }
            hin[s] = v;
        }
        const short *wl = w;
        const float *bl = bias;
#pragma unroll
        for (int layer = 0; layer < NH; ++layer) {
            const int S = layer == 0 ? S0 : MLP_T;
            f32x4 acc[MLP_T];
#pragma unroll
            for (int t = 0; t < MLP_T; ++t) {
                f32x4 c;
#pragma unroll
                for (int j = 0; j < 4; ++j) c[j] = bl[16 * t + 4 * q + j];
#pragma unroll
                for (int s = 0; s < MLP_T; ++s)
                    if (s < S) {
                        const bf16x4 a = *(const bf16x4 *)(wl + ((t * S + s) * 64 + lane) * 4);
                        c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, hin[s], c, 0, 0, 0);
                    }
                acc[t] = c;
            }
#pragma unroll
            for (int t = 0; t < MLP_T; ++t) {
                bf16x4 v;
#pragma unroll
                for (int j = 0; j < 4; ++j) v[j] = to_bf16(silu_f(acc[t][j]));
                hin[t] = v;
            }
            wl += MLP_T * S * 256;
            bl += MLP_W;
        }
        f32x4 c;
#pragma unroll
        for (int j = 0; j < 4; ++j) c[j] = (4 * q + j < dout) ? bl[4 * q + j] : 0.0f;
#pragma unroll
        for (int s = 0; s < MLP_T; ++s) {
            const bf16x4 a = *(const bf16x4 *)(wl + (s * 64 + lane) * 4);
            c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, hin[s], c, 0, 0, 0);
        }
        if (atom < n)
#pragma unroll
            for (int j = 0; j < 4; ++j)
                if (4 * q + j < dout) y[(size_t)atom * dout + 4 * q + j] = c[j];
    }
}

// ---------------------------------------------------------------------------- FP32 training
// C (M x N) = op(A) op(B) (+ bias row) with element (m, k) of op(A) at A[m * sam + k * sak] and
// (k, n) of op(B) at B[k * sbk + n * sbn]; 16 x 16 output tile per 256-thread block, K in LDS
// chunks of 16, fixed summation order (deterministic).  act: 0 none, 1 silu (z kept in Z).
__global__ __launch_bounds__(256) void k_gemm(int M, int N, int K, const float *__restrict__ A, long sam, long sak,
                                              const float *__restrict__ B, long sbk, long sbn, const float *__restrict__ bias,
                                              float *__restrict__ C, float *__restrict__ Z, int act) {
    __shared__ float As[16][17], Bs[16][17];
    const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
    const int m = blockIdx.y * 16 + ty, nn = blockIdx.x * 16 + tx;
    float acc = 0.0f;
    for (int k0 = 0; k0 < K; k0 += 16) {
        const int ka = k0 + tx, kb = k0 + ty;
        As[ty][tx] = (m < M && ka < K) ? A[m * sam + ka * sak] : 0.0f;
        Bs[ty][tx] = (kb < K && nn < N) ? B[kb * sbk + nn * sbn] : 0.0f;
        __syncthreads();
#pragma unroll
        for (int k = 0; k < 16; ++k) acc += As[ty][k] * Bs[k][tx];
        __syncthreads();
    }
    if (m < M && nn < N) {
        if (bias) acc += bias[nn];
        if (Z) Z[(size_t)m * N + nn] = acc;
        C[(size_t)m * N + nn] = act == 1 ? silu_f(acc) : acc;
    }
}

// delta (M x N) *= silu'(Z)
__global__ void k_dsilu_mul(size_t count, const float *__restrict__ Z, float *__restrict__ delta) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const float z = Z[i];
    const float s = 1.0f / (1.0f + __expf(-z));
    delta[i] *= s * (1.0f + z * (1.0f - s));
}

// delta = 2 (y - t) / (M N) and per-block partial sums of (y - t)^2 (FP64).
__global__ __launch_bounds__(256) void k_mse_grad(size_t count, const float *__restrict__ y, const float *__restrict__ t,
                                                  float *__restrict__ delta, double *__restrict__ part) {
    __shared__ double lds[16];
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    double e[1] = {0.0};
    if (i < count) {
        const float d = y[i] - t[i];
        delta[i] = 2.0f * d / (float)count;
        e[0] = (double)d * d;
    }
    block_sum<1>(e, lds);
    if (threadIdx.x == 0) part[blockIdx.x] = e[0];
}

// db[n] = sum_m delta[m][n]: one 256-thread block per column n, strided rows, FP64 block sum
// (fixed order).
__global__ __launch_bounds__(256) void k_colsum(int M, int N, const float *__restrict__ delta, float *__restrict__ db) {
    __shared__ double lds[16];
    const int nn = blockIdx.x;
    double v[1] = {0.0};
    for (int m = threadIdx.x; m < M; m += blockDim.x) v[0] += (double)delta[(size_t)m * N + nn];
    block_sum<1>(v, lds);
    if (threadIdx.x == 0) db[nn] = (float)v[0];
}

// Adam (Kingma and Ba, 2015) on `count` parameters: c1 = 1 - b1^step, c2 = 1 - b2^step.
__global__ void k_adam(size_t count, float *__restrict__ w, const float *__restrict__ g, float *__restrict__ m,
                       float *__restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

}  // namespace pgm
