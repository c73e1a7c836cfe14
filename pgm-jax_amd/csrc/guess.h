// guess.h -- the learned CG initial guess for rigid 3-site molecules (water): per-atom features in
// a local frame, the guess assembled from the network output, and the harvest of training rows.
//
// Each atom gets an orthonormal frame from its molecule at the current step (rows e1, e2, e3):
//   O:  e1 = unit((h1 - o) + (h2 - o)), e2 = unit(h1 - h2 made orthogonal to e1), e3 = e1 x e2;
//   Hk: e1 = unit(hk - o), e2 = unit((h_other - o) made orthogonal to e1), e3 = e1 x e2.
// The guess is a linear predictor over the last GUESS_K converged dipoles, its coefficients fitted
// by least squares to the dipoles the solves produce, plus a per-atom network correction:
//   x0 = sum_k w_k mu_{t-1-k} + sigma R^T y.
// Network features (GUESS_D = 16 per atom): the last four converged dipoles rotated into the frame
// and divided by the dipole scale (12), 1 for oxygen / 0 for hydrogen, zeros; its target is the
// linear predictor's error in the frame, in units of that error's RMS (sigma).
#pragma once

#include "common.h"

namespace pgm {

constexpr int GUESS_D = 16;
constexpr int GUESS_OUT = 3;

PGM_DEV void unit3(double *v) {
    const double s = 1.0 / sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    v[0] *= s;
    v[1] *= s;
    v[2] *= s;
}

// frame rows of atom a (0 O, 1 H1, 2 H2) of the molecule with atoms at o, h1, h2
PGM_DEV void water_frame(const double *o, const double *h1, const double *h2, int a, double (&R)[9]) {
    double e1[3], e2[3];
    if (a == 0) {
        for (int k = 0; k < 3; ++k) {
            e1[k] = (h1[k] - o[k]) + (h2[k] - o[k]);
            e2[k] = h1[k] - h2[k];
        }
    } else {
        const double *hh = a == 1 ? h1 : h2, *ho = a == 1 ? h2 : h1;
        for (int k = 0; k < 3; ++k) {
            e1[k] = hh[k] - o[k];
            e2[k] = ho[k] - o[k];
        }
    }
    unit3(e1);
    const double d = e1[0] * e2[0] + e1[1] * e2[1] + e1[2] * e2[2];
    for (int k = 0; k < 3; ++k) e2[k] -= d * e1[k];
    unit3(e2);
    R[0] = e1[0], R[1] = e1[1], R[2] = e1[2];
    R[3] = e2[0], R[4] = e2[1], R[5] = e2[2];
    R[6] = e1[1] * e2[2] - e1[2] * e2[1];
    R[7] = e1[2] * e2[0] - e1[0] * e2[2];
    R[8] = e1[0] * e2[1] - e1[1] * e2[0];
}

// The guess keeps its own ring of the last GUESS_K converged dipoles (slot (head - 1 - k) mod
// GUESS_K holds mu_{t-1-k}); ring_at returns component c of atom i of mu_{t-1-k}.
constexpr int GUESS_K = 8;
PGM_DEV double ring_at(const double *ring, int n, int head, int k, int i, int c) {
    const int slot = (head - 1 - k + 2 * GUESS_K) % GUESS_K;
    return ring[((size_t)slot * n + i) * 3 + c];
}

// ring slot `slot` = mu (3 n doubles)
__global__ void k_ring_push(int n3, const double *__restrict__ mu, double *__restrict__ ring, int slot) {
    const int o = blockIdx.x * blockDim.x + threadIdx.x;
    if (o < n3) ring[(size_t)slot * n3 + o] = mu[o];
}

// Features (n x GUESS_D, FP32) and frames (n x 9, FP64) at positions pos from the four newest
// ring entries; inv_scale = 1 / dipole scale.  One thread per atom; n = 3 nmol.
__global__ void k_guess_features(int n, const double *__restrict__ pos, const double *__restrict__ ring, int head,
                                 double inv_scale, float *__restrict__ feat, double *__restrict__ frames) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int m = i / 3, a = i - 3 * m;
    const double *o = pos + 9 * m;
    double R[9];
    water_frame(o, o + 3, o + 6, a, R);
#pragma unroll
    for (int k = 0; k < 9; ++k) frames[(size_t)9 * i + k] = R[k];
    float *f = feat + (size_t)GUESS_D * i;
#pragma unroll
    for (int h = 0; h < 4; ++h) {
        const double mu0 = ring_at(ring, n, head, h, i, 0), mu1 = ring_at(ring, n, head, h, i, 1), mu2 = ring_at(ring, n, head, h, i, 2);
#pragma unroll
        for (int r = 0; r < 3; ++r) f[3 * h + r] = (float)(inv_scale * (R[3 * r] * mu0 + R[3 * r + 1] * mu1 + R[3 * r + 2] * mu2));
    }
    f[12] = a == 0 ? 1.0f : 0.0f;
    f[13] = f[14] = f[15] = 0.0f;
}

struct GuessCoef {
    double w[GUESS_K];  // linear predictor over the ring, newest first
};

PGM_DEV void guess_linear(const double *ring, int n, int head, const GuessCoef &cf, int i, double (&v)[3]) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        double s = 0.0;
#pragma unroll
        for (int k = 0; k < GUESS_K; ++k) s += cf.w[k] * ring_at(ring, n, head, k, i, c);
        v[c] = s;
    }
}

// x0 = sum_k w_k mu_{t-1-k} (+ sigma R^T y with use_net), written as real4 (d, 0).
template <typename real>
__global__ void k_guess_x0(int n, const double *__restrict__ ring, int head, GuessCoef cf, const double *__restrict__ frames,
                           const float *__restrict__ y, double sigma, int use_net, real4<real> *__restrict__ x0) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double v[3];
    guess_linear(ring, n, head, cf, i, v);
    if (use_net) {
        const double *R = frames + (size_t)9 * i;
#pragma unroll
        for (int k = 0; k < 3; ++k)
            v[k] += sigma * (R[k] * y[3 * (size_t)i] + R[3 + k] * y[3 * (size_t)i + 1] + R[6 + k] * y[3 * (size_t)i + 2]);
    }
    real4<real> o;
    o.x = (real)v[0];
    o.y = (real)v[1];
    o.z = (real)v[2];
    o.w = 0;
    x0[i] = o;
}

// After the solve of step t (mu = mu_t, the ring still ends at mu_{t-1}), per block:
//   [0..35]  the Gram matrix sum mu_{t-1-j} . mu_{t-1-k} (upper triangle, j <= k),
//   [36..43] sum mu_{t-1-j} . mu_t,
//   [44] |mu_t - lin|^2, [45] |mu_t - mu4|^2, [46] |mu_t - lin - sigma R^T y|^2 (with have_y),
// and, with row0 >= 0, the training rows (features, R (mu_t - lin) / sigma) at (row0 + i) % cap.
constexpr int GUESS_STATS = 47;
__global__ __launch_bounds__(256) void k_guess_stats(int n, const double *__restrict__ ring, int head, GuessCoef lin, GuessCoef mu4,
                                                     const double *__restrict__ mu, const double *__restrict__ frames,
                                                     const float *__restrict__ y, double sigma, int have_y,
                                                     const float *__restrict__ feat, long long row0, int cap,
                                                     float *__restrict__ bx, float *__restrict__ by, double *__restrict__ part) {
    __shared__ double lds[GUESS_STATS * 4];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double acc[GUESS_STATS];
#pragma unroll
    for (int k = 0; k < GUESS_STATS; ++k) acc[k] = 0.0;
    if (i < n) {
        double h[GUESS_K][3], m[3];
#pragma unroll
        for (int k = 0; k < GUESS_K; ++k)
#pragma unroll
            for (int c = 0; c < 3; ++c) h[k][c] = ring_at(ring, n, head, k, i, c);
#pragma unroll
        for (int c = 0; c < 3; ++c) m[c] = mu[3 * (size_t)i + c];
        int q = 0;
#pragma unroll
        for (int j = 0; j < GUESS_K; ++j)
#pragma unroll
            for (int k = j; k < GUESS_K; ++k) acc[q++] += h[j][0] * h[k][0] + h[j][1] * h[k][1] + h[j][2] * h[k][2];
#pragma unroll
        for (int j = 0; j < GUESS_K; ++j) acc[36 + j] = h[j][0] * m[0] + h[j][1] * m[1] + h[j][2] * m[2];
        double dl[3], d4[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            double sl = 0.0, s4 = 0.0;
#pragma unroll
            for (int k = 0; k < GUESS_K; ++k) {
                sl += lin.w[k] * h[k][c];
                s4 += mu4.w[k] * h[k][c];
            }
            dl[c] = m[c] - sl;
            d4[c] = m[c] - s4;
        }
        acc[44] = dl[0] * dl[0] + dl[1] * dl[1] + dl[2] * dl[2];
        acc[45] = d4[0] * d4[0] + d4[1] * d4[1] + d4[2] * d4[2];
        const double *R = frames + (size_t)9 * i;
        double t[3];
#pragma unroll
        for (int r = 0; r < 3; ++r) t[r] = R[3 * r] * dl[0] + R[3 * r + 1] * dl[1] + R[3 * r + 2] * dl[2];
        if (have_y) {
            double e = 0.0;
#pragma unroll
            for (int r = 0; r < 3; ++r) {
                const double d = t[r] - sigma * (double)y[3 * (size_t)i + r];
                e += d * d;
            }
            acc[46] = e;
        }
        if (row0 >= 0) {
            const size_t row = (size_t)((row0 + i) % cap);
#pragma unroll
            for (int k = 0; k < GUESS_D; ++k) bx[row * GUESS_D + k] = feat[(size_t)GUESS_D * i + k];
#pragma unroll
            for (int r = 0; r < 3; ++r) by[row * GUESS_OUT + r] = (float)(t[r] / sigma);
        }
    }
    // block reduction of the 47 sums in groups of 4 (LDS of 4 x 47 x 4 waves)
    const int lane = threadIdx.x % WAVE, wv = threadIdx.x / WAVE;
#pragma unroll
    for (int k = 0; k < GUESS_STATS; ++k) acc[k] = wave_sum(acc[k]);
    if (lane == 0)
        for (int k = 0; k < GUESS_STATS; ++k) lds[k * 4 + wv] = acc[k];
    __syncthreads();
    if (threadIdx.x < GUESS_STATS) {
        double s = 0.0;
        for (int w = 0; w < (int)(blockDim.x / WAVE); ++w) s += lds[threadIdx.x * 4 + w];
        part[(size_t)blockIdx.x * GUESS_STATS + threadIdx.x] = s;
    }
}

// Gather a minibatch of training rows: xb[r] = bx[idx[r]], yb[r] = by[idx[r]].
__global__ void k_guess_batch(int rows, const int *__restrict__ idx, const float *__restrict__ bx, const float *__restrict__ by,
                              float *__restrict__ xb, float *__restrict__ yb) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= rows) return;
    const size_t s = (size_t)idx[r];
#pragma unroll
    for (int k = 0; k < GUESS_D; ++k) xb[(size_t)r * GUESS_D + k] = bx[s * GUESS_D + k];
#pragma unroll
    for (int k = 0; k < GUESS_OUT; ++k) yb[(size_t)r * GUESS_OUT + k] = by[s * GUESS_OUT + k];
}

// err partials of the network: sum |y_i - target_i|^2 (targets at rows (row0 + i) % cap).
__global__ __launch_bounds__(256) void k_guess_err(int n, const float *__restrict__ y, const float *__restrict__ by,
                                                   long long row0, int cap, double *__restrict__ part) {
    __shared__ double lds[16];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double e[1] = {0.0};
    if (i < n) {
        const size_t row = (size_t)((row0 + i) % cap);
#pragma unroll
        for (int k = 0; k < GUESS_OUT; ++k) {
            const double d = (double)y[(size_t)GUESS_OUT * i + k] - (double)by[row * GUESS_OUT + k];
            e[0] += d * d;
        }
    }
    block_sum<1>(e, lds);
    if (threadIdx.x == 0) part[blockIdx.x] = e[0];
}

// by rows (row0 + i) % cap, i < n: *= f
__global__ void k_guess_scale_rows(int n, long long row0, int cap, float f, float *__restrict__ by) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const size_t row = (size_t)((row0 + i) % cap);
#pragma unroll
    for (int k = 0; k < GUESS_OUT; ++k) by[row * GUESS_OUT + k] *= f;
}

}  // namespace pgm
