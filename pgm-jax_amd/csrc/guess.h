// guess.h -- the learned CG initial guess for rigid 3-site molecules (water): per-atom features in
// a local frame, the guess assembled from the network output, and the harvest of training rows.
//
// Each atom gets an orthonormal frame from its molecule at the current step (rows e1, e2, e3):
//   O:  e1 = unit((h1 - o) + (h2 - o)), e2 = unit(h1 - h2 made orthogonal to e1), e3 = e1 x e2;
//   Hk: e1 = unit(hk - o), e2 = unit((h_other - o) made orthogonal to e1), e3 = e1 x e2.
// Features (GUESS_D = 16 per atom): the last four converged dipoles rotated into the frame and
// divided by the dipole scale (12), 1 for oxygen / 0 for hydrogen, zeros.  The network predicts
// the correction to the fourth-order extrapolation in the frame, in units of `sigma`:
//   x0 = mu4 + sigma R^T y,   mu4 = 4 mu_{t-1} - 6 mu_{t-2} + 4 mu_{t-3} - mu_{t-4}.
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

// Features (n x GUESS_D, FP32) and frames (n x 9, FP64) at positions pos from the history hist
// (4, n, 3) newest first; inv_scale = 1 / dipole scale.  One thread per atom; n = 3 nmol.
__global__ void k_guess_features(int n, const double *__restrict__ pos, const double *__restrict__ hist, double inv_scale,
                                 float *__restrict__ feat, double *__restrict__ frames) {
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
        const double *mu = hist + ((size_t)h * n + i) * 3;
#pragma unroll
        for (int r = 0; r < 3; ++r) f[3 * h + r] = (float)(inv_scale * (R[3 * r] * mu[0] + R[3 * r + 1] * mu[1] + R[3 * r + 2] * mu[2]));
    }
    f[12] = a == 0 ? 1.0f : 0.0f;
    f[13] = f[14] = f[15] = 0.0f;
}

// x0 = mu4 + sigma R^T y (y: n x GUESS_OUT network outputs), written as real4 (d, 0).
template <typename real>
__global__ void k_guess_apply(int n, const double *__restrict__ hist, const double *__restrict__ frames,
                              const float *__restrict__ y, double sigma, real4<real> *__restrict__ x0) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double *R = frames + (size_t)9 * i;
    double v[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const double p = 4.0 * hist[3 * (size_t)i + k] - 6.0 * hist[3 * ((size_t)n + i) + k] +
                         4.0 * hist[3 * (2 * (size_t)n + i) + k] - hist[3 * (3 * (size_t)n + i) + k];
        v[k] = p + sigma * (R[k] * y[3 * (size_t)i] + R[3 + k] * y[3 * (size_t)i + 1] + R[6 + k] * y[3 * (size_t)i + 2]);
    }
    real4<real> o;
    o.x = (real)v[0];
    o.y = (real)v[1];
    o.z = (real)v[2];
    o.w = 0;
    x0[i] = o;
}

// Training rows: features of step t (from the history before the solve) and the target
// (R mu_t - R mu4) / sigma, appended to the ring buffer at rows (row0 + i) % cap.  Per-block
// partial sums of |target|^2 (the guess error of mu4 in sigma units) for the judge.
__global__ __launch_bounds__(256) void k_guess_harvest(int n, const float *__restrict__ feat, const double *__restrict__ frames,
                                                       const double *__restrict__ hist_before, const double *__restrict__ mu,
                                                       double inv_sigma, long long row0, int cap, float *__restrict__ bx,
                                                       float *__restrict__ by, double *__restrict__ part) {
    __shared__ double lds[16];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double e[1] = {0.0};
    if (i < n) {
        const size_t row = (size_t)((row0 + i) % cap);
#pragma unroll
        for (int k = 0; k < GUESS_D; ++k) bx[row * GUESS_D + k] = feat[(size_t)GUESS_D * i + k];
        const double *R = frames + (size_t)9 * i;
        double d[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const double p = 4.0 * hist_before[3 * (size_t)i + k] - 6.0 * hist_before[3 * ((size_t)n + i) + k] +
                             4.0 * hist_before[3 * (2 * (size_t)n + i) + k] - hist_before[3 * (3 * (size_t)n + i) + k];
            d[k] = mu[3 * (size_t)i + k] - p;
        }
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            const double t = inv_sigma * (R[3 * r] * d[0] + R[3 * r + 1] * d[1] + R[3 * r + 2] * d[2]);
            by[row * GUESS_OUT + r] = (float)t;
            e[0] += t * t;
        }
    }
    block_sum<1>(e, lds);
    if (threadIdx.x == 0) part[blockIdx.x] = e[0];
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
