// ewald.h -- the exact Ewald reciprocal sum of Gaussian charges and dipoles (FP64), the
// alternative to PME used by the reference models PeriodicPGM / PeriodicModel
// (pgm_jax/ewald.py:317-333, periodic.py).
//
//   U_rec = sum_{k in half space} f_k |S_k|^2,  f_k = (4 pi / V) exp(-k^2 / (4 b0^2)) / k^2,
//   S_k   = sum_i (q_i + i k.d_i) exp(i k.r_i) = A_k + i B_k,
//   A_k   = sum_i q_i c_ik - (k.d_i) s_ik,   B_k = sum_i q_i s_ik + (k.d_i) c_ik,
//   dU/dd_i = sum_k 2 f_k (B_k c_ik - A_k s_ik) k,
//   dU/dr_i = sum_k 2 f_k (A_k (-q_i s_ik - (k.d_i) c_ik) + B_k (q_i c_ik - (k.d_i) s_ik)) k.
//
// Two O(N K) passes, both FP64 with sincos per (atom, k): the structure factors (one thread per
// k-vector, atoms staged through LDS) and the gathers (one thread per atom, k-vectors and S_k
// staged through LDS).  Deterministic: fixed summation orders, per-block partials.
#pragma once

#include "common.h"

namespace pgm {

constexpr int EW_TILE = 256;

// S_k for every k (kvec[k] = (kx, ky, kz, f_k)); per-block partial sums of f_k |S_k|^2.
// WITHQ: the charges dq.w enter S; otherwise dipoles only.
template <typename real, bool WITHQ>
__global__ __launch_bounds__(EW_TILE) void k_ewald_sk(int K, const double4 *__restrict__ kvec, int n,
                                                      const double *__restrict__ pos, const real4<real> *__restrict__ dq,
                                                      double2 *__restrict__ S, double *__restrict__ epart,
                                                      const int *__restrict__ conv_it, int iter) {
    if (conv_it && iter >= *conv_it) return;
    __shared__ double sp[EW_TILE][3];
    __shared__ double sd[EW_TILE][4];
    __shared__ double lds[16];
    const int k = blockIdx.x * EW_TILE + threadIdx.x;
    double4 kv = make_double4(0, 0, 0, 0);
    if (k < K) kv = kvec[k];
    double A = 0.0, B = 0.0;
    for (int t0 = 0; t0 < n; t0 += EW_TILE) {
        __syncthreads();
        const int a = t0 + threadIdx.x;
        if (a < n) {
            sp[threadIdx.x][0] = pos[3 * a];
            sp[threadIdx.x][1] = pos[3 * a + 1];
            sp[threadIdx.x][2] = pos[3 * a + 2];
            const real4<real> d = dq[a];
            sd[threadIdx.x][0] = (double)d.x;
            sd[threadIdx.x][1] = (double)d.y;
            sd[threadIdx.x][2] = (double)d.z;
            sd[threadIdx.x][3] = WITHQ ? (double)d.w : 0.0;
        }
        __syncthreads();
        const int m = min(EW_TILE, n - t0);
        if (k < K)
            for (int j = 0; j < m; ++j) {
                const double ph = kv.x * sp[j][0] + kv.y * sp[j][1] + kv.z * sp[j][2];
                double s, c;
                sincos(ph, &s, &c);
                const double kd = kv.x * sd[j][0] + kv.y * sd[j][1] + kv.z * sd[j][2];
                const double q = sd[j][3];
                A += q * c - kd * s;
                B += q * s + kd * c;
            }
    }
    double e[1] = {0.0};
    if (k < K) {
        S[k] = make_double2(A, B);
        e[0] = kv.w * (A * A + B * B);
    }
    block_sum<1>(e, lds);
    if (threadIdx.x == 0 && epart) epart[blockIdx.x] = e[0];
}

// dU_rec/dd_i (and with DR, dU_rec/dr_i) for atoms i, FP64.
template <typename real, bool DR>
PGM_DEV void ewald_atom(int i, int K, const double4 *__restrict__ kvec, const double2 *__restrict__ S,
                        const double *__restrict__ pos, const real4<real> *__restrict__ dq, double4 (*sk)[2],
                        double (&gd)[3], double (&gr)[3]) {
    double r[3] = {0, 0, 0}, d[3] = {0, 0, 0}, q = 0.0;
    if (i >= 0) {
        r[0] = pos[3 * i];
        r[1] = pos[3 * i + 1];
        r[2] = pos[3 * i + 2];
        const real4<real> v = dq[i];
        d[0] = (double)v.x;
        d[1] = (double)v.y;
        d[2] = (double)v.z;
        q = (double)v.w;
    }
    for (int a = 0; a < 3; ++a) gd[a] = gr[a] = 0.0;
    for (int t0 = 0; t0 < K; t0 += EW_TILE) {
        __syncthreads();
        const int kk = t0 + threadIdx.x;
        if (kk < K) {
            const double4 kv = kvec[kk];
            const double2 s = S[kk];
            sk[threadIdx.x][0] = kv;
            sk[threadIdx.x][1] = make_double4(2.0 * kv.w * s.x, 2.0 * kv.w * s.y, 0.0, 0.0);
        }
        __syncthreads();
        const int m = min(EW_TILE, K - t0);
        if (i >= 0)
            for (int j = 0; j < m; ++j) {
                const double4 kv = sk[j][0];
                const double wA = sk[j][1].x, wB = sk[j][1].y;
                const double ph = kv.x * r[0] + kv.y * r[1] + kv.z * r[2];
                double s, c;
                sincos(ph, &s, &c);
                const double cd = wB * c - wA * s;
                gd[0] += cd * kv.x;
                gd[1] += cd * kv.y;
                gd[2] += cd * kv.z;
                if (DR) {
                    const double kd = kv.x * d[0] + kv.y * d[1] + kv.z * d[2];
                    const double cr = wA * (-q * s - kd * c) + wB * (q * c - kd * s);
                    gr[0] += cr * kv.x;
                    gr[1] += cr * kv.y;
                    gr[2] += cr * kv.z;
                }
            }
    }
}

// The CG's field evaluation with the Ewald reciprocal part: rec = dU_rec/dd, then the combination
// of `mode` (combine_out, shared with the PME gather).
template <typename real>
__global__ __launch_bounds__(EW_TILE) void k_ewald_gather_combine(int n, int K, const double4 *__restrict__ kvec,
                                                                  const double2 *__restrict__ S, const double *__restrict__ pos,
                                                                  const real4<real> *__restrict__ dq,
                                                                  const real4<real> *__restrict__ row, const real *__restrict__ alpha,
                                                                  const real4<real> *__restrict__ extra, real cself, int mode,
                                                                  real4<real> *__restrict__ out, double *__restrict__ part,
                                                                  const int *__restrict__ conv_it, int iter) {
    if (conv_it && iter >= *conv_it) return;
    __shared__ double4 sk[EW_TILE][2];
    __shared__ double lds[16];
    const int i0 = blockIdx.x * EW_TILE + threadIdx.x;
    const int i = i0 < n ? i0 : -1;
    double gd[3], gr[3];
    ewald_atom<real, false>(i, K, kvec, S, pos, dq, sk, gd, gr);
    double acc[1] = {0.0};
    if (i >= 0) combine_out<real>(mode, i, (real)gd[0], (real)gd[1], (real)gd[2], dq, row, alpha, extra, cself, out, acc[0]);
    if (part) {
        block_sum<1>(acc, lds);
        if (threadIdx.x == 0) part[blockIdx.x] = acc[0];
    }
}

// Forces pass: rec_dr = dU_rec/dr, gp = KE (rowf + dU_rec/dd - c_self d), rec_dd (optional).
template <typename real>
__global__ __launch_bounds__(EW_TILE) void k_ewald_gather_forces(int n, int K, const double4 *__restrict__ kvec,
                                                                 const double2 *__restrict__ S, const double *__restrict__ pos,
                                                                 const real4<real> *__restrict__ dq, const double *__restrict__ dd,
                                                                 const double *__restrict__ rowf, double cself,
                                                                 double *__restrict__ rec_dr, double *__restrict__ gp,
                                                                 double *__restrict__ rec_dd) {
    __shared__ double4 sk[EW_TILE][2];
    const int i0 = blockIdx.x * EW_TILE + threadIdx.x;
    const int i = i0 < n ? i0 : -1;
    double gd[3], gr[3];
    ewald_atom<real, true>(i, K, kvec, S, pos, dq, sk, gd, gr);
    if (i < 0) return;
#pragma unroll
    for (int b = 0; b < 3; ++b) {
        rec_dr[3 * i + b] = gr[b];
        gp[3 * i + b] = KE * (rowf[3 * i + b] + gd[b] - cself * dd[3 * i + b]);
        if (rec_dd) rec_dd[3 * i + b] = gd[b];
    }
}

// Strain derivative of the k-space factors: per block the six sums of
// f_k |S_k|^2 2 (1/(4 b0^2) + 1/k^2) k_a k_b (xx yy zz xy xz yz) and E_rec = sum f_k |S_k|^2.
__global__ __launch_bounds__(256) void k_ewald_virial(int K, const double4 *__restrict__ kvec, const double2 *__restrict__ S,
                                                      double inv4b2, double *__restrict__ vpart) {
    __shared__ double lds[7 * 16];
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    double acc[7] = {0, 0, 0, 0, 0, 0, 0};
    if (k < K) {
        const double4 kv = kvec[k];
        const double2 s = S[k];
        const double e = kv.w * (s.x * s.x + s.y * s.y);
        const double k2 = kv.x * kv.x + kv.y * kv.y + kv.z * kv.z;
        const double c = 2.0 * e * (inv4b2 + 1.0 / k2);
        acc[0] = c * kv.x * kv.x;
        acc[1] = c * kv.y * kv.y;
        acc[2] = c * kv.z * kv.z;
        acc[3] = c * kv.x * kv.y;
        acc[4] = c * kv.x * kv.z;
        acc[5] = c * kv.y * kv.z;
        acc[6] = e;
    }
    block_sum<7>(acc, lds);
    if (threadIdx.x == 0)
        for (int k2i = 0; k2i < 7; ++k2i) vpart[7 * blockIdx.x + k2i] = acc[k2i];
}

}  // namespace pgm
