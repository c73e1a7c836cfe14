// dense.h -- the dense induction solve of pgm_jax/solver.py:29-62 (solve_linear_induction):
// (T + diag(1/alpha)) mu = F with T (n, n, 3, 3) the dipole-dipole tensors (zero diagonal blocks),
// by Jacobi-preconditioned conjugate gradients in FP64 (the matrix is symmetric positive definite
// for a physical polarizable system).  One workgroup per atom row reads its 9n-double slab of T
// once per matvec, coalesced; reductions are per-block partials in a fixed order.
#pragma once

#include "common.h"

namespace pgm {

// y_i = sum_j T_ij x_j + x_i / alpha_i for atom i = blockIdx.x; part[i] = x_i . y_i.
__global__ __launch_bounds__(256) void k_dense_matvec(int n, const double *__restrict__ T, const double *__restrict__ alpha,
                                                      const double *__restrict__ x, double *__restrict__ y,
                                                      double *__restrict__ part) {
    __shared__ double lds[3 * 16];
    const int i = blockIdx.x;
    const double *row = T + (size_t)i * n * 9;
    double s[3] = {0.0, 0.0, 0.0};
    for (int j = threadIdx.x; j < n; j += blockDim.x) {
        const double *t = row + (size_t)j * 9;
        const double x0 = x[3 * j], x1 = x[3 * j + 1], x2 = x[3 * j + 2];
        s[0] += t[0] * x0 + t[1] * x1 + t[2] * x2;
        s[1] += t[3] * x0 + t[4] * x1 + t[5] * x2;
        s[2] += t[6] * x0 + t[7] * x1 + t[8] * x2;
    }
    block_sum<3>(s, lds);
    if (threadIdx.x == 0) {
        const double ia = 1.0 / alpha[i];
        double d = 0.0;
        for (int a = 0; a < 3; ++a) {
            const double v = s[a] + x[3 * i + a] * ia;
            y[3 * i + a] = v;
            d += x[3 * i + a] * v;
        }
        part[i] = d;
    }
}

// Vector updates of one CG iteration (3n entries, one thread each) with block partials:
//   mode 0 (start):  r = F - y, z = alpha r, p = z;      partials (r.z, r.r)
//   mode 1 (update): x += al p, r -= al y, z = alpha r;   partials (r.z, r.r)
//   mode 2 (search): p = z + be p
//   mode 3 (guess):  x = alpha F;                        partials (F.F, 0)
__global__ __launch_bounds__(256) void k_dense_update(int n3, int mode, double al, double be, const double *__restrict__ F,
                                                      const double *__restrict__ alpha, const double *__restrict__ y,
                                                      double *__restrict__ x, double *__restrict__ r, double *__restrict__ z,
                                                      double *__restrict__ p, double *__restrict__ part) {
    __shared__ double lds[2 * 16];
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    double s[2] = {0.0, 0.0};
    if (k < n3) {
        const double a = alpha[k / 3];
        if (mode == 0) {
            const double rr = F[k] - y[k];
            r[k] = rr;
            z[k] = a * rr;
            p[k] = a * rr;
            s[0] = a * rr * rr;
            s[1] = rr * rr;
        } else if (mode == 1) {
            x[k] += al * p[k];
            const double rr = r[k] - al * y[k];
            r[k] = rr;
            z[k] = a * rr;
            s[0] = a * rr * rr;
            s[1] = rr * rr;
        } else if (mode == 2) {
            p[k] = z[k] + be * p[k];
        } else {
            x[k] = a * F[k];
            s[0] = F[k] * F[k];
        }
    }
    if (mode != 2) {
        block_sum<2>(s, lds);
        if (threadIdx.x == 0) {
            part[2 * blockIdx.x] = s[0];
            part[2 * blockIdx.x + 1] = s[1];
        }
    }
}

// out[c] = sum of np partials with stride `stride` at offset c (c < stride), one block.
__global__ void k_dense_reduce(const double *__restrict__ part, int np, int stride, double *__restrict__ out) {
    __shared__ double lds[16];
    for (int c = 0; c < stride; ++c) {
        double v[1] = {0.0};
        for (int b = threadIdx.x; b < np; b += blockDim.x) v[0] += part[stride * b + c];
        block_sum<1>(v, lds);
        if (threadIdx.x == 0) out[c] = v[0];
    }
}

}  // namespace pgm
