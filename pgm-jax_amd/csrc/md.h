// md.h -- the FP64 MD step around the force field: velocity Verlet with SHAKE/RATTLE for rigid
// 3-site molecules (water), kinetic energy, and the displacement test of the neighbour rows.
//
// Restates the constrained step of pgm_jax/md/flexible.py:776-799 with the rigid-water cluster of
// pgm_jax/md/constraints.py:254-299 (three distance constraints O-H1, O-H2, H1-H2 per molecule,
// Newton iterations on the multipliers with closed-form 3x3 solves, FP64 throughout).  One thread
// per molecule; atoms of a molecule are consecutive.
#pragma once

#include "common.h"

namespace pgm {

// 3x3 solve by Cramer's rule (A row-major), returns false when singular.
PGM_DEV bool solve3(const double *A, const double *b, double *x) {
    const double det = A[0] * (A[4] * A[8] - A[5] * A[7]) - A[1] * (A[3] * A[8] - A[5] * A[6]) + A[2] * (A[3] * A[7] - A[4] * A[6]);
    if (fabs(det) < 1e-300) return false;
    const double id = 1.0 / det;
    x[0] = id * (b[0] * (A[4] * A[8] - A[5] * A[7]) - A[1] * (b[1] * A[8] - A[5] * b[2]) + A[2] * (b[1] * A[7] - A[4] * b[2]));
    x[1] = id * (A[0] * (b[1] * A[8] - A[5] * b[2]) - b[0] * (A[3] * A[8] - A[5] * A[6]) + A[2] * (A[3] * b[2] - b[1] * A[6]));
    x[2] = id * (A[0] * (A[4] * b[2] - b[1] * A[7]) - A[1] * (A[3] * b[2] - b[1] * A[6]) + b[0] * (A[3] * A[7] - A[4] * A[6]));
    return true;
}

// constraint c joins atoms (CA[c], CB[c]) of the molecule
__constant__ const int CA[3] = {0, 0, 1};
__constant__ const int CB[3] = {1, 2, 2};

// Half kick, drift and SHAKE (positions), velocities corrected by the constraint displacement / dt.
// d2[3 m + c] = squared constraint lengths.
__global__ void k_md_first(int nmol, double *__restrict__ pos, double *__restrict__ vel, const double *__restrict__ f,
                           const double *__restrict__ minv, const double *__restrict__ d2, double dt, int *__restrict__ fail) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= nmol) return;
    const int a0 = 3 * m;
    // molecule-local coordinates (origin at atom 0's old position): the constraint residuals then
    // resolve to ~1e-16 of the bond length whatever the absolute coordinates
    double org[3], r0[3][3], r1[3][3], v[3][3], im[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) org[k] = pos[3 * a0 + k];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        im[a] = minv[a0 + a];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            r0[a][k] = pos[3 * (a0 + a) + k] - org[k];
            v[a][k] = vel[3 * (a0 + a) + k] + 0.5 * dt * f[3 * (a0 + a) + k] * im[a];
            r1[a][k] = r0[a][k] + dt * v[a][k];
        }
    }
    double s[3][3];  // old constraint vectors r_A - r_B
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int k = 0; k < 3; ++k) s[c][k] = r0[CA[c]][k] - r0[CB[c]][k];
    // displacement of atom a: sum_c lam_c G(a, c) s_c / m_a, G = +1 for CA, -1 for CB
    double lam[3] = {0, 0, 0};
    bool ok = false;
    for (int it = 0; it < 50; ++it) {
        double x[3][3], sig[3], J[9];
        double rr[3][3];
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                double dsum = 0.0;
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const double g = (CA[c] == a) ? 1.0 : ((CB[c] == a) ? -1.0 : 0.0);
                    dsum += lam[c] * g * s[c][k];
                }
                rr[a][k] = r1[a][k] + dsum * im[a];
            }
        double worst = 0.0;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
#pragma unroll
            for (int k = 0; k < 3; ++k) x[c][k] = rr[CA[c]][k] - rr[CB[c]][k];
            sig[c] = x[c][0] * x[c][0] + x[c][1] * x[c][1] + x[c][2] * x[c][2] - d2[3 * m + c];
            worst = fmax(worst, fabs(sig[c]) / d2[3 * m + c]);
        }
        if (worst < 1e-14) {
            ok = true;
            break;
        }
        // J[c][c'] = 2 x_c . (G(A_c, c')/m_A - G(B_c, c')/m_B) s_c'
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int cp = 0; cp < 3; ++cp) {
                const double gA = (CA[cp] == CA[c]) ? 1.0 : ((CB[cp] == CA[c]) ? -1.0 : 0.0);
                const double gB = (CA[cp] == CB[c]) ? 1.0 : ((CB[cp] == CB[c]) ? -1.0 : 0.0);
                const double w = gA * im[CA[c]] - gB * im[CB[c]];
                J[3 * c + cp] = 2.0 * w * (x[c][0] * s[cp][0] + x[c][1] * s[cp][1] + x[c][2] * s[cp][2]);
            }
        double rhs[3] = {-sig[0], -sig[1], -sig[2]}, dl[3];
        if (!solve3(J, rhs, dl)) break;
#pragma unroll
        for (int c = 0; c < 3; ++c) lam[c] += dl[c];
    }
    if (!ok) atomicAdd(fail, 1);
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            double dsum = 0.0;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const double g = (CA[c] == a) ? 1.0 : ((CB[c] == a) ? -1.0 : 0.0);
                dsum += lam[c] * g * s[c][k];
            }
            const double dr = dsum * im[a];
            pos[3 * (a0 + a) + k] = org[k] + (r1[a][k] + dr);
            vel[3 * (a0 + a) + k] = v[a][k] + dr / dt;
        }
}

// Half kick and RATTLE (velocities tangent to the constraints); partial sums of m v^2 per block.
__global__ __launch_bounds__(256) void k_md_second(int nmol, const double *__restrict__ pos, double *__restrict__ vel,
                                                   const double *__restrict__ f, const double *__restrict__ minv,
                                                   const double *__restrict__ mass, double dt, double *__restrict__ part) {
    __shared__ double lds[32];
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    double ke[1] = {0.0};
    if (m < nmol) {
        const int a0 = 3 * m;
        double v[3][3], s[3][3], im[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            im[a] = minv[a0 + a];
#pragma unroll
            for (int k = 0; k < 3; ++k) v[a][k] = vel[3 * (a0 + a) + k] + 0.5 * dt * f[3 * (a0 + a) + k] * im[a];
        }
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int k = 0; k < 3; ++k) s[c][k] = pos[3 * (a0 + CA[c]) + k] - pos[3 * (a0 + CB[c]) + k];
        // K mu = -(v_A - v_B) . s_c,   K[c][c'] = s_c . s_c' (G(A_c, c')/m_A - G(B_c, c')/m_B)
        double K[9], rhs[3], mu[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            rhs[c] = -((v[CA[c]][0] - v[CB[c]][0]) * s[c][0] + (v[CA[c]][1] - v[CB[c]][1]) * s[c][1] +
                       (v[CA[c]][2] - v[CB[c]][2]) * s[c][2]);
#pragma unroll
            for (int cp = 0; cp < 3; ++cp) {
                const double gA = (CA[cp] == CA[c]) ? 1.0 : ((CB[cp] == CA[c]) ? -1.0 : 0.0);
                const double gB = (CA[cp] == CB[c]) ? 1.0 : ((CB[cp] == CB[c]) ? -1.0 : 0.0);
                const double w = gA * im[CA[c]] - gB * im[CB[c]];
                K[3 * c + cp] = w * (s[c][0] * s[cp][0] + s[c][1] * s[cp][1] + s[c][2] * s[cp][2]);
            }
        }
        if (!solve3(K, rhs, mu)) mu[0] = mu[1] = mu[2] = 0.0;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const double mass_a = mass[a0 + a];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                double dsum = 0.0;
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const double g = (CA[c] == a) ? 1.0 : ((CB[c] == a) ? -1.0 : 0.0);
                    dsum += mu[c] * g * s[c][k];
                }
                const double vv = v[a][k] + dsum * im[a];
                vel[3 * (a0 + a) + k] = vv;
                ke[0] += mass_a * vv * vv;
            }
        }
    }
    block_sum<1>(ke, lds);
    if (threadIdx.x == 0) part[blockIdx.x] = ke[0];
}

// Largest displacement since the rows were built (block partials of the maximum).
__global__ __launch_bounds__(256) void k_max_disp(int n, const double *__restrict__ pos, const double *__restrict__ ref,
                                                  double *__restrict__ part) {
    __shared__ double lds[32];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double d = 0.0;
    if (i < n) {
        const double dx = pos[3 * i] - ref[3 * i], dy = pos[3 * i + 1] - ref[3 * i + 1], dz = pos[3 * i + 2] - ref[3 * i + 2];
        d = sqrt(dx * dx + dy * dy + dz * dz);
    }
    d = block_max(d, lds);
    if (threadIdx.x == 0) part[blockIdx.x] = d;
}

__global__ void k_scale3(int n3, double *__restrict__ v, double f) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n3) v[i] *= f;
}

// One block: out = reduction of np partials (sum or max).
__global__ void k_reduce_partials(const double *__restrict__ part, int np, int is_max, double *__restrict__ out) {
    __shared__ double lds[32];
    double v = 0.0;
    for (int b = threadIdx.x; b < np; b += blockDim.x) v = is_max ? fmax(v, part[b]) : v + part[b];
    if (is_max) {
        v = block_max(v, lds);
    } else {
        double a[1] = {v};
        block_sum<1>(a, lds);
        v = a[0];
    }
    if (threadIdx.x == 0) out[0] = v;
}

}  // namespace pgm
