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

PGM_DEV double gsign(int a, int c) { return (CA[c] == a) ? 1.0 : ((CB[c] == a) ? -1.0 : 0.0); }

// SHAKE: given old positions r0 and unconstrained new positions r1 (molecule-local coordinates),
// find multipliers lam with |r1 + disp|^2 constraints met to 1e-14 relative, disp_a =
// sum_c lam_c G(a, c) s_c / m_a (s_c the old constraint vectors).  Returns false without
// convergence in 50 Newton iterations.
PGM_DEV bool shake3(const double (&r0)[3][3], const double (&r1)[3][3], const double (&im)[3], const double *d2m,
                    double (&disp)[3][3]) {
    double s[3][3];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int k = 0; k < 3; ++k) s[c][k] = r0[CA[c]][k] - r0[CB[c]][k];
    double lam[3] = {0, 0, 0};
    bool ok = false;
    for (int it = 0; it < 50; ++it) {
        double x[3][3], sig[3], J[9], rr[3][3];
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                double dsum = 0.0;
#pragma unroll
                for (int c = 0; c < 3; ++c) dsum += lam[c] * gsign(a, c) * s[c][k];
                rr[a][k] = r1[a][k] + dsum * im[a];
            }
        double worst = 0.0;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
#pragma unroll
            for (int k = 0; k < 3; ++k) x[c][k] = rr[CA[c]][k] - rr[CB[c]][k];
            sig[c] = x[c][0] * x[c][0] + x[c][1] * x[c][1] + x[c][2] * x[c][2] - d2m[c];
            worst = fmax(worst, fabs(sig[c]) / d2m[c]);
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
                const double w = gsign(CA[c], cp) * im[CA[c]] - gsign(CB[c], cp) * im[CB[c]];
                J[3 * c + cp] = 2.0 * w * (x[c][0] * s[cp][0] + x[c][1] * s[cp][1] + x[c][2] * s[cp][2]);
            }
        double rhs[3] = {-sig[0], -sig[1], -sig[2]}, dl[3];
        if (!solve3(J, rhs, dl)) break;
#pragma unroll
        for (int c = 0; c < 3; ++c) lam[c] += dl[c];
    }
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            double dsum = 0.0;
#pragma unroll
            for (int c = 0; c < 3; ++c) dsum += lam[c] * gsign(a, c) * s[c][k];
            disp[a][k] = dsum * im[a];
        }
    return ok;
}

// RATTLE: project velocities v tangent to the constraints of positions r (s_c = r_A - r_B):
// K mu = -(v_A - v_B) . s_c,  K[c][c'] = s_c . s_c' (G(A_c, c')/m_A - G(B_c, c')/m_B).
PGM_DEV void rattle3(const double (&r)[3][3], double (&v)[3][3], const double (&im)[3]) {
    double s[3][3];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int k = 0; k < 3; ++k) s[c][k] = r[CA[c]][k] - r[CB[c]][k];
    double K[9], rhs[3], mu[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        rhs[c] = -((v[CA[c]][0] - v[CB[c]][0]) * s[c][0] + (v[CA[c]][1] - v[CB[c]][1]) * s[c][1] +
                   (v[CA[c]][2] - v[CB[c]][2]) * s[c][2]);
#pragma unroll
        for (int cp = 0; cp < 3; ++cp) {
            const double w = gsign(CA[c], cp) * im[CA[c]] - gsign(CB[c], cp) * im[CB[c]];
            K[3 * c + cp] = w * (s[c][0] * s[cp][0] + s[c][1] * s[cp][1] + s[c][2] * s[cp][2]);
        }
    }
    if (!solve3(K, rhs, mu)) mu[0] = mu[1] = mu[2] = 0.0;
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            double dsum = 0.0;
#pragma unroll
            for (int c = 0; c < 3; ++c) dsum += mu[c] * gsign(a, c) * s[c][k];
            v[a][k] += dsum * im[a];
        }
}

// ---------------------------------------------------------------------------- Philox4x32-10
// Counter-based generator of Salmon, Moraes, Dror and Shaw ("Parallel random numbers: as easy as
// 1, 2, 3", SC 2011): ten rounds of two 32x32->64 multiplications, key bumped by the Weyl
// constants.  Normals by Box-Muller from pairs of uniforms in (0, 1).
struct U4 {
    uint32_t x, y, z, w;
};
PGM_DEV U4 philox4x32_10(U4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c.x;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c.z;
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        c = U4{hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0};
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return c;
}
PGM_DEV double u01(uint32_t x) { return ((double)x + 0.5) * 2.3283064365386963e-10; }  // (x + 1/2) / 2^32
// Four standard normals from counter (a, b, c, 0) and the 64-bit seed.
PGM_DEV void normals4(uint32_t a, uint32_t b, uint32_t c, uint64_t seed, double (&g)[4]) {
    const U4 u = philox4x32_10(U4{a, b, c, 0u}, (uint32_t)seed, (uint32_t)(seed >> 32));
    const double r0 = sqrt(-2.0 * log(u01(u.x))), t0 = 6.283185307179586 * u01(u.y);
    const double r1 = sqrt(-2.0 * log(u01(u.z))), t1 = 6.283185307179586 * u01(u.w);
    g[0] = r0 * cos(t0);
    g[1] = r0 * sin(t0);
    g[2] = r1 * cos(t1);
    g[3] = r1 * sin(t1);
}

// Minibatch rows for training: idx[r] = Philox(r, counter, 0x5eed) mod filled.
__global__ void k_sample_rows(int rows, long long filled, uint64_t seed, uint32_t counter, int *__restrict__ idx) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= rows) return;
    const U4 u = philox4x32_10(U4{(uint32_t)r, counter, 0x5eedu, 0u}, (uint32_t)seed, (uint32_t)(seed >> 32));
    const uint64_t w = ((uint64_t)u.x << 32) | u.y;
    idx[r] = (int)(w % (uint64_t)filled);
}

// Velocity Verlet, first half (NVE): half kick, drift by dt, SHAKE; velocities corrected by the
// constraint displacement / dt (the closing kick's RATTLE projects them).  d2[3 m + c] = squared
// constraint lengths.
__global__ void k_md_first(int nmol, double *__restrict__ pos, double *__restrict__ vel, const double *__restrict__ f,
                           const double *__restrict__ minv, const double *__restrict__ d2, double dt, int *__restrict__ fail) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= nmol) return;
    const int a0 = 3 * m;
    // molecule-local coordinates (origin at atom 0's old position): the constraint residuals then
    // resolve to ~1e-16 of the bond length whatever the absolute coordinates
    double org[3], r0[3][3], r1[3][3], v[3][3], im[3], disp[3][3];
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
    if (!shake3(r0, r1, im, d2 + 3 * m, disp)) atomicAdd(fail, 1);
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            pos[3 * (a0 + a) + k] = org[k] + (r1[a][k] + disp[a][k]);
            vel[3 * (a0 + a) + k] = v[a][k] + disp[a][k] / dt;
        }
}

// g-BAOAB first part (pgm_jax/md/flexible.py:776-799 with the O step of
// pgm_jax/md/thermostats.py:195-242): half kick; drift dt/2 + SHAKE, v = (q1 - q)/(dt/2), RATTLE;
// O: v <- c v + sqrt(kT (1 - c^2) / m) xi, RATTLE; drift dt/2 + SHAKE, v = (q2 - q1)/(dt/2).
// xi from Philox with counter (molecule, step, k) and the seed.  heat partials: the O step's
// change of sum m v^2 / 2 per block.
__global__ __launch_bounds__(256) void k_md_first_langevin(int nmol, double *__restrict__ pos, double *__restrict__ vel,
                                                           const double *__restrict__ f, const double *__restrict__ minv,
                                                           const double *__restrict__ d2, double dt, double c, double kT,
                                                           uint64_t seed, uint32_t step, int *__restrict__ fail,
                                                           double *__restrict__ heat_part) {
    __shared__ double lds[32];
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    double dq[1] = {0.0};
    if (m < nmol) {
        const int a0 = 3 * m;
        const double h = 0.5 * dt;
        double org[3], r0[3][3], r1[3][3], v[3][3], im[3], disp[3][3];
#pragma unroll
        for (int k = 0; k < 3; ++k) org[k] = pos[3 * a0 + k];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            im[a] = minv[a0 + a];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                r0[a][k] = pos[3 * (a0 + a) + k] - org[k];
                v[a][k] = vel[3 * (a0 + a) + k] + 0.5 * dt * f[3 * (a0 + a) + k] * im[a];
                r1[a][k] = r0[a][k] + h * v[a][k];
            }
        }
        bool ok = shake3(r0, r1, im, d2 + 3 * m, disp);
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                r1[a][k] += disp[a][k];
                v[a][k] += disp[a][k] / h;
            }
        rattle3(r1, v, im);
        double g[12];
        {
            double t[4];
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                normals4((uint32_t)m, step, (uint32_t)q, seed, t);
#pragma unroll
                for (int k = 0; k < 4; ++k) g[4 * q + k] = t[k];
            }
        }
        double e0 = 0.0, e1 = 0.0;
        const double sq = sqrt(kT * (1.0 - c * c));
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const double ma = 1.0 / im[a], sm = sq * sqrt(im[a]);
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                e0 += ma * v[a][k] * v[a][k];
                v[a][k] = c * v[a][k] + sm * g[3 * a + k];
            }
        }
        rattle3(r1, v, im);
        double r2[3][3];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const double ma = 1.0 / im[a];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                e1 += ma * v[a][k] * v[a][k];
                r2[a][k] = r1[a][k] + h * v[a][k];
            }
        }
        ok = shake3(r1, r2, im, d2 + 3 * m, disp) && ok;
        if (!ok) atomicAdd(fail, 1);
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                pos[3 * (a0 + a) + k] = org[k] + (r2[a][k] + disp[a][k]);
                vel[3 * (a0 + a) + k] = v[a][k] + disp[a][k] / h;
            }
        dq[0] = 0.5 * (e1 - e0);
    }
    block_sum<1>(dq, lds);
    if (threadIdx.x == 0) heat_part[blockIdx.x] = dq[0];
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
        double v[3][3], r[3][3], im[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            im[a] = minv[a0 + a];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                v[a][k] = vel[3 * (a0 + a) + k] + 0.5 * dt * f[3 * (a0 + a) + k] * im[a];
                r[a][k] = pos[3 * (a0 + a) + k];
            }
        }
        rattle3(r, v, im);
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const double mass_a = mass[a0 + a];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                vel[3 * (a0 + a) + k] = v[a][k];
                ke[0] += mass_a * v[a][k] * v[a][k];
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

// One block: acc[0] += sum of np partials.
__global__ void k_accumulate_partials(const double *__restrict__ part, int np, double *__restrict__ acc) {
    __shared__ double lds[32];
    double v[1] = {0.0};
    for (int b = threadIdx.x; b < np; b += blockDim.x) v[0] += part[b];
    block_sum<1>(v, lds);
    if (threadIdx.x == 0) acc[0] += v[0];
}

// One block: out = reduction of np partials (sum or max).
// host_out (pinned, coherent host memory; may be null): the result also written there.
__global__ void k_reduce_partials(const double *__restrict__ part, int np, int is_max, double *__restrict__ out,
                                  double *host_out = nullptr) {
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
    if (threadIdx.x == 0) {
        out[0] = v;
        if (host_out) *host_out = v;
    }
}

}  // namespace pgm
