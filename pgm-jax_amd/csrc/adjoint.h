// adjoint.h -- the backward pass of the induced-dipole solve (the custom VJP of
// pgm_jax/md/forcefield.py:1853-1915).
//
// From A mu = b(theta), a cotangent mu_bar of the dipoles gives theta_bar = lam . d(b - A mu)/dtheta
// at fixed mu, with A lam = mu_bar (A symmetric: one more CG with the same operator, x0 = 0, no
// peek).  With E(mu) = U(q, p + mu) + sum |mu|^2 / (2 alpha) and b - A mu = -dE/dmu:
//   q_bar_i     = -phi_i(lam)                         the potential of the dipoles lam at atom i
//   alpha_bar_i = mu_i . lam_i / alpha_i^2
//   p_bar_i     = field_i(0, lam)                     (then projected on the covalent frames)
//   R_bar_i     = 2 R_i sum_k a_ik^3 d/da [e_ik along d -> d + eps lam]
// a_ik = (2 (R_i^2 + R_k^2))^(-1/2).  Only the Gaussian part of the real-space kernel depends on a:
// dG_n/da = (2/sqrt(pi)) (2 a^2)^n exp(-a^2 r^2) (G_0 = erf(a r)/r, G_n = -(1/r) dG_(n-1)/dr).
// Everything here is FP64 whatever the CG's type; the kernels run once per backward pass.
#pragma once

#include "cg.h"
#include "common.h"
#include "pme.h"

namespace pgm {

// CG normaliser of the adjoint: mean |alpha mu_bar| + 1e-300 over the 3N components.  One block.
__global__ void k_adjoint_norm(int n, const double *__restrict__ mu_bar, const double *__restrict__ alpha, CGState *st) {
    __shared__ double lds[16];
    double s[1] = {0.0};
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const double a = alpha[i];
        s[0] += fabs(a * mu_bar[3 * i]) + fabs(a * mu_bar[3 * i + 1]) + fabs(a * mu_bar[3 * i + 2]);
    }
    block_sum<1>(s, lds);
    if (threadIdx.x == 0) st->norm = s[0] / (3.0 * n) + 1e-300;
}

// Real-space part over the rows (both parts of every compacted row: near from the front, far from
// the back), RL lanes per atom, i's own row only (each pair is in both rows, so no scatter):
//   phi_i = sum_k G1 lam_k . x,                      x = r_i - r_k
//   rbar_i = 2 R_i sum_k a^3 [(q_i lam_k.x - q_k lam_i.x) H1 - H2 (lam_i.x d_k.x + d_i.x lam_k.x)
//                             + H1 (lam_i . d_k + d_i . lam_k)],   H_n = dG_n/da
// with d = p + mu; electrostatic terms stop at the elec cutoff (r^2 < rce2).
constexpr int ADJ_RL = 8;
__global__ __launch_bounds__(256) void k_adjoint_rows(int n, const int *__restrict__ rows, const int *__restrict__ count,
                                                      const int *__restrict__ count_near, int cap,
                                                      const double4 *__restrict__ xr, Box box, double beta, double rce2,
                                                      const double *__restrict__ q, const double *__restrict__ lam,
                                                      const double *__restrict__ p, const double *__restrict__ mu,
                                                      double *__restrict__ phi, double *__restrict__ rbar) {
    const int lane = threadIdx.x % ADJ_RL;
    const int i = blockIdx.x * (blockDim.x / ADJ_RL) + threadIdx.x / ADJ_RL;
    if (i >= n) return;  // whole groups of ADJ_RL lanes leave together (blockDim a multiple of ADJ_RL)
    const double4 a = xr[i];
    const double qi = q[i];
    const double li[3] = {lam[3 * i], lam[3 * i + 1], lam[3 * i + 2]};
    const double di[3] = {p[3 * i] + mu[3 * i], p[3 * i + 1] + mu[3 * i + 1], p[3 * i + 2] + mu[3 * i + 2]};
    const int c = count[i], cn = count_near[i];
    const int *row = rows + (size_t)i * cap;
    double ph = 0.0, s = 0.0;
    for (int e = lane; e < c; e += ADJ_RL) {
        const int j = e < cn ? row[e] : row[cap - (c - cn) + (e - cn)];
        const double4 b = xr[j];
        double x = a.x - b.x, y = a.y - b.y, z = a.z - b.z;
        min_image(x, y, z, box);
        const double r2 = x * x + y * y + z * z;
        if (r2 >= rce2) continue;
        const double aij = 1.0 / sqrt(2.0 * (a.w + b.w));
        double G[2];
        pair_G<double, 2>(aij, beta, r2, G);
        const double lk[3] = {lam[3 * j], lam[3 * j + 1], lam[3 * j + 2]};
        const double dk[3] = {p[3 * j] + mu[3 * j], p[3 * j + 1] + mu[3 * j + 1], p[3 * j + 2] + mu[3 * j + 2]};
        const double lkx = lk[0] * x + lk[1] * y + lk[2] * z;
        const double lix = li[0] * x + li[1] * y + li[2] * z;
        const double dkx = dk[0] * x + dk[1] * y + dk[2] * z;
        const double dix = di[0] * x + di[1] * y + di[2] * z;
        ph += G[1] * lkx;
        const double f = 2.0 * aij * aij;
        const double H1 = TWO_OVER_SQRTPI * exp(-aij * aij * r2) * f, H2 = H1 * f;
        const double de = (qi * lkx - q[j] * lix) * H1 - H2 * (lix * dkx + dix * lkx) +
                          H1 * (li[0] * dk[0] + li[1] * dk[1] + li[2] * dk[2] + di[0] * lk[0] + di[1] * lk[1] + di[2] * lk[2]);
        s += aij * aij * aij * de;
    }
    ph = group_sum<ADJ_RL>(ph);
    s = group_sum<ADJ_RL>(s);
    if (lane == 0) {
        phi[i] = ph;
        rbar[i] = 2.0 * sqrt(a.w) * s;
    }
}

// q_bar_i = -(phi_real_i + sum_stencil theta1 theta2 theta3 phi_grid): the reciprocal potential of
// the spread dipoles at atom i (pme.py:378-388 with the plain splines), one thread per atom.
template <typename real, int P>
__global__ __launch_bounds__(256) void k_adjoint_potential(int n, const double *__restrict__ pos, PMEGeom g,
                                                           const real *__restrict__ grid, const double *__restrict__ phi_real,
                                                           double *__restrict__ q_bar) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int base[3];
    real f[3];
    stencil<real, P>(pos, i, g, base, f);
    real t1[P], t2[P], t3[P], dt[P];
    bspline<real, P, false>(f[0], t1, dt, nullptr);
    bspline<real, P, false>(f[1], t2, dt, nullptr);
    bspline<real, P, false>(f[2], t3, dt, nullptr);
    const int K1 = g.K[0], K2 = g.K[1], K3 = g.K[2];
    double acc = 0.0;
#pragma unroll
    for (int ix = 0; ix < P; ++ix) {
        int gx = base[0] + ix;
        gx -= gx >= K1 ? K1 : 0;
        real sy = 0;
#pragma unroll
        for (int iy = 0; iy < P; ++iy) {
            int gy = base[1] + iy;
            gy -= gy >= K2 ? K2 : 0;
            const real *line = grid + ((size_t)gx * K2 + gy) * K3;
            real sz = 0;
#pragma unroll
            for (int iz = 0; iz < P; ++iz) {
                int gz = base[2] + iz;
                gz -= gz >= K3 ? K3 : 0;
                sz += line[gz] * t3[iz];
            }
            sy += sz * t2[iy];
        }
        acc += (double)(sy * t1[ix]);
    }
    q_bar[i] = -(phi_real[i] + acc);
}

// alpha_bar_i = mu_i . lam_i / alpha_i^2 (0 for a non-polarizable atom, whose mu and lam are 0).
__global__ void k_adjoint_alpha(int n, const double *__restrict__ mu, const double *__restrict__ lam,
                                const double *__restrict__ alpha, double *__restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double a = alpha[i];
    const double ml = mu[3 * i] * lam[3 * i] + mu[3 * i + 1] * lam[3 * i + 1] + mu[3 * i + 2] * lam[3 * i + 2];
    out[i] = a != 0.0 ? ml / (a * a) : 0.0;
}

// cov_bar_m = p_bar[i_m] . u_m, u_m the unit vector from atom i_m to atom j_m (minimum image): the
// covalent dipole p_i = sum_m c_m u_m is linear in c_m.  Entry k of the owner-sorted list (ci, cj)
// is the caller's covalent dipole order[k].
__global__ void k_adjoint_cov(int ncov, const double *__restrict__ pos, Box box, const int *__restrict__ ci,
                              const int *__restrict__ cj, const int *__restrict__ order, const double *__restrict__ p_bar,
                              double *__restrict__ out) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= ncov) return;
    const int i = ci[k], j = cj[k];
    double vx = pos[3 * j] - pos[3 * i], vy = pos[3 * j + 1] - pos[3 * i + 1], vz = pos[3 * j + 2] - pos[3 * i + 2];
    min_image(vx, vy, vz, box);
    const double il = 1.0 / sqrt(vx * vx + vy * vy + vz * vz);
    out[order[k]] = (p_bar[3 * i] * vx + p_bar[3 * i + 1] * vy + p_bar[3 * i + 2] * vz) * il;
}

}  // namespace pgm
