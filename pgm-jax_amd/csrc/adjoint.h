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


// ---------------------------------------------------------------------------- parameter cotangents
// The force cotangent f_bar enters through second derivatives: f_bar . dF/dtheta at fixed mu is
// -D_v (dE/dtheta) and the induced dipoles' cotangent gains KE D_v field, D_v the derivative along
// r -> r + eps v, v = f_bar (F = -dE/dr at fixed mu).  These kernels carry (value, D_v value) pairs:
// forward-mode derivatives along v of the same sums, with the pair set and cutoff masks of the
// geometry held fixed, as autodiff of the reference's masked sums has them (forcefield.py:2232-2292).
struct dual {
    double v, d;
};
PGM_DEV dual operator+(dual a, dual b) { return {a.v + b.v, a.d + b.d}; }
PGM_DEV dual operator-(dual a, dual b) { return {a.v - b.v, a.d - b.d}; }
PGM_DEV dual operator*(dual a, dual b) { return {a.v * b.v, a.v * b.d + a.d * b.v}; }
PGM_DEV dual operator*(double s, dual a) { return {s * a.v, s * a.d}; }
PGM_DEV dual dot3(const dual *a, const dual *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

// D_v p_i of the covalent dipoles p_i = sum_m c_m u_m, u = w / |w|, w = r_j - r_i (minimum image):
// D_v u = (D_v w - u (u . D_v w)) / |w|.  One thread per atom over its owner-sorted list.
__global__ void k_perm_tangent(int n, const double *__restrict__ pos, const double *__restrict__ v, Box box,
                               const int *__restrict__ cstart, const int *__restrict__ cj, const double *__restrict__ cc,
                               double *__restrict__ pdot) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s[3] = {0.0, 0.0, 0.0};
    for (int m = cstart[i]; m < cstart[i + 1]; ++m) {
        const int j = cj[m];
        double wx = pos[3 * j] - pos[3 * i], wy = pos[3 * j + 1] - pos[3 * i + 1], wz = pos[3 * j + 2] - pos[3 * i + 2];
        min_image(wx, wy, wz, box);
        const double il = 1.0 / sqrt(wx * wx + wy * wy + wz * wz);
        const double u[3] = {wx * il, wy * il, wz * il};
        const double wd[3] = {v[3 * j] - v[3 * i], v[3 * j + 1] - v[3 * i + 1], v[3 * j + 2] - v[3 * i + 2]};
        const double uw = u[0] * wd[0] + u[1] * wd[1] + u[2] * wd[2];
#pragma unroll
        for (int a = 0; a < 3; ++a) s[a] += cc[m] * (wd[a] - u[a] * uw) * il;
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) pdot[3 * i + a] = s[a];
}

// Real-space sums over atom i's row (both parts of the compacted row), with their D_v, FP64:
//   sq    = sum_k q_k G0 + (d_k . x) G1                                     (dU_pair/dq_i)
//   field = sum_k q_k G1 x + G2 (d_k . x) x - G1 d_k                        (-dU_pair/dd_i)
//   sR    = sum_k a^3 [q_i q_k H0 + (q_i d_k.x - q_k d_i.x) H1 - H2 (d_i.x)(d_k.x) + H1 d_i.d_k]
//   srh   = sum_k 12 eps (s^12 - s^6) / rmin,   sse = sum_k sqrt(eps_k) (s^12 - 2 s^6)      (LJ)
// x = r_i - r_k, G_n the screened kernels (D_v G_n = -G_(n+1) x.D_v x), H_n = dG_n/da =
// (2/sqrt(pi)) (2a^2)^n exp(-a^2 r^2) (D_v H_n = -2 a^2 x.D_v x H_n), s = rmin / r, d = p + mu
// (D_v d = D_v p).  Electrostatic terms inside the elec cutoff, LJ between molecules inside the van
// der Waals cutoff (with eps = 0 too: d/d sqrt(eps_i) of eps_i eps_k is sqrt(eps_k)).
// out[16 i + ...]: (value, D_v) of sq, sR, srh, sse, field x, y, z; v == nullptr: D_v = 0.
__global__ __launch_bounds__(256) void k_param_rows(int n, const int *__restrict__ rows, const int *__restrict__ count,
                                                    const int *__restrict__ count_near, int cap, const double4 *__restrict__ xr,
                                                    Box box, double beta, double rce2, double rcv2, const double *__restrict__ q,
                                                    const double4 *__restrict__ ljp, const double *__restrict__ d,
                                                    const double *__restrict__ ddot, const double *__restrict__ v,
                                                    double *__restrict__ out) {
    const int lane = threadIdx.x % ADJ_RL;
    const int i = blockIdx.x * (blockDim.x / ADJ_RL) + threadIdx.x / ADJ_RL;
    if (i >= n) return;  // whole groups of ADJ_RL lanes leave together
    const double4 a = xr[i];
    const double qi = q[i];
    const double4 lji = ljp[i];
    dual di[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) di[c] = {d[3 * i + c], v ? ddot[3 * i + c] : 0.0};
    const int cnt = count[i], cn = count_near[i];
    const int *row = rows + (size_t)i * cap;
    dual sq{0, 0}, sR{0, 0}, srh{0, 0}, sse{0, 0}, fx{0, 0}, fy{0, 0}, fz{0, 0};
    for (int e = lane; e < cnt; e += ADJ_RL) {
        const int k = e < cn ? row[e] : row[cap - (cnt - cn) + (e - cn)];
        const double4 b = xr[k];
        double x0 = a.x - b.x, x1 = a.y - b.y, x2 = a.z - b.z;
        min_image(x0, x1, x2, box);
        const double r2 = x0 * x0 + x1 * x1 + x2 * x2;
        dual x[3] = {{x0, 0.0}, {x1, 0.0}, {x2, 0.0}};
        if (v)
#pragma unroll
            for (int c = 0; c < 3; ++c) x[c].d = v[3 * i + c] - v[3 * k + c];
        const double xd = x[0].v * x[0].d + x[1].v * x[1].d + x[2].v * x[2].d;
        if (r2 < rce2) {
            const double aij = 1.0 / sqrt(2.0 * (a.w + b.w));
            double G[4];
            pair_G<double, 4>(aij, beta, r2, G);
            const dual G0{G[0], -G[1] * xd}, G1{G[1], -G[2] * xd}, G2{G[2], -G[3] * xd};
            const double f2 = 2.0 * aij * aij;
            const double h0 = TWO_OVER_SQRTPI * exp(-aij * aij * r2);
            const dual H0{h0, -f2 * xd * h0}, H1 = f2 * H0, H2 = f2 * H1;
            const double qk = q[k];
            dual dk[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) dk[c] = {d[3 * k + c], v ? ddot[3 * k + c] : 0.0};
            const dual dkx = dot3(dk, x), dix = dot3(di, x), didk = dot3(di, dk);
            sq = sq + qk * G0 + dkx * G1;
            const dual radial = qk * G1 + G2 * dkx;
            fx = fx + radial * x[0] - G1 * dk[0];
            fy = fy + radial * x[1] - G1 * dk[1];
            fz = fz + radial * x[2] - G1 * dk[2];
            const dual dea = (qi * qk) * H0 + (qi * dkx - qk * dix) * H1 - H2 * dix * dkx + H1 * didk;
            sR = sR + (aij * aij * aij) * dea;
        }
        const double4 ljk = ljp[k];
        if ((int)ljk.z != (int)lji.z && r2 < rcv2) {
            const double rmin = lji.x + ljk.x, eps = lji.y * ljk.y;
            const dual ir2{1.0 / r2, -2.0 * xd / (r2 * r2)};
            const dual s2 = (rmin * rmin) * ir2;
            const dual s6 = s2 * s2 * s2;
            const dual s12 = s6 * s6;
            if (eps != 0.0) srh = srh + (12.0 * eps / rmin) * (s12 - s6);
            sse = sse + ljk.y * (s12 - 2.0 * s6);
        }
    }
    double vals[14] = {sq.v, sq.d, sR.v, sR.d, srh.v, srh.d, sse.v, sse.d, fx.v, fx.d, fy.v, fy.d, fz.v, fz.d};
#pragma unroll
    for (int k = 0; k < 14; ++k) vals[k] = group_sum<ADJ_RL>(vals[k]);
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < 14; ++k) out[16 * (size_t)i + k] = vals[k];
}

// Per-axis spline weights of atom i as (value, D_v) pairs: T[a][k] = theta, D[a][k] = dtheta/dw with
// D_v w_a = v . escale column a (w = K frac(r H^-1)).
template <int P>
PGM_DEV void dual_splines(const double *__restrict__ pos, const double *__restrict__ v, int i, const PMEGeom &g, int *base,
                          dual (*T)[P], dual (*D)[P]) {
    double f[3];
    stencil<double, P>(pos, i, g, base, f);
    const double *es = g.escale;
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
        double t[P], dt[P], d2t[P];
        bspline<double, P, true>(f[ax], t, dt, d2t);
        const double wd = v ? v[3 * i] * es[ax] + v[3 * i + 1] * es[3 + ax] + v[3 * i + 2] * es[6 + ax] : 0.0;
#pragma unroll
        for (int k = 0; k < P; ++k) {
            T[ax][k] = {t[k], dt[k] * wd};
            D[ax][k] = {dt[k], d2t[k] * wd};
        }
    }
}

// D_v of the PME grid Q of charges q and dipoles d (pme.py:306-337): the weight
// q t1 t2 t3 + e1 d1 t2 t3 + e2 t1 d2 t3 + e3 t1 t2 d3 (e = d . escale) differentiated along v
// through the splines and through d (D_v d = ddot); one thread per atom, atomic adds.
template <typename real, int P>
__global__ __launch_bounds__(256) void k_spread_tangent(int n, const double *__restrict__ pos, const double *__restrict__ v,
                                                        PMEGeom g, const double *__restrict__ q, const double *__restrict__ d,
                                                        const double *__restrict__ ddot, real *__restrict__ grid) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int base[3];
    dual T[3][P], D[3][P];
    dual_splines<P>(pos, v, i, g, base, T, D);
    const double *es = g.escale;
    dual e[3];
#pragma unroll
    for (int b = 0; b < 3; ++b) {
        e[b].v = d[3 * i] * es[b] + d[3 * i + 1] * es[3 + b] + d[3 * i + 2] * es[6 + b];
        e[b].d = ddot[3 * i] * es[b] + ddot[3 * i + 1] * es[3 + b] + ddot[3 * i + 2] * es[6 + b];
    }
    const dual qi{q[i], 0.0};
    const int K1 = g.K[0], K2 = g.K[1], K3 = g.K[2];
    for (int ix = 0; ix < P; ++ix) {
        int gx = base[0] + ix;
        gx -= gx >= K1 ? K1 : 0;
        const dual ax = qi * T[0][ix] + e[0] * D[0][ix];  // charge + dipole along x
        const dual bx = e[1] * T[0][ix], cx = e[2] * T[0][ix];
        for (int iy = 0; iy < P; ++iy) {
            int gy = base[1] + iy;
            gy -= gy >= K2 ? K2 : 0;
            const dual axy = ax * T[1][iy] + bx * D[1][iy];
            const dual cxy = cx * T[1][iy];
            real *line = grid + ((size_t)gx * K2 + gy) * K3;
#pragma unroll
            for (int iz = 0; iz < P; ++iz) {
                int gz = base[2] + iz;
                gz -= gz >= K3 ? K3 : 0;
                const dual w = axy * T[2][iz] + cxy * D[2][iz];
                atomicAdd(line + gz, (real)w.d);
            }
        }
    }
}

// Gathers of the potential phi (grid) with its D_v (gridd, the transform of the tangent spread),
// one thread per atom: phi_i = sum t1 t2 t3 phi and dU_rec/dd_i = escale (sum d1 t2 t3 phi, sum t1 d2
// t3 phi, sum t1 t2 d3 phi) as (value, D_v) pairs: out[8 i + ...] = phi, D_v phi, dU/dd (3), D_v dU/dd (3).
template <typename real, int P>
__global__ __launch_bounds__(256) void k_gather_tangent(int n, const double *__restrict__ pos, const double *__restrict__ v,
                                                        PMEGeom g, const real *__restrict__ grid, const real *__restrict__ gridd,
                                                        double *__restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int base[3];
    dual T[3][P], D[3][P];
    dual_splines<P>(pos, v, i, g, base, T, D);
    const int K1 = g.K[0], K2 = g.K[1], K3 = g.K[2];
    dual pot{0, 0}, g1{0, 0}, g2{0, 0}, g3{0, 0};
    for (int ix = 0; ix < P; ++ix) {
        int gx = base[0] + ix;
        gx -= gx >= K1 ? K1 : 0;
        dual s_tt{0, 0}, s_dt{0, 0}, s_td{0, 0};
        for (int iy = 0; iy < P; ++iy) {
            int gy = base[1] + iy;
            gy -= gy >= K2 ? K2 : 0;
            const size_t lo = ((size_t)gx * K2 + gy) * K3;
            dual lt{0, 0}, ld{0, 0};
#pragma unroll
            for (int iz = 0; iz < P; ++iz) {
                int gz = base[2] + iz;
                gz -= gz >= K3 ? K3 : 0;
                const dual ph{(double)grid[lo + gz], gridd ? (double)gridd[lo + gz] : 0.0};
                lt = lt + ph * T[2][iz];
                ld = ld + ph * D[2][iz];
            }
            s_tt = s_tt + lt * T[1][iy];
            s_dt = s_dt + lt * D[1][iy];
            s_td = s_td + ld * T[1][iy];
        }
        pot = pot + s_tt * T[0][ix];
        g1 = g1 + s_tt * D[0][ix];
        g2 = g2 + s_dt * T[0][ix];
        g3 = g3 + s_td * T[0][ix];
    }
    const double *es = g.escale;
    double *o = out + 8 * (size_t)i;
    o[0] = pot.v;
    o[1] = pot.d;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const dual r = es[3 * a] * g1 + es[3 * a + 1] * g2 + es[3 * a + 2] * g3;
        o[2 + a] = r.v;
        o[5 + a] = r.d;
    }
}

// Per-atom parameter cotangents without the adjoint solve's part, and the dipoles' effective
// cotangent for it (ebar dE/dtheta - D_v dE/dtheta, mu_bar + KE D_v field):
//   q:       KE [sq + phi_rec - 2 beta q / sqrt(pi) - pi Q / (V beta^2)]
//   R:       -2 R KE sR,   alpha: -KE |mu|^2 / (2 alpha^2),   R*/2, sqrt(eps): srh, sse + the tail's
//   field    = field_rows - dU_rec/dd + c_self d  (kept in fld[6 i ..] with its D_v for the covalent dipoles)
__global__ void k_vjp_atoms(int n, const double *__restrict__ rowsum, const double *__restrict__ rec,
                            const double *__restrict__ q, const double4 *__restrict__ xr, const double *__restrict__ alpha,
                            const double *__restrict__ mu, const double *__restrict__ d, const double *__restrict__ ddot,
                            const double *__restrict__ mu_bar, const double *__restrict__ tail, double ebar, double qconst,
                            double qscale, double cself, double *__restrict__ q_bar, double *__restrict__ r_bar,
                            double *__restrict__ a_bar, double *__restrict__ rh_bar, double *__restrict__ se_bar,
                            double *__restrict__ fld, double *__restrict__ mu_eff) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double *s = rowsum + 16 * (size_t)i, *rc = rec + 8 * (size_t)i;
    const double dEq = KE * (s[0] + rc[0] + qscale * q[i] + qconst), ddEq = KE * (s[1] + rc[1]);
    q_bar[i] = ebar * dEq - ddEq;
    const double m2R = -2.0 * sqrt(xr[i].w) * KE;
    r_bar[i] = ebar * m2R * s[2] - m2R * s[3];
    const double al = alpha[i];
    const double mm = mu[3 * i] * mu[3 * i] + mu[3 * i + 1] * mu[3 * i + 1] + mu[3 * i + 2] * mu[3 * i + 2];
    a_bar[i] = al != 0.0 ? ebar * (-KE * mm / (2.0 * al * al)) : 0.0;
    rh_bar[i] = ebar * (s[4] + (tail ? tail[2 * i] : 0.0)) - s[5];
    se_bar[i] = ebar * (s[6] + (tail ? tail[2 * i + 1] : 0.0)) - s[7];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const double f = s[8 + 2 * a] - rc[2 + a] + cself * d[3 * i + a];
        const double fd = s[9 + 2 * a] - rc[5 + a] + cself * (ddot ? ddot[3 * i + a] : 0.0);
        fld[6 * (size_t)i + a] = f;
        fld[6 * (size_t)i + 3 + a] = fd;
        mu_eff[3 * i + a] = (mu_bar ? mu_bar[3 * i + a] : 0.0) + KE * fd;
    }
}

// cov_bar_m = ebar dE/dc_m - D_v dE/dc_m with dE/dc_m = -KE field_i . u_m (i the owner):
// -KE ebar field . u + KE (D_v field . u + field . D_v u).  Entry k of the owner-sorted list is the
// caller's covalent dipole order[k].
__global__ void k_vjp_cov(int ncov, const double *__restrict__ pos, const double *__restrict__ v, Box box,
                          const int *__restrict__ ci, const int *__restrict__ cj, const int *__restrict__ order,
                          const double *__restrict__ fld, double ebar, double *__restrict__ out) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= ncov) return;
    const int i = ci[k], j = cj[k];
    double wx = pos[3 * j] - pos[3 * i], wy = pos[3 * j + 1] - pos[3 * i + 1], wz = pos[3 * j + 2] - pos[3 * i + 2];
    min_image(wx, wy, wz, box);
    const double il = 1.0 / sqrt(wx * wx + wy * wy + wz * wz);
    const double u[3] = {wx * il, wy * il, wz * il};
    double wd[3] = {0.0, 0.0, 0.0};
    if (v)
        for (int a = 0; a < 3; ++a) wd[a] = v[3 * j + a] - v[3 * i + a];
    const double uw = u[0] * wd[0] + u[1] * wd[1] + u[2] * wd[2];
    const double *f = fld + 6 * (size_t)i;
    double fu = 0.0, dfu = 0.0;
    for (int a = 0; a < 3; ++a) {
        fu += f[a] * u[a];
        dfu += f[3 + a] * u[a] + f[a] * (wd[a] - u[a] * uw) * il;
    }
    out[order[k]] = -KE * ebar * fu + KE * dfu;
}

// z = x + y (n doubles).
__global__ void k_add_vec(int n, const double *__restrict__ x, const double *__restrict__ y, double *__restrict__ z) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) z[i] = x[i] + y[i];
}

// y += x (n doubles).
__global__ void k_add(int n, const double *__restrict__ x, double *__restrict__ y) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] += x[i];
}

}  // namespace pgm
