// virial.h -- the molecular strain derivative dE/d eps at fixed induced dipoles
// (pgm_jax/md/forcefield.py:2628-2695, which takes it by autodiff of the whole energy through
// full_strain_derivative, :608-662; written out here).
//
// Under x -> x + (com eps^T)[mol], H -> H (1 + eps)^T (molecules translated with their centres of
// mass, dipoles held fixed):
//   pairs:  W += sum de/dx (x) X,  X = x - o_i + o_k, o = r - com        (k_pair_forces<VIRIAL>)
//   PME:    fractional coordinates move by -o eps^T H^-1, e = d H^-1 K by -d eps^T H^-1 K:
//           W -= KE sum (dU_rec/dr (x) o + dU_rec/dd (x) d)               (k_virial_atoms)
//           plus the influence-function term                              (k_convolve_virial)
//   background and LJ tail scale as 1/V (host constants).
// All sums are FP64 per-block partials reduced in a fixed order: the result is deterministic.
#pragma once

#include "common.h"

namespace pgm {

// Centres of mass, one thread per molecule over its atoms in index order (molecules whole: no
// minimum image, as pgm_jax/md/box.py:340-359 requires).
__global__ void k_com(int nmol, const int *__restrict__ mol_ptr, const int *__restrict__ mol_atoms,
                      const double *__restrict__ pos, const double *__restrict__ mass, double *__restrict__ com) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= nmol) return;
    double sx = 0, sy = 0, sz = 0, sm = 0;
    for (int k = mol_ptr[m]; k < mol_ptr[m + 1]; ++k) {
        const int i = mol_atoms[k];
        const double w = mass[i];
        sx += w * pos[3 * i];
        sy += w * pos[3 * i + 1];
        sz += w * pos[3 * i + 2];
        sm += w;
    }
    const double inv = sm != 0.0 ? 1.0 / sm : 0.0;
    com[3 * m] = sx * inv;
    com[3 * m + 1] = sy * inv;
    com[3 * m + 2] = sz * inv;
}

// o_i = r_i - com[mol_i].
__global__ void k_offsets(int n, const double *__restrict__ pos, const int *__restrict__ mol,
                          const double *__restrict__ com, double *__restrict__ off) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int m = mol[i];
#pragma unroll
    for (int a = 0; a < 3; ++a) off[3 * i + a] = pos[3 * i + a] - com[3 * m + a];
}

// Per block: sum_i dU_rec/dr_i (x) o_i + dU_rec/dd_i (x) d_i (row-major 3x3, without KE).
__global__ __launch_bounds__(256) void k_virial_atoms(int n, const double *__restrict__ rec_dr,
                                                      const double *__restrict__ off, const double *__restrict__ rec_dd,
                                                      const double *__restrict__ dd, double *__restrict__ part) {
    __shared__ double lds[9 * 16];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double w[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (i < n) {
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int b = 0; b < 3; ++b)
                w[3 * a + b] = rec_dr[3 * i + a] * off[3 * i + b] + rec_dd[3 * i + a] * dd[3 * i + b];
    }
    block_sum<9>(w, lds);
    if (threadIdx.x == 0)
        for (int k = 0; k < 9; ++k) part[9 * blockIdx.x + k] = w[k];
}

// Atomic scaling strains the covalent dipole frames: per block sum_m t_m (x) v_m over covalent
// dipoles m (owner i = ci[m], partner j = cj[m]), v = min_image(r_j - r_i), u = v/|v|,
// t = c_m (g - (g.u) u)/|v| with g = gp_i the total dE/dd of the owner (KE included).
__global__ __launch_bounds__(256) void k_virial_cov(int ncov, const double *__restrict__ pos, Box box,
                                                    const int *__restrict__ ci, const int *__restrict__ cj,
                                                    const double *__restrict__ cc, const double *__restrict__ gp,
                                                    double *__restrict__ part) {
    __shared__ double lds[9 * 16];
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    double w[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (m < ncov) {
        const int i = ci[m], j = cj[m];
        double v[3] = {pos[3 * j] - pos[3 * i], pos[3 * j + 1] - pos[3 * i + 1], pos[3 * j + 2] - pos[3 * i + 2]};
        min_image(v[0], v[1], v[2], box);
        const double L = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        const double u[3] = {v[0] / L, v[1] / L, v[2] / L};
        const double g[3] = {gp[3 * i], gp[3 * i + 1], gp[3 * i + 2]};
        const double gu = g[0] * u[0] + g[1] * u[1] + g[2] * u[2];
        const double f = cc[m] / L;
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int b = 0; b < 3; ++b) w[3 * a + b] = f * (g[a] - gu * u[a]) * v[b];
    }
    block_sum<9>(w, lds);
    if (threadIdx.x == 0)
        for (int k = 0; k < 9; ++k) part[9 * blockIdx.x + k] = w[k];
}

// One block: W = pair/2 + KE (influence - E_rec I) - KE atoms + cov - (E_bg + 2 E_tail) I  (out[9]).
// pair: np_pair blocks x 9, infl: np_conv blocks x 7 (xx yy zz xy xz yz E), atoms: np_atom x 9,
// cov: np_cov x 9 (0 blocks under molecular scaling).
__global__ void k_virial_final(const double *__restrict__ pair, int np_pair, const double *__restrict__ infl, int np_conv,
                               const double *__restrict__ atoms, int np_atom, const double *__restrict__ cov, int np_cov,
                               double diag_const, double *__restrict__ out) {
    __shared__ double lds[32];
    double W[9];
    for (int k = 0; k < 9; ++k) {
        double v = 0.0;
        for (int b = threadIdx.x; b < np_pair; b += blockDim.x) v += pair[9 * b + k];
        double a1[1] = {v};
        block_sum<1>(a1, lds);
        double u = 0.0;
        for (int b = threadIdx.x; b < np_atom; b += blockDim.x) u += atoms[9 * b + k];
        double a2[1] = {u};
        block_sum<1>(a2, lds);
        double t = 0.0;
        for (int b = threadIdx.x; b < np_cov; b += blockDim.x) t += cov[9 * b + k];
        double a3[1] = {t};
        block_sum<1>(a3, lds);
        W[k] = 0.5 * a1[0] - KE * a2[0] + a3[0];
    }
    double I[7];
    for (int k = 0; k < 7; ++k) {
        double v = 0.0;
        for (int b = threadIdx.x; b < np_conv; b += blockDim.x) v += infl[7 * b + k];
        double a1[1] = {v};
        block_sum<1>(a1, lds);
        I[k] = a1[0];
    }
    if (threadIdx.x == 0) {
        const double e = I[6];
        const int sym[9] = {0, 3, 4, 3, 1, 5, 4, 5, 2};
        for (int k = 0; k < 9; ++k) {
            double v = W[k] + KE * I[sym[k]];
            if (k % 4 == 0) v += -KE * e + diag_const;
            out[k] = v;
        }
    }
}

}  // namespace pgm
