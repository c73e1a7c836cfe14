// cg.h -- induced-dipole conjugate gradients on the device, covalent dipole frames, force and
// energy assembly.
//
// CG (pgm_jax/md/forcefield.py:1714-1821): Jacobi preconditioner z = alpha r, flexible
// Polak-Ribiere beta, converged when max|alpha r| / norm <= tol or after max_iter iterations, then
// the peek step mu = x + peek alpha r.  Vectors are real4 (x, y, z, 0) in the compute precision;
// every inner product and the residual maximum are FP64 block partials, reduced in a fixed order
// by every block of the next kernel (deterministic, no atomics).  Iteration k's kernels run only
// while k < *conv_it; the update kernel of iteration k sets *conv_it = k + 1 when it converges, so
// the host can queue iterations ahead of the convergence test.
//
// Beads: every vector kernel runs on a (blocks per bead, beads) grid, and each bead has its own
// CGState, rz and err rows (stride rzs) and partials, so the beads' solves are independent CGs
// (bead-wise FP64 reductions) advanced by the same launches; a converged bead's blocks return.
#pragma once

#include "common.h"

namespace pgm {

// A uniform external field in internal units (e/nm^2; zero when off), or KE times it, by value.
struct Ext3 {
    double x, y, z;
};

struct CGState {
    int conv_it;  // first iteration index that must not run (INT_MAX while unconverged)
    unsigned seq; // sequence number of the solve (set by k_cg_start_finish; stamps the host mirror)
    double norm;  // mean |alpha b| (convergence normaliser)
};
static_assert(sizeof(CGState) == CG_STRIDE * sizeof(int), "CGState is CG_STRIDE ints");

// The CG state after iteration k - 1 of bead b for the host: one 64-bit store into pinned host
// memory (slot k % 4) of the solve's sequence number, k and conv_it (clamped to 0xffff, which stands
// for "not converged"), so the host learns of convergence without a copy and an event per check.
PGM_DEV void cg_mirror(unsigned long long *mirror, int nbead, int b, unsigned seq, int k, int conv) {
    if (!mirror) return;
    const unsigned long long c16 = conv < 0xffff ? (unsigned long long)conv : 0xffffull;
    mirror[(size_t)(k & 3) * nbead + b] = ((unsigned long long)seq << 32) | ((unsigned long long)(k & 0xffff) << 16) | c16;
}

// Sum of np partials (stride `stride`, component c) by the whole block, fixed order; result in all threads.
PGM_DEV double block_reduce_partials(const double *__restrict__ part, int np, int stride, int c, double *lds, bool is_max) {
    double v = 0.0;
    for (int b = threadIdx.x; b < np; b += blockDim.x) {
        const double x = part[b * stride + c];
        v = is_max ? fmax(v, x) : v + x;
    }
    double r;
    if (is_max) {
        r = block_max(v, lds);
    } else {
        double a[1] = {v};
        block_sum<1>(a, lds);
        r = a[0];
    }
    __syncthreads();
    if (threadIdx.x == 0) lds[31] = r;
    __syncthreads();
    r = lds[31];
    __syncthreads();
    return r;
}

// ---------------------------------------------------------------------------- small vector kernels
// x0 = sum_k c_k hist[k] (predictor polynomial over the converged dipoles, newest first).
template <typename real>
__global__ void k_predict(int n, const double *__restrict__ hist, double c0, double c1, double c2, double c3,
                          real4<real> *__restrict__ x0) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const size_t s = (size_t)n * 3;
    real v[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const size_t o = (size_t)3 * i + c;
        v[c] = (real)(c0 * hist[o] + c1 * hist[s + o] + c2 * hist[2 * s + o] + c3 * hist[3 * s + o]);
    }
    real4<real> r;
    r.x = v[0];
    r.y = v[1];
    r.z = v[2];
    r.w = 0;
    x0[i] = r;
}

// x0 = 0 where alpha = 0: atoms without polarizability keep a zero induced dipole
// (forcefield.py:767-780).
template <typename real>
__global__ void k_mask_alpha0(int n, const real *__restrict__ alpha, real4<real> *__restrict__ x0) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || alpha[i] != real(0)) return;
    real4<real> z;
    z.x = z.y = z.z = z.w = real(0);
    x0[i] = z;
}

// x0 = mu (the initial guess of a trial solve: the dipoles passed in, no predictor).
template <typename real>
__global__ void k_mu_to_x0(int n, const double *__restrict__ mu, real4<real> *__restrict__ x0) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    real4<real> r;
    r.x = (real)mu[3 * i];
    r.y = (real)mu[3 * i + 1];
    r.z = (real)mu[3 * i + 2];
    r.w = 0;
    x0[i] = r;
}

// dq = (p + x0, q or 0): the dipoles and charges a field evaluation spreads.
template <typename real>
__global__ void k_pack_dq(int n, const double *__restrict__ p, const real4<real> *__restrict__ x0, const double *__restrict__ q,
                          real4<real> *__restrict__ dq) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    real4<real> r;
    r.x = (real)p[3 * i];
    r.y = (real)p[3 * i + 1];
    r.z = (real)p[3 * i + 2];
    if (x0) {
        const real4<real> a = x0[i];
        r.x += a.x;
        r.y += a.y;
        r.z += a.z;
    }
    r.w = q ? (real)q[i] : real(0);
    dq[i] = r;
}

// x = alpha * b
template <typename real>
__global__ void k_scale_alpha(int n, const real4<real> *__restrict__ b, const real *__restrict__ alpha, real4<real> *__restrict__ x) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const real a = alpha[i];
    real4<real> v = b[i];
    v.x *= a;
    v.y *= a;
    v.z *= a;
    v.w = 0;
    x[i] = v;
}

// real (n) -> double (n)
template <typename real> __global__ void k_to_double(size_t n, const real *__restrict__ v, double *__restrict__ out) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (double)v[i];
}

// real4 (N) -> double (N, 3)
template <typename real>
__global__ void k_to_double3(int n, const real4<real> *__restrict__ v, double *__restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const real4<real> a = v[i];
    out[3 * i] = a.x;
    out[3 * i + 1] = a.y;
    out[3 * i + 2] = a.z;
}

// norm = sum(partials) / (3 N) + 1e-300 (mean |alpha b|); one block per bead (np partials each).
__global__ void k_set_norm(const double *__restrict__ part, int np, double inv3n, CGState *st) {
    __shared__ double lds[32];
    const double s = block_reduce_partials(part + (size_t)blockIdx.x * np, np, 1, 0, lds, false);
    if (threadIdx.x == 0) st[blockIdx.x].norm = s * inv3n + 1e-300;
}

__global__ void k_set_norm_value(CGState *st, double norm) {
    if (threadIdx.x == 0) st->norm = norm;
}

// ---------------------------------------------------------------------------- CG
// Start: z = alpha r, p = z, x = x0 (0 when x0 is null), partials (r . z, max |alpha r|).
template <typename real>
__global__ __launch_bounds__(256) void k_cg_start(int nper, const real4<real> *__restrict__ r, const real *__restrict__ alpha,
                                                  real4<real> *__restrict__ z, real4<real> *__restrict__ p,
                                                  double *__restrict__ part, const real4<real> *__restrict__ x0,
                                                  real4<real> *__restrict__ x) {
    __shared__ double lds[32];
    const int li = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y * nper + li;
    const int pb = blockIdx.y * gridDim.x + blockIdx.x;
    double s[1] = {0.0};
    double m = 0.0;
    if (li < nper) {
        const real a = alpha[i];
        const real4<real> rr = r[i];
        real4<real> zz;
        zz.x = rr.x * a;
        zz.y = rr.y * a;
        zz.z = rr.z * a;
        zz.w = 0;
        z[i] = zz;
        p[i] = zz;
        x[i] = x0 ? x0[i] : real4<real>{};
        s[0] = (double)(rr.x * zz.x) + (double)(rr.y * zz.y) + (double)(rr.z * zz.z);
        m = fmax(fmax(fabs((double)zz.x), fabs((double)zz.y)), fabs((double)zz.z));
    }
    block_sum<1>(s, lds);
    m = block_max(m, lds + 16);
    if (threadIdx.x == 0) {
        part[2 * pb] = s[0];
        part[2 * pb + 1] = m;
    }
}

// rz[0], err[0], conv_it (one block per bead; np partials per bead, rows of rzs).
__global__ void k_cg_start_finish(const double *__restrict__ part, int np, CGState *st, double *__restrict__ rz,
                                  double *__restrict__ err, int rzs, double tol, int max_iter,
                                  unsigned long long *mirror, unsigned seq) {
    __shared__ double lds[32];
    const int b = blockIdx.x;
    part += (size_t)b * np * 2;
    st += b;
    rz += (size_t)b * rzs;
    err += (size_t)b * rzs;
    const double s = block_reduce_partials(part, np, 2, 0, lds, false);
    const double m = block_reduce_partials(part, np, 2, 1, lds, true);
    if (threadIdx.x == 0) {
        rz[0] = s;
        const double e = m / st->norm;
        err[0] = e;
        st->conv_it = (e > tol && 0 < max_iter) ? 0x7fffffff : 0;
        st->seq = seq;
        cg_mirror(mirror, gridDim.x, b, seq, 0, st->conv_it);
    }
}

// Iteration k, first half: al = rz_k / max(p . Ap, 1e-300) (rounded to real); x += al p;
// r' = r - al Ap; z' = alpha r'; partials (z' . (r' - r), r' . z', max |alpha r'|).
template <typename real>
__global__ __launch_bounds__(256) void k_cg_update1(int nper, const double *__restrict__ pap_part, int np,
                                                    const double *__restrict__ rz, int rzs, int k, real4<real> *__restrict__ x,
                                                    real4<real> *__restrict__ r, real4<real> *__restrict__ z,
                                                    const real4<real> *__restrict__ p, const real4<real> *__restrict__ Ap,
                                                    const real *__restrict__ alpha, double *__restrict__ part,
                                                    const CGState *__restrict__ st) {
    const int b = blockIdx.y;
    if (k >= st[b].conv_it) return;
    __shared__ double lds[32];
    const double pap = block_reduce_partials(pap_part + (size_t)b * np, np, 1, 0, lds, false);
    const real al = (real)(rz[(size_t)b * rzs + k] / fmax(pap, 1e-300));
    const int li = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = b * nper + li;
    const int pb = b * gridDim.x + blockIdx.x;
    double s[2] = {0.0, 0.0};
    double m = 0.0;
    if (li < nper) {
        const real4<real> pp = p[i], ap = Ap[i], r0 = r[i];
        real4<real> xx = x[i];
        xx.x += al * pp.x;
        xx.y += al * pp.y;
        xx.z += al * pp.z;
        x[i] = xx;
        real4<real> r1;
        r1.x = r0.x - al * ap.x;
        r1.y = r0.y - al * ap.y;
        r1.z = r0.z - al * ap.z;
        r1.w = 0;
        r[i] = r1;
        const real a = alpha[i];
        real4<real> z1;
        z1.x = r1.x * a;
        z1.y = r1.y * a;
        z1.z = r1.z * a;
        z1.w = 0;
        z[i] = z1;
        s[0] = (double)(z1.x * (r1.x - r0.x)) + (double)(z1.y * (r1.y - r0.y)) + (double)(z1.z * (r1.z - r0.z));
        s[1] = (double)(r1.x * z1.x) + (double)(r1.y * z1.y) + (double)(r1.z * z1.z);
        m = fmax(fmax(fabs((double)z1.x), fabs((double)z1.y)), fabs((double)z1.z));
    }
    block_sum<2>(s, lds);
    m = block_max(m, lds + 16);
    if (threadIdx.x == 0) {
        part[3 * pb] = s[0];
        part[3 * pb + 1] = s[1];
        part[3 * pb + 2] = m;
    }
}

// Iteration k, second half: beta = z'.(r' - r) / max(rz_k, 1e-300) (rounded to real); p = z' + beta p.
// Block 0 of each bead records rz[k+1], err[k+1] and, when converged or out of iterations,
// conv_it = k + 1, and mirrors the state to the host (slot k + 1; also when the iteration is a no-op).
template <typename real>
__global__ __launch_bounds__(256) void k_cg_update2(int nper, const double *__restrict__ part, int np, double *__restrict__ rz,
                                                    double *__restrict__ err, int rzs, int k, const real4<real> *__restrict__ z,
                                                    real4<real> *__restrict__ p, CGState *st, double tol, int max_iter,
                                                    unsigned long long *mirror) {
    const int b = blockIdx.y;
    const int conv0 = st[b].conv_it;
    if (k >= conv0) {
        if (blockIdx.x == 0 && threadIdx.x == 0) cg_mirror(mirror, gridDim.y, b, st[b].seq, k + 1, conv0);
        return;
    }
    __shared__ double lds[32];
    part += (size_t)b * np * 3;
    rz += (size_t)b * rzs;
    err += (size_t)b * rzs;
    const double s0 = block_reduce_partials(part, np, 3, 0, lds, false);
    const double s1 = block_reduce_partials(part, np, 3, 1, lds, false);
    const double mx = block_reduce_partials(part, np, 3, 2, lds, true);
    const real beta = (real)(s0 / fmax(rz[k], 1e-300));
    const int li = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = b * nper + li;
    if (li < nper) {
        const real4<real> zz = z[i];
        real4<real> pp = p[i];
        pp.x = zz.x + beta * pp.x;
        pp.y = zz.y + beta * pp.y;
        pp.z = zz.z + beta * pp.z;
        pp.w = 0;
        p[i] = pp;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        rz[k + 1] = s1;
        const double e = mx / st[b].norm;
        err[k + 1] = e;
        const int conv = (e > tol && k + 1 < max_iter) ? conv0 : k + 1;
        if (conv != conv0) st[b].conv_it = conv;
        cg_mirror(mirror, gridDim.y, b, st[b].seq, k + 1, conv);
    }
}

// ---------------------------------------------------------------------------- molecule-block preconditioner
// Optional (3-site molecules): z = M_I^-1 r over the nine components of molecule I, with
// M_I = diag(1/alpha_p) + the dipole tensors B1 I - B2 x x^T (x = r_p - r_k, minimum image) between
// its atoms for the full Gaussian kernel erf(a r)/r -- the direct plus reciprocal intramolecular
// coupling of the operator, periodic images left out.  The CG is the same flexible PCG with another
// symmetric positive definite preconditioner; the convergence test stays max|alpha r| / norm.
// M_I^-1 (81 reals, row major) per molecule by Gauss-Jordan in FP64, remade with the neighbour rows
// (the molecules turn little in between, and any such blocks are a valid preconditioner).
template <typename real>
__global__ __launch_bounds__(128) void k_mol_precond(int nmol, const double4 *__restrict__ xr, const double *__restrict__ alpha, Box box,
                              real *__restrict__ minv) {
    const int I = blockIdx.x * blockDim.x + threadIdx.x;
    if (I >= nmol) return;
    // every loop unrolled, so the 9 x 9 matrix stays in registers (no scratch)
    double M[9][9];
#pragma unroll
    for (int r = 0; r < 9; ++r)
#pragma unroll
        for (int c = 0; c < 9; ++c) M[r][c] = 0.0;
    double4 at[3];
#pragma unroll
    for (int p = 0; p < 3; ++p) at[p] = xr[3 * I + p];
#pragma unroll
    for (int p = 0; p < 3; ++p) {
        const double ia = 1.0 / alpha[3 * I + p];
#pragma unroll
        for (int c = 0; c < 3; ++c) M[3 * p + c][3 * p + c] = ia;
    }
#pragma unroll
    for (int p = 0; p < 3; ++p)
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            if (p == k) continue;
            const double4 a = at[p], b = at[k];
            double x[3] = {a.x - b.x, a.y - b.y, a.z - b.z};
            min_image(x[0], x[1], x[2], box);
            const double r2 = x[0] * x[0] + x[1] * x[1] + x[2] * x[2];
            const double ir = 1.0 / sqrt(r2), r = r2 * ir;
            double B[3];
            one_B<double, 3>(1.0 / sqrt(2.0 * (a.w + b.w)), r, ir, ir * ir, B);
#pragma unroll
            for (int u = 0; u < 3; ++u)
#pragma unroll
                for (int v = 0; v < 3; ++v) M[3 * p + u][3 * k + v] = (u == v ? B[1] : 0.0) - B[2] * x[u] * x[v];
        }
    // in-place Gauss-Jordan inversion (M symmetric positive definite: no pivoting)
#pragma unroll
    for (int col = 0; col < 9; ++col) {
        const double d = 1.0 / M[col][col];
        M[col][col] = 1.0;
#pragma unroll
        for (int c = 0; c < 9; ++c) M[col][c] *= d;
#pragma unroll
        for (int r = 0; r < 9; ++r) {
            if (r == col) continue;
            const double f = M[r][col];
            M[r][col] = 0.0;
#pragma unroll
            for (int c = 0; c < 9; ++c) M[r][c] -= f * M[col][c];
        }
    }
    real *out = minv + (size_t)81 * I;
#pragma unroll
    for (int r = 0; r < 9; ++r)
#pragma unroll
        for (int c = 0; c < 9; ++c) out[9 * r + c] = (real)M[r][c];
}

// z of atom p (0..2) of the molecule whose nine residual components are sr[0..8]: rows 3p..3p+2 of M^-1.
template <typename real>
PGM_DEV real4<real> mol_apply(const real *__restrict__ Mi, int p, const real *sr) {
    real v[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        real s = real(0);
#pragma unroll
        for (int q = 0; q < 9; ++q) s += Mi[9 * (3 * p + c) + q] * sr[q];
        v[c] = s;
    }
    real4<real> z;
    z.x = v[0];
    z.y = v[1];
    z.z = v[2];
    z.w = 0;
    return z;
}

constexpr int MOLP_BLOCK = 192;  // 64 whole molecules per workgroup

// k_cg_start with the molecule blocks: z = M^-1 r, p = z, x = x0 (0 when x0 is null), partials
// (r . z, max |alpha r|).
template <typename real>
__global__ __launch_bounds__(MOLP_BLOCK) void k_cg_start_mol(int nper, const real4<real> *__restrict__ r,
                                                             const real *__restrict__ alpha, const real *__restrict__ minv,
                                                             real4<real> *__restrict__ z, real4<real> *__restrict__ p,
                                                             double *__restrict__ part, const real4<real> *__restrict__ x0,
                                                             real4<real> *__restrict__ x) {
    __shared__ real sr[3 * MOLP_BLOCK];
    __shared__ double lds[32];
    const int li = blockIdx.x * MOLP_BLOCK + threadIdx.x;
    const int i = blockIdx.y * nper + li;
    const int pb = blockIdx.y * gridDim.x + blockIdx.x;
    const bool in = li < nper;
    real4<real> rr;
    rr.x = rr.y = rr.z = rr.w = real(0);
    if (in) rr = r[i];
    sr[3 * threadIdx.x] = rr.x;
    sr[3 * threadIdx.x + 1] = rr.y;
    sr[3 * threadIdx.x + 2] = rr.z;
    __syncthreads();
    double s[1] = {0.0};
    double m = 0.0;
    if (in) {
        const int q = threadIdx.x / 3, pa = threadIdx.x - 3 * q;
        const real4<real> zz = mol_apply<real>(minv + (size_t)81 * (i / 3), pa, sr + 9 * q);
        z[i] = zz;
        p[i] = zz;
        x[i] = x0 ? x0[i] : real4<real>{};
        const real a = alpha[i];
        s[0] = (double)(rr.x * zz.x) + (double)(rr.y * zz.y) + (double)(rr.z * zz.z);
        m = fmax(fmax(fabs((double)(a * rr.x)), fabs((double)(a * rr.y))), fabs((double)(a * rr.z)));
    }
    block_sum<1>(s, lds);
    m = block_max(m, lds + 16);
    if (threadIdx.x == 0) {
        part[2 * pb] = s[0];
        part[2 * pb + 1] = m;
    }
}

// k_cg_update1 with the molecule blocks: x += al p, r' = r - al Ap, z' = M^-1 r', partials
// (z' . (r' - r), r' . z', max |alpha r'|).
template <typename real>
__global__ __launch_bounds__(MOLP_BLOCK) void k_cg_update1_mol(int nper, const double *__restrict__ pap_part, int np,
                                                               const double *__restrict__ rz, int rzs, int k,
                                                               real4<real> *__restrict__ x, real4<real> *__restrict__ r,
                                                               real4<real> *__restrict__ z, const real4<real> *__restrict__ p,
                                                               const real4<real> *__restrict__ Ap, const real *__restrict__ alpha,
                                                               const real *__restrict__ minv, double *__restrict__ part,
                                                               const CGState *__restrict__ st) {
    const int b = blockIdx.y;
    if (k >= st[b].conv_it) return;
    __shared__ real sr[3 * MOLP_BLOCK];
    __shared__ double lds[32];
    const double pap = block_reduce_partials(pap_part + (size_t)b * np, np, 1, 0, lds, false);
    const real al = (real)(rz[(size_t)b * rzs + k] / fmax(pap, 1e-300));
    const int li = blockIdx.x * MOLP_BLOCK + threadIdx.x;
    const int i = b * nper + li;
    const int pb = b * gridDim.x + blockIdx.x;
    const bool in = li < nper;
    real4<real> r0, r1;
    r0.x = r0.y = r0.z = r0.w = real(0);
    r1 = r0;
    if (in) {
        const real4<real> pp = p[i], ap = Ap[i];
        r0 = r[i];
        real4<real> xx = x[i];
        xx.x += al * pp.x;
        xx.y += al * pp.y;
        xx.z += al * pp.z;
        x[i] = xx;
        r1.x = r0.x - al * ap.x;
        r1.y = r0.y - al * ap.y;
        r1.z = r0.z - al * ap.z;
        r1.w = 0;
        r[i] = r1;
    }
    sr[3 * threadIdx.x] = r1.x;
    sr[3 * threadIdx.x + 1] = r1.y;
    sr[3 * threadIdx.x + 2] = r1.z;
    __syncthreads();
    double s[2] = {0.0, 0.0};
    double m = 0.0;
    if (in) {
        const int q = threadIdx.x / 3, pa = threadIdx.x - 3 * q;
        const real4<real> z1 = mol_apply<real>(minv + (size_t)81 * (i / 3), pa, sr + 9 * q);
        z[i] = z1;
        const real a = alpha[i];
        s[0] = (double)(z1.x * (r1.x - r0.x)) + (double)(z1.y * (r1.y - r0.y)) + (double)(z1.z * (r1.z - r0.z));
        s[1] = (double)(r1.x * z1.x) + (double)(r1.y * z1.y) + (double)(r1.z * z1.z);
        m = fmax(fmax(fabs((double)(a * r1.x)), fabs((double)(a * r1.y))), fabs((double)(a * r1.z)));
    }
    block_sum<2>(s, lds);
    m = block_max(m, lds + 16);
    if (threadIdx.x == 0) {
        part[3 * pb] = s[0];
        part[3 * pb + 1] = s[1];
        part[3 * pb + 2] = m;
    }
}

// Per bead b: iterations it_b = min(conv_it_b, launched) and the final residual err_b[it_b]
// (bead_it, bead_res); stats slot = the largest of the residuals (kept on the device).  One block.
__global__ void k_record_residual(int nbead, const double *__restrict__ err, int rzs, const CGState *__restrict__ st,
                                  int launched, int *__restrict__ bead_it, double *__restrict__ bead_res,
                                  double *__restrict__ slot) {
    __shared__ double lds[32];
    double m = 0.0;
    for (int b = threadIdx.x; b < nbead; b += blockDim.x) {
        const int it = min(st[b].conv_it, launched);
        const double e = err[(size_t)b * rzs + it];
        bead_it[b] = it;
        bead_res[b] = e;
        m = fmax(m, e);
    }
    m = block_max(m, lds);
    if (threadIdx.x == 0) slot[0] = m;
}

// mu = x + peek alpha r (FP64 out).
template <typename real>
__global__ void k_cg_finish(int n, const real4<real> *__restrict__ x, const real4<real> *__restrict__ r,
                            const real *__restrict__ alpha, double peek, double *__restrict__ mu) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const real4<real> xx = x[i], rr = r[i];
    const real a = alpha[i];
    const real pk = (real)peek;
    if (peek != 0.0) {
        mu[3 * i] = (double)(xx.x + pk * rr.x * a);
        mu[3 * i + 1] = (double)(xx.y + pk * rr.y * a);
        mu[3 * i + 2] = (double)(xx.z + pk * rr.z * a);
    } else {
        mu[3 * i] = (double)xx.x;
        mu[3 * i + 1] = (double)xx.y;
        mu[3 * i + 2] = (double)xx.z;
    }
}

// The end of a solve in one launch (k_cg_finish, then the pushes of mu into the learned guess's
// ring, slot `ring_slot`, and into the predictor history, each skipped when its pointer is null,
// then k_record_residual by block 0).
template <typename real>
__global__ __launch_bounds__(256) void k_cg_finish_all(int n, const real4<real> *__restrict__ x, const real4<real> *__restrict__ r,
                                                       const real *__restrict__ alpha, double peek, double *__restrict__ mu,
                                                       double *__restrict__ ring, int ring_slot, double *__restrict__ hist,
                                                       int nbead, const double *__restrict__ err, int rzs,
                                                       const CGState *__restrict__ st, int launched, int *__restrict__ bead_it,
                                                       double *__restrict__ bead_res, double *__restrict__ slot) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const real4<real> xx = x[i], rr = r[i];
        const real a = alpha[i];
        const real pk = (real)peek;
        double m[3];
        if (peek != 0.0) {
            m[0] = (double)(xx.x + pk * rr.x * a);
            m[1] = (double)(xx.y + pk * rr.y * a);
            m[2] = (double)(xx.z + pk * rr.z * a);
        } else {
            m[0] = (double)xx.x;
            m[1] = (double)xx.y;
            m[2] = (double)xx.z;
        }
        const size_t s3 = 3 * (size_t)n;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const size_t o = 3 * (size_t)i + c;
            mu[o] = m[c];
            if (ring) ring[(size_t)ring_slot * s3 + o] = m[c];
            if (hist) {
                hist[3 * s3 + o] = hist[2 * s3 + o];
                hist[2 * s3 + o] = hist[s3 + o];
                hist[s3 + o] = hist[o];
                hist[o] = m[c];
            }
        }
    }
    if (blockIdx.x == 0) {  // k_record_residual
        __shared__ double lds[32];
        double mx = 0.0;
        for (int b = threadIdx.x; b < nbead; b += blockDim.x) {
            const int it = min(st[b].conv_it, launched);
            const double e = err[(size_t)b * rzs + it];
            bead_it[b] = it;
            bead_res[b] = e;
            mx = fmax(mx, e);
        }
        mx = block_max(mx, lds);
        if (threadIdx.x == 0) slot[0] = mx;
    }
}

// hist = [mu, hist[0], hist[1], hist[2]] (in place, newest first).
__global__ void k_hist_push(int n3, const double *__restrict__ mu, double *__restrict__ hist) {
    const int o = blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= n3) return;
    const size_t s = (size_t)n3;
    hist[3 * s + o] = hist[2 * s + o];
    hist[2 * s + o] = hist[s + o];
    hist[s + o] = hist[o];
    hist[o] = mu[o];
}

// ---------------------------------------------------------------------------- covalent frames
// p_i = sum over i's covalent dipoles m of c_m unit(min_image(r_j - r_i))  (forcefield.py:1032-1042)
__global__ void k_perm_dipoles(int n, const double *__restrict__ pos, Box box, const int *__restrict__ cstart,
                               const int *__restrict__ cj, const double *__restrict__ cc, double *__restrict__ p) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double px = 0, py = 0, pz = 0;
    for (int m = cstart[i]; m < cstart[i + 1]; ++m) {
        const int j = cj[m];
        double vx = pos[3 * j] - pos[3 * i], vy = pos[3 * j + 1] - pos[3 * i + 1], vz = pos[3 * j + 2] - pos[3 * i + 2];
        min_image(vx, vy, vz, box);
        const double il = 1.0 / sqrt(vx * vx + vy * vy + vz * vz);
        px += cc[m] * vx * il;
        py += cc[m] * vy * il;
        pz += cc[m] * vz * il;
    }
    p[3 * i] = px;
    p[3 * i + 1] = py;
    p[3 * i + 2] = pz;
}

// d = p + mu (FP64) and dq = (d, q) in real; partials (sum |d|^2, sum |mu|^2 / (2 alpha)).
template <typename real>
__global__ __launch_bounds__(256) void k_energy_pack(int nper, const double *__restrict__ p, const double *__restrict__ mu,
                                                     const double *__restrict__ q, const double *__restrict__ alpha,
                                                     double *__restrict__ dd, real4<real> *__restrict__ dq,
                                                     double *__restrict__ part) {
    __shared__ double lds[32];
    const int li = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y * nper + li;
    const int pb = blockIdx.y * gridDim.x + blockIdx.x;
    double s[2] = {0.0, 0.0};
    if (li < nper) {
        double d[3];
        double m2 = 0.0;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const double m = mu ? mu[3 * i + c] : 0.0;
            d[c] = p[3 * i + c] + m;
            dd[3 * i + c] = d[c];
            m2 += m * m;
        }
        real4<real> v;
        v.x = (real)d[0];
        v.y = (real)d[1];
        v.z = (real)d[2];
        v.w = (real)q[i];
        dq[i] = v;
        s[0] = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
        s[1] = alpha[i] != 0.0 ? m2 / (2.0 * alpha[i]) : 0.0;  // 0 for atoms without polarizability
    }
    block_sum<2>(s, lds);
    if (threadIdx.x == 0) {
        part[2 * pb] = s[0];
        part[2 * pb + 1] = s[1];
    }
}

// forces = -(KE g_el + g_lj + KE rec_dr + VJP_p(gp)); the pull-back through the covalent frames of
// p_i = sum c unit(v), v = r_j - r_i: t_m = c_m (gp_i - (gp_i . u) u) / |v|, +t_m on j, -t_m on i,
// gathered per atom from the covalent dipoles it owns (cstart) and those pointing at it (jstart).
__global__ void k_forces_final(int n, const double *__restrict__ pos, Box box, const double *__restrict__ g_el,
                               const double *__restrict__ g_lj, const double *__restrict__ rec_dr,
                               const double *__restrict__ gp, const int *__restrict__ cstart, const int *__restrict__ ci,
                               const int *__restrict__ cj, const double *__restrict__ cc, const int *__restrict__ jstart,
                               const int *__restrict__ jlist, double *__restrict__ forces, const double *__restrict__ q = nullptr,
                               Ext3 kf = {}) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= n) return;
    double v[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) v[c] = KE * g_el[3 * a + c] + g_lj[3 * a + c] + KE * rec_dr[3 * a + c];
    if (q) {  // a uniform external field pulls every charge: force KE q F0 (kf = KE F0)
        v[0] -= q[a] * kf.x;
        v[1] -= q[a] * kf.y;
        v[2] -= q[a] * kf.z;
    }
    for (int pass = 0; pass < 2; ++pass) {
        const int m0 = pass == 0 ? cstart[a] : jstart[a];
        const int m1 = pass == 0 ? cstart[a + 1] : jstart[a + 1];
        for (int s = m0; s < m1; ++s) {
            const int m = pass == 0 ? s : jlist[s];
            const int i = ci[m], j = cj[m];
            double ux = pos[3 * j] - pos[3 * i], uy = pos[3 * j + 1] - pos[3 * i + 1], uz = pos[3 * j + 2] - pos[3 * i + 2];
            min_image(ux, uy, uz, box);
            const double L = sqrt(ux * ux + uy * uy + uz * uz);
            ux /= L;
            uy /= L;
            uz /= L;
            const double gx = gp[3 * i], gy = gp[3 * i + 1], gz = gp[3 * i + 2];
            const double gu = gx * ux + gy * uy + gz * uz;
            const double sgn = pass == 0 ? -1.0 : 1.0;
            const double f = sgn * cc[m] / L;
            v[0] += f * (gx - gu * ux);
            v[1] += f * (gy - gu * uy);
            v[2] += f * (gz - gu * uz);
        }
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) forces[3 * a + c] = -v[c];
}

// Energy assembly (one block per bead, out[4 b .. 4 b + 3] = elec, vdw, total, u_rec): parts =
// {pair rows (e_el, e_lj) x npair, (sum|d|^2, u_pol) x npack, u_rec x nconv} per bead; consts =
// {KE * (u_self_q + u_bg), e_tail, c_self}.
//
// With a uniform external field (mpart not null): the field's dipole M = sum of the nm partials of
// k_field_dipole per bead + offset, its energy -KE F0 . M (kf = KE F0) added to the total, and
// (e_field, M) in fout[4 b .. 4 b + 3] (forcefield.py:2075-2097).
__global__ void k_energy_final(const double *__restrict__ pair_part, int npair, const double *__restrict__ pack_part, int npack,
                               const double *__restrict__ conv_part, int nconv, double ke_const, double tail, double cself,
                               double *__restrict__ out, const double *__restrict__ mpart = nullptr, int nm = 0,
                               Ext3 kf = {}, Ext3 off = {}, double *__restrict__ fout = nullptr) {
    __shared__ double lds[8 * 16];
    const int b = blockIdx.x;
    const double2 *pp = reinterpret_cast<const double2 *>(pair_part) + (size_t)b * npair;
    const double2 *pk = reinterpret_cast<const double2 *>(pack_part) + (size_t)b * npack;
    conv_part += (size_t)b * nconv;
    out += 4 * b;
    // one pass over the three partial arrays, five sums per thread, then one block sum
    double v[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    const int nt = blockDim.x;
    for (int k0 = threadIdx.x; k0 < npair; k0 += 4 * nt) {
        double2 x[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) x[u] = k0 + u * nt < npair ? pp[k0 + u * nt] : make_double2(0.0, 0.0);
#pragma unroll
        for (int u = 0; u < 4; ++u) v[0] += x[u].x, v[1] += x[u].y;
    }
    for (int k = threadIdx.x; k < npack; k += nt) {
        const double2 x = pk[k];
        v[2] += x.x, v[3] += x.y;
    }
    for (int k = threadIdx.x; k < nconv; k += nt) v[4] += conv_part[k];
    double m[3] = {0.0, 0.0, 0.0};
    if (mpart)
        for (int k = threadIdx.x; k < nm; k += nt)
#pragma unroll
            for (int c = 0; c < 3; ++c) m[c] += mpart[3 * ((size_t)b * nm + k) + c];
    double w[8] = {v[0], v[1], v[2], v[3], v[4], m[0], m[1], m[2]};
    block_sum<8>(w, lds);
    const double e_el = w[0], e_lj = w[1], d2 = w[2], upol = w[3], urec = w[4];
    if (threadIdx.x == 0) {
        const double elec = KE * 0.5 * e_el + KE * (urec - 0.5 * cself * d2 + upol) + ke_const;
        const double vdw = 0.5 * e_lj + tail;
        double efield = 0.0;
        if (mpart) {
            const double M[3] = {w[5] + off.x, w[6] + off.y, w[7] + off.z};
            efield = -(kf.x * M[0] + kf.y * M[1] + kf.z * M[2]);
            fout[4 * b] = efield;
            fout[4 * b + 1] = M[0];
            fout[4 * b + 2] = M[1];
            fout[4 * b + 3] = M[2];
        }
        out[0] = elec;
        out[1] = vdw;
        out[2] = elec + vdw + efield;
        out[3] = urec;
    }
}

// Per-block partials (grid: blocks per bead, beads) of the dipole a uniform external field acts on,
// sum q r + sum d with d = p + mu (FP64; molecules whole, forcefield.py:2090-2097).
__global__ __launch_bounds__(256) void k_field_dipole(int nper, const double *__restrict__ pos, const double *__restrict__ q,
                                                      const double *__restrict__ dd, double *__restrict__ part) {
    __shared__ double lds[3 * 16];
    const int li = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y * nper + li;
    const int pb = blockIdx.y * gridDim.x + blockIdx.x;
    double m[3] = {0.0, 0.0, 0.0};
    if (li < nper)
#pragma unroll
        for (int c = 0; c < 3; ++c) m[c] = q[i] * pos[3 * i + c] + dd[3 * i + c];
    block_sum<3>(m, lds);
    if (threadIdx.x == 0)
#pragma unroll
        for (int c = 0; c < 3; ++c) part[3 * pb + c] = m[c];
}

}  // namespace pgm
