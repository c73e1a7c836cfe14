// pme.h -- smooth PME of Gaussian charges and point dipoles: spread, convolution, gathers.
//
// Per atom: u = r H^-1 (FP64), w = K frac(u), base = floor(w), f = w - base (rounded to real);
// theta / dtheta (and for forces d2theta) of order P are recomputed from f in registers by every
// kernel (no per-atom spline storage).  Grid point (base + k) mod K gets weight theta[k].
//   Q(g) = sum_i q_i t1 t2 t3 + e_i1 d1 t2 t3 + e_i2 t1 d2 t3 + e_i3 t1 t2 d3,  e_i = d_i H^-1 K
// (pgm_jax/md/pme.py:306-337).  phi = dE_rec/dQ comes from r2c FFT -> multiply by
// Gc = 2 G / w3 -> c2r FFT (unnormalised), pme.py:373-376.
#pragma once

#include "common.h"

namespace pgm {

struct PMEGeom {
    double hinv[9];      // H^-1
    double escale[9];    // H^-1 with column b scaled by K_b: e = d . escale
    int K[3];
};

template <typename real, int P>
PGM_DEV void stencil(const double *__restrict__ pos, int i, const PMEGeom &g, int *base, real *f) {
    const double x = pos[3 * i], y = pos[3 * i + 1], z = pos[3 * i + 2];
    const double *hi = g.hinv;
    const double u[3] = {x * hi[0] + y * hi[3] + z * hi[6], x * hi[1] + y * hi[4] + z * hi[7], x * hi[2] + y * hi[5] + z * hi[8]};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const double w = (u[a] - floor(u[a])) * (double)g.K[a];
        const double b = floor(w);
        base[a] = (int)b;
        f[a] = (real)(w - b);
    }
}

template <typename real>
PGM_DEV void scaled_dipole(const real4<real> &d, const PMEGeom &g, real *e) {
#pragma unroll
    for (int b = 0; b < 3; ++b) e[b] = d.x * (real)g.escale[b] + d.y * (real)g.escale[3 + b] + d.z * (real)g.escale[6 + b];
}

template <typename T> __global__ void k_fill(size_t n, T *__restrict__ p, T v) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Spread: one thread per (atom, x offset); 36 atomic adds each (v1 of the spread; deterministic
// order is not guaranteed).  dq = (d, q); WITHQ includes the charges.
template <typename real, int P, bool WITHQ>
__global__ __launch_bounds__(256) void k_spread(int n, const double *__restrict__ pos, const real4<real> *__restrict__ dq,
                                                PMEGeom g, real *__restrict__ grid, const int *__restrict__ conv_it, int iter) {
    if (conv_it && iter >= *conv_it) return;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = t / P, ix = t - i * P;
    if (i >= n) return;
    int base[3];
    real f[3];
    stencil<real, P>(pos, i, g, base, f);
    real t1[P], d1[P], t2[P], d2[P], t3[P], d3[P];
    bspline<real, P, false>(f[0], t1, d1, nullptr);
    bspline<real, P, false>(f[1], t2, d2, nullptr);
    bspline<real, P, false>(f[2], t3, d3, nullptr);
    const real4<real> d = dq[i];
    real e[3];
    scaled_dipole<real>(d, g, e);
    const real q = WITHQ ? d.w : real(0);
    const real a1 = q * t1[ix] + e[0] * d1[ix];
    const real b1 = e[1] * t1[ix];
    const real c1 = e[2] * t1[ix];
    const int K1 = g.K[0], K2 = g.K[1], K3 = g.K[2];
    int gx = base[0] + ix;
    gx -= gx >= K1 ? K1 : 0;
#pragma unroll
    for (int iy = 0; iy < P; ++iy) {
        const real A = a1 * t2[iy] + b1 * d2[iy];
        const real C = c1 * t2[iy];
        int gy = base[1] + iy;
        gy -= gy >= K2 ? K2 : 0;
        real *line = grid + ((size_t)gx * K2 + gy) * K3;
#pragma unroll
        for (int iz = 0; iz < P; ++iz) {
            int gz = base[2] + iz;
            gz -= gz >= K3 ? K3 : 0;
            unsafeAtomicAdd(line + gz, A * t3[iz] + C * d3[iz]);
        }
    }
}

// ---------------------------------------------------------------------------- tiled spread
// Atomic-free, deterministic spread.  The grid is cut into tiles (T_a points per axis, every tile at
// least P-1 wide, at least two tiles per axis); every atom belongs to the tile holding its stencil
// base.  Atoms are sorted by tile once per geometry (k_pme_bin + radix sort + k_pme_splines, which
// also stores the spline weights in sorted order).  One wave per tile accumulates its atoms' full
// P^3 stencils into an LDS block of (T+P-1)^3 points: one atom at a time, the 64 lanes take 64
// distinct stencil points, so plain LDS read-add-write is race free and the order is fixed.  The
// blocks are written out and k_merge_table sums, for every grid point, the (up to 8) blocks that
// cover it in a fixed order.
struct PMETiles {
    int T[3];    // tile width per axis (the last tile may be narrower, >= P-1)
    int nt[3];   // tiles per axis
    int B[3];    // LDS block extent per axis: T + P - 1
    int ntiles;  // nt0 nt1 nt2
    int bsize;   // B0 B1 B2
};

// The spread works on tiles of T8 points per axis, the gather on tiles of T16 = 2 T8 (each the
// union of 2 x 2 x 2 spread tiles).  Atoms are sorted once per geometry by the fine key
// key = t16 index * 8 + octant, so both the spread tiles (one key each) and the gather tiles (eight
// consecutive keys) are contiguous ranges of the sorted order.
PGM_DEV int fine_key(const int (&u8)[3], const PMETiles &t16) {
    const int oct = ((u8[0] & 1) << 2) | ((u8[1] & 1) << 1) | (u8[2] & 1);
    return (((u8[0] >> 1) * t16.nt[1] + (u8[1] >> 1)) * t16.nt[2] + (u8[2] >> 1)) * 8 + oct;
}

// key = bead * t8.ntiles + fine key of each atom's stencil base (nper atoms per bead); idx = atom index.
__global__ void k_pme_bin(int n, int nper, const double *__restrict__ pos, PMEGeom g, PMETiles t8, PMETiles t16,
                          int *__restrict__ key, int *__restrict__ idx) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int base[3];
    double f[3];
    stencil<double, 4>(pos, i, g, base, f);
    int u8[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const int bb = base[a] >= g.K[a] ? base[a] - g.K[a] : base[a];
        const int u = bb / t8.T[a];
        u8[a] = u < t8.nt[a] ? u : t8.nt[a] - 1;
    }
    key[i] = (i / nper) * t8.ntiles + fine_key(u8, t16);
    idx[i] = i;
}

// Spline weights in sorted order: spl[s] = (t0[P], d0[P], t1[P], d1[P], t2[P], d2[P]) of atom
// sidx[s]; loc8[s] / loc16[s] = the stencil base relative to its spread / gather tile start,
// packed (lx | ly << 10 | lz << 20).
template <typename real, int P>
__global__ void k_pme_splines(int n, const double *__restrict__ pos, PMEGeom g, PMETiles t8, const int *__restrict__ sidx,
                              real *__restrict__ spl, int *__restrict__ loc8, int *__restrict__ loc16) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const int i = sidx[s];
    int base[3];
    real f[3];
    stencil<real, P>(pos, i, g, base, f);
    int l8[3], l16[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const int bb = base[a] >= g.K[a] ? base[a] - g.K[a] : base[a];
        int u = bb / t8.T[a];
        u = u < t8.nt[a] ? u : t8.nt[a] - 1;
        l8[a] = bb - u * t8.T[a];
        l16[a] = l8[a] + (u & 1) * t8.T[a];
        real th[P], dth[P];
        bspline<real, P, false>(f[a], th, dth, nullptr);
#pragma unroll
        for (int k = 0; k < P; ++k) {
            spl[(size_t)s * 6 * P + (2 * a) * P + k] = th[k];
            spl[(size_t)s * 6 * P + (2 * a + 1) * P + k] = dth[k];
        }
    }
    loc8[s] = l8[0] | (l8[1] << 10) | (l8[2] << 20);
    loc16[s] = l16[0] | (l16[1] << 10) | (l16[2] << 20);
}

// The atom ranges of every spread tile and every gather tile from the sorted keys, in one launch and
// without clearing first.  Thread s in [0, n] looks at the keys either side of position s (key -1
// before the first, the tile count after the last): position s is the start of every tile k with
// key[s-1] < k <= key[s] and the end of tile k - 1, so every tile's start and end is written once
// and an empty tile gets an empty range.  Spread tiles are the keys, gather tiles key >> shift.
__global__ void k_tile_bounds(int n, const int *__restrict__ skey, int n8, int *__restrict__ start8, int *__restrict__ end8,
                              int n16, int shift, int *__restrict__ start16, int *__restrict__ end16) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s > n) return;
    const int p8 = s > 0 ? skey[s - 1] : -1, c8 = s < n ? skey[s] : n8;
    for (int k = p8 + 1; k <= c8; ++k) {
        if (k < n8) start8[k] = s;
        if (k > 0 && k <= n8) end8[k - 1] = s;
    }
    const int p16 = s > 0 ? skey[s - 1] >> shift : -1, c16 = s < n ? skey[s] >> shift : n16;
    for (int k = p16 + 1; k <= c16; ++k) {
        if (k < n16) start16[k] = s;
        if (k > 0 && k <= n16) end16[k - 1] = s;
    }
}

// Tiled spread: one wave per spread tile (fine key; block b * t.ntiles + key for bead b) accumulates
// the stencils of its atoms into an LDS block of (T8 + P - 1)^3 points, written to `blocks` (merged
// by k_merge_table).  The atoms'
// spline weights, scaled dipoles, charges and local offsets are staged through LDS in chunks of
// SPREAD_CHUNK (one parallel load instead of a chain of dependent loads per atom); the 64 lanes
// cover the P^3 stencil points of one atom at a time, so no two lanes touch one point and the sum
// is free of atomics and bitwise reproducible (an LDS float add, ds_add_f32, in place of the read and
// write measured three times slower).
constexpr int SPREAD_CHUNK = 8;  // atoms staged at a time: 10 KB of LDS per wave, 16 waves (all 8^3 tiles of 128^3) per CU

template <typename real, int P>
PGM_DEV constexpr int spread_lds_atoms() {
    return SPREAD_CHUNK * (6 * P + 4) * (int)sizeof(real) + SPREAD_CHUNK * (int)sizeof(int);
}

template <typename real, int P, bool WITHQ>
__global__ __launch_bounds__(64) void k_spread_tiles(PMETiles t, PMEGeom g, const int *__restrict__ tstart,
                                                     const int *__restrict__ tend, const int *__restrict__ sidx,
                                                     const real *__restrict__ spl, const int *__restrict__ loc,
                                                     const real4<real> *__restrict__ dq, real *__restrict__ blocks,
                                                     const int *__restrict__ conv_it, int iter) {
    if (cg_done(conv_it, blockIdx.x / t.ntiles, iter)) return;
    extern __shared__ __align__(16) unsigned char lds_raw[];
    constexpr int SW = 6 * P + 4;  // staged values per atom: splines, e0 e1 e2 q
    real *sv = reinterpret_cast<real *>(lds_raw);
    int *sl = reinterpret_cast<int *>(sv + SPREAD_CHUNK * SW);
    real *blk = reinterpret_cast<real *>(lds_raw + spread_lds_atoms<real, P>());
    const int lane = threadIdx.x;
    const int key = blockIdx.x;
    const int BY = t.B[1], BZ = t.B[2];
    for (int k = lane; k < t.bsize; k += WAVE) blk[k] = real(0);
    const int s0 = tstart[key], s1 = tend[key];
    // this lane's stencil points p = lane + 64 u (the last pass only where p < P^3): axis indices
    // and block offsets, the same for every atom
    constexpr int NP = P * P * P, NPASS = (NP + WAVE - 1) / WAVE;
    static_assert((NPASS - 1) * WAVE < NP, "every pass but the last is full");
    int pix[NPASS], piy[NPASS], piz[NPASS], poff[NPASS];
    const bool plast = lane + (NPASS - 1) * WAVE < NP;
#pragma unroll
    for (int u = 0; u < NPASS; ++u) {
        const int p = (u < NPASS - 1 || plast) ? lane + u * WAVE : 0;
        pix[u] = p / (P * P);
        piy[u] = (p / P) % P;
        piz[u] = p % P;
        poff[u] = (pix[u] * BY + piy[u]) * BZ + piz[u];
    }
    for (int c0 = s0; c0 < s1; c0 += SPREAD_CHUNK) {
        const int m = min(SPREAD_CHUNK, s1 - c0);
        __syncthreads();
        for (int k = lane; k < m * 6 * P; k += WAVE) {  // (theta, dtheta) pairs per axis point
            const int a = k / (6 * P), v = k - a * 6 * P;
            const int ax = v / (2 * P), w = v - ax * 2 * P, which = w / P, i = w - which * P;
            sv[a * SW + ax * 2 * P + 2 * i + which] = spl[(size_t)(c0 + a) * 6 * P + v];
        }
        if (lane < m) {
            const real4<real> d = dq[sidx[c0 + lane]];
            real e[3];
            scaled_dipole<real>(d, g, e);
            sv[lane * SW + 6 * P] = e[0];
            sv[lane * SW + 6 * P + 1] = e[1];
            sv[lane * SW + 6 * P + 2] = e[2];
            sv[lane * SW + 6 * P + 3] = WITHQ ? d.w : real(0);
            sl[lane] = loc[c0 + lane];
        }
        __syncthreads();
        for (int a = 0; a < m; ++a) {
            const real *v = sv + a * SW;
            const int lc = sl[a];
            const int lx = lc & 1023, ly = (lc >> 10) & 1023, lz = lc >> 20;
            const real e0 = v[6 * P], e1 = v[6 * P + 1], e2 = v[6 * P + 2], q = v[6 * P + 3];
            const int base = (lx * BY + ly) * BZ + lz;
            real w[NPASS], old[NPASS];
#pragma unroll
            for (int u = 0; u < NPASS; ++u) {
                const real tx = v[2 * pix[u]], dx = v[2 * pix[u] + 1];
                const real ty = v[2 * P + 2 * piy[u]], dy = v[2 * P + 2 * piy[u] + 1];
                const real tz = v[4 * P + 2 * piz[u]], dz = v[4 * P + 2 * piz[u] + 1];
                w[u] = (q * tx + e0 * dx) * ty * tz + tx * (e1 * dy * tz + e2 * ty * dz);
            }
#pragma unroll
            for (int u = 0; u < NPASS; ++u)
                if (u < NPASS - 1 || plast) old[u] = blk[base + poff[u]];
#pragma unroll
            for (int u = 0; u < NPASS; ++u)
                if (u < NPASS - 1 || plast) blk[base + poff[u]] = old[u] + w[u];
        }
    }
    __syncthreads();
    real *out = blocks + (size_t)key * t.bsize;
    for (int k = lane; k < t.bsize; k += WAVE) out[k] = blk[k];
}

// Block merge: grid point value = sum of the spread blocks covering it -- along each axis its own
// spread tile (local l) and, when l <= P-2, the previous one (local l + width of that tile), in a
// fixed order (deterministic).  Table driven: for grid coordinate g along axis a, tab_a[2 g] is the offset (in reals) of the
// point in the block of its own spread tile's axis part -- A_a(u) bsize + l stride_a with
// A_0(u) = ((u >> 1) nt16_1 nt16_2 8 + (u & 1) 4), A_1(u) = ((u >> 1) nt16_2 8 + (u & 1) 2),
// A_2(u) = ((u >> 1) 8 + (u & 1)), strides B1 B2, B2, 1 -- and tab_a[2 g + 1] the same for the
// previous tile (local l + its width) when l <= P-2, else -1.  Both the fine key and the in-block
// offset are sums over the axes, so a grid point's up to 8 block values are
// blocks[bead base + X + Y + Z] over the axes' candidates, summed in a fixed order.
// Launch: x = z-line (blockDim.x points), y = K2 rows, z = K1 planes times beads.
template <typename real>
__global__ __launch_bounds__(128) void k_merge_table(int K1, int K2, int K3, size_t bead_blocks, const int *__restrict__ tab,
                                                     const real *__restrict__ blocks, real *__restrict__ grid,
                                                     const int *__restrict__ conv_it, int iter) {
    const int gz = blockIdx.x * blockDim.x + threadIdx.x;
    const int gy = blockIdx.y;
    const int gx = blockIdx.z % K1, bead = blockIdx.z / K1;
    if (gz >= K3 || cg_done(conv_it, bead, iter)) return;
    const int *tx = tab, *ty = tab + 2 * K1, *tz = tab + 2 * (K1 + K2);
    const int X[2] = {tx[2 * gx], tx[2 * gx + 1]}, Y[2] = {ty[2 * gy], ty[2 * gy + 1]}, Z[2] = {tz[2 * gz], tz[2 * gz + 1]};
    const real *b = blocks + bead_blocks * bead;
    real acc = real(0);
#pragma unroll
    for (int i0 = 0; i0 < 2; ++i0)
#pragma unroll
        for (int i1 = 0; i1 < 2; ++i1)
#pragma unroll
            for (int i2 = 0; i2 < 2; ++i2)
                if (X[i0] >= 0 && Y[i1] >= 0 && Z[i2] >= 0) acc += b[(size_t)X[i0] + Y[i1] + Z[i2]];
    grid[(((size_t)bead * K1 + gx) * K2 + gy) * K3 + gz] = acc;
}

// Convolution on the half spectrum: F[Q] *= Gc; with ENERGY also sum (w3/2) Gc |F[Q]|^2 per block
// (= sum G |F[Q]|^2 with G of pme.py:243-267), FP64.  Grid (blocks per bead, beads): bead b's
// transform at cgrid + 2 b nc, its partials at epart[b * gridDim.x + blockIdx.x].
template <typename real, bool ENERGY>
__global__ __launch_bounds__(256) void k_convolve(size_t nc, int K3h, int K3even, real *__restrict__ cgrid,
                                                  const real *__restrict__ Gc, double *__restrict__ epart,
                                                  const int *__restrict__ conv_it, int iter) {
    if (cg_done(conv_it, blockIdx.y, iter)) return;
    __shared__ double lds[16];
    const size_t s = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    cgrid += 2 * nc * blockIdx.y;
    double e[1] = {0.0};
    if (s < nc) {
        const real gc = Gc[s];
        const real re = cgrid[2 * s], im = cgrid[2 * s + 1];
        if (ENERGY) {
            const int k3 = (int)(s % (size_t)K3h);
            const double w = (k3 == 0 || (K3even && k3 == K3h - 1)) ? 0.5 : 1.0;
            e[0] = w * (double)gc * ((double)re * re + (double)im * im);
        }
        cgrid[2 * s] = re * gc;
        cgrid[2 * s + 1] = im * gc;
    }
    if (ENERGY) {
        block_sum<1>(e, lds);
        if (threadIdx.x == 0) epart[blockIdx.y * gridDim.x + blockIdx.x] = e[0];
    }
}

// Reciprocal-space part of the strain derivative that comes from the influence function: for
// G(m*, V) = exp(-pi^2 |m*|^2 / beta^2) / (2 pi V |m*|^2) B(m), with m* = H^-1 m moving as
// -eps^T m* and V as (1 + tr eps), dG/d eps_ab = G [2 (pi^2/beta^2 + 1/|m*|^2) m*_a m*_b - delta_ab].
// Per block: the six sums of w Gc |F Q|^2 2 (pi^2/beta^2 + 1/|m*|^2) m*_a m*_b (xx yy zz xy xz yz)
// and E_rec = sum w Gc |F Q|^2, from the unconvolved transform (w = 1/2 on the k3 = 0 and Nyquist
// planes of the half spectrum, 1 elsewhere).  cgrid is only read.
struct RecipGeom {
    double hinv[9];   // H^-1: m* = H^-1 (m1, m2, m3)
    int K[3];
    double pb2;       // pi^2 / beta^2
};

template <typename real>
__global__ __launch_bounds__(256) void k_convolve_virial(size_t nc, RecipGeom rg, const real *__restrict__ cgrid,
                                                         const real *__restrict__ Gc, double *__restrict__ vpart) {
    __shared__ double lds[7 * 16];
    const size_t s = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    double acc[7] = {0, 0, 0, 0, 0, 0, 0};
    if (s < nc) {
        const int K1 = rg.K[0], K2 = rg.K[1], K3 = rg.K[2], K3h = K3 / 2 + 1;
        const int k3 = (int)(s % (size_t)K3h);
        const int k2 = (int)((s / (size_t)K3h) % (size_t)K2);
        const int k1 = (int)(s / ((size_t)K3h * K2));
        const double f1 = k1 <= (K1 - 1) / 2 ? k1 : k1 - K1;
        const double f2 = k2 <= (K2 - 1) / 2 ? k2 : k2 - K2;
        const double f3 = k3;
        const double *h = rg.hinv;
        const double mx = h[0] * f1 + h[1] * f2 + h[2] * f3;
        const double my = h[3] * f1 + h[4] * f2 + h[5] * f3;
        const double mz = h[6] * f1 + h[7] * f2 + h[8] * f3;
        const double msq = mx * mx + my * my + mz * mz;
        const double re = (double)cgrid[2 * s], im = (double)cgrid[2 * s + 1];
        const double w = (k3 == 0 || (K3 % 2 == 0 && k3 == K3h - 1)) ? 0.5 : 1.0;
        const double e = w * (double)Gc[s] * (re * re + im * im);
        if (msq > 0.0) {
            const double c = 2.0 * e * (rg.pb2 + 1.0 / msq);
            acc[0] = c * mx * mx;
            acc[1] = c * my * my;
            acc[2] = c * mz * mz;
            acc[3] = c * mx * my;
            acc[4] = c * mx * mz;
            acc[5] = c * my * mz;
            acc[6] = e;
        }
    }
    block_sum<7>(acc, lds);
    if (threadIdx.x == 0)
        for (int k = 0; k < 7; ++k) vpart[7 * blockIdx.x + k] = acc[k];
}

// Gather of the dipole gradient (pme.py:378-388): g_a = sum phi * (derivative spline along a),
// dU/dd = escale . g; then combined with the row field according to `mode`:
//   0 FIELD   out = -(row + rec - c_self d)                          [+ sum |alpha out| per block]
//   1 AV      out = d / alpha + row + rec - c_self d                 [+ sum d . out per block]
//   2 FUSED   out = -(row + rec - c_self d) - x0 / alpha             (d = p + x0, extra = x0)
//   3 PLAIN   out = b - (d / alpha + row + rec - c_self d)           (d = x0, extra = b)
//   4 REC     out = dU_rec/dd alone
enum { MODE_FIELD = 0, MODE_AV = 1, MODE_FUSED = 2, MODE_PLAIN = 3, MODE_REC = 4 };

// out[i] from rec = dU_rec/dd (rx, ry, rz) and the row field, by `mode` (see above); acc gets the
// per-atom contribution to the block partial of MODE_FIELD / MODE_AV.
template <typename real>
PGM_DEV void combine_out(int mode, int i, real rx, real ry, real rz, const real4<real> *__restrict__ dq,
                         const real4<real> *__restrict__ row, const real *__restrict__ alpha,
                         const real4<real> *__restrict__ extra, real cself, real4<real> *__restrict__ out, double &acc) {
    const real4<real> d = dq[i];
    const real4<real> rw = row[i];
    const real tx = rw.x + rx - cself * d.x;  // dU/dd (direct + reciprocal + self)
    const real ty = rw.y + ry - cself * d.y;
    const real tz = rw.z + rz - cself * d.z;
    real4<real> o;
    o.w = 0;
    const real ia = alpha[i] != real(0) ? real(1) / alpha[i] : real(0);  // 0: not polarizable
    if (mode == MODE_FIELD) {
        o.x = -tx;
        o.y = -ty;
        o.z = -tz;
        const real a = alpha[i];
        acc = fabs((double)(a * o.x)) + fabs((double)(a * o.y)) + fabs((double)(a * o.z));
    } else if (mode == MODE_AV) {
        o.x = d.x * ia + tx;
        o.y = d.y * ia + ty;
        o.z = d.z * ia + tz;
        acc = (double)(d.x * o.x) + (double)(d.y * o.y) + (double)(d.z * o.z);
    } else if (mode == MODE_FUSED) {
        const real4<real> x0 = extra[i];
        o.x = -tx - x0.x * ia;
        o.y = -ty - x0.y * ia;
        o.z = -tz - x0.z * ia;
    } else if (mode == MODE_REC) {
        o.x = rx;
        o.y = ry;
        o.z = rz;
    } else {
        const real4<real> b = extra[i];
        o.x = b.x - (d.x * ia + tx);
        o.y = b.y - (d.y * ia + ty);
        o.z = b.z - (d.z * ia + tz);
    }
    out[i] = o;
}

// Tiled gather + combine: one 512-thread workgroup per grid tile loads the potential on the
// tile's (T + P - 1)^3 block (periodic wrap) into LDS once, then every atom binned to the tile
// (two neighbouring threads per atom, sorted order s, atom sidx[s]: the first (P+1)/2 x planes of
// its stencil and the rest, summed with one shuffle) sums its stencil values from LDS with the
// spline weights stored by k_pme_splines, and the first thread's combine_out writes its result.  Per-block partials
// of MODE_FIELD / MODE_AV go to part[tile].  Block b * t.ntiles + tile serves tile `tile` of bead
// b (potential at phi + b K1 K2 K3).
constexpr int GATHER_THREADS = 512;
template <typename real, int P>
__global__ __launch_bounds__(GATHER_THREADS) void k_gather_tiles(PMETiles t, PMEGeom g, const int *__restrict__ tstart,
                                                      const int *__restrict__ tend, const int *__restrict__ sidx,
                                                      const real *__restrict__ spl, const int *__restrict__ loc,
                                                      const real *__restrict__ phi, const real4<real> *__restrict__ dq,
                                                      const real4<real> *__restrict__ row, const real *__restrict__ alpha,
                                                      const real4<real> *__restrict__ extra, real cself, int mode,
                                                      real4<real> *__restrict__ out, double *__restrict__ part,
                                                      const int *__restrict__ conv_it, int iter) {
    const int tile = blockIdx.x, tid = threadIdx.x;
    const int bead = tile / t.ntiles, lt = tile - bead * t.ntiles;
    if (cg_done(conv_it, bead, iter)) return;
    extern __shared__ __align__(16) unsigned char lds_raw[];
    real *blk = reinterpret_cast<real *>(lds_raw);
    __shared__ double lds[16];
    const int BX = t.B[0], BY = t.B[1], BZ = t.B[2];
    const int u2 = lt % t.nt[2], u1 = (lt / t.nt[2]) % t.nt[1], u0 = lt / (t.nt[2] * t.nt[1]);
    const int o0 = u0 * t.T[0], o1 = u1 * t.T[1], o2 = u2 * t.T[2];
    const int K1 = g.K[0], K2 = g.K[1], K3 = g.K[2];
    phi += (size_t)bead * K1 * K2 * K3;
    // block load, four loads in flight per thread before the LDS stores
    const int nth = blockDim.x, bs = t.bsize;
    for (int k0 = tid; k0 < bs; k0 += 4 * nth) {
        real v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int k = k0 + u * nth < bs ? k0 + u * nth : 0;
            const int bz = k % BZ, by = (k / BZ) % BY, bx = k / (BZ * BY);
            int gx = o0 + bx, gy = o1 + by, gz = o2 + bz;
            gx -= gx >= K1 ? K1 : 0;
            gy -= gy >= K2 ? K2 : 0;
            gz -= gz >= K3 ? K3 : 0;
            v[u] = phi[((size_t)gx * K2 + gy) * K3 + gz];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (k0 + u * nth < bs) blk[k0 + u * nth] = v[u];
    }
    (void)BX;
    __syncthreads();
    const int s0 = tstart[tile], s1 = tend[tile];
    double acc[1] = {0.0};
    const int half = tid & 1;
    constexpr int PH = (P + 1) / 2;  // x planes of the first thread of a pair
    for (int s = s0 + (tid >> 1); s < s1; s += blockDim.x >> 1) {  // both threads of a pair: same s
        const real *v = spl + (size_t)s * 6 * P;
        real t1[P], d1[P], t2[P], d2[P], t3[P], d3[P];
#pragma unroll
        for (int k = 0; k < P; ++k) {
            t1[k] = v[k];
            d1[k] = v[P + k];
            t2[k] = v[2 * P + k];
            d2[k] = v[3 * P + k];
            t3[k] = v[4 * P + k];
            d3[k] = v[5 * P + k];
        }
        const int lc = loc[s];
        const int lx = lc & 1023, ly = (lc >> 10) & 1023, lz = lc >> 20;
        real g1 = 0, g2 = 0, g3 = 0;
#pragma unroll
        for (int ixh = 0; ixh < PH; ++ixh) {
            const int ix = half ? PH + ixh : ixh;
            if (ix >= P) break;
            // spline weights picked with compile-time indices (registers, not scratch)
            const int ih = PH + ixh < P ? PH + ixh : P - 1;
            const real d1x = half ? d1[ih] : d1[ixh], t1x = half ? t1[ih] : t1[ixh];
            real s_tt = 0, s_dt = 0, s_td = 0;
#pragma unroll
            for (int iy = 0; iy < P; ++iy) {
                const real *line = blk + ((lx + ix) * BY + (ly + iy)) * BZ + lz;
                real lt = 0, ld = 0;
#pragma unroll
                for (int iz = 0; iz < P; ++iz) {
                    const real w = line[iz];
                    lt += w * t3[iz];
                    ld += w * d3[iz];
                }
                s_tt += lt * t2[iy];
                s_dt += lt * d2[iy];
                s_td += ld * t2[iy];
            }
            g1 += s_tt * d1x;
            g2 += s_dt * t1x;
            g3 += s_td * t1x;
        }
        g1 += __shfl_xor(g1, 1, WAVE);
        g2 += __shfl_xor(g2, 1, WAVE);
        g3 += __shfl_xor(g3, 1, WAVE);
        if (half) continue;
        const double *es = g.escale;
        const real rx = (real)es[0] * g1 + (real)es[1] * g2 + (real)es[2] * g3;
        const real ry = (real)es[3] * g1 + (real)es[4] * g2 + (real)es[5] * g3;
        const real rz = (real)es[6] * g1 + (real)es[7] * g2 + (real)es[8] * g3;
        double a = 0.0;
        combine_out<real>(mode, sidx[s], rx, ry, rz, dq, row, alpha, extra, cself, out, a);
        acc[0] += a;
    }
    if (part) {
        block_sum<1>(acc, lds);
        if (tid == 0) part[tile] = acc[0];
    }
}

template <typename real, int P>
__global__ __launch_bounds__(256) void k_gather_combine(int n, const double *__restrict__ pos, PMEGeom g,
                                                        const real *__restrict__ phi, const real4<real> *__restrict__ dq,
                                                        const real4<real> *__restrict__ row, const real *__restrict__ alpha,
                                                        const real4<real> *__restrict__ extra, real cself, int mode,
                                                        real4<real> *__restrict__ out, double *__restrict__ part,
                                                        const int *__restrict__ conv_it, int iter) {
    if (conv_it && iter >= *conv_it) return;
    __shared__ double lds[16];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double acc[1] = {0.0};
    if (i < n) {
        int base[3];
        real f[3];
        stencil<real, P>(pos, i, g, base, f);
        real t1[P], d1[P], t2[P], d2[P], t3[P], d3[P];
        bspline<real, P, false>(f[0], t1, d1, nullptr);
        bspline<real, P, false>(f[1], t2, d2, nullptr);
        bspline<real, P, false>(f[2], t3, d3, nullptr);
        const int K1 = g.K[0], K2 = g.K[1], K3 = g.K[2];
        real g1 = 0, g2 = 0, g3 = 0;
#pragma unroll
        for (int ix = 0; ix < P; ++ix) {
            int gx = base[0] + ix;
            gx -= gx >= K1 ? K1 : 0;
            real s_tt = 0, s_dt = 0, s_td = 0;  // sums over (iy, iz) of phi t2 t3, phi d2 t3, phi t2 d3
#pragma unroll
            for (int iy = 0; iy < P; ++iy) {
                int gy = base[1] + iy;
                gy -= gy >= K2 ? K2 : 0;
                const real *line = phi + ((size_t)gx * K2 + gy) * K3;
                real lt = 0, ld = 0;
#pragma unroll
                for (int iz = 0; iz < P; ++iz) {
                    int gz = base[2] + iz;
                    gz -= gz >= K3 ? K3 : 0;
                    const real v = line[gz];
                    lt += v * t3[iz];
                    ld += v * d3[iz];
                }
                s_tt += lt * t2[iy];
                s_dt += lt * d2[iy];
                s_td += ld * t2[iy];
            }
            g1 += s_tt * d1[ix];
            g2 += s_dt * t1[ix];
            g3 += s_td * t1[ix];
        }
        const double *es = g.escale;
        const real rx = (real)es[0] * g1 + (real)es[1] * g2 + (real)es[2] * g3;
        const real ry = (real)es[3] * g1 + (real)es[4] * g2 + (real)es[5] * g3;
        const real rz = (real)es[6] * g1 + (real)es[7] * g2 + (real)es[8] * g3;
        combine_out<real>(mode, i, rx, ry, rz, dq, row, alpha, extra, cself, out, acc[0]);
    }
    if (part) {
        block_sum<1>(acc, lds);
        if (threadIdx.x == 0) part[blockIdx.x] = acc[0];
    }
}

// Force gather: dU_rec/dd and dU_rec/dr at fixed q, d (the autodiff of pme.py:390-411 written
// out, with the second spline derivative), FP64 outputs; also gp = KE (rowf + dU_rec/dd - c_self d),
// the total dE/dd of the atom for the covalent-frame pull-back.  rec_dd (optional) receives
// dU_rec/dd itself (for the strain derivative).
template <typename real, int P>
__global__ __launch_bounds__(256) void k_gather_forces(int n, int nper, const double *__restrict__ pos, PMEGeom g,
                                                       const real *__restrict__ phi, const real4<real> *__restrict__ dq,
                                                       const double *__restrict__ dd, const double *__restrict__ rowf,
                                                       double cself, double *__restrict__ rec_dr, double *__restrict__ gp,
                                                       double *__restrict__ rec_dd) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int base[3];
    real f[3];
    stencil<real, P>(pos, i, g, base, f);
    real t[3][P], dt[3][P], d2t[3][P];
#pragma unroll
    for (int a = 0; a < 3; ++a) bspline<real, P, true>(f[a], t[a], dt[a], d2t[a]);
    const real4<real> d = dq[i];
    real e[3];
    scaled_dipole<real>(d, g, e);
    const real q = d.w;
    const int K1 = g.K[0], K2 = g.K[1], K3 = g.K[2];
    phi += (size_t)(i / nper) * K1 * K2 * K3;
    // phi-weighted sums S[wx][wy][wz] with w = 0 (theta), 1 (dtheta), 2 (d2theta), needed:
    // gd_a = S(d along a), gw_a = q S(d along a) + sum_b e_b S(d along a, d along b / d2 if a == b)
    double s100 = 0, s010 = 0, s001 = 0, s200 = 0, s020 = 0, s002 = 0, s110 = 0, s101 = 0, s011 = 0;
#pragma unroll
    for (int ix = 0; ix < P; ++ix) {
        int gx = base[0] + ix;
        gx -= gx >= K1 ? K1 : 0;
        real a_0 = 0, a_y = 0, a_z = 0, a_yy = 0, a_zz = 0, a_yz = 0;
#pragma unroll
        for (int iy = 0; iy < P; ++iy) {
            int gy = base[1] + iy;
            gy -= gy >= K2 ? K2 : 0;
            const real *line = phi + ((size_t)gx * K2 + gy) * K3;
            real l0 = 0, l1 = 0, l2 = 0;
#pragma unroll
            for (int iz = 0; iz < P; ++iz) {
                int gz = base[2] + iz;
                gz -= gz >= K3 ? K3 : 0;
                const real v = line[gz];
                l0 += v * t[2][iz];
                l1 += v * dt[2][iz];
                l2 += v * d2t[2][iz];
            }
            a_0 += l0 * t[1][iy];
            a_y += l0 * dt[1][iy];
            a_z += l1 * t[1][iy];
            a_yy += l0 * d2t[1][iy];
            a_zz += l2 * t[1][iy];
            a_yz += l1 * dt[1][iy];
        }
        s100 += (double)(a_0 * dt[0][ix]);
        s010 += (double)(a_y * t[0][ix]);
        s001 += (double)(a_z * t[0][ix]);
        s200 += (double)(a_0 * d2t[0][ix]);
        s020 += (double)(a_yy * t[0][ix]);
        s002 += (double)(a_zz * t[0][ix]);
        s110 += (double)(a_y * dt[0][ix]);
        s101 += (double)(a_z * dt[0][ix]);
        s011 += (double)(a_yz * t[0][ix]);
    }
    const double gd[3] = {s100, s010, s001};
    const double qq = (double)q, e0 = (double)e[0], e1 = (double)e[1], e2 = (double)e[2];
    const double gw[3] = {qq * s100 + e0 * s200 + e1 * s110 + e2 * s101, qq * s010 + e0 * s110 + e1 * s020 + e2 * s011,
                          qq * s001 + e0 * s101 + e1 * s011 + e2 * s002};
    const double *es = g.escale;
#pragma unroll
    for (int b = 0; b < 3; ++b) {
        const double rdd = es[3 * b] * gd[0] + es[3 * b + 1] * gd[1] + es[3 * b + 2] * gd[2];
        rec_dr[3 * i + b] = es[3 * b] * gw[0] + es[3 * b + 1] * gw[1] + es[3 * b + 2] * gw[2];
        gp[3 * i + b] = KE * (rowf[3 * i + b] + rdd - cself * dd[3 * i + b]);
        if (rec_dd) rec_dd[3 * i + b] = rdd;
    }
}

}  // namespace pgm
