// pairs.h -- neighbour rows and the real-space pair kernels of the pGM force field.
//
// Rows: for every atom i, the partners j with |min_image(r_i - r_j)| below a cutoff, stored as
// int32 in row-major (N, cap) order; every pair appears in both rows (no scatter, no atomics: the
// per-atom sums are deterministic).  Rows are built from a cell list every few steps with the
// cutoff + skin, and compacted to the cutoff at every force evaluation.
//
// The force pass uses rowl<real>() lanes per atom (lane l takes entries l, l + lanes, ...), so one
// wave64 covers 64/lanes atoms; the partial sums are reduced across the lanes with shuffles: 16 in
// FP32 (measured 7 % faster than 8), 8 in FP64 (3 % faster than 16).  The
// displacement is formed in FP64 from the FP64 positions (exact for intramolecular pairs whatever
// the absolute coordinates) and then rounded to `real`.
#pragma once

#include "common.h"

namespace pgm {

constexpr int ROWL = 16;  // the most lanes per atom of the force pass (buffer sizes)
template <typename real> constexpr int rowl() { return sizeof(real) == 4 ? 16 : 8; }

// Packed per-atom geometry xr: x, y, z [nm] and R^2 (Gaussian radius squared) [nm^2].
// ---------------------------------------------------------------------------- cells
// Cell of every atom in fractional coordinates of the box; key = bead * ncell + (c0 * n1 + c1) * n2 + c2
// (nper atoms per bead).
__global__ void k_cell_keys(int n, int nper, const double *__restrict__ pos, Box box, int3 nc, int *__restrict__ key,
                            int *__restrict__ atom) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double x = pos[3 * i], y = pos[3 * i + 1], z = pos[3 * i + 2];
    const double *hi = box.hinv;  // u_b = sum_a r_a hinv[a][b]
    double u[3] = {x * hi[0] + y * hi[3] + z * hi[6], x * hi[1] + y * hi[4] + z * hi[7], x * hi[2] + y * hi[5] + z * hi[8]};
    int c[3];
    const int ncs[3] = {nc.x, nc.y, nc.z};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        double f = u[a] - floor(u[a]);
        int k = (int)(f * ncs[a]);
        c[a] = k < 0 ? 0 : (k >= ncs[a] ? ncs[a] - 1 : k);
    }
    key[i] = (i / nper) * (nc.x * nc.y * nc.z) + (c[0] * nc.y + c[1]) * nc.z + c[2];
    atom[i] = i;
}

__global__ void k_fill_int(int n, int *__restrict__ p, int v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// cell_start[c] / cell_end[c] from the sorted keys.
__global__ void k_cell_bounds(int n, const int *__restrict__ skey, int *__restrict__ start, int *__restrict__ end) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const int k = skey[s];
    if (s == 0 || skey[s - 1] != k) start[k] = s;
    if (s == n - 1 || skey[s + 1] != k) end[k] = s + 1;
}

// Full rows within rc_list (cutoff + skin) by a search of the neighbouring cells (27, or every
// distinct cell along an axis with fewer than 3 cells) of the atom's own bead.  One thread per atom;
// the partners come in cell order and, within a cell, in atom order (deterministic).
__global__ void k_build_rows(int n, int nper, const double *__restrict__ pos, Box box, int3 nc, const int *__restrict__ satom,
                             const int *__restrict__ cstart, const int *__restrict__ cend, double rc2, int cap,
                             int *__restrict__ rows, int *__restrict__ count, int *__restrict__ maxcount) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double xi = pos[3 * i], yi = pos[3 * i + 1], zi = pos[3 * i + 2];
    const double *hi = box.hinv;
    const double u[3] = {xi * hi[0] + yi * hi[3] + zi * hi[6], xi * hi[1] + yi * hi[4] + zi * hi[7],
                         xi * hi[2] + yi * hi[5] + zi * hi[8]};
    const int ncs[3] = {nc.x, nc.y, nc.z};
    int c[3], lo[3], hiO[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        double f = u[a] - floor(u[a]);
        int k = (int)(f * ncs[a]);
        c[a] = k < 0 ? 0 : (k >= ncs[a] ? ncs[a] - 1 : k);
        lo[a] = ncs[a] >= 3 ? -1 : 0;
        hiO[a] = ncs[a] >= 3 ? 1 : ncs[a] - 1;
    }
    int cnt = 0;
    int *row = rows + (size_t)i * cap;
    const int cell0 = (i / nper) * (nc.x * nc.y * nc.z);
    for (int o0 = lo[0]; o0 <= hiO[0]; ++o0) {
        const int c0 = (c[0] + o0 + ncs[0]) % ncs[0];
        for (int o1 = lo[1]; o1 <= hiO[1]; ++o1) {
            const int c1 = (c[1] + o1 + ncs[1]) % ncs[1];
            for (int o2 = lo[2]; o2 <= hiO[2]; ++o2) {
                const int c2 = (c[2] + o2 + ncs[2]) % ncs[2];
                const int cell = cell0 + (c0 * nc.y + c1) * nc.z + c2;
                const int s0 = cstart[cell], s1 = cend[cell];
                for (int s = s0; s < s1; ++s) {
                    const int j = satom[s];
                    if (j == i) continue;
                    double dx = xi - pos[3 * j], dy = yi - pos[3 * j + 1], dz = zi - pos[3 * j + 2];
                    min_image(dx, dy, dz, box);
                    if (dx * dx + dy * dy + dz * dz < rc2) {
                        if (cnt < cap) row[cnt] = j;
                        ++cnt;
                    }
                }
            }
        }
    }
    count[i] = cnt;
    atomicMax(maxcount, cnt);
}

// The same rows built cell by cell: one workgroup per cell (bead * ncell + cell, atoms in sorted
// order), its atoms the owners (one thread each, in chunks of blockDim), the neighbouring cells'
// positions staged through LDS in chunks of blockDim and tested by every owner.  Cells and atoms are
// visited in k_build_rows's order; each position is read once per workgroup instead of once per
// owner.  With at least three cells per axis (cell heights >= the list cutoff) the only image of j
// within the cutoff of i lies in the neighbour cell as visited, so positions are wrapped into the
// box (r - floor(r H^-1) H, the cell's own fractional coordinates) and the neighbour cell's lattice
// shift is added when staging: one subtraction per pair instead of a minimum image.  Fewer cells
// per axis keep the per-pair minimum image.
__global__ __launch_bounds__(256) void k_build_rows_cells(int ncell_bead, const double *__restrict__ pos, Box box, int3 nc,
                                                          const int *__restrict__ satom, const int *__restrict__ cstart,
                                                          const int *__restrict__ cend, double rc2, int cap,
                                                          int *__restrict__ rows, int *__restrict__ count,
                                                          int *__restrict__ maxcount) {
    __shared__ double4 sp[256];  // staged candidate: position (shifted) and index (as a double)
    const int cell = blockIdx.x;
    const int bead = cell / ncell_bead, lc = cell - bead * ncell_bead;
    const int own0 = cstart[cell], own1 = cend[cell];
    if (own1 <= own0) return;  // empty cell (uniform)
    const int ncs[3] = {nc.x, nc.y, nc.z};
    const int c[3] = {lc / (nc.y * nc.z), (lc / nc.z) % nc.y, lc % nc.z};
    int lo[3], hiO[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        lo[a] = ncs[a] >= 3 ? -1 : 0;
        hiO[a] = ncs[a] >= 3 ? 1 : ncs[a] - 1;
    }
    const bool shifted = ncs[0] >= 3 && ncs[1] >= 3 && ncs[2] >= 3;
    // position wrapped into the box: r - sum_a floor(u_a) h_a, u = r H^-1
    auto wrapped = [&](int j, double &x, double &y, double &z) {
        x = pos[3 * j];
        y = pos[3 * j + 1];
        z = pos[3 * j + 2];
        if (!shifted) return;
        const double *hi = box.hinv, *h = box.h;
        const double n0 = floor(x * hi[0] + y * hi[3] + z * hi[6]), n1 = floor(x * hi[1] + y * hi[4] + z * hi[7]),
                     n2 = floor(x * hi[2] + y * hi[5] + z * hi[8]);
        x -= n0 * h[0] + n1 * h[3] + n2 * h[6];
        y -= n0 * h[1] + n1 * h[4] + n2 * h[7];
        z -= n0 * h[2] + n1 * h[5] + n2 * h[8];
    };
    for (int o0s = own0; o0s < own1; o0s += blockDim.x) {
        const int s = o0s + threadIdx.x;
        const bool owner = s < own1;
        const int i = owner ? satom[s] : 0;
        double xi = 0.0, yi = 0.0, zi = 0.0;
        if (owner) wrapped(i, xi, yi, zi);
        int cnt = 0;
        int *row = rows + (size_t)i * cap;
        for (int q0 = lo[0]; q0 <= hiO[0]; ++q0) {
            const int c0 = (c[0] + q0 + ncs[0]) % ncs[0];
            for (int q1 = lo[1]; q1 <= hiO[1]; ++q1) {
                const int c1 = (c[1] + q1 + ncs[1]) % ncs[1];
                for (int q2 = lo[2]; q2 <= hiO[2]; ++q2) {
                    const int c2 = (c[2] + q2 + ncs[2]) % ncs[2];
                    const int nb = bead * ncell_bead + (c0 * nc.y + c1) * nc.z + c2;
                    const int s0 = cstart[nb], s1 = cend[nb];
                    // lattice shift of the neighbour cell as visited (wrap across a box face)
                    double shx = 0.0, shy = 0.0, shz = 0.0;
                    if (shifted) {
                        const int q[3] = {q0, q1, q2};
#pragma unroll
                        for (int a = 0; a < 3; ++a) {
                            const int w = c[a] + q[a] < 0 ? -1 : (c[a] + q[a] >= ncs[a] ? 1 : 0);
                            shx += w * box.h[3 * a];
                            shy += w * box.h[3 * a + 1];
                            shz += w * box.h[3 * a + 2];
                        }
                    }
                    for (int t0 = s0; t0 < s1; t0 += blockDim.x) {
                        const int m = min((int)blockDim.x, s1 - t0);
                        __syncthreads();
                        if ((int)threadIdx.x < m) {
                            const int j = satom[t0 + threadIdx.x];
                            double x, y, z;
                            wrapped(j, x, y, z);
                            sp[threadIdx.x] = make_double4(x + shx, y + shy, z + shz, (double)j);
                        }
                        __syncthreads();
                        if (owner) {
                            if (shifted) {
                                // eight candidates' distances at a time (their LDS reads in flight
                                // together), then the hits appended in candidate order
                                for (int u0 = 0; u0 < m; u0 += 8) {
                                    double d2[8];
                                    int jj[8];
#pragma unroll
                                    for (int k = 0; k < 8; ++k) {
                                        const int u = u0 + k < m ? u0 + k : m - 1;
                                        const double4 c4 = sp[u];
                                        jj[k] = (int)c4.w;
                                        const double dx = xi - c4.x, dy = yi - c4.y, dz = zi - c4.z;
                                        d2[k] = dx * dx + dy * dy + dz * dz;
                                    }
#pragma unroll
                                    for (int k = 0; k < 8; ++k)
                                        if (u0 + k < m && d2[k] < rc2 && jj[k] != i) {
                                            if (cnt < cap) row[cnt] = jj[k];
                                            ++cnt;
                                        }
                                }
                            } else {
                                for (int u = 0; u < m; ++u) {
                                    const double4 c4 = sp[u];
                                    const int j = (int)c4.w;
                                    if (j == i) continue;
                                    double dx = xi - c4.x, dy = yi - c4.y, dz = zi - c4.z;
                                    min_image(dx, dy, dz, box);
                                    if (dx * dx + dy * dy + dz * dz < rc2) {
                                        if (cnt < cap) row[cnt] = j;
                                        ++cnt;
                                    }
                                }
                            }
                        }
                    }
                }
            }
        }
        if (owner) count[i] = cnt;
        // the largest count: one atomic per wave, and only above the value already there (a
        // same-address atomic per atom serialises in one L2 channel)
        int m = owner ? cnt : 0;
#pragma unroll
        for (int o = WAVE / 2; o > 0; o >>= 1) m = max(m, __shfl_xor(m, o, WAVE));
        if (threadIdx.x % WAVE == 0 && m > __hip_atomic_load(maxcount, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
            atomicMax(maxcount, m);
    }
}

// Compacted rows within rc: one wave per atom, ballot + prefix count, one pass.  "Near" partners
// (those for which a series branch of the radial kernels may be taken: (a r)^2 or (beta r)^2
// below 1.05, FP64) fill the row from the front, the others from the back (row[cap - 1 - k]);
// count_near[i] and count[i] - count_near[i] give the two parts, so the kernels run the far part on
// the closed form only (no divergent series).  Entry order within each part is the build order.
__global__ __launch_bounds__(256) void k_compact_rows(int n, const double4 *__restrict__ xr, Box box,
                                                      const int *__restrict__ rows_full, const int *__restrict__ count_full,
                                                      int cap_full, double rc2, double beta2, int cap, int *__restrict__ rows,
                                                      int *__restrict__ count, int *__restrict__ count_near) {
    const int lane = threadIdx.x % WAVE;
    const int i = blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE;
    if (i >= n) return;
    const double4 a = xr[i];
    const int cf = count_full[i] < cap_full ? count_full[i] : cap_full;
    const int *rf = rows_full + (size_t)i * cap_full;
    int *rc = rows + (size_t)i * cap;
    int on = 0, of = 0;
    for (int e0 = 0; e0 < cf; e0 += WAVE) {
        const int e = e0 + lane;
        bool kn = false, kf = false;
        int j = 0;
        if (e < cf) {
            j = rf[e];
            const double4 b = xr[j];
            double dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
            min_image(dx, dy, dz, box);
            const double r2 = dx * dx + dy * dy + dz * dz;
            const bool near = r2 < 1.05 * 2.0 * (a.w + b.w) || beta2 * r2 < 1.05;
            kn = r2 < rc2 && near;
            kf = r2 < rc2 && !near;
        }
        const unsigned long long below = (1ull << lane) - 1ull;
        const unsigned long long mn = __ballot(kn), mf = __ballot(kf);
        const int pn = on + __popcll(mn & below), pf = of + __popcll(mf & below);
        if (kn && pn + of < cap) rc[pn] = j;                 // the far part starts at cap - of
        if (kf && on + pf < cap) rc[cap - 1 - pf] = j;
        on += __popcll(mn);
        of += __popcll(mf);
    }
    if (lane == 0) {
        const int tot = on + of < cap ? on + of : cap;
        count_near[i] = on < tot ? on : tot;
        count[i] = tot;
    }
}

// Sliced-ELL layout of the rows used by the CG matvec: entry e of atom i lives at
// ell(e, i) = (e / RL) * (n * RL) + i * RL + e % RL, so RL lanes per atom, 64 / RL atoms per wave,
// read 64 consecutive words per load.
constexpr int RL = 4;
PGM_DEV size_t ell(int e, int i, int n) { return (size_t)(e / RL) * ((size_t)n * RL) + (size_t)i * RL + (e % RL); }

// Pair tensors of the compacted rows, once per geometry: partner index, G1 and G2 of every entry
// (the dipole field needs only these) in the ELL layout, near part (series branches) first, then
// the far part (closed form).  RL lanes per atom.
template <typename real>
__global__ __launch_bounds__(256) void k_pair_tensors(int n, const int *__restrict__ rows, const int *__restrict__ count,
                                                      const int *__restrict__ count_near, int cap,
                                                      const double4 *__restrict__ xr, Box box, real beta, double rce2,
                                                      int *__restrict__ jell, real *__restrict__ G1, real *__restrict__ G2) {
    const int lane = threadIdx.x % RL;
    const int i = blockIdx.x * (blockDim.x / RL) + threadIdx.x / RL;
    if (i >= n) return;
    const double4 a = xr[i];
    const int c = count[i], cn = count_near[i];
    const size_t o = (size_t)i * cap;
    for (int e = lane; e < c; e += RL) {
        const int j = e < cn ? rows[o + e] : rows[o + cap - 1 - (e - cn)];
        const double4 b = xr[j];
        double dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
        min_image(dx, dy, dz, box);
        const real x = (real)dx, y = (real)dy, z = (real)dz;
        const real aij = rsqrt_(real(2) * (real)(a.w + b.w));
        real G[3] = {real(0), real(0), real(0)};
        if (dx * dx + dy * dy + dz * dz < rce2) {  // electrostatics within the elec cutoff only
            if (e < cn)
                pair_G<real, 3>(aij, beta, x * x + y * y + z * z, G);
            else
                pair_G_far<real, 3>(aij, beta, x * x + y * y + z * z, G);
        }
        const size_t q = ell(e, i, n);
        jell[q] = j;
        G1[q] = G[1];
        G2[q] = G[2];
    }
}

// ---------------------------------------------------------------------------- field rows
// out_i = sum_k de_ik/dd_i = sum_k (-q_k G1 - G2 (d_k . x)) x + G1 d_k, x = r_i - r_k
// (minus the direct-space field at atom i; pgm_jax/md/forcefield.py:1395-1402).
// dq holds (d_x, d_y, d_z, q) per atom; WITHQ adds the charge term.
template <typename real, bool WITHQ>
PGM_DEV void row_field_entry(const double4 &a, const double4 &b, const real4<real> &dk, const Box &box, real g1, real g2,
                             real &ex, real &ey, real &ez) {
    double dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
    min_image(dx, dy, dz, box);
    const real x = (real)dx, y = (real)dy, z = (real)dz;
    const real dkx = dk.x * x + dk.y * y + dk.z * z;
    real cf = -g2 * dkx;
    if (WITHQ) cf -= dk.w * g1;
    ex += cf * x + g1 * dk.x;
    ey += cf * y + g1 * dk.y;
    ez += cf * z + g1 * dk.z;
}

// out_i = sum_k de_ik/dd_i = sum_k (-q_k G1 - G2 (d_k . x)) x + G1 d_k, x = r_i - r_k
// (minus the direct-space field at atom i; pgm_jax/md/forcefield.py:1395-1402), with the partner
// indices and G1, G2 precomputed by k_pair_tensors (ELL layout, RL lanes per atom).  dq holds
// (d_x, d_y, d_z, q) per atom; WITHQ adds the charge term.
template <typename real, bool WITHQ>
__global__ __launch_bounds__(256) void k_row_field(int n, int nper, const int *__restrict__ jell, const int *__restrict__ count,
                                                   const real *__restrict__ G1, const real *__restrict__ G2,
                                                   const double4 *__restrict__ xr, const real4<real> *__restrict__ dq, Box box,
                                                   real4<real> *__restrict__ out, const int *__restrict__ conv_it, int iter,
                                                   Ext3 ext = {}) {
    const int lane = threadIdx.x % RL;
    const int i = blockIdx.x * (blockDim.x / RL) + threadIdx.x / RL;
    if (i >= n || cg_done(conv_it, i / nper, iter)) return;
    const double4 a = xr[i];
    const int c = count[i];
    real ex = 0, ey = 0, ez = 0;
    // Entries in batches of RF_UNROLL per lane: the partner indices and kernels of the batch are
    // loaded first, then the partners' positions and dipoles, so the dependent loads of the batch
    // are in flight together instead of one chain per entry.  Slots past the row's end are masked
    // (own index, zero kernels); the summation order per lane is unchanged.
    constexpr int RF_UNROLL = 4;
    for (int e0 = lane; e0 < c; e0 += RF_UNROLL * RL) {
        int j[RF_UNROLL];
        real g1[RF_UNROLL], g2[RF_UNROLL];
#pragma unroll
        for (int u = 0; u < RF_UNROLL; ++u) {
            const int e = e0 + u * RL;
            const bool ok = e < c;
            const size_t q = ell(ok ? e : 0, i, n);
            j[u] = ok ? jell[q] : i;
            g1[u] = ok ? G1[q] : real(0);
            g2[u] = ok ? G2[q] : real(0);
        }
        double4 b[RF_UNROLL];
        real4<real> d[RF_UNROLL];
#pragma unroll
        for (int u = 0; u < RF_UNROLL; ++u) {
            b[u] = xr[j[u]];
            d[u] = dq[j[u]];
        }
#pragma unroll
        for (int u = 0; u < RF_UNROLL; ++u) row_field_entry<real, WITHQ>(a, b[u], d[u], box, g1[u], g2[u], ex, ey, ez);
    }
    ex = group_sum<RL>(ex);
    ey = group_sum<RL>(ey);
    ez = group_sum<RL>(ez);
    if (lane == 0) {
        real4<real> v;
        v.x = ex;
        v.y = ey;
        v.z = ez;
        v.w = 0;
        if (WITHQ) {  // a uniform external field: its dU/dd is -ext
            v.x -= (real)ext.x;
            v.y -= (real)ext.y;
            v.z -= (real)ext.z;
        }
        out[i] = v;
    }
}

// ---------------------------------------------------------------------------- pair forces
// Per atom i, over its row (forcefield.py:2294-2360, :2116-2146):
//   e_ik  = q_i q_k G0 + (q_i d_k.x - q_k d_i.x) G1 - G2 (d_i.x)(d_k.x) + G1 d_i.d_k
//   de/dx = x [-q_i q_k G1 - t G2 + G3 (d_i.x)(d_k.x) - G2 d_i.d_k] + q_i G1 d_k - q_k G1 d_i
//           - G2 (d_i (d_k.x) + d_k (d_i.x)),   t = q_i d_k.x - q_k d_i.x
//   LJ (intermolecular only): eps (s^12 - 2 s^6), s = rmin/r, (1/r) de/dr = 12 eps (s^6 - s^12)/r^2,
//   one FP64 reciprocal of r^2 per pair; the partner's (rmin/2, sqrt eps, molecule) in one 32-byte
//   load (ljp).
// Electrostatic terms stop at the elec cutoff (r^2 < rce2), LJ at the van der Waals cutoff (r^2 < rcv2;
// pgm_jax/md/forcefield.py:1197-1238, 2355): the rows span both.
// Outputs (FP64): g_el = sum de/dx, g_lj, the row field sum de/dd_i, and per-block energy sums
// (each pair counted in both rows: the caller halves them).  With VIRIAL, also the per-block sums
// of (KE de/dx + de_LJ/dx) (x) X, X = x - o_i + o_k with o the offsets of the atoms from their
// molecules' centres of mass: the pair part of the molecular strain derivative (again twice).
#define PAIR_FORCE_ENTRY(GFUN) \
            const double4 b = xr[j]; \
            double dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z; \
            min_image(dx, dy, dz, box); \
            const real x = (real)dx, y = (real)dy, z = (real)dz; \
            const real r2 = x * x + y * y + z * z; \
            const double r2d = dx * dx + dy * dy + dz * dz; \
            const real aij = rsqrt_(real(2) * (real)(a.w + b.w)); \
            real G[4] = {real(0), real(0), real(0), real(0)}; \
            if (r2d < rce2) GFUN<real, 4>(aij, beta, r2, G); \
            const real4<real> dk = dq[j]; \
            const real qk = dk.w; \
            const real dix = di.x * x + di.y * y + di.z * z; \
            const real dkx = dk.x * x + dk.y * y + dk.z * z; \
            const real didk = di.x * dk.x + di.y * dk.y + di.z * dk.z; \
            const real t = qi * dkx - qk * dix; \
            const real ep = qi * qk * G[0] + t * G[1] - G[2] * dix * dkx + G[1] * didk; \
            const real radial = -qi * qk * G[1] - t * G[2] + G[3] * dix * dkx - G[2] * didk; \
            const real qiG1 = qi * G[1], qkG1 = qk * G[1]; \
            const real ge0 = radial * x + qiG1 * dk.x - qkG1 * di.x - G[2] * (di.x * dkx + dk.x * dix); \
            const real ge1 = radial * y + qiG1 * dk.y - qkG1 * di.y - G[2] * (di.y * dkx + dk.y * dix); \
            const real ge2 = radial * z + qiG1 * dk.z - qkG1 * di.z - G[2] * (di.z * dkx + dk.z * dix); \
            gx += (double)ge0; \
            gy += (double)ge1; \
            gz += (double)ge2; \
            ee += (double)ep; \
            const real cfield = -qk * G[1] - G[2] * dkx; \
            fx += (double)(cfield * x + G[1] * dk.x); \
            fy += (double)(cfield * y + G[1] * dk.y); \
            fz += (double)(cfield * z + G[1] * dk.z); \
            const double4 ljk = ljp[j]; \
            const double eps = sei * ljk.y; \
            double l0 = 0.0, l1 = 0.0, l2 = 0.0; \
            if ((int)ljk.z != mi && eps != 0.0 && r2d < rcv2) { \
                const double rmin = rhi + ljk.x; \
                const double ir2 = 1.0 / r2d; \
                const double s2 = rmin * rmin * ir2; \
                const double s6 = s2 * s2 * s2; \
                el += eps * (s6 * s6 - 2.0 * s6); \
                const double dl = eps * 12.0 * (s6 - s6 * s6) * ir2; \
                l0 = dl * dx; \
                l1 = dl * dy; \
                l2 = dl * dz; \
            } \
            lx += l0; \
            ly += l1; \
            lz += l2; \
            if (VIRIAL) { \
                const double X[3] = {dx - oi[0] + off[3 * j], dy - oi[1] + off[3 * j + 1], dz - oi[2] + off[3 * j + 2]}; \
                const double v[3] = {KE * (double)ge0 + l0, KE * (double)ge1 + l1, KE * (double)ge2 + l2}; \
                _Pragma("unroll") for (int p_ = 0; p_ < 3; ++p_) \
                _Pragma("unroll") for (int q_ = 0; q_ < 3; ++q_) w[3 * p_ + q_] += v[p_] * X[q_]; \
            }

// Grid (blocks per bead, beads): partials per block of one bead, at blockIdx.y * gridDim.x + blockIdx.x.
template <typename real, bool VIRIAL>
__global__ __launch_bounds__(256) void k_pair_forces(int nper, const int *__restrict__ rows, const int *__restrict__ count,
                                                     const int *__restrict__ count_near, int cap,
                                                     const double4 *__restrict__ xr, const real4<real> *__restrict__ dq, Box box,
                                                     real beta, double rce2, double rcv2, const double4 *__restrict__ ljp,
                                                     double *__restrict__ g_el,
                                                     double *__restrict__ g_lj, double *__restrict__ rowf,
                                                     double *__restrict__ epart, const double *__restrict__ off,
                                                     double *__restrict__ wpart, Ext3 ext = {}) {
    __shared__ double lds[9 * 16];
    constexpr int RL_ = rowl<real>();
    const int lane = threadIdx.x % RL_;
    const int li = blockIdx.x * (blockDim.x / RL_) + threadIdx.x / RL_;
    const int i = blockIdx.y * nper + li;
    const int pb = blockIdx.y * gridDim.x + blockIdx.x;
    double esum[2] = {0.0, 0.0};
    double w[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (li < nper) {
        const double4 a = xr[i];
        const real4<real> di = dq[i];
        const real qi = di.w;
        const double4 lji = ljp[i];
        const double rhi = lji.x, sei = lji.y;
        const int mi = (int)lji.z;
        const int c = count[i];
        const int *row = rows + (size_t)i * cap;
        double oi[3] = {0.0, 0.0, 0.0};
        if (VIRIAL) {
            oi[0] = off[3 * i];
            oi[1] = off[3 * i + 1];
            oi[2] = off[3 * i + 2];
        }
        double gx = 0, gy = 0, gz = 0, lx = 0, ly = 0, lz = 0, fx = 0, fy = 0, fz = 0, ee = 0, el = 0;
        const int cn = count_near[i];
        for (int e = lane; e < cn; e += RL_) {
            const int j = row[e];
            PAIR_FORCE_ENTRY(pair_G)
        }
        for (int e = cap - (c - cn) + lane; e < cap; e += RL_) {
            const int j = row[e];
            PAIR_FORCE_ENTRY(pair_G_far)
        }
        gx = group_sum<RL_>(gx);
        gy = group_sum<RL_>(gy);
        gz = group_sum<RL_>(gz);
        lx = group_sum<RL_>(lx);
        ly = group_sum<RL_>(ly);
        lz = group_sum<RL_>(lz);
        fx = group_sum<RL_>(fx);
        fy = group_sum<RL_>(fy);
        fz = group_sum<RL_>(fz);
        if (lane == 0) {
            g_el[3 * i] = gx;
            g_el[3 * i + 1] = gy;
            g_el[3 * i + 2] = gz;
            g_lj[3 * i] = lx;
            g_lj[3 * i + 1] = ly;
            g_lj[3 * i + 2] = lz;
            rowf[3 * i] = fx - ext.x;  // a uniform external field: its dU/dd is -ext
            rowf[3 * i + 1] = fy - ext.y;
            rowf[3 * i + 2] = fz - ext.z;
        }
        esum[0] = ee;
        esum[1] = el;
    }
    block_sum<2>(esum, lds);
    if (threadIdx.x == 0) {
        epart[2 * pb] = esum[0];
        epart[2 * pb + 1] = esum[1];
    }
    if (VIRIAL) {
        block_sum<9>(w, lds);
        if (threadIdx.x == 0)
            for (int k = 0; k < 9; ++k) wpart[9 * pb + k] = w[k];
    }
}

// ---------------------------------------------------------------------------- molecule-blocked field rows
// For systems of 3-site molecules (atoms 3I, 3I+1, 3I+2) the CG row field runs on molecule pairs:
// record (I, J) holds the 9 atom pairs (p of I, k of J) as G1[3p+k], G2[3p+k] (zero beyond the
// cutoff and for p == k when J == I), so one lane fetches J's three positions and dipoles once for
// nine pair terms instead of one gather per pair.  Records of molecule I sit at
// mell(e, I) = (e / RLM) nmol RLM + I RLM + e % RLM for its RLM lanes.  The record's word is
// J | code << 26: code = the lattice shifts (n_a + 1 in two bits per axis) shared by every pair
// within the cutoff, or MOL_OWN_IMAGES when they differ (the pairs then take min_image each).
constexpr int MOL_OWN_IMAGES = 63;
constexpr int MOL_REC = 18;  // reals per record: G1 of the 9 pairs, then G2
PGM_DEV size_t mell(int e, int I, int nmol, int rlm) {
    return (size_t)(e / rlm) * ((size_t)nmol * rlm) + (size_t)I * rlm + (e % rlm);
}

// Candidate partner molecules of molecule I: the molecules of the atoms in the full (cutoff + skin)
// rows of I's three atoms, deduplicated with a bitmap over the bead's molecules in LDS (`words`
// words per wave) and written in increasing order.  One wave per molecule, no block barriers.
__global__ __launch_bounds__(256) void k_mol_rows(int nmol, int nmol_bead, const int *__restrict__ rows_full,
                                                  const int *__restrict__ count_full, int cap_full, int mcap,
                                                  int *__restrict__ mrows, int *__restrict__ mcount, int *__restrict__ maxcount) {
    extern __shared__ unsigned int mol_bits[];
    const int lane = threadIdx.x % WAVE, w = threadIdx.x / WAVE;
    const int I = blockIdx.x * (blockDim.x / WAVE) + w;
    if (I >= nmol) return;
    const int words = (nmol_bead + 31) / 32;
    unsigned int *bits = mol_bits + (size_t)w * words;
    for (int k = lane; k < words; k += WAVE) bits[k] = 0u;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    const int m0 = (I / nmol_bead) * nmol_bead;
    // the three rows in rounds of four entries per lane and atom, all twelve loads issued first
    int cnt3[3], cmax = 0;
#pragma unroll
    for (int p = 0; p < 3; ++p) {
        cnt3[p] = min(count_full[3 * I + p], cap_full);
        cmax = max(cmax, cnt3[p]);
    }
    for (int e0 = 0; e0 < cmax; e0 += 4 * WAVE) {
        int v[3][4];
#pragma unroll
        for (int p = 0; p < 3; ++p)
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int e = e0 + lane + u * WAVE;
                v[p][u] = e < cnt3[p] ? rows_full[(size_t)(3 * I + p) * cap_full + e] : -1;
            }
        // consecutive entries of a row are often the atoms of one molecule (and its neighbours in
        // index): a lane skips a molecule its left neighbour sets, which keeps most same-word LDS
        // atomics out of one instruction
#pragma unroll
        for (int p = 0; p < 3; ++p)
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int J = v[p][u] >= 0 ? v[p][u] / 3 - m0 : -1;
                const int Jl = __shfl_up(J, 1, WAVE);
                if (J >= 0 && (lane == 0 || Jl != J)) atomicOr(&bits[J >> 5], 1u << (J & 31));
            }
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();
    const int per = (words + WAVE - 1) / WAVE;
    const int w0 = min(words, lane * per), w1 = min(words, w0 + per);
    int cnt = 0;
    for (int k = w0; k < w1; ++k) cnt += __popc(bits[k]);
    int incl = cnt;
#pragma unroll
    for (int o = 1; o < WAVE; o <<= 1) {
        const int t = __shfl_up(incl, o, WAVE);
        if (lane >= o) incl += t;
    }
    int pos = incl - cnt;
    const int total = __shfl(incl, WAVE - 1, WAVE);
    int *mrow = mrows + (size_t)I * mcap;
    for (int k = w0; k < w1; ++k) {
        unsigned int b = bits[k];
        while (b) {
            const int t = __ffs(b) - 1;
            b &= b - 1u;
            if (pos < mcap) mrow[pos] = m0 + 32 * k + t;
            ++pos;
        }
    }
    if (lane == 0) {
        mcount[I] = total;
        if (total > __hip_atomic_load(maxcount, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) atomicMax(maxcount, total);
    }
}

// Records of the molecule blocks, once per geometry (the molecule form of k_compact_rows +
// k_pair_tensors): every candidate J of I with at least one atom pair inside the electrostatics
// cutoff (r^2 < rce2; pairs beyond it hold zeros, and rce2 = rc2 without a split cutoff) is kept, in candidate
// order, its G1 / G2 by pair_G from the FP64 minimum-image displacement rounded to real.  One wave
// per molecule, lanes over candidates, ballot compaction.  Divergence of pair_G's series branches is
// kept out of the common path: the closed form (both screenings with (a r)^2, (beta r)^2 >= 1) is
// evaluated for every pair at once, the other pairs are marked and then evaluated one per lane per
// round, and the own molecule's record (J == I, six close pairs) is computed by six lanes together.
//
// With ROWS the same pass also writes the compacted atom rows of the force pass (what
// k_compact_rows makes from the full atom rows): for atom 3I + p every 3J + k with r^2 < rc^2,
// the near pairs (r^2 < 1.05 * 2 (R_i^2 + R_j^2) or beta^2 r^2 < 1.05, FP64) from the front of the
// row and the others from the back, in (k, candidate) order per wave chunk; count / count_near
// as k_compact_rows gives them.  (Used for FP32; the FP64 variant keeps its registers for the
// tensors and leaves the atom rows to k_compact_rows.)
// FP32: four waves per SIMD (at most 128 VGPRs, no spills), 11 % faster than the three that the
// unbounded register allocation gives (319 against 360 us at 32k waters); five spill, and so does
// the FP64 variant at four.
template <typename real, bool ROWS>
__global__ __launch_bounds__(256, sizeof(real) == 4 ? 4 : 1) void k_mol_tensors(int nmol, const int *__restrict__ mrows, const int *__restrict__ mcount_full,
                                                     int mcap_full, int mcap, int rlm, const double4 *__restrict__ xr, Box box,
                                                     double rc2, double rce2, real beta, int *__restrict__ mj, real *__restrict__ mg,
                                                     int *__restrict__ mcount, int *__restrict__ rows, int acap,
                                                     int *__restrict__ acount, int *__restrict__ acount_near, double beta2) {
    const int lane = threadIdx.x % WAVE;
    const int I = blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE;
    if (I >= nmol) return;
    const double4 a[3] = {xr[3 * I], xr[3 * I + 1], xr[3 * I + 2]};
    const int cf = min(mcount_full[I], mcap_full);
    const int *cand = mrows + (size_t)I * mcap_full;
    int kept = 0, self_slot = -1;
    int on[3] = {0, 0, 0}, of[3] = {0, 0, 0};  // atom rows: near / far entries so far
    const unsigned long long below = (1ull << lane) - 1ull;
    for (int e0 = 0; e0 < cf; e0 += WAVE) {
        const int e = e0 + lane;
        bool keep = false, own = false;
        int J = 0, code = MOL_OWN_IMAGES;
        unsigned near = 0u, akeep = 0u, anear = 0u;
        real G1[9], G2[9];
        double4 b[3];
#pragma unroll
        for (int t = 0; t < 9; ++t) G1[t] = G2[t] = real(0);
#pragma unroll
        for (int k = 0; k < 3; ++k) b[k] = make_double4(0.0, 0.0, 0.0, 0.0);
        if (e < cf) {
            J = cand[e];
            own = J == I;
#pragma unroll
            for (int k = 0; k < 3; ++k) b[k] = xr[3 * J + k];
            int n0[3] = {0, 0, 0};
            bool first = true, same = true;
#pragma unroll
            for (int p = 0; p < 3; ++p)
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    if (own && p == k) continue;
                    double dx = a[p].x - b[k].x, dy = a[p].y - b[k].y, dz = a[p].z - b[k].z;
                    int n[3];
                    min_image_n(dx, dy, dz, box, n);
                    const double r2d = dx * dx + dy * dy + dz * dz;
                    if (r2d >= rc2) continue;
                    if (ROWS) {
                        akeep |= 1u << (3 * p + k);
                        if (r2d < 1.05 * 2.0 * (a[p].w + b[k].w) || beta2 * r2d < 1.05) anear |= 1u << (3 * p + k);
                    }
                    if (r2d >= rce2) continue;  // records hold the electrostatic pairs only
                    keep = true;
                    if (own) continue;  // the own molecule's pairs come after the compaction
                    if (first) {
                        n0[0] = n[0], n0[1] = n[1], n0[2] = n[2];
                        first = false;
                    } else if (n[0] != n0[0] || n[1] != n0[1] || n[2] != n0[2]) {
                        same = false;
                    }
                    const real x = (real)dx, y = (real)dy, z = (real)dz;
                    const real r2 = x * x + y * y + z * z;
                    const real aij = rsqrt_(real(2) * (real)(a[p].w + b[k].w));
                    if (aij * aij * r2 >= real(1) && beta * beta * r2 >= real(1)) {
                        real G[3];
                        pair_G_far<real, 3>(aij, beta, r2, G);
                        G1[3 * p + k] = G[1];
                        G2[3 * p + k] = G[2];
                    } else {
                        near |= 1u << (3 * p + k);
                    }
                }
            const bool small = n0[0] >= -1 && n0[0] <= 1 && n0[1] >= -1 && n0[1] <= 1 && n0[2] >= -1 && n0[2] <= 1;
            if (!own && same && small) code = (n0[0] + 1) | ((n0[1] + 1) << 2) | ((n0[2] + 1) << 4);
        }
        // pairs off the closed form: one per lane per round (the series branches run on useful lanes)
        while (__ballot(near != 0u)) {
            if (near) {
                const int t = __ffs(near) - 1;
                near &= near - 1u;
                const int p = t / 3, k = t - 3 * p;
                const double4 ap = xr[3 * I + p], bk = xr[3 * J + k];  // reloaded (cached) rather than selected
                double dx = ap.x - bk.x, dy = ap.y - bk.y, dz = ap.z - bk.z;
                min_image(dx, dy, dz, box);
                const real x = (real)dx, y = (real)dy, z = (real)dz;
                const real aij = rsqrt_(real(2) * (real)(ap.w + bk.w));
                real G[3];
                pair_G<real, 3>(aij, beta, x * x + y * y + z * z, G);
#pragma unroll
                for (int u = 0; u < 9; ++u)
                    if (u == t) {
                        G1[u] = G[1];
                        G2[u] = G[2];
                    }
            }
        }
        if (ROWS) {  // atom rows of the force pass
#pragma unroll
            for (int p = 0; p < 3; ++p) {
                int *rw = rows + (size_t)(3 * I + p) * acap;
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    const unsigned bit = 1u << (3 * p + k);
                    const bool kn = (akeep & bit) && (anear & bit), kf = (akeep & bit) && !(anear & bit);
                    const unsigned long long mn = __ballot(kn), mf = __ballot(kf);
                    const int pn = on[p] + __popcll(mn & below), pf = of[p] + __popcll(mf & below);
                    if (kn && pn + of[p] < acap) rw[pn] = 3 * J + k;  // the far part starts at acap - of
                    if (kf && on[p] + pf < acap) rw[acap - 1 - pf] = 3 * J + k;
                    on[p] += __popcll(mn);
                    of[p] += __popcll(mf);
                }
            }
        }
        const unsigned long long m = __ballot(keep);
        const int slot_e = kept + __popcll(m & below);
        if (keep && slot_e < mcap) {
            const size_t q = mell(slot_e, I, nmol, rlm);
            mj[q] = (int)((unsigned)J | ((unsigned)code << 26));
            real *g = mg + q * MOL_REC;
            if (own) {  // the diagonal (p == k) only; the six pairs follow
                g[0] = g[4] = g[8] = real(0);
                g[9] = g[13] = g[17] = real(0);
                self_slot = slot_e;
            } else {
#pragma unroll
                for (int t = 0; t < 9; ++t) {
                    g[t] = G1[t];
                    g[9 + t] = G2[t];
                }
            }
        }
        kept += __popcll(m);
    }
    // the own molecule's record: lanes 0..5 take the pairs (0,1) (0,2) (1,0) (1,2) (2,0) (2,1)
    const unsigned long long hs = __ballot(self_slot >= 0);
    if (hs) {
        const int sslot = __shfl(self_slot, (int)__ffsll((long long)hs) - 1, WAVE);
        if (lane < 6) {
            const int p = lane / 2, k = (lane % 2) + (lane % 2 >= p ? 1 : 0);
            const double4 ap = xr[3 * I + p], ak = xr[3 * I + k];
            double dx = ap.x - ak.x, dy = ap.y - ak.y, dz = ap.z - ak.z;
            min_image(dx, dy, dz, box);
            real G[3] = {real(0), real(0), real(0)};
            if (dx * dx + dy * dy + dz * dz < rce2) {
                const real x = (real)dx, y = (real)dy, z = (real)dz;
                const real aij = rsqrt_(real(2) * (real)(ap.w + ak.w));
                pair_G<real, 3>(aij, beta, x * x + y * y + z * z, G);
            }
            real *g = mg + mell(sslot, I, nmol, rlm) * MOL_REC;
            g[3 * p + k] = G[1];
            g[9 + 3 * p + k] = G[2];
        }
    }
    if (lane == 0) mcount[I] = kept < mcap ? kept : mcap;
    if (ROWS && lane < 3) {
        const int o = on[lane], f = of[lane];
        const int tot = o + f < acap ? o + f : acap;
        acount_near[3 * I + lane] = o < tot ? o : tot;
        acount[3 * I + lane] = tot;
    }
}

// Row field of the molecule blocks: out_i for the three atoms of I (the same sum as k_row_field over
// the pairs within rc, in molecule-pair order).  RLM lanes per molecule.
template <typename real, bool WITHQ, int RLM>
__global__ __launch_bounds__(256) void k_row_field_mol(int nmol, int nmol_bead, const int *__restrict__ mj,
                                                       const real *__restrict__ mg, const int *__restrict__ mcount,
                                                       const double4 *__restrict__ xr, const real4<real> *__restrict__ dq,
                                                       Box box, real4<real> *__restrict__ out, const int *__restrict__ conv_it,
                                                       int iter, Ext3 ext = {}) {
    const int lane = threadIdx.x % RLM;
    const int I = blockIdx.x * (blockDim.x / RLM) + threadIdx.x / RLM;
    if (I >= nmol || cg_done(conv_it, I / nmol_bead, iter)) return;
    const double4 a[3] = {xr[3 * I], xr[3 * I + 1], xr[3 * I + 2]};
    real acc[9];
#pragma unroll
    for (int t = 0; t < 9; ++t) acc[t] = real(0);
    const int c = mcount[I];
    for (int e = lane; e < c; e += RLM) {
        const size_t q = mell(e, I, nmol, RLM);
        const unsigned word = (unsigned)mj[q];
        const int J = (int)(word & ((1u << 26) - 1u)), code = (int)(word >> 26);
        real G[MOL_REC];
        if (sizeof(real) == 4) {
            const float2 *g2 = reinterpret_cast<const float2 *>(mg + q * MOL_REC);
#pragma unroll
            for (int t = 0; t < MOL_REC / 2; ++t) {
                const float2 v = g2[t];
                G[2 * t] = (real)v.x;
                G[2 * t + 1] = (real)v.y;
            }
        } else {
            const double2 *g2 = reinterpret_cast<const double2 *>(mg + q * MOL_REC);
#pragma unroll
            for (int t = 0; t < MOL_REC / 2; ++t) {
                const double2 v = g2[t];
                G[2 * t] = (real)v.x;
                G[2 * t + 1] = (real)v.y;
            }
        }
        double4 b[3];
        real4<real> d[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            b[k] = xr[3 * J + k];
            d[k] = dq[3 * J + k];
        }
        const int n[3] = {(code & 3) - 1, ((code >> 2) & 3) - 1, ((code >> 4) & 3) - 1};
#pragma unroll
        for (int p = 0; p < 3; ++p)
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                double dx = a[p].x - b[k].x, dy = a[p].y - b[k].y, dz = a[p].z - b[k].z;
                if (code == MOL_OWN_IMAGES)
                    min_image(dx, dy, dz, box);
                else
                    apply_shift(dx, dy, dz, box, n);
                const real x = (real)dx, y = (real)dy, z = (real)dz;
                const real g1 = G[3 * p + k], g2 = G[9 + 3 * p + k];
                const real dkx = d[k].x * x + d[k].y * y + d[k].z * z;
                real cf = -g2 * dkx;
                if (WITHQ) cf -= d[k].w * g1;
                acc[3 * p] += cf * x + g1 * d[k].x;
                acc[3 * p + 1] += cf * y + g1 * d[k].y;
                acc[3 * p + 2] += cf * z + g1 * d[k].z;
            }
    }
#pragma unroll
    for (int t = 0; t < 9; ++t) acc[t] = group_sum<RLM>(acc[t]);
    if (lane == 0)
#pragma unroll
        for (int p = 0; p < 3; ++p) {
            real4<real> v;
            v.x = acc[3 * p];
            v.y = acc[3 * p + 1];
            v.z = acc[3 * p + 2];
            v.w = 0;
            if (WITHQ) {  // a uniform external field: its dU/dd is -ext
                v.x -= (real)ext.x;
                v.y -= (real)ext.y;
                v.z -= (real)ext.z;
            }
            out[3 * I + p] = v;
        }
}

// ---------------------------------------------------------------------------- molecule-blocked force pass
// The force pass of k_pair_forces (the same per-atom sums g_el, g_lj, rowf and per-block energy
// partials, each pair again counted from both sides) on the molecule-block records, for 3-site
// molecules when the van der Waals cutoff does not exceed the electrostatic one (every LJ pair then
// lies in a record) and no strain derivative is asked for.  FL lanes per atom i = 3I + p walk the
// records of its molecule I and take the three pairs (p, k) of each: J's positions, dipoles, charges
// and LJ parameters are the same loads for the three atoms of I (one cache line for the lanes of
// a wave), and the record's G1, G2 are reused.  On the closed-form branch ((a r)^2, (b r)^2 >= 1)
// the other two radial functions follow from them and the two Gaussians c_a = 2a/sqrt(pi)
// exp(-a^2 r^2), c_b likewise: G0 = r^2 G1 + (c_a - c_b) and G3 = (5 G2 - ((2a^2)^2 c_a -
// (2b^2)^2 c_b)) / r^2, the recursion B_n = ((2n-1) B_(n-1) - (2a^2)^(n-1) c) / r^2 of the closed
// form read down and up (no erfc).  Pairs off it (a few per atom: its own molecule, hydrogen bonds)
// are queued per lane in LDS and take pair_G (series) after the record loop, one per lane per round,
// so that the wave runs the series branch a few times per atom and not in most record rounds.  The
// terms of a record are summed in `real`, then added to FP64 sums.  Grid (blocks per bead, beads).
#define MOL_FORCE_TERMS(G0_, G1_, G2_, G3_)                                                            \
    {                                                                                                  \
        const real dix = di.x * x + di.y * y + di.z * z;                                               \
        const real dkx = dk.x * x + dk.y * y + dk.z * z;                                               \
        const real didk = di.x * dk.x + di.y * dk.y + di.z * dk.z;                                     \
        const real qk = dk.w;                                                                          \
        const real t = qi * dkx - qk * dix;                                                            \
        pe += qi * qk * (G0_) + t * (G1_) - (G2_) * dix * dkx + (G1_) * didk;                          \
        const real radial = -qi * qk * (G1_) - t * (G2_) + (G3_) * dix * dkx - (G2_) * didk;           \
        const real qiG1 = qi * (G1_), qkG1 = qk * (G1_);                                               \
        pg0 += radial * x + qiG1 * dk.x - qkG1 * di.x - (G2_) * (di.x * dkx + dk.x * dix);             \
        pg1 += radial * y + qiG1 * dk.y - qkG1 * di.y - (G2_) * (di.y * dkx + dk.y * dix);             \
        pg2 += radial * z + qiG1 * dk.z - qkG1 * di.z - (G2_) * (di.z * dkx + dk.z * dix);             \
        const real cfield = -qk * (G1_) - (G2_) * dkx;                                                 \
        pf0 += cfield * x + (G1_) * dk.x;                                                              \
        pf1 += cfield * y + (G1_) * dk.y;                                                              \
        pf2 += cfield * z + (G1_) * dk.z;                                                              \
    }

constexpr int MOL_FORCE_LANES = 8;
constexpr int MOL_NEAR_QUEUE = 8;  // queued series pairs per lane (more are evaluated at once)
// Launched with 256 threads per block (the LDS queue has 256 columns).  FP32: at most 128 VGPRs
// (four waves per SIMD); FP64 unbounded.
template <typename real, int FL>
__global__ __launch_bounds__(256, sizeof(real) == 4 ? 4 : 1) void k_pair_forces_mol(int nper, int nmol, int rlm, const int *__restrict__ mj,
                                                         const real *__restrict__ mg, const int *__restrict__ mcount,
                                                         const double4 *__restrict__ xr, const real4<real> *__restrict__ dq,
                                                         Box box, real beta, double rce2, double rcv2,
                                                         const double *__restrict__ rh, const double *__restrict__ se,
                                                         double *__restrict__ g_el, double *__restrict__ g_lj,
                                                         double *__restrict__ rowf, double *__restrict__ epart, Ext3 ext = {}) {
    __shared__ double lds[2 * 16];
    __shared__ int nq[MOL_NEAR_QUEUE * 256];  // (record index << 2 | k) of the lane's series pairs
    const int lane = threadIdx.x % FL;
    const int li = blockIdx.x * (blockDim.x / FL) + threadIdx.x / FL;
    const int i = blockIdx.y * nper + li;
    const int pb = blockIdx.y * gridDim.x + blockIdx.x;
    int *queue = nq + threadIdx.x;  // entry u at queue[256 u]
    double esum[2] = {0.0, 0.0};
    if (li < nper) {
        const int I = i / 3, p = i - 3 * I;
        const double4 a = xr[i];
        const real4<real> di = dq[i];
        const real qi = di.w;
        const double rhi = rh[i], sei = se[i];
        const real fb = real(2) * beta * beta;
        double gx = 0, gy = 0, gz = 0, lx = 0, ly = 0, lz = 0, fx = 0, fy = 0, fz = 0, ee = 0, el = 0;
        const int c = mcount[I];
        int qlen = 0;
        // the pair (p, k) of record e on the series branch
        auto near_pair = [&](int e, int k, real &pe, real &pg0, real &pg1, real &pg2, real &pf0, real &pf1, real &pf2) {
            const unsigned word = (unsigned)mj[mell(e, I, nmol, rlm)];
            const int J = (int)(word & ((1u << 26) - 1u)), code = (int)(word >> 26);
            const int n[3] = {(code & 3) - 1, ((code >> 2) & 3) - 1, ((code >> 4) & 3) - 1};
            const double4 b = xr[3 * J + k];
            const real4<real> dk = dq[3 * J + k];
            double dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
            if (code == MOL_OWN_IMAGES)
                min_image(dx, dy, dz, box);
            else
                apply_shift(dx, dy, dz, box, n);
            const real x = (real)dx, y = (real)dy, z = (real)dz;
            const real aij = rsqrt_(real(2) * (real)(a.w + b.w));
            real G[4];
            pair_G<real, 4>(aij, beta, x * x + y * y + z * z, G);
            MOL_FORCE_TERMS(G[0], G[1], G[2], G[3])
        };
        for (int e = lane; e < c; e += FL) {
            const size_t q = mell(e, I, nmol, rlm);
            const unsigned word = (unsigned)mj[q];
            const int J = (int)(word & ((1u << 26) - 1u)), code = (int)(word >> 26);
            const bool own = J == I;
            const real *gr = mg + q * MOL_REC + 3 * p;
            const real G1[3] = {gr[0], gr[1], gr[2]}, G2[3] = {gr[9], gr[10], gr[11]};
            const int n[3] = {(code & 3) - 1, ((code >> 2) & 3) - 1, ((code >> 4) & 3) - 1};
            real pg0 = 0, pg1 = 0, pg2 = 0, pf0 = 0, pf1 = 0, pf2 = 0, pe = 0;
            unsigned nearm = 0u;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                if (own && k == p) continue;
                const double4 b = xr[3 * J + k];
                const real4<real> dk = dq[3 * J + k];
                double dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
                if (code == MOL_OWN_IMAGES)
                    min_image(dx, dy, dz, box);
                else
                    apply_shift(dx, dy, dz, box, n);
                const double r2d = dx * dx + dy * dy + dz * dz;
                const double eps = sei * se[3 * J + k];
                if (!own && eps != 0.0 && r2d < rcv2) {  // LJ, FP64 as in k_pair_forces
                    const double rmin = rhi + rh[3 * J + k];
                    const double ir2 = 1.0 / r2d;
                    const double s2 = rmin * rmin * ir2;
                    const double s6 = s2 * s2 * s2;
                    el += eps * (s6 * s6 - 2.0 * s6);
                    const double dl = eps * 12.0 * (s6 - s6 * s6) * ir2;
                    lx += dl * dx;
                    ly += dl * dy;
                    lz += dl * dz;
                }
                if (r2d >= rce2) continue;
                const real x = (real)dx, y = (real)dy, z = (real)dz;
                const real r2 = x * x + y * y + z * z;
                const real aij = rsqrt_(real(2) * (real)(a.w + b.w));
                const real sa = aij * aij * r2, sb = beta * beta * r2;
                if (sa >= real(1) && sb >= real(1)) {
                    const real ir = rsqrt_(r2), ir2 = ir * ir;
                    const real ca = real(TWO_OVER_SQRTPI) * aij * exp_(-sa);
                    const real cb = real(TWO_OVER_SQRTPI) * beta * exp_(-sb);
                    const real fa = real(2) * aij * aij;
                    const real g1 = G1[k], g2 = G2[k];
                    const real g0 = r2 * g1 + (ca - cb);
                    const real g3 = (real(5) * g2 - (ca * fa * fa - cb * fb * fb)) * ir2;
                    MOL_FORCE_TERMS(g0, g1, g2, g3)
                } else {
                    nearm |= 1u << k;
                }
            }
            while (nearm) {  // queue the series pairs (evaluate at once when the queue is full)
                const int k = __ffs(nearm) - 1;
                nearm &= nearm - 1u;
                if (qlen < MOL_NEAR_QUEUE)
                    queue[256 * qlen++] = (e << 2) | k;
                else
                    near_pair(e, k, pe, pg0, pg1, pg2, pf0, pf1, pf2);
            }
            gx += (double)pg0;
            gy += (double)pg1;
            gz += (double)pg2;
            fx += (double)pf0;
            fy += (double)pf1;
            fz += (double)pf2;
            ee += (double)pe;
        }
        for (int u = 0; __ballot(u < qlen); ++u) {  // the queued series pairs, one per lane per round
            if (u < qlen) {
                const int v = queue[256 * u];
                real pg0 = 0, pg1 = 0, pg2 = 0, pf0 = 0, pf1 = 0, pf2 = 0, pe = 0;
                near_pair(v >> 2, v & 3, pe, pg0, pg1, pg2, pf0, pf1, pf2);
                gx += (double)pg0;
                gy += (double)pg1;
                gz += (double)pg2;
                fx += (double)pf0;
                fy += (double)pf1;
                fz += (double)pf2;
                ee += (double)pe;
            }
        }
        gx = group_sum<FL>(gx);
        gy = group_sum<FL>(gy);
        gz = group_sum<FL>(gz);
        lx = group_sum<FL>(lx);
        ly = group_sum<FL>(ly);
        lz = group_sum<FL>(lz);
        fx = group_sum<FL>(fx);
        fy = group_sum<FL>(fy);
        fz = group_sum<FL>(fz);
        if (lane == 0) {
            g_el[3 * i] = gx;
            g_el[3 * i + 1] = gy;
            g_el[3 * i + 2] = gz;
            g_lj[3 * i] = lx;
            g_lj[3 * i + 1] = ly;
            g_lj[3 * i + 2] = lz;
            rowf[3 * i] = fx - ext.x;  // a uniform external field: its dU/dd is -ext
            rowf[3 * i + 1] = fy - ext.y;
            rowf[3 * i + 2] = fz - ext.z;
        }
        esum[0] = ee;
        esum[1] = el;
    }
    block_sum<2>(esum, lds);
    if (threadIdx.x == 0) {
        epart[2 * pb] = esum[0];
        epart[2 * pb + 1] = esum[1];
    }
}
#undef MOL_FORCE_TERMS

}  // namespace pgm
