// fft.h -- the PME reciprocal-space convolution of one field (or energy) evaluation as three fused kernels.
//
// phi = c2r(Gc * r2c(Q)) over a K1 x K2 x K3 grid (z = K3 fastest; half spectrum of K3h = K3/2 + 1
// points along z, the layout [x][y][kz] of the library's FFT plans), pme.py:373-376, with each K a
// power of two (8..256):
//   k_pme_fwd_plane  one workgroup per x-plane: the spread-block merge of the plane (the sums of
//                    k_merge_table), the real transform along z (a K3/2-point complex transform of
//                    (x_2m, x_2m+1) and the split into the half spectrum), the transform along y;
//   k_pme_x_conv     per 16 (y, kz) columns: the transform along x, the product with Gc, the
//                    inverse transform along x (with ENERGY also the sum of w Gc |F Q|^2);
//   k_pme_bwd_plane  one workgroup per x-plane: the inverse along y, the merge of the half
//                    spectrum into a K3/2-point complex sequence, the inverse along z -> real plane.
// Every transform runs in LDS as radix-16/8/4/2 Stockham stages (autosort, in place: a stage reads all
// its inputs into registers, then writes) with twiddles exp(-2 pi i t / N) tabulated in FP64 on the
// host.  Unnormalised in both directions, like the library plans it replaces.
#pragma once

#include "common.h"

namespace pgm {

template <typename real> struct vec2;
template <> struct vec2<float> { using type = float2; };
template <> struct vec2<double> { using type = double2; };
template <typename real> using cplx = typename vec2<real>::type;

template <typename C> PGM_DEV C cmk(decltype(C{}.x) a, decltype(C{}.x) b) {
    C r;
    r.x = a;
    r.y = b;
    return r;
}
template <typename C> PGM_DEV C cadd(C a, C b) { return cmk<C>(a.x + b.x, a.y + b.y); }
template <typename C> PGM_DEV C csub(C a, C b) { return cmk<C>(a.x - b.x, a.y - b.y); }
template <typename C> PGM_DEV C cmul(C a, C b) { return cmk<C>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
template <typename C> PGM_DEV C cconj(C a) { return cmk<C>(a.x, -a.y); }
// a * (-i) forward, a * (+i) inverse
template <bool INV, typename C> PGM_DEV C cmi(C a) { return INV ? cmk<C>(-a.y, a.x) : cmk<C>(a.y, -a.x); }

// Radix-R DFTs in registers, y_m' = sum_m v_m exp(-+2 pi i m m' / R).
template <bool INV, typename C> PGM_DEV void dft2(C *v) {
    const C a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
}
template <bool INV, typename C> PGM_DEV void dft4(C &v0, C &v1, C &v2, C &v3) {
    const C a0 = cadd(v0, v2), a1 = csub(v0, v2), a2 = cadd(v1, v3), a3 = cmi<INV>(csub(v1, v3));
    v0 = cadd(a0, a2);
    v2 = csub(a0, a2);
    v1 = cadd(a1, a3);
    v3 = csub(a1, a3);
}
template <bool INV, typename C> PGM_DEV void dft8(C *v) {
    using R = decltype(C{}.x);
    C e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6], o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
    dft4<INV>(e0, e1, e2, e3);
    dft4<INV>(o0, o1, o2, o3);
    const R h = R(0.70710678118654752440);
    // W8^1 o1, W8^2 o2, W8^3 o3 with W8 = exp(-+ i pi / 4)
    const C w1 = INV ? cmk<C>(h * (o1.x - o1.y), h * (o1.x + o1.y)) : cmk<C>(h * (o1.x + o1.y), h * (o1.y - o1.x));
    const C w2 = cmi<INV>(o2);
    const C w3 = INV ? cmk<C>(-h * (o3.x + o3.y), h * (o3.x - o3.y)) : cmk<C>(h * (o3.y - o3.x), -h * (o3.x + o3.y));
    v[0] = cadd(e0, o0);
    v[4] = csub(e0, o0);
    v[1] = cadd(e1, w1);
    v[5] = csub(e1, w1);
    v[2] = cadd(e2, w2);
    v[6] = csub(e2, w2);
    v[3] = cadd(e3, w3);
    v[7] = csub(e3, w3);
}
// 16 = 4 x 4: A[m0][k1] = DFT4_m1(v[4 m1 + m0]), times W16^(m0 k1), then DFT4 over m0 gives
// y[k1 + 4 k2].
template <bool INV, typename C> PGM_DEV void dft16(C *v) {
    using R = decltype(C{}.x);
#pragma unroll
    for (int m0 = 0; m0 < 4; ++m0) dft4<INV>(v[m0], v[4 + m0], v[8 + m0], v[12 + m0]);
    // A[m0][k1] sits at v[4 k1 + m0]; twiddles W16^(m0 k1), m0 k1 in {1, 2, 3, 4, 6, 9}
    const R c1 = R(0.92387953251128675613), s1 = R(0.38268343236508977173), h = R(0.70710678118654752440);
    const R sg = INV ? R(1) : R(-1);  // W16^t = cos(2 pi t / 16) + sg i sin(2 pi t / 16)
    auto tw = [&](C a, R c, R sn) { return cmk<C>(a.x * c - a.y * sg * sn, a.x * sg * sn + a.y * c); };
    v[4 * 1 + 1] = tw(v[5], c1, s1);    // t = 1
    v[4 * 1 + 2] = tw(v[6], h, h);      // t = 2
    v[4 * 1 + 3] = tw(v[7], s1, c1);    // t = 3
    v[4 * 2 + 1] = tw(v[9], h, h);      // t = 2
    v[4 * 2 + 2] = cmi<INV>(v[10]);     // t = 4
    v[4 * 2 + 3] = tw(v[11], -h, h);    // t = 6
    v[4 * 3 + 1] = tw(v[13], s1, c1);   // t = 3
    v[4 * 3 + 2] = tw(v[14], -h, h);    // t = 6
    v[4 * 3 + 3] = tw(v[15], -c1, -s1); // t = 9
    C y[16];
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) {
        C a0 = v[4 * k1], a1 = v[4 * k1 + 1], a2 = v[4 * k1 + 2], a3 = v[4 * k1 + 3];
        dft4<INV>(a0, a1, a2, a3);
        y[k1] = a0;
        y[k1 + 4] = a1;
        y[k1 + 8] = a2;
        y[k1 + 12] = a3;
    }
#pragma unroll
    for (int k = 0; k < 16; ++k) v[k] = y[k];
}
template <int R, bool INV, typename C> PGM_DEV void dft(C *v) {
    if constexpr (R == 16) dft16<INV>(v);
    else if constexpr (R == 8) dft8<INV>(v);
    else if constexpr (R == 4) dft4<INV>(v[0], v[1], v[2], v[3]);
    else dft2<INV>(v);
}

// Radix of the next Stockham stage for a remaining length n (a power of two): 16 for n = 16 and
// n >= 128, else 8 while it divides n, then 4, then 2 (128 = 16 x 8, 256 = 16 x 16, 64 = 8 x 8,
// 32 = 8 x 4).
PGM_DEV int fft_radix(int n) {
    if (n == 16 || (n >= 128 && n % 16 == 0)) return 16;
    return n % 8 == 0 ? 8 : (n % 4 == 0 ? 4 : 2);
}

// One Stockham stage (radix R, sub-transform length L -> L R) of nlines length-N transforms held in
// LDS: element e of line l at buf[l * ls + e * es].  Butterfly j of a line (N/R of them) reads
// elements j + m N/R, twiddles them by W_N^(m (j mod L) N/(L R)), transforms and writes element
// (j / L) L R + (j mod L) + m' L.  Lines are handled in rounds of whole lines, each thread holding
// at most IT butterflies, so a round's writes never meet another round's reads.  Both barriers
// are reached by every thread of the workgroup.
template <int R, bool INV, int IT, typename C>
PGM_DEV void fft_stage(C *buf, int N, int L, int nlines, int es, int ls, const C *__restrict__ tw) {
    const int tid = threadIdx.x, nth = blockDim.x;
    const int nb = N / R;
    const int lpr = max(1, (IT * nth) / nb);
    const int tws = N / (L * R);
    const bool contiguous = es == 1;
    for (int l0 = 0; l0 < nlines; l0 += lpr) {
        const int nl = min(lpr, nlines - l0);
        const int items = nl * nb;
        C v[IT][R];
        int base[IT], jj[IT];
#pragma unroll
        for (int it = 0; it < IT; ++it) {
            const int item = tid + it * nth;
            base[it] = -1;
            if (item < items) {
                const int line = contiguous ? item / nb : item % nl;
                const int j = contiguous ? item - line * nb : item / nl;
                const int k = j & (L - 1);
                base[it] = (l0 + line) * ls;
                jj[it] = j;
#pragma unroll
                for (int m = 0; m < R; ++m) {
                    C x = buf[base[it] + (j + m * nb) * es];
                    if (m > 0 && k > 0) {
                        const C w = tw[m * k * tws];
                        x = cmul(x, INV ? cconj(w) : w);
                    }
                    v[it][m] = x;
                }
                dft<R, INV>(v[it]);
            }
        }
        __syncthreads();
#pragma unroll
        for (int it = 0; it < IT; ++it) {
            if (base[it] >= 0) {
                const int j = jj[it], k = j & (L - 1);
                const int o = (j - k) * R + k;
#pragma unroll
                for (int m = 0; m < R; ++m) buf[base[it] + (o + m * L) * es] = v[it][m];
            }
        }
        __syncthreads();
    }
}

// Full length-N transform (forward: exp(-2 pi i ...), inverse: exp(+...), unnormalised) of nlines
// lines in LDS; tw = W_N^t for t < N.
template <bool INV, typename C> PGM_DEV void fft_lds(C *buf, int N, int nlines, int es, int ls, const C *__restrict__ tw) {
    for (int L = 1; L < N;) {
        const int R = fft_radix(N / L);
        if (R == 16) fft_stage<16, INV, 1>(buf, N, L, nlines, es, ls, tw);
        else if (R == 8) fft_stage<8, INV, 2>(buf, N, L, nlines, es, ls, tw);
        else if (R == 4) fft_stage<4, INV, 2>(buf, N, L, nlines, es, ls, tw);
        else fft_stage<2, INV, 4>(buf, N, L, nlines, es, ls, tw);
        L *= R;
    }
}

// Offsets (in complex entries) of the twiddle tables W_N^t, t < N, for N = K3/2, K3, K2, K1.
struct PFFTGeom {
    int K1, K2, K3, M, K3h;
    int twM, twK3, twK2, twK1;
};

// Threads of a plane workgroup at most: 1024 (FP32), 512 (FP64: twice the registers per value).
template <typename real> constexpr int pfft_plane_threads() { return sizeof(real) == 4 ? 1024 : 512; }

// Global -> LDS copy of n values, four loads in flight per thread before the stores.
template <typename T> PGM_DEV void copy_in(T *__restrict__ dst, const T *__restrict__ src, int n) {
    const int nth = blockDim.x;
    int p = threadIdx.x;
    for (; p + 3 * nth < n; p += 4 * nth) {
        const T a = src[p], b = src[p + nth], c = src[p + 2 * nth], d = src[p + 3 * nth];
        dst[p] = a;
        dst[p + nth] = b;
        dst[p + 2 * nth] = c;
        dst[p + 3 * nth] = d;
    }
    for (; p < n; p += nth) dst[p] = src[p];
}

// LDS bytes of a plane workgroup: the plane and the twiddles of lengths K3/2, K3, K2.
template <typename real> inline size_t pfft_plane_lds(const PFFTGeom &f) {
    return sizeof(cplx<real>) * ((size_t)f.K2 * f.K3h + f.M + f.K3 + f.K2);
}

// Forward plane: merge the spread blocks of plane x (bead b), r2c along z, transform along y, and
// store the plane's K2 K3h complex values at spec[(b K1 + x) K2 K3h ...].  Launch: K1 B blocks,
// pfft_plane_lds bytes of dynamic LDS.  The twiddle tables of lengths K3/2, K3, K2 lie one after
// the other from tw (offsets twM = 0, twK3 = K3/2, twK2 = K3/2 + K3).
template <typename real>
__global__ __launch_bounds__(pfft_plane_threads<real>()) void k_pme_fwd_plane(PFFTGeom f, size_t bead_blocks, const int *__restrict__ tab,
                                                        const real *__restrict__ blocks, cplx<real> *__restrict__ spec,
                                                        const cplx<real> *__restrict__ tw, const int *__restrict__ conv_it,
                                                        int iter) {
    using C = cplx<real>;
    const int gx = blockIdx.x % f.K1, bead = blockIdx.x / f.K1;
    if (cg_done(conv_it, bead, iter)) return;
    extern __shared__ __align__(16) unsigned char lds_raw[];
    C *pl = reinterpret_cast<C *>(lds_raw);
    real *plr = reinterpret_cast<real *>(lds_raw);
    const int K1 = f.K1, K2 = f.K2, K3 = f.K3, M = f.M, K3h = f.K3h;
    C *tws = pl + K2 * K3h;
    copy_in(tws, tw, M + K3 + K2);
    const int *tx = tab, *ty = tab + 2 * K1, *tz = tab + 2 * (K1 + K2);
    const int X0 = tx[2 * gx], X1 = tx[2 * gx + 1];
    const real *b = blocks + bead_blocks * bead;
    const int lg3 = __ffs(K3) - 1;
    const int nth = blockDim.x, npt = K2 * K3;
    // 1. merge (fixed order per point: the table candidates x, y, z in turn), four points per
    //    thread per pass with all their loads issued first; real line y at plr[2 y K3h ...]
    for (int p0 = threadIdx.x; p0 < npt; p0 += 4 * nth) {
        real v[4][8];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int p = p0 + u * nth;
            const int y = (p < npt ? p : 0) >> lg3, z = p & (K3 - 1);
            const int X[2] = {X0, X1}, Y[2] = {ty[2 * y], ty[2 * y + 1]}, Z[2] = {tz[2 * z], tz[2 * z + 1]};
#pragma unroll
            for (int i0 = 0; i0 < 2; ++i0)
#pragma unroll
                for (int i1 = 0; i1 < 2; ++i1)
#pragma unroll
                    for (int i2 = 0; i2 < 2; ++i2) {
                        const bool ok = p < npt && X[i0] >= 0 && Y[i1] >= 0 && Z[i2] >= 0;
                        v[u][4 * i0 + 2 * i1 + i2] = ok ? b[(size_t)X[i0] + Y[i1] + Z[i2]] : real(0);
                    }
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int p = p0 + u * nth;
            if (p < npt) {
                real acc = real(0);
#pragma unroll
                for (int c = 0; c < 8; ++c) acc += v[u][c];
                plr[2 * (p >> lg3) * K3h + (p & (K3 - 1))] = acc;
            }
        }
    }
    __syncthreads();
    // 2. M-point transforms of z_m = x_2m + i x_2m+1 along every line
    fft_lds<false>(pl, M, K2, 1, K3h, tws);
    // 3. half spectrum: X_k = E_k + W_K3^k O_k, E_k = (Z_k + conj Z_M-k) / 2, O_k = (Z_k - conj Z_M-k) / 2i
    const C *wk = tws + M;
    const int half = M / 2 + 1;
    for (int p = threadIdx.x; p < K2 * half; p += nth) {
        const int y = p / half, k = p - y * half;
        C *ln = pl + y * K3h;
        if (k == 0) {
            const C z0 = ln[0];
            ln[0] = cmk<C>(z0.x + z0.y, real(0));
            ln[M] = cmk<C>(z0.x - z0.y, real(0));
        } else {
            const int kk = M - k;
            const C a = ln[k], c = ln[kk];
            const C w = wk[k];
            // k
            C e = cmk<C>(real(0.5) * (a.x + c.x), real(0.5) * (a.y - c.y));
            C o = cmk<C>(real(0.5) * (a.y + c.y), real(-0.5) * (a.x - c.x));
            const C xk = cadd(e, cmul(w, o));
            // M - k: W_K3^(M-k) = -conj(W_K3^k)
            e = cmk<C>(real(0.5) * (c.x + a.x), real(0.5) * (c.y - a.y));
            o = cmk<C>(real(0.5) * (c.y + a.y), real(-0.5) * (c.x - a.x));
            const C xkk = cadd(e, cmul(cmk<C>(-w.x, w.y), o));
            ln[k] = xk;
            if (kk != k) ln[kk] = xkk;
        }
    }
    __syncthreads();
    // 4. K2-point transforms along y (columns kz: stride K3h)
    fft_lds<false>(pl, K2, K3h, K3h, 1, tws + M + K3);
    // 5. store the plane
    C *out = spec + ((size_t)bead * K1 + gx) * (size_t)(K2 * K3h);
    for (int p = threadIdx.x; p < K2 * K3h; p += nth) out[p] = pl[p];
}

constexpr int PFFT_COLS = 16;  // (y, kz) columns per workgroup of k_pme_x_conv

// LDS bytes of k_pme_x_conv: 16 columns of K1 values and the K1 twiddles.
template <typename real> inline size_t pfft_xconv_lds(const PFFTGeom &f) {
    return sizeof(cplx<real>) * ((size_t)f.K1 * PFFT_COLS + f.K1);
}

// Along x: columns c = y K3h + kz in [c0, c0 + 16) of bead b: transform, multiply by Gc, inverse.
// With ENERGY also the FP64 sum of w Gc |F Q|^2 over the block's columns (w = 1/2 on the kz = 0 and
// Nyquist planes of the half spectrum, 1 elsewhere: the sum of k_convolve's ENERGY branch) to
// epart[b * gridDim.x + blockIdx.x].  Launch: (ceil(K2 K3h / 16), B) blocks of 256 threads,
// pfft_xconv_lds bytes of dynamic LDS; twx = the K1 twiddles.
template <typename real, bool ENERGY = false>
__global__ __launch_bounds__(256) void k_pme_x_conv(PFFTGeom f, cplx<real> *__restrict__ spec, const real *__restrict__ Gc,
                                                    const cplx<real> *__restrict__ twx, const int *__restrict__ conv_it, int iter,
                                                    double *__restrict__ epart = nullptr) {
    using C = cplx<real>;
    const int bead = blockIdx.y;
    if (cg_done(conv_it, bead, iter)) return;
    __shared__ double red[16];
    extern __shared__ __align__(16) unsigned char lds_raw[];
    C *col = reinterpret_cast<C *>(lds_raw);
    const int K1 = f.K1;
    C *tws = col + K1 * PFFT_COLS;
    copy_in(tws, twx, K1);
    const int ncol = f.K2 * f.K3h;
    const int c0 = blockIdx.x * PFFT_COLS;
    const int nc = min(PFFT_COLS, ncol - c0);
    C *s = spec + (size_t)bead * K1 * ncol;
    const int nth = blockDim.x, nv = K1 * PFFT_COLS;
    for (int p0 = threadIdx.x; p0 < nv; p0 += 4 * nth) {
        C v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int p = p0 + u * nth, x = p / PFFT_COLS, cc = p - x * PFFT_COLS;
            v[u] = (p < nv && cc < nc) ? s[(size_t)x * ncol + c0 + cc] : cmk<C>(real(0), real(0));
        }
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (p0 + u * nth < nv) col[p0 + u * nth] = v[u];
    }
    __syncthreads();
    fft_lds<false>(col, K1, nc, PFFT_COLS, 1, tws);
    double e[1] = {0.0};
    for (int p0 = threadIdx.x; p0 < nv; p0 += 4 * nth) {
        real g[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int p = p0 + u * nth, x = p / PFFT_COLS, cc = p - x * PFFT_COLS;
            g[u] = (p < nv && cc < nc) ? Gc[(size_t)x * ncol + c0 + cc] : real(0);
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int p = p0 + u * nth;
            if (p < nv) {
                const C v = col[p];
                if (ENERGY) {  // g = 0 on the padding columns
                    const int kz = (c0 + (p % PFFT_COLS)) % f.K3h;
                    const double w = (kz == 0 || (f.K3 % 2 == 0 && kz == f.K3h - 1)) ? 0.5 : 1.0;
                    e[0] += w * (double)g[u] * ((double)v.x * v.x + (double)v.y * v.y);
                }
                col[p] = cmk<C>(v.x * g[u], v.y * g[u]);
            }
        }
    }
    if (ENERGY) {
        block_sum<1>(e, red);
        if (threadIdx.x == 0) epart[(size_t)bead * gridDim.x + blockIdx.x] = e[0];
    }
    __syncthreads();
    fft_lds<true>(col, K1, nc, PFFT_COLS, 1, tws);
    for (int p = threadIdx.x; p < nv; p += nth) {
        const int x = p / PFFT_COLS, cc = p - x * PFFT_COLS;
        if (cc < nc) s[(size_t)x * ncol + c0 + cc] = col[p];
    }
}

// Inverse plane: inverse along y, then the real inverse along z: Z_k = A_k + i B_k with
// A_k = X_k + conj X_M-k, B_k = (X_k - conj X_M-k) W_K3^-k (imaginary parts of X_0 and X_M
// dropped), the inverse M-point transform gives (x_2m, x_2m+1); the real plane goes to
// grid[(b K1 + x) K2 K3 ...].  Launch: K1 B blocks, pfft_plane_lds bytes of dynamic LDS, tw as
// for the forward plane.
template <typename real>
__global__ __launch_bounds__(pfft_plane_threads<real>()) void k_pme_bwd_plane(PFFTGeom f, const cplx<real> *__restrict__ spec, real *__restrict__ grid,
                                                        const cplx<real> *__restrict__ tw, const int *__restrict__ conv_it,
                                                        int iter) {
    using C = cplx<real>;
    const int gx = blockIdx.x % f.K1, bead = blockIdx.x / f.K1;
    if (cg_done(conv_it, bead, iter)) return;
    extern __shared__ __align__(16) unsigned char lds_raw[];
    C *pl = reinterpret_cast<C *>(lds_raw);
    const real *plr = reinterpret_cast<const real *>(lds_raw);
    const int K1 = f.K1, K2 = f.K2, K3 = f.K3, M = f.M, K3h = f.K3h;
    C *tws = pl + K2 * K3h;
    copy_in(tws, tw, M + K3 + K2);
    copy_in(pl, spec + ((size_t)bead * K1 + gx) * (size_t)(K2 * K3h), K2 * K3h);
    __syncthreads();
    fft_lds<true>(pl, K2, K3h, K3h, 1, tws + M + K3);
    const C *wk = tws + M;
    const int half = M / 2 + 1;
    for (int p = threadIdx.x; p < K2 * half; p += blockDim.x) {
        const int y = p / half, k = p - y * half;
        C *ln = pl + y * K3h;
        if (k == 0) {
            const real x0 = ln[0].x, xm = ln[M].x;
            ln[0] = cmk<C>(x0 + xm, x0 - xm);
        } else {
            const int kk = M - k;
            const C a = ln[k], c = ln[kk];
            const C w = wk[k];
            // k: A = a + conj c, B = (a - conj c) conj(W^k); Z = A + i B
            C A = cmk<C>(a.x + c.x, a.y - c.y);
            C Bv = cmul(cmk<C>(a.x - c.x, a.y + c.y), cconj(w));
            const C zk = cmk<C>(A.x - Bv.y, A.y + Bv.x);
            // M - k: W^-(M-k) = conj(-conj W^k) = -W^k
            A = cmk<C>(c.x + a.x, c.y - a.y);
            Bv = cmul(cmk<C>(c.x - a.x, c.y + a.y), cmk<C>(-w.x, -w.y));
            const C zkk = cmk<C>(A.x - Bv.y, A.y + Bv.x);
            ln[k] = zk;
            if (kk != k) ln[kk] = zkk;
        }
    }
    __syncthreads();
    fft_lds<true>(pl, M, K2, 1, K3h, tws);
    real *out = grid + ((size_t)bead * K1 + gx) * (size_t)(K2 * K3);
    const int lg3 = __ffs(K3) - 1;
    for (int p = threadIdx.x; p < K2 * K3; p += blockDim.x) {
        const int y = p >> lg3, z = p & (K3 - 1);
        out[p] = plr[2 * y * K3h + z];
    }
}

}  // namespace pgm
