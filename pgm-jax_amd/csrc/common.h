// common.h -- shared device helpers of libpgm_hip: precision traits, box geometry, wave/block
// reductions and the Gaussian-screened radial kernels of the pGM pair terms.
//
// gfx950 (CDNA4): wave64.  Every kernel is templated on `real` (float: mixed precision, double:
// the plain FP64 path); positions, minimum images and all reductions are FP64.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define PGM_DEV __device__ __forceinline__

namespace pgm {

constexpr int WAVE = 64;
constexpr double KE = 138.935458;                 // Coulomb constant, kJ/mol nm / e^2 (CODATA)
constexpr double TWO_OVER_SQRTPI = 1.1283791670955126;
constexpr double SQRT_PI = 1.7724538509055159;
constexpr int MAX_ORDER = 8;

template <typename real> struct vec4;
template <> struct vec4<float> { using type = float4; };
template <> struct vec4<double> { using type = double4; };
template <typename real> using real4 = typename vec4<real>::type;

// Beads: B independent copies of one system (ring-polymer beads, replicas) stacked along the atom
// axis, atom i of bead b at b * nper + i.  Kernels that write per-block partial sums run on a
// (blocks per bead, B) grid so that no block straddles two beads; the CG state is one record of
// CG_STRIDE ints per bead (conv_it first), and iteration `iter` of a bead runs while
// iter < conv_it of that bead.
constexpr int CG_STRIDE = 4;
PGM_DEV bool cg_done(const int *conv, int bead, int iter) { return conv && iter >= conv[bead * CG_STRIDE]; }

// Reduced lower-triangular box: lattice vectors as rows, h[3a+b].
struct Box {
    double h[9];
    double hinv[9];   // H^-1 (lower triangular)
    double inv_diag[3];
};

// Minimum image of a reduced box: sequential reduction along c, b, a (nearest integers).
PGM_DEV void min_image(double &x, double &y, double &z, const Box &b) {
    double n = rint(z * b.inv_diag[2]);
    x -= n * b.h[6];
    y -= n * b.h[7];
    z -= n * b.h[8];
    n = rint(y * b.inv_diag[1]);
    x -= n * b.h[3];
    y -= n * b.h[4];
    n = rint(x * b.inv_diag[0]);
    x -= n * b.h[0];
}

// min_image that also returns the lattice shifts it applied (n[a] times lattice row a subtracted).
PGM_DEV void min_image_n(double &x, double &y, double &z, const Box &b, int (&n)[3]) {
    double m = rint(z * b.inv_diag[2]);
    x -= m * b.h[6];
    y -= m * b.h[7];
    z -= m * b.h[8];
    n[2] = (int)m;
    m = rint(y * b.inv_diag[1]);
    x -= m * b.h[3];
    y -= m * b.h[4];
    n[1] = (int)m;
    m = rint(x * b.inv_diag[0]);
    x -= m * b.h[0];
    n[0] = (int)m;
}

// The same shifts applied in min_image's order (bitwise the min_image result when n is its own).
PGM_DEV void apply_shift(double &x, double &y, double &z, const Box &b, const int (&n)[3]) {
    x -= n[2] * b.h[6];
    y -= n[2] * b.h[7];
    z -= n[2] * b.h[8];
    x -= n[1] * b.h[3];
    y -= n[1] * b.h[4];
    x -= n[0] * b.h[0];
}

// ---------------------------------------------------------------------------- reductions
template <typename T> PGM_DEV T wave_sum(T v) {
#pragma unroll
    for (int o = WAVE / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, WAVE);
    return v;
}
template <typename T> PGM_DEV T wave_max(T v) {
#pragma unroll
    for (int o = WAVE / 2; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, WAVE));
    return v;
}
// Sum over groups of G consecutive lanes (G a power of two <= 64); every lane gets its group's sum.
template <int G, typename T> PGM_DEV T group_sum(T v) {
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, WAVE);
    return v;
}

// Block sum of NV doubles; result valid in thread 0.  blockDim.x a multiple of 64, <= 1024.
template <int NV> PGM_DEV void block_sum(double (&v)[NV], double *lds /* NV * 16 */) {
    const int lane = threadIdx.x % WAVE, w = threadIdx.x / WAVE, nw = blockDim.x / WAVE;
#pragma unroll
    for (int k = 0; k < NV; ++k) v[k] = wave_sum(v[k]);
    __syncthreads();
    if (lane == 0)
        for (int k = 0; k < NV; ++k) lds[k * 16 + w] = v[k];
    __syncthreads();
    if (threadIdx.x == 0)
        for (int k = 0; k < NV; ++k) {
            double s = 0.0;
            for (int i = 0; i < nw; ++i) s += lds[k * 16 + i];
            v[k] = s;
        }
}
PGM_DEV double block_max(double v, double *lds) {
    const int lane = threadIdx.x % WAVE, w = threadIdx.x / WAVE, nw = blockDim.x / WAVE;
    v = wave_max(v);
    __syncthreads();
    if (lane == 0) lds[w] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = 0.0;
        for (int i = 0; i < nw; ++i) m = fmax(m, lds[i]);
        v = m;
    }
    return v;
}

// ---------------------------------------------------------------------------- radial kernels
// B_n of phi_a(r) = erf(a r)/r: B0 = phi, B_n = -(1/r) dB_{n-1}/dr (grad_x B_n = -B_{n+1} x).
// Series for s = (a r)^2 < 1: B_n = (2a/sqrt(pi)) (2a^2)^n sum_{m<15} (-s)^m / (m! (2m+2n+1)).
// Closed form: B_n = [(2n-1) B_{n-1} - (2a^2)^(n-1) c] / r^2, c = (2a/sqrt(pi)) exp(-a^2 r^2).
// For a Gaussian pair G_n = B_n[a] - B_n[b0]; when both screenings are on the closed-form branch
// the difference is formed from erfc (G0 = (erfc(b0 r) - erfc(a r))/r), which keeps float32
// accurate where erf(a r)/r and erf(b0 r)/r nearly cancel.
PGM_DEV float erfc_(float x) { return erfcf(x); }
PGM_DEV double erfc_(double x) { return erfc(x); }
PGM_DEV float erf_(float x) { return erff(x); }
PGM_DEV double erf_(double x) { return erf(x); }
PGM_DEV float exp_(float x) { return expf(x); }
PGM_DEV double exp_(double x) { return exp(x); }
PGM_DEV float sqrt_(float x) { return sqrtf(x); }
PGM_DEV double sqrt_(double x) { return sqrt(x); }
PGM_DEV float rsqrt_(float x) { return rsqrtf(x); }
PGM_DEV double rsqrt_(double x) { return 1.0 / sqrt(x); }

// erfc(x) for x >= 1 from e = exp(-x^2) and ix = 1/x, which the caller has anyway (the same
// Gaussian enters the recursion below):  erfc(x) = e h(ix^2) ix / sqrt(pi), where h(u) =
// sqrt(pi) x exp(x^2) erfc(x) at u = 1/x^2 is written h = 1 - u R(u) with R a (4,3) rational fitted
// by this build (relative least squares on x in [1, 1e4] against the float64 erfcx): |h error| <=
// 1.3e-8 in exact arithmetic, 9.4e-8 evaluated in FP32 -- below the 1e-6 relative error that
// rounding x^2 to FP32 already puts into exp(-x^2) near x = 3.6.  FP64 evaluations keep the
// library erfc.
constexpr float ERFC_RP[5] = {0.49999627f, 3.05839218f, 3.49413551f, 0.24121188f, -0.02171751f};
constexpr float ERFC_RQ[4] = {1.0f, 7.61614911f, 14.68195645f, 6.73569184f};
PGM_DEV float erfc_ge1(float e, float ix) {
    const float u = ix * ix;
    const float P = (((ERFC_RP[4] * u + ERFC_RP[3]) * u + ERFC_RP[2]) * u + ERFC_RP[1]) * u + ERFC_RP[0];
    const float Q = ((ERFC_RQ[3] * u + ERFC_RQ[2]) * u + ERFC_RQ[1]) * u + ERFC_RQ[0];
    const float h = 1.0f - u * (P * __builtin_amdgcn_rcpf(Q));
    return e * h * ix * float(TWO_OVER_SQRTPI / 2.0);
}

// G_0 = (erfc(b r) - erfc(a r)) / r on the closed-form branch ((a r)^2, (b r)^2 >= 1) from the two
// Gaussians ea = exp(-(a r)^2), eb = exp(-(b r)^2) and ir = 1/r.
PGM_DEV float closed_G0(float a, float b, float, float ir, float ea, float eb) {
    return (erfc_ge1(eb, ir * __builtin_amdgcn_rcpf(b)) - erfc_ge1(ea, ir * __builtin_amdgcn_rcpf(a))) * ir;
}
PGM_DEV double closed_G0(double a, double b, double r, double ir, double, double) {
    return (erfc(b * r) - erfc(a * r)) * ir;
}

template <typename real, int NMAX> PGM_DEV void series_B(real a, real s, real *B) {
    real S[NMAX];
#pragma unroll
    for (int n = 0; n < NMAX; ++n) S[n] = real(0);
    real t = real(1);
#pragma unroll
    for (int m = 0; m < 15; ++m) {
#pragma unroll
        for (int n = 0; n < NMAX; ++n) S[n] += t * real(1.0 / double(2 * m + 2 * n + 1));
        t *= -s * real(1.0 / double(m + 1));
    }
    real c = real(TWO_OVER_SQRTPI) * a;
    const real f = real(2) * a * a;
#pragma unroll
    for (int n = 0; n < NMAX; ++n) {
        B[n] = c * S[n];
        c *= f;
    }
}

template <typename real, int NMAX> PGM_DEV void closed_B(real a, real r, real ir, real ir2, real *B) {
    const real c = real(TWO_OVER_SQRTPI) * a * exp_(-(a * r) * (a * r));
    const real f = real(2) * a * a;
    B[0] = erf_(a * r) * ir;
    real fc = c;
#pragma unroll
    for (int n = 1; n < NMAX; ++n) {
        B[n] = (real(2 * n - 1) * B[n - 1] - fc) * ir2;
        fc *= f;
    }
}

template <typename real, int NMAX> PGM_DEV void one_B(real a, real r, real ir, real ir2, real *B) {
    const real s = (a * r) * (a * r);
    if (s < real(1))
        series_B<real, NMAX>(a, s, B);
    else
        closed_B<real, NMAX>(a, r, ir, ir2, B);
}

// G_n when both (a r)^2 and (b r)^2 are >= 1 (the closed-form difference only).
template <typename real, int NMAX> PGM_DEV void pair_G_far(real a, real b, real r2, real *G) {
    const real ir = rsqrt_(r2);
    const real r = r2 * ir;
    const real ir2 = ir * ir;
    const real sa = a * a * r2, sb = b * b * r2;
    const real ea = exp_(-sa), eb = exp_(-sb);
    const real ca = real(TWO_OVER_SQRTPI) * a * ea;
    const real cb = real(TWO_OVER_SQRTPI) * b * eb;
    G[0] = closed_G0(a, b, r, ir, ea, eb);
    const real fa = real(2) * a * a, fb = real(2) * b * b;
    real pa = ca, pb = cb;
#pragma unroll
    for (int n = 1; n < NMAX; ++n) {
        G[n] = (real(2 * n - 1) * G[n - 1] - (pa - pb)) * ir2;
        pa *= fa;
        pb *= fb;
    }
}

// G_n = B_n[erf(a r)/r] - B_n[erf(b r)/r], n < NMAX.
template <typename real, int NMAX> PGM_DEV void pair_G(real a, real b, real r2, real *G) {
    const real ir = rsqrt_(r2);
    const real r = r2 * ir;
    const real ir2 = ir * ir;
    const real sa = a * a * r2, sb = b * b * r2;
    if (sa >= real(1) && sb >= real(1)) {
        const real ea = exp_(-sa), eb = exp_(-sb);
        const real ca = real(TWO_OVER_SQRTPI) * a * ea;
        const real cb = real(TWO_OVER_SQRTPI) * b * eb;
        G[0] = closed_G0(a, b, r, ir, ea, eb);
        const real fa = real(2) * a * a, fb = real(2) * b * b;
        real pa = ca, pb = cb;
#pragma unroll
        for (int n = 1; n < NMAX; ++n) {
            G[n] = (real(2 * n - 1) * G[n - 1] - (pa - pb)) * ir2;
            pa *= fa;
            pb *= fb;
        }
    } else {
        real Ba[NMAX], Bb[NMAX];
        one_B<real, NMAX>(a, r, ir, ir2, Ba);
        one_B<real, NMAX>(b, r, ir, ir2, Bb);
#pragma unroll
        for (int n = 0; n < NMAX; ++n) G[n] = Ba[n] - Bb[n];
    }
}

// ---------------------------------------------------------------------------- B-splines
// Cardinal B-spline weights of order P at offset f in [0, 1): th[k] = M_P(f + P - 1 - k), the
// weight of grid point floor(w) + k, and their derivatives d/df (recursion raised one order at a
// time, derivatives from the order P-1 values).  With D2, also the second derivatives (order P-2).
template <typename real, int P, bool D2>
PGM_DEV void bspline(real f, real *th, real *dth, real *d2th) {
    real d[P];
#pragma unroll
    for (int k = 0; k < P; ++k) d[k] = real(0);
    d[0] = real(1) - f;
    d[1] = f;
    real m2[P], m1[P];
#pragma unroll
    for (int j = 3; j <= P; ++j) {
        if (j == P - 1) {
#pragma unroll
            for (int k = 0; k < P; ++k) m2[k] = d[k];
        }
        if (j == P) {
#pragma unroll
            for (int k = 0; k < P; ++k) m1[k] = d[k];
        }
        const real div = real(1.0 / double(j - 1));
        d[j - 1] = div * f * d[j - 2];
#pragma unroll
        for (int k = 1; k < j - 1; ++k) d[j - k - 1] = div * ((f + real(k)) * d[j - k - 2] + (real(j - k) - f) * d[j - k - 1]);
        d[0] = div * (real(1) - f) * d[0];
    }
    if (P == 3) {
#pragma unroll
        for (int k = 0; k < P; ++k) m1[k] = real(0);
    }
#pragma unroll
    for (int k = 0; k < P; ++k) {
        th[k] = d[k];
        const real a = k >= 1 ? m1[k - 1] : real(0);
        const real b = k < P - 1 ? m1[k] : real(0);
        dth[k] = a - b;
    }
    if (D2) {
#pragma unroll
        for (int k = 0; k < P; ++k) {
            const real a = k >= 2 ? m2[k - 2] : real(0);
            const real b = (k >= 1 && k - 1 < P - 2) ? m2[k - 1] : real(0);
            const real c = k < P - 2 ? m2[k] : real(0);
            d2th[k] = a - real(2) * b + c;
        }
    }
}

}  // namespace pgm
