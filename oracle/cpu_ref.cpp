// cpu_ref.cpp -- TEST INFRASTRUCTURE AND CPU BASELINE ONLY (never linked or loaded by the product).
//
// A C++17 / OpenMP float64 restatement of the pGM + LJ + smooth-PME force field, its Jacobi CG
// induced-dipole solve with the mu4 predictor, and the rigid-water velocity-Verlet SHAKE/RATTLE
// step: the same algorithm, step for step, as the numpy oracle in this directory, which follows the
// reference (pgm_jax/md/forcefield.py, pme.py, kernels.py and the constrained step).  Pinned by the
// CPU tests against that oracle (energies, forces, dipoles, CG iterations).  bench.py times it on
// every host core as the CPU baseline of the metric.
//
//   kernels    md/kernels.py:46-118      B_n of erf(a r)/r: 15-term series for (a r)^2 < 1, recursion else
//   pairs      md/forcefield.py:1089-1238, 1318-1338   rows within the cutoff (both directions),
//              G_n = B_n[a_ij] - B_n[beta] (zero beyond elec_cutoff), LJ within cutoff
//   PME        md/pme.py:57-411          order-p B-splines (OpenMM recursion), moduli, influence,
//              spread of charges and dipoles, r2c-equivalent FFT, convolution, gathers
//   solve      md/forcefield.py:1694-2008   A v = v/alpha - field(0, v); Jacobi PCG, flexible
//              Polak-Ribiere beta, max|alpha r| / norm <= tol, peek; mu4 predictor, fused residual
//   energy     md/forcefield.py:2011-2045, 2294-2430   pair, reciprocal, self, background,
//              polarisation and LJ energies; forces incl. the covalent-dipole frames; LJ tail
//   step       md/constraints.py:254-299 (SHAKE Newton on the three distances, RATTLE)
//
// Parallelism: atoms' rows are full (every pair in both rows) so each atom sums its own row; the
// spread writes thread-private grids that are summed afterwards; FFT lines run in parallel.
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <vector>

namespace {

using cd = std::complex<double>;
constexpr double KE = 138.935458;  // kJ/mol nm / e^2
constexpr double SQRT_PI = 1.7724538509055160272981674833411;
constexpr double KB = 0.0083144626181532;

// ------------------------------------------------------------------ radial kernels
// B_0..B_3 of erf(a r)/r (kernels.py:46-118).
static void erf_B(double a, double r, double *B) {
    const double s = a * a * r * r;
    if (s < 1.0) {
        double t[15];
        t[0] = 1.0;
        for (int m = 1; m < 15; ++m) t[m] = t[m - 1] * (-s) / m;
        const double c0 = 2.0 * a / SQRT_PI;
        double f = 1.0;
        for (int n = 0; n < 4; ++n) {
            double acc = 0.0;
            for (int m = 14; m >= 0; --m) acc += t[m] / (2 * m + 2 * n + 1);
            B[n] = c0 * f * acc;
            f *= 2.0 * a * a;
        }
    } else {
        const double c = 2.0 * a / SQRT_PI * std::exp(-s);
        const double ir2 = 1.0 / (r * r);
        B[0] = std::erf(a * r) / r;
        double f = 1.0;
        for (int n = 1; n < 4; ++n) {
            B[n] = ((2 * n - 1) * B[n - 1] - f * c) * ir2;
            f *= 2.0 * a * a;
        }
    }
}

// ------------------------------------------------------------------ FFT (mixed radix, recursive)
struct FFT1 {
    int n = 0;
    std::vector<int> fac;
    std::vector<cd> W;  // W[j] = exp(-2 pi i j / n)
    void init(int n_) {
        n = n_;
        fac.clear();
        int m = n;
        for (int p : {4, 2, 3, 5})
            while (m % p == 0) {
                fac.push_back(p);
                m /= p;
            }
        for (int p = 7; m > 1; p += 2)
            while (m % p == 0) {
                fac.push_back(p);
                m /= p;
            }
        W.resize(n);
        for (int j = 0; j < n; ++j) W[j] = std::polar(1.0, -2.0 * M_PI * j / n);
    }
    // y[0..len) = DFT_len of x[0], x[xs], ... (forward sign), factors from fac[fi]
    void rec(int len, const cd *x, int xs, cd *y, int fi, int wstep) const {
        if (len == 1) {
            y[0] = x[0];
            return;
        }
        const int p = fac[fi], m = len / p;
        for (int r = 0; r < p; ++r) rec(m, x + (size_t)r * xs, xs * p, y + (size_t)r * m, fi + 1, wstep * p);
        if (p == 2) {  // radix-2 and radix-4 butterflies written out
            for (int k = 0; k < m; ++k) {
                const cd a = y[k], b = y[(size_t)m + k] * W[(size_t)k * wstep];
                y[k] = a + b;
                y[(size_t)m + k] = a - b;
            }
            return;
        }
        if (p == 4) {
            for (int k = 0; k < m; ++k) {
                const cd a0 = y[k], a1 = y[(size_t)m + k] * W[(size_t)k * wstep], a2 = y[2 * (size_t)m + k] * W[2 * (size_t)k * wstep],
                         a3 = y[3 * (size_t)m + k] * W[3 * (size_t)k * wstep];
                const cd s02 = a0 + a2, d02 = a0 - a2, s13 = a1 + a3, d13 = a1 - a3;
                const cd jd13(d13.imag(), -d13.real());  // -i (a1 - a3)
                y[k] = s02 + s13;
                y[(size_t)m + k] = d02 + jd13;
                y[2 * (size_t)m + k] = s02 - s13;
                y[3 * (size_t)m + k] = d02 - jd13;
            }
            return;
        }
        cd t[16], o[16];
        for (int k = 0; k < m; ++k) {
            for (int r = 0; r < p; ++r) t[r] = y[(size_t)r * m + k] * W[(size_t)r * k * wstep];
            for (int q = 0; q < p; ++q) {
                cd s = 0.0;
                for (int r = 0; r < p; ++r) s += t[r] * W[(size_t)((r * q) % p) * m * wstep];
                o[q] = s;
            }
            for (int q = 0; q < p; ++q) y[(size_t)k + (size_t)q * m] = o[q];
        }
    }
    void forward(const cd *x, cd *y) const { rec(n, x, 1, y, 0, 1); }
};

// 3D transform of a complex grid K1 x K2 x K3 (row-major), forward (sign -1) or unnormalised inverse.
static void fft3(cd *g, const int K[3], const FFT1 f[3], bool inverse) {
    if (inverse)
        for (size_t i = 0; i < (size_t)K[0] * K[1] * K[2]; ++i) g[i] = std::conj(g[i]);
    const size_t stride[3] = {(size_t)K[1] * K[2], (size_t)K[2], 1};
    for (int ax = 2; ax >= 0; --ax) {
        const int L = K[ax];
        const int o1 = ax == 0 ? 1 : 0, o2 = ax == 2 ? 1 : 2;
        const int n1 = K[o1], n2 = K[o2];
#pragma omp parallel
        {
            std::vector<cd> a(L), b(L);
#pragma omp for schedule(static)
            for (long long l = 0; l < (long long)n1 * n2; ++l) {
                const size_t base = (size_t)(l / n2) * stride[o1] + (size_t)(l % n2) * stride[o2];
                for (int t = 0; t < L; ++t) a[t] = g[base + t * stride[ax]];
                f[ax].forward(a.data(), b.data());
                for (int t = 0; t < L; ++t) g[base + t * stride[ax]] = b[t];
            }
        }
    }
    if (inverse)
        for (size_t i = 0; i < (size_t)K[0] * K[1] * K[2]; ++i) g[i] = std::conj(g[i]);
}

// ------------------------------------------------------------------ B-splines (pme.py:57-131)
static void bspline(double f, int P, double *th, double *dth, double *d2th) {
    double data[3][9];  // orders P-2, P-1, P (index 0, 1, 2)
    double d[9] = {0};
    d[0] = 1.0 - f;
    d[1] = f;
    auto save = [&](int order) {
        if (order == P - 2) std::memcpy(data[0], d, sizeof(d));
        if (order == P - 1) std::memcpy(data[1], d, sizeof(d));
        if (order == P) std::memcpy(data[2], d, sizeof(d));
    };
    save(2);
    for (int j = 3; j <= P; ++j) {
        const double div = 1.0 / (j - 1.0);
        d[j - 1] = div * f * d[j - 2];
        for (int k = 1; k < j - 1; ++k) d[j - k - 1] = div * ((f + k) * d[j - k - 2] + (j - k - f) * d[j - k - 1]);
        d[0] = div * (1.0 - f) * d[0];
        save(j);
    }
    auto at = [&](int which, int k, int n) { return (k >= 0 && k < n) ? data[which][k] : 0.0; };
    for (int k = 0; k < P; ++k) {
        th[k] = data[2][k];
        dth[k] = at(1, k - 1, P - 1) - at(1, k, P - 1);
        d2th[k] = at(0, k - 2, P - 2) - 2.0 * at(0, k - 1, P - 2) + at(0, k, P - 2);
    }
}

static std::vector<double> moduli(int K, int P) {
    double th[9], dth[9], d2th[9];
    bspline(0.0, P, th, dth, d2th);
    std::vector<double> mod(K);
    for (int m = 0; m < K; ++m) {
        double c = 0.0, s = 0.0;
        for (int j = 0; j < P; ++j) {
            const double arg = 2.0 * M_PI * m * j / K;
            c += th[j] * std::cos(arg);
            s += th[j] * std::sin(arg);
        }
        mod[m] = c * c + s * s;
    }
    for (int i = 0; i < K; ++i)
        if (mod[i] < 1e-7) mod[i] = 0.5 * (mod[(i - 1 + K) % K] + mod[(i + 1) % K]);
    return mod;
}

// wall-clock seconds per section, summed over calls (cref_timings)
static double g_t[6] = {0, 0, 0, 0, 0, 0};  // rows, spread+FFT+convolution, gathers, row field, energy/forces, other
struct Timer {
    int k;
    double t0;
    explicit Timer(int k_) : k(k_), t0(omp_get_wtime()) {}
    ~Timer() { g_t[k] += omp_get_wtime() - t0; }
};

struct Entry {
    int j;
    double x, y, z, G[4];
};

struct Ref {
    int n = 0, nmol = 0;
    std::vector<double> q, R, alpha, rh, se, masses;
    std::vector<int> mol, ci, cj;
    std::vector<double> cc;
    double H[9], Hi[9], V = 0;
    double rc = 0, rce = 0, rcp = 0, beta = 0, tol = 1e-5, peek = 0.65;
    int K[3], P = 6, max_iter = 50, lj_lrc = 1;
    FFT1 fft[3];
    std::vector<double> Gf;  // influence on the full grid (with moduli)
    // per geometry
    std::vector<int> rs;  // row starts (n + 1)
    std::vector<Entry> rows;
    std::vector<int> base;  // 3 per atom
    std::vector<double> th, dth, d2th;  // (n, 3, P)
    std::vector<cd> grid;
    std::vector<double> phi;
    int nthreads = 1;
    // predictor state (mu4) and the last solve
    std::vector<double> hist;  // 4 x n x 3 newest first
    int count = 0;
    double norm = 1.0;
    int last_iters = 0;
    double last_err = 0;

    void min_image(double &dx, double &dy, double &dz) const {
        double s = std::nearbyint(dz / H[8]);
        dx -= s * H[6], dy -= s * H[7], dz -= s * H[8];
        s = std::nearbyint(dy / H[4]);
        dx -= s * H[3], dy -= s * H[4], dz -= s * H[5];
        s = std::nearbyint(dx / H[0]);
        dx -= s * H[0], dy -= s * H[1], dz -= s * H[2];
    }

    void set_box(const double *Hn) {
        std::memcpy(H, Hn, sizeof(H));
        const double a = H[0], b = H[1], c = H[2], d = H[3], e = H[4], f = H[5], g = H[6], h = H[7], i = H[8];
        const double det = a * (e * i - f * h) - b * (d * i - f * g) + c * (d * h - e * g);
        Hi[0] = (e * i - f * h) / det, Hi[1] = (c * h - b * i) / det, Hi[2] = (b * f - c * e) / det;
        Hi[3] = (f * g - d * i) / det, Hi[4] = (a * i - c * g) / det, Hi[5] = (c * d - a * f) / det;
        Hi[6] = (d * h - e * g) / det, Hi[7] = (b * g - a * h) / det, Hi[8] = (a * e - b * d) / det;
        V = std::fabs(det);
        // influence G_raw(m) B(m) on the full grid; reciprocal vectors are the rows of H^-T
        std::vector<double> m1 = moduli(K[0], P), m2 = moduli(K[1], P), m3 = moduli(K[2], P);
        Gf.assign((size_t)K[0] * K[1] * K[2], 0.0);
        for (int x = 0; x < K[0]; ++x) {
            const double fa = x <= (K[0] - 1) / 2 ? x : x - K[0];
            for (int y = 0; y < K[1]; ++y) {
                const double fb = y <= (K[1] - 1) / 2 ? y : y - K[1];
                for (int z = 0; z < K[2]; ++z) {
                    const double fc = z <= (K[2] - 1) / 2 ? z : z - K[2];
                    const double mx = fa * Hi[0] + fb * Hi[1] + fc * Hi[2];
                    const double my = fa * Hi[3] + fb * Hi[4] + fc * Hi[5];
                    const double mz = fa * Hi[6] + fb * Hi[7] + fc * Hi[8];
                    const double msq = mx * mx + my * my + mz * mz;
                    double G = 0.0;
                    if (msq > 0) G = std::exp(-M_PI * M_PI * msq / (beta * beta)) / (2.0 * M_PI * V * msq);
                    Gf[((size_t)x * K[1] + y) * K[2] + z] = G / (m1[x] * m2[y] * m3[z]);
                }
            }
        }
    }

    // rows within rcp (both directions) with G within rce
    void build_rows(const double *pos) {
        Timer tm(0);
        // cells along each lattice direction by plane height
        const double *h = H;
        double hgt[3];
        const double cr[3][3] = {{h[4] * h[8] - h[5] * h[7], h[5] * h[6] - h[3] * h[8], h[3] * h[7] - h[4] * h[6]},
                                 {h[7] * h[2] - h[8] * h[1], h[8] * h[0] - h[6] * h[2], h[6] * h[1] - h[7] * h[0]},
                                 {h[1] * h[5] - h[2] * h[4], h[2] * h[3] - h[0] * h[5], h[0] * h[4] - h[1] * h[3]}};
        int nc[3];
        for (int a = 0; a < 3; ++a) {
            hgt[a] = V / std::sqrt(cr[a][0] * cr[a][0] + cr[a][1] * cr[a][1] + cr[a][2] * cr[a][2]);
            nc[a] = std::max(1, (int)std::floor(hgt[a] / rcp));
        }
        const int ncell = nc[0] * nc[1] * nc[2];
        std::vector<int> cell(n), cstart(ncell + 1, 0), order(n);
        for (int i = 0; i < n; ++i) {
            int c3[3];
            for (int a = 0; a < 3; ++a) {
                double u = pos[3 * i] * Hi[a] + pos[3 * i + 1] * Hi[3 + a] + pos[3 * i + 2] * Hi[6 + a];
                u -= std::floor(u);
                c3[a] = std::min(nc[a] - 1, (int)(u * nc[a]));
            }
            cell[i] = (c3[0] * nc[1] + c3[1]) * nc[2] + c3[2];
            cstart[cell[i] + 1]++;
        }
        for (int c = 0; c < ncell; ++c) cstart[c + 1] += cstart[c];
        {
            std::vector<int> fill(cstart.begin(), cstart.end() - 1);
            for (int i = 0; i < n; ++i) order[fill[cell[i]]++] = i;
        }
        // two passes over the same candidates: count, then fill (no per-row allocations)
        rs.assign(n + 1, 0);
        auto scan = [&](int i, Entry *out) -> int {
            const int c = cell[i];
            const int c3[3] = {c / (nc[1] * nc[2]), (c / nc[2]) % nc[1], c % nc[2]};
            int nb[3][3], nn[3];
            for (int a = 0; a < 3; ++a) {
                nn[a] = 0;
                for (int d = -1; d <= 1; ++d) {
                    const int v = ((c3[a] + d) % nc[a] + nc[a]) % nc[a];
                    bool dup = false;
                    for (int u = 0; u < nn[a]; ++u) dup |= nb[a][u] == v;
                    if (!dup) nb[a][nn[a]++] = v;
                }
            }
            int cnt = 0;
            for (int u0 = 0; u0 < nn[0]; ++u0)
                for (int u1 = 0; u1 < nn[1]; ++u1)
                    for (int u2 = 0; u2 < nn[2]; ++u2) {
                        const int cc2 = (nb[0][u0] * nc[1] + nb[1][u1]) * nc[2] + nb[2][u2];
                        for (int s = cstart[cc2]; s < cstart[cc2 + 1]; ++s) {
                            const int j = order[s];
                            if (j == i) continue;
                            double dx = pos[3 * i] - pos[3 * j], dy = pos[3 * i + 1] - pos[3 * j + 1],
                                   dz = pos[3 * i + 2] - pos[3 * j + 2];
                            min_image(dx, dy, dz);
                            if (dx * dx + dy * dy + dz * dz >= rcp * rcp) continue;
                            if (out) out[cnt] = Entry{j, dx, dy, dz, {0, 0, 0, 0}};
                            ++cnt;
                        }
                    }
            return cnt;
        };
#pragma omp parallel for schedule(dynamic, 256)
        for (int i = 0; i < n; ++i) rs[i + 1] = scan(i, nullptr);
        for (int i = 0; i < n; ++i) rs[i + 1] += rs[i];
        rows.resize(rs[n]);
#pragma omp parallel for schedule(dynamic, 256)
        for (int i = 0; i < n; ++i) {
            scan(i, rows.data() + rs[i]);
            std::sort(rows.begin() + rs[i], rows.begin() + rs[i + 1], [](const Entry &a, const Entry &b) { return a.j < b.j; });
        }
        // kernels once per unordered pair (entries j > i), then copied into the mirror entries
#pragma omp parallel for schedule(dynamic, 256)
        for (int i = 0; i < n; ++i)
            for (int e = rs[i]; e < rs[i + 1]; ++e) {
                Entry &E = rows[e];
                if (E.j < i) continue;
                const double r = std::sqrt(E.x * E.x + E.y * E.y + E.z * E.z);
                if (r >= rce) continue;
                const double a = 1.0 / std::sqrt(2.0 * (R[i] * R[i] + R[E.j] * R[E.j]));
                double Ba[4], Bb[4];
                erf_B(a, r, Ba);
                erf_B(beta, r, Bb);
                for (int k = 0; k < 4; ++k) E.G[k] = Ba[k] - Bb[k];
            }
#pragma omp parallel for schedule(dynamic, 256)
        for (int i = 0; i < n; ++i)
            for (int e = rs[i]; e < rs[i + 1]; ++e) {
                Entry &E = rows[e];
                if (E.j > i) continue;
                const Entry *b = rows.data() + rs[E.j], *f = rows.data() + rs[E.j + 1];
                const Entry *m = std::lower_bound(b, f, i, [](const Entry &u, int v) { return u.j < v; });
                for (int k = 0; k < 4; ++k) E.G[k] = m->G[k];
            }
    }

    void pme_setup(const double *pos) {
        base.resize(3 * (size_t)n);
        th.resize(3 * (size_t)n * P), dth.resize(3 * (size_t)n * P), d2th.resize(3 * (size_t)n * P);
#pragma omp parallel for schedule(static)
        for (int i = 0; i < n; ++i)
            for (int a = 0; a < 3; ++a) {
                const double u = pos[3 * i] * Hi[a] + pos[3 * i + 1] * Hi[3 + a] + pos[3 * i + 2] * Hi[6 + a];
                const double w = (u - std::floor(u)) * K[a];
                const double b = std::floor(w);
                base[3 * i + a] = (int)b;
                const size_t o = ((size_t)i * 3 + a) * P;
                bspline(w - b, P, &th[o], &dth[o], &d2th[o]);
            }
    }

    // e = d . H^-1 . diag(K)
    void escale(const double *d, double *e) const {
        for (int b = 0; b < 3; ++b) e[b] = (d[0] * Hi[b] + d[1] * Hi[3 + b] + d[2] * Hi[6 + b]) * K[b];
    }

    // phi = dE_rec/dQ on the grid for charges q (nullable) and dipoles d; returns E_rec
    double potential(const double *q, const double *d) {
        Timer tm(1);
        const size_t K3 = (size_t)K[0] * K[1] * K[2];
        std::vector<double> Q(K3 * nthreads, 0.0);
#pragma omp parallel
        {
            double *mine = Q.data() + K3 * omp_get_thread_num();
#pragma omp for schedule(static)
            for (int i = 0; i < n; ++i) {
                double e[3];
                escale(d + 3 * i, e);
                const double qi = q ? q[i] : 0.0;
                const double *t1 = &th[(size_t)i * 3 * P], *t2 = t1 + P, *t3 = t2 + P;
                const double *s1 = &dth[(size_t)i * 3 * P], *s2 = s1 + P, *s3 = s2 + P;
                for (int a = 0; a < P; ++a) {
                    const int xa = (base[3 * i] + a) % K[0];
                    const double a1 = qi * t1[a] + e[0] * s1[a], b1 = e[1] * t1[a], c1 = e[2] * t1[a];
                    for (int b = 0; b < P; ++b) {
                        const int yb = (base[3 * i + 1] + b) % K[1];
                        const double v12 = a1 * t2[b] + b1 * s2[b], w12 = c1 * t2[b];
                        double *row = mine + ((size_t)xa * K[1] + yb) * K[2];
                        for (int c = 0; c < P; ++c) row[(base[3 * i + 2] + c) % K[2]] += v12 * t3[c] + w12 * s3[c];
                    }
                }
            }
        }
        grid.resize(K3);
#pragma omp parallel for schedule(static)
        for (long long x = 0; x < (long long)K3; ++x) {
            double s = 0.0;
            for (int t = 0; t < nthreads; ++t) s += Q[K3 * t + x];
            grid[x] = cd(s, 0.0);
        }
        fft3(grid.data(), K, fft, false);
        double E = 0.0;
#pragma omp parallel for reduction(+ : E) schedule(static)
        for (long long x = 0; x < (long long)K3; ++x) {
            E += Gf[x] * std::norm(grid[x]);
            grid[x] *= Gf[x];
        }
        fft3(grid.data(), K, fft, true);
        phi.resize(K3);
#pragma omp parallel for schedule(static)
        for (long long x = 0; x < (long long)K3; ++x) phi[x] = 2.0 * grid[x].real();
        return E;
    }

    // dE_rec/dd (N,3) and, if gr, dE_rec/dr (N,3), from phi
    void pme_gradients(const double *q, const double *d, double *gd, double *gr) {
        Timer tm(2);
#pragma omp parallel for schedule(static)
        for (int i = 0; i < n; ++i) {
            const double *t[3], *s[3], *u[3];
            for (int a = 0; a < 3; ++a) {
                t[a] = &th[((size_t)i * 3 + a) * P];
                s[a] = &dth[((size_t)i * 3 + a) * P];
                u[a] = &d2th[((size_t)i * 3 + a) * P];
            }
            // contractions over the stencil: sum P w1 w2 w3 for the weight combinations needed
            double g[3] = {0, 0, 0};                 // first derivatives (dipole gradient)
            double h2[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};  // second derivatives (a, b)
            for (int a = 0; a < P; ++a) {
                const int xa = (base[3 * i] + a) % K[0];
                for (int b = 0; b < P; ++b) {
                    const int yb = (base[3 * i + 1] + b) % K[1];
                    const double *row = &phi[((size_t)xa * K[1] + yb) * K[2]];
                    double S0 = 0, S1 = 0, S2 = 0;  // sum_c row * t3, * s3, * u3
                    for (int c = 0; c < P; ++c) {
                        const double v = row[(base[3 * i + 2] + c) % K[2]];
                        S0 += v * t[2][c];
                        S1 += v * s[2][c];
                        S2 += v * u[2][c];
                    }
                    g[0] += s[0][a] * t[1][b] * S0;
                    g[1] += t[0][a] * s[1][b] * S0;
                    g[2] += t[0][a] * t[1][b] * S1;
                    if (gr) {
                        h2[0][0] += u[0][a] * t[1][b] * S0;
                        h2[1][1] += t[0][a] * u[1][b] * S0;
                        h2[2][2] += t[0][a] * t[1][b] * S2;
                        h2[0][1] += s[0][a] * s[1][b] * S0;
                        h2[0][2] += s[0][a] * t[1][b] * S1;
                        h2[1][2] += t[0][a] * s[1][b] * S1;
                    }
                }
            }
            // dE/dd_a = sum_b g_b e_scale[a, b], e_scale = H^-1 diag(K)
            for (int a = 0; a < 3; ++a) gd[3 * i + a] = g[0] * Hi[3 * a] * K[0] + g[1] * Hi[3 * a + 1] * K[1] + g[2] * Hi[3 * a + 2] * K[2];
            if (gr) {
                h2[1][0] = h2[0][1], h2[2][0] = h2[0][2], h2[2][1] = h2[1][2];
                double e[3];
                escale(d + 3 * i, e);
                const double qi = q ? q[i] : 0.0;
                double dw[3];
                for (int a = 0; a < 3; ++a) {
                    dw[a] = qi * g[a];
                    for (int b = 0; b < 3; ++b) dw[a] += e[b] * h2[a][b];
                }
                for (int a = 0; a < 3; ++a)
                    gr[3 * i + a] = dw[0] * Hi[3 * a] * K[0] + dw[1] * Hi[3 * a + 1] * K[1] + dw[2] * Hi[3 * a + 2] * K[2];
            }
        }
    }

    // minus the direct-space field: sum_k (-q_k G1 - G2 d_k.x) x + G1 d_k (forcefield.py:1395-1402)
    void row_field(const double *qq, const double *d, double *out) const {
        Timer tm(3);
#pragma omp parallel for schedule(dynamic, 256)
        for (int i = 0; i < n; ++i) {
            double ex = 0, ey = 0, ez = 0;
            for (int e = rs[i]; e < rs[i + 1]; ++e) {
                const Entry &E = rows[e];
                const double *dk = d + 3 * E.j;
                double cf = -E.G[2] * (dk[0] * E.x + dk[1] * E.y + dk[2] * E.z);
                if (qq) cf -= qq[E.j] * E.G[1];
                ex += cf * E.x + E.G[1] * dk[0];
                ey += cf * E.y + E.G[1] * dk[1];
                ez += cf * E.z + E.G[1] * dk[2];
            }
            out[3 * i] = ex, out[3 * i + 1] = ey, out[3 * i + 2] = ez;
        }
    }

    double c_self() const { return 4.0 * beta * beta * beta / (3.0 * SQRT_PI); }

    // total field -(row + PME - c_self d)
    void field(const double *qq, const double *d, double *out) {
        std::vector<double> rf(3 * (size_t)n), gd(3 * (size_t)n);
        row_field(qq, d, rf.data());
        potential(qq, d);
        pme_gradients(qq, d, gd.data(), nullptr);
        const double cs = c_self();
#pragma omp parallel for schedule(static)
        for (long long k = 0; k < 3LL * n; ++k) out[k] = -(rf[k] + gd[k] - cs * d[k]);
    }

    void apply_A(const double *v, double *out) {
        field(nullptr, v, out);
#pragma omp parallel for schedule(static)
        for (long long k = 0; k < 3LL * n; ++k) out[k] = (alpha[k / 3] > 0 ? v[k] / alpha[k / 3] : 0.0) - out[k];
    }

    void perm_dipoles(const double *pos, double *p) const {
        std::fill(p, p + 3 * (size_t)n, 0.0);
        for (size_t m = 0; m < ci.size(); ++m) {
            double dx = pos[3 * cj[m]] - pos[3 * ci[m]], dy = pos[3 * cj[m] + 1] - pos[3 * ci[m] + 1],
                   dz = pos[3 * cj[m] + 2] - pos[3 * ci[m] + 2];
            min_image(dx, dy, dz);
            const double L = std::sqrt(dx * dx + dy * dy + dz * dz);
            p[3 * ci[m]] += cc[m] * dx / L, p[3 * ci[m] + 1] += cc[m] * dy / L, p[3 * ci[m] + 2] += cc[m] * dz / L;
        }
    }

    static double dot(const std::vector<double> &a, const std::vector<double> &b) {
        double s = 0.0;
#pragma omp parallel for reduction(+ : s) schedule(static)
        for (long long k = 0; k < (long long)a.size(); ++k) s += a[k] * b[k];
        return s;
    }

    // predictor + CG (forcefield.py:1714-2008 as the oracle's solve); mu out
    void solve(const double *p, double *mu) {
        const size_t n3 = 3 * (size_t)n;
        std::vector<double> x(n3), r(n3), b(n3), tmp(n3), dsum(n3);
        const bool have = count >= 4;
        if (have)
            for (size_t k = 0; k < n3; ++k)
                x[k] = 4 * hist[k] - 6 * hist[n3 + k] + 4 * hist[2 * n3 + k] - hist[3 * n3 + k];
        if (have && count % 1000 != 0) {  // fused residual: field(q, p + x0) - x0/alpha
            for (size_t k = 0; k < n3; ++k) dsum[k] = p[k] + x[k];
            field(q.data(), dsum.data(), r.data());
            for (size_t k = 0; k < n3; ++k) r[k] -= alpha[k / 3] > 0 ? x[k] / alpha[k / 3] : 0.0;
        } else {
            field(q.data(), p, b.data());
            double s = 0.0;
            for (size_t k = 0; k < n3; ++k) s += std::fabs(alpha[k / 3] * b[k]);
            norm = s / n3 + 1e-300;
            if (!have)
                for (size_t k = 0; k < n3; ++k) x[k] = alpha[k / 3] * b[k];
            apply_A(x.data(), tmp.data());
            for (size_t k = 0; k < n3; ++k) r[k] = b[k] - tmp[k];
        }
        std::vector<double> z(n3), pp(n3), Ap(n3), rn(n3), zn(n3);
        for (size_t k = 0; k < n3; ++k) z[k] = pp[k] = r[k] * alpha[k / 3];
        double rz = dot(r, z);
        auto errf = [&](const std::vector<double> &rr) {
            double m = 0.0;
            for (size_t k = 0; k < n3; ++k) m = std::max(m, std::fabs(rr[k] * alpha[k / 3]));
            return m / norm;
        };
        int it = 0;
        double err = errf(r);
        while (err > tol && it < max_iter) {
            apply_A(pp.data(), Ap.data());
            const double al = rz / std::max(dot(pp, Ap), 1e-300);
            for (size_t k = 0; k < n3; ++k) {
                x[k] += al * pp[k];
                rn[k] = r[k] - al * Ap[k];
                zn[k] = rn[k] * alpha[k / 3];
            }
            double num = 0.0;
            for (size_t k = 0; k < n3; ++k) num += zn[k] * (rn[k] - r[k]);
            const double be = num / std::max(rz, 1e-300);
            for (size_t k = 0; k < n3; ++k) pp[k] = zn[k] + be * pp[k];
            rz = dot(rn, zn);
            r.swap(rn);
            ++it;
            err = errf(r);
        }
        for (size_t k = 0; k < n3; ++k) mu[k] = x[k] + peek * r[k] * alpha[k / 3];
        // history push (newest first)
        std::memmove(hist.data() + n3, hist.data(), 3 * n3 * sizeof(double));
        std::memcpy(hist.data(), mu, n3 * sizeof(double));
        count += 1;
        last_iters = it;
        last_err = err;
    }

    double lj_tail() const {
        if (!lj_lrc) return 0.0;
        double Mk[7] = {0};
        for (int i = 0; i < n; ++i) {
            double f = 1.0;
            for (int k = 0; k < 7; ++k) {
                Mk[k] += se[i] * f;
                f *= rh[i];
            }
        }
        const double binom[7] = {1, 6, 15, 20, 15, 6, 1};
        double SB = 0.0;
        for (int k = 0; k < 7; ++k) SB += 2.0 * binom[k] * Mk[k] * Mk[6 - k];
        return -2.0 * M_PI * SB / (3.0 * V * rc * rc * rc);
    }

    // energies {elec, vdw} and forces at fixed mu (forcefield.py:2362-2430)
    void energy_forces(const double *pos, const double *mu, double *en, double *forces) {
        const size_t n3 = 3 * (size_t)n;
        std::vector<double> p(n3), d(n3), g_el(n3), g_lj(n3), dEdd(n3), rec_dd(n3), rec_dr(n3);
        perm_dipoles(pos, p.data());
        for (size_t k = 0; k < n3; ++k) d[k] = p[k] + mu[k];
        double se_sum = 0.0, sl_sum = 0.0;
#pragma omp parallel for reduction(+ : se_sum, sl_sum) schedule(dynamic, 256)
        for (int i = 0; i < n; ++i) {
            const double qi = q[i];
            const double *di = &d[3 * i];
            double gx = 0, gy = 0, gz = 0, lx = 0, ly = 0, lz = 0, fx = 0, fy = 0, fz = 0;
            for (int e = rs[i]; e < rs[i + 1]; ++e) {
                const Entry &E = rows[e];
                const int j = E.j;
                const double *dk = &d[3 * j];
                const double qk = q[j], x = E.x, y = E.y, z = E.z;
                const double *G = E.G;
                const double dix = di[0] * x + di[1] * y + di[2] * z, dkx = dk[0] * x + dk[1] * y + dk[2] * z;
                const double didk = di[0] * dk[0] + di[1] * dk[1] + di[2] * dk[2];
                const double t = qi * dkx - qk * dix;
                se_sum += 0.5 * (qi * qk * G[0] + t * G[1] - G[2] * dix * dkx + G[1] * didk);
                const double radial = -qi * qk * G[1] - t * G[2] + G[3] * dix * dkx - G[2] * didk;
                gx += radial * x + qi * G[1] * dk[0] - qk * G[1] * di[0] - G[2] * (di[0] * dkx + dk[0] * dix);
                gy += radial * y + qi * G[1] * dk[1] - qk * G[1] * di[1] - G[2] * (di[1] * dkx + dk[1] * dix);
                gz += radial * z + qi * G[1] * dk[2] - qk * G[1] * di[2] - G[2] * (di[2] * dkx + dk[2] * dix);
                const double cf = -qk * G[1] - G[2] * dkx;
                fx += cf * x + G[1] * dk[0], fy += cf * y + G[1] * dk[1], fz += cf * z + G[1] * dk[2];
                const double eps = se[i] * se[j];
                const double r2 = x * x + y * y + z * z;
                if (mol[i] != mol[j] && eps != 0.0 && r2 < rc * rc) {
                    const double rmin = rh[i] + rh[j];
                    const double s2 = rmin * rmin / r2, s6 = s2 * s2 * s2;
                    sl_sum += 0.5 * eps * (s6 * s6 - 2.0 * s6);
                    const double dl = eps * 12.0 * (s6 - s6 * s6) / r2;
                    lx += dl * x, ly += dl * y, lz += dl * z;
                }
            }
            g_el[3 * i] = gx, g_el[3 * i + 1] = gy, g_el[3 * i + 2] = gz;
            g_lj[3 * i] = lx, g_lj[3 * i + 1] = ly, g_lj[3 * i + 2] = lz;
            dEdd[3 * i] = KE * fx, dEdd[3 * i + 1] = KE * fy, dEdd[3 * i + 2] = KE * fz;
        }
        const double u_rec = potential(q.data(), d.data());
        pme_gradients(q.data(), d.data(), rec_dd.data(), rec_dr.data());
        double sq = 0, sq2 = 0, dd = 0, pol = 0;
        for (int i = 0; i < n; ++i) {
            sq += q[i], sq2 += q[i] * q[i];
            for (int a = 0; a < 3; ++a) {
                dd += d[3 * i + a] * d[3 * i + a];
                if (alpha[i] > 0) pol += mu[3 * i + a] * mu[3 * i + a] / (2.0 * alpha[i]);
            }
        }
        const double cs = c_self();
        const double u_self = -(beta / SQRT_PI) * sq2 - 0.5 * cs * dd;
        const double u_bg = -M_PI * sq * sq / (2.0 * V * beta * beta);
        std::vector<double> gdd(n3);
        for (size_t k = 0; k < n3; ++k) gdd[k] = dEdd[k] + KE * (rec_dd[k] - cs * d[k]);
        // pull-back of dE/dd through the covalent-dipole frames
        std::vector<double> vjp(n3, 0.0);
        for (size_t m = 0; m < ci.size(); ++m) {
            double v[3] = {pos[3 * cj[m]] - pos[3 * ci[m]], pos[3 * cj[m] + 1] - pos[3 * ci[m] + 1],
                           pos[3 * cj[m] + 2] - pos[3 * ci[m] + 2]};
            min_image(v[0], v[1], v[2]);
            const double L = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
            const double u[3] = {v[0] / L, v[1] / L, v[2] / L};
            const double *gi = &gdd[3 * ci[m]];
            const double gu = gi[0] * u[0] + gi[1] * u[1] + gi[2] * u[2];
            for (int a = 0; a < 3; ++a) {
                const double tv = cc[m] * (gi[a] - gu * u[a]) / L;
                vjp[3 * cj[m] + a] += tv;
                vjp[3 * ci[m] + a] -= tv;
            }
        }
#pragma omp parallel for schedule(static)
        for (long long k = 0; k < (long long)n3; ++k) forces[k] = -(KE * g_el[k] + g_lj[k] + KE * rec_dr[k] + vjp[k]);
        en[0] = KE * se_sum + KE * (u_rec + u_self + u_bg + pol);
        en[1] = sl_sum + lj_tail();
    }

    void compute(const double *pos, double *mu, double *en, double *forces) {
        build_rows(pos);
        pme_setup(pos);
        std::vector<double> p(3 * (size_t)n);
        perm_dipoles(pos, p.data());
        solve(p.data(), mu);
        energy_forces(pos, mu, en, forces);
    }
};

}  // namespace

extern "C" {

void *cref_create(int n, const double *q, const double *radius, const double *alpha, const double *rh, const double *se,
                  const int *mol, int ncov, const int *ci, const int *cj, const double *cc, const double *H, double cutoff,
                  double elec_cutoff, double beta, const int *grid, int order, int lj_lrc, double tol, int max_iter,
                  double peek, int threads) {
    Ref *r = new Ref();
    r->n = n;
    r->q.assign(q, q + n), r->R.assign(radius, radius + n), r->alpha.assign(alpha, alpha + n);
    r->rh.assign(rh, rh + n), r->se.assign(se, se + n), r->mol.assign(mol, mol + n);
    r->ci.assign(ci, ci + ncov), r->cj.assign(cj, cj + ncov), r->cc.assign(cc, cc + ncov);
    r->rc = cutoff;
    r->rce = elec_cutoff > 0 ? elec_cutoff : cutoff;
    r->rcp = std::max(r->rc, r->rce);
    r->beta = beta;
    for (int a = 0; a < 3; ++a) r->K[a] = grid[a];
    r->P = order, r->lj_lrc = lj_lrc, r->tol = tol, r->max_iter = max_iter, r->peek = peek;
    if (threads > 0) omp_set_num_threads(threads);
    r->nthreads = omp_get_max_threads();
    for (int a = 0; a < 3; ++a) r->fft[a].init(r->K[a]);
    r->hist.assign(12 * (size_t)n, 0.0);
    r->set_box(H);
    return r;
}

void cref_destroy(void *h) { delete (Ref *)h; }

// seconds spent so far in: rows (cell list + pair kernels), spread + FFTs + convolution, PME
// gathers, row field, (unused), (unused); reset with reset != 0
void cref_timings(double *out, int reset) {
    for (int k = 0; k < 6; ++k) out[k] = g_t[k];
    if (reset)
        for (int k = 0; k < 6; ++k) g_t[k] = 0.0;
}

int cref_threads(void *h) { return ((Ref *)h)->nthreads; }

// one force evaluation with the mu4 predictor state of the context; energies {elec, vdw}
int cref_compute(void *h, const double *pos, double *mu, double *energy, double *forces, int *iterations,
                 double *residual) {
    Ref *r = (Ref *)h;
    r->compute(pos, mu, energy, forces);
    if (iterations) *iterations = r->last_iters;
    if (residual) *residual = r->last_err;
    return 0;
}

// nsteps NVE velocity-Verlet steps with SHAKE / RATTLE of consecutive 3-site molecules (the
// oracle's step); forces hold the forces at pos on entry and exit; report[0..3] = epot, ekin,
// total CG iterations, largest CG residual
int cref_md(void *h, double *pos, double *vel, double *forces, const double *masses, const double *d2, int nsteps, double dt,
            double *mu, double *report) {
    Ref *r = (Ref *)h;
    const int M = r->n / 3;
    const int CA[3] = {0, 0, 1}, CB[3] = {1, 2, 2};
    auto g = [&](int a, int c) { return CA[c] == a ? 1.0 : (CB[c] == a ? -1.0 : 0.0); };
    double en[2] = {0, 0};
    long long cg = 0;
    double res = 0.0;
    for (int st = 0; st < nsteps; ++st) {
        std::vector<double> r0(pos, pos + 3 * (size_t)r->n);
#pragma omp parallel for schedule(static)
        for (int m = 0; m < M; ++m) {
            double v[3][3], x0[3][3], x1[3][3], minv[3];
            for (int a = 0; a < 3; ++a) {
                const int i = 3 * m + a;
                minv[a] = 1.0 / masses[i];
                for (int k = 0; k < 3; ++k) {
                    v[a][k] = vel[3 * i + k] + 0.5 * dt * forces[3 * i + k] * minv[a];
                    x0[a][k] = r0[3 * i + k] - r0[3 * (3 * m) + k];
                    x1[a][k] = x0[a][k] + dt * v[a][k];
                }
            }
            // SHAKE: Newton on sigma_c = |x_c|^2 - d_c^2 with displacements sum_c lam_c g(a,c) s_c / m_a
            double s[3][3], lam[3] = {0, 0, 0}, disp[3][3] = {{0}};
            for (int c = 0; c < 3; ++c)
                for (int k = 0; k < 3; ++k) s[c][k] = x0[CA[c]][k] - x0[CB[c]][k];
            for (int it = 0; it < 50; ++it) {
                double rr[3][3], xc[3][3], sig[3];
                for (int a = 0; a < 3; ++a)
                    for (int k = 0; k < 3; ++k) {
                        disp[a][k] = 0.0;
                        for (int c = 0; c < 3; ++c) disp[a][k] += g(a, c) * lam[c] * s[c][k];
                        disp[a][k] *= minv[a];
                        rr[a][k] = x1[a][k] + disp[a][k];
                    }
                double worst = 0.0;
                for (int c = 0; c < 3; ++c) {
                    double q2 = 0.0;
                    for (int k = 0; k < 3; ++k) {
                        xc[c][k] = rr[CA[c]][k] - rr[CB[c]][k];
                        q2 += xc[c][k] * xc[c][k];
                    }
                    sig[c] = q2 - d2[3 * m + c];
                    worst = std::max(worst, std::fabs(sig[c]) / d2[3 * m + c]);
                }
                if (worst < 1e-14) break;
                double J[3][3];
                for (int c = 0; c < 3; ++c)
                    for (int cp = 0; cp < 3; ++cp) {
                        const double w = g(CA[c], cp) * minv[CA[c]] - g(CB[c], cp) * minv[CB[c]];
                        J[c][cp] = 2.0 * w * (xc[c][0] * s[cp][0] + xc[c][1] * s[cp][1] + xc[c][2] * s[cp][2]);
                    }
                // solve J dl = -sig (Cramer)
                const double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) - J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                                   J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
                double dl[3];
                for (int c = 0; c < 3; ++c) {
                    double A[3][3];
                    for (int u = 0; u < 3; ++u)
                        for (int w = 0; w < 3; ++w) A[u][w] = w == c ? -sig[u] : J[u][w];
                    dl[c] = (A[0][0] * (A[1][1] * A[2][2] - A[1][2] * A[2][1]) - A[0][1] * (A[1][0] * A[2][2] - A[1][2] * A[2][0]) +
                             A[0][2] * (A[1][0] * A[2][1] - A[1][1] * A[2][0])) / det;
                }
                for (int c = 0; c < 3; ++c) lam[c] += dl[c];
            }
            for (int a = 0; a < 3; ++a) {
                const int i = 3 * m + a;
                for (int k = 0; k < 3; ++k) {
                    pos[3 * i + k] = x1[a][k] + disp[a][k] + r0[3 * (3 * m) + k];
                    vel[3 * i + k] = v[a][k] + disp[a][k] / dt;
                }
            }
        }
        r->compute(pos, mu, en, forces);
        cg += r->last_iters;
        res = std::max(res, r->last_err);
        // second half kick and RATTLE
#pragma omp parallel for schedule(static)
        for (int m = 0; m < M; ++m) {
            double v[3][3], x[3][3], minv[3], s[3][3];
            for (int a = 0; a < 3; ++a) {
                const int i = 3 * m + a;
                minv[a] = 1.0 / masses[i];
                for (int k = 0; k < 3; ++k) {
                    v[a][k] = vel[3 * i + k] + 0.5 * dt * forces[3 * i + k] * minv[a];
                    x[a][k] = pos[3 * i + k];
                }
            }
            for (int c = 0; c < 3; ++c)
                for (int k = 0; k < 3; ++k) s[c][k] = x[CA[c]][k] - x[CB[c]][k];
            double Km[3][3], rhs[3];
            for (int c = 0; c < 3; ++c) {
                for (int cp = 0; cp < 3; ++cp) {
                    const double w = g(CA[c], cp) * minv[CA[c]] - g(CB[c], cp) * minv[CB[c]];
                    Km[c][cp] = w * (s[c][0] * s[cp][0] + s[c][1] * s[cp][1] + s[c][2] * s[cp][2]);
                }
                rhs[c] = -((v[CA[c]][0] - v[CB[c]][0]) * s[c][0] + (v[CA[c]][1] - v[CB[c]][1]) * s[c][1] +
                           (v[CA[c]][2] - v[CB[c]][2]) * s[c][2]);
            }
            const double det = Km[0][0] * (Km[1][1] * Km[2][2] - Km[1][2] * Km[2][1]) - Km[0][1] * (Km[1][0] * Km[2][2] - Km[1][2] * Km[2][0]) +
                               Km[0][2] * (Km[1][0] * Km[2][1] - Km[1][1] * Km[2][0]);
            double mu3[3];
            for (int c = 0; c < 3; ++c) {
                double A[3][3];
                for (int u = 0; u < 3; ++u)
                    for (int w = 0; w < 3; ++w) A[u][w] = w == c ? rhs[u] : Km[u][w];
                mu3[c] = (A[0][0] * (A[1][1] * A[2][2] - A[1][2] * A[2][1]) - A[0][1] * (A[1][0] * A[2][2] - A[1][2] * A[2][0]) +
                          A[0][2] * (A[1][0] * A[2][1] - A[1][1] * A[2][0])) / det;
            }
            for (int a = 0; a < 3; ++a) {
                const int i = 3 * m + a;
                for (int k = 0; k < 3; ++k) {
                    double dv = 0.0;
                    for (int c = 0; c < 3; ++c) dv += g(a, c) * mu3[c] * s[c][k];
                    vel[3 * i + k] = v[a][k] + dv * minv[a];
                }
            }
        }
    }
    double ke = 0.0;
    for (int i = 0; i < r->n; ++i)
        for (int k = 0; k < 3; ++k) ke += 0.5 * masses[i] * vel[3 * i + k] * vel[3 * i + k];
    report[0] = en[0] + en[1];
    report[1] = ke;
    report[2] = (double)cg;
    report[3] = res;
    return 0;
}

}  // extern "C"
