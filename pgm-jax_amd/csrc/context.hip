// context.hip -- the C ABI of libpgm_hip (include/pgm_hip.h): context, rows, PME plans, the
// induced-dipole solve and the energy/force pass, dispatched to the float (mixed) or double
// instantiation of the kernels.
#include <hipcub/hipcub.hpp>
#include <hip/hip_ext.h>
#include <rocfft/rocfft.h>

#include <algorithm>
#include <functional>
#include <mutex>
#include <climits>
#include <random>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "cg.h"
#include "common.h"
#include "pairs.h"
#include "md.h"
#include "pme.h"
#include "fft.h"
#include "virial.h"
#include "ewald.h"
#include "dense.h"
#include "mlp.h"
#include "guess.h"
#include "adjoint.h"
#include "pgm_hip.h"

using namespace pgm;

static thread_local std::string g_err;

#define HIP_TRY(x)                                                                         \
    do {                                                                                   \
        hipError_t e_ = (x);                                                               \
        if (e_ != hipSuccess) {                                                            \
            g_err = std::string(#x) + ": " + hipGetErrorString(e_);                        \
            return PGM_ERR_HIP;                                                            \
        }                                                                                  \
    } while (0)
#define FFT_TRY(x)                                                                         \
    do {                                                                                   \
        rocfft_status s_ = (x);                                                            \
        if (s_ != rocfft_status_success) {                                                 \
            g_err = std::string(#x) + ": rocfft status " + std::to_string((int)s_);        \
            return PGM_ERR_FFT;                                                            \
        }                                                                                  \
    } while (0)
#define PGM_TRY(x)                                                                         \
    do {                                                                                   \
        int r_ = (x);                                                                      \
        if (r_ < 0) return r_;                                                             \
    } while (0)
#define LAUNCH_CHECK()                                                                     \
    do {                                                                                   \
        hipError_t e_ = hipGetLastError();                                                 \
        if (e_ != hipSuccess) {                                                            \
            g_err = std::string("kernel launch: ") + hipGetErrorString(e_) + " at line " + \
                    std::to_string(__LINE__);                                              \
            return PGM_ERR_HIP;                                                            \
        }                                                                                  \
    } while (0)

static inline int nblk(long long n, int b) { return (int)((n + b - 1) / b); }

#define ORDER_SWITCH(P, ...)                                          \
    switch (P) {                                                      \
        case 4: { constexpr int PO = 4; __VA_ARGS__; } break;         \
        case 5: { constexpr int PO = 5; __VA_ARGS__; } break;         \
        case 6: { constexpr int PO = 6; __VA_ARGS__; } break;         \
        case 7: { constexpr int PO = 7; __VA_ARGS__; } break;         \
        case 8: { constexpr int PO = 8; __VA_ARGS__; } break;         \
        default: g_err = "PME order must be 4..8"; return PGM_ERR_ARG; \
    }

struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    int ensure(size_t b) {
        if (b <= bytes && p) return PGM_OK;
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
        if (b == 0) return PGM_OK;
        hipError_t e = hipMalloc(&p, b);
        if (e != hipSuccess) {
            g_err = std::string("hipMalloc(") + std::to_string(b) + "): " + hipGetErrorString(e);
            return PGM_ERR_HIP;
        }
        bytes = b;
        return PGM_OK;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    template <typename T> T *as() const { return (T *)p; }
};

struct pgm_ctx {
    int n = 0;
    // beads: nbead copies of one system of nper atoms stacked along the atom axis (n = nbead nper);
    // each has its own rows, PME grid, CG state and energies, all advanced by the same launches
    int nbead = 1, nper = 0;
    pgm_settings s{};
    // cutoffs [nm]: electrostatic pair terms within rc_e, LJ within rc_v (the settings' cutoff), rows
    // (cell lists, candidates, compacted force-pass rows) within rc_pair = max(rc_e, rc_v)
    double rc_e = 0, rc_v = 0, rc_pair = 0;
    hipStream_t stream = nullptr;
    hipStream_t side = nullptr;           // second stream: the pair rows run beside the PME chain
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    hipEvent_t ev_rec = nullptr;  // the molecule records of the last evaluation, when they ran on the side stream
    bool rec_pending = false;     // ev_rec not yet waited for on the context stream
    hipEvent_t ev_read[4] = {nullptr, nullptr, nullptr, nullptr};
    bool mixed = false;
    size_t rsize = 8;  // sizeof(real)
    // box
    double H[9]{};
    double volume = 0;
    Box box{};
    PMEGeom pg{};
    int3 nc{1, 1, 1};
    // topology (host copies and device arrays)
    std::vector<double> h_q, h_R2, h_alpha, h_rh, h_se;
    std::vector<int> h_mol;
    double sum_q = 0, sum_q2 = 0;
    double lj_sb = 0;  // sum_{i,j} 2 eps_ij rmin_ij^6 (LJ tail)
    DevBuf q, alpha64, alpha_r, rh, se, mol, R2;
    DevBuf ljp;  // (rmin/2, sqrt eps, molecule, 0) per atom, double4 (the force pass)
    DevBuf cstart, ci, cj, cc, jstart, jlist;
    DevBuf corder;  // caller index of entry k of the owner-sorted covalent list
    std::vector<int> h_corder;
    // geometry: the positions of the current evaluation (caller's device buffer) and xr
    const double *pos = nullptr;
    DevBuf xr;
    // rows
    int cap_full = 0, cap = 0;
    bool rows_valid = false;
    DevBuf rows_full, count_full, rows, count, count_near, maxcount, jell, G1, G2;
    // CG row field layout: atom rows in sliced ELL (jell, G1, G2; RL lanes per atom) or, for 3-site
    // molecules, molecule blocks (mj, mg; rlm lanes per molecule) from the candidate molecules mrows
    bool mol3 = false;            // every bead is consecutive 3-atom molecules
    bool alpha0 = false;          // some atom has polarizability 0 (its induced dipole stays 0)
    bool atom_rows_fresh = false; // rows / count / count_near compacted for the current geometry
    // CG iterations as hipGraphs (pgm_set_graphs): one executable per iteration index k, captured on
    // first use and replayed; dropped whenever a launch argument they hold can change
    bool use_graphs = false;
    std::vector<hipGraphExec_t> it_graphs;
    hipStream_t gstream = nullptr;              // capture / replay stream when the context runs on the null stream
    hipEvent_t ev_g0 = nullptr, ev_g1 = nullptr;  // ordering between the two
    bool molp = false;            // molecule-block CG preconditioner (settings.precond 1)
    bool pinv_stale = true;       // its blocks are remade at the next evaluation after a row build
    DevBuf pinv;                  // its inverse 9x9 blocks, 81 reals per molecule
    int layout = 0;               // 0 atom rows, 1 molecule blocks
    int rlm = 4;
    int mcap_full = 0, mcap = 0;
    DevBuf mrows, mcount_full, mcount, mj, mg;
    DevBuf key, atom, skey, satom, cell_start, cell_end, sort_tmp;
    long long entries = 0;
    // reciprocal space: PME, or (ewald) the exact Ewald sum over the half-space index set m
    // (fixed at the reference box) with kvec = (k, f_k) recomputed per box
    bool ewald = false;
    int nk = 0;
    std::vector<int> h_m;
    DevBuf kvec, sk;
    // PME
    size_t ngrid = 0, ncplx = 0;
    DevBuf grid, cgrid, Gc, fft_work;
    rocfft_plan fwd = nullptr, bwd = nullptr;
    rocfft_execution_info info = nullptr;
    // fused transform of the CG field evaluations (fft.h): available for power-of-two grids with
    // tiled spreading; pme_fused selects it over the library plans
    bool fused_ok = false, pme_fused = false;
    // where a CG field evaluation forks the row field onto the side stream: 0 before the spread,
    // 1 after the spread, 2 after the forward planes, 3 before the spread with the forward planes
    // waiting for it (2 and 3: fused transform only); chosen by pgm_autotune
    int fork_stage = 0;
    PFFTGeom pfg{};
    DevBuf pfft_tw;
    // tiled spread (atoms sorted by grid tile)
    // spread tiles of T8 points per axis (tiles8, one fine key each) inside gather tiles of 2 T8
    // (tiles); tstart/tend by fine key, tstart16/tend16 by gather tile
    bool tiled = false;
    int tile_cap = 8;  // largest spread-tile width tried per axis (pgm_autotune lowers it)
    PMETiles tiles{}, tiles8{};
    DevBuf pkey, pidx, pskey, psidx, tstart, tend, tstart16, tend16, spl, loc, loc16, blocks, pme_sort_tmp, merge_tab;
    // vectors (real4 N) and FP64 (N, 3) work arrays
    DevBuf vx, vr, vz, vp, vAp, vb, vx0, vrow, vdq, vtmp;
    DevBuf pperm, dd, g_el, g_lj, rowf, rec_dr, gp;
    DevBuf partA, partB, partC, pair_part, pack_part, conv_part;
    DevBuf st, rz, err, energy;    // CGState, rz and err rows (stride rzs), energy: 4 per bead then aux slots
    int rzs = 0;
    DevBuf bead_it, bead_res;      // iterations and final residual of every bead's last solve
    int *h_pin_int = nullptr;      // 16 + nbead ints (bead iterations from 16)
    CGState *h_pin_conv = nullptr; // CG state read-backs of the pipelined CG (4 slots of nbead states)
    unsigned long long *h_mirror = nullptr;  // CG states written by the update kernels (cg_mirror), 4 slots of nbead,
                                             // then the MD step's largest displacement (a double, h_disp)
    unsigned cg_seq = 0;                     // sequence number of the current CG solve (the mirror's stamp)
    // uniform static external field (pgm_set_efield): F0 in e/nm^2, the dipole offset in e nm; the
    // field's dipole partials and (e_field, M) per bead of the last evaluation
    bool efield_on = false;
    Ext3 F0{}, foff{};
    DevBuf fpart, fout;
    double *h_pin_dbl = nullptr;   // 32 + 5 nbead doubles (bead energies from 32, residuals after them)
    int last_iters = 4;
    int last_adj_iters = 4;  // iterations of the last adjoint solve (read-back hint)
    DevBuf adj;              // FP64 work arrays of the adjoint: lam, p_bar (3N each), phi, q_bar, R_bar, alpha_bar (N each)
    DevBuf adj2;             // FP64 work arrays of the parameter cotangents (compute_vjp_impl)
    bool norm_on_device = false;  // CGState.norm holds the normaliser of the last unfused solve
    DevBuf stats;                 // per-solve final residuals (ring of stats_cap)
    int stats_cap = 4096;
    long long stats_slot = 0;
    // strain derivative: molecules as CSR (atoms of molecule m in index order), centres of mass,
    // offsets from them, dU_rec/dd and per-block partials
    int nmol_all = 0;
    DevBuf mol_ptr, mol_atoms, com, off, rec_dd, vir_pair, vir_conv, vir_atoms, vir_cov, vir_out;
    int ncov = 0;
    // live kernel timing of the CG matvecs (pgm_set_timing): HIP event pairs around the row field
    // (side stream), the spread (+ merge) and the gather (main stream)
    struct Timed {
        std::vector<hipEvent_t> ev;  // 2 per launch
        int used = 0;
        double sum_ms = 0.0;
        long long count = 0;
    };
    bool timing = false;
    Timed timed[3];
    // learned CG initial guess (predictor 4): network GUESS_D -> 128 x nh -> 3, harvest ring, Adam
    struct Guess {
        bool on = false;        // configured
        bool trained = false;   // at least one training call
        int mode = 0;           // the selector's choice: 0 mu4, 1 fitted linear predictor, 2 linear + network
        int nh = 3, every = 10, cap = 0;
        long long nparams = 0, rows = 0;
        unsigned long long seed = 0;
        int adam = 0, head = 0, nring = 0, nfit = 0;
        double scale = 1.0, sigma = 0.0, err_mu4 = 0.0, err_lin = 0.0, err_net = 0.0;
        double gram[GUESS_K * GUESS_K] = {0}, rhs[GUESS_K] = {0};
        GuessCoef lin{{4.0, -6.0, 4.0, -1.0, 0.0, 0.0, 0.0, 0.0}};
        bool have_y = false;    // y holds the network output of the current solve
        DevBuf params, grads, m, v, packed, bias, feat, frames, y, bx, by, idx, xb, yb, acts, zs, delta, delta2, part, scal, ring;
    } guess;
    // MD (rigid 3-site molecules)
    DevBuf pos_build, minv, mass, d2, md_part, md_fail, heat;
    int nmol = 0;
    double friction = 0.0, kT = 0.0;  // Langevin O step (friction 0: NVE)
    unsigned long long seed = 0;
    bool md_ready = false;
    int rebuilds = 0;
};

// Device slots after the per-bead energies: [0] sum m v^2 of the MD step, [1] largest displacement
// since the row build.
static double *aux_slots(const pgm_ctx *c) { return c->energy.as<double>() + 4 * c->nbead; }
// The largest displacement since the row build, written by the MD step's reduction into pinned host memory.
static double *h_disp(const pgm_ctx *c) { return reinterpret_cast<double *>(c->h_mirror + 4 * (size_t)c->nbead); }

static void drop_graphs(pgm_ctx *c) {
    for (hipGraphExec_t &g : c->it_graphs)
        if (g) (void)hipGraphExecDestroy(g);
    c->it_graphs.clear();
}

extern "C" const char *pgm_last_error(void) { return g_err.c_str(); }
extern "C" int pgm_version(void) { return 1; }

// ---------------------------------------------------------------------------- host helpers
static void host_bspline(double f, int P, double *th) {
    std::vector<double> d(P, 0.0);
    d[0] = 1.0 - f;
    d[1] = f;
    for (int j = 3; j <= P; ++j) {
        const double div = 1.0 / (j - 1.0);
        d[j - 1] = div * f * d[j - 2];
        for (int k = 1; k < j - 1; ++k) d[j - k - 1] = div * ((f + k) * d[j - k - 2] + (j - k - f) * d[j - k - 1]);
        d[0] = div * (1.0 - f) * d[0];
    }
    for (int k = 0; k < P; ++k) th[k] = d[k];
}

// |b(m)|^-2 denominators (pgm_jax/md/pme.py:106-131)
static std::vector<double> moduli(int K, int P) {
    std::vector<double> th(P), mod(K);
    host_bspline(0.0, P, th.data());
    for (int m = 0; m < K; ++m) {
        double c = 0, s = 0;
        for (int j = 0; j < P; ++j) {
            const double a = 2.0 * M_PI * m * j / K;
            c += th[j] * std::cos(a);
            s += th[j] * std::sin(a);
        }
        mod[m] = c * c + s * s;
    }
    for (int m = 0; m < K; ++m)
        if (mod[m] < 1e-7) mod[m] = 0.5 * (mod[(m - 1 + K) % K] + mod[(m + 1) % K]);
    return mod;
}

static void inv3(const double *H, double *Hi) {
    const double a = H[0], b = H[1], c = H[2], d = H[3], e = H[4], f = H[5], g = H[6], h = H[7], i = H[8];
    const double det = a * (e * i - f * h) - b * (d * i - f * g) + c * (d * h - e * g);
    Hi[0] = (e * i - f * h) / det;
    Hi[1] = (c * h - b * i) / det;
    Hi[2] = (b * f - c * e) / det;
    Hi[3] = (f * g - d * i) / det;
    Hi[4] = (a * i - c * g) / det;
    Hi[5] = (c * d - a * f) / det;
    Hi[6] = (d * h - e * g) / det;
    Hi[7] = (b * g - a * h) / det;
    Hi[8] = (a * e - b * d) / det;
}

// sum_{i,j} 2 eps_ij rmin_ij^6 by the binomial moments of rmin/2 (pgm_jax/lj.py:87-114): fixed by the
// topology, computed once at creation.
static double lj_sum_b(const pgm_ctx *c) {
    double M[7] = {0};
    for (int i = 0; i < c->nper; ++i) {
        double p = c->h_se[i];
        for (int k = 0; k < 7; ++k) {
            M[k] += p;
            p *= c->h_rh[i];
        }
    }
    static const double C6[7] = {1, 6, 15, 20, 15, 6, 1};
    double SB = 0;
    for (int k = 0; k < 7; ++k) SB += C6[k] * M[k] * M[6 - k];
    return 2.0 * SB;
}

// E_lrc = -2 pi / (3 V rc^3) sum_{i,j} 2 eps_ij rmin_ij^6
static double lj_tail(const pgm_ctx *c) {
    return -2.0 * M_PI * c->lj_sb / (3.0 * c->volume * c->s.cutoff * c->s.cutoff * c->s.cutoff);
}

template <typename T> static int upload(DevBuf &b, const T *h, size_t count, hipStream_t st) {
    PGM_TRY(b.ensure(std::max<size_t>(count, 1) * sizeof(T)));
    if (count) HIP_TRY(hipMemcpyAsync(b.p, h, count * sizeof(T), hipMemcpyHostToDevice, st));
    return PGM_OK;
}

// ---------------------------------------------------------------------------- box
// The influence function of the convolution, Gc = 2 G_raw B(m) on the half spectrum (K1, K2,
// K3/2 + 1) without the half-space weights (pgm_jax/md/pme.py:243-267), for the box whose inverse
// is Hi (row-major) and volume c->volume.
static void influence_host(const pgm_ctx *c, const double *Hi, std::vector<double> &G) {
    const int K1 = c->s.grid[0], K2 = c->s.grid[1], K3 = c->s.grid[2], K3h = K3 / 2 + 1;
    std::vector<double> m1 = moduli(K1, c->s.order), m2 = moduli(K2, c->s.order), m3 = moduli(K3, c->s.order);
    const double R[9] = {Hi[0], Hi[3], Hi[6], Hi[1], Hi[4], Hi[7], Hi[2], Hi[5], Hi[8]};  // rows: reciprocal vectors (H^-T)
    const double beta = c->s.ewald_beta;
    G.assign((size_t)K1 * K2 * K3h, 0.0);
    for (int a = 0; a < K1; ++a) {
        const double fa = a <= (K1 - 1) / 2 ? a : a - K1;  // fftfreq
        for (int b = 0; b < K2; ++b) {
            const double fb = b <= (K2 - 1) / 2 ? b : b - K2;
            for (int k = 0; k < K3h; ++k) {
                const double mx = fa * R[0] + fb * R[3] + k * R[6];
                const double my = fa * R[1] + fb * R[4] + k * R[7];
                const double mz = fa * R[2] + fb * R[5] + k * R[8];
                const double msq = mx * mx + my * my + mz * mz;
                double g = 0.0;
                if (msq > 0) g = std::exp(-M_PI * M_PI * msq / (beta * beta)) / (2.0 * M_PI * c->volume * msq);
                G[((size_t)a * K2 + b) * K3h + k] = 2.0 * g / (m1[a] * m2[b] * m3[k]);
            }
        }
    }
}

static int set_box(pgm_ctx *c, const double *H) {
    drop_graphs(c);
    for (int k = 0; k < 9; ++k) c->H[k] = H[k];
    if (std::fabs(H[1]) + std::fabs(H[2]) + std::fabs(H[5]) > 1e-12 * std::fabs(H[0])) {
        g_err = "box must be lower triangular (lattice vectors as rows)";
        return PGM_ERR_ARG;
    }
    double Hi[9];
    inv3(H, Hi);
    for (int k = 0; k < 9; ++k) c->box.h[k] = H[k];
    for (int k = 0; k < 9; ++k) c->box.hinv[k] = Hi[k];
    for (int a = 0; a < 3; ++a) c->box.inv_diag[a] = 1.0 / H[4 * a];
    c->volume = std::fabs(H[0] * H[4] * H[8]);
    for (int k = 0; k < 9; ++k) c->pg.hinv[k] = Hi[k];
    for (int b = 0; b < 3; ++b)
        for (int a = 0; a < 3; ++a) c->pg.escale[3 * b + a] = Hi[3 * b + a] * c->s.grid[a];
    for (int a = 0; a < 3; ++a) c->pg.K[a] = c->s.grid[a];
    // cell grid: plane heights / (row cutoff + skin)
    const double rl = c->rc_pair + c->s.skin;
    double ax[3] = {H[0], H[1], H[2]}, bx[3] = {H[3], H[4], H[5]}, cx[3] = {H[6], H[7], H[8]};
    auto cross = [](const double *u, const double *v, double *w) {
        w[0] = u[1] * v[2] - u[2] * v[1];
        w[1] = u[2] * v[0] - u[0] * v[2];
        w[2] = u[0] * v[1] - u[1] * v[0];
    };
    double w[3];
    int ncs[3];
    const double *vec[3][2] = {{bx, cx}, {cx, ax}, {ax, bx}};
    for (int a = 0; a < 3; ++a) {
        cross(vec[a][0], vec[a][1], w);
        const double h = c->volume / std::sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
        if (2.0 * rl > h * (1.0 + 1e-12) && rl > 0.5 * h) {
            g_err = "cutoff + skin exceeds half the box height";
            return PGM_ERR_ARG;
        }
        ncs[a] = std::max(1, (int)std::floor(h / rl));
    }
    c->nc = make_int3(ncs[0], ncs[1], ncs[2]);
    std::vector<double> G;
    influence_host(c, Hi, G);
    if (c->mixed) {
        std::vector<float> Gf(G.begin(), G.end());
        PGM_TRY(upload(c->Gc, Gf.data(), Gf.size(), c->stream));
    } else {
        PGM_TRY(upload(c->Gc, G.data(), G.size(), c->stream));
    }
    if (c->ewald) {  // k = 2 pi H^-1 m, f_k = (4 pi / V) exp(-k^2 / (4 beta^2)) / k^2 (pgm_jax/ewald.py:317-333)
        const double beta = c->s.ewald_beta;
        std::vector<double4> kv((size_t)std::max(c->nk, 1));
        for (int k = 0; k < c->nk; ++k) {
            const double m0 = c->h_m[3 * k], m1 = c->h_m[3 * k + 1], m2 = c->h_m[3 * k + 2];
            const double kx = 2.0 * M_PI * (Hi[0] * m0 + Hi[1] * m1 + Hi[2] * m2);
            const double ky = 2.0 * M_PI * (Hi[3] * m0 + Hi[4] * m1 + Hi[5] * m2);
            const double kz = 2.0 * M_PI * (Hi[6] * m0 + Hi[7] * m1 + Hi[8] * m2);
            const double k2 = kx * kx + ky * ky + kz * kz;
            kv[k] = make_double4(kx, ky, kz, (4.0 * M_PI / c->volume) * std::exp(-k2 / (4.0 * beta * beta)) / k2);
        }
        PGM_TRY(upload(c->kvec, kv.data(), kv.size(), c->stream));
    }
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->rows_valid = false;
    return PGM_OK;
}

// Tile widths of the spread: per axis the smallest T in 8..16 (else 7..P-1) with at least two tiles
// and a last tile at least P-1 wide; no such T: the atomic spread is used.
static size_t spread_lds_bytes(const pgm_ctx *c) {
    const int P = c->s.order;
    return (size_t)SPREAD_CHUNK * (6 * P + 4) * c->rsize + SPREAD_CHUNK * sizeof(int) + c->rsize * (size_t)c->tiles8.bsize;
}

// Raise the dynamic-LDS limit of the tiled spread / gather instantiations of this context.
template <typename real, int P> static int allow_lds_p(size_t spread, size_t gather) {
    HIP_TRY(hipFuncSetAttribute((const void *)k_spread_tiles<real, P, true>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)spread));
    HIP_TRY(hipFuncSetAttribute((const void *)k_spread_tiles<real, P, false>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)spread));
    HIP_TRY(hipFuncSetAttribute((const void *)k_gather_tiles<real, P>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)gather));
    return PGM_OK;
}
static int allow_lds(pgm_ctx *c, size_t spread, size_t gather) {
    if (spread > 160 * 1024 || gather > 160 * 1024) {
        g_err = "PME tile block exceeds the LDS";
        return PGM_ERR_ARG;
    }
    int r = PGM_OK;
    if (c->mixed) {
        ORDER_SWITCH(c->s.order, r = (allow_lds_p<float, PO>(spread, gather)));
    } else {
        ORDER_SWITCH(c->s.order, r = (allow_lds_p<double, PO>(spread, gather)));
    }
    return r;
}

// Spread tiles: per axis the largest T8 in tile_cap (8 unless tuned)..P-1 with at least three tiles
// and a last tile at least P-1 wide; gather tiles of T16 = 2 T8 (the union of two spread tiles per axis).  No such T8: the
// atomic spread and the per-atom gather are used.
static int setup_tiles(pgm_ctx *c) {
    drop_graphs(c);
    const int P = c->s.order;
    PMETiles t8{}, t16{};
    bool ok = true;
    for (int a = 0; a < 3; ++a) {
        const int K = c->s.grid[a];
        int best = 0;
        for (int T = c->tile_cap; T >= P - 1 && !best; --T) {
            const int nt = (K + T - 1) / T;
            const int last = K - (nt - 1) * T;
            if (nt >= 3 && last >= P - 1) best = T;
        }
        if (!best) {
            ok = false;
            break;
        }
        t8.T[a] = best;
        t8.nt[a] = (K + best - 1) / best;
        t8.B[a] = best + P - 1;
        t16.T[a] = 2 * best;
        t16.nt[a] = (t8.nt[a] + 1) / 2;
        t16.B[a] = 2 * best + P - 1;
    }
    c->tiled = ok;
    if (!ok) return PGM_OK;
    t16.ntiles = t16.nt[0] * t16.nt[1] * t16.nt[2];
    t16.bsize = t16.B[0] * t16.B[1] * t16.B[2];
    t8.ntiles = 8 * t16.ntiles;  // by fine key (octants past the grid stay empty)
    t8.bsize = t8.B[0] * t8.B[1] * t8.B[2];
    c->tiles = t16;
    c->tiles8 = t8;
    const int n = c->n;
    const size_t B = (size_t)c->nbead;
    PGM_TRY(c->pkey.ensure(sizeof(int) * n));
    PGM_TRY(c->pidx.ensure(sizeof(int) * n));
    PGM_TRY(c->pskey.ensure(sizeof(int) * n));
    PGM_TRY(c->psidx.ensure(sizeof(int) * n));
    PGM_TRY(c->loc.ensure(sizeof(int) * n));
    PGM_TRY(c->loc16.ensure(sizeof(int) * n));
    PGM_TRY(c->tstart.ensure(sizeof(int) * B * t8.ntiles));
    PGM_TRY(c->tend.ensure(sizeof(int) * B * t8.ntiles));
    PGM_TRY(c->tstart16.ensure(sizeof(int) * B * t16.ntiles));
    PGM_TRY(c->tend16.ensure(sizeof(int) * B * t16.ntiles));
    PGM_TRY(c->spl.ensure(c->rsize * 6 * P * (size_t)n));
    PGM_TRY(c->blocks.ensure(c->rsize * B * t8.ntiles * t8.bsize));
    {  // per-axis offsets of the block merge (k_merge_table, k_pme_fwd_plane)
        std::vector<int> tab;
        for (int a = 0; a < 3; ++a) {
            const int K = c->s.grid[a], T = t8.T[a], nt = t8.nt[a];
            const int stride = a == 0 ? t8.B[1] * t8.B[2] : (a == 1 ? t8.B[2] : 1);
            auto A = [&](int u) {
                const int hi = u >> 1, lo = u & 1;
                if (a == 0) return hi * t16.nt[1] * t16.nt[2] * 8 + lo * 4;
                if (a == 1) return hi * t16.nt[2] * 8 + lo * 2;
                return hi * 8 + lo;
            };
            for (int g = 0; g < K; ++g) {
                const int u = std::min(g / T, nt - 1), l = g - u * T;
                tab.push_back(A(u) * t8.bsize + l * stride);
                if (l <= P - 2) {
                    const int up = u > 0 ? u - 1 : nt - 1;
                    const int width = up < nt - 1 ? T : K - (nt - 1) * T;
                    tab.push_back(A(up) * t8.bsize + (l + width) * stride);
                } else {
                    tab.push_back(-1);
                }
            }
        }
        PGM_TRY(upload(c->merge_tab, tab.data(), tab.size(), c->stream));
    }
    // gather partials come per gather tile: the partial buffers must hold B ntiles entries
    for (DevBuf *b : {&c->partA, &c->partB}) PGM_TRY(b->ensure(std::max(b->bytes, sizeof(double) * 4 * B * t16.ntiles)));
    // dynamic LDS of the tiled spread and gather beyond the default limit
    const size_t lds_spread = spread_lds_bytes(c), lds_gather = c->rsize * (size_t)t16.bsize;
    PGM_TRY(allow_lds(c, lds_spread, lds_gather));
    return PGM_OK;
}

// Sort the atoms by fine key and store their spline weights (once per geometry).
template <typename real> static int pme_bin(pgm_ctx *c, const double *pos) {
    if (!c->tiled) return PGM_OK;
    const int n = c->n;
    hipStream_t st = c->stream;
    const PMETiles &t8 = c->tiles8, &t16 = c->tiles;
    const int nk8 = c->nbead * t8.ntiles, nk16 = c->nbead * t16.ntiles;
    k_pme_bin<<<nblk(n, 256), 256, 0, st>>>(n, c->nper, pos, c->pg, t8, t16, c->pkey.as<int>(), c->pidx.as<int>());
    LAUNCH_CHECK();
    int bits = 1;
    while ((1 << bits) < nk8) ++bits;
    size_t tmp = 0;
    HIP_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, c->pkey.as<int>(), c->pskey.as<int>(), c->pidx.as<int>(),
                                               c->psidx.as<int>(), n, 0, bits, st));
    PGM_TRY(c->pme_sort_tmp.ensure(tmp));
    HIP_TRY(hipcub::DeviceRadixSort::SortPairs(c->pme_sort_tmp.p, tmp, c->pkey.as<int>(), c->pskey.as<int>(), c->pidx.as<int>(),
                                               c->psidx.as<int>(), n, 0, bits, st));
    k_tile_bounds<<<nblk(n + 1, 256), 256, 0, st>>>(n, c->pskey.as<int>(), nk8, c->tstart.as<int>(), c->tend.as<int>(), nk16,
                                                         3, c->tstart16.as<int>(), c->tend16.as<int>());
    LAUNCH_CHECK();
    ORDER_SWITCH(c->s.order, (k_pme_splines<real, PO><<<nblk(n, 256), 256, 0, st>>>(n, pos, c->pg, t8, c->psidx.as<int>(),
                                                                                    c->spl.as<real>(), c->loc.as<int>(),
                                                                                    c->loc16.as<int>())));
    LAUNCH_CHECK();
    return PGM_OK;
}

static int make_plans(pgm_ctx *c) {
    static bool setup_done = false;
    if (!setup_done) {
        FFT_TRY(rocfft_setup());
        setup_done = true;
    }
    // one batched plan per direction over the beads' grids: real grids of K1 K2 K3 points and half
    // spectra of K1 K2 (K3/2 + 1) complex points, each bead's after the previous one
    const size_t K1 = c->s.grid[0], K2 = c->s.grid[1], K3 = c->s.grid[2], K3h = K3 / 2 + 1;
    const size_t len[3] = {K3, K2, K1};
    const size_t rstr[3] = {1, K3, K3 * K2}, cstr[3] = {1, K3h, K3h * K2};
    const rocfft_precision prec = c->mixed ? rocfft_precision_single : rocfft_precision_double;
    const size_t nb = (size_t)c->nbead;
    rocfft_plan_description df = nullptr, db = nullptr;
    FFT_TRY(rocfft_plan_description_create(&df));
    FFT_TRY(rocfft_plan_description_create(&db));
    int rc = PGM_OK;
    if (rocfft_plan_description_set_data_layout(df, rocfft_array_type_real, rocfft_array_type_hermitian_interleaved, nullptr,
                                                nullptr, 3, rstr, c->ngrid, 3, cstr, c->ncplx) != rocfft_status_success ||
        rocfft_plan_description_set_data_layout(db, rocfft_array_type_hermitian_interleaved, rocfft_array_type_real, nullptr,
                                                nullptr, 3, cstr, c->ncplx, 3, rstr, c->ngrid) != rocfft_status_success ||
        rocfft_plan_create(&c->fwd, rocfft_placement_notinplace, rocfft_transform_type_real_forward, prec, 3, len, nb, df) !=
            rocfft_status_success ||
        rocfft_plan_create(&c->bwd, rocfft_placement_notinplace, rocfft_transform_type_real_inverse, prec, 3, len, nb, db) !=
            rocfft_status_success) {
        g_err = "rocFFT plan creation failed";
        rc = PGM_ERR_FFT;
    }
    rocfft_plan_description_destroy(df);
    rocfft_plan_description_destroy(db);
    if (rc < 0) return rc;
    size_t w1 = 0, w2 = 0;
    FFT_TRY(rocfft_plan_get_work_buffer_size(c->fwd, &w1));
    FFT_TRY(rocfft_plan_get_work_buffer_size(c->bwd, &w2));
    PGM_TRY(c->fft_work.ensure(std::max(w1, w2)));
    FFT_TRY(rocfft_execution_info_create(&c->info));
    if (c->fft_work.bytes) FFT_TRY(rocfft_execution_info_set_work_buffer(c->info, c->fft_work.p, c->fft_work.bytes));
    FFT_TRY(rocfft_execution_info_set_stream(c->info, c->stream));
    return PGM_OK;
}

// Fused transform of the CG field evaluations (fft.h): every grid axis a power of two in 8..256,
// the plane of K2 (K3/2 + 1) complex values within 160 KB of LDS, no Ewald sum.  Twiddles
// W_N^t = exp(-2 pi i t / N) for N = K3/2, K3, K2, K1 in FP64, rounded to the context's precision.
template <typename real> static int fused_lds(pgm_ctx *c) {
    const int plane = (int)pfft_plane_lds<real>(c->pfg), xc = (int)pfft_xconv_lds<real>(c->pfg);
    HIP_TRY(hipFuncSetAttribute((const void *)k_pme_fwd_plane<real>, hipFuncAttributeMaxDynamicSharedMemorySize, plane));
    HIP_TRY(hipFuncSetAttribute((const void *)k_pme_bwd_plane<real>, hipFuncAttributeMaxDynamicSharedMemorySize, plane));
    HIP_TRY(hipFuncSetAttribute((const void *)k_pme_x_conv<real>, hipFuncAttributeMaxDynamicSharedMemorySize, xc));
    HIP_TRY(hipFuncSetAttribute((const void *)k_pme_x_conv<real, true>, hipFuncAttributeMaxDynamicSharedMemorySize, xc));
    std::vector<real> tw;
    int *offs[4] = {&c->pfg.twM, &c->pfg.twK3, &c->pfg.twK2, &c->pfg.twK1};
    const int lens[4] = {c->pfg.M, c->pfg.K3, c->pfg.K2, c->pfg.K1};
    for (int a = 0; a < 4; ++a) {
        *offs[a] = (int)(tw.size() / 2);
        for (int t = 0; t < lens[a]; ++t) {
            const double ang = 2.0 * M_PI * (double)t / (double)lens[a];
            tw.push_back((real)std::cos(ang));
            tw.push_back((real)-std::sin(ang));
        }
    }
    PGM_TRY(upload(c->pfft_tw, tw.data(), tw.size(), c->stream));
    return PGM_OK;
}

static int setup_fused(pgm_ctx *c) {
    c->fused_ok = false;
    c->pme_fused = false;
    if (c->ewald) return PGM_OK;
    const int *K = c->s.grid;
    for (int a = 0; a < 3; ++a)
        if (K[a] < 8 || K[a] > 256 || (K[a] & (K[a] - 1))) return PGM_OK;
    const PFFTGeom f{K[0], K[1], K[2], K[2] / 2, K[2] / 2 + 1, 0, 0, 0, 0};
    const size_t plane = c->mixed ? pfft_plane_lds<float>(f) : pfft_plane_lds<double>(f);
    if (plane > 160 * 1024) return PGM_OK;
    c->pfg = f;
    PGM_TRY(c->mixed ? fused_lds<float>(c) : fused_lds<double>(c));
    c->fused_ok = true;
    c->pme_fused = true;
    return PGM_OK;
}

// ---------------------------------------------------------------------------- rows
__global__ void k_pack_xr(int n, const double *__restrict__ pos, const double *__restrict__ R2, double4 *__restrict__ xr) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    xr[i] = make_double4(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2], R2[i]);
}

static int alloc_rows(pgm_ctx *c, int cap) {
    drop_graphs(c);
    c->cap_full = cap;
    c->cap = cap;
    PGM_TRY(c->rows_full.ensure((size_t)c->n * cap * sizeof(int)));
    PGM_TRY(c->rows.ensure((size_t)c->n * cap * sizeof(int)));
    const size_t ell_n = (size_t)((cap + RL - 1) / RL) * RL * c->n;
    PGM_TRY(c->jell.ensure(ell_n * sizeof(int)));
    PGM_TRY(c->G1.ensure(ell_n * c->rsize));
    PGM_TRY(c->G2.ensure(ell_n * c->rsize));
    return PGM_OK;
}

// Molecule-block buffers for candidate capacity mcap_full (records rounded up to rlm lanes).
static int alloc_mol(pgm_ctx *c, int mcap_full) {
    drop_graphs(c);
    const int nmol = c->n / 3;
    c->mcap_full = mcap_full;
    c->mcap = (mcap_full + c->rlm - 1) / c->rlm * c->rlm;
    PGM_TRY(c->mrows.ensure(sizeof(int) * (size_t)nmol * mcap_full));
    PGM_TRY(c->mcount_full.ensure(sizeof(int) * nmol));
    PGM_TRY(c->mcount.ensure(sizeof(int) * nmol));
    PGM_TRY(c->mj.ensure(sizeof(int) * (size_t)nmol * c->mcap));
    PGM_TRY(c->mg.ensure(c->rsize * MOL_REC * (size_t)nmol * c->mcap));
    return PGM_OK;
}

// Candidate partner molecules from the full atom rows (molecule-block layout only).
static int build_mol_rows(pgm_ctx *c) {
    const int nmol = c->n / 3, nmol_bead = c->nper / 3;
    hipStream_t st = c->stream;
    const size_t lds = sizeof(unsigned int) * 4 * (size_t)((nmol_bead + 31) / 32);
    for (int attempt = 0; attempt < 3; ++attempt) {
        k_fill_int<<<1, 64, 0, st>>>(1, c->maxcount.as<int>(), 0);
        k_mol_rows<<<nblk(nmol, 4), 256, lds, st>>>(nmol, nmol_bead, c->rows_full.as<int>(), c->count_full.as<int>(), c->cap_full,
                                                    c->mcap_full, c->mrows.as<int>(), c->mcount_full.as<int>(), c->maxcount.as<int>());
        LAUNCH_CHECK();
        HIP_TRY(hipMemcpyAsync(c->h_pin_int, c->maxcount.p, sizeof(int), hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        const int mc = c->h_pin_int[0];
        if (mc <= c->mcap_full) return PGM_OK;
        PGM_TRY(alloc_mol(c, (int)(mc * 1.2) + 8));
    }
    g_err = "molecule row capacity could not be sized";
    return PGM_ERR_STATE;
}

static int build_rows(pgm_ctx *c, const double *pos) {
    const int n = c->n;
    hipStream_t st = c->stream;
    const int ncell = c->nbead * c->nc.x * c->nc.y * c->nc.z;
    PGM_TRY(c->cell_start.ensure(sizeof(int) * ncell));
    PGM_TRY(c->cell_end.ensure(sizeof(int) * ncell));
    k_cell_keys<<<nblk(n, 256), 256, 0, st>>>(n, c->nper, pos, c->box, c->nc, c->key.as<int>(), c->atom.as<int>());
    LAUNCH_CHECK();
    int bits = 1;
    while ((1 << bits) < ncell) ++bits;
    size_t tmp = 0;
    HIP_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, c->key.as<int>(), c->skey.as<int>(), c->atom.as<int>(),
                                               c->satom.as<int>(), n, 0, bits, st));
    PGM_TRY(c->sort_tmp.ensure(tmp));
    HIP_TRY(hipcub::DeviceRadixSort::SortPairs(c->sort_tmp.p, tmp, c->key.as<int>(), c->skey.as<int>(), c->atom.as<int>(),
                                               c->satom.as<int>(), n, 0, bits, st));
    k_fill_int<<<nblk(ncell, 256), 256, 0, st>>>(ncell, c->cell_start.as<int>(), 0);
    k_fill_int<<<nblk(ncell, 256), 256, 0, st>>>(ncell, c->cell_end.as<int>(), 0);
    k_cell_bounds<<<nblk(n, 256), 256, 0, st>>>(n, c->skey.as<int>(), c->cell_start.as<int>(), c->cell_end.as<int>());
    LAUNCH_CHECK();
    const double rl = c->rc_pair + c->s.skin;
    for (int attempt = 0; attempt < 3; ++attempt) {
        k_fill_int<<<1, 64, 0, st>>>(1, c->maxcount.as<int>(), 0);
        k_build_rows_cells<<<ncell, 256, 0, st>>>(ncell / c->nbead, pos, c->box, c->nc, c->satom.as<int>(), c->cell_start.as<int>(),
                                                  c->cell_end.as<int>(), rl * rl, c->cap_full, c->rows_full.as<int>(),
                                                  c->count_full.as<int>(), c->maxcount.as<int>());
        LAUNCH_CHECK();
        HIP_TRY(hipMemcpyAsync(c->h_pin_int, c->maxcount.p, sizeof(int), hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        const int mc = c->h_pin_int[0];
        if (mc <= c->cap_full) {
            if (c->layout == 1) PGM_TRY(build_mol_rows(c));
            c->rows_valid = true;
            c->pinv_stale = true;
            PGM_TRY(c->pos_build.ensure(sizeof(double) * 3 * (size_t)n));
            HIP_TRY(hipMemcpyAsync(c->pos_build.p, pos, sizeof(double) * 3 * (size_t)n, hipMemcpyDeviceToDevice, st));
            c->rebuilds += 1;
            return PGM_OK;
        }
        PGM_TRY(alloc_rows(c, (int)(mc * 1.2) + 16));
    }
    g_err = "row capacity could not be sized";
    return PGM_ERR_STATE;
}

// Atom rows compacted to the cutoff (rows, count, count_near) for the current geometry (xr packed).
static int compact_atom_rows(pgm_ctx *c) {
    const int n = c->n;
    k_compact_rows<<<nblk(n, 4), 256, 0, c->stream>>>(n, c->xr.as<double4>(), c->box, c->rows_full.as<int>(),
                                                      c->count_full.as<int>(), c->cap_full, c->rc_pair * c->rc_pair,
                                                      c->s.ewald_beta * c->s.ewald_beta, c->cap, c->rows.as<int>(),
                                                      c->count.as<int>(), c->count_near.as<int>());
    LAUNCH_CHECK();
    c->atom_rows_fresh = true;
    return PGM_OK;
}

// Pack positions and compact the rows to the cutoff (every force evaluation).  With molecule
// blocks the records are the compacted form; the atom rows are compacted only when asked for.
// defer_records (the solve path): with molecule blocks the record pass runs on the side stream, where
// the first field evaluation's row field follows it in stream order, while the context stream
// goes on with the binning, the permanent dipoles and the PME chain; later consumers wait for ev_rec
// (rec_pending).
static int prepare(pgm_ctx *c, const double *pos, bool defer_records = false) {
    if (c->rec_pending) {  // records of an earlier evaluation still pending on the side stream
        HIP_TRY(hipStreamWaitEvent(c->stream, c->ev_rec, 0));
        c->rec_pending = false;
    }
    if (!c->rows_valid) PGM_TRY(build_rows(c, pos));
    // Captured CG iterations of the per-atom PME and Ewald paths hold the position pointer as a
    // kernel argument: positions at another address need new captures.
    if (pos != c->pos) drop_graphs(c);
    c->pos = pos;
    const int n = c->n;
    hipStream_t st = c->stream;
    k_pack_xr<<<nblk(n, 256), 256, 0, st>>>(n, pos, c->R2.as<double>(), c->xr.as<double4>());
    LAUNCH_CHECK();
    c->atom_rows_fresh = false;
    if (c->layout != 1) PGM_TRY(compact_atom_rows(c));
    if (c->molp && c->pinv_stale) {  // blocks of the geometry at the row build (any SPD blocks precondition)
        c->pinv_stale = false;
        const int nmol = n / 3;
        if (c->mixed)
            k_mol_precond<float><<<nblk(nmol, 128), 128, 0, st>>>(nmol, c->xr.as<double4>(), c->alpha64.as<double>(), c->box,
                                                                  c->pinv.as<float>());
        else
            k_mol_precond<double><<<nblk(nmol, 128), 128, 0, st>>>(nmol, c->xr.as<double4>(), c->alpha64.as<double>(), c->box,
                                                                   c->pinv.as<double>());
        LAUNCH_CHECK();
    }
    auto bin = [&]() -> int { return c->mixed ? pme_bin<float>(c, pos) : pme_bin<double>(c, pos); };
    // the PME binning (key, sort, tile bounds, splines) depends on the positions only: with molecule
    // blocks it runs on the side stream beside the record pass (enqueued after it, so that the record
    // pass starts while the host still enqueues the binning's dozen launches), joined before return
    const bool fork_bin = c->layout == 1 && c->tiled && !c->ewald;
    const bool side_records = fork_bin && defer_records;
    if (fork_bin) HIP_TRY(hipEventRecord(c->ev_fork, st));
    if (side_records) HIP_TRY(hipStreamWaitEvent(c->side, c->ev_fork, 0));
    if (c->layout == 1) {
        const int nmol = n / 3;
        const double rc2 = c->rc_pair * c->rc_pair, rce2 = c->rc_e * c->rc_e;
        // The force pass runs on these records unless LJ reaches beyond the electrostatic cutoff; then
        // (FP32) the same pass compacts the atom rows of the force pass (k_compact_rows not needed).
        // Other users of the atom rows (strain derivative, parameter gradients) compact them when asked.
        const double beta2 = c->s.ewald_beta * c->s.ewald_beta;
        hipStream_t rs = side_records ? c->side : st;
        const bool rows_too = c->mixed && c->rc_v > c->rc_e;
        if (rows_too)
            k_mol_tensors<float, true><<<nblk(nmol, 4), 256, 0, rs>>>(nmol, c->mrows.as<int>(), c->mcount_full.as<int>(), c->mcap_full,
                                                               c->mcap, c->rlm, c->xr.as<double4>(), c->box, rc2, rce2,
                                                               (float)c->s.ewald_beta, c->mj.as<int>(), c->mg.as<float>(),
                                                               c->mcount.as<int>(), c->rows.as<int>(), c->cap,
                                                               c->count.as<int>(), c->count_near.as<int>(), beta2);
        else if (c->mixed)
            k_mol_tensors<float, false><<<nblk(nmol, 4), 256, 0, rs>>>(
                nmol, c->mrows.as<int>(), c->mcount_full.as<int>(), c->mcap_full, c->mcap, c->rlm, c->xr.as<double4>(), c->box,
                rc2, rce2, (float)c->s.ewald_beta, c->mj.as<int>(), c->mg.as<float>(), c->mcount.as<int>(), nullptr, 0, nullptr,
                nullptr, beta2);
        else
            k_mol_tensors<double, false><<<nblk(nmol, 4), 256, 0, rs>>>(
                nmol, c->mrows.as<int>(), c->mcount_full.as<int>(), c->mcap_full, c->mcap, c->rlm, c->xr.as<double4>(), c->box,
                rc2, rce2, c->s.ewald_beta, c->mj.as<int>(), c->mg.as<double>(), c->mcount.as<int>(), nullptr, 0, nullptr, nullptr,
                beta2);
        c->atom_rows_fresh = rows_too;
        if (side_records) {
            LAUNCH_CHECK();
            HIP_TRY(hipEventRecord(c->ev_rec, c->side));
            c->rec_pending = true;
        }
    } else if (c->mixed)
        k_pair_tensors<float><<<nblk((long long)n * RL, 256), 256, 0, st>>>(
            n, c->rows.as<int>(), c->count.as<int>(), c->count_near.as<int>(), c->cap, c->xr.as<double4>(), c->box,
            (float)c->s.ewald_beta, c->rc_e * c->rc_e, c->jell.as<int>(), c->G1.as<float>(), c->G2.as<float>());
    else
        k_pair_tensors<double><<<nblk((long long)n * RL, 256), 256, 0, st>>>(
            n, c->rows.as<int>(), c->count.as<int>(), c->count_near.as<int>(), c->cap, c->xr.as<double4>(), c->box,
            c->s.ewald_beta, c->rc_e * c->rc_e, c->jell.as<int>(), c->G1.as<double>(), c->G2.as<double>());
    LAUNCH_CHECK();
    if (side_records) return bin();  // on the context stream, beside the records
    if (fork_bin) {
        HIP_TRY(hipStreamWaitEvent(c->side, c->ev_fork, 0));
        c->stream = c->side;
        const int rb = bin();
        c->stream = st;
        PGM_TRY(rb);
        HIP_TRY(hipEventRecord(c->ev_join, c->side));
        HIP_TRY(hipStreamWaitEvent(st, c->ev_join, 0));
        return PGM_OK;
    }
    if (c->ewald) return PGM_OK;
    return bin();
}

// Live timing: fold the recorded event pairs into the sums (waits for them), then free the slots.
// Launches queued past the CG's convergence return at once (a few microseconds); a launch counts
// only if it took at least a fifth of the longest one of its batch.
static int harvest_timing(pgm_ctx *c) {
    for (auto &t : c->timed) {
        std::vector<float> d((size_t)t.used);
        float mx = 0.f;
        for (int k = 0; k < t.used; ++k) {
            HIP_TRY(hipEventSynchronize(t.ev[2 * k + 1]));
            HIP_TRY(hipEventElapsedTime(&d[k], t.ev[2 * k], t.ev[2 * k + 1]));
            mx = std::max(mx, d[k]);
        }
        for (int k = 0; k < t.used; ++k)
            if (d[k] >= 0.2f * mx) {
                t.sum_ms += d[k];
                t.count += 1;
            }
        t.used = 0;
    }
    return PGM_OK;
}

static constexpr int TIMED_CAP = 1024;

// The event pair of the next launch of kernel slot `which` ({nullptr, nullptr} when timing is off).
// The timed kernels are launched with the pair attached (hipExtLaunchKernel: the kernel's own start
// and end), so the timing adds no marker packets to the stream: a separate hipEventRecord before and
// after each timed launch put a wait of several microseconds between the kernels of the CG chain.
struct TimedPair {
    hipEvent_t e0 = nullptr, e1 = nullptr;
};
static TimedPair timing_pair(pgm_ctx *c, int which) {
    if (!c->timing) return {};
    auto &t = c->timed[which];
    if (t.ev.empty()) {
        t.ev.resize(2 * TIMED_CAP);
        for (auto &e : t.ev)
            if (hipEventCreate(&e) != hipSuccess) return {};
    }
    if (t.used == TIMED_CAP && harvest_timing(c) < 0) return {};
    const int k = t.used++;
    return {t.ev[2 * k], t.ev[2 * k + 1]};
}

// Launch kernel k, with the event pair attached when it is set.
template <typename... KArgs, typename... Args>
static void launch_timed(TimedPair tp, void (*k)(KArgs...), dim3 grid, dim3 block, size_t lds, hipStream_t st, Args... args) {
    if (tp.e0)
        hipExtLaunchKernelGGL(k, grid, block, (std::uint32_t)lds, st, tp.e0, tp.e1, 0u, args...);
    else
        k<<<grid, block, lds, st>>>(args...);
}

// ---------------------------------------------------------------------------- PME dispatch
// The fused transform of the merged spread blocks: phi = c2r(Gc r2c(Q)) into c->grid (fft.h); with
// epart also the reciprocal energy partials, conv_partials(c) per bead.
template <typename real>
static int fused_transform(pgm_ctx *c, const int *conv, int iter, hipStream_t st, const std::function<int()> *fork = nullptr,
                           double *epart = nullptr) {
    using C = cplx<real>;
    const PFFTGeom &f = c->pfg;
    const int B = c->nbead;
    const size_t plane = pfft_plane_lds<real>(f);
    const int nth = f.K2 * f.K3 >= 8192 ? pfft_plane_threads<real>() : 256;
    const C *tw = c->pfft_tw.as<C>();
    if (fork && c->fork_stage == 3) HIP_TRY(hipStreamWaitEvent(st, c->ev_join, 0));  // planes after the row field
    k_pme_fwd_plane<real><<<f.K1 * B, nth, plane, st>>>(f, (size_t)c->tiles8.ntiles * c->tiles8.bsize, c->merge_tab.as<int>(),
                                                         c->blocks.as<real>(), c->cgrid.as<C>(), tw, conv, iter);
    if (fork && c->fork_stage == 2) PGM_TRY((*fork)());
    const dim3 gx(nblk((long long)f.K2 * f.K3h, PFFT_COLS), B);
    if (epart)
        k_pme_x_conv<real, true><<<gx, 256, pfft_xconv_lds<real>(f), st>>>(f, c->cgrid.as<C>(), c->Gc.as<real>(), tw + f.twK1,
                                                                          conv, iter, epart);
    else
        k_pme_x_conv<real><<<gx, 256, pfft_xconv_lds<real>(f), st>>>(f, c->cgrid.as<C>(), c->Gc.as<real>(), tw + f.twK1, conv,
                                                                     iter);
    k_pme_bwd_plane<real><<<f.K1 * B, nth, plane, st>>>(f, c->cgrid.as<C>(), c->grid.as<real>(), tw, conv, iter);
    LAUNCH_CHECK();
    return PGM_OK;
}

// The PME transform runs fused (fft.h) when the grid allows it and no strain derivative is asked for.
static bool use_fused(const pgm_ctx *c, bool vir) { return c->tiled && c->pme_fused && c->fused_ok && !vir; }

// Reciprocal energy partials per bead that spread_fft(energy = true) leaves in c->conv_part (vir:
// with the strain derivative).
static int conv_partials(const pgm_ctx *c, bool vir) {
    if (c->ewald) return nblk(c->nk, EW_TILE);
    return use_fused(c, vir) ? nblk((long long)c->pfg.K2 * c->pfg.K3h, PFFT_COLS) : nblk((long long)c->ncplx, 256);
}

// fork: called once the spread (fork_stage 1 and 4) or the forward planes (fork_stage 2) are enqueued.
template <typename real>
static int spread_fft(pgm_ctx *c, const real4<real> *dq, bool withq, bool energy, const int *conv, int iter,
                      double *vir_part = nullptr, const std::function<int()> *fork = nullptr) {
    const int n = c->n;
    hipStream_t st = c->stream;
    const int P = c->s.order;
    const TimedPair tp = energy ? TimedPair{} : timing_pair(c, 1);  // the spread kernel
    const int B = c->nbead;
    if (c->tiled) {
        const PMETiles &t = c->tiles8;
        const size_t lds = spread_lds_bytes(c);
        if (withq) {
            ORDER_SWITCH(P, launch_timed(tp, k_spread_tiles<real, PO, true>, dim3(B * t.ntiles), dim3(WAVE), lds, st, t, c->pg,
                                         c->tstart.as<int>(), c->tend.as<int>(), c->psidx.as<int>(), c->spl.as<real>(),
                                         c->loc.as<int>(), dq, c->blocks.as<real>(), conv, iter));
        } else {
            ORDER_SWITCH(P, launch_timed(tp, k_spread_tiles<real, PO, false>, dim3(B * t.ntiles), dim3(WAVE), lds, st, t, c->pg,
                                         c->tstart.as<int>(), c->tend.as<int>(), c->psidx.as<int>(), c->spl.as<real>(),
                                         c->loc.as<int>(), dq, c->blocks.as<real>(), conv, iter));
        }
        LAUNCH_CHECK();
        if (use_fused(c, vir_part != nullptr)) {  // merge + transforms + Gc fused (fft.h)
            if (fork && (c->fork_stage == 1 || c->fork_stage == 4)) PGM_TRY((*fork)());
            return fused_transform<real>(c, conv, iter, st, fork, energy ? c->conv_part.as<double>() : nullptr);
        }
        if (fork && (c->fork_stage == 1 || c->fork_stage == 4)) PGM_TRY((*fork)());
        const int *K = c->s.grid;
        k_merge_table<real><<<dim3(nblk(K[2], 128), K[1], K[0] * B), 128, 0, st>>>(
            K[0], K[1], K[2], (size_t)t.ntiles * t.bsize, c->merge_tab.as<int>(), c->blocks.as<real>(), c->grid.as<real>(), conv,
            iter);
    } else {
    k_fill<real><<<nblk((long long)c->ngrid, 256), 256, 0, st>>>(c->ngrid, c->grid.as<real>(), real(0));
    const dim3 gs(nblk((long long)n * P, 256));
    if (withq) {
        ORDER_SWITCH(P, launch_timed(tp, k_spread<real, PO, true>, gs, dim3(256), 0, st, n, c->pos, dq, c->pg, c->grid.as<real>(),
                                     conv, iter));
    } else {
        ORDER_SWITCH(P, launch_timed(tp, k_spread<real, PO, false>, gs, dim3(256), 0, st, n, c->pos, dq, c->pg,
                                     c->grid.as<real>(), conv, iter));
    }
    }
    LAUNCH_CHECK();
    void *in[1] = {c->grid.p}, *out[1] = {c->cgrid.p};
    FFT_TRY(rocfft_execute(c->fwd, in, out, c->info));
    const int K3h = c->s.grid[2] / 2 + 1;
    if (vir_part) {  // influence-function part of the strain derivative, from the unconvolved transform
        RecipGeom rg{};
        for (int k = 0; k < 9; ++k) rg.hinv[k] = c->pg.hinv[k];
        for (int a = 0; a < 3; ++a) rg.K[a] = c->s.grid[a];
        rg.pb2 = M_PI * M_PI / (c->s.ewald_beta * c->s.ewald_beta);
        k_convolve_virial<real><<<nblk((long long)c->ncplx, 256), 256, 0, st>>>(c->ncplx, rg, c->cgrid.as<real>(),
                                                                                  c->Gc.as<real>(), vir_part);
        LAUNCH_CHECK();
    }
    const dim3 cgridB(nblk((long long)c->ncplx, 256), B);
    if (energy)
        k_convolve<real, true><<<cgridB, 256, 0, st>>>(c->ncplx, K3h, c->s.grid[2] % 2 == 0, c->cgrid.as<real>(), c->Gc.as<real>(),
                                                       c->conv_part.as<double>(), conv, iter);
    else
        k_convolve<real, false><<<cgridB, 256, 0, st>>>(c->ncplx, K3h, c->s.grid[2] % 2 == 0, c->cgrid.as<real>(),
                                                        c->Gc.as<real>(), nullptr, conv, iter);
    LAUNCH_CHECK();
    void *in2[1] = {c->cgrid.p}, *out2[1] = {c->grid.p};
    FFT_TRY(rocfft_execute(c->bwd, in2, out2, c->info));
    return PGM_OK;
}

template <typename real> static int spread_fft_only(pgm_ctx *c, const real4<real> *dq) {
    const bool timing = c->timing;
    c->timing = false;
    const int r = spread_fft<real>(c, dq, false, false, nullptr, -1);
    c->timing = timing;
    return r;
}

// One field evaluation of the dipoles (and charges) dq: spread -> FFT -> convolve -> FFT, the row
// field, then the gather combined with the row field according to `mode` (pme.h).
static double c_self(const pgm_ctx *c) { return 4.0 * std::pow(c->s.ewald_beta, 3) / (3.0 * SQRT_PI); }

// Number of per-block partials per bead a field evaluation writes (MODE_FIELD / MODE_AV).
static int gather_parts(const pgm_ctx *c) {
    if (c->ewald) return nblk(c->n, EW_TILE);
    return c->tiled ? c->tiles.ntiles : nblk(c->n, 256);
}

// Grid of a per-atom kernel with per-block partials: (blocks per bead, beads).
static dim3 bead_grid(const pgm_ctx *c, long long per_bead_threads) { return dim3(nblk(per_bead_threads, 256), c->nbead); }

// PME gather + combine (tiled through LDS when the grid is tiled).
template <typename real>
static int gather_combine(pgm_ctx *c, const real4<real> *dq, int mode, const real4<real> *extra, real4<real> *out, double *part,
                          const int *conv, int iter, TimedPair tp = {}) {
    const int n = c->n;
    hipStream_t st = c->stream;
    const real cself = (real)c_self(c);
    if (c->tiled) {
        const PMETiles &t = c->tiles;
        ORDER_SWITCH(c->s.order, launch_timed(tp, k_gather_tiles<real, PO>, dim3(c->nbead * t.ntiles), dim3(GATHER_THREADS),
                                              c->rsize * t.bsize, st, t, c->pg, c->tstart16.as<int>(), c->tend16.as<int>(),
                                              c->psidx.as<int>(), c->spl.as<real>(), c->loc16.as<int>(), c->grid.as<real>(), dq,
                                              c->vrow.as<real4<real>>(), c->alpha_r.as<real>(), extra, cself, mode, out, part,
                                              conv, iter));
    } else {
        ORDER_SWITCH(c->s.order, launch_timed(tp, k_gather_combine<real, PO>, dim3(nblk(n, 256)), dim3(256), 0, st, n, c->pos,
                                              c->pg, c->grid.as<real>(), dq, c->vrow.as<real4<real>>(), c->alpha_r.as<real>(),
                                              extra, cself, mode, out, part, conv, iter));
    }
    LAUNCH_CHECK();
    return PGM_OK;
}

// The CG row field of dq into vrow on stream st, in the context's layout.
template <typename real, bool WITHQ>
static int row_field_launch(pgm_ctx *c, const real4<real> *dq, const int *conv, int iter, hipStream_t st, TimedPair tp = {},
                            Ext3 ext = {}) {
    using R4 = real4<real>;
    if (c->layout == 1) {
        const int nmol = c->n / 3, nb = c->nper / 3;
        const int *mj = c->mj.as<int>(), *mc = c->mcount.as<int>();
        const real *mg = c->mg.as<real>();
        const double4 *xr = c->xr.as<double4>();
        R4 *out = c->vrow.as<R4>();
        const dim3 b(256);
        switch (c->rlm) {
            case 4: launch_timed(tp, k_row_field_mol<real, WITHQ, 4>, dim3(nblk((long long)nmol * 4, 256)), b, 0, st, nmol, nb, mj, mg, mc, xr, dq, c->box, out, conv, iter, ext); break;
            case 8: launch_timed(tp, k_row_field_mol<real, WITHQ, 8>, dim3(nblk((long long)nmol * 8, 256)), b, 0, st, nmol, nb, mj, mg, mc, xr, dq, c->box, out, conv, iter, ext); break;
            default: launch_timed(tp, k_row_field_mol<real, WITHQ, 16>, dim3(nblk((long long)nmol * 16, 256)), b, 0, st, nmol, nb, mj, mg, mc, xr, dq, c->box, out, conv, iter, ext); break;
        }
    } else {
        launch_timed(tp, k_row_field<real, WITHQ>, dim3(nblk((long long)c->n * RL, 256)), dim3(256), 0, st, c->n, c->nper,
                     c->jell.as<int>(), c->count.as<int>(), c->G1.as<real>(), c->G2.as<real>(), c->xr.as<double4>(), dq, c->box,
                     c->vrow.as<R4>(), conv, iter, ext);
    }
    LAUNCH_CHECK();
    return PGM_OK;
}

// ext: a uniform external field added to the field of every atom (evaluations with charges only).
template <typename real>
static int field_eval(pgm_ctx *c, const real4<real> *dq, bool withq, int mode, const real4<real> *extra, real4<real> *out,
                      double *part, int iter, Ext3 ext = {}) {
    const int n = c->n;
    hipStream_t st = c->stream;
    const int *conv = iter >= 0 ? &c->st.as<CGState>()->conv_it : nullptr;
    // fork: the row field on the side stream, concurrently with spread -> FFT -> convolve -> FFT
    // (from the start, or after the spread / the forward planes: fork_stage)
    // stage 4: no side stream; the row field runs on st after the spread, and no events are recorded
    const bool serial = c->fork_stage == 4 && !c->ewald;
    const std::function<int()> fork = [&]() -> int {
        if (serial) {
            if (c->rec_pending) {  // the records of this geometry come from the side stream
                HIP_TRY(hipStreamWaitEvent(st, c->ev_rec, 0));
                c->rec_pending = false;
            }
            if (withq) return row_field_launch<real, true>(c, dq, conv, iter, st, timing_pair(c, 0), ext);
            return row_field_launch<real, false>(c, dq, conv, iter, st, timing_pair(c, 0));
        }
        HIP_TRY(hipEventRecord(c->ev_fork, st));
        HIP_TRY(hipStreamWaitEvent(c->side, c->ev_fork, 0));
        const TimedPair tp = timing_pair(c, 0);
        if (withq)
            PGM_TRY((row_field_launch<real, true>(c, dq, conv, iter, c->side, tp, ext)));
        else
            PGM_TRY((row_field_launch<real, false>(c, dq, conv, iter, c->side, tp)));
        HIP_TRY(hipEventRecord(c->ev_join, c->side));
        return PGM_OK;
    };
    // stage 3: forked at the start, joined before the forward planes (the row field beside the spread)
    const bool fused = c->tiled && !c->ewald && c->pme_fused && c->fused_ok;
    const bool late = c->tiled && !c->ewald && (c->fork_stage == 1 || c->fork_stage == 4 || (c->fork_stage == 2 && fused));
    if (!late) PGM_TRY(fork());
    const real cself = (real)(4.0 * std::pow(c->s.ewald_beta, 3) / (3.0 * SQRT_PI));
    if (c->ewald) {
        const dim3 gk(nblk(c->nk, EW_TILE)), bk(EW_TILE);
        const TimedPair tsk = timing_pair(c, 1);
        if (withq)
            launch_timed(tsk, k_ewald_sk<real, true>, gk, bk, 0, st, c->nk, c->kvec.as<double4>(), n, c->pos, dq,
                         c->sk.as<double2>(), (double *)nullptr, conv, iter);
        else
            launch_timed(tsk, k_ewald_sk<real, false>, gk, bk, 0, st, c->nk, c->kvec.as<double4>(), n, c->pos, dq,
                         c->sk.as<double2>(), (double *)nullptr, conv, iter);
        LAUNCH_CHECK();
        HIP_TRY(hipStreamWaitEvent(st, c->ev_join, 0));
        launch_timed(timing_pair(c, 2), k_ewald_gather_combine<real>, dim3(nblk(n, EW_TILE)), bk, 0, st, n, c->nk,
                     c->kvec.as<double4>(), c->sk.as<double2>(), c->pos, dq, c->vrow.as<real4<real>>(), c->alpha_r.as<real>(),
                     extra, cself, mode, out, part, conv, iter);
        LAUNCH_CHECK();
        return PGM_OK;
    }
    PGM_TRY(spread_fft<real>(c, dq, withq, false, conv, iter, nullptr, late || (fused && c->fork_stage == 3) ? &fork : nullptr));
    if (!serial) {
        HIP_TRY(hipStreamWaitEvent(st, c->ev_join, 0));
        c->rec_pending = false;  // the row field followed the record pass on the side stream
    }
    PGM_TRY(gather_combine<real>(c, dq, mode, extra, out, part, conv, iter, timing_pair(c, 2)));
    return PGM_OK;
}


// ---------------------------------------------------------------------------- learned guess (host)
// Layer sizes GUESS_D -> 128 (x nh) -> GUESS_OUT; flat FP32 parameters: per layer W (in x out, row
// major) then b (out).
static std::vector<int> guess_sizes(const pgm_ctx *c) {
    std::vector<int> sz{GUESS_D};
    for (int l = 0; l < c->guess.nh; ++l) sz.push_back(MLP_W);
    sz.push_back(GUESS_OUT);
    return sz;
}

static int guess_pack(pgm_ctx *c) {
    auto &G = c->guess;
    const std::vector<int> sz = guess_sizes(c);
    hipStream_t st = c->stream;
    long long off = 0, poff = 0, boff = 0;
    for (size_t l = 0; l + 1 < sz.size(); ++l) {
        const int in = sz[l], out = sz[l + 1], S = (in + 15) / 16, T = (out + 15) / 16;
        k_mlp_pack<<<nblk((long long)T * S * 256, 256), 256, 0, st>>>(in, out, S, T, G.params.as<float>() + off,
                                                                       G.packed.as<short>() + poff);
        LAUNCH_CHECK();
        HIP_TRY(hipMemcpyAsync(G.bias.as<float>() + boff, G.params.as<float>() + off + (long long)in * out, sizeof(float) * out,
                               hipMemcpyDeviceToDevice, st));
        off += (long long)in * out + out;
        poff += (long long)T * S * 256;
        boff += out;
    }
    return PGM_OK;
}

template <int NH> static int guess_infer_nh(pgm_ctx *c, const float *x, int n, float *y) {
    constexpr size_t lds = sizeof(short) * mlp_packed_shorts<1, NH>();
    static bool attr = false;
    if (!attr) {
        HIP_TRY(hipFuncSetAttribute((const void *)k_mlp_infer<1, NH>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        attr = true;
    }
    const int tiles = (n + 15) / 16;
    const int grid = std::max(1, std::min(512, (tiles + 3) / 4));
    k_mlp_infer<1, NH><<<grid, 256, lds, c->stream>>>(n, GUESS_D, GUESS_OUT, c->guess.packed.as<short>(), c->guess.bias.as<float>(),
                                                     x, y);
    LAUNCH_CHECK();
    return PGM_OK;
}

// BF16 inference of n feature rows x (device) into y (device, n x GUESS_OUT).
static int guess_infer(pgm_ctx *c, const float *x, int n, float *y) {
    switch (c->guess.nh) {
        case 1: return guess_infer_nh<1>(c, x, n, y);
        case 2: return guess_infer_nh<2>(c, x, n, y);
        case 3: return guess_infer_nh<3>(c, x, n, y);
        case 4: return guess_infer_nh<4>(c, x, n, y);
        default: g_err = "learned guess: 1..4 hidden layers"; return PGM_ERR_ARG;
    }
}

static void gemm(hipStream_t st, int M, int N, int K, const float *A, long sam, long sak, const float *B, long sbk, long sbn,
                 const float *bias, float *C, float *Z, int act) {
    k_gemm<<<dim3((N + 15) / 16, (M + 15) / 16), 256, 0, st>>>(M, N, K, A, sam, sak, B, sbk, sbn, bias, C, Z, act);
}

// FP32 forward of B rows x (B x GUESS_D) keeping pre-activations (zs) and activations (acts) of
// the hidden layers; output y (B x GUESS_OUT).
static int guess_forward(pgm_ctx *c, const float *x, int B, float *y) {
    auto &G = c->guess;
    const std::vector<int> sz = guess_sizes(c);
    hipStream_t st = c->stream;
    const float *h = x;
    long long off = 0;
    for (size_t l = 0; l + 1 < sz.size(); ++l) {
        const int in = sz[l], out = sz[l + 1];
        const float *W = G.params.as<float>() + off, *b = W + (long long)in * out;
        const bool last = l + 2 == sz.size();
        float *Z = last ? nullptr : G.zs.as<float>() + (size_t)l * B * MLP_W;
        float *A = last ? y : G.acts.as<float>() + (size_t)l * B * MLP_W;
        gemm(st, B, out, in, h, in, 1, W, out, 1, b, A, Z, last ? 0 : 1);
        h = A;
        off += (long long)in * out + out;
    }
    LAUNCH_CHECK();
    return PGM_OK;
}

// Gradients of the mean squared error of the last forward (rows x, targets t) into grads; per-block
// partial sums of the squared error into part.
static int guess_backward(pgm_ctx *c, const float *x, const float *t, int B, const float *y) {
    auto &G = c->guess;
    const std::vector<int> sz = guess_sizes(c);
    const int L = (int)sz.size() - 1;
    hipStream_t st = c->stream;
    std::vector<long long> off(L + 1, 0);
    for (int l = 0; l < L; ++l) off[l + 1] = off[l] + (long long)sz[l] * sz[l + 1] + sz[l + 1];
    const size_t cnt = (size_t)B * GUESS_OUT;
    float *dcur = G.delta.as<float>(), *dnext = G.delta2.as<float>();
    k_mse_grad<<<nblk((long long)cnt, 256), 256, 0, st>>>(cnt, y, t, dcur, G.part.as<double>());
    for (int l = L - 1; l >= 0; --l) {
        const int in = sz[l], out = sz[l + 1];
        const float *hin = l == 0 ? x : G.acts.as<float>() + (size_t)(l - 1) * B * MLP_W;
        float *gW = G.grads.as<float>() + off[l], *gb = gW + (long long)in * out;
        gemm(st, in, out, B, hin, 1, in, dcur, out, 1, nullptr, gW, nullptr, 0);        // dW = h^T delta
        k_colsum<<<out, 256, 0, st>>>(B, out, dcur, gb);                                 // db = sum delta
        if (l > 0) {
            const float *W = G.params.as<float>() + off[l];
            gemm(st, B, in, out, dcur, out, 1, W, 1, out, nullptr, dnext, nullptr, 0);   // delta W^T
            k_dsilu_mul<<<nblk((long long)B * in, 256), 256, 0, st>>>((size_t)B * in, G.zs.as<float>() + (size_t)(l - 1) * B * MLP_W,
                                                                     dnext);
            std::swap(dcur, dnext);
        }
    }
    LAUNCH_CHECK();
    return PGM_OK;
}

static int guess_alloc_batch(pgm_ctx *c, int B) {
    auto &G = c->guess;
    PGM_TRY(G.xb.ensure(sizeof(float) * (size_t)B * GUESS_D));
    PGM_TRY(G.yb.ensure(sizeof(float) * (size_t)B * GUESS_OUT));
    PGM_TRY(G.idx.ensure(sizeof(int) * (size_t)B));
    PGM_TRY(G.acts.ensure(sizeof(float) * (size_t)G.nh * B * MLP_W));
    PGM_TRY(G.zs.ensure(sizeof(float) * (size_t)G.nh * B * MLP_W));
    PGM_TRY(G.delta.ensure(sizeof(float) * (size_t)B * MLP_W));
    PGM_TRY(G.delta2.ensure(sizeof(float) * (size_t)B * MLP_W));
    PGM_TRY(G.y.ensure(sizeof(float) * (size_t)std::max(B, c->n) * GUESS_OUT));
    PGM_TRY(G.part.ensure(sizeof(double) * (nblk((long long)std::max(B * GUESS_OUT, c->n), 256) + 16)));
    return PGM_OK;
}

static const GuessCoef MU4_COEF{{4.0, -6.0, 4.0, -1.0, 0.0, 0.0, 0.0, 0.0}};

// Before the solve (predictor 4, ring full): features of every atom, the network's output when
// trained, and x0 by the selector's mode: the fitted linear predictor (+ the network), or mu4.
// x0 of the learned guess.  The features and the network run when the selector uses the network
// (mode 2) or when this solve is harvested (the harvest scores the network and stores features).
template <typename real> static int guess_x0(pgm_ctx *c, const double *pos, real4<real> *x0, bool harvest) {
    auto &G = c->guess;
    const int n = c->n;
    hipStream_t st = c->stream;
    G.have_y = false;
    if (harvest || G.mode == 2) {
        k_guess_features<<<nblk(n, 256), 256, 0, st>>>(n, pos, G.ring.as<double>(), G.head, 1.0 / G.scale, G.feat.as<float>(),
                                                       G.frames.as<double>());
        LAUNCH_CHECK();
    }
    if (G.trained && (harvest || G.mode == 2)) {
        PGM_TRY(guess_infer(c, G.feat.as<float>(), n, G.y.as<float>()));
        G.have_y = true;
    }
    const GuessCoef &cf = G.mode >= 1 ? G.lin : MU4_COEF;
    k_guess_x0<real><<<nblk(n, 256), 256, 0, st>>>(n, G.ring.as<double>(), G.head, cf, G.frames.as<double>(), G.y.as<float>(),
                                                  G.sigma, G.mode == 2 && G.have_y ? 1 : 0, x0);
    LAUNCH_CHECK();
    return PGM_OK;
}

// Least squares (Gram + ridge) for the linear predictor; Gaussian elimination with pivoting.
static bool solve_small(int K, const double *A0, const double *b0, double *x) {
    std::vector<double> A(A0, A0 + K * K), b(b0, b0 + K);
    double tr = 0.0;
    for (int k = 0; k < K; ++k) tr += A[k * K + k];
    for (int k = 0; k < K; ++k) A[k * K + k] += 1e-13 * tr / K;
    for (int col = 0; col < K; ++col) {
        int piv = col;
        for (int r = col + 1; r < K; ++r)
            if (std::fabs(A[r * K + col]) > std::fabs(A[piv * K + col])) piv = r;
        if (std::fabs(A[piv * K + col]) < 1e-300) return false;
        if (piv != col) {
            for (int k = 0; k < K; ++k) std::swap(A[col * K + k], A[piv * K + k]);
            std::swap(b[col], b[piv]);
        }
        for (int r = col + 1; r < K; ++r) {
            const double f = A[r * K + col] / A[col * K + col];
            for (int k = col; k < K; ++k) A[r * K + k] -= f * A[col * K + k];
            b[r] -= f * b[col];
        }
    }
    for (int r = K - 1; r >= 0; --r) {
        double v = b[r];
        for (int k = r + 1; k < K; ++k) v -= A[r * K + k] * x[k];
        x[r] = v / A[r * K + r];
    }
    return true;
}

// After the solve of a harvest step (ring full, before the ring push): statistics of the ring
// against mu_t (one sync), the refit of the linear predictor, the network's training rows, and the
// selector's choice for the following steps.
static int guess_harvest(pgm_ctx *c, const pgm_induction *ind) {
    auto &G = c->guess;
    const int n = c->n;
    hipStream_t st = c->stream;
    const int nb = nblk(n, 256);
    PGM_TRY(G.part.ensure(sizeof(double) * (size_t)GUESS_STATS * nb));
    PGM_TRY(G.scal.ensure(sizeof(double) * GUESS_STATS));
    const long long row0 = G.sigma > 0.0 ? G.rows : -1;
    k_guess_stats<<<nb, 256, 0, st>>>(n, G.ring.as<double>(), G.head, G.lin, MU4_COEF, ind->mu, G.frames.as<double>(),
                                      G.y.as<float>(), G.sigma, G.have_y ? 1 : 0, G.feat.as<float>(), row0, G.cap, G.bx.as<float>(),
                                      G.by.as<float>(), G.part.as<double>());
    k_dense_reduce<<<1, 256, 0, st>>>(G.part.as<double>(), nb, GUESS_STATS, G.scal.as<double>());
    LAUNCH_CHECK();
    double h[GUESS_STATS];
    HIP_TRY(hipMemcpyAsync(h, G.scal.p, sizeof(h), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    const double inv = 1.0 / (3.0 * n);
    G.err_lin = h[44] * inv;
    G.err_mu4 = h[45] * inv;
    G.err_net = G.have_y ? h[46] * inv : 0.0;
    // selector for the following steps (errors of this harvest, measured before the refit)
    int mode = 0;
    if (G.nfit > 0 && G.err_lin < 0.8 * G.err_mu4) mode = 1;
    if (mode == 1 && G.have_y && G.err_net < 0.8 * G.err_lin) mode = 2;
    G.mode = mode;
    if (row0 >= 0) G.rows += n;
    if (G.sigma <= 0.0) G.sigma = std::sqrt(std::max(G.err_lin, 1e-300));
    // refit: exponentially weighted normal equations over the harvests
    int q = 0;
    for (int j = 0; j < GUESS_K; ++j)
        for (int k = j; k < GUESS_K; ++k, ++q) {
            G.gram[j * GUESS_K + k] = 0.9 * G.gram[j * GUESS_K + k] + h[q];
            G.gram[k * GUESS_K + j] = G.gram[j * GUESS_K + k];
        }
    for (int j = 0; j < GUESS_K; ++j) G.rhs[j] = 0.9 * G.rhs[j] + h[36 + j];
    double w[GUESS_K];
    if (solve_small(GUESS_K, G.gram, G.rhs, w)) {
        for (int k = 0; k < GUESS_K; ++k) G.lin.w[k] = w[k];
        G.nfit += 1;
    }
    return PGM_OK;
}

// ---------------------------------------------------------------------------- induced dipoles
// Predictor, initial residual (fused or plain) and CG (pgm_jax/md/forcefield.py:1917-2008, :1714-1821).
// The host queues iterations without waiting: up to the iteration count of the last solve it does
// not look at the device at all; for the last two before that count, after queueing iteration k it
// waits for the convergence state written by iteration k-1 (iteration k keeps the GPU busy
// meanwhile, and turns into no-ops if k-1 converged); past the count, iteration k is queued only
// once k-1 has reported that it did not converge.  With `quiet` (the MD loop) nothing else is read back: the
// residual and the normaliser stay on the device (stats slot `slot`).  With beads every bead runs
// its own CG in the same launches; the host waits for the slowest, and iterations / residual are
// the largest over the beads (per bead: bead_it, bead_res).  Bead normalisers stay on the device.
// The CG iterations after the start kernels, on c->stream: iteration k evaluates A p (the field of
// p) and runs the two update kernels, which return at once after convergence; the converged flags
// are read back from about the expected count on, so the host rarely waits.  With `graphs`, one
// captured hipGraph per iteration index.  Returns the iteration count (launched_out: the iterations
// queued) or a negative status.
template <typename real>
static int cg_iterate(pgm_ctx *c, double tol, int expect, bool graphs, int *launched_out) {
    const int B = c->nbead, nper = c->nper;
    hipStream_t st = c->stream;
    using R4 = real4<real>;
    const dim3 gB = bead_grid(c, nper);
    const dim3 gM(nblk(nper, MOLP_BLOCK), c->nbead);
    const int npb = c->molp ? nblk(nper, MOLP_BLOCK) : nblk(nper, 256);
    R4 *x = c->vx.as<R4>(), *r = c->vr.as<R4>(), *z = c->vz.as<R4>(), *p = c->vp.as<R4>(), *Ap = c->vAp.as<R4>();
    CGState *S = c->st.as<CGState>();
    double *rz = c->rz.as<double>(), *err = c->err.as<double>();
    const int max_iter = c->s.max_iter;
    int launched = 0;
    int conv = 0x7fffffff;  // last conv_it seen on the host (largest over the beads)
    // The update kernel mirrors the CG states to pinned host memory (cg_mirror, stamped with the
    // solve's sequence number, which k_cg_start_finish keeps in the device state, so that captured
    // iterations stamp the current solve too, and the slot); iteration counts beyond 16 bits take
    // a copy and an event per check instead.
    const bool mirror = max_iter < 0xffff;
    const unsigned seq = c->cg_seq;
    auto read_conv = [&](int k) -> int {  // queue the read-back of the CG states after iteration k-1 (slot k)
        if (mirror) return PGM_OK;
        HIP_TRY(hipMemcpyAsync(c->h_pin_conv + (size_t)(k & 3) * B, S, sizeof(CGState) * B, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipEventRecord(c->ev_read[k & 3], st));
        return PGM_OK;
    };
    auto wait_conv = [&](int k, int &out) -> int {  // the slowest bead's conv_it after iteration k-1 (slot k)
        int m = 0;
        if (!mirror) {
            HIP_TRY(hipEventSynchronize(c->ev_read[k & 3]));
            for (int b2 = 0; b2 < B; ++b2) m = std::max(m, c->h_pin_conv[(size_t)(k & 3) * B + b2].conv_it);
            out = m;
            return PGM_OK;
        }
        const unsigned long long want = ((unsigned long long)seq << 32) | ((unsigned long long)(k & 0xffff) << 16);
        for (int b2 = 0; b2 < B; ++b2) {
            volatile unsigned long long *w = c->h_mirror + (size_t)(k & 3) * B + b2;
            unsigned long long v = *w;
            for (long polls = 1; (v & ~0xffffull) != want; ++polls) {
                if (polls % 4096 == 0) {  // the stream idle or failed with the state still missing: stop
                    const hipError_t q = hipStreamQuery(st);
                    if (q != hipSuccess && q != hipErrorNotReady) {
                        g_err = std::string("CG state wait: ") + hipGetErrorString(q);
                        return PGM_ERR_HIP;
                    }
                    if (q == hipSuccess && ((v = *w) & ~0xffffull) != want) {
                        g_err = "CG state wait: the stream finished without writing the state";
                        return PGM_ERR_HIP;
                    }
                }
                __builtin_ia32_pause();
                v = *w;
            }
            const int cv = (int)(v & 0xffffull);
            m = std::max(m, cv == 0xffff ? 0x7fffffff : cv);
        }
        out = m;
        return PGM_OK;
    };
    // iteration k on c->stream: the field of p, then the two update kernels (all early-exit once converged)
    auto iteration = [&](int k) -> int {
        hipStream_t cs = c->stream;
        PGM_TRY(field_eval<real>(c, p, false, MODE_AV, nullptr, Ap, c->partA.as<double>(), k));
        if (c->molp)
            k_cg_update1_mol<real><<<gM, MOLP_BLOCK, 0, cs>>>(nper, c->partA.as<double>(), gather_parts(c), rz, c->rzs, k, x, r, z,
                                                               p, Ap, c->alpha_r.as<real>(), c->pinv.as<real>(),
                                                               c->partB.as<double>(), S);
        else
            k_cg_update1<real><<<gB, 256, 0, cs>>>(nper, c->partA.as<double>(), gather_parts(c), rz, c->rzs, k, x, r, z, p, Ap,
                                                   c->alpha_r.as<real>(), c->partB.as<double>(), S);
        k_cg_update2<real><<<gB, 256, 0, cs>>>(nper, c->partB.as<double>(), npb, rz, err, c->rzs, k, z, p, S, tol,
                                               max_iter, mirror ? c->h_mirror : nullptr);
        LAUNCH_CHECK();
        return PGM_OK;
    };
    hipStream_t gs = st ? st : c->gstream;  // the null stream cannot be captured: a private stream, ordered by events
    PGM_TRY(read_conv(0));  // converged before any iteration?
    while (launched < max_iter) {
        const int k = launched;
        if (k >= expect && k > 0) {
            // past the expected count: queue iteration k only once iteration k-1 has reported (a
            // few microseconds of idle GPU against a whole iteration of no-op launches)
            PGM_TRY(wait_conv(k, conv));
            if (conv <= k) break;
        }
        if (graphs) {
            if ((int)c->it_graphs.size() <= k) c->it_graphs.resize(k + 1, nullptr);
            if (gs != st) {
                HIP_TRY(hipEventRecord(c->ev_g0, st));
                HIP_TRY(hipStreamWaitEvent(gs, c->ev_g0, 0));
            }
            if (!c->it_graphs[k]) {  // capture iteration k once (on gs)
                if (gs != st) {
                    c->stream = gs;
                    FFT_TRY(rocfft_execution_info_set_stream(c->info, gs));
                }
                HIP_TRY(hipStreamBeginCapture(gs, hipStreamCaptureModeRelaxed));
                const int rc = iteration(k);
                hipGraph_t g = nullptr;
                const hipError_t ec = hipStreamEndCapture(gs, &g);
                if (gs != st) {
                    c->stream = st;
                    FFT_TRY(rocfft_execution_info_set_stream(c->info, st));
                }
                if (rc < 0) {
                    if (g) (void)hipGraphDestroy(g);
                    return rc;
                }
                if (ec != hipSuccess) {
                    g_err = std::string("hipStreamEndCapture: ") + hipGetErrorString(ec);
                    return PGM_ERR_HIP;
                }
                hipGraphExec_t ex = nullptr;
                const hipError_t ei = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
                (void)hipGraphDestroy(g);
                if (ei != hipSuccess) {
                    g_err = std::string("hipGraphInstantiate: ") + hipGetErrorString(ei);
                    return PGM_ERR_HIP;
                }
                c->it_graphs[k] = ex;
            }
            HIP_TRY(hipGraphLaunch(c->it_graphs[k], gs));
            if (gs != st) {
                HIP_TRY(hipEventRecord(c->ev_g1, gs));
                HIP_TRY(hipStreamWaitEvent(st, c->ev_g1, 0));
            }
        } else {
            PGM_TRY(iteration(k));
        }
        launched = k + 1;
        // slot `launched` is waited on at the next pass (launched >= expect - 2) or after the loop;
        // earlier reads would only add a copy and its gap to the queue
        if (launched >= expect - 2 || launched == max_iter) PGM_TRY(read_conv(launched));
        if ((launched >= expect - 1 && launched <= expect) || launched == max_iter) {
            // wait for the state after iteration k-1 (read slot k) while iteration k runs
            PGM_TRY(wait_conv(k, conv));
            if (conv <= k) break;  // every bead converged before iteration k: k was a no-op
        }
    }
    if (conv > launched) {  // not yet seen converged: the state after the last launched iteration decides
        PGM_TRY(wait_conv(launched, conv));
    }
    *launched_out = launched;
    return std::min(conv, launched);
}

template <typename real>
// trial: the CG starts from ind->mu (no predictor, the right-hand side and its norm evaluated
// afresh), and neither the predictor history nor the learned guess's records change
// (PGMForceField.energy, forcefield.py:2562-2626).
static int solve_impl(pgm_ctx *c, pgm_induction *ind, int *iterations, double *residual, bool quiet = false,
                      bool trial = false) {
    const int n = c->n, B = c->nbead, nper = c->nper;
    hipStream_t st = c->stream;
    using R4 = real4<real>;
    const int nb = nblk(n, 256);
    const dim3 gB = bead_grid(c, nper);
    const dim3 gM(nblk(nper, MOLP_BLOCK), c->nbead);           // molecule-block preconditioner kernels
    const int npb = c->molp ? nblk(nper, MOLP_BLOCK) : nblk(nper, 256);  // start / update1 partials per bead
    R4 *x = c->vx.as<R4>(), *r = c->vr.as<R4>(), *z = c->vz.as<R4>(), *p = c->vp.as<R4>(), *Ap = c->vAp.as<R4>();
    R4 *b = c->vb.as<R4>(), *x0 = c->vx0.as<R4>(), *dq = c->vdq.as<R4>();
    const double *q = c->q.as<double>();
    CGState *S = c->st.as<CGState>();
    double *rz = c->rz.as<double>(), *err = c->err.as<double>();
    // predictor coefficients (newest first)
    double cf[4] = {0, 0, 0, 0};
    int K = 0;
    switch (c->s.predictor) {
        case 1: cf[0] = 3; cf[1] = -3; cf[2] = 1; K = 3; break;
        case 2: cf[0] = 4; cf[1] = -6; cf[2] = 4; cf[3] = -1; K = 4; break;
        case 3: cf[0] = 1; K = 1; break;
        default: K = 0;
    }
    const bool learned = !trial && c->s.predictor == 4 && c->guess.on;
    if (c->s.predictor == 4) {  // learned guess; mu4 / previous-step until the histories are there
        if (ind->count >= 4) {
            K = 4;
        } else {
            cf[0] = 1;
            K = 1;
        }
    }
    const bool have = trial || (K > 0 && ind->count >= K);
    const bool use_net = learned && c->guess.nring >= GUESS_K;
    if (trial) {
        k_mu_to_x0<real><<<nb, 256, 0, st>>>(n, ind->mu, x0);
        LAUNCH_CHECK();
    } else if (use_net) {
        PGM_TRY(guess_x0<real>(c, c->pos, x0, ind->count % c->guess.every == 0));
        if (c->alpha0) {  // the network's guess keeps non-polarizable atoms at 0
            k_mask_alpha0<real><<<nb, 256, 0, st>>>(n, c->alpha_r.as<real>(), x0);
            LAUNCH_CHECK();
        }
    } else if (have) {
        if (c->s.predictor == 4 && K == 4) {
            cf[0] = 4, cf[1] = -6, cf[2] = 4, cf[3] = -1;
        }
        k_predict<real><<<nb, 256, 0, st>>>(n, ind->hist, cf[0], cf[1], cf[2], cf[3], x0);
        LAUNCH_CHECK();
    }
    const bool fused = !trial && have && c->s.fused && (c->s.norm_refresh <= 0 || ind->count % c->s.norm_refresh != 0) &&
                       (B == 1 || c->norm_on_device);
    if (fused) {
        k_pack_dq<real><<<nb, 256, 0, st>>>(n, c->pperm.as<double>(), x0, q, dq);
        LAUNCH_CHECK();
        PGM_TRY(field_eval<real>(c, dq, true, MODE_FUSED, x0, r, nullptr, -1, c->F0));
        if (B == 1 && (!quiet || !c->norm_on_device)) k_set_norm_value<<<1, 64, 0, st>>>(S, ind->norm);
    } else {
        k_pack_dq<real><<<nb, 256, 0, st>>>(n, c->pperm.as<double>(), nullptr, q, dq);
        LAUNCH_CHECK();
        PGM_TRY(field_eval<real>(c, dq, true, MODE_FIELD, nullptr, b, c->partA.as<double>(), -1, c->F0));
        k_set_norm<<<B, 256, 0, st>>>(c->partA.as<double>(), gather_parts(c), 1.0 / (3.0 * nper), S);
        if (!have) k_scale_alpha<real><<<nb, 256, 0, st>>>(n, b, c->alpha_r.as<real>(), x0);
        LAUNCH_CHECK();
        PGM_TRY(field_eval<real>(c, x0, false, MODE_PLAIN, b, r, nullptr, -1));
    }
    // A trial solve tests convergence against its own geometry's normaliser but leaves the state's
    // as it was (forcefield.py:2612-2624): the next solve reloads ind->norm.
    c->norm_on_device = !trial;
    if (c->molp)  // x = x0 in the start kernel
        k_cg_start_mol<real><<<gM, MOLP_BLOCK, 0, st>>>(nper, r, c->alpha_r.as<real>(), c->pinv.as<real>(), z, p,
                                                         c->partB.as<double>(), x0, x);
    else
        k_cg_start<real><<<gB, 256, 0, st>>>(nper, r, c->alpha_r.as<real>(), z, p, c->partB.as<double>(), x0, x);
    k_cg_start_finish<<<B, 256, 0, st>>>(c->partB.as<double>(), npb, S, rz, err, c->rzs, c->s.dipole_tol, c->s.max_iter,
                                         c->h_mirror, ++c->cg_seq);
    LAUNCH_CHECK();
    const int max_iter = c->s.max_iter;
    int launched = 0;
    const int it = cg_iterate<real>(c, c->s.dipole_tol, std::max(1, std::min(c->last_iters, max_iter)),
                                    c->use_graphs && !c->timing, &launched);
    if (it < 0) return it;
    c->last_iters = std::max(1, it);
    double *slot = c->stats.as<double>() + (c->stats_slot % c->stats_cap);
    auto &G = c->guess;
    if (use_net && ind->count % c->guess.every == 0) {  // a harvest reads the ring before mu joins it
        k_cg_finish<real><<<nb, 256, 0, st>>>(n, x, r, c->alpha_r.as<real>(), c->s.peek, ind->mu);
        PGM_TRY(guess_harvest(c, ind));
        if (learned) k_ring_push<<<nblk(3 * n, 256), 256, 0, st>>>(3 * n, ind->mu, G.ring.as<double>(), G.head);
        if (!trial) k_hist_push<<<nblk(3 * n, 256), 256, 0, st>>>(3 * n, ind->mu, ind->hist);
        k_record_residual<<<1, 256, 0, st>>>(B, err, c->rzs, S, launched, c->bead_it.as<int>(), c->bead_res.as<double>(), slot);
    } else {
        k_cg_finish_all<real><<<nb, 256, 0, st>>>(n, x, r, c->alpha_r.as<real>(), c->s.peek, ind->mu,
                                                  learned ? G.ring.as<double>() : nullptr, G.head, trial ? nullptr : ind->hist, B,
                                                  err, c->rzs, S, launched, c->bead_it.as<int>(), c->bead_res.as<double>(),
                                                  slot);
    }
    LAUNCH_CHECK();
    if (learned) {  // the guess's own history ring
        G.head = (G.head + 1) % GUESS_K;
        G.nring = std::min(G.nring + 1, GUESS_K);
    }
    c->stats_slot += 1;
    if (!trial) ind->count += 1;
    if (iterations) *iterations = it;
    if (!quiet) {
        HIP_TRY(hipMemcpyAsync(c->h_pin_dbl, slot, sizeof(double), hipMemcpyDeviceToHost, st));
        HIP_TRY(hipMemcpyAsync(c->h_pin_dbl + 1, &S->norm, sizeof(double), hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        if (residual) *residual = c->h_pin_dbl[0];
        if (!trial) ind->norm = c->h_pin_dbl[1];
        return (it >= max_iter && c->h_pin_dbl[0] > c->s.dipole_tol) ? PGM_NOT_CONVERGED : PGM_OK;
    }
    return (it >= max_iter) ? PGM_NOT_CONVERGED : PGM_OK;
}

// ---------------------------------------------------------------------------- energy and forces
// Energies of the last evaluation, summed over the beads (per bead in h_pin_dbl[32 + 4 b ..]).
static int read_energy(pgm_ctx *c, pgm_energy *energy) {
    double *h = c->h_pin_dbl + 32;
    HIP_TRY(hipMemcpyAsync(h, c->energy.p, 4 * sizeof(double) * c->nbead, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (energy) {
        energy->elec = energy->vdw = energy->total = 0.0;
        for (int b = 0; b < c->nbead; ++b) {
            energy->elec += h[4 * b];
            energy->vdw += h[4 * b + 1];
            energy->total += h[4 * b + 2];
        }
    }
    return PGM_OK;
}

template <typename real>
static int energy_forces_impl(pgm_ctx *c, const double *mu, pgm_energy *energy, double *forces, bool quiet = false,
                              const double *masses = nullptr, double *strain = nullptr, bool molecular = true) {
    const int n = c->n, B = c->nbead, nper = c->nper;
    hipStream_t st = c->stream;
    using R4 = real4<real>;
    const int nb = nblk(n, 256);
    const int npb = nblk(nper, 256);  // pack partials per bead
    R4 *dq = c->vdq.as<R4>();
    if (c->rec_pending) {  // the force pass reads the records (or atom rows) the record pass wrote
        HIP_TRY(hipStreamWaitEvent(st, c->ev_rec, 0));
        c->rec_pending = false;
    }
    k_energy_pack<real><<<bead_grid(c, nper), 256, 0, st>>>(nper, c->pperm.as<double>(), mu, c->q.as<double>(),
                                                            c->alpha64.as<double>(), c->dd.as<double>(), dq,
                                                            c->pack_part.as<double>());
    LAUNCH_CHECK();
    const bool vir = strain != nullptr;
    if (vir && c->efield_on) {
        g_err = "the strain derivative with an external field is not built";
        return PGM_ERR_STATE;
    }
    const Ext3 ext = c->F0;  // zero without a field
    const Ext3 kf{KE * c->F0.x, KE * c->F0.y, KE * c->F0.z};
    if (c->efield_on) {  // the field's dipole sum q r + sum (p + mu), partials per bead
        k_field_dipole<<<bead_grid(c, nper), 256, 0, st>>>(nper, c->pos, c->q.as<double>(), c->dd.as<double>(),
                                                           c->fpart.as<double>());
        LAUNCH_CHECK();
    }
    // the force pass on the molecule-block records (no strain derivative; every LJ pair in a record)
    // or on the atom rows
    const bool mol_forces = c->layout == 1 && !vir && c->rc_v <= c->rc_e;
    if (!mol_forces && !c->atom_rows_fresh) PGM_TRY(compact_atom_rows(c));
    const int npair = mol_forces ? nblk((long long)nper * MOL_FORCE_LANES, 256)
                                 : nblk((long long)nper * rowl<real>(), 256);  // pair partials per bead
    const dim3 gpair(npair, B);
    const int nconv = conv_partials(c, vir);
    if (vir) {
        PGM_TRY(c->com.ensure(sizeof(double) * 3 * (size_t)c->nmol_all));
        PGM_TRY(c->off.ensure(sizeof(double) * 3 * (size_t)n));
        PGM_TRY(c->rec_dd.ensure(sizeof(double) * 3 * (size_t)n));
        PGM_TRY(c->vir_pair.ensure(sizeof(double) * 9 * (size_t)npair));
        PGM_TRY(c->vir_conv.ensure(sizeof(double) * 7 * (size_t)nconv));
        PGM_TRY(c->vir_atoms.ensure(sizeof(double) * 9 * (size_t)nb));
        PGM_TRY(c->vir_cov.ensure(sizeof(double) * 9 * (size_t)(nblk(c->ncov, 256) + 1)));
        PGM_TRY(c->vir_out.ensure(sizeof(double) * 9));
        if (molecular) {
            k_com<<<nblk(c->nmol_all, 256), 256, 0, st>>>(c->nmol_all, c->mol_ptr.as<int>(), c->mol_atoms.as<int>(), c->pos,
                                                           masses, c->com.as<double>());
            LAUNCH_CHECK();
            k_offsets<<<nb, 256, 0, st>>>(n, c->pos, c->mol.as<int>(), c->com.as<double>(), c->off.as<double>());
        } else {  // atomic scaling: every atom its own centre
            k_fill<double><<<nblk(3LL * n, 256), 256, 0, st>>>(3 * (size_t)n, c->off.as<double>(), 0.0);
        }
        LAUNCH_CHECK();
        k_pair_forces<real, true><<<gpair, 256, 0, st>>>(
            nper, c->rows.as<int>(), c->count.as<int>(), c->count_near.as<int>(), c->cap, c->xr.as<double4>(), dq, c->box,
            (real)c->s.ewald_beta, c->rc_e * c->rc_e, c->rc_v * c->rc_v, c->ljp.as<double4>(), c->g_el.as<double>(),
            c->g_lj.as<double>(), c->rowf.as<double>(), c->pair_part.as<double>(), c->off.as<double>(),
            c->vir_pair.as<double>());
    } else {
        // without the strain derivative the pair pass runs on the side stream beside the PME chain,
        // joined before the force gather (which reads its row field)
        hipStream_t ps = st;
        if (!c->ewald) {
            HIP_TRY(hipEventRecord(c->ev_fork, st));
            HIP_TRY(hipStreamWaitEvent(c->side, c->ev_fork, 0));
            ps = c->side;
        }
        if (mol_forces)
            k_pair_forces_mol<real, MOL_FORCE_LANES><<<gpair, 256, 0, ps>>>(
                nper, n / 3, c->rlm, c->mj.as<int>(), c->mg.as<real>(), c->mcount.as<int>(), c->xr.as<double4>(), dq, c->box,
                (real)c->s.ewald_beta, c->rc_e * c->rc_e, c->rc_v * c->rc_v, c->rh.as<double>(), c->se.as<double>(),
                c->g_el.as<double>(), c->g_lj.as<double>(), c->rowf.as<double>(), c->pair_part.as<double>(), ext);
        else
            k_pair_forces<real, false><<<gpair, 256, 0, ps>>>(
                nper, c->rows.as<int>(), c->count.as<int>(), c->count_near.as<int>(), c->cap, c->xr.as<double4>(), dq, c->box,
                (real)c->s.ewald_beta, c->rc_e * c->rc_e, c->rc_v * c->rc_v, c->ljp.as<double4>(), c->g_el.as<double>(),
                c->g_lj.as<double>(), c->rowf.as<double>(), c->pair_part.as<double>(), nullptr, nullptr, ext);
        if (!c->ewald) HIP_TRY(hipEventRecord(c->ev_join, c->side));
    }
    LAUNCH_CHECK();
    if (c->ewald) {
        k_ewald_sk<real, true><<<nconv, EW_TILE, 0, st>>>(c->nk, c->kvec.as<double4>(), n, c->pos, dq, c->sk.as<double2>(),
                                                          c->conv_part.as<double>(), nullptr, -1);
        LAUNCH_CHECK();
        if (vir) {
            k_ewald_virial<<<nconv, 256, 0, st>>>(c->nk, c->kvec.as<double4>(), c->sk.as<double2>(),
                                                   0.25 / (c->s.ewald_beta * c->s.ewald_beta), c->vir_conv.as<double>());
            LAUNCH_CHECK();
        }
        k_ewald_gather_forces<real><<<nblk(n, EW_TILE), EW_TILE, 0, st>>>(
            n, c->nk, c->kvec.as<double4>(), c->sk.as<double2>(), c->pos, dq, c->dd.as<double>(), c->rowf.as<double>(),
            c_self(c), c->rec_dr.as<double>(), c->gp.as<double>(), vir ? c->rec_dd.as<double>() : nullptr);
    } else {
        PGM_TRY(spread_fft<real>(c, dq, true, true, nullptr, -1, vir ? c->vir_conv.as<double>() : nullptr));
        if (!vir) HIP_TRY(hipStreamWaitEvent(st, c->ev_join, 0));  // the pair pass's row field
        ORDER_SWITCH(c->s.order, (k_gather_forces<real, PO><<<nb, 256, 0, st>>>(
                                     n, nper, c->pos, c->pg, c->grid.as<real>(), dq, c->dd.as<double>(), c->rowf.as<double>(),
                                     c_self(c), c->rec_dr.as<double>(), c->gp.as<double>(),
                                     vir ? c->rec_dd.as<double>() : nullptr)));
    }
    LAUNCH_CHECK();
    if (vir) {
        k_virial_atoms<<<nb, 256, 0, st>>>(n, c->rec_dr.as<double>(), c->off.as<double>(), c->rec_dd.as<double>(),
                                           c->dd.as<double>(), c->vir_atoms.as<double>());
        LAUNCH_CHECK();
        const double b0 = c->s.ewald_beta;
        const double e_bg = -KE * M_PI * c->sum_q * c->sum_q / (2.0 * c->volume * b0 * b0);
        const double tail = c->s.lj_lrc ? lj_tail(c) : 0.0;
        const int ncb = molecular ? 0 : nblk(c->ncov, 256);
        if (ncb > 0) {
            k_virial_cov<<<ncb, 256, 0, st>>>(c->ncov, c->pos, c->box, c->ci.as<int>(), c->cj.as<int>(), c->cc.as<double>(),
                                               c->gp.as<double>(), c->vir_cov.as<double>());
            LAUNCH_CHECK();
        }
        k_virial_final<<<1, 256, 0, st>>>(c->vir_pair.as<double>(), npair, c->vir_conv.as<double>(), nconv,
                                          c->vir_atoms.as<double>(), nb, c->vir_cov.as<double>(), ncb, -e_bg - 2.0 * tail,
                                          c->vir_out.as<double>());
        LAUNCH_CHECK();
        HIP_TRY(hipMemcpyAsync(c->h_pin_dbl + 4, c->vir_out.p, 9 * sizeof(double), hipMemcpyDeviceToHost, st));
    }
    if (forces) {
        k_forces_final<<<nb, 256, 0, st>>>(n, c->pos, c->box, c->g_el.as<double>(), c->g_lj.as<double>(), c->rec_dr.as<double>(),
                                           c->gp.as<double>(), c->cstart.as<int>(), c->ci.as<int>(), c->cj.as<int>(),
                                           c->cc.as<double>(), c->jstart.as<int>(), c->jlist.as<int>(), forces,
                                           c->efield_on ? c->q.as<double>() : nullptr, kf);
        LAUNCH_CHECK();
    }
    const double b0 = c->s.ewald_beta;
    const double ke_const = KE * (-(b0 / SQRT_PI) * c->sum_q2 - M_PI * c->sum_q * c->sum_q / (2.0 * c->volume * b0 * b0));
    const double tail = c->s.lj_lrc ? lj_tail(c) : 0.0;
    k_energy_final<<<B, 256, 0, st>>>(c->pair_part.as<double>(), npair, c->pack_part.as<double>(), npb, c->conv_part.as<double>(),
                                      nconv, ke_const, tail, c_self(c), c->energy.as<double>(),
                                      c->efield_on ? c->fpart.as<double>() : nullptr, npb, kf, c->foff, c->fout.as<double>());
    LAUNCH_CHECK();
    if (quiet) return PGM_OK;  // energies stay in c->energy
    PGM_TRY(read_energy(c, energy));
    if (vir)
        for (int k = 0; k < 9; ++k) strain[k] = c->h_pin_dbl[4 + k];
    return PGM_OK;
}

// ---------------------------------------------------------------------------- C ABI
extern "C" int pgm_ctx_create(pgm_ctx **out, const pgm_topology *top, const pgm_settings *s, const double *H, void *stream) {
    if (!out || !top || !s || !H || top->n_atoms <= 0) {
        g_err = "pgm_ctx_create: null argument or no atoms";
        return PGM_ERR_ARG;
    }
    if (s->order < 4 || s->order > 8 || s->grid[0] < s->order || s->grid[1] < s->order || s->grid[2] < s->order) {
        g_err = "pgm_ctx_create: PME order must be 4..8 and each grid size at least the order";
        return PGM_ERR_ARG;
    }
    if (!(s->cutoff > 0) || s->skin < 0 || !(s->ewald_beta > 0) || s->max_iter < 0 || s->max_iter > 100000) {
        g_err = "pgm_ctx_create: invalid cutoff, skin, ewald_beta or max_iter";
        return PGM_ERR_ARG;
    }
    if (!(s->elec_cutoff >= 0)) {
        g_err = "pgm_ctx_create: elec_cutoff must be >= 0 (0: the cutoff)";
        return PGM_ERR_ARG;
    }
    if (s->recip != 0 && s->recip != 1) {
        g_err = "pgm_ctx_create: recip must be 0 (PME) or 1 (Ewald)";
        return PGM_ERR_ARG;
    }
    if (s->recip == 1 && !(s->k_tol > 0 && s->k_tol < 1)) {
        g_err = "pgm_ctx_create: Ewald needs 0 < k_tol < 1";
        return PGM_ERR_ARG;
    }
    const int nbead = s->nbead > 1 ? s->nbead : 1;
    if (nbead > 1 && (s->recip != 0 || s->predictor == 4)) {
        g_err = "pgm_ctx_create: beads run with smooth PME and a polynomial or previous-step predictor";
        return PGM_ERR_ARG;
    }
    if ((long long)nbead * top->n_atoms > (1LL << 26)) {
        g_err = "pgm_ctx_create: more than 2^26 atoms over all beads";
        return PGM_ERR_ARG;
    }
    pgm_ctx *c = new pgm_ctx();
    c->nbead = nbead;
    c->nper = top->n_atoms;
    c->n = nbead * top->n_atoms;
    c->s = *s;
    c->rc_v = s->cutoff;
    c->rc_e = s->elec_cutoff > 0 ? s->elec_cutoff : s->cutoff;
    c->rc_pair = std::max(c->rc_e, c->rc_v);
    c->ewald = s->recip == 1;
    if (c->ewald) {  // half-space index set with 0 < |k| < kcut at the creation box (pgm_jax/ewald.py:129-158)
        const double kcut = 2.0 * s->ewald_beta * std::sqrt(-std::log(s->k_tol));
        double Hi[9], hmax = 0.0;
        inv3(H, Hi);
        for (int a = 0; a < 3; ++a) hmax = std::max(hmax, std::sqrt(H[3 * a] * H[3 * a] + H[3 * a + 1] * H[3 * a + 1] + H[3 * a + 2] * H[3 * a + 2]));
        const int Mx = (int)std::ceil(kcut * hmax / (2.0 * M_PI)) + 1;
        for (int a = -Mx; a <= Mx; ++a)
            for (int b = -Mx; b <= Mx; ++b)
                for (int d = -Mx; d <= Mx; ++d) {
                    if (!(a > 0 || (a == 0 && b > 0) || (a == 0 && b == 0 && d > 0))) continue;
                    double k2 = 0.0;
                    for (int j = 0; j < 3; ++j) {
                        const double kj = 2.0 * M_PI * (Hi[3 * j] * a + Hi[3 * j + 1] * b + Hi[3 * j + 2] * d);
                        k2 += kj * kj;
                    }
                    if (k2 > 0.0 && std::sqrt(k2) < kcut) {
                        c->h_m.push_back(a);
                        c->h_m.push_back(b);
                        c->h_m.push_back(d);
                    }
                }
        c->nk = (int)(c->h_m.size() / 3);
        if (c->nk == 0) {
            g_err = "pgm_ctx_create: k_tol leaves no k-vector";
            delete c;
            return PGM_ERR_ARG;
        }
    }
    c->stream = (hipStream_t)stream;
    c->mixed = s->precision == 1;
    c->rsize = c->mixed ? sizeof(float) : sizeof(double);
    const int n = c->n, nper = c->nper;
    // per-atom parameters of one bead, repeated for every bead (molecule indices shifted by the
    // bead's molecule count)
    int nmol_bead = 0;
    for (int i = 0; i < nper; ++i) nmol_bead = std::max(nmol_bead, top->mol[i] + 1);
    for (int b = 0; b < nbead; ++b) {
        c->h_q.insert(c->h_q.end(), top->q, top->q + nper);
        c->h_alpha.insert(c->h_alpha.end(), top->alpha, top->alpha + nper);
        c->h_rh.insert(c->h_rh.end(), top->lj_rmin_half, top->lj_rmin_half + nper);
        c->h_se.insert(c->h_se.end(), top->lj_sqrt_eps, top->lj_sqrt_eps + nper);
        for (int i = 0; i < nper; ++i) c->h_mol.push_back(top->mol[i] + b * nmol_bead);
    }
    c->h_R2.resize(n);
    for (int i = 0; i < n; ++i) c->h_R2[i] = top->radius[i % nper] * top->radius[i % nper];
    for (int i = 0; i < nper; ++i) {  // per bead
        c->sum_q += c->h_q[i];
        c->sum_q2 += c->h_q[i] * c->h_q[i];
        if (!(c->h_alpha[i] >= 0)) {
            g_err = "pgm_ctx_create: polarizabilities must be >= 0 (0: not polarizable)";
            delete c;
            return PGM_ERR_ARG;
        }
        if (c->h_alpha[i] == 0) c->alpha0 = true;
    }
    c->lj_sb = lj_sum_b(c);
    auto fail = [&](int r) {
        pgm_ctx_destroy(c);
        return r;
    };
    int r;
    hipStream_t st = c->stream;
    if ((r = upload(c->q, c->h_q.data(), n, st)) < 0) return fail(r);
    if ((r = upload(c->alpha64, c->h_alpha.data(), n, st)) < 0) return fail(r);
    if ((r = upload(c->rh, c->h_rh.data(), n, st)) < 0) return fail(r);
    if ((r = upload(c->se, c->h_se.data(), n, st)) < 0) return fail(r);
    if ((r = upload(c->mol, c->h_mol.data(), n, st)) < 0) return fail(r);
    {  // (rmin/2, sqrt eps, molecule, 0) per atom for the force pass
        std::vector<double> lj(4 * (size_t)n);
        for (int i = 0; i < n; ++i) {
            lj[4 * (size_t)i] = c->h_rh[i];
            lj[4 * (size_t)i + 1] = c->h_se[i];
            lj[4 * (size_t)i + 2] = (double)c->h_mol[i];
            lj[4 * (size_t)i + 3] = 0.0;
        }
        if ((r = upload(c->ljp, lj.data(), lj.size(), st)) < 0) return fail(r);
    }
    if ((r = upload(c->R2, c->h_R2.data(), n, st)) < 0) return fail(r);
    {  // molecules as CSR over atom indices (for centres of mass)
        int mmax = -1;
        for (int i = 0; i < n; ++i) {
            if (top->mol[i % nper] < 0 || top->mol[i % nper] >= nper) {
                g_err = "pgm_ctx_create: molecule index out of range [0, n_atoms)";
                return fail(PGM_ERR_ARG);
            }
            mmax = std::max(mmax, c->h_mol[i]);
        }
        c->nmol_all = mmax + 1;
        std::vector<int> mp(c->nmol_all + 1, 0), ma(n);
        for (int i = 0; i < n; ++i) mp[c->h_mol[i] + 1]++;
        for (int m = 0; m < c->nmol_all; ++m) mp[m + 1] += mp[m];
        std::vector<int> fill(mp.begin(), mp.end() - 1);
        for (int i = 0; i < n; ++i) ma[fill[c->h_mol[i]]++] = i;
        if ((r = upload(c->mol_ptr, mp.data(), mp.size(), st)) < 0) return fail(r);
        if ((r = upload(c->mol_atoms, ma.data(), ma.size(), st)) < 0) return fail(r);
    }
    if (c->mixed) {
        std::vector<float> af(c->h_alpha.begin(), c->h_alpha.end());
        if ((r = upload(c->alpha_r, af.data(), n, st)) < 0) return fail(r);
    } else {
        if ((r = upload(c->alpha_r, c->h_alpha.data(), n, st)) < 0) return fail(r);
    }
    // covalent dipoles: CSR by owner atom (cstart over m sorted by cov_i) and by target (jstart/jlist);
    // one bead's list, repeated with atom offsets for every bead
    const int nc1 = top->n_cov, nc = nbead * top->n_cov;
    c->ncov = nc;
    std::vector<int> order(nc1), ci(nc), cj(nc), cs(n + 1, 0), js(n + 1, 0), jl(nc);
    std::vector<double> cc(nc);
    for (int m = 0; m < nc1; ++m) {
        if (top->cov_i[m] < 0 || top->cov_i[m] >= nper || top->cov_j[m] < 0 || top->cov_j[m] >= nper) {
            g_err = "pgm_ctx_create: covalent dipole atom out of range";
            return fail(PGM_ERR_ARG);
        }
        order[m] = m;
    }
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return top->cov_i[a] < top->cov_i[b]; });
    for (int k = 0; k < nc; ++k) {
        const int m = order[k % nc1], off = (k / nc1) * nper;
        ci[k] = top->cov_i[m] + off;
        cj[k] = top->cov_j[m] + off;
        cc[k] = top->cov_c[m];
        cs[ci[k] + 1]++;
        js[cj[k] + 1]++;
    }
    for (int i = 0; i < n; ++i) {
        cs[i + 1] += cs[i];
        js[i + 1] += js[i];
    }
    std::vector<int> fillj(js.begin(), js.end() - 1);
    for (int k = 0; k < nc; ++k) jl[fillj[cj[k]]++] = k;
    if ((r = upload(c->cstart, cs.data(), n + 1, st)) < 0) return fail(r);
    if ((r = upload(c->ci, ci.data(), nc, st)) < 0) return fail(r);
    if ((r = upload(c->cj, cj.data(), nc, st)) < 0) return fail(r);
    if ((r = upload(c->cc, cc.data(), nc, st)) < 0) return fail(r);
    if ((r = upload(c->jstart, js.data(), n + 1, st)) < 0) return fail(r);
    if ((r = upload(c->jlist, jl.data(), nc, st)) < 0) return fail(r);
    c->h_corder.resize(nc);
    for (int k = 0; k < nc; ++k) c->h_corder[k] = order[k % nc1] + (k / nc1) * nc1;
    if ((r = upload(c->corder, c->h_corder.data(), nc, st)) < 0) return fail(r);
    // buffers
    const size_t r4 = 4 * c->rsize * (size_t)n, d3 = sizeof(double) * 3 * (size_t)n;
    DevBuf *v4[] = {&c->vx, &c->vr, &c->vz, &c->vp, &c->vAp, &c->vb, &c->vx0, &c->vrow, &c->vdq, &c->vtmp};
    for (DevBuf *b : v4)
        if ((r = b->ensure(r4)) < 0) return fail(r);
    DevBuf *v3[] = {&c->pperm, &c->dd, &c->g_el, &c->g_lj, &c->rowf, &c->rec_dr, &c->gp};
    for (DevBuf *b : v3)
        if ((r = b->ensure(d3)) < 0) return fail(r);
    if ((r = c->xr.ensure(sizeof(double4) * n)) < 0) return fail(r);
    const long long nbmax = (long long)nbead * (nblk((long long)nper * ROWL, 64) + 16);
    DevBuf *parts[] = {&c->partA, &c->partB, &c->partC, &c->pair_part, &c->pack_part};
    for (DevBuf *b : parts)
        if ((r = b->ensure(sizeof(double) * 4 * nbmax)) < 0) return fail(r);
    c->rzs = s->max_iter + 2;
    if ((r = c->st.ensure(sizeof(CGState) * nbead)) < 0) return fail(r);
    if ((r = c->rz.ensure(sizeof(double) * c->rzs * nbead)) < 0) return fail(r);
    if ((r = c->err.ensure(sizeof(double) * c->rzs * nbead)) < 0) return fail(r);
    if ((r = c->energy.ensure(sizeof(double) * (4 * nbead + 4))) < 0) return fail(r);
    if ((r = c->bead_it.ensure(sizeof(int) * nbead)) < 0) return fail(r);
    if ((r = c->bead_res.ensure(sizeof(double) * nbead)) < 0) return fail(r);
    if ((r = c->stats.ensure(sizeof(double) * c->stats_cap)) < 0) return fail(r);
    if ((r = c->maxcount.ensure(sizeof(int))) < 0) return fail(r);
    DevBuf *ib[] = {&c->key, &c->atom, &c->skey, &c->satom, &c->count_full, &c->count, &c->count_near};
    for (DevBuf *b : ib)
        if ((r = b->ensure(sizeof(int) * n)) < 0) return fail(r);
    if (hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_rec, hipEventDisableTiming) != hipSuccess) {
        g_err = "stream/event creation failed";
        return fail(PGM_ERR_HIP);
    }
    for (int k = 0; k < 4; ++k)
        if (hipEventCreateWithFlags(&c->ev_read[k], hipEventDisableTiming) != hipSuccess) {
            g_err = "event creation failed";
            return fail(PGM_ERR_HIP);
        }
    if (hipHostMalloc((void **)&c->h_pin_int, (16 + nbead) * sizeof(int), hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void **)&c->h_pin_dbl, (32 + 5 * (size_t)nbead) * sizeof(double), hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void **)&c->h_pin_conv, 4 * (size_t)nbead * sizeof(CGState), hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void **)&c->h_mirror, (4 * (size_t)nbead + 1) * sizeof(unsigned long long), hipHostMallocCoherent) !=
            hipSuccess) {
        g_err = "hipHostMalloc failed";
        return fail(PGM_ERR_HIP);
    }
    c->ngrid = (size_t)s->grid[0] * s->grid[1] * s->grid[2];
    c->ncplx = (size_t)s->grid[0] * s->grid[1] * (s->grid[2] / 2 + 1);
    if ((r = c->grid.ensure(c->rsize * c->ngrid * nbead)) < 0) return fail(r);
    if ((r = c->cgrid.ensure(2 * c->rsize * c->ncplx * nbead)) < 0) return fail(r);
    if ((r = c->conv_part.ensure(sizeof(double) * nbead *
                                 (std::max({nblk((long long)c->ncplx, 256), nblk(c->nk, EW_TILE),
                                            nblk((long long)s->grid[1] * (s->grid[2] / 2 + 1), PFFT_COLS)}) + 16))) < 0)
        return fail(r);
    if (c->ewald && (r = c->sk.ensure(sizeof(double2) * std::max(c->nk, 1))) < 0) return fail(r);
    if ((r = setup_tiles(c)) < 0) return fail(r);
    if (nbead > 1 && !c->tiled) {
        g_err = "pgm_ctx_create: beads need a tiled PME grid (per axis at least three tiles of P-1..8 points)";
        return fail(PGM_ERR_ARG);
    }
    {  // molecule blocks of the CG row field when every molecule is 3 consecutive atoms, the
       // candidate bitmap of a bead fits 64 KB of LDS per block (k_mol_rows) and every molecule
       // index of all beads fits the 26-bit partner field of a record word (k_mol_tensors);
       // otherwise the atom rows
        bool m3 = nper % 3 == 0;
        for (int i = 0; i < nper && m3; ++i) m3 = top->mol[i] == i / 3;
        c->mol3 = m3 && (nper / 3 + 31) / 32 <= 4096 && (long long)c->n / 3 < (1LL << 26);
        c->layout = c->mol3 ? 1 : 0;
        if (c->layout == 1 && (r = alloc_mol(c, 64)) < 0) return fail(r);
    }
    if (s->precond != 0 && s->precond != 1) {
        g_err = "pgm_ctx_create: precond must be 0 (Jacobi) or 1 (molecule blocks)";
        return fail(PGM_ERR_ARG);
    }
    if (s->precond == 1 && c->alpha0) {
        g_err = "pgm_ctx_create: the molecule-block preconditioner needs every polarizability > 0";
        return fail(PGM_ERR_ARG);
    }
    if (s->precond == 1) {
        bool m3 = nper % 3 == 0;
        for (int i = 0; i < nper && m3; ++i) m3 = top->mol[i] == i / 3;
        if (!m3) {
            g_err = "pgm_ctx_create: the molecule-block preconditioner needs consecutive 3-atom molecules";
            return fail(PGM_ERR_ARG);
        }
        c->molp = true;
        if ((r = c->pinv.ensure(c->rsize * 81 * (size_t)(n / 3))) < 0) return fail(r);
    }
    if ((r = set_box(c, H)) < 0) return fail(r);
    // initial row capacity from the density (a sphere of the row cutoff + skin), grown on overflow
    const double rl = c->rc_pair + s->skin;
    const int cap0 = (int)(1.25 * 4.0 / 3.0 * M_PI * rl * rl * rl * c->nper / c->volume) + 32;
    if ((r = alloc_rows(c, std::min(cap0, c->nper))) < 0) return fail(r);
    if ((r = make_plans(c)) < 0) return fail(r);
    if ((r = setup_fused(c)) < 0) return fail(r);
    *out = c;
    return PGM_OK;
}

extern "C" int pgm_ctx_destroy(pgm_ctx *c) {
    if (!c) return PGM_OK;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    else (void)hipDeviceSynchronize();
    if (c->side) (void)hipStreamSynchronize(c->side);
    if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
    if (c->ev_join) (void)hipEventDestroy(c->ev_join);
    if (c->ev_rec) (void)hipEventDestroy(c->ev_rec);
    for (int k = 0; k < 4; ++k)
        if (c->ev_read[k]) (void)hipEventDestroy(c->ev_read[k]);
    for (auto &t : c->timed)
        for (hipEvent_t e : t.ev)
            if (e) (void)hipEventDestroy(e);
    drop_graphs(c);
    if (c->gstream) {
        (void)hipStreamSynchronize(c->gstream);
        (void)hipStreamDestroy(c->gstream);
    }
    if (c->ev_g0) (void)hipEventDestroy(c->ev_g0);
    if (c->ev_g1) (void)hipEventDestroy(c->ev_g1);
    if (c->side) (void)hipStreamDestroy(c->side);
    if (c->fwd) rocfft_plan_destroy(c->fwd);
    if (c->bwd) rocfft_plan_destroy(c->bwd);
    if (c->info) rocfft_execution_info_destroy(c->info);
    DevBuf *all[] = {&c->q, &c->alpha64, &c->alpha_r, &c->rh, &c->se, &c->mol, &c->ljp, &c->R2, &c->cstart, &c->ci, &c->cj, &c->cc,
                     &c->jstart, &c->jlist, &c->corder, &c->adj, &c->adj2, &c->xr, &c->rows_full, &c->count_full, &c->rows, &c->count, &c->count_near, &c->maxcount, &c->jell, &c->G1, &c->G2, &c->key,
                     &c->atom, &c->skey, &c->satom, &c->cell_start, &c->cell_end, &c->sort_tmp, &c->grid, &c->cgrid, &c->Gc,
                     &c->fft_work, &c->vx, &c->vr, &c->vz, &c->vp, &c->vAp, &c->vb, &c->vx0, &c->vrow, &c->vdq, &c->vtmp,
                     &c->pperm, &c->dd, &c->g_el, &c->g_lj, &c->rowf, &c->rec_dr, &c->gp, &c->partA, &c->partB, &c->partC,
                     &c->pair_part, &c->pack_part, &c->conv_part, &c->st, &c->rz, &c->err, &c->energy, &c->bead_it, &c->bead_res,
                     &c->mrows, &c->mcount_full, &c->mcount, &c->mj, &c->mg, &c->pinv,
                     &c->pos_build, &c->minv, &c->mass, &c->d2, &c->md_part, &c->md_fail, &c->heat,
                     &c->pkey, &c->pidx, &c->pskey, &c->psidx, &c->tstart, &c->tend, &c->tstart16, &c->tend16, &c->spl, &c->loc,
                     &c->loc16, &c->blocks, &c->merge_tab, &c->pfft_tw,
                     &c->pme_sort_tmp, &c->stats, &c->mol_ptr, &c->mol_atoms, &c->com, &c->off, &c->rec_dd,
                     &c->vir_pair, &c->vir_conv, &c->vir_atoms, &c->vir_cov, &c->vir_out, &c->kvec, &c->sk,
                     &c->guess.params, &c->guess.grads, &c->guess.m, &c->guess.v, &c->guess.packed, &c->guess.bias,
                     &c->guess.feat, &c->guess.frames, &c->guess.y, &c->guess.bx, &c->guess.by, &c->guess.idx, &c->guess.xb,
                     &c->guess.yb, &c->guess.acts, &c->guess.zs, &c->guess.delta, &c->guess.delta2, &c->guess.part,
                     &c->guess.scal, &c->guess.ring};
    for (DevBuf *b : all) b->release();
    if (c->h_pin_int) (void)hipHostFree(c->h_pin_int);
    if (c->h_pin_conv) (void)hipHostFree(c->h_pin_conv);
    if (c->h_mirror) (void)hipHostFree(c->h_mirror);
    if (c->h_pin_dbl) (void)hipHostFree(c->h_pin_dbl);
    delete c;
    return PGM_OK;
}

extern "C" int pgm_set_stream(pgm_ctx *c, void *stream) {
    if (!c) return PGM_ERR_ARG;
    drop_graphs(c);
    c->stream = (hipStream_t)stream;
    if (c->info) FFT_TRY(rocfft_execution_info_set_stream(c->info, c->stream));
    return PGM_OK;
}

extern "C" int pgm_set_box(pgm_ctx *c, const double *H) {
    if (!c || !H) return PGM_ERR_ARG;
    return set_box(c, H);
}

extern "C" int pgm_build_rows(pgm_ctx *c, const double *pos) {
    if (!c || !pos) return PGM_ERR_ARG;
    return build_rows(c, pos);
}

extern "C" long long pgm_row_entries(pgm_ctx *c) {
    if (!c || !c->rows_valid) return -1;
    if (!c->atom_rows_fresh && compact_atom_rows(c) < 0) return -1;
    std::vector<int> cnt(c->n);
    if (hipMemcpyAsync(cnt.data(), c->count.p, sizeof(int) * c->n, hipMemcpyDeviceToHost, c->stream) != hipSuccess) return -1;
    if (hipStreamSynchronize(c->stream) != hipSuccess) return -1;
    long long s = 0;
    for (int v : cnt) s += v;
    return s;
}

extern "C" long long pgm_row_blocks(pgm_ctx *c) {
    if (!c || !c->rows_valid) return -1;
    if (c->layout != 1) return 0;
    const int nmol = c->n / 3;
    std::vector<int> cnt(nmol);
    if (hipMemcpyAsync(cnt.data(), c->mcount.p, sizeof(int) * nmol, hipMemcpyDeviceToHost, c->stream) != hipSuccess) return -1;
    if (hipStreamSynchronize(c->stream) != hipSuccess) return -1;
    long long s = 0;
    for (int v : cnt) s += v;
    return s;
}

extern "C" int pgm_perm_dipoles(pgm_ctx *c, const double *pos, double *p) {
    if (!c || !pos || !p) return PGM_ERR_ARG;
    k_perm_dipoles<<<nblk(c->n, 256), 256, 0, c->stream>>>(c->n, pos, c->box, c->cstart.as<int>(), c->cj.as<int>(),
                                                            c->cc.as<double>(), p);
    LAUNCH_CHECK();
    return PGM_OK;
}

template <typename real> static int field_api(pgm_ctx *c, const double *pos, int withq, const double *d, double *field) {
    PGM_TRY(prepare(c, pos));
    using R4 = real4<real>;
    const int nb = nblk(c->n, 256);
    k_pack_dq<real><<<nb, 256, 0, c->stream>>>(c->n, d, nullptr, withq ? c->q.as<double>() : nullptr, c->vdq.as<R4>());
    LAUNCH_CHECK();
    PGM_TRY(field_eval<real>(c, c->vdq.as<R4>(), withq != 0, MODE_FIELD, nullptr, c->vtmp.as<R4>(), nullptr, -1));
    k_to_double3<real><<<nb, 256, 0, c->stream>>>(c->n, c->vtmp.as<R4>(), field);
    LAUNCH_CHECK();
    HIP_TRY(hipStreamSynchronize(c->stream));
    return PGM_OK;
}

extern "C" int pgm_field(pgm_ctx *c, const double *pos, int with_charges, const double *d, double *field) {
    if (!c || !pos || !d || !field) return PGM_ERR_ARG;
    return c->mixed ? field_api<float>(c, pos, with_charges, d, field) : field_api<double>(c, pos, with_charges, d, field);
}

template <typename real>
static int pme_api(pgm_ctx *c, const double *pos, int withq, const double *d, double *e_rec, double *grad_d) {
    using R4 = real4<real>;
    if (pos != c->pos) drop_graphs(c);
    c->pos = pos;
    PGM_TRY(pme_bin<real>(c, pos));
    const int nb = nblk(c->n, 256);
    k_pack_dq<real><<<nb, 256, 0, c->stream>>>(c->n, d, nullptr, withq ? c->q.as<double>() : nullptr, c->vdq.as<R4>());
    LAUNCH_CHECK();
    PGM_TRY(spread_fft<real>(c, c->vdq.as<R4>(), withq != 0, true, nullptr, -1));
    if (grad_d) {
        ORDER_SWITCH(c->s.order, (k_gather_combine<real, PO><<<nb, 256, 0, c->stream>>>(
                                     c->n, pos, c->pg, c->grid.as<real>(), c->vdq.as<R4>(), c->vrow.as<R4>(),
                                     c->alpha_r.as<real>(), nullptr, real(0), MODE_REC, c->vtmp.as<R4>(), nullptr, nullptr, -1)));
        k_to_double3<real><<<nb, 256, 0, c->stream>>>(c->n, c->vtmp.as<R4>(), grad_d);
        LAUNCH_CHECK();
    }
    // energy: sum of the convolution partials
    k_energy_final<<<1, 256, 0, c->stream>>>(c->pair_part.as<double>(), 0, c->pack_part.as<double>(), 0, c->conv_part.as<double>(),
                                             conv_partials(c, false), 0.0, 0.0, 0.0, c->energy.as<double>());
    LAUNCH_CHECK();
    HIP_TRY(hipMemcpyAsync(c->h_pin_dbl, c->energy.p, 4 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (e_rec) *e_rec = c->h_pin_dbl[3];
    return PGM_OK;
}

static int single_system(const pgm_ctx *c, const char *what) {
    if (c->nbead == 1) return PGM_OK;
    g_err = std::string(what) + ": one system per context (this context holds beads)";
    return PGM_ERR_STATE;
}

extern "C" int pgm_pme_energy(pgm_ctx *c, const double *pos, int with_charges, const double *d, double *e_rec, double *grad_d) {
    if (!c || !pos || !d) return PGM_ERR_ARG;
    PGM_TRY(single_system(c, "pgm_pme_energy"));
    return c->mixed ? pme_api<float>(c, pos, with_charges, d, e_rec, grad_d) : pme_api<double>(c, pos, with_charges, d, e_rec, grad_d);
}

// PME.spread (pgm_jax/md/pme.py:306-337): the grid Q of the topology charges (with_charges) and
// the dipoles d, merged from the spread tiles (or spread per atom on untiled grids).
template <typename real> static int spread_api(pgm_ctx *c, const double *pos, int withq, const double *d, double *Q) {
    using R4 = real4<real>;
    if (pos != c->pos) drop_graphs(c);
    c->pos = pos;
    PGM_TRY(pme_bin<real>(c, pos));
    hipStream_t st = c->stream;
    const int n = c->n, P = c->s.order;
    k_pack_dq<real><<<nblk(n, 256), 256, 0, st>>>(n, d, nullptr, withq ? c->q.as<double>() : nullptr, c->vdq.as<R4>());
    LAUNCH_CHECK();
    const R4 *dq = c->vdq.as<R4>();
    const int *conv = nullptr;
    const int iter = -1;
    if (c->tiled) {
        const PMETiles &t = c->tiles8;
        const size_t lds = spread_lds_bytes(c);
        if (withq) {
            ORDER_SWITCH(P, (k_spread_tiles<real, PO, true><<<t.ntiles, WAVE, lds, st>>>(
                                t, c->pg, c->tstart.as<int>(), c->tend.as<int>(), c->psidx.as<int>(), c->spl.as<real>(),
                                c->loc.as<int>(), dq, c->blocks.as<real>(), conv, iter)));
        } else {
            ORDER_SWITCH(P, (k_spread_tiles<real, PO, false><<<t.ntiles, WAVE, lds, st>>>(
                                t, c->pg, c->tstart.as<int>(), c->tend.as<int>(), c->psidx.as<int>(), c->spl.as<real>(),
                                c->loc.as<int>(), dq, c->blocks.as<real>(), conv, iter)));
        }
        LAUNCH_CHECK();
        const int *K = c->s.grid;
        k_merge_table<real><<<dim3(nblk(K[2], 128), K[1], K[0]), 128, 0, st>>>(
            K[0], K[1], K[2], (size_t)t.ntiles * t.bsize, c->merge_tab.as<int>(), c->blocks.as<real>(), c->grid.as<real>(), conv,
            iter);
    } else {
        k_fill<real><<<nblk((long long)c->ngrid, 256), 256, 0, st>>>(c->ngrid, c->grid.as<real>(), real(0));
        const dim3 gs(nblk((long long)n * P, 256));
        if (withq) {
            ORDER_SWITCH(P, (k_spread<real, PO, true><<<gs, 256, 0, st>>>(n, pos, dq, c->pg, c->grid.as<real>(), conv, iter)));
        } else {
            ORDER_SWITCH(P, (k_spread<real, PO, false><<<gs, 256, 0, st>>>(n, pos, dq, c->pg, c->grid.as<real>(), conv, iter)));
        }
    }
    LAUNCH_CHECK();
    k_to_double<real><<<nblk((long long)c->ngrid, 256), 256, 0, st>>>(c->ngrid, c->grid.as<real>(), Q);
    LAUNCH_CHECK();
    return PGM_OK;
}

extern "C" int pgm_pme_spread(pgm_ctx *c, const double *pos, int with_charges, const double *d, double *Q) {
    if (!c || !pos || !d || !Q) return PGM_ERR_ARG;
    PGM_TRY(single_system(c, "pgm_pme_spread"));
    if (c->ewald) {
        g_err = "pgm_pme_spread: PME reciprocal space only (this context holds the exact Ewald sum)";
        return PGM_ERR_STATE;
    }
    return c->mixed ? spread_api<float>(c, pos, with_charges, d, Q) : spread_api<double>(c, pos, with_charges, d, Q);
}

// PME.influence (pgm_jax/md/pme.py:243-267) of the context's box: G (K1, K2, K3/2 + 1) with the
// half-space weights (1 on the planes m3 = 0 and m3 = K3/2, 2 elsewhere), float64, host.
extern "C" int pgm_pme_influence(pgm_ctx *c, double *G) {
    if (!c || !G) return PGM_ERR_ARG;
    if (c->ewald) {
        g_err = "pgm_pme_influence: PME reciprocal space only (this context holds the exact Ewald sum)";
        return PGM_ERR_STATE;
    }
    std::vector<double> Gc;
    influence_host(c, c->box.hinv, Gc);
    const int K3 = c->s.grid[2], K3h = K3 / 2 + 1;
    for (size_t i = 0; i < Gc.size(); ++i) {
        const int k = (int)(i % K3h);
        const bool single = k == 0 || (K3 % 2 == 0 && k == K3 / 2);
        G[i] = single ? 0.5 * Gc[i] : Gc[i];
    }
    return PGM_OK;
}

template <typename real>
static int compute_api(pgm_ctx *c, const double *pos, pgm_induction *ind, pgm_energy *energy, double *forces, int *iterations,
                       double *residual, bool quiet = false) {
    PGM_TRY(prepare(c, pos, true));
    k_perm_dipoles<<<nblk(c->n, 256), 256, 0, c->stream>>>(c->n, pos, c->box, c->cstart.as<int>(), c->cj.as<int>(),
                                                            c->cc.as<double>(), c->pperm.as<double>());
    LAUNCH_CHECK();
    const int rs = solve_impl<real>(c, ind, iterations, residual, quiet);
    if (rs < 0) return rs;
    PGM_TRY(energy_forces_impl<real>(c, ind->mu, energy, forces, quiet));
    return rs;
}

extern "C" int pgm_compute(pgm_ctx *c, const double *pos, pgm_induction *ind, pgm_energy *energy, double *forces, int *iterations,
                           double *residual) {
    if (!c || !pos || !ind || !ind->mu || !ind->hist) return PGM_ERR_ARG;
    return c->mixed ? compute_api<float>(c, pos, ind, energy, forces, iterations, residual)
                    : compute_api<double>(c, pos, ind, energy, forces, iterations, residual);
}

template <typename real>
static int energy_api(pgm_ctx *c, const double *pos, pgm_induction *ind, pgm_energy *energy, int *iterations, double *residual) {
    const int pr = prepare(c, pos);
    if (pr < 0) return pr;
    k_perm_dipoles<<<nblk(c->n, 256), 256, 0, c->stream>>>(c->n, pos, c->box, c->cstart.as<int>(), c->cj.as<int>(),
                                                            c->cc.as<double>(), c->pperm.as<double>());
    LAUNCH_CHECK();
    const int rs = solve_impl<real>(c, ind, iterations, residual, false, true);
    if (rs < 0) return rs;
    PGM_TRY(energy_forces_impl<real>(c, ind->mu, energy, nullptr));
    return rs != PGM_OK ? rs : pr;
}

extern "C" int pgm_trial_energy(pgm_ctx *c, const double *pos, pgm_induction *ind, pgm_energy *energy, int *iterations,
                                double *residual) {
    if (!c || !pos || !ind || !ind->mu) return PGM_ERR_ARG;
    return c->mixed ? energy_api<float>(c, pos, ind, energy, iterations, residual)
                    : energy_api<double>(c, pos, ind, energy, iterations, residual);
}

extern "C" int pgm_bead_count(const pgm_ctx *c) { return c ? c->nbead : PGM_ERR_ARG; }

extern "C" int pgm_set_graphs(pgm_ctx *c, int on) {
    if (!c) return PGM_ERR_ARG;
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (on && !c->gstream) {
        HIP_TRY(hipStreamCreateWithFlags(&c->gstream, hipStreamNonBlocking));
        HIP_TRY(hipEventCreateWithFlags(&c->ev_g0, hipEventDisableTiming));
        HIP_TRY(hipEventCreateWithFlags(&c->ev_g1, hipEventDisableTiming));
    }
    if (!on) drop_graphs(c);
    c->use_graphs = on != 0;
    return PGM_OK;
}

extern "C" int pgm_set_row_layout(pgm_ctx *c, int layout, int lanes) {
    if (!c || (layout != 0 && layout != 1) || (layout == 1 && lanes != 4 && lanes != 8 && lanes != 16)) {
        g_err = "pgm_set_row_layout: layout 0 (atom rows) or 1 (molecule blocks, 4, 8 or 16 lanes)";
        return PGM_ERR_ARG;
    }
    if (layout == 1 && !c->mol3) {
        g_err = "pgm_set_row_layout: molecule blocks need consecutive 3-atom molecules";
        return PGM_ERR_ARG;
    }
    HIP_TRY(hipStreamSynchronize(c->stream));
    drop_graphs(c);  // captured iterations hold the previous layout's kernels and row tensors
    c->layout = layout;
    if (layout == 1) {
        c->rlm = lanes;
        PGM_TRY(alloc_mol(c, std::max(c->mcap_full, 64)));
    }
    c->rows_valid = false;
    return PGM_OK;
}

extern "C" int pgm_row_layout(const pgm_ctx *c, int *layout, int *lanes) {
    if (!c) return PGM_ERR_ARG;
    if (layout) *layout = c->layout;
    if (lanes) *lanes = c->layout == 1 ? c->rlm : RL;
    return PGM_OK;
}

extern "C" int pgm_bead_results(pgm_ctx *c, pgm_energy *energies, int *iterations, double *residuals) {
    if (!c) return PGM_ERR_ARG;
    const int B = c->nbead;
    hipStream_t st = c->stream;
    double *h = c->h_pin_dbl + 32;
    HIP_TRY(hipMemcpyAsync(h, c->energy.p, 4 * sizeof(double) * B, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(h + 4 * B, c->bead_res.p, sizeof(double) * B, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(c->h_pin_int + 16, c->bead_it.p, sizeof(int) * B, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    for (int b = 0; b < B; ++b) {
        if (energies) {
            energies[b].elec = h[4 * b];
            energies[b].vdw = h[4 * b + 1];
            energies[b].total = h[4 * b + 2];
        }
        if (iterations) iterations[b] = c->h_pin_int[16 + b];
        if (residuals) residuals[b] = h[4 * B + b];
    }
    return PGM_OK;
}

template <typename real>
static int ef_api(pgm_ctx *c, const double *pos, const double *mu, pgm_energy *energy, double *forces,
                  const double *masses = nullptr, double *strain = nullptr, bool molecular = true) {
    PGM_TRY(prepare(c, pos));
    k_perm_dipoles<<<nblk(c->n, 256), 256, 0, c->stream>>>(c->n, pos, c->box, c->cstart.as<int>(), c->cj.as<int>(),
                                                            c->cc.as<double>(), c->pperm.as<double>());
    LAUNCH_CHECK();
    return energy_forces_impl<real>(c, mu, energy, forces, false, masses, strain, molecular);
}

extern "C" int pgm_energy_forces(pgm_ctx *c, const double *pos, const double *mu, pgm_energy *energy, double *forces) {
    if (!c || !pos || !mu) return PGM_ERR_ARG;
    return c->mixed ? ef_api<float>(c, pos, mu, energy, forces) : ef_api<double>(c, pos, mu, energy, forces);
}

extern "C" int pgm_strain_derivative(pgm_ctx *c, const double *pos, const double *mu, const double *masses, int molecular,
                                     pgm_energy *energy, double *forces, double *strain) {
    if (!c || !pos || !mu || !strain || (molecular && !masses)) {
        g_err = "pgm_strain_derivative: null argument";
        return PGM_ERR_ARG;
    }
    PGM_TRY(single_system(c, "pgm_strain_derivative"));
    return c->mixed ? ef_api<float>(c, pos, mu, energy, forces, masses, strain, molecular != 0)
                    : ef_api<double>(c, pos, mu, energy, forces, masses, strain, molecular != 0);
}

// ---------------------------------------------------------------------------- backward pass of the solve
// The custom VJP of PGMForceField._solve (forcefield.py:1853-1915): the adjoint CG A lam = mu_bar
// (x0 = 0, no peek, to `tol`, the same operator and preconditioner as the forward solve), then lam
// pulled back through the residual b - A mu at the dipoles mu (adjoint.h).  Every output pointer
// may be null.
template <typename real>
static int dipole_vjp_impl(pgm_ctx *c, const double *pos, const double *mu, const double *mu_bar, double tol, double *lambda,
                           double *q_bar, double *radius_bar, double *alpha_bar, double *p_bar, double *cov_bar,
                           int *iterations, double *residual) {
    using R4 = real4<real>;
    const int n = c->n, nper = c->nper;
    hipStream_t st = c->stream;
    const int nb = nblk(n, 256);
    PGM_TRY(prepare(c, pos));
    k_perm_dipoles<<<nb, 256, 0, st>>>(n, pos, c->box, c->cstart.as<int>(), c->cj.as<int>(), c->cc.as<double>(),
                                       c->pperm.as<double>());
    LAUNCH_CHECK();
    PGM_TRY(c->adj.ensure(sizeof(double) * 10 * (size_t)n));
    double *lam = c->adj.as<double>(), *pb = lam + 3 * (size_t)n, *phi = pb + 3 * (size_t)n;
    double *qb = q_bar ? q_bar : phi + n, *rb = radius_bar ? radius_bar : phi + 2 * (size_t)n;
    double *ab = alpha_bar ? alpha_bar : phi + 3 * (size_t)n;
    // adjoint CG: r = mu_bar, x = 0, normaliser mean |alpha mu_bar|
    R4 *x = c->vx.as<R4>(), *r = c->vr.as<R4>(), *z = c->vz.as<R4>(), *p = c->vp.as<R4>();
    CGState *S = c->st.as<CGState>();
    const int npb = c->molp ? nblk(nper, MOLP_BLOCK) : nblk(nper, 256);
    k_adjoint_norm<<<1, 256, 0, st>>>(n, mu_bar, c->alpha64.as<double>(), S);
    k_mu_to_x0<real><<<nb, 256, 0, st>>>(n, mu_bar, r);
    LAUNCH_CHECK();
    if (c->molp)  // x = 0 in the start kernel
        k_cg_start_mol<real><<<dim3(nblk(nper, MOLP_BLOCK), 1), MOLP_BLOCK, 0, st>>>(
            nper, r, c->alpha_r.as<real>(), c->pinv.as<real>(), z, p, c->partB.as<double>(), nullptr, x);
    else
        k_cg_start<real><<<bead_grid(c, nper), 256, 0, st>>>(nper, r, c->alpha_r.as<real>(), z, p, c->partB.as<double>(),
                                                              nullptr, x);
    k_cg_start_finish<<<1, 256, 0, st>>>(c->partB.as<double>(), npb, S, c->rz.as<double>(), c->err.as<double>(), c->rzs, tol,
                                         c->s.max_iter, c->h_mirror, ++c->cg_seq);
    LAUNCH_CHECK();
    c->norm_on_device = false;  // CGState.norm now holds the adjoint's normaliser
    int launched = 0;
    const int it = cg_iterate<real>(c, tol, std::max(1, std::min(c->last_adj_iters, c->s.max_iter)), false, &launched);
    if (it < 0) return it;
    c->last_adj_iters = std::max(1, it);
    k_cg_finish<real><<<nb, 256, 0, st>>>(n, x, r, c->alpha_r.as<real>(), 0.0, lam);
    LAUNCH_CHECK();
    // real-space potential and radius terms (FP64, atom rows)
    if (!c->atom_rows_fresh) PGM_TRY(compact_atom_rows(c));
    k_adjoint_rows<<<nblk((long long)n * ADJ_RL, 256), 256, 0, st>>>(
        n, c->rows.as<int>(), c->count.as<int>(), c->count_near.as<int>(), c->cap, c->xr.as<double4>(), c->box, c->s.ewald_beta,
        c->rc_e * c->rc_e, c->q.as<double>(), lam, c->pperm.as<double>(), mu, phi, rb);
    k_adjoint_alpha<<<nb, 256, 0, st>>>(n, mu, lam, c->alpha64.as<double>(), ab);
    LAUNCH_CHECK();
    // p_bar = field(0, lam)
    R4 *dq = c->vdq.as<R4>();
    k_pack_dq<real><<<nb, 256, 0, st>>>(n, lam, nullptr, nullptr, dq);
    LAUNCH_CHECK();
    PGM_TRY(field_eval<real>(c, dq, false, MODE_FIELD, nullptr, c->vtmp.as<R4>(), nullptr, -1));
    k_to_double3<real><<<nb, 256, 0, st>>>(n, c->vtmp.as<R4>(), pb);
    LAUNCH_CHECK();
    // q_bar: the reciprocal potential of lam at the atoms plus the real-space part
    PGM_TRY(spread_fft_only<real>(c, dq));
    ORDER_SWITCH(c->s.order, (k_adjoint_potential<real, PO><<<nb, 256, 0, st>>>(n, pos, c->pg, c->grid.as<real>(), phi, qb)));
    if (cov_bar && c->ncov > 0)
        k_adjoint_cov<<<nblk(c->ncov, 256), 256, 0, st>>>(c->ncov, pos, c->box, c->ci.as<int>(), c->cj.as<int>(),
                                                          c->corder.as<int>(), pb, cov_bar);
    LAUNCH_CHECK();
    if (lambda) HIP_TRY(hipMemcpyAsync(lambda, lam, sizeof(double) * 3 * (size_t)n, hipMemcpyDeviceToDevice, st));
    if (p_bar) HIP_TRY(hipMemcpyAsync(p_bar, pb, sizeof(double) * 3 * (size_t)n, hipMemcpyDeviceToDevice, st));
    HIP_TRY(hipMemcpyAsync(c->h_pin_dbl, c->err.as<double>() + it, sizeof(double), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    if (iterations) *iterations = it;
    if (residual) *residual = c->h_pin_dbl[0];
    return (it >= c->s.max_iter && c->h_pin_dbl[0] > tol) ? PGM_NOT_CONVERGED : PGM_OK;
}

extern "C" int pgm_dipole_vjp(pgm_ctx *c, const double *pos, const double *mu, const double *mu_bar, double tol, double *lambda,
                              double *q_bar, double *radius_bar, double *alpha_bar, double *p_bar, double *cov_bar,
                              int *iterations, double *residual) {
    if (!c || !pos || !mu || !mu_bar || !(tol > 0)) {
        g_err = "pgm_dipole_vjp: null argument or tol <= 0";
        return PGM_ERR_ARG;
    }
    PGM_TRY(single_system(c, "pgm_dipole_vjp"));
    if (c->efield_on) {
        g_err = "pgm_dipole_vjp: the derivatives with an external field are not built";
        return PGM_ERR_STATE;
    }
    if (c->ewald) {
        g_err = "pgm_dipole_vjp: PME reciprocal space only (this context holds the exact Ewald sum)";
        return PGM_ERR_STATE;
    }
    return c->mixed ? dipole_vjp_impl<float>(c, pos, mu, mu_bar, tol, lambda, q_bar, radius_bar, alpha_bar, p_bar, cov_bar,
                                             iterations, residual)
                    : dipole_vjp_impl<double>(c, pos, mu, mu_bar, tol, lambda, q_bar, radius_bar, alpha_bar, p_bar, cov_bar,
                                              iterations, residual);
}

// The library transform of the grid in place (rocFFT r2c, x Gc, c2r): phi of the grid's Q, one bead.
template <typename real> static int library_transform(pgm_ctx *c) {
    hipStream_t st = c->stream;
    void *in[1] = {c->grid.p}, *out[1] = {c->cgrid.p};
    FFT_TRY(rocfft_execute(c->fwd, in, out, c->info));
    k_convolve<real, false><<<dim3(nblk((long long)c->ncplx, 256), 1), 256, 0, st>>>(
        c->ncplx, c->s.grid[2] / 2 + 1, c->s.grid[2] % 2 == 0, c->cgrid.as<real>(), c->Gc.as<real>(), nullptr, nullptr, -1);
    LAUNCH_CHECK();
    void *in2[1] = {c->cgrid.p}, *out2[1] = {c->grid.p};
    FFT_TRY(rocfft_execute(c->bwd, in2, out2, c->info));
    return PGM_OK;
}

// Parameter cotangents of (E, F, mu) of the differentiable force field (forcefield.py:1853-1915,
// 2489-2560 with differentiable=True): ebar dE/dtheta + f_bar . dF/dtheta + mu_bar . dmu/dtheta with
// f_bar . dF/dtheta = -D_v dE/dtheta + (KE D_v field) . dmu/dtheta (v = f_bar; adjoint.h), the
// dmu/dtheta terms by the adjoint solve of mu_bar + KE D_v field.
template <typename real>
static int compute_vjp_impl(pgm_ctx *c, const double *pos, const double *mu, double ebar, const double *f_bar,
                            const double *mu_bar, double tol, double *q_bar, double *radius_bar, double *alpha_bar,
                            double *rh_bar, double *se_bar, double *cov_bar, int *iterations, double *residual) {
    using R4 = real4<real>;
    const int n = c->n, ncov = c->ncov;
    const size_t N = (size_t)n;
    hipStream_t st = c->stream;
    const int nb = nblk(n, 256);
    PGM_TRY(prepare(c, pos));
    k_perm_dipoles<<<nb, 256, 0, st>>>(n, pos, c->box, c->cstart.as<int>(), c->cj.as<int>(), c->cc.as<double>(),
                                       c->pperm.as<double>());
    LAUNCH_CHECK();
    // work arrays: d, D_v p (3N each), row sums (16N), reciprocal sums (8N), field and D_v field (6N),
    // mu_eff (3N), tail (2N), 5 N-vectors of cotangents without the solve part, the solve's 4 + n_cov,
    // a copy of the potential grid
    const size_t ngrid_d = (c->ngrid * sizeof(real) + sizeof(double) - 1) / sizeof(double);
    PGM_TRY(c->adj2.ensure(sizeof(double) * (50 * N + 2 * (size_t)ncov + ngrid_d + 16)));
    double *d = c->adj2.as<double>(), *pdot = d + 3 * N, *rowsum = pdot + 3 * N, *rec = rowsum + 16 * N;
    double *fld = rec + 8 * N, *mu_eff = fld + 6 * N, *tail = mu_eff + 3 * N, *own = tail + 2 * N;
    double *sol = own + 5 * N, *cov_tmp = sol + 4 * N, *cov_sol = cov_tmp + ncov;
    real *gridP = reinterpret_cast<real *>(cov_sol + ncov);
    if (reinterpret_cast<char *>(gridP + c->ngrid) > static_cast<char *>(c->adj2.p) + c->adj2.bytes) {
        g_err = "pgm_compute_vjp: work array layout exceeds its allocation";
        return PGM_ERR_STATE;
    }
    double *qb = q_bar ? q_bar : own, *rb = radius_bar ? radius_bar : own + N, *ab = alpha_bar ? alpha_bar : own + 2 * N;
    double *rhb = rh_bar ? rh_bar : own + 3 * N, *seb = se_bar ? se_bar : own + 4 * N;
    const double *v = f_bar;
    k_add_vec<<<nblk(3 * (long long)n, 256), 256, 0, st>>>(3 * n, c->pperm.as<double>(), mu, d);
    if (v) k_perm_tangent<<<nb, 256, 0, st>>>(n, pos, v, c->box, c->cstart.as<int>(), c->cj.as<int>(), c->cc.as<double>(), pdot);
    LAUNCH_CHECK();
    // real-space sums over the atom rows
    if (!c->atom_rows_fresh) PGM_TRY(compact_atom_rows(c));
    k_param_rows<<<nblk((long long)n * ADJ_RL, 256), 256, 0, st>>>(
        n, c->rows.as<int>(), c->count.as<int>(), c->count_near.as<int>(), c->cap, c->xr.as<double4>(), c->box, c->s.ewald_beta,
        c->rc_e * c->rc_e, c->rc_v * c->rc_v, c->q.as<double>(), c->ljp.as<double4>(), d, pdot, v, rowsum);
    LAUNCH_CHECK();
    // reciprocal potential of (q, d) and its D_v
    R4 *dq = c->vdq.as<R4>();
    k_pack_dq<real><<<nb, 256, 0, st>>>(n, d, nullptr, c->q.as<double>(), dq);
    LAUNCH_CHECK();
    {
        const bool timing = c->timing;
        c->timing = false;
        const int rs = spread_fft<real>(c, dq, true, false, nullptr, -1);
        c->timing = timing;
        if (rs < 0) return rs;
    }
    HIP_TRY(hipMemcpyAsync(gridP, c->grid.p, c->ngrid * sizeof(real), hipMemcpyDeviceToDevice, st));
    if (v) {
        k_fill<real><<<nblk((long long)c->ngrid, 256), 256, 0, st>>>(c->ngrid, c->grid.as<real>(), real(0));
        ORDER_SWITCH(c->s.order, (k_spread_tangent<real, PO><<<nb, 256, 0, st>>>(n, pos, v, c->pg, c->q.as<double>(), d, pdot,
                                                                                 c->grid.as<real>())));
        LAUNCH_CHECK();
        PGM_TRY(library_transform<real>(c));
    }
    ORDER_SWITCH(c->s.order, (k_gather_tangent<real, PO><<<nb, 256, 0, st>>>(n, pos, v, c->pg, gridP,
                                                                             v ? c->grid.as<real>() : nullptr, rec)));
    LAUNCH_CHECK();
    // LJ tail: dE_lrc/d(R*/2), dE_lrc/d sqrt(eps) per atom (E_lrc = -2 pi SB / (3 V rc^3), SB = 2 sum_k
    // C(6,k) M_k M_(6-k), M_k = sum sqrt(eps) (R*/2)^k)
    std::vector<double> ht;
    if (ebar != 0.0 && c->s.lj_lrc) {
        double M[7] = {0};
        for (int i = 0; i < c->nper; ++i) {
            double pw = c->h_se[i];
            for (int k = 0; k < 7; ++k) {
                M[k] += pw;
                pw *= c->h_rh[i];
            }
        }
        static const double C6[7] = {1, 6, 15, 20, 15, 6, 1};
        const double c0 = -2.0 * M_PI / (3.0 * c->volume * c->s.cutoff * c->s.cutoff * c->s.cutoff);
        ht.assign(2 * N, 0.0);
        for (int i = 0; i < n; ++i) {
            const double rh = c->h_rh[i], se = c->h_se[i];
            double pk = 1.0, pk1 = 0.0;  // rh^k, rh^(k-1)
            for (int k = 0; k < 7; ++k) {
                const double dSB = 4.0 * C6[k] * M[6 - k];
                ht[2 * i + 1] += c0 * dSB * pk;
                if (k) ht[2 * i] += c0 * dSB * k * se * pk1;
                pk1 = pk;
                pk *= rh;
            }
        }
        HIP_TRY(hipMemcpyAsync(tail, ht.data(), sizeof(double) * 2 * N, hipMemcpyHostToDevice, st));
    }
    const double b0 = c->s.ewald_beta;
    k_vjp_atoms<<<nb, 256, 0, st>>>(n, rowsum, rec, c->q.as<double>(), c->xr.as<double4>(), c->alpha64.as<double>(), mu, d,
                                    v ? pdot : nullptr, mu_bar, ht.empty() ? nullptr : tail, ebar,
                                    -M_PI * c->sum_q / (c->volume * b0 * b0), -2.0 * b0 / SQRT_PI, c_self(c), qb, rb, ab, rhb,
                                    seb, fld, mu_eff);
    if (ncov > 0)
        k_vjp_cov<<<nblk(ncov, 256), 256, 0, st>>>(ncov, pos, v, c->box, c->ci.as<int>(), c->cj.as<int>(), c->corder.as<int>(),
                                                    fld, ebar, cov_bar ? cov_bar : cov_tmp);
    LAUNCH_CHECK();
    int it = 0;
    double res = 0.0;
    int rc = PGM_OK;
    if (v || mu_bar) {  // the induced dipoles' part: the adjoint solve of mu_eff
        rc = dipole_vjp_impl<real>(c, pos, mu, mu_eff, tol, nullptr, sol, sol + N, sol + 2 * N, nullptr, cov_sol, &it, &res);
        if (rc < 0) return rc;
        k_add<<<nb, 256, 0, st>>>(n, sol, qb);
        k_add<<<nb, 256, 0, st>>>(n, sol + N, rb);
        k_add<<<nb, 256, 0, st>>>(n, sol + 2 * N, ab);
        if (cov_bar && ncov > 0) k_add<<<nblk(ncov, 256), 256, 0, st>>>(ncov, cov_sol, cov_bar);
        LAUNCH_CHECK();
    }
    HIP_TRY(hipStreamSynchronize(st));
    if (iterations) *iterations = it;
    if (residual) *residual = res;
    return rc;
}

extern "C" int pgm_compute_vjp(pgm_ctx *c, const double *pos, const double *mu, double e_bar, const double *f_bar,
                               const double *mu_bar, double tol, double *q_bar, double *radius_bar, double *alpha_bar,
                               double *rmin_half_bar, double *sqrt_eps_bar, double *cov_bar, int *iterations, double *residual) {
    if (!c || !pos || !mu || !(tol > 0) || !std::isfinite(e_bar)) {
        g_err = "pgm_compute_vjp: null argument, tol <= 0 or e_bar not finite";
        return PGM_ERR_ARG;
    }
    PGM_TRY(single_system(c, "pgm_compute_vjp"));
    if (c->efield_on) {
        g_err = "pgm_compute_vjp: the derivatives with an external field are not built";
        return PGM_ERR_STATE;
    }
    if (c->ewald) {
        g_err = "pgm_compute_vjp: PME reciprocal space only (this context holds the exact Ewald sum)";
        return PGM_ERR_STATE;
    }
    return c->mixed ? compute_vjp_impl<float>(c, pos, mu, e_bar, f_bar, mu_bar, tol, q_bar, radius_bar, alpha_bar, rmin_half_bar,
                                              sqrt_eps_bar, cov_bar, iterations, residual)
                    : compute_vjp_impl<double>(c, pos, mu, e_bar, f_bar, mu_bar, tol, q_bar, radius_bar, alpha_bar, rmin_half_bar,
                                               sqrt_eps_bar, cov_bar, iterations, residual);
}

// ---------------------------------------------------------------------------- external field
// The uniform static field of forcefield.py:2048-2097 (md/efield.py, the constant-field kind): F0 =
// E [V/nm] x FARADAY / KE in the model's e/nm^2 joins the permanent field of every atom in the
// induction equations, the energy gains -KE F0 . M with M = sum q r + sum (p + mu) + offset, every
// charge feels KE q F0 and every covalent dipole the torque of the field.
constexpr double FARADAY_KJ = 96.48533212331002;  // kJ/mol per (e V)

extern "C" int pgm_set_efield(pgm_ctx *c, const double *E_vnm, const double *offset) {
    if (!c) return PGM_ERR_ARG;
    if (c->ewald && E_vnm) {
        g_err = "pgm_set_efield: PME reciprocal space only (this context holds the exact Ewald sum)";
        return PGM_ERR_STATE;
    }
    if (!E_vnm) {
        c->efield_on = false;
        c->F0 = Ext3{};
        c->foff = Ext3{};
        return PGM_OK;
    }
    for (int k = 0; k < 3; ++k)
        if (!std::isfinite(E_vnm[k]) || (offset && !std::isfinite(offset[k]))) {
            g_err = "pgm_set_efield: non-finite field or offset";
            return PGM_ERR_ARG;
        }
    const double s = FARADAY_KJ / KE;
    PGM_TRY(c->fpart.ensure(sizeof(double) * 3 * (size_t)c->nbead * nblk(c->nper, 256)));
    PGM_TRY(c->fout.ensure(sizeof(double) * 4 * (size_t)c->nbead));
    HIP_TRY(hipMemsetAsync(c->fout.p, 0, sizeof(double) * 4 * (size_t)c->nbead, c->stream));
    c->F0 = Ext3{s * E_vnm[0], s * E_vnm[1], s * E_vnm[2]};
    c->foff = offset ? Ext3{offset[0], offset[1], offset[2]} : Ext3{};
    c->efield_on = true;
    return PGM_OK;
}

extern "C" int pgm_field_result(pgm_ctx *c, double *e_field, double *dipole) {
    if (!c || !e_field || !dipole) return PGM_ERR_ARG;
    if (!c->efield_on) {
        g_err = "pgm_field_result: no external field is set";
        return PGM_ERR_ARG;
    }
    const int B = c->nbead;
    std::vector<double> h(4 * (size_t)B);
    HIP_TRY(hipMemcpyAsync(h.data(), c->fout.p, sizeof(double) * 4 * B, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    for (int b = 0; b < B; ++b) {
        e_field[b] = h[4 * b];
        for (int k = 0; k < 3; ++k) dipole[3 * b + k] = h[4 * b + 1 + k];
    }
    return PGM_OK;
}

// ---------------------------------------------------------------------------- parameters
// New per-atom parameters and covalent-dipole strengths (host arrays of one bead; NULL: unchanged):
// the parameter pytree of PGMForceField.compute(..., params) (forcefield.py:2489-2560, System.expand
// of pgm_jax/system.py:561-577).  Every quantity derived from them is refreshed: device copies, the
// charge sums of the self and background terms, the LJ tail's moment sum, the packed LJ records, the
// molecule-block preconditioner (rebuilt at the next evaluation).  Synchronises.
extern "C" int pgm_set_params(pgm_ctx *c, const double *q, const double *radius, const double *alpha,
                              const double *lj_rmin_half, const double *lj_sqrt_eps, const double *cov_c) {
    if (!c) return PGM_ERR_ARG;
    const int n = c->n, nper = c->nper;
    for (int i = 0; i < nper; ++i) {
        if ((q && !std::isfinite(q[i])) || (radius && !(radius[i] > 0 && std::isfinite(radius[i]))) ||
            (alpha && !(alpha[i] >= 0 && std::isfinite(alpha[i]))) || (lj_rmin_half && !std::isfinite(lj_rmin_half[i])) ||
            (lj_sqrt_eps && !std::isfinite(lj_sqrt_eps[i]))) {
            g_err = "pgm_set_params: atom " + std::to_string(i) +
                    ": parameters must be finite, radii > 0 and polarizabilities >= 0";
            return PGM_ERR_ARG;
        }
    }
    const int nc1 = c->nbead ? c->ncov / c->nbead : 0;
    if (cov_c)
        for (int m = 0; m < nc1; ++m)
            if (!std::isfinite(cov_c[m])) {
                g_err = "pgm_set_params: covalent dipole strengths must be finite";
                return PGM_ERR_ARG;
            }
    hipStream_t st = c->stream;
    HIP_TRY(hipStreamSynchronize(st));  // earlier uploads from the host copies have completed
    if (q) {
        c->sum_q = c->sum_q2 = 0.0;
        for (int i = 0; i < n; ++i) c->h_q[i] = q[i % nper];
        for (int i = 0; i < nper; ++i) {
            c->sum_q += q[i];
            c->sum_q2 += q[i] * q[i];
        }
        PGM_TRY(upload(c->q, c->h_q.data(), n, st));
    }
    if (radius) {
        for (int i = 0; i < n; ++i) c->h_R2[i] = radius[i % nper] * radius[i % nper];
        PGM_TRY(upload(c->R2, c->h_R2.data(), n, st));
    }
    std::vector<float> af;
    if (alpha) {
        c->alpha0 = false;
        for (int i = 0; i < n; ++i) {
            c->h_alpha[i] = alpha[i % nper];
            if (c->h_alpha[i] == 0) c->alpha0 = true;
        }
        PGM_TRY(upload(c->alpha64, c->h_alpha.data(), n, st));
        if (c->mixed) {
            af.assign(c->h_alpha.begin(), c->h_alpha.end());
            PGM_TRY(upload(c->alpha_r, af.data(), n, st));
        } else {
            PGM_TRY(upload(c->alpha_r, c->h_alpha.data(), n, st));
        }
        c->pinv_stale = true;
    }
    std::vector<double> lj;
    if (lj_rmin_half || lj_sqrt_eps) {
        lj.resize(4 * (size_t)n);
        for (int i = 0; i < n; ++i) {
            if (lj_rmin_half) c->h_rh[i] = lj_rmin_half[i % nper];
            if (lj_sqrt_eps) c->h_se[i] = lj_sqrt_eps[i % nper];
            lj[4 * (size_t)i] = c->h_rh[i];
            lj[4 * (size_t)i + 1] = c->h_se[i];
            lj[4 * (size_t)i + 2] = (double)c->h_mol[i];
            lj[4 * (size_t)i + 3] = 0.0;
        }
        c->lj_sb = lj_sum_b(c);
        PGM_TRY(upload(c->rh, c->h_rh.data(), n, st));
        PGM_TRY(upload(c->se, c->h_se.data(), n, st));
        PGM_TRY(upload(c->ljp, lj.data(), lj.size(), st));
    }
    std::vector<double> cc;
    if (cov_c && c->ncov > 0) {
        cc.resize(c->ncov);
        for (int k = 0; k < c->ncov; ++k) cc[k] = cov_c[c->h_corder[k] % nc1];
        PGM_TRY(upload(c->cc, cc.data(), c->ncov, st));
    }
    HIP_TRY(hipStreamSynchronize(st));
    return PGM_OK;
}

// ---------------------------------------------------------------------------- learned guess (ABI)
extern "C" int pgm_guess_config(pgm_ctx *c, int hidden_layers, double dipole_scale, int harvest_every, int capacity,
                                unsigned long long seed) {
    if (!c || hidden_layers < 1 || hidden_layers > 4 || !(dipole_scale > 0) || harvest_every < 1 || capacity < c->n) {
        g_err = "pgm_guess_config: 1..4 hidden layers, dipole_scale > 0, harvest_every >= 1, capacity >= n_atoms";
        return PGM_ERR_ARG;
    }
    PGM_TRY(single_system(c, "pgm_guess_config"));
    if (c->n % 3 != 0) {
        g_err = "pgm_guess_config: the learned guess needs consecutive 3-site molecules";
        return PGM_ERR_ARG;
    }
    for (int i = 0; i < c->n; ++i)
        if (c->h_mol[i] != i / 3) {
            g_err = "pgm_guess_config: the learned guess needs consecutive 3-site molecules";
            return PGM_ERR_ARG;
        }
    auto &G = c->guess;
    for (DevBuf *b : {&G.params, &G.grads, &G.m, &G.v, &G.packed, &G.bias, &G.feat, &G.frames, &G.y, &G.bx, &G.by, &G.idx,
                      &G.xb, &G.yb, &G.acts, &G.zs, &G.delta, &G.delta2, &G.part, &G.scal, &G.ring})
        b->release();
    G = pgm_ctx::Guess();
    G.nh = hidden_layers;
    G.scale = dipole_scale;
    G.every = harvest_every;
    G.cap = capacity;
    G.seed = seed;
    const std::vector<int> sz = guess_sizes(c);
    long long np = 0, packed = 0, nbias = 0;
    for (size_t l = 0; l + 1 < sz.size(); ++l) {
        np += (long long)sz[l] * sz[l + 1] + sz[l + 1];
        packed += (long long)((sz[l + 1] + 15) / 16) * ((sz[l] + 15) / 16) * 256;
        nbias += sz[l + 1];
    }
    G.nparams = np;
    for (DevBuf *b : {&G.params, &G.grads, &G.m, &G.v}) PGM_TRY(b->ensure(sizeof(float) * np));
    PGM_TRY(G.packed.ensure(sizeof(short) * packed));
    PGM_TRY(G.bias.ensure(sizeof(float) * (nbias + 16)));
    PGM_TRY(G.feat.ensure(sizeof(float) * (size_t)c->n * GUESS_D));
    PGM_TRY(G.frames.ensure(sizeof(double) * (size_t)c->n * 9));
    PGM_TRY(G.ring.ensure(sizeof(double) * (size_t)GUESS_K * 3 * c->n));
    PGM_TRY(G.bx.ensure(sizeof(float) * (size_t)capacity * GUESS_D));
    PGM_TRY(G.by.ensure(sizeof(float) * (size_t)capacity * GUESS_OUT));
    PGM_TRY(G.scal.ensure(sizeof(double) * 4));
    PGM_TRY(guess_alloc_batch(c, 1024));
    // LeCun-normal weights (Box-Muller on a 64-bit Mersenne twister keyed by seed), zero biases
    std::vector<float> h((size_t)np, 0.0f);
    std::mt19937_64 rng(seed);
    auto normal = [&]() {
        const double u1 = ((double)(rng() >> 11) + 0.5) * (1.0 / 9007199254740992.0);
        const double u2 = ((double)(rng() >> 11) + 0.5) * (1.0 / 9007199254740992.0);
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    };
    long long off = 0;
    for (size_t l = 0; l + 1 < sz.size(); ++l) {
        const double sd = 1.0 / std::sqrt((double)sz[l]);
        for (long long k = 0; k < (long long)sz[l] * sz[l + 1]; ++k) h[(size_t)(off + k)] = (float)(sd * normal());
        off += (long long)sz[l] * sz[l + 1] + sz[l + 1];
    }
    hipStream_t st = c->stream;
    HIP_TRY(hipMemcpyAsync(G.params.p, h.data(), sizeof(float) * np, hipMemcpyHostToDevice, st));
    k_fill<float><<<nblk(np, 256), 256, 0, st>>>((size_t)np, G.m.as<float>(), 0.0f);
    k_fill<float><<<nblk(np, 256), 256, 0, st>>>((size_t)np, G.v.as<float>(), 0.0f);
    LAUNCH_CHECK();
    PGM_TRY(guess_pack(c));
    HIP_TRY(hipStreamSynchronize(st));
    G.on = true;
    return PGM_OK;
}

extern "C" int pgm_guess_params(pgm_ctx *c, float *host, long long count, int set) {
    if (!c || !c->guess.on || !host || count != c->guess.nparams) {
        g_err = "pgm_guess_params: configure first; count must equal the number of parameters";
        return PGM_ERR_ARG;
    }
    hipStream_t st = c->stream;
    if (set) {
        HIP_TRY(hipMemcpyAsync(c->guess.params.p, host, sizeof(float) * count, hipMemcpyHostToDevice, st));
        PGM_TRY(guess_pack(c));
        c->guess.trained = true;
    } else {
        HIP_TRY(hipMemcpyAsync(host, c->guess.params.p, sizeof(float) * count, hipMemcpyDeviceToHost, st));
    }
    HIP_TRY(hipStreamSynchronize(st));
    return PGM_OK;
}

extern "C" long long pgm_guess_nparams(pgm_ctx *c) { return c && c->guess.on ? c->guess.nparams : -1; }

extern "C" int pgm_guess_forward(pgm_ctx *c, const float *x, int n, float *y, int bf16) {
    if (!c || !c->guess.on || !x || !y || n <= 0) return PGM_ERR_ARG;
    if (bf16) {
        PGM_TRY(guess_infer(c, x, n, y));
    } else {
        PGM_TRY(guess_alloc_batch(c, n));
        PGM_TRY(guess_forward(c, x, n, y));
    }
    HIP_TRY(hipStreamSynchronize(c->stream));
    return PGM_OK;
}

extern "C" int pgm_guess_grad(pgm_ctx *c, const float *x, const float *t, int n, float *grads, double *loss) {
    if (!c || !c->guess.on || !x || !t || !grads || n <= 0) return PGM_ERR_ARG;
    auto &G = c->guess;
    hipStream_t st = c->stream;
    PGM_TRY(guess_alloc_batch(c, n));
    PGM_TRY(guess_forward(c, x, n, G.y.as<float>()));
    PGM_TRY(guess_backward(c, x, t, n, G.y.as<float>()));
    const int np = nblk((long long)n * GUESS_OUT, 256);
    k_reduce_partials<<<1, 256, 0, st>>>(G.part.as<double>(), np, 0, G.scal.as<double>() + 2);
    LAUNCH_CHECK();
    HIP_TRY(hipMemcpyAsync(grads, G.grads.p, sizeof(float) * G.nparams, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(c->h_pin_dbl + 22, G.scal.as<double>() + 2, sizeof(double), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    if (loss) *loss = c->h_pin_dbl[22] / ((double)n * GUESS_OUT);
    return PGM_OK;
}

extern "C" int pgm_guess_train(pgm_ctx *c, int steps, int batch, double lr, double *loss) {
    if (!c || !c->guess.on || steps < 0 || batch < 16 || !(lr > 0)) return PGM_ERR_ARG;
    auto &G = c->guess;
    const long long filled = std::min<long long>(G.rows, G.cap);
    if (filled <= 0) {
        g_err = "pgm_guess_train: no harvested rows yet";
        return PGM_ERR_STATE;
    }
    hipStream_t st = c->stream;
    PGM_TRY(guess_alloc_batch(c, batch));
    k_fill<double><<<1, 64, 0, st>>>(1, G.scal.as<double>() + 3, 0.0);
    const double b1 = 0.9, b2 = 0.999;
    for (int k = 0; k < steps; ++k) {
        k_sample_rows<<<nblk(batch, 256), 256, 0, st>>>(batch, filled, G.seed, (uint32_t)G.adam, G.idx.as<int>());
        k_guess_batch<<<nblk(batch, 256), 256, 0, st>>>(batch, G.idx.as<int>(), G.bx.as<float>(), G.by.as<float>(),
                                                         G.xb.as<float>(), G.yb.as<float>());
        PGM_TRY(guess_forward(c, G.xb.as<float>(), batch, G.y.as<float>()));
        PGM_TRY(guess_backward(c, G.xb.as<float>(), G.yb.as<float>(), batch, G.y.as<float>()));
        k_accumulate_partials<<<1, 256, 0, st>>>(G.part.as<double>(), nblk((long long)batch * GUESS_OUT, 256), G.scal.as<double>() + 3);
        G.adam += 1;
        const float c1 = (float)(1.0 - std::pow(b1, G.adam)), c2 = (float)(1.0 - std::pow(b2, G.adam));
        k_adam<<<nblk(G.nparams, 256), 256, 0, st>>>((size_t)G.nparams, G.params.as<float>(), G.grads.as<float>(), G.m.as<float>(),
                                                     G.v.as<float>(), (float)lr, (float)b1, (float)b2, 1e-8f, c1, c2);
        LAUNCH_CHECK();
    }
    PGM_TRY(guess_pack(c));
    HIP_TRY(hipMemcpyAsync(c->h_pin_dbl + 23, G.scal.as<double>() + 3, sizeof(double), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    if (loss) *loss = steps > 0 ? c->h_pin_dbl[23] / ((double)steps * batch * GUESS_OUT) : 0.0;
    G.trained = true;
    return PGM_OK;
}

// out[16]: rows harvested, sigma, squared errors per component at the last harvest of mu4, the
// linear predictor and linear + network, mode (0 mu4, 1 linear, 2 linear + network), trained,
// fits, and the eight linear coefficients (newest first).
extern "C" int pgm_guess_status(pgm_ctx *c, double *out) {
    if (!c || !out) return PGM_ERR_ARG;
    const auto &G = c->guess;
    out[0] = (double)G.rows;
    out[1] = G.sigma;
    out[2] = G.err_mu4;
    out[3] = G.err_lin;
    out[4] = G.err_net;
    out[5] = G.mode;
    out[6] = G.trained ? 1.0 : 0.0;
    out[7] = G.nfit;
    for (int k = 0; k < GUESS_K; ++k) out[8 + k] = G.lin.w[k];
    return PGM_OK;
}

// ---------------------------------------------------------------------------- live kernel timing
extern "C" int pgm_set_timing(pgm_ctx *c, int on) {
    if (!c) return PGM_ERR_ARG;
    PGM_TRY(harvest_timing(c));
    for (auto &t : c->timed) {
        t.sum_ms = 0.0;
        t.count = 0;
    }
    c->timing = on != 0;
    return PGM_OK;
}

extern "C" int pgm_kernel_timing(pgm_ctx *c, double *avg_ms, long long *launches) {
    if (!c || !avg_ms || !launches) return PGM_ERR_ARG;
    PGM_TRY(harvest_timing(c));
    for (int k = 0; k < 3; ++k) {
        launches[k] = c->timed[k].count;
        avg_ms[k] = c->timed[k].count ? c->timed[k].sum_ms / (double)c->timed[k].count : 0.0;
    }
    return PGM_OK;
}

// ---------------------------------------------------------------------------- dense induction solve
// The CG of the dense solve on caller-provided device buffers: y, r, z, p (3n doubles each), part
// (2 max(n, blocks) doubles), out (2 doubles), and a host array h of 4 doubles for the scalars.
static int dense_solve(int n, const double *T, const double *alpha, const double *F, double *mu, double tol, int max_iter,
                       hipStream_t st, int *iterations, double *residual, double *y, double *r, double *z, double *p,
                       double *part, double *out, double *h) {
    const int n3 = 3 * n, nb = nblk(n3, 256);
    auto scalars = [&](int np, int stride) -> int {
        k_dense_reduce<<<1, 256, 0, st>>>(part, np, stride, out);
        if (hipGetLastError() != hipSuccess || hipMemcpyAsync(h, out, stride * sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess) {
            g_err = "pgm_solve_linear_induction: device error";
            return PGM_ERR_HIP;
        }
        return PGM_OK;
    };
    double *x = mu;
    // x0 = alpha F, r = F - A x0, z = alpha r, p = z
    k_dense_update<<<nb, 256, 0, st>>>(n3, 3, 0.0, 0.0, F, alpha, nullptr, x, nullptr, nullptr, nullptr, part);
    PGM_TRY(scalars(nb, 2));
    const double fnorm = std::sqrt(h[0]);
    k_dense_matvec<<<n, 256, 0, st>>>(n, T, alpha, x, y, part);
    k_dense_update<<<nb, 256, 0, st>>>(n3, 0, 0.0, 0.0, F, alpha, y, x, r, z, p, part);
    PGM_TRY(scalars(nb, 2));
    double rz = h[0], rr = h[1];
    int it = 0;
    // stopping rule of jax.scipy.sparse.linalg.cg: |r| <= tol |F|
    while (it < max_iter && std::sqrt(rr) > tol * fnorm) {
        k_dense_matvec<<<n, 256, 0, st>>>(n, T, alpha, p, y, part);
        PGM_TRY(scalars(n, 1));
        const double al = rz / std::max(h[0], 1e-300);
        k_dense_update<<<nb, 256, 0, st>>>(n3, 1, al, 0.0, F, alpha, y, x, r, z, p, part);
        PGM_TRY(scalars(nb, 2));
        const double be = h[0] / std::max(rz, 1e-300);
        rz = h[0];
        rr = h[1];
        k_dense_update<<<nb, 256, 0, st>>>(n3, 2, 0.0, be, F, alpha, nullptr, x, nullptr, z, p, nullptr);
        ++it;
    }
    if (hipStreamSynchronize(st) != hipSuccess) {
        g_err = "pgm_solve_linear_induction: device error";
        return PGM_ERR_HIP;
    }
    if (iterations) *iterations = it;
    if (residual) *residual = fnorm > 0 ? std::sqrt(rr) / fnorm : 0.0;
    return std::sqrt(rr) > tol * fnorm ? PGM_NOT_CONVERGED : PGM_OK;
}

static bool dense_args_ok(int n, const double *T, const double *alpha, const double *F, double *mu, double tol, int max_iter) {
    if (n <= 0 || !T || !alpha || !F || !mu || !(tol > 0) || max_iter < 0) {
        g_err = "pgm_solve_linear_induction: invalid argument";
        return false;
    }
    return true;
}

extern "C" int pgm_solve_linear_induction(int n, const double *T, const double *alpha, const double *F, double *mu,
                                          double tol, int max_iter, void *stream, int *iterations, double *residual) {
    if (!dense_args_ok(n, T, alpha, F, mu, tol, max_iter)) return PGM_ERR_ARG;
    const int n3 = 3 * n, nb = nblk(n3, 256);
    DevBuf y, r, z, p, part, out;
    double *h = nullptr;
    auto cleanup = [&](int rc) {
        if (h) (void)hipHostFree(h);
        for (DevBuf *b : {&y, &r, &z, &p, &part, &out}) b->release();
        return rc;
    };
    int rc;
    for (DevBuf *b : {&y, &r, &z, &p})
        if ((rc = b->ensure(sizeof(double) * n3)) < 0) return cleanup(rc);
    if ((rc = part.ensure(sizeof(double) * 2 * std::max(n, nb))) < 0) return cleanup(rc);
    if ((rc = out.ensure(sizeof(double) * 2)) < 0) return cleanup(rc);
    if (hipHostMalloc((void **)&h, 4 * sizeof(double), hipHostMallocDefault) != hipSuccess) {
        g_err = "hipHostMalloc failed";
        return cleanup(PGM_ERR_HIP);
    }
    return cleanup(dense_solve(n, T, alpha, F, mu, tol, max_iter, (hipStream_t)stream, iterations, residual, y.as<double>(),
                               r.as<double>(), z.as<double>(), p.as<double>(), part.as<double>(), out.as<double>(), h));
}

// The dense solve without allocations: its work vectors live in `scratch`, a file-scope device
// array of fixed size that every GPU holds once for the life of the process.  Host code must never
// pass `scratch` itself to a kernel or a copy (in host code the name is a stand-in, not a device
// address): its address on each device comes from hipGetSymbolAddress, taken once per device with
// that device current and kept in scratch_dev[].  Every vector is written before it is read, so
// nothing needs clearing between calls; calls that overlap on one GPU share the array and race.
static constexpr size_t align32(size_t v) { return (v + 31) / 32 * 32; }
static constexpr size_t DENSE_SCRATCH_DOUBLES =
    4 * align32(3 * (size_t)PGM_DENSE_NOALLOC_MAX_ATOMS) + align32(2 * (size_t)PGM_DENSE_NOALLOC_MAX_ATOMS) + 32;
static __device__ __align__(256) double scratch[DENSE_SCRATCH_DOUBLES];
static constexpr int SCRATCH_MAX_DEVICES = 64;
static double *scratch_dev[SCRATCH_MAX_DEVICES] = {};
static std::mutex scratch_mu;

static int scratch_address(double **out) {
    int dev = -1;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= SCRATCH_MAX_DEVICES) {
        g_err = "pgm_solve_linear_induction_noalloc: device index beyond 64";
        return PGM_ERR_ARG;
    }
    std::lock_guard<std::mutex> lock(scratch_mu);
    if (!scratch_dev[dev]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) {
            g_err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
            return PGM_ERR_HIP;
        }
        scratch_dev[dev] = (double *)ptr;
    }
    *out = scratch_dev[dev];
    return PGM_OK;
}

extern "C" int pgm_solve_linear_induction_noalloc(int n, const double *T, const double *alpha, const double *F, double *mu,
                                                  double tol, int max_iter, void *stream, int *iterations, double *residual) {
    if (!dense_args_ok(n, T, alpha, F, mu, tol, max_iter)) return PGM_ERR_ARG;
    if (n > PGM_DENSE_NOALLOC_MAX_ATOMS) {
        g_err = "pgm_solve_linear_induction_noalloc: n = " + std::to_string(n) + " atoms exceeds the scratch limit of " +
                std::to_string(PGM_DENSE_NOALLOC_MAX_ATOMS);
        return PGM_ERR_ARG;
    }
    double *base = nullptr;
    PGM_TRY(scratch_address(&base));
    const size_t v = align32(3 * (size_t)n);
    double *y = base, *r = base + v, *z = base + 2 * v, *p = base + 3 * v, *part = base + 4 * v;
    double *out = part + align32(2 * (size_t)n);  // part holds 2 max(n, blocks) = 2n doubles (blocks <= n)
    double h[4] = {0.0, 0.0, 0.0, 0.0};
    return dense_solve(n, T, alpha, F, mu, tol, max_iter, (hipStream_t)stream, iterations, residual, y, r, z, p, part, out, h);
}

// ---------------------------------------------------------------------------- MD
extern "C" int pgm_md_setup(pgm_ctx *c, const double *masses, const double *pos) {
    if (!c || !masses || !pos) return PGM_ERR_ARG;
    const int n = c->n;
    if (n % 3 != 0) {
        g_err = "pgm_md_setup: rigid MD needs 3-site molecules";
        return PGM_ERR_ARG;
    }
    for (int i = 0; i < n; ++i)
        if (c->h_mol[i] != i / 3 || !(masses[i] > 0)) {
            g_err = "pgm_md_setup: molecules must be consecutive 3-atom groups with positive masses";
            return PGM_ERR_ARG;
        }
    c->nmol = n / 3;
    std::vector<double> minv(n);
    for (int i = 0; i < n; ++i) minv[i] = 1.0 / masses[i];
    PGM_TRY(upload(c->mass, masses, n, c->stream));
    PGM_TRY(upload(c->minv, minv.data(), n, c->stream));
    // constraint lengths from the given geometry (squared, per molecule)
    std::vector<double> hp(3 * (size_t)n), d2(3 * (size_t)c->nmol);
    HIP_TRY(hipMemcpyAsync(hp.data(), pos, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    const int A[3] = {0, 0, 1}, B[3] = {1, 2, 2};
    for (int m = 0; m < c->nmol; ++m)
        for (int k = 0; k < 3; ++k) {
            double s = 0;
            for (int x = 0; x < 3; ++x) {
                const double d = hp[3 * (3 * m + A[k]) + x] - hp[3 * (3 * m + B[k]) + x];
                s += d * d;
            }
            d2[3 * m + k] = s;
        }
    PGM_TRY(upload(c->d2, d2.data(), d2.size(), c->stream));
    PGM_TRY(c->md_part.ensure(sizeof(double) * (nblk(n, 256) + 16)));
    PGM_TRY(c->md_fail.ensure(sizeof(int) * 4));
    k_fill_int<<<1, 64, 0, c->stream>>>(1, c->md_fail.as<int>(), 0);
    LAUNCH_CHECK();
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->md_ready = true;
    return PGM_OK;
}

static int md_kinetic(pgm_ctx *c, double *ke) {
    const int nb = nblk(c->nmol, 256);
    k_reduce_partials<<<1, 256, 0, c->stream>>>(c->md_part.as<double>(), nb, 0, aux_slots(c));
    LAUNCH_CHECK();
    HIP_TRY(hipMemcpyAsync(c->h_pin_dbl + 8, aux_slots(c), sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    *ke = 0.5 * c->h_pin_dbl[8];
    return PGM_OK;
}

// First half of a step: NVE (half kick, drift, SHAKE) or, with a Langevin friction, the g-BAOAB
// half kick / drift / O / drift with SHAKE and RATTLE.
static int md_first(pgm_ctx *c, pgm_md_state *s, double dt) {
    hipStream_t st = c->stream;
    if (c->friction > 0.0) {
        const int nb = nblk(c->nmol, 256);
        k_md_first_langevin<<<nb, 256, 0, st>>>(c->nmol, s->pos, s->vel, s->forces, c->minv.as<double>(), c->d2.as<double>(),
                                                dt, std::exp(-c->friction * dt), c->kT, c->seed, (uint32_t)s->step,
                                                c->md_fail.as<int>(), c->md_part.as<double>());
        LAUNCH_CHECK();
        k_accumulate_partials<<<1, 256, 0, st>>>(c->md_part.as<double>(), nb, c->heat.as<double>());
    } else {
        k_md_first<<<nblk(c->nmol, 128), 128, 0, st>>>(c->nmol, s->pos, s->vel, s->forces, c->minv.as<double>(),
                                                       c->d2.as<double>(), dt, c->md_fail.as<int>());
    }
    LAUNCH_CHECK();
    return PGM_OK;
}

extern "C" int pgm_md_set_thermostat(pgm_ctx *c, double friction, double temperature, unsigned long long seed) {
    if (!c || !(friction >= 0.0) || !(temperature >= 0.0)) {
        g_err = "pgm_md_set_thermostat: friction and temperature must be >= 0";
        return PGM_ERR_ARG;
    }
    c->friction = friction;
    c->kT = 0.0083144626181532 * temperature;
    c->seed = seed;
    return PGM_OK;
}

extern "C" int pgm_md_run(pgm_ctx *c, pgm_md_state *s, int nsteps, double dt, pgm_md_report *rep) {
    if (!c || !s || !s->pos || !s->vel || !s->forces || nsteps < 0 || !(dt > 0)) return PGM_ERR_ARG;
    if (!c->md_ready) {
        g_err = "pgm_md_run: call pgm_md_setup first";
        return PGM_ERR_STATE;
    }
    const int n = c->n;
    hipStream_t st = c->stream;
    pgm_energy e{};
    int it = 0;
    double res = 0.0;
    long long it_total = 0;
    int it_max = 0;
    const int rebuilds0 = c->rebuilds;
    const long long slot0 = c->stats_slot;
    PGM_TRY(c->heat.ensure(sizeof(double)));
    k_fill<double><<<1, 64, 0, st>>>(1, c->heat.as<double>(), 0.0);
    LAUNCH_CHECK();
    if (!s->have_forces) {
        const int r = c->mixed ? compute_api<float>(c, s->pos, &s->ind, &e, s->forces, &it, &res, true)
                               : compute_api<double>(c, s->pos, &s->ind, &e, s->forces, &it, &res, true);
        if (r < 0) return r;
        s->have_forces = 1;
    }
    int status = PGM_OK;
    // The displacement test of the rows is read one step late (asynchronously), so it fires at
    // half the skin minus a margin that covers one step's motion.
    const double trigger = 0.5 * c->s.skin - 0.01;
    bool disp_pending = false;
    if (nsteps > 0 && c->rows_valid) {
        // Earlier calls (or the caller) may have moved the atoms since the row build: the first step
        // of every call tests the displacement of the current positions, read back synchronously.
        k_max_disp<<<nblk(n, 256), 256, 0, st>>>(n, s->pos, c->pos_build.as<double>(), c->md_part.as<double>());
        k_reduce_partials<<<1, 256, 0, st>>>(c->md_part.as<double>(), nblk(n, 256), 1, aux_slots(c) + 1);
        LAUNCH_CHECK();
        HIP_TRY(hipMemcpyAsync(c->h_pin_dbl + 9, aux_slots(c) + 1, sizeof(double), hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        disp_pending = true;
    }
    for (int k = 0; k < nsteps; ++k) {
        if (disp_pending) {  // the value written last step is in: the CG waited on later work of the stream
            const double disp = k == 0 ? c->h_pin_dbl[9] : *(volatile double *)h_disp(c);
            if (!c->rows_valid || disp > trigger) {
                PGM_TRY(md_first(c, s, dt));
                PGM_TRY(build_rows(c, s->pos));
                goto forces;
            }
        }
        PGM_TRY(md_first(c, s, dt));
    forces:
        k_max_disp<<<nblk(n, 256), 256, 0, st>>>(n, s->pos, c->pos_build.as<double>(), c->md_part.as<double>());
        k_reduce_partials<<<1, 256, 0, st>>>(c->md_part.as<double>(), nblk(n, 256), 1, aux_slots(c) + 1, h_disp(c));
        LAUNCH_CHECK();
        disp_pending = true;
        {
            const int rebuilds_before = c->rebuilds;
            const int r = c->mixed ? compute_api<float>(c, s->pos, &s->ind, &e, s->forces, &it, &res, true)
                                   : compute_api<double>(c, s->pos, &s->ind, &e, s->forces, &it, &res, true);
            if (r < 0) return r;
            if (r == PGM_NOT_CONVERGED) status = r;
            // rows rebuilt inside (they were invalid, e.g. after a box change): the displacement just
            // queued was measured against the previous build, and the new build is at these positions
            if (c->rebuilds != rebuilds_before) disp_pending = false;
        }
        it_total += it;
        it_max = std::max(it_max, it);
        k_md_second<<<nblk(c->nmol, 256), 256, 0, st>>>(c->nmol, s->pos, s->vel, s->forces, c->minv.as<double>(),
                                                        c->mass.as<double>(), dt, c->md_part.as<double>());
        LAUNCH_CHECK();
        s->step += 1;
    }
    double ke = 0.0;
    if (nsteps > 0) PGM_TRY(md_kinetic(c, &ke));
    PGM_TRY(read_energy(c, &e));
    // residuals of the solves of this call (device ring), the normaliser, SHAKE failures
    const long long nsolve = std::min<long long>(c->stats_slot - slot0, c->stats_cap);
    double res_max = 0.0;
    if (nsolve > 0) {
        std::vector<double> ring((size_t)c->stats_cap);
        HIP_TRY(hipMemcpy(ring.data(), c->stats.p, sizeof(double) * ring.size(), hipMemcpyDeviceToHost));
        for (long long j = c->stats_slot - nsolve; j < c->stats_slot; ++j) res_max = std::max(res_max, ring[(size_t)(j % c->stats_cap)]);
    }
    HIP_TRY(hipMemcpyAsync(c->h_pin_dbl + 1, &c->st.as<CGState>()->norm, sizeof(double), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(c->h_pin_int + 4, c->md_fail.p, sizeof(int), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(c->h_pin_dbl + 2, c->heat.p, sizeof(double), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    s->ind.norm = c->h_pin_dbl[1];
    if (c->h_pin_int[4] != 0) {
        g_err = "SHAKE did not converge for some molecule";
        return PGM_ERR_STATE;
    }
    if (rep) {
        rep->epot = e.total;
        rep->elec = e.elec;
        rep->vdw = e.vdw;
        rep->ekin = ke;
        rep->etot = e.total + ke;
        const double dof = 3.0 * n - 3.0 * c->nmol - 3.0 * c->nbead;  // rigid molecules, each bead's centre of mass
        rep->temperature = 2.0 * ke / (dof * 0.0083144626181532);
        rep->cg_iterations_total = it_total;
        rep->cg_iterations_max = it_max;
        rep->cg_iterations_last = it;
        rep->cg_residual_max = res_max;
        rep->row_rebuilds = c->rebuilds - rebuilds0;
        rep->heat = c->h_pin_dbl[2];
    }
    return status;
}

extern "C" int pgm_md_rescale(pgm_ctx *c, pgm_md_state *s, double factor) {
    if (!c || !s || !s->vel || !(factor > 0)) return PGM_ERR_ARG;
    k_scale3<<<nblk(3LL * c->n, 256), 256, 0, c->stream>>>(3 * c->n, s->vel, factor);
    LAUNCH_CHECK();
    return PGM_OK;
}

// ---------------------------------------------------------------------------- copy probe
__global__ void k_copy_probe(size_t n4, const float4 *__restrict__ a, float4 *__restrict__ b) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n4) b[i] = a[i];
}

extern "C" int pgm_copy_probe(void *dst, const void *src, size_t bytes, void *stream) {
    if (!dst || !src || bytes % 16 != 0) {
        g_err = "pgm_copy_probe: device buffers and a multiple of 16 bytes";
        return PGM_ERR_ARG;
    }
    const size_t n4 = bytes / 16;
    k_copy_probe<<<(unsigned)((n4 + 255) / 256), 256, 0, (hipStream_t)stream>>>(n4, (const float4 *)src, (float4 *)dst);
    LAUNCH_CHECK();
    return PGM_OK;
}

// Launch floor: `reps` back-to-back launches of a one-workgroup kernel that does nothing, between two
// HIP events on `stream`; *us = microseconds per launch (the time below which no kernel of a stream
// can go, whatever its bytes).
__global__ void k_empty() {}

extern "C" int pgm_launch_floor(void *stream, int reps, double *us) {
    if (reps <= 0 || !us) return PGM_ERR_ARG;
    hipStream_t st = (hipStream_t)stream;
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    for (int r = 0; r < 16; ++r) k_empty<<<1, 64, 0, st>>>();
    HIP_TRY(hipEventRecord(e0, st));
    for (int r = 0; r < reps; ++r) k_empty<<<1, 64, 0, st>>>();
    HIP_TRY(hipEventRecord(e1, st));
    HIP_TRY(hipEventSynchronize(e1));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    LAUNCH_CHECK();
    *us = 1e3 * (double)ms / reps;
    return PGM_OK;
}

// ---------------------------------------------------------------------------- kernel timing
// Average duration [ms] of each hot kernel over `reps` launches, HIP events on the context stream:
// out[0] block merge (tiled spread; grid clear otherwise), [1] spread (dipoles), [2] r2c FFT, [3] convolution, [4] c2r FFT, [5] row field
// (the CG matvec's pair part), [6] gather + combine, [7] one whole CG matvec (all of the above).
// With the fused transform: [0] 0 (the merge is part of [2]), [2] forward planes (merge, z, y),
// [3] x transform, Gc and inverse, [4] inverse planes (y, z).
template <typename real> static int time_impl(pgm_ctx *c, const double *pos, int reps, double *ms) {
    using R4 = real4<real>;
    if (c->ewald) {
        g_err = "pgm_time_kernels: PME kernels only (this context uses the Ewald sum)";
        return PGM_ERR_STATE;
    }
    PGM_TRY(prepare(c, pos));
    hipStream_t st = c->stream;
    const int n = c->n;
    k_perm_dipoles<<<nblk(n, 256), 256, 0, st>>>(n, pos, c->box, c->cstart.as<int>(), c->cj.as<int>(), c->cc.as<double>(),
                                                 c->pperm.as<double>());
    k_pack_dq<real><<<nblk(n, 256), 256, 0, st>>>(n, c->pperm.as<double>(), nullptr, nullptr, c->vp.as<R4>());
    LAUNCH_CHECK();
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    const R4 *dq = c->vp.as<R4>();
    const real beta = (real)c->s.ewald_beta;
    const real cself = (real)c_self(c);
    const int K3h = c->s.grid[2] / 2 + 1;
    auto timed = [&](int slot, auto &&launch) -> int {
        launch();  // warm-up
        HIP_TRY(hipEventRecord(e0, st));
        for (int r = 0; r < reps; ++r) launch();
        HIP_TRY(hipEventRecord(e1, st));
        HIP_TRY(hipEventSynchronize(e1));
        float t = 0.f;
        HIP_TRY(hipEventElapsedTime(&t, e0, e1));
        ms[slot] = (double)t / reps;
        return PGM_OK;
    };
    int rc = PGM_OK;
    if (c->tiled) {
        const PMETiles &t = c->tiles8;
        const size_t lds = spread_lds_bytes(c);
        ORDER_SWITCH(c->s.order, rc = timed(1, [&] {
                                     k_spread_tiles<real, PO, false><<<c->nbead * t.ntiles, WAVE, lds, st>>>(
                                         t, c->pg, c->tstart.as<int>(), c->tend.as<int>(), c->psidx.as<int>(), c->spl.as<real>(),
                                         c->loc.as<int>(), dq, c->blocks.as<real>(), nullptr, -1);
                                 }));
        if (rc < 0) return rc;
        const int *K = c->s.grid;
        rc = timed(0, [&] {
            k_merge_table<real><<<dim3(nblk(K[2], 128), K[1], K[0] * c->nbead), 128, 0, st>>>(
                K[0], K[1], K[2], (size_t)t.ntiles * t.bsize, c->merge_tab.as<int>(), c->blocks.as<real>(), c->grid.as<real>(),
                nullptr, -1);
        });
        if (rc < 0) return rc;
    } else {
        rc = timed(0, [&] { k_fill<real><<<nblk((long long)c->ngrid, 256), 256, 0, st>>>(c->ngrid, c->grid.as<real>(), real(0)); });
        if (rc < 0) return rc;
        ORDER_SWITCH(c->s.order, rc = timed(1, [&] {
                                     k_spread<real, PO, false><<<nblk((long long)n * PO, 256), 256, 0, st>>>(
                                         n, pos, dq, c->pg, c->grid.as<real>(), nullptr, -1);
                                 }));
        if (rc < 0) return rc;
    }
    if (c->pme_fused && c->fused_ok && c->tiled) {  // merge fused into the forward plane kernel
        using C = cplx<real>;
        const PFFTGeom &f = c->pfg;
        const size_t plane = pfft_plane_lds<real>(f);
        const int nth = f.K2 * f.K3 >= 8192 ? pfft_plane_threads<real>() : 256;
        const C *tw = c->pfft_tw.as<C>();
        ms[0] = 0.0;
        rc = timed(2, [&] {
            k_pme_fwd_plane<real><<<f.K1 * c->nbead, nth, plane, st>>>(f, (size_t)c->tiles8.ntiles * c->tiles8.bsize,
                                                                        c->merge_tab.as<int>(), c->blocks.as<real>(),
                                                                        c->cgrid.as<C>(), tw, nullptr, -1);
        });
        if (rc < 0) return rc;
        rc = timed(3, [&] {
            k_pme_x_conv<real><<<dim3(nblk((long long)f.K2 * f.K3h, PFFT_COLS), c->nbead), 256, pfft_xconv_lds<real>(f), st>>>(
                f, c->cgrid.as<C>(), c->Gc.as<real>(), tw + f.twK1, nullptr, -1);
        });
        if (rc < 0) return rc;
        rc = timed(4, [&] {
            k_pme_bwd_plane<real><<<f.K1 * c->nbead, nth, plane, st>>>(f, c->cgrid.as<C>(), c->grid.as<real>(), tw, nullptr, -1);
        });
        if (rc < 0) return rc;
    } else {
        void *in[1] = {c->grid.p}, *out[1] = {c->cgrid.p};
        rc = timed(2, [&] { rocfft_execute(c->fwd, in, out, c->info); });
        if (rc < 0) return rc;
        rc = timed(3, [&] {
            k_convolve<real, false><<<dim3(nblk((long long)c->ncplx, 256), c->nbead), 256, 0, st>>>(
                c->ncplx, K3h, c->s.grid[2] % 2 == 0, c->cgrid.as<real>(), c->Gc.as<real>(), nullptr, nullptr, -1);
        });
        if (rc < 0) return rc;
        void *in2[1] = {c->cgrid.p}, *out2[1] = {c->grid.p};
        rc = timed(4, [&] { rocfft_execute(c->bwd, in2, out2, c->info); });
        if (rc < 0) return rc;
    }
    rc = timed(5, [&] { (void)row_field_launch<real, false>(c, dq, nullptr, -1, st); });
    if (rc < 0) return rc;
    rc = timed(6, [&] { (void)gather_combine<real>(c, dq, MODE_AV, nullptr, c->vAp.as<R4>(), c->partA.as<double>(), nullptr, -1); });
    if (rc < 0) return rc;
    rc = timed(7, [&] { field_eval<real>(c, dq, false, MODE_AV, nullptr, c->vAp.as<R4>(), c->partA.as<double>(), -1); });
    if (rc < 0) return rc;
    LAUNCH_CHECK();
    HIP_TRY(hipEventDestroy(e0));
    HIP_TRY(hipEventDestroy(e1));
    return PGM_OK;
}

// ---------------------------------------------------------------------------- autotuner
// Time the CG matvec's pieces for every applicable shape on this context's own system and keep the
// fastest: the row-field layout (molecule blocks with 4, 8 or 16 lanes per molecule, atom rows) by
// the row field alone, then the spread-tile width (the largest valid width up to 8, 7, 6, ... P-1
// per axis; gather tiles are 2 x 2 x 2 spread tiles) by spread + merge + gather.
template <typename real> static int autotune_impl(pgm_ctx *c, const double *pos, int reps, double *out) {
    using R4 = real4<real>;
    hipStream_t st = c->stream;
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    auto time_ms = [&](auto &&launch, double &ms) -> int {
        launch();
        HIP_TRY(hipEventRecord(e0, st));
        for (int r = 0; r < reps; ++r) launch();
        HIP_TRY(hipEventRecord(e1, st));
        HIP_TRY(hipEventSynchronize(e1));
        float t = 0.f;
        HIP_TRY(hipEventElapsedTime(&t, e0, e1));
        ms = (double)t / reps;
        return PGM_OK;
    };
    auto dipoles = [&]() -> int {  // permanent dipoles as the spread / field input
        k_perm_dipoles<<<nblk(c->n, 256), 256, 0, st>>>(c->n, pos, c->box, c->cstart.as<int>(), c->cj.as<int>(),
                                                         c->cc.as<double>(), c->pperm.as<double>());
        k_pack_dq<real><<<nblk(c->n, 256), 256, 0, st>>>(c->n, c->pperm.as<double>(), nullptr, nullptr, c->vp.as<R4>());
        LAUNCH_CHECK();
        return PGM_OK;
    };
    // row-field layouts
    struct Lay {
        int layout, lanes;
    };
    std::vector<Lay> lays;
    if (c->mol3)
        for (int l : {4, 8, 16}) lays.push_back({1, l});
    lays.push_back({0, RL});
    double best_row = 1e300;
    Lay best_lay = lays.front();
    std::vector<double> row_ms;
    auto use_layout = [&](const Lay &L) -> int {
        c->layout = L.layout;
        if (L.layout == 1) {
            c->rlm = L.lanes;
            PGM_TRY(alloc_mol(c, std::max(c->mcap_full, 64)));
        }
        c->rows_valid = false;
        return PGM_OK;
    };
    for (const Lay &L : lays) {
        PGM_TRY(use_layout(L));
        PGM_TRY(prepare(c, pos));
        PGM_TRY(dipoles());
        double ms = 0;
        PGM_TRY(time_ms([&] { (void)row_field_launch<real, false>(c, c->vp.as<R4>(), nullptr, -1, st); }, ms));
        row_ms.push_back(ms);
        if (ms < best_row) best_row = ms, best_lay = L;
    }
    PGM_TRY(use_layout(best_lay));
    // spread-tile widths
    double best_pme = 1e300;
    int best_cap = c->tile_cap;
    if (c->tiled && !c->ewald) {
        std::vector<int> tried;  // T8 triples already timed (a lower cap can give the same tiles)
        for (int cap = 8; cap >= c->s.order - 1; --cap) {
            c->tile_cap = cap;
            PGM_TRY(setup_tiles(c));
            if (!c->tiled) continue;
            const int key = (c->tiles8.T[0] * 16 + c->tiles8.T[1]) * 16 + c->tiles8.T[2];
            if (std::find(tried.begin(), tried.end(), key) != tried.end()) continue;
            tried.push_back(key);
            PGM_TRY(prepare(c, pos));
            PGM_TRY(dipoles());
            double ms = 0;
            PGM_TRY(time_ms([&] {
                (void)spread_fft_only<real>(c, c->vp.as<R4>());
                (void)gather_combine<real>(c, c->vp.as<R4>(), MODE_AV, nullptr, c->vAp.as<R4>(), c->partA.as<double>(), nullptr, -1);
            }, ms));
            if (ms < best_pme) best_pme = ms, best_cap = cap;
        }
        c->tile_cap = best_cap;
        PGM_TRY(setup_tiles(c));
    }
    // reciprocal transform: the fused kernels against the library plans, with the chosen tiles
    double other_pme = 0.0;
    if (c->tiled && c->fused_ok) {
        PGM_TRY(prepare(c, pos));
        PGM_TRY(dipoles());
        double ms[2] = {0.0, 0.0};
        for (int fz = 0; fz < 2; ++fz) {
            c->pme_fused = fz == 1;
            PGM_TRY(time_ms([&] {
                (void)spread_fft_only<real>(c, c->vp.as<R4>());
                (void)gather_combine<real>(c, c->vp.as<R4>(), MODE_AV, nullptr, c->vAp.as<R4>(), c->partA.as<double>(), nullptr, -1);
            }, ms[fz]));
        }
        c->pme_fused = ms[1] <= ms[0];
        best_pme = ms[c->pme_fused ? 1 : 0];
        other_pme = ms[c->pme_fused ? 0 : 1];
        drop_graphs(c);
    }
    // fork point of the row field in a whole CG field evaluation (row field beside the PME chain),
    // jointly with the row layout among those within 10 % of the fastest row field alone: the
    // layouts load the GPU differently beside the PME chain
    double best_fork = 1e300;
    int fork_best = 0;
    if (c->tiled && !c->ewald) {
        const bool fz = c->pme_fused && c->fused_ok;
        Lay fork_lay = best_lay;
        for (size_t li = 0; li < lays.size(); ++li) {
            if (row_ms[li] > 1.1 * best_row) continue;
            PGM_TRY(use_layout(lays[li]));
            PGM_TRY(prepare(c, pos));
            PGM_TRY(dipoles());
            for (int fs = 0; fs < 5; ++fs) {
                if (!fz && (fs == 2 || fs == 3)) continue;
                c->fork_stage = fs;
                double ms = 0;
                PGM_TRY(time_ms([&] {
                    (void)field_eval<real>(c, c->vp.as<R4>(), false, MODE_AV, nullptr, c->vAp.as<R4>(), c->partA.as<double>(), -1);
                }, ms));
                if (ms < best_fork) best_fork = ms, fork_best = fs, fork_lay = lays[li];
            }
            drop_graphs(c);
        }
        for (size_t li = 0; li < lays.size(); ++li)
            if (lays[li].layout == fork_lay.layout && lays[li].lanes == fork_lay.lanes) best_row = row_ms[li];
        PGM_TRY(use_layout(fork_lay));
        c->fork_stage = fork_best;
    }
    c->rows_valid = false;
    PGM_TRY(prepare(c, pos));
    HIP_TRY(hipStreamSynchronize(st));
    HIP_TRY(hipEventDestroy(e0));
    HIP_TRY(hipEventDestroy(e1));
    out[0] = c->layout;
    out[1] = c->layout == 1 ? c->rlm : RL;
    for (int a = 0; a < 3; ++a) out[2 + a] = c->tiled ? c->tiles8.T[a] : 0;
    out[5] = best_row;
    out[6] = best_pme < 1e300 ? best_pme : 0.0;
    out[7] = (double)lays.size();
    out[8] = c->pme_fused && c->fused_ok ? 1.0 : 0.0;
    out[9] = other_pme;
    out[10] = c->fork_stage;
    out[11] = best_fork < 1e300 ? best_fork : 0.0;
    return PGM_OK;
}

extern "C" int pgm_set_pme_transform(pgm_ctx *c, int fused) {
    if (!c) return PGM_ERR_ARG;
    if (fused && !(c->fused_ok && c->tiled)) {
        g_err = "pgm_set_pme_transform: the fused transform needs power-of-two grids of 8..256 points per axis, tiled "
                "spreading and a plane of K2 (K3/2 + 1) complex values within 160 KB";
        return PGM_ERR_ARG;
    }
    drop_graphs(c);
    c->pme_fused = fused != 0;
    if (!c->pme_fused && (c->fork_stage == 2 || c->fork_stage == 3)) c->fork_stage = c->fork_stage == 2 ? 1 : 0;
    return PGM_OK;
}

extern "C" int pgm_set_row_fork(pgm_ctx *c, int stage) {
    if (!c || stage < 0 || stage > 4) return PGM_ERR_ARG;
    if ((stage == 2 || stage == 3) && !(c->pme_fused && c->fused_ok && c->tiled)) {
        g_err = "pgm_set_row_fork: stages 2 and 3 need the fused transform";
        return PGM_ERR_ARG;
    }
    drop_graphs(c);
    c->fork_stage = stage;
    return PGM_OK;
}

extern "C" int pgm_pme_transform(const pgm_ctx *c, int *fused, int *available) {
    if (!c) return PGM_ERR_ARG;
    if (fused) *fused = c->pme_fused && c->fused_ok && c->tiled ? 1 : 0;
    if (available) *available = c->fused_ok && c->tiled ? 1 : 0;
    return PGM_OK;
}

extern "C" int pgm_autotune(pgm_ctx *c, const double *pos, int reps, double *report) {
    if (!c || !pos || !report || reps <= 0) return PGM_ERR_ARG;
    for (int k = 0; k < 12; ++k) report[k] = 0.0;
    return c->mixed ? autotune_impl<float>(c, pos, reps, report) : autotune_impl<double>(c, pos, reps, report);
}

extern "C" int pgm_time_kernels(pgm_ctx *c, const double *pos, int reps, double *ms) {
    if (!c || !pos || !ms || reps <= 0) return PGM_ERR_ARG;
    return c->mixed ? time_impl<float>(c, pos, reps, ms) : time_impl<double>(c, pos, reps, ms);
}
