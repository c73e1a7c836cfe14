/*
 * pgm_hip.h -- C ABI of the MI355X (gfx950) pGM electrostatics library (libpgm_hip.so).
 *
 * One context holds the static topology of one system (per-atom charges, Gaussian radii,
 * polarizabilities, LJ parameters, molecules, covalent dipoles), the settings, the PME grid with
 * its rocFFT plans, the neighbour rows and every scratch buffer.  Positions, forces, dipoles and
 * the induction state are caller-owned DEVICE buffers (float64, row-major (N, 3)).  Every call is
 * issued on the context's stream and returns without synchronising unless stated.  No C++
 * exception crosses this boundary; every function returns a status code:
 *
 *   PGM_OK 0, PGM_ROW_OVERFLOW 1 (neighbour-row capacity exceeded: results invalid, call
 *   pgm_build_rows again -- it grows the capacity), PGM_NOT_CONVERGED 3 (CG stopped at max_iter;
 *   results valid, residual reported), negative values are errors (pgm_last_error() describes them).
 *
 * Units: nm, e, e nm, nm^3, kJ/mol, kJ/mol/nm.  Boxes are reduced lower-triangular with the
 * lattice vectors as rows (H[3*a + b] = component b of vector a).
 *
 * Each entry point names the reference interface (pgm-jax) it replaces.
 */
#ifndef PGM_HIP_H
#define PGM_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PGM_OK 0
#define PGM_ROW_OVERFLOW 1
#define PGM_NOT_CONVERGED 3
#define PGM_ERR_ARG -1
#define PGM_ERR_HIP -2
#define PGM_ERR_FFT -3
#define PGM_ERR_STATE -4

typedef struct pgm_ctx pgm_ctx;

/* Per-atom parameters (host arrays, float64 / int32), the expand() output of the reference's
 * System (pgm_jax/system.py:561-577); covalent dipole m puts c[m] e nm on atom cov_i[m] along
 * unit(r_{cov_j[m]} - r_{cov_i[m]}) (pgm_jax/md/forcefield.py:1032-1042). */
typedef struct {
    int n_atoms;
    const double *q, *radius, *alpha, *lj_rmin_half, *lj_sqrt_eps;  /* alpha >= 0; 0: not polarizable (its
                                                                      * induced dipole stays 0, forcefield.py:767-780) */
    const int *mol;
    int n_cov;
    const int *cov_i, *cov_j;
    const double *cov_c;
} pgm_topology;

/* The MDSettings fields of the hot path (pgm_jax/md/forcefield.py:142-454). */
typedef struct {
    double cutoff;      /* pair cutoff rc [nm] (Cutoffs.cutoff) */
    double skin;        /* neighbour-row skin [nm] (NeighborList.skin) */
    double ewald_beta;  /* [1/nm] (PMESettings.ewald_beta) */
    int grid[3];        /* PME grid K1, K2, K3 (PMESettings.grid) */
    int order;          /* B-spline order (PMESettings.order), 4..8 */
    int lj_lrc;         /* LJ long-range correction (Terms.lj_lrc) */
    double dipole_tol;  /* max|alpha r| / mean|alpha b| (Induction.tol) */
    int max_iter;       /* Induction.max_iter */
    int predictor;      /* 0 none, 1 mu3, 2 mu4, 3 previous step (Induction.predictor), 4 learned (pgm_guess_*) */
    int fused;          /* fused initial residual (Induction.fused) */
    int norm_refresh;   /* Induction.norm_refresh */
    double peek;        /* Induction.peek */
    int precision;      /* 0 double, 1 mixed (MDSettings.precision) */
    int recip;          /* reciprocal space: 0 smooth PME, 1 exact Ewald sum (PeriodicPGM, ewald.py:186-470) */
    double k_tol;       /* Ewald: k-vectors with exp(-k^2/(4 beta^2)) >= k_tol, fixed at the creation box */
    int nbead;          /* beads (0 or 1: one system): copies of the topology advanced together (PGMBeads) */
    int precond;        /* CG preconditioner: 0 Jacobi z = alpha r (the reference's), 1 the inverse 9x9 block of
                           each 3-site molecule (intramolecular coupling; same convergence test) */
    double elec_cutoff; /* real-space electrostatics cutoff [nm] (Cutoffs.elec_cutoff, pmemd's es_cutoff); 0: the
                           same as cutoff.  Electrostatic pair terms (CG row field, energies, forces, strain) stop
                           at elec_cutoff, LJ at cutoff (pgm_jax/md/forcefield.py:44-58, :1197-1238); the rows
                           span the larger of the two */
} pgm_settings;

/* Induction state (pgm_jax/md/forcefield.py:679-713 InductionState), caller-owned:
 * mu (N,3) f64 device, hist (4,N,3) f64 device (newest first), count and norm on the host. */
typedef struct {
    double *mu;
    double *hist;
    int count;
    double norm;
} pgm_induction;

/* Energy parts [kJ/mol] (Result.energy of forcefield.py:716-746). */
typedef struct {
    double elec, vdw, total;
} pgm_energy;

const char *pgm_last_error(void);
int pgm_version(void);

/* PGMForceField.__init__ (pgm_jax/md/forcefield.py:868-1000): build the context. `stream` is a
 * hipStream_t (NULL: the null stream).  H is the box used for the cell grid and influence function. */
int pgm_ctx_create(pgm_ctx **out, const pgm_topology *top, const pgm_settings *s, const double *H, void *stream);
int pgm_ctx_destroy(pgm_ctx *ctx);
int pgm_set_stream(pgm_ctx *ctx, void *stream);

/* PME.influence + box change (pgm_jax/md/pme.py:243-267): new box, influence function, cells. */
int pgm_set_box(pgm_ctx *ctx, const double *H);

/* Neighbour rows (replaces the neighbour list + PGMForceField._rows/_compact_parts,
 * pgm_jax/md/neighbors.py:525-558, forcefield.py:1140-1186): every pair within cutoff + skin of
 * the device positions pos (N,3).  Returns PGM_ROW_OVERFLOW after growing the capacity. */
int pgm_build_rows(pgm_ctx *ctx, const double *pos);

/* Number of stored row entries (pairs counted twice) within cutoff at the last compute. */
long long pgm_row_entries(pgm_ctx *ctx);

/* Number of molecule-block records (molecule pairs with a pair inside the cutoff, each direction) at
 * the last evaluation; 0 with atom rows, -1 before the first. */
long long pgm_row_blocks(pgm_ctx *ctx);

/* New parameters for the next evaluations: the parameter pytree of PGMForceField.compute(...,
 * params) (forcefield.py:2489-2560) expanded per atom (System.expand, pgm_jax/system.py:561-577).
 * Host f64 arrays of one bead (n_atoms; cov_c: n_cov in the topology's order), NULL: unchanged.
 * Radii must be > 0, polarizabilities >= 0, everything finite (PGM_ERR_ARG otherwise, nothing
 * changed).  Synchronises. */
int pgm_set_params(pgm_ctx *ctx, const double *q, const double *radius, const double *alpha, const double *lj_rmin_half,
                   const double *lj_sqrt_eps, const double *cov_c);

/* A uniform static external field for the next evaluations: the constant-field kind of the
 * efield=(E, offset) argument of PGMForceField.compute / energy / energy_fixed_mu
 * (forcefield.py:2048-2097, :2489-2560; md/efield.py:372-404).  E (3, V/nm, host) joins the
 * permanent field of every atom in the induction equations (as F0 = E FARADAY / KE), every charge
 * feels KE q F0 and every covalent dipole its torque, and the energy gains
 * field = -E . M FARADAY with M = sum q r + sum (p + mu) + offset [e nm] (offset: 3 host doubles,
 * NULL: zero); pgm_energy.total includes it.  E NULL: no field (the field-free kernels' results,
 * bit for bit).  Strain derivatives and the VJPs refuse a context with a field (PGM_ERR_STATE). */
int pgm_set_efield(pgm_ctx *ctx, const double *E_vnm, const double *offset);

/* The field's part of the last pgm_compute / pgm_trial_energy / pgm_energy_forces: e_field
 * (nbead, kJ/mol; Result.energy["field"]) and the total dipole M (nbead x 3, e nm; Result.dipole),
 * host arrays.  PGM_ERR_ARG without a field.  Synchronises. */
int pgm_field_result(pgm_ctx *ctx, double *e_field, double *dipole);

/* PGMForceField.perm_dipoles (forcefield.py:1032-1042): p (N,3) f64 device. */
int pgm_perm_dipoles(pgm_ctx *ctx, const double *pos, double *p);

/* PGMForceField._field (forcefield.py:1428-1433): the total field -dU/dd (N,3) f64 device of
 * charges (q_scale * topology charges, q_scale 0 or 1) and dipoles d (N,3) f64 device. */
int pgm_field(pgm_ctx *ctx, const double *pos, int with_charges, const double *d, double *field);

/* PME.energy / PME.grad_dipoles (pgm_jax/md/pme.py:339-411) of the topology charges times
 * q_scale and dipoles d: E_rec [e^2/nm] (host) and dU_rec/dd (N,3) device (may be NULL). */
int pgm_pme_energy(pgm_ctx *ctx, const double *pos, int with_charges, const double *d, double *e_rec,
                   double *grad_d);

/* PME.spread (pgm_jax/md/pme.py:306-337): the grid Q (K1, K2, K3) f64 device of the topology
 * charges times q_scale (with_charges 0 or 1) and the dipoles d (N,3) f64 device at pos, in the
 * working precision (FP32 mixed) converted to f64.  One system per context, PME reciprocal space. */
int pgm_pme_spread(pgm_ctx *ctx, const double *pos, int with_charges, const double *d, double *Q);

/* PME.influence (pgm_jax/md/pme.py:243-267) of the context's box: G (K1, K2, K3/2 + 1) f64 host,
 * exp(-pi^2 |m|^2 / beta^2) / (2 pi V |m|^2) times the spline moduli and the half-space weights
 * (1 on the planes m3 = 0 and m3 = K3/2, 2 elsewhere), 0 at m = 0 [1/nm]. */
int pgm_pme_influence(pgm_ctx *ctx, double *G);

/* PGMForceField.compute (forcefield.py:2489-2560): rows compaction, permanent dipoles, PME setup,
 * predictor + CG induced dipoles (updates ind), energies and forces (N,3) f64 device.
 * iterations and residual as Result.iterations / Result.residual. */
int pgm_compute(pgm_ctx *ctx, const double *pos, pgm_induction *ind, pgm_energy *energy, double *forces,
                int *iterations, double *residual);

/* PGMForceField.energy (forcefield.py:2562-2626), the Monte Carlo barostat's trial energy: the
 * induced dipoles solved by CG from ind->mu (no predictor; the right-hand side and the
 * convergence norm evaluated afresh), written back to ind->mu; ind->hist, count and norm are
 * left as they are, as is the learned guess's record.  Energies only (no forces). */
int pgm_trial_energy(pgm_ctx *ctx, const double *pos, pgm_induction *ind, pgm_energy *energy, int *iterations,
                     double *residual);

/* Beads (settings.nbead > 1; PGMBeads.compute over the force beads, pgm_jax/md/pimd.py:1104-1172,
 * with every bead's force field in the same launches instead of vmap / lax.map chunks,
 * pimd.py:1068-1102): positions, forces, mu and hist are (nbead, N, 3) with N = top->n_atoms; every
 * bead has its own rows, PME grid and CG (bead-wise FP64 reductions).  pgm_compute then returns the
 * energies summed over the beads and the largest iteration count and residual; pgm_bead_results
 * gives each bead's energies, CG iterations and final residual of the last evaluation (host arrays
 * of nbead, each may be NULL).  Strain, the learned guess and the Ewald sum take one system. */
int pgm_bead_count(const pgm_ctx *ctx);
int pgm_bead_results(pgm_ctx *ctx, pgm_energy *energies, int *iterations, double *residuals);

/* CG iterations as hipGraphs (no counterpart in the reference): with `on`, iteration k of every
 * solve (the row field on the side stream, the PME chain, the gather, the two update kernels) is
 * captured once into a graph and replayed with hipGraphLaunch; the host's pacing of the iterations
 * and the results are unchanged.  Graphs are dropped when the box, row capacities, tiles or layout
 * change, and not used while live kernel timing is on.  A context on the null stream (which cannot be
 * captured) replays them on a private stream ordered after and before its work by events. */
int pgm_set_graphs(pgm_ctx *ctx, int on);

/* Layout of the CG row field (the data of PGMForceField._row_field, forcefield.py:1395-1402): 0 atom
 * rows (partner index, G1, G2 per pair in sliced ELL), 1 molecule blocks (contexts of consecutive
 * 3-atom molecules: per molecule pair the 9 atom pairs' G1, G2, `lanes` = 4, 8 or 16 lanes per
 * molecule; the default where it applies).  Same sums in another order; rows are rebuilt at the
 * next evaluation.  pgm_row_layout reports the layout and its lanes. */
int pgm_set_row_layout(pgm_ctx *ctx, int layout, int lanes);
int pgm_row_layout(const pgm_ctx *ctx, int *layout, int *lanes);

/* Reciprocal transform of the CG field, energy and force evaluations (the FFT / convolution / inverse FFT of
 * pme.py:373-376): 0 the library's batched r2c / c2r plans with a separate convolution kernel, 1
 * the fused kernels (the spread-block merge, the real transform along z and the transform along y
 * per grid plane; along x with the product by the influence function and its inverse; the inverse
 * planes), the default where available: every grid axis a power of two in 8..256, tiled spreading,
 * K2 (K3/2 + 1) complex values within 160 KB; in an energy evaluation the x-transform kernel also
 * sums the reciprocal energy.  Strain evaluations keep the library plans.  pgm_pme_transform reports the transform in use and whether the fused one is available. */
int pgm_set_pme_transform(pgm_ctx *ctx, int fused);
int pgm_pme_transform(const pgm_ctx *ctx, int *fused, int *available);

/* Where a CG field evaluation starts the row field on its side stream (no counterpart in the
 * reference, whose field is one XLA program): 0 before the PME spread, 1 after it, 2 after the
 * forward planes of the fused transform, 3 before the spread with the forward planes waiting for
 * the row field; 4: no side stream (the row field after the spread on the context's stream, no
 * fork or join events).  Only the overlap changes, not the results; the autotuner picks the
 * fastest. */
int pgm_set_row_fork(pgm_ctx *ctx, int stage);

/* Autotuner (no counterpart in the reference): times, on this context's system at positions pos
 * (device), every applicable row-field layout (molecule blocks with 4, 8, 16 lanes; atom rows),
 * every valid PME spread-tile width (8 down to order - 1 points per axis; gather tiles twice that),
 * both reciprocal transforms and the point where a CG field evaluation starts the row field on its
 * side stream (before the spread, after it, after the forward planes of the fused transform) or
 * runs it on the context's stream (jointly with the row layouts within 10 % of the fastest row
 * field, by the time of a whole field evaluation),
 * `reps` launches each, and keeps the fastest.  report (host, 12 doubles): layout, lanes, spread
 * tile width per axis (3), row-field ms, spread + transform + gather ms, layouts tried, fused
 * transform chosen (1) or not (0), the other transform's ms (0 if not timed), fork point (0..4,
 * pgm_set_row_fork), whole field-evaluation ms.  Synchronises. */
int pgm_autotune(pgm_ctx *ctx, const double *pos, int reps, double *report);

/* PGMForceField._energy_forces / energy_fixed_mu (forcefield.py:2249-2292, :2362-2430): energy
 * and forces at fixed induced dipoles mu (N,3) f64 device. */
int pgm_energy_forces(pgm_ctx *ctx, const double *pos, const double *mu, pgm_energy *energy, double *forces);

/* PGMForceField.strain_derivative (forcefield.py:2628-2695, with the tail impulse -E_lrc I when
 * lj_lrc) and PeriodicModel.strain_derivative (periodic.py:236-251, which adds it back through
 * virial_derivative): dE/d eps (3,3) row-major [kJ/mol] on the host at fixed induced dipoles mu.
 * molecular != 0: molecules translated with their centres of mass (masses (N,) f64 device;
 * molecules whole); 0: atomic scaling (masses may be NULL; the covalent dipole frames strain).
 * Also returns the energy and forces of the same pass (forces may be NULL). */
int pgm_strain_derivative(pgm_ctx *ctx, const double *pos, const double *mu, const double *masses, int molecular,
                          pgm_energy *energy, double *forces, double *strain);

/* The backward pass of the differentiable solve (the custom VJP of PGMForceField._solve,
 * forcefield.py:1853-1915): the adjoint CG A lambda = mu_bar (the forward operator and
 * preconditioner, x0 = 0, no peek, until max |alpha r| / mean |alpha mu_bar| <= tol, at most
 * max_iter iterations), then lambda pulled back through the residual b - A mu at the dipoles mu:
 *   q_bar = -phi(lambda)          (N)    the potential of the dipoles lambda at every atom
 *   radius_bar                    (N)    through a_ij = (2 (R_i^2 + R_j^2))^(-1/2) of the pair kernels
 *   alpha_bar = mu . lambda / alpha^2 (N)
 *   p_bar = field(0, lambda)      (N,3)  the cotangent of the permanent dipoles
 *   cov_bar                       (n_cov) p_bar projected on each covalent dipole's bond direction
 * in the units of the solve (field and potential without the Coulomb constant, as `pgm_field`).
 * pos, mu (the solved dipoles), mu_bar and every output: f64 device; any output may be NULL.  One
 * system per context, PME reciprocal space.  Returns PGM_NOT_CONVERGED (results valid) when the
 * adjoint CG stops at max_iter.  Synchronises. */
int pgm_dipole_vjp(pgm_ctx *ctx, const double *pos, const double *mu, const double *mu_bar, double tol, double *lambda,
                   double *q_bar, double *radius_bar, double *alpha_bar, double *p_bar, double *cov_bar, int *iterations,
                   double *residual);

/* Parameter cotangents of the differentiable force field's outputs (PGMForceField.compute with
 * differentiable=True and a parameter pytree, forcefield.py:1853-1915, 2489-2560): for a loss with
 * dL/dE = e_bar (the total energy, host scalar), dL/dF = f_bar (N,3) and dL/dmu = mu_bar (N,3) at
 * the induced dipoles mu solved at pos, the per-atom cotangents
 *   theta_bar = e_bar dE/dtheta + f_bar . dF/dtheta + mu_bar . dmu/dtheta
 * of the charges, Gaussian radii, polarizabilities, LJ rmin/2 and sqrt(eps) (N each) and of the
 * covalent-dipole strengths (n_cov).  The forces' part takes the derivatives of dE/dtheta and of
 * the field along f_bar (forward mode, the pair set and cutoff masks held fixed, as autodiff of the
 * reference's masked sums); the dipoles' part is the adjoint solve of mu_bar + KE D_f_bar field to
 * `tol` (pgm_dipole_vjp).  f_bar and mu_bar may be NULL (zero); any output may be NULL.  Summing
 * the cotangents over the atoms of each tying key gives the gradient of the reference's parameter
 * table.  Device f64 arrays; one system per context, PME reciprocal space.  Synchronises. */
int pgm_compute_vjp(pgm_ctx *ctx, const double *pos, const double *mu, double e_bar, const double *f_bar,
                    const double *mu_bar, double tol, double *q_bar, double *radius_bar, double *alpha_bar,
                    double *rmin_half_bar, double *sqrt_eps_bar, double *cov_bar, int *iterations, double *residual);

/* solver.solve_linear_induction (pgm_jax/solver.py:29-62): mu (n,3) f64 device solving
 * (T + diag(1/alpha)) mu = F for T (n,n,3,3) f64 device (zero diagonal blocks), alpha (n), F (n,3),
 * by FP64 Jacobi-preconditioned CG from alpha F until |r| <= tol |F| (the reference solves it by
 * dense LU).  No context; runs on `stream` (hipStream_t) and synchronises it. */
int pgm_solve_linear_induction(int n, const double *T, const double *alpha, const double *F, double *mu, double tol,
                               int max_iter, void *stream, int *iterations, double *residual);

/* No counterpart in the reference.  pgm_solve_linear_induction without any allocation, for callers
 * that cannot allocate device or pinned memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc): the same arguments, kernels and results, with the CG work
 * vectors in a fixed device array that the library holds once per GPU for the life of the process
 * (about 1.8 MB).  n may not exceed PGM_DENSE_NOALLOC_MAX_ATOMS (PGM_ERR_ARG beyond it).  Calls
 * that overlap on one GPU share that array and race: serialise them. */
#define PGM_DENSE_NOALLOC_MAX_ATOMS 16384
int pgm_solve_linear_induction_noalloc(int n, const double *T, const double *alpha, const double *F, double *mu,
                                       double tol, int max_iter, void *stream, int *iterations, double *residual);

/* MD state (caller-owned device buffers, float64 (N,3)) and the report of pgm_md_run. */
typedef struct {
    double *pos, *vel, *forces;
    pgm_induction ind;
    int have_forces; /* forces hold the forces at pos (set by pgm_md_run after the first evaluation) */
    long long step;
} pgm_md_state;

typedef struct {
    double epot, elec, vdw, ekin, etot, temperature;
    long long cg_iterations_total;
    int cg_iterations_max, cg_iterations_last;
    double cg_residual_max;
    int row_rebuilds;
    double heat; /* Langevin: the O steps' change of kinetic energy [kJ/mol] over the call */
} pgm_md_report;

/* Rigid-molecule MD set-up (replaces the rigid engine's constraints, pgm_jax/md/constraints.py:254-299
 * with MDTopology.rigid): consecutive 3-site molecules, masses (N) host [amu], constraint lengths
 * taken from the device positions pos (N,3). */
int pgm_md_setup(pgm_ctx *ctx, const double *masses, const double *pos);

/* nsteps of NVE velocity Verlet with SHAKE/RATTLE in FP64 (Integrator._step / FlexibleIntegrator._step,
 * pgm_jax/md/integrate.py:743-761, md/flexible.py:776-799), one force evaluation (pgm_compute) per
 * step, rows rebuilt when an atom moved by more than skin/2 (MDEngine's list update). */
int pgm_md_run(pgm_ctx *ctx, pgm_md_state *state, int nsteps, double dt, pgm_md_report *report);

/* Langevin thermostat (Langevin(friction).apply, pgm_jax/md/thermostats.py:195-242, in the g-BAOAB
 * step of md/flexible.py:776-799): friction [1/ps] (0: NVE), temperature [K], seed of the
 * counter-based noise (Philox4x32-10 keyed by seed, counter (molecule, step, k)).  Applies to the
 * following pgm_md_run calls. */
int pgm_md_set_thermostat(pgm_ctx *ctx, double friction, double temperature, unsigned long long seed);

/* Velocity rescaling v *= factor (constraints are preserved): the velocity-rescaling thermostat
 * used to equilibrate (pgm_jax/md/thermostats.py's role for this engine). */
int pgm_md_rescale(pgm_ctx *ctx, pgm_md_state *state, double factor);

/* Measurement (no reference counterpart): average duration [ms] of the hot kernels over reps
 * launches, timed with HIP events on the context stream, for the dipoles p(pos):
 * ms[0] block merge of the tiled spread (grid clear without tiles), [1] spread, [2] r2c FFT, [3] convolution, [4] c2r FFT, [5] row field,
 * [6] gather + combine, [7] one whole dipole-field matvec.  ms must hold 8 doubles. */
int pgm_time_kernels(pgm_ctx *ctx, const double *pos, int reps, double *ms);

/* Learned CG initial guess (no reference counterpart; MDSettings.predictor "learned" = 4): a linear
 * predictor over the last 8 converged dipoles whose coefficients are fitted by least squares to the
 * dipoles the solves produce (refitted every harvest_every solves), plus a per-atom MLP correction
 * (16 features: the last four dipoles in a local frame of the molecule and an oxygen flag;
 * hidden_layers x 128 SiLU; 3 outputs) trained on the linear predictor's errors harvested into a
 * ring of `capacity` rows.  BF16 inference on the matrix cores, FP32 training (Adam).  The selector,
 * at every harvest: the linear predictor is used while its error is below 0.8 x that of mu4, the
 * network on top of it while the sum's error is below 0.8 x the linear predictor's; else mu4.  The
 * CG always runs to the same tolerance.  Needs consecutive 3-site molecules.
 *   pgm_guess_config: set up (weights LeCun normal from seed); dipole_scale [e nm] scales inputs.
 *   pgm_guess_params: get (set = 0) or set (set = 1) the flat FP32 network parameters (per layer W
 *     (in x out) then b), host array of pgm_guess_nparams() floats.
 *   pgm_guess_forward: y (n x 3) of feature rows x (n x 16), device FP32; bf16 != 0: the matrix-core
 *     inference kernel, else the FP32 training forward.
 *   pgm_guess_grad: gradients (host, flat) and loss of the mean squared error of rows x vs t (n x 3).
 *   pgm_guess_train: `steps` Adam steps on minibatches of `batch` harvested rows; mean loss.
 *   pgm_guess_status: out[16] = rows harvested, sigma, squared errors per component at the last
 *     harvest of mu4, the linear predictor and linear + network, mode (0 mu4, 1 linear, 2 linear
 *     + network), trained, number of fits, the eight linear coefficients (newest first). */
int pgm_guess_config(pgm_ctx *ctx, int hidden_layers, double dipole_scale, int harvest_every, int capacity,
                     unsigned long long seed);
long long pgm_guess_nparams(pgm_ctx *ctx);
int pgm_guess_params(pgm_ctx *ctx, float *host, long long count, int set);
int pgm_guess_forward(pgm_ctx *ctx, const float *x, int n, float *y, int bf16);
int pgm_guess_grad(pgm_ctx *ctx, const float *x, const float *t, int n, float *grads, double *loss);
int pgm_guess_train(pgm_ctx *ctx, int steps, int batch, double lr, double *loss);
int pgm_guess_status(pgm_ctx *ctx, double *out);

/* Measurement (no reference counterpart): live timing of the CG matvec kernels.  With on != 0,
 * every following dipole-field evaluation records HIP events around its row field (on the side
 * stream), its reciprocal spread (tiled spread + merge, or the Ewald structure factors) and its
 * gather; pgm_set_timing also clears the sums.  pgm_kernel_timing waits for the recorded events
 * and returns the average duration [ms] and the number of launches of [0] row field,
 * [1] spread, [2] gather. */
int pgm_set_timing(pgm_ctx *ctx, int on);
int pgm_kernel_timing(pgm_ctx *ctx, double *avg_ms, long long *launches);

/* Measurement hook (no counterpart in the reference): a plain copy of `bytes` (a multiple of 16)
 * between device buffers on `stream`, the known-size kernel that calibrates HBM traffic counters. */
int pgm_copy_probe(void *dst, const void *src, size_t bytes, void *stream);

/* Measurement hook, no counterpart in the reference: microseconds per launch of `reps` back-to-back
 * launches of an empty one-workgroup kernel on `stream` (hipStream_t), between HIP events. */
int pgm_launch_floor(void *stream, int reps, double *us);

#ifdef __cplusplus
}
#endif
#endif /* PGM_HIP_H */
