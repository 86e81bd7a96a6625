/*
 * lsqed.h -- C-ABI of the MI355X matrix-free Sz-sector Hamiltonian and
 * Krylov eigensolver (liblsqed.so).
 *
 * The boundary replaces the PETSc matrix that LSQuantumED builds and hands to
 * SLEPc: `Hamiltonian<L>::hamilt` (include/hamiltonian.hpp:187), created as
 * MATMPIAIJ in `MatInit` (include/hamiltonian.tmpl_func.hpp:8-56) and passed
 * to `Solver::SolverInit` (src/solver.cpp:7-24).  A PETSc MATSHELL whose
 * MATOP_MULT calls lsqed_hv() and MATOP_GET_DIAGONAL calls lsqed_diag() is
 * the drop-in on the reference side.
 *
 * Conventions
 *  - Every call returns an int status: LSQED_OK (0) or a negative LSQED_ERR_*
 *    code (mirrors PetscErrorCode); lsqed_last_error() gives the message of
 *    the last failure on the calling thread.
 *  - Device buffers are caller-owned, FP64, allocated on the current HIP
 *    device; vectors are dense in basis-rank (colex) order.
 *  - `stream` may be NULL (the legacy default stream); kernels are enqueued
 *    asynchronously on it.  Calls that return host scalars synchronise it.
 *  - States are bit strings (bit i = spin i up), ranks are colex combinadic
 *    indices r(s) = sum_k C(p_k, k) as in HamiltHelper::search_index
 *    (src/hamiltonian.cpp:9-27).  N <= 62.
 */
#ifndef LSQED_H
#define LSQED_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ihipStream_t* lsqed_stream_t; /* == hipStream_t */

enum {
  LSQED_OK = 0,
  LSQED_ERR_ARG = -1,     /* invalid argument (std::invalid_argument in the reference) */
  LSQED_ERR_HIP = -2,     /* a HIP runtime call failed */
  LSQED_ERR_OOM = -3,     /* device memory plan does not fit */
  LSQED_ERR_NOCONV = -4,  /* reserved */
  LSQED_ERR_STATE = -5    /* handle used in the wrong state */
};

/* Lattice kinds of src/lattice.cpp. */
enum { LSQED_CHAIN1D = 0, LSQED_SQUARE2D = 1, LSQED_HONEYCOMB2D = 2 };

/* Build flags: which of the reference's build calls have been made.
 *  DIAG     -> Hamiltonian::build_diag      (include/hamiltonian.tmpl_func.hpp:129-225)
 *  OFFDIAG  -> Hamiltonian::build_off_diag  (tmpl:229-344)
 *  DISORDER -> Hamiltonian::build_disorder  (tmpl:97-125; needs fields) */
enum { LSQED_BUILD_DIAG = 1, LSQED_BUILD_OFFDIAG = 2, LSQED_BUILD_DISORDER = 4 };

typedef struct lsqed_ham lsqed_ham;

/* Ordered forward-neighbour bonds of one kind (0 = nn, 1 = nnn) as
 * chain1D/square2D/honeycomb2D::get_nn/get_nnn list them (src/lattice.cpp:
 * 27-53, 86-136, 172-264), flattened in the loop order of
 * HamiltHelper::get_coup_elems (tmpl:350-387).  bi/bj may be NULL to query
 * the count.  Errors on lx*ly != N (lattice.cpp:67-68) and on odd honeycomb
 * sides (lattice.cpp:153-154). */
int lsqed_lattice_bonds(int lattice, int nspins, int lx, int ly, int kind, int* bi, int* bj, int* count);

/* Sector size C(N, nup) (Basis::compute_size, src/basis.cpp:38-46). */
int64_t lsqed_sector_dim(int nspins, int nup);

typedef struct {
  int nspins;         /* N, even and > 0 (src/environment.cpp:31-35) */
  int nup;            /* up spins = total_mag + N/2 (src/basis.cpp:66-71) */
  int n_nn, n_nnn;    /* bond counts */
  const int* nn_i;    /* host arrays, reference emission order */
  const int* nn_j;
  const int* nnn_i;
  const int* nnn_j;
  double J1, D1, J2, D2; /* Hamiltonian<L>(env, basis, lattice, J1, D1, J2, D2), tmpl:69-91 */
} lsqed_params;

/* Replaces Hamiltonian<L>::Hamiltonian + MatInit + prealloc (tmpl:8-91,
 * :432-531): records the model and uploads the rank tables.  Nothing of
 * size D is stored.  Build flags start at DIAG|OFFDIAG. */
int lsqed_ham_create(const lsqed_params* p, lsqed_ham** out);
int lsqed_ham_destroy(lsqed_ham* h);
int64_t lsqed_ham_dim(const lsqed_ham* h);
int lsqed_ham_nspins(const lsqed_ham* h);
int lsqed_ham_nup(const lsqed_ham* h);

/* Which build calls are in effect (LSQED_BUILD_* bits). */
int lsqed_ham_set_build(lsqed_ham* h, int flags);

/* The model as recorded: couplings (Hamiltonian<L>::hamilt_J1/D1/J2/D2,
 * include/hamiltonian.hpp:271-295), build flags and bond counts per kind.
 * Any output may be NULL.  Used by host code that formats the operator
 * (MatView) from the handle alone. */
int lsqed_ham_get_params(const lsqed_ham* h, double* J1, double* D1, double* J2, double* D2, int* flags, int* n_nn,
                         int* n_nnn);

/* Hamiltonian<L>::build_disorder(env, hi) (tmpl:97-125): records the N
 * longitudinal fields (host array, copied) and sets LSQED_BUILD_DISORDER;
 * the diagonal becomes D1/4*c_nn + D2/4*c_nnn + 1/2*sum_i(+-h_i).
 * hi == NULL clears the fields and the flag. */
int lsqed_ham_set_fields(lsqed_ham* h, const double* hi);

/* New couplings (J1, D1, J2, D2) for an existing model, as if the
 * Hamiltonian<L>(env, basis, lattice, J1, D1, J2, D2) constructor (tmpl:69-91)
 * had been called with them; lattice, sector and build flags are kept.  Used
 * by the differentiable E0 (finite differences in J) instead of rebuilding
 * the handle. */
int lsqed_ham_set_couplings(lsqed_ham* h, double J1, double D1, double J2, double D2);

/* y = H x (the MatMult of EPSSolve, src/solver.cpp:45).  x and y: D doubles
 * on the device, must not alias. */
int lsqed_hv(lsqed_ham* h, const double* x, double* y, lsqed_stream_t stream);

/* y = H x on complex vectors, as a complex-scalar PETSc build stores them
 * (README.md:18 of the reference requires complex PETSc; the MatMult of
 * EPSSolve, src/solver.cpp:45): x and y are D complex128 entries each,
 * interleaved (re, im) on the device, must not alias.  H is real, so both
 * parts are gathered together and each part equals lsqed_hv of that part
 * (same kernel arithmetic).  Uses the specialised row split when that is the
 * handle's kernel (lsqed_ham_get_hv_kernel == 3), else the chunk-table
 * kernel. */
int lsqed_hv_c128(lsqed_ham* h, const double* x, double* y, lsqed_stream_t stream);

/* H.v kernel choice: 0 = automatic (specialised row split when N <= 40 and
 * hipRTC is available, else the row split when N <= 32, else the generic
 * chunk-table kernel), 1 = generic chunk-table kernel, 2 = row-split kernel
 * (N <= 32), 3 = row split specialised to the lattice (compiled once per
 * bond structure with hipRTC; 64-bit row bases above N = 32).  No
 * counterpart in the reference (whose operator is an assembled matrix); used
 * by the tests to check every kernel and by the tuner.
 * Setting LSQED_JIT=0 in the environment keeps automatic mode off hipRTC. */
int lsqed_ham_set_hv_kernel(lsqed_ham* h, int variant);
/* The kernel lsqed_hv will use (1, 2 or 3), resolving automatic mode. */
int lsqed_ham_get_hv_kernel(lsqed_ham* h);
/* A 64-bit fingerprint of the H.v that lsqed_hv launches: kernel, grid and
 * model, the order in which rows are dealt to blocks (row-split kernels), and
 * for the specialised kernel the hash of its generated source (bonds,
 * couplings, build mode, tuning) and block size.  No counterpart in
 * the reference; measurement records (hardware counters of one launch) are
 * keyed by it, so that a changed kernel does not inherit an old record.  For
 * the built-in kernels (1, 2) it covers the launch and the model, not the
 * compiled code. */
int lsqed_ham_hv_fingerprint(lsqed_ham* h, uint64_t* fp);

/* Diagonal of H exactly as the reference stores it (MatGetDiagonal,
 * tests/unit_tests/build_diag/chain_1d/test_build_diag60.cpp:29): entries the
 * reference never sets are 0.  d: D doubles (device). */
int lsqed_diag(lsqed_ham* h, double* d, lsqed_stream_t stream);
/* The same diagonal into D complex128 entries (imaginary parts 0), for
 * MATOP_GET_DIAGONAL of a complex-scalar PETSc. */
int lsqed_diag_c128(lsqed_ham* h, double* d, lsqed_stream_t stream);

/* Column lists of rows [row0, row0+nrows) for one bond kind, in the
 * emission order of HamiltHelper::get_coup_elems (tmpl:350-387,
 * src/hamiltonian.cpp:72-99): cols[r*max_cols + c], ncol[r].  Device
 * buffers.  max_cols >= number of bonds of that kind. */
int lsqed_export_cols(lsqed_ham* h, int kind, int64_t row0, int64_t nrows, int max_cols, int64_t* cols,
                      int32_t* ncol, lsqed_stream_t stream);

/* Unrank rows [row0, row0+n) to states (BasisHelper::search_elem +
 * Basis::generate_int_basis, src/basis.cpp:9-36, :192-215) and rank states
 * (HamiltHelper::search_index, src/hamiltonian.cpp:9-27).  Device buffers.
 * lsqed_rank writes -1 for a state outside the handle's sector (a bit set at
 * or above N, or a number of set bits other than nup). */
int lsqed_unrank(lsqed_ham* h, int64_t row0, int64_t n, uint64_t* states, lsqed_stream_t stream);
int lsqed_rank(lsqed_ham* h, const uint64_t* states, int64_t n, int64_t* ranks, lsqed_stream_t stream);

/* ---------------------------------------------------------------- solver -- */
/* EPS settings of Solver::SolverInit (src/solver.cpp:7-24): HEP,
 * Krylov-Schur, smallest real; EPSSetDimensions(nev, ncv, mpd).  Values <= 0
 * select the SLEPc defaults: ncv = min(D, max(2 nev, nev+15)), mpd = ncv,
 * tol = 1e-8, maxit = max(100, 2 D / ncv). */
typedef struct {
  int nev, ncv, mpd;
  double tol;
  int maxit;
  uint64_t seed;  /* start vector: splitmix64 hash of (seed, row), uniform in [-1, 1) */
  /* Multiplicity check (no counterpart in SLEPc): once nev values have
   * converged they are locked and the iteration continues from a random
   * vector orthogonal to them until one more value converges, so that copies
   * of a degenerate level that the single start vector did not reach are
   * found; repeated while it changes the nev lowest, at most this many
   * rounds.  <= 0: off -- SLEPc's stopping rule, which the reference's tests
   * pin (tests/unit_tests/solver: one copy per level of a diagonal H). */
  int multiplicity_rounds;
} lsqed_eps_opts;

typedef struct lsqed_eps lsqed_eps;

/* Solver::SolverInit + Solver::solve (src/solver.cpp:7-52): device-resident
 * thick-restart Lanczos (Krylov-Schur for a symmetric operator) with full
 * CGS2 reorthogonalisation -- SLEPc's refined form (three passes over the
 * basis per step) below 2^16 states, a delayed form (two passes) from 2^16
 * states when ncv <= 32 (LSQED_KS_DELAYED=0/1 forces either); convergence
 * when ||r_i|| / |theta_i| < tol (EPS_CONV_REL).  Allocates (ncv+1)
 * D-vectors of workspace, and with the delayed form one more when the device
 * keeps 1 GiB free beyond it (the update pass then writes the next residual
 * to a column it does not read; results are the same either way).  On
 * sectors of 2^24 states or more the solver runs its products on a copy of
 * the model with the sites renumbered so that the specialised kernel's low
 * block holds more bonds (4x8 J1-J2: 34.2 -> 29.3 ms per product), when that
 * gains at least 4 bonds; lsqed_eps_solution returns eigenvectors in the
 * model's own basis order (LSQED_RELABEL=0 keeps the numbering, =1 applies it
 * on any sector). */
int lsqed_eps_create(lsqed_ham* h, const lsqed_eps_opts* o, lsqed_eps** out);
int lsqed_eps_solve(lsqed_eps* e, int* nconv, lsqed_stream_t stream);
/* Solver::solution (src/solver.cpp:58-68): eigenpair i (ascending) and the
 * relative error ||H x - lambda x|| / |lambda| (EPS_ERROR_RELATIVE).  x may
 * be NULL; otherwise D doubles on the device (unit norm). */
int lsqed_eps_solution(lsqed_eps* e, int i, double* lambda, double* x, double* error, lsqed_stream_t stream);
int lsqed_eps_iterations(const lsqed_eps* e);
/* Device bytes the solver holds (basis columns and scratch): memory planning
 * for sectors of several GB per vector, and which basis layout was taken.
 * No counterpart in the reference (PETSc/SLEPc report memory per object
 * through their logging). */
int64_t lsqed_eps_workspace_bytes(const lsqed_eps* e);
int lsqed_eps_destroy(lsqed_eps* e); /* Solver::SolverClean (src/solver.cpp:28-32) */

/* Plain three-vector Lanczos for the ground energy only (no
 * reorthogonalisation, no eigenvector): for sectors where the (ncv+1)-vector
 * workspace does not fit (N=36).  work: 3*D doubles on the device.
 * Stops when the lowest Ritz value changes by < tol relative over 5 steps
 * or after maxit steps. */
int lsqed_lanczos_ground(lsqed_ham* h, int maxit, double tol, uint64_t seed, double* work, double* e0, int* steps,
                         lsqed_stream_t stream);

/* One three-term Lanczos step (the unit of the steps/s metric), on
 * unnormalised vectors carried with scale factors (true v = sq*q,
 * true v_prev = sp*p):
 *   y = H q;  alpha = sq^2 (q.y);  p <- sq*y - alpha*sq*q - beta*sp*p;
 *   beta' = ||p||.  The next Lanczos vector is p / beta'.
 * Traffic: one H.v (with the alpha dot fused) plus one pass reading y, q, p
 * and writing p.  q, p, y: D doubles on the device. */
int lsqed_lanczos_step(lsqed_ham* h, const double* q, double sq, double* p, double sp, double beta, double* y,
                       double* alpha, double* beta_next, lsqed_stream_t stream);

/* lsqed_lanczos_step without any device allocation: no counterpart in the
 * reference; for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * kernels and results as lsqed_lanczos_step; the partial sums go to a static
 * device array of 2^22 doubles (32 MiB per device, held until the process
 * ends), found once per device with hipGetSymbolAddress.  It calls no
 * hipMalloc/hipFree, so the model's H.v plan must already exist: call
 * lsqed_hv or lsqed_lanczos_step on the handle once first (and again after
 * set_fields / set_couplings / set_build / set_hv_kernel), else
 * LSQED_ERR_STATE.  A sector whose step needs more than 2^22 partial sums
 * (H.v blocks + reduction blocks + 8; N >= 38 with the specialised kernel) is
 * rejected with LSQED_ERR_ARG.  Every call on one device shares the array:
 * calls that overlap on one device (several host threads or streams) race. */
int lsqed_lanczos_step_noalloc(lsqed_ham* h, const double* q, double sq, double* p, double sp, double beta,
                               double* y, double* alpha, double* beta_next, lsqed_stream_t stream);

/* ----------------------------------------------------------- observables -- */
/* All pairs <S^z_i S^z_j> (SzSz::ExpectVal, src/spinOperators/correlation.cpp
 * :86-139), <S^+_i S^-_j> (SpSm, :154-224) for a real state psi (device,
 * D doubles).  out: N*N doubles on the host, row-major [i*N + j]. */
int lsqed_obs_szsz(lsqed_ham* h, const double* psi, double* out, lsqed_stream_t stream);
int lsqed_obs_spsm(lsqed_ham* h, const double* psi, double* out, lsqed_stream_t stream);
/* <S^z_i> for all i: out N doubles (host). */
int lsqed_obs_sz(lsqed_ham* h, const double* psi, double* out, lsqed_stream_t stream);
/* DSF amplitudes a[n*N + i] = <n| S_i |0> as Phys::DynStructFactor forms them
 * (include/phys.tmpl_func.hpp:166-186) with Sqz (0.5*(+-1), same sector,
 * spinOPft.tmpl_func.hpp:16-60) when h1 == h0, or Sqp (S^+_i into the sector
 * of h1 with one more up spin, :77-143).  psi0: D0 doubles, evecs: nev
 * column-major D1-vectors (device).  out: nev*N doubles (host). */
int lsqed_obs_dsf(lsqed_ham* h0, lsqed_ham* h1, const double* psi0, const double* evecs, int nev, double* out,
                  lsqed_stream_t stream);

/* ------------------------------------------------------- batched solver -- */
/* Ground states of many models at once: the loop of examples/
 * main_disorder.cpp:82-120 (per realisation build_disorder, SolverInit with
 * nev = 1, solve, solution) as one batched three-term Lanczos.  The models
 * may differ in N (<= 32), lattice, couplings and fields.  Every Lanczos step
 * of the whole batch is one fixed launch sequence over a task list of
 * (model, row range) pairs; steps are replayed from a captured hipGraph and
 * the host retires converged models between replays.  Convergence: the
 * lowest Ritz pair of the tridiagonal matrix has |beta_m s_m| / |theta| < tol
 * (the EPS_CONV_REL test of the Krylov-Schur solver).
 *   maxit            Lanczos steps per model (<= 0: 300)
 *   tol              (<= 0: 1e-9)
 *   seed             start vectors (0: 1)
 *   lanes_per_state  0: autotune the launch shape (rows per task, lanes per
 *                    basis state, blocks per CU) by timing it on the batch,
 *                    cached per batch shape; 1, 2, 4: fixed lanes, default tile
 *   keep_vectors     1: a second, identical pass builds the ground-state
 *                    vectors (one more D-vector per model)
 *   hv_kernel        0: autotune chooses; 1: chunk-table kernel over the task
 *                    list; 2: row split specialised per lattice group (hipRTC,
 *                    couplings and fields per model at run time) */
typedef struct {
  int maxit;
  double tol;
  uint64_t seed;
  int lanes_per_state;
  int keep_vectors;
  int hv_kernel;
} lsqed_batch_opts;

typedef struct lsqed_batch lsqed_batch;

/* Device bytes lsqed_batch_create allocates for the vectors (3 or 4 doubles
 * per basis state of every model); -1 on invalid input. */
int64_t lsqed_batch_workspace_bytes(int n, lsqed_ham* const* hams, const lsqed_batch_opts* o);
/* The models are referenced, not copied: set_fields / set_couplings /
 * set_build on them between solves take effect at the next solve. */
int lsqed_batch_create(int n, lsqed_ham* const* hams, const lsqed_batch_opts* o, lsqed_batch** out);
/* e0: n doubles (host), the lowest eigenvalue of each model. */
int lsqed_batch_solve(lsqed_batch* b, double* e0, lsqed_stream_t stream);
/* Lanczos steps, relative residual and convergence flag of model k. */
int lsqed_batch_info(const lsqed_batch* b, int k, int* steps, double* resid, int* converged);
/* Launch shape in use (the autotuner's choice): H.v kernel (1 chunk table,
 * 2 specialised row split), rows per task, lanes per basis state, blocks per
 * CU, whether replays run from a captured hipGraph (1) or are enqueued
 * launch by launch (0), and the tuner's time of one Lanczos step of the
 * whole batch with every model active (ms).  Any pointer may be NULL. */
int lsqed_batch_schedule(const lsqed_batch* b, int* hv_kernel, int* rows_per_task, int* lanes_per_state,
                         int* blocks_per_cu, int* graph, double* step_ms);
/* Unit-norm ground state of model k into psi (D_k doubles, device). */
int lsqed_batch_ground_state(lsqed_batch* b, int k, double* psi, lsqed_stream_t stream);
/* Hellmann-Feynman derivatives of every model's E0 from its ground state,
 * out[k*LSQED_BATCH_OBS_STRIDE + j] (host):
 *   j < N : <S^z_j>                                    (= dE0/dh_j)
 *   32    : <psi|psi> before normalisation
 *   33/35 : sum over nn / nnn bonds of <S^z_i S^z_j>    (= dE0/dD1, dE0/dD2)
 *   34/36 : sum over nn / nnn bonds of <(S^+_i S^-_j + S^-_i S^+_j)/2>
 *                                                       (= dE0/dJ1, dE0/dJ2)
 * The correlators are SzSz/SpSm::ExpectVal (src/spinOperators/
 * correlation.cpp:86-224) summed over the Hamiltonian's bonds. */
#define LSQED_BATCH_OBS_STRIDE 40
int lsqed_batch_observables(lsqed_batch* b, double* out, lsqed_stream_t stream);
int lsqed_batch_destroy(lsqed_batch* b);

/* ----------------------------------------------------------- measurement -- */
/* Streaming-bandwidth probes, no counterpart in the reference: the measured
 * ceiling reported beside the nominal HBM peak.  kind 0: STREAM triad
 * a = b + 0.5 c (24 B per element); kind 1: read-only stream of b (8 B per
 * element; n even; part: 2048 doubles of block sums).  kinds 2 and 3
 * calibrate the HBM byte counters on scattered reads: a[i] = the first double
 * of piece (i * 1000003 mod n) of b, pieces of 16 doubles (kind 2: one per
 * 128-byte line, b holds 16 n doubles) or 8 doubles (kind 3: one per 64-byte
 * half line, b holds 8 n doubles); n < 2^34 and not a multiple of 1000003
 * (prime), so every piece is read once.
 * Device buffers. */
int lsqed_stream_probe(int kind, double* a, const double* b, const double* c, int64_t n, double* part,
                       lsqed_stream_t stream);

/* ---------------------------------------------------------------- misc ---- */
/* Tools::RandomDisorder (src/tools/random_disorder.cpp:33-59): the
 * reproducible seed list (mt19937_64(1) -> uniform_int(0,1000)) and the
 * fields h_i = uniform_real(min, max) from mt19937_64(seed). */
int lsqed_rd_seeds(int rep, int64_t* seeds);
int lsqed_rd_fields(int64_t seed, int nspins, double lo, double hi, double* out);

const char* lsqed_last_error(void);
const char* lsqed_version(void);

#ifdef __cplusplus
}
#endif
#endif /* LSQED_H */
