/*
 * shci_gpu.h — C ABI of the MI355X-native SHCI variational step.
 *
 * Plain C types only (pointers, sizes, status codes).  Determinants are
 * passed as `words` 64-bit integers each (words = 2*norb/64 + 1, the
 * reference's EffDetLen), bit k = spin-orbital k, even bits alpha, odd bits
 * beta, word 0 first — the in-memory layout of the reference's
 * `Determinant::repr` (SHCI/Determinants.h:166-200).  Every call is ordered
 * on the handle's HIP stream and returns when its results are in the caller's
 * buffers.  A handle is not re-entrant: one host thread per handle.
 *
 * Reference interfaces replaced (sanshar/Dice):
 *   readIntegrals + twoIntHeatBath(SHM)::constructClass  SHCI/integral.cpp:43,228; SHCI/integral.h:125
 *   Determinant::Energy / Hij                           SHCI/Determinants.cpp:252,564
 *   getDeterminantsVariationalApprox + sort/unique      SHCI/SHCIgetdeterminants.cpp:586; SHCI/SHCIbasics.cpp:969-1077
 *   PopulateHelpers + MakeHfromSMHelpers2               SHCI/SHCImakeHamiltonian.cpp:290,486
 *   Hmult2::operator()(x, y)                            SHCI/Hmult.h:52
 *   davidson(Hmult2&, x0, diag, maxCopies, tol, ...)    SHCI/Davidson.cpp:111
 *   SHCIbasics::DoVariational                           SHCI/SHCIbasics.cpp:783
 *   main (`Dice input.dat`, variational part)           SHCI/SHCI.cpp:157-446
 */
#ifndef SHCI_GPU_H
#define SHCI_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  SHCI_OK = 0,
  SHCI_BAD_INPUT = 1,      /* malformed file / argument (reference: message + exit) */
  SHCI_OOM = 2,            /* device allocation failed */
  SHCI_NO_CONVERGENCE = 3, /* Davidson iteration cap (reference: Davidson.cpp:305-309) */
  SHCI_HIP_ERROR = 4,      /* HIP runtime error */
  SHCI_UNSUPPORTED = 5     /* keyword / size outside this implementation */
} shci_status;

/* value precision of the stored Hamiltonian */
typedef enum {
  SHCI_PREC_FP64 = 0, /* FP64 values: the plain drop-in */
  SHCI_PREC_FP32 = 1  /* FP32 values, FP64 accumulation, FP64 residual refinement */
} shci_precision;

typedef struct shci_handle shci_handle;

/* Read an FCIDUMP and build the screening tables on the current device.
 * Replaces readIntegrals (SHCI/integral.cpp:43-225) and
 * twoIntHeatBath(SHM)::constructClass (SHCI/integral.h:125-153,
 * SHCI/integral.cpp:228-389), as called from SHCI/SHCI.cpp:213,248-252. */
shci_status shci_create(const char* fcidump_path, shci_handle** out);
/* The same from integrals in memory, as SHCIbasics::DoVariational takes them
 * (SHCI/SHCIbasics.h:89-93: oneInt, twoInt, irrep, coreE, nelec) instead of
 * a file: h1 = norb x norb one-electron integrals, row-major (its lower
 * triangle is used and mirrored, as an FCIDUMP lists each pair once);
 * eri8 = the two-electron integrals (pq|rs) packed 8-fold in the reference's
 * twoInt order (SHCI/integral.h:65-102): element A(A+1)/2 + B for pair
 * indices A >= B, pair(p,q) = p(p+1)/2 + q for p >= q (PySCF's
 * ao2mo.restore(8, eri, norb) layout), npair(npair+1)/2 values with
 * npair = norb(norb+1)/2; core_energy = the constant term; orbsym = norb
 * irreps or NULL (all 1).  norb <= 63. */
shci_status shci_create_from_integrals(int norb, int nelec, const double* h1, const double* eri8, double core_energy,
                                      const int* orbsym, shci_handle** out);
shci_status shci_destroy(shci_handle* h);
/* last error message of the handle (or of the last failed shci_create when h is NULL) */
const char* shci_last_error(const shci_handle* h);

int shci_norb(const shci_handle* h);
int shci_nelec(const shci_handle* h);
int shci_det_words(const shci_handle* h);
double shci_core_energy(const shci_handle* h);

/* Hamiltonian value precision for subsequent shci_build_hamiltonian calls. */
shci_status shci_set_precision(shci_handle* h, int precision);
/* Restrict excitations as the ncore / nact keywords do (SHCI/SHCIgetdeterminants.cpp:649,694,711)
 * and add RAS restrictions (`restrict min max orbs...`, SHCI/OccRestrictions.cpp,
 * SHCI/input.cpp:267-276: to the variational and the PT space). */
shci_status shci_set_active_space(shci_handle* h, int ncore, int nact);
shci_status shci_add_restriction(shci_handle* h, int min_elec, int max_elec, const int* spatial_orbs, int n);

/* ---- determinant primitives ---------------------------------------------------------- */
/* Determinant::Energy with coreE = 0 (SHCI/Determinants.cpp:252-300). */
shci_status shci_det_energy(shci_handle* h, const int64_t* dets, int64_t n, double* energies);
/* Hij(bra[k], ket[k]) (SHCI/Determinants.cpp:564-631); 0 beyond two excitations. */
shci_status shci_hij(shci_handle* h, const int64_t* bra, const int64_t* ket, int64_t n, double* out);

/* ---- variational space --------------------------------------------------------------- */
/* Replace the handle's determinant list (order kept; duplicates are an error). */
shci_status shci_set_dets(shci_handle* h, const int64_t* dets, int64_t n);
shci_status shci_num_dets(const shci_handle* h, int64_t* n);
/* Copy up to cap determinants out (in variational-space order). */
shci_status shci_get_dets(shci_handle* h, int64_t* out, int64_t cap);

/* Heat-bath screening of every held determinant with coefficient bound
 * cmax[i] against eps1 (getDeterminantsVariationalApprox, SHCI/SHCIgetdeterminants.cpp:586-739),
 * followed by sort + unique of the candidates not already held
 * (SHCI/SHCIbasics.cpp:984-1055).  The new determinants, in the reference's
 * signed-word order, are staged; *n_new receives their count. */
shci_status shci_screen(shci_handle* h, const double* cmax, double eps1, int64_t* n_new);
/* Copy the staged new determinants out (up to cap). */
shci_status shci_get_new_dets(shci_handle* h, int64_t* out, int64_t cap);
/* Append the staged determinants after the held ones (SHCI/SHCIbasics.cpp:1062-1077). */
shci_status shci_append_new(shci_handle* h);

/* ---- Hamiltonian ----------------------------------------------------------------------- */
/* Build H over the held determinants: the diagonal (Determinant::Energy) and
 * every pair one or two excitations apart, explicit zeros included
 * (SHCImakeHamiltonian::PopulateHelpers + MakeHfromSMHelpers2,
 * SHCI/SHCImakeHamiltonian.cpp:290-408,486-715).  *nnz_lower receives the
 * number of stored elements of the reference's lower-triangular rows
 * (diagonal + J < I). */
shci_status shci_build_hamiltonian(shci_handle* h, int64_t* nnz_lower);
/* The reference's row layout for rows [row_begin, n): row I = diagonal, then
 * every J < I, columns ascending after the diagonal.  rowptr has n-row_begin+1
 * entries relative to the first requested row. */
shci_status shci_get_hamiltonian_rows(shci_handle* h, int64_t row_begin, int64_t* rowptr, int32_t* cols,
                                      double* vals, int64_t cap);

/* ---- sigma and Davidson --------------------------------------------------------------- */
/* y[:, r] = H x[:, r] for r < k; x, y host, column-major n x k
 * (Hmult2::operator(), SHCI/Hmult.h:52-150). */
shci_status shci_sigma(shci_handle* h, const double* x, double* y, int k);
/* Device-pointer variant on the handle's stream (stream = NULL) for callers
 * that keep vectors resident; returns after enqueueing.  ld >= n; d_y must
 * not overlap d_x (SHCI_BAD_INPUT). */
shci_status shci_sigma_device(shci_handle* h, const double* d_x, double* d_y, int k, int64_t ld);
/* shci_sigma_device without any device allocation.  No reference
 * counterpart: it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * Hamiltonian must already be built (shci_build_hamiltonian).  The permuted
 * x and work counters live in one fixed scratch region per GPU (64 MiB of x:
 * with split columns n * vectors-per-pass up to 4,194,304 in FP64, where a
 * pass takes 1, 2, 4 or 8 of the k vectors; with half storage the permuted x
 * and the 8-byte fixed-point sums, n up to about 2 million in FP64 with
 * split columns), held until the process ends;
 * a larger request returns SHCI_UNSUPPORTED before anything is enqueued, as
 * do time-reversal and direct mode.  Every call on one GPU shares that
 * region: calls that overlap on one GPU (other handles or streams) race.
 * Results are those of shci_sigma_device, bit for bit. */
shci_status shci_sigma_device_noalloc(shci_handle* h, const double* d_x, double* d_y, int k, int64_t ld);
/* Davidson with the reference's semantics (SHCI/Davidson.cpp:111-348):
 * x0 host n x nroots column-major guess in, eigenvectors out; energies out
 * (coreE not added); *iterations = Davidson iteration count.  With the
 * davidson_block option (shci_set_options) the block Davidson runs instead
 * and max_copies is ignored. */
shci_status shci_davidson(shci_handle* h, double* x0, int nroots, int max_copies, double tol, double* energies,
                          int* iterations);
/* Deterministic Epstein-Nesbet PT2 of one root over the held determinants
 * (SHCIbasics::DoPerturbativeDeterministic, SHCI/SHCIbasics.cpp:453-728, with
 * the candidates of SHCIgetdeterminants::getDeterminantsDeterministicPT,
 * SHCI/SHCIgetdeterminants.cpp:40-171): ci = the root's n coefficients
 * (host), e0 = its variational energy without the core energy; *e_pt = the
 * PT energy (add e0 and the core energy for the reference's PTEnergy);
 * psi1_norm / n_external (optional) = sum (num / (e0 - E_a))^2 and the
 * number of external determinants.  The PT space honours the restrictions
 * of shci_add_restriction (the "restrict" keyword), not ncore / nact. */
shci_status shci_pt_deterministic(shci_handle* h, const double* ci, double e0, double eps2, double* e_pt,
                                  double* psi1_norm, int64_t* n_external);
/* RDMs of <bra| ... |ket> over the held determinants (bra, ket: n host
 * coefficients each; SHCIrdm::EvaluateOneRDM and SHCIrdm::EvaluateRDM,
 * SHCI/SHCIrdm.cpp:996-1100 and :863-995, whose connections this call
 * builds as shci_build_hamiltonian does):
 *   one_rdm[P*n+R]           = sum_s <a+_{P s} a_{R s}>          (n = orbitals)
 *   two_rdm[((P*n+Q)*n+R)*n+S] = sum_s,t <a+_{P s} a+_{Q t} a_{S t} a_{R s}>
 *                              (the reference's s2RDM(P*n+Q, R*n+S))
 * and, when not null, spin_one_rdm[p*2n+r] = <a+_p a_r> over spin-orbitals
 * and spin_two_rdm[pair(p,q)*npair + pair(r,s)] = <a+_p a+_q a_s a_r> for
 * p > q, r > s (pair(p,q) = p(p+1)/2 + q, npair = 2n(2n+1)/2; the other
 * entries are zero), the reference's twoRDM of DoSpinRDM. */
shci_status shci_rdm(shci_handle* h, const double* bra, const double* ket, double* one_rdm, double* two_rdm,
                     double* spin_one_rdm, double* spin_two_rdm);
/* Semistochastic EN-PT2 of one root over the held determinants
 * (SHCIbasics::DoPerturbativeStochastic2SingleListDoubleEpsilon2AllTogether,
 * SHCI/SHCIbasics.cpp:55-452, with the alias sampling of
 * SHCI/SHCIsampledeterminants.cpp:37-181): the deterministic PT at
 * eps2_large plus sampled iterations of sample_n draws (C library rand()
 * after srand(seed)) until the standard error is below target_error or
 * n_iter iterations ran.  *e_pt = the reference function's value (the mean
 * correction; the unaveraged sum when n_iter runs out); per_iter (cap
 * entries, optional) = each iteration's estimate; *iterations, *error
 * (optional) = iterations run and the last standard error. */
shci_status shci_pt_semistochastic(shci_handle* h, const double* ci, double e0, double eps2, double eps2_large,
                                   int sample_n, int n_iter, double target_error, unsigned seed, double* e_pt,
                                   double* per_iter, int cap, int* iterations, double* error);

/* ---- whole variational step ----------------------------------------------------------- */
/* Run the variational step of `Dice input.dat` (SHCI/SHCI.cpp:157-446,
 * SHCIbasics::DoVariational SHCI/SHCIbasics.cpp:783-1378): reads the input
 * file and the FCIDUMP it names, writes <prefix>/shci.e, and prints the
 * reference's tables to log_path (stdout when NULL or empty).  With the
 * `deterministic` keyword (and nblocks 1) the deterministic PT step follows
 * (SHCI/SHCI.cpp:706-721), otherwise with sampleN set and nPTiter != 0 the
 * semistochastic one (SHCI/SHCI.cpp:722-741); shci.e then holds
 * E_var + E_PT per root.
 * precision as shci_precision.  The result handle holds the final space for
 * inspection. */
shci_status shci_run(const char* input_path, const char* log_path, int precision, shci_handle** out);

/* Schedule and tolerances of shci_do_variational: the fields of the
 * reference's schedule (SHCI/input.h:35-194) that DoVariational reads. */
typedef struct {
  int nroots;
  int n_iter;                /* number of variational iterations: entries of eps1 */
  const double* eps1;        /* eps1 of each iteration (the reference's expanded schd.epsilon1) */
  double davidson_tol;       /* final tight Davidson (davidsonTol) */
  double davidson_tol_loose; /* the iterations' Davidson (davidsonTolLoose) */
  double dE;                 /* convergence of the lowest energy between iterations */
  unsigned seed;             /* srand(seed) of the random Davidson guesses (SHCI/SHCI.cpp:199) */
  int ncore, nact;           /* excitation limits (0 and >= norb for none) */
  int reserved[4];
} shci_var_params;
/* SHCIbasics::DoVariational (SHCI/SHCIbasics.cpp:783-1378, declared at
 * SHCI/SHCIbasics.h:89-93, called from SHCI/SHCI.cpp:422 and
 * SHCI/MBE.cpp:441-446) on in-memory data: the variational space starts as
 * dets (ndets x words, unique), with coefficients ci (ndets x nroots
 * column-major; NULL: every determinant 1 and the lowest-energy one
 * + ndets, normalised, root columns >= 1 zero, as SHCI/SHCI.cpp:395-399);
 * the iterations select, build and diagonalise as shci_run does (method
 * options from shci_set_options; no files, no table, no PT).  energies =
 * nroots variational energies with the core energy added.  The handle then
 * holds the final space: shci_num_dets / shci_get_dets, shci_run_ci,
 * shci_run_iteration. */
shci_status shci_do_variational(shci_handle* h, const shci_var_params* p, const int64_t* dets, int64_t ndets,
                                const double* ci, double* energies);

/* Options of shci_run_ex and shci_set_options (zero-initialise, then set
 * fields; all zero = the reference's method in FP64). */
typedef struct {
  int precision;      /* shci_precision */
  int profile_sigma;  /* nonzero: time every sigma launch with HIP events on the handle's stream */
  int davidson_block; /* nonzero: block Davidson (all unconverged roots corrected per iteration, the new
                         block's sigma as one SpMM) in place of the reference's single-vector Davidson
                         (Davidson.cpp:111-348); up to 8 roots */
  int autotune;       /* nonzero: choose the sigma launch shape (rows per work item, static or persistent
                         workgroups) by timing on each Hamiltonian whose size has doubled */
  int hip_graphs;     /* nonzero (with davidson_block): replay block-Davidson iterations from captured
                         hipGraphs */
  int learned;        /* learned screening form (shci_run / shci_run_ex only).  3: the first two
                         iterations screen in full and log training samples, an MLP trained on them then
                         predicts log10(|c| |v|) for the double-excitation candidates (heat-bath value
                         >= eps / 10) and only those predicted above log10(eps1) - 0.8 (environment
                         SHCI_LEARNED_MARGIN) are tested exactly; singles stay exact; when a sample shows
                         that more than half would be tested, that iteration and the later ones screen in
                         full.  1: the same, gated: after the second logging iteration the form turns
                         itself off (no training, full screening) unless ranking that iteration's candidate
                         records at 2 ns each would cost less than its exact screen.  2: as 3 with the
                         untrained network (control) */
  int direct_sigma;   /* nonzero: no stored Hamiltonian; every sigma evaluates the elements from the string
                         tables (HmultDirect, SHCI/Hmult.h:153-382).  shci_run_ex also turns it on, with
                         davidsonDirect's rules (SHCI/Davidson.cpp:353-590), for the `directdavidson`
                         keyword */
  int half_storage;   /* > 0: each connected pair stored once, in row max(I, J) (the reference's
                         lower-triangular rows, SHCI/SHCImakeHamiltonian.cpp:549-711), sigma applying
                         each stored element to both rows (SHCI/Hmult.h:110-121; on a tiled copy, the
                         transposed half summed in 64-bit fixed point, so in a fixed result); < 0: both
                         triangles stored; 0: the library's choice (environment SHCI_STORAGE=half/full,
                         else half storage, the fastest layout measured for every form).  A build
                         switches to half storage by itself when the two-triangle Hamiltonian does not
                         fit */
} shci_run_options;
/* shci_run with options; same files, log and result handle. */
shci_status shci_run_ex(const char* input_path, const char* log_path, const shci_run_options* options,
                        shci_handle** out);
/* Method options for the piecewise calls on h (precision, profile_sigma,
 * davidson_block, autotune, hip_graphs); a precision change resets the
 * Hamiltonian like shci_set_precision.  No reference counterpart: the
 * reference has one method. */
shci_status shci_set_options(shci_handle* h, const shci_run_options* options);

/* ---- candidate-ranking MLP of the learned screening form ------------------------------
 * No reference counterpart (the reference screens with the heat-bath lists
 * alone).  A network din -> hidden -> hidden -> 1 with GELU, FP32 parameters
 * in the flat layout W1 [hidden x din], b1, W2 [hidden x hidden], b2,
 * W3 [hidden], b3; matrix products on bf16 MFMA with FP32 accumulation.
 * Inputs X are n x din row-major FP32, host buffers; these calls use their
 * own stream on the current device (errors: shci_last_error(NULL)). */
int64_t shci_mlp_param_count(int din, int hidden);
/* Xavier-uniform weights, zero biases */
shci_status shci_mlp_init(float* params, int din, int hidden, uint64_t seed);
shci_status shci_mlp_forward(const float* params, int din, int hidden, const float* X, int64_t n, float* y);
/* loss = 0.5 mean((y - t)^2) and its gradient (parameter layout) */
shci_status shci_mlp_loss_grad(const float* params, int din, int hidden, const float* X, const float* t, int64_t n,
                               float* loss, float* grad);
/* Adam (lr, betas 0.9 / 0.999) over shuffled minibatches for `epochs` passes; params updated in place */
shci_status shci_mlp_train(float* params, int din, int hidden, const float* X, const float* t, int64_t n, int epochs,
                           int batch, float lr, uint64_t seed, float* first_loss, float* last_loss);

/* learned screening records of the last run: one per ranked iteration */
int shci_run_learned_count(const shci_handle* h);
shci_status shci_run_learned(const shci_handle* h, int k, int* iteration, int64_t* candidates, int64_t* evaluated,
                             int64_t* accepted, double* sampled_recall, double* sampled_fraction, int* fallback,
                             double* seconds);
/* training of the learned form's network in the last run */
shci_status shci_run_learned_training(const shci_handle* h, int64_t* samples, float* first_loss, float* last_loss,
                                      double* seconds);

/* per-iteration records of the last shci_run on this handle */
int shci_run_iterations(const shci_handle* h);
shci_status shci_run_iteration(const shci_handle* h, int it, int64_t* ndets, double* energy0, int* davidson_iters,
                               double* seconds);
/* wall time of each phase of iteration `it` (screening + merge, Hamiltonian build, Davidson) and the
 * element count of the reference's lower-triangular rows after it */
shci_status shci_run_iteration_phases(const shci_handle* h, int it, double* screen_s, double* hamiltonian_s,
                                      double* davidson_s, int64_t* nnz_lower);
/* wall time of the final tight Davidson of the last run (seconds) */
double shci_run_final_davidson_seconds(const shci_handle* h);
/* sigma launches timed since the run started (profile_sigma): count, summed kernel time (HIP
 * events) and summed algorithmic bytes (stored values + indices, row offsets, diagonal, x and y once) */
shci_status shci_profile_sigma_read(shci_handle* h, int64_t* calls, double* milliseconds, double* bytes);
shci_status shci_run_energies(const shci_handle* h, double* energies, int* nroots);
/* E_var + E_PT per root when the run did the deterministic PT step (*nroots = 0 otherwise) */
shci_status shci_run_pt_energies(const shci_handle* h, double* energies, int* nroots);
/* final coefficients, n x nroots column-major */
shci_status shci_run_ci(shci_handle* h, double* out, int64_t cap);

#ifdef __cplusplus
}
#endif
#endif /* SHCI_GPU_H */
