/*
 * itp2d.h -- C ABI of the MI355X (gfx950) imaginary-time-propagation engine.
 *
 * Plain C: opaque handles, pointers and sizes, int status codes. No C++
 * exceptions cross this boundary; the C++ facade (itp2d_amd/host/itp2d.hpp)
 * maps the codes back to the reference's exception types.
 *
 * Each entry point names the reference interface it replaces (paths under
 * src/ of the reference). All work is queued on the context's one HIP stream;
 * calls that return host-visible results synchronise that stream.
 *
 * Field layout in HBM: [N][sy][sx] interleaved complex (re, im), the same
 * order as the reference's StateArray (statearray.cpp:21-75); double for
 * ITP_FP64 sets, float for ITP_FP32 sets. Host-side buffers passed to this API
 * are always interleaved complex double.
 */
#ifndef ITP2D_H
#define ITP2D_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. The three eigen errors keep the reference's semantics:
 * NonPositiveEigenvalue / NonNormalEigenvalue (stateset.cpp:164-173,
 * exceptions.hpp:55-86) and EigensolverError (eigensolver.hpp:80-84). */
enum itp_status {
  ITP_OK = 0,
  ITP_E_INVALID = 1,
  ITP_E_NONPOSITIVE_EIG = 2,
  ITP_E_NONNORMAL_EIG = 3,
  ITP_E_EIGENSOLVER = 4,
  ITP_E_HIP = 5,
  ITP_E_UNSUPPORTED = 6,
  ITP_E_NOMEM = 7
};

/* BoundaryType (itp2d_common.hpp:50) */
enum itp_boundary { ITP_PERIODIC = 0, ITP_DIRICHLET = 1 };

/* enum Transform, same order (transformer.hpp:27-28) */
enum itp_transform_kind {
  ITP_FFT = 0, ITP_iFFT, ITP_FFTx, ITP_iFFTx, ITP_FFTy, ITP_iFFTy,
  ITP_DST, ITP_iDST, ITP_DSTx, ITP_iDSTx, ITP_DSTy, ITP_iDSTy,
  ITP_DCT, ITP_iDCT, ITP_DCTx, ITP_iDCTx, ITP_DCTy, ITP_iDCTy
};

enum itp_precision { ITP_FP64 = 0, ITP_FP32 = 1 };

typedef struct itp_ctx itp_ctx;
typedef struct itp_grid itp_grid;
typedef struct itp_potential itp_potential;
typedef struct itp_stateset itp_stateset;
typedef struct itp_split itp_split;
typedef struct itp_hamiltonian itp_hamiltonian;

/* ---- context: one device, one stream (replaces the process-wide OpenMP
 * setup, itpsystem.cpp:51) ---- */
int itp_ctx_create(int device, itp_ctx** out);
/* Objects created on a context may be destroyed before or after its handle;
 * the context itself (stream, cached device memory) goes with the last of
 * them. A destroyed object's device memory is kept by its context for the
 * objects created after it (up to 4 GiB per context; reused memory reads as
 * zero): no hipFree, which waits for every stream of the device, between the
 * instances a batch worker runs one after another. */
void itp_ctx_destroy(itp_ctx* ctx);
const char* itp_last_error(const itp_ctx* ctx);
int itp_ctx_synchronize(itp_ctx* ctx);
/* Device properties used by measurement code: number of CUs. */
int itp_ctx_num_cus(const itp_ctx* ctx);

/* ---- grid: DataLayout (datalayout.hpp:30-54, datalayout.cpp:21-36) plus the
 * Transformer's k tables and normalisation factors (transformer.cpp:66-90).
 * dx = lenx/sx as Parameters::get_grid_delta (parameters.hpp:97). ---- */
int itp_grid_create(itp_ctx* ctx, size_t sx, size_t sy, double dx, int boundary, itp_grid** out);
void itp_grid_destroy(itp_grid* g);
/* Transformer::fft_kx/fft_ky/dsct_kx/dsct_ky (transformer.hpp:36-45); any pointer may be NULL. */
int itp_grid_kvalues(const itp_grid* g, double* fft_kx, double* fft_ky, double* dsct_kx, double* dsct_ky);
/* Transformer::normalization_factor(Transform) (transformer.hpp:94-130) */
double itp_grid_normalization_factor(const itp_grid* g, int kind);
/* DataLayout::get_posx/get_posy (datalayout.hpp:56-61) */
int itp_grid_positions(const itp_grid* g, double* posx, double* posy);

/* Transformer::transform(comp*, Transform) (transformer.hpp:132-158) applied to
 * nfields fields already on the device (interleaved complex double). */
int itp_transform_device(itp_grid* g, void* fields, size_t nfields, int kind);
/* Same on host buffers (upload, transform, download). */
int itp_transform_host(itp_grid* g, double* fields, size_t nfields, int kind);

/* ---- potential: Potential (potential.hpp:31-58). V is [sy][sx] doubles;
 * V == NULL is the null potential (Potential::is_null). ---- */
int itp_potential_create(itp_grid* g, const double* V, itp_potential** out);
void itp_potential_destroy(itp_potential* p);
/* A named potential type of the reference (parse_potential_description and
 * the PotentialType classes, potentialtypes.cpp:56-500, potentialtypes.hpp:
 * 72-382): "zero", "harmonic" (A, x0, y0), "elliptic" (Ax, Ay),
 * "prettyhardsquare" (e), "softpentagon", "henonheiles"/"henon" (a, b),
 * "gaussian"/"gaussianblob" (A, w, x0, y0), "quartic"/"quarticoscillator" (b),
 * "squareoscillator" (alpha), "power"/"poweroscillator" (a, A),
 * "ring"/"ringpotential" (R, w, e, A_asym, w_asym), "cosh"/"coshpotential"
 * (A, L), "softstadium" (R, L, V, a, b), "powerstadium" (R, L, a), with the
 * reference's accepted parameter counts, defaults and validity checks
 * (ITP_E_INVALID otherwise). Evaluated on the device on the cell-centred grid
 * (Potential::init_values, potential.cpp:54-63). "zero" is the null
 * potential until impurities are added. */
int itp_potential_create_type(itp_grid* g, const char* name, const double* params, size_t nparams,
                              itp_potential** out);
/* Host-side parameter resolution of itp_potential_create_type: the kind
 * index (order of the list above) and 5 resolved parameters (defaults filled
 * in; stadium lengths halved as the reference stores them). */
int itp_potential_type_params(const char* name, const double* params, size_t nparams, int* kind, double* resolved);
/* Impurities added to V (SpatialImpurities::add_noise with the ImpurityType
 * formulas, noise.cpp:25-110), given as realised positions and parameters
 * (the reference draws them from its own RNG):
 * ITP_IMP_GAUSSIAN  data [count][4] = x, y, A, w: += A exp(-r^2 / (2 w^2));
 * ITP_IMP_COULOMB   data [count][2] = x, y; kind_params = {e, alpha}:
 *                   += alpha r^-e;
 * ITP_IMP_HEMISPHERE data [count][2]; kind_params = {A, radius}:
 *                   += A sqrt(1 - r^2 / radius^2) inside the radius;
 * ITP_IMP_DELTA     data [count][3] = x, y, A: += A at the nearest grid point.
 * Add impurities before creating splits or Hamiltonians on the potential
 * (their multiplier tables are made from V at creation). */
enum itp_impurity_kind { ITP_IMP_GAUSSIAN = 0, ITP_IMP_COULOMB, ITP_IMP_HEMISPHERE, ITP_IMP_DELTA };
int itp_potential_add_impurities(itp_potential* p, int kind, const double* data, size_t count,
                                 const double* kind_params);
/* The potential's values ([sy][sx]; zeros for the null potential). */
int itp_potential_values(const itp_potential* p, double* V);

/* ---- state set: StateSet storage (stateset.cpp:26-48); ITP_FP64 or ITP_FP32
 * fields. Host buffers are [N][sy][sx] interleaved complex double. ---- */
int itp_stateset_create(itp_grid* g, size_t N, int precision, itp_stateset** out);
void itp_stateset_destroy(itp_stateset* s);
int itp_stateset_upload(itp_stateset* s, const double* interleaved);
int itp_stateset_download(itp_stateset* s, double* interleaved);
/* Device pointer of the current field buffer (changes after orthonormalize and propagate). */
void* itp_stateset_device_ptr(itp_stateset* s);
size_t itp_stateset_size(const itp_stateset* s);

/* ---- evolution operator: MultiProductSplit (multiproductsplit.cpp:23-58) built
 * from SecondOrderSplit members (secondordersplit.cpp:21-69), ExpKinetic
 * (expkinetic.cpp:33-197) and ExpPotential (exppotential.cpp:28-63). ---- */
int itp_split_create(itp_potential* pot, int halforder, double eps, double B, itp_split** out);
void itp_split_destroy(itp_split* sp);
/* EvolutionOperator::set_time_step (operators.hpp:65-69; multiproductsplit.cpp:43-47) */
int itp_split_set_time_step(itp_split* sp, double eps);
/* Effective half order (1 for a null potential, multiproductsplit.cpp:24). */
int itp_split_halforder(const itp_split* sp);
/* ITPSystem::propagate: T applied to every state (itpsystem.cpp:243-256);
 * Operator::operate (operators.hpp:33-52) batched over the whole set. */
int itp_propagate(itp_split* sp, itp_stateset* s);

/* itp_propagate without any allocation: no counterpart in the reference. For
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc). The intermediate
 * fields live in a fixed device scratch array of
 * itp_propagate_noalloc_max_bytes() bytes (64 MiB) that each device holds for
 * the life of the process; a set needing more (N * sx * sy * 16 bytes for
 * ITP_FP64, 8 for ITP_FP32) is rejected with ITP_E_UNSUPPORTED before any
 * work is queued. Results equal itp_propagate's. Every call on one device
 * shares that device's array: calls that overlap on one device (two streams,
 * two host threads) race. */
int itp_propagate_noalloc(itp_split* sp, itp_stateset* s);
size_t itp_propagate_noalloc_max_bytes(void);

/* ---- StateSet::orthonormalize (stateset.cpp:136-232): overlap matrix,
 * Hermitian eigendecomposition, eigenvalue checks, Psi <- Psi U Lambda^(-1/2).
 * On ITP_E_NONPOSITIVE_EIG / ITP_E_NONNORMAL_EIG the states are left as they
 * were and *bad_n, *bad_lambda (nullable) name the offending eigenvalue in
 * ascending order, like NonPositiveEigenvalue(n, eval, ...). ---- */
int itp_orthonormalize(itp_stateset* s, size_t* bad_n, double* bad_lambda);
/* Cholesky orthonormalization Psi <- Psi R^-1 with S = R^H R (no counterpart
 * in the reference, which goes through the eigendecomposition of S,
 * stateset.cpp:136-232): the same span as itp_orthonormalize, another basis
 * (the CholeskyQR step of the mixed path; the states are not sorted by the
 * overlap eigenvalues). When the factorization meets a pivot below 1e-5
 * (ITP_FP32 sets) or 1e-12 (ITP_FP64) of its diagonal, the eigendecomposition
 * route of itp_orthonormalize runs instead (*used_eigen = 1, nullable) with
 * its error codes. */
int itp_orthonormalize_cholesky(itp_stateset* s, int* used_eigen, size_t* bad_n, double* bad_lambda);
/* Condition estimate of the overlap matrix of the last
 * itp_orthonormalize_cholesky: max R_ii^2 / min R_ii^2 of its factor (0 if it
 * fell back to the eigendecomposition or none ran). */
int itp_stateset_cholesky_condition(const itp_stateset* s, double* kappa);
/* R_ii^2 (i < N) of the last itp_orthonormalize_cholesky's factor, in the
 * states' order: the squared norms of the successive Gram-Schmidt residuals.
 * ITP_E_INVALID if the last call fell back to the eigendecomposition. */
int itp_stateset_cholesky_diagonal(const itp_stateset* s, double* d);
/* FP64 Rayleigh quotients E[n] = Re<psi_n|H psi_n> / <psi_n|psi_n> of the
 * states of src (either precision), evaluated in FP64 on a copy made into dst
 * (an ITP_FP64 set of the same grid and size, whose fields are overwritten with
 * src's). No counterpart in the reference: the mixed path's check of its FP32
 * states against FP64 energies. */
int itp_rayleigh_quotients(itp_hamiltonian* h, const itp_stateset* src, itp_stateset* dst, double* E);
/* Overlap matrix S_ij = <psi_i|psi_j> as computed by itp_orthonormalize
 * (stateset.cpp:145-154); out is N*N complex, column-major, full Hermitian. */
int itp_overlap(itp_stateset* s, double* out);

/* ---- Hamiltonian H = Kinetic + Potential (itpsystem.cpp:117-119;
 * kinetic.cpp:25-142) and Operator::mean_and_standard_deviation
 * (operators.cpp:40-50) for every state, H applied twice literally.
 * E[n] = Re<psi_n|H psi_n>, sigma[n] = Re sqrt(<psi_n|H H psi_n> - <H>^2), by
 * slot n (unsorted, like the tuples before the sort at itpsystem.cpp:496). ---- */
int itp_hamiltonian_create(itp_potential* pot, double B, itp_hamiltonian** out);
void itp_hamiltonian_destroy(itp_hamiltonian* h);
int itp_energies(itp_hamiltonian* h, itp_stateset* s, double* E, double* sigma);
/* Complex <H> and <H^2> per slot (interleaved re, im; each nullable). */
int itp_energies_complex(itp_hamiltonian* h, itp_stateset* s, double* mean, double* meansq);

/* ---- mixed-precision path (no counterpart in the reference; it completes
 * the loop above for FP32 sets) ---- */
/* Rayleigh-Ritz: orthonormalize (as itp_orthonormalize), W = H Psi, the
 * H-matrix Psi^H W (split-BF16x3 MFMA for FP32 sets, FP64 MFMA for FP64 sets),
 * its FP64 eigendecomposition, Psi <- Psi U. E[n] are the ascending Ritz values
 * and sigma[n] = ||H psi_n - E_n psi_n|| the residual norms (the energy
 * deviation of operators.cpp:40-50 for an eigenpair estimate). Slot n of the
 * set then holds the Ritz vector of E[n]. E, sigma nullable. */
int itp_rayleigh_ritz(itp_hamiltonian* h, itp_stateset* s, double* E, double* sigma);
/* Fills the set with seeded uniform noise, the initial states of
 * StateSet::init_to_gaussian_noise (stateset.cpp:124-131) drawn as
 * itp_init_states draws them, generated on the device: element values are
 * identical to
 * itp_init_states'; the per-state normalisation sums in another order, so the
 * scale may differ in the last bits. Advances *rng_state past the 2 N M draws
 * used, as itp_init_states does. */
int itp_stateset_init_noise(itp_stateset* s, uint64_t* rng_state, double dx);
/* Copies the fields of src into dst (same grid and size), converting the
 * precision if the two differ. */
int itp_stateset_convert(itp_stateset* dst, const itp_stateset* src);
/* Ascending overlap eigenvalues of the set's last orthonormalization
 * (lambda_n of stateset.cpp:158-177); N doubles. */
int itp_stateset_overlap_eigenvalues(const itp_stateset* s, double* lam);

/* ---- single-operator hooks (Operator::operate on host fields, for parity
 * checks; fields are [n][sy][sx] interleaved complex double, in place) ---- */
/* ExpKinetic(e, B, tr, bt, coeff, prefactor) (expkinetic.hpp:38-39) */
int itp_apply_expkinetic_host(itp_grid* g, double B, double e, double coeff, double prefactor,
                              double* fields, size_t n);
/* ExpPotential(pot, e, coeff, prefactor) (exppotential.hpp:33) */
int itp_apply_exppotential_host(itp_potential* p, double e, double coeff, double prefactor,
                                double* fields, size_t n);
/* SecondOrderSplit(pot, e, B, tr, bt, prefactor, exponent) (secondordersplit.cpp:21) */
int itp_apply_secondorder_host(itp_potential* p, double B, double e, double prefactor, int exponent,
                               double* fields, size_t n);
/* H = T + V applied once (OperatorSum::operate, operatorsum.cpp:61-86) */
int itp_apply_hamiltonian_host(itp_hamiltonian* h, double* fields, size_t n);
/* EigenSolver::solve (eigensolver.hpp:80-84) on the GPU: A is n*n complex
 * column-major (upper triangle read); on return A holds eigenvectors in
 * columns and w the eigenvalues in ascending order, like zheev('V','U'). */
int itp_eigh_host(itp_ctx* ctx, size_t n, double* A, double* w);
/* Eigenvalues only (ascending, w[n]) of the same input, by Householder
 * tridiagonalization and bisection on the GPU: the route of the eigenvalue
 * check (itp_ctx_set_eig_check_min). Test hook; A is not changed. */
int itp_eigvalsh_host(itp_ctx* ctx, size_t n, const double* A, double* w);

/* Seeded initial states: re, im ~ U[-1,1] from a splitmix64 stream, each state
 * normalised (StateSet::init_to_gaussian_noise, stateset.cpp:124-131, with
 * uniform noise). Host-side setup, not part of the timed loop. */
int itp_init_states(uint64_t* rng_state, size_t sx, size_t sy, double dx, size_t N, double* out);

/* ---- timing of the last calls on the context's stream (milliseconds, from
 * HIP events recorded around each kernel group; 0 if not recorded). ---- */
enum itp_phase { ITP_PHASE_PROPAGATE = 0, ITP_PHASE_OVERLAP, ITP_PHASE_EIGENSOLVE,
                 ITP_PHASE_LINCOMB, ITP_PHASE_ENERGIES, ITP_NUM_PHASES };
int itp_ctx_enable_timing(itp_ctx* ctx, int enable);
/* Accumulated event time per phase since the last reset (ms). */
int itp_ctx_phase_times(itp_ctx* ctx, double* ms, long long* counts);
int itp_ctx_reset_phase_times(itp_ctx* ctx);
/* Eigensolver statistics since the last reset: calls and outer sweeps. */
int itp_ctx_eig_stats(itp_ctx* ctx, long long* calls, long long* sweeps);
/* Upper bound on the element sweeps spent on each block sub-problem per round
 * of the block Jacobi eigensolver (default 1). */
int itp_ctx_set_eig_inner(itp_ctx* ctx, int inner);
/* Rotation of FP32 sets (mixed path): 1 (default) split-BF16x3 products on the
 * bf16 matrix cores, 0 FP32 matrix cores; both accumulate in FP32. */
int itp_ctx_set_fp32_rotation(itp_ctx* ctx, int kind);
/* Workgroups of the cooperative eigensolver launch (default 0: one per 16
 * super-tile updates of a round, at least 64, at most what is co-resident). */
int itp_ctx_set_eig_grid(itp_ctx* ctx, int workgroups);
/* Polls of a block-Jacobi grid barrier before the wait gives up (default
 * 2^24). A wait that gives up makes every workgroup of the launch leave it, and
 * the call returns ITP_E_EIGENSOLVER with the states untouched. 0 makes every
 * workgroup but the last to arrive give up at once (used to test that path). */
int itp_ctx_set_eig_spin_limit(itp_ctx* ctx, unsigned long long polls);
/* Pivot threshold of itp_orthonormalize_cholesky relative to the diagonal
 * (0, the default: 1e-5 for ITP_FP32 sets, 1e-12 for ITP_FP64). */
int itp_ctx_set_cholesky_tol(itp_ctx* ctx, double tol);
/* Eigensolver of the overlap matrix (replaces zheev, eigensolver.hpp:80-84):
 * 0 (default) block Jacobi sweeps until every unconverged off-diagonal element
 * is below half of its diagonal gap, then the refinement iteration
 * X <- X + X E on n x n matrix products (falls back to plain Jacobi if it does
 * not converge); 1 block Jacobi only. */
int itp_ctx_set_eig_mode(itp_ctx* ctx, int mode);
/* Hand-off threshold of mode 0: the Jacobi sweeps stop once every unconverged
 * |a_ij| is below ratio * |a_ii - a_jj| (default 0.5). */
int itp_ctx_set_eig_handoff(itp_ctx* ctx, double ratio);
/* Calls finished by the refinement iteration and calls that fell back to Jacobi. */
int itp_ctx_oa_stats(itp_ctx* ctx, long long* refined, long long* fallbacks);
/* Cluster start of mode 0 (tried before any Jacobi sweep): the pairs with
 * |a_ij| > tau |a_ii - a_jj| are joined into clusters of at most 64 indices,
 * each diagonalised exactly, and the refinement iteration starts from those
 * rotations; with more than 8192 such pairs, a larger cluster or no
 * convergence the block-Jacobi route runs as before. tau in [0, 0.5); 0 turns
 * the cluster start off (default 0.1). No counterpart in the reference, which
 * calls zheev (eigensolver.hpp:80-84): same eigenpairs, another route. */
int itp_ctx_set_eig_cluster(itp_ctx* ctx, double tau);
/* Launch-shape autotuning (default off): with it on, the first overlap of each
 * shape (precision, grid points, states padded to 64) times the split-K chunk
 * counts of the Gram product (the formula's, which fills the last wave, and
 * 3/4, 1/2, 1/4 of it) and every later set of that shape uses the fastest.
 * The tuned shapes are the process's "wisdom" (shared by all its contexts and
 * used by those with autotuning on), exported and imported as text lines
 * "gram <bytes per real: 4|8> <grid points> <padded states> <chunks>" (the
 * counterpart of the reference's FFTW wisdom file, itp2d.cpp:85-93, 124-128).
 * Only the launch shape changes: a chunk count changes the order of the Gram
 * product's partial sums, so results agree to rounding, not bit for bit. */
int itp_ctx_set_autotune(itp_ctx* ctx, int on);
/* Writes the wisdom text into out (cap bytes, NUL-terminated, may be NULL) and
 * its length into *len. */
int itp_ctx_wisdom_export(const itp_ctx* ctx, char* out, size_t cap, size_t* len);
/* Adds the shapes of a wisdom text (as exported) to the process's wisdom; a
 * malformed line is rejected with ITP_E_INVALID and nothing is added. */
int itp_ctx_wisdom_import(itp_ctx* ctx, const char* text);
/* Shapes tuned by measurement in this context. */
int itp_ctx_tunings(const itp_ctx* ctx);
/* 1: the fixed launch sequences of a propagation (every sweep of the split
 * members), of an overlap (Gram product and reduction) and of a Cholesky
 * factorization are captured as HIP graphs on their second call with the same
 * objects and buffers and replayed afterwards; 0 (default): launched one by one
 * (the graphs kept so far are released). Results are the same. While any
 * context has graphs on, allocations, frees and device-wide synchronisations
 * of the library wait for captures in progress in other threads. */
int itp_ctx_set_graphs(itp_ctx* ctx, int on);
/* Overlap matrices of at least n states (default 1024) that a Cholesky
 * factorization does not show positive definite (a pivot below 1e-12 of its
 * diagonal) get their eigenvalues first, by Householder tridiagonalization and
 * bisection, and the reference's checks (stateset.cpp:164-173) run on those: a
 * rejected matrix returns ITP_E_NONPOSITIVE_EIG / ITP_E_NONNORMAL_EIG without
 * the eigenvectors being computed (the eigenvalues reported are the check's).
 * No counterpart in the reference, whose zheev computes both. */
int itp_ctx_set_eig_check_min(itp_ctx* ctx, int n);
/* Eigenvalue checks run and matrices they rejected. */
int itp_ctx_eig_check_stats(itp_ctx* ctx, long long* checks, long long* rejected);
/* out[4]: calls finished by the cluster start, clusters solved, largest
 * cluster, calls that fell back to the block-Jacobi route. */
int itp_ctx_cluster_stats(itp_ctx* ctx, long long* out);

#ifdef __cplusplus
}
#endif

#endif /* ITP2D_H */
