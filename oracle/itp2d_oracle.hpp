// FP64 CPU oracle for the itp2d imaginary-time-propagation loop.
//
// TEST INFRASTRUCTURE ONLY. This library is the checker: only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and
// never as the thing measured or shipped. The product path (itp2d_amd) does
// not link, load or call anything here.
//
// It restates, in plain C++ with its own FFT and its own Hermitian
// eigensolver, the algorithm of the reference's hot path:
//   grid                  datalayout.cpp:21-36, parameters.hpp:97
//   transforms            transformer.cpp:66-151, transformer.hpp:132-158
//   pointwise / BLAS-1    state.hpp:180-240
//   ExpPotential          exppotential.cpp:55-63, exppotential.hpp:56-65
//   ExpKinetic            expkinetic.cpp:65-197
//   Kinetic, Potential    kinetic.cpp:25-142, potential.hpp:52-58
//   OperatorSum/Product   operatorsum.cpp:61-86, operatorproduct.cpp:48-52
//   SecondOrderSplit      secondordersplit.cpp:21-69
//   MultiProductSplit     multiproductsplit.cpp:23-58
//   orthonormalize        stateset.cpp:136-232, eigensolver.hpp:57-84
//   energies              operators.cpp:40-50, itpsystem.cpp:473-511
//   ITP loop              itpsystem.cpp:202-233,291-372,407-470
//   convergence tests     convergence.cpp:75-200
// Each function below cites the lines it follows.
#pragma once
#include <cstddef>
#include <cstdint>

extern "C" {

// Transform kinds, same order as the reference's enum Transform
// (transformer.hpp:27-28).
enum orc_transform_kind {
  ORC_FFT = 0, ORC_iFFT, ORC_FFTx, ORC_iFFTx, ORC_FFTy, ORC_iFFTy,
  ORC_DST, ORC_iDST, ORC_DSTx, ORC_iDSTx, ORC_DSTy, ORC_iDSTy,
  ORC_DCT, ORC_iDCT, ORC_DCTx, ORC_iDCTx, ORC_DCTy, ORC_iDCTy
};

// 1D real-to-real kinds with FFTW's names (transformer.cpp:112-115).
enum orc_r2r_kind { ORC_RODFT10 = 0, ORC_RODFT01 = 1, ORC_REDFT10 = 2, ORC_REDFT01 = 3 };

enum orc_boundary { ORC_PERIODIC = 0, ORC_DIRICHLET = 1 };

// Convergence test kinds (convergence.cpp:28-49).
enum orc_test_kind { ORC_TEST_NONE = 0, ORC_TEST_ONESTEP, ORC_TEST_ABSCHANGE,
                     ORC_TEST_RELCHANGE, ORC_TEST_ABSSTDEV, ORC_TEST_RELSTDEV };

struct orc_test { int kind; double p0, p1; };

struct orc_params {
  size_t sx, sy;
  double lenx;             // grid length along x; dx = lenx/sx (parameters.hpp:97)
  int boundary;            // orc_boundary
  size_t N;                // states propagated
  size_t n_wanted;         // states needed to converge
  size_t ignore_lowest;
  int halforder;
  double B;
  const double* eps_values; size_t n_eps;
  double eps_divisor;
  int exhaust_eps;
  int max_steps;
  double min_time_step;
  int recover;
  uint64_t seed;           // initial-state noise seed (also used on recovery)
  orc_test timestep_test, final_test;
};

// Error codes shared with the product's C-ABI meaning.
enum { ORC_OK = 0, ORC_E_INVALID = 1, ORC_E_NONPOSITIVE_EIG = 2,
       ORC_E_NONNORMAL_EIG = 3, ORC_E_EIGENSOLVER = 4 };

int orc_num_threads(void);
void orc_set_num_threads(int n);

// 1D transforms on a strided complex vector (interleaved re,im doubles).
int orc_fft1d(double* data, size_t n, size_t stride, int sign);
int orc_r2r1d(double* data, size_t n, size_t stride, int kind);
// Direct O(n^2) evaluation of the FFTW definitions, for pinning the fast ones.
int orc_r2r1d_direct(double* data, size_t n, size_t stride, int kind);

// In-place 2D transform of one field [sy][sx] complex, like Transformer::transform.
int orc_transform(size_t sx, size_t sy, double* data, int kind);
// Grid tables: posx[sx], posy[sy], fft_kx[sx], fft_ky[sy], dsct_kx[sx], dsct_ky[sy],
// norms[6] = FFT, FFTx, FFTy, DSCT, DSCTx, DSCTy normalisation factors.
int orc_grid_tables(size_t sx, size_t sy, double dx, double* posx, double* posy,
                    double* fft_kx, double* fft_ky, double* dsct_kx, double* dsct_ky,
                    double* norms);

// Single-operator applications on nstates fields [nstates][sy][sx] (in place).
// V may be NULL (null potential).
int orc_apply_exppotential(size_t sx, size_t sy, double dx, const double* V, double e,
                           double coeff, double prefactor, double* states, size_t nstates);
int orc_apply_expkinetic(size_t sx, size_t sy, double dx, int boundary, double B, double e,
                         double coeff, double prefactor, double* states, size_t nstates);
int orc_apply_kinetic(size_t sx, size_t sy, double dx, int boundary, double B,
                      double* states, size_t nstates);
int orc_apply_hamiltonian(size_t sx, size_t sy, double dx, int boundary, double B,
                          const double* V, double* states, size_t nstates);
int orc_apply_secondorder(size_t sx, size_t sy, double dx, int boundary, double B,
                          const double* V, double e, double prefactor, int exponent,
                          double* states, size_t nstates);
int orc_propagate(size_t sx, size_t sy, double dx, int boundary, double B, const double* V,
                  int halforder, double eps, double* states, size_t nstates);
// out[t] = MPS coefficients c_t (multiproductsplit.cpp:49-58); returns effective halforder.
int orc_mps_coefficients(int halforder, double* out);

// Hermitian eigensolver (cyclic Jacobi). A: n x n complex, column-major, upper
// triangle read. On return A holds eigenvectors in columns (ascending eigenvalue
// order) and w the eigenvalues, like zheev('V','U').
int orc_eigh(size_t n, double* A, double* w);

// StateSet::orthonormalize on [N][sy][sx]; on failure *bad_n, *bad_lambda are set.
int orc_orthonormalize(size_t sx, size_t sy, double dx, double* states, size_t N,
                       size_t* bad_n, double* bad_lambda);
// Overlap matrix S (column-major, full Hermitian) only.
int orc_overlap(size_t sx, size_t sy, double dx, const double* states, size_t N, double* S);
// Per-state mean and standard deviation of H (unsorted, by slot): E[n], sigma[n];
// mean_c / meansq_c (nullable) receive the complex <H> and <H^2>.
int orc_energies(size_t sx, size_t sy, double dx, int boundary, double B, const double* V,
                 const double* states, size_t N, double* E, double* sigma,
                 double* mean_c, double* meansq_c);
// max |<psi_n|psi_k> - delta_nk| (stateset.cpp:250-269)
double orc_how_orthonormal(size_t sx, size_t sy, double dx, const double* states, size_t N);

// Seeded initial states: re, im ~ U[-1,1] from a splitmix64 stream, then each
// state normalised (stateset.cpp:124-131 with uniform instead of Gaussian noise).
int orc_init_states(uint64_t* rng_state, size_t sx, size_t sy, double dx, size_t N, double* out);

// The ITP driver.
void* orc_system_create(const orc_params* p, const double* V, const double* init_states);
void orc_system_destroy(void* sys);
int orc_system_step(void* sys);
// Runs until finished; returns number of steps done.
int orc_system_run(void* sys);
int orc_system_finished(void* sys);
int orc_system_error_flag(void* sys);
int orc_system_total_steps(void* sys);
double orc_system_eps(void* sys);
size_t orc_system_num_converged(void* sys);
size_t orc_system_num_timestep_converged(void* sys);
// sorted energies / deviations / slot indices of the latest step (length N)
int orc_system_sorted(void* sys, double* E, double* sigma, size_t* index);
// State pointwise operations (kind: 0 multiply, 1 imaginary multiply with the
// x shift, 2 multiply by a y-table, 3 multiply and add) and dot/norm
int orc_pointwise(int kind, size_t sx, size_t sy, double* state, const double* values, const double* addstate);
double orc_dot_norm(size_t M, double dx, const double* a, const double* b, double* out);
// final-convergence flag of each sorted position
int orc_system_converged(void* sys, int* flags);
int orc_system_states(void* sys, double* out);
double orc_system_total_time(void* sys);
int orc_system_history_len(void* sys);

}  // extern "C"
