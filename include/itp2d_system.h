/*
 * itp2d_system.h -- C ABI of the ITP driver (the reference's ITPSystem,
 * itpsystem.hpp:55-156) running on the gfx950 engine of itp2d.h.
 *
 * The loop semantics follow itpsystem.cpp:202-511 and convergence.cpp:75-200:
 * sorted (E, sigma, slot) tuples, sticky time-step convergence flags, the eps
 * schedule (list, then division), the max-steps and min-eps bail-outs and the
 * optional recovery after a non-positive overlap eigenvalue.
 */
#ifndef ITP2D_SYSTEM_H
#define ITP2D_SYSTEM_H

#include <stddef.h>
#include <stdint.h>

#include "itp2d.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ConvergenceTest kinds (convergence.cpp:28-49) */
enum itps_test_kind { ITPS_TEST_NONE = 0, ITPS_TEST_ONESTEP, ITPS_TEST_ABSCHANGE,
                      ITPS_TEST_RELCHANGE, ITPS_TEST_ABSSTDEV, ITPS_TEST_RELSTDEV };

typedef struct itps_test { int kind; double p0, p1; } itps_test;

/* The Parameters fields the loop consumes (parameters.hpp, defaults at
 * parameters.cpp:25-55). dx = lenx/sx (parameters.hpp:97). */
typedef struct itps_params {
  size_t sx, sy;
  double lenx;
  int boundary;
  size_t N;
  size_t n_wanted;
  size_t ignore_lowest;
  int halforder;
  double B;
  const double* eps_values; size_t n_eps;
  double eps_divisor;
  int exhaust_eps;
  int max_steps;
  double min_time_step;
  int recover;
  uint64_t seed;            /* initial-state noise seed, continued on recovery */
  itps_test timestep_test, final_test;
  int precision;            /* ITP_FP64: the reference algorithm in FP64.
                               ITP_FP32: the mixed path (no counterpart in the
                               reference). FP32 phase: FP32 fields and sweeps,
                               split-BF16x3 overlap, orthonormalization by
                               CholeskyQR (FP64 Cholesky factor and inverse on
                               the GPU, split-BF16x3 rotation), and every
                               checked step (mixed_check_every) a Rayleigh-Ritz step
                               (the overlap eigendecomposition, as the
                               reference orthonormalizes) followed by the FP64
                               Rayleigh quotients of the states
                               (itp_rayleigh_quotients). The FP32 phase ends
                               when those energies change by less than
                               mixed_tol (relative, per step) for every wanted
                               state between consecutive checks, or stop improving for
                               3 checks (the FP32 floor). Then the reference
                               algorithm continues in FP64 on the checked FP64
                               copy until the configured tests pass (the
                               reference's energies belong to the fixed point of
                               the split propagator, so parity needs that fixed
                               point). Uses the first eps for the FP32 phase. */
  double mixed_tol;         /* 0: the time-step test's tolerance if it is
                               relchange, else 1e-9 */
  double accuracy;          /* error bound of itps_rayleigh_ritz; 0: 1e-8 */
  int mixed_check_every;    /* FP32 steps per check of the mixed path; 0:
                               adaptive (first check once the log-determinant of the
                               leading overlap from the CholeskyQR factors
                               changes by < 1e-5 per step or stops falling,
                               not before 16 steps and at the latest after
                               64; then the step count predicted from the
                               decay of the energy changes, within 2..16) */
  int mixed_skip_ortho;     /* mixed path: 1 (default) lets CholeskyQR steps
                               be skipped when mixed_tol >= 1e-10, while the
                               last factor's condition per propagation step,
                               raised to the number of propagations sharing
                               one orthonormalization, stays below 1e3 (the
                               span is the same); -1 orthonormalizes every
                               step */
} itps_params;

typedef struct itps_system itps_system;

/* V may be NULL (null potential); init_states ([N][sy][sx] complex double) may
 * be NULL, in which case seeded noise is used (itp_stateset_init_noise, the
 * values of itp_init_states generated on the device). */
int itps_create(itp_ctx* ctx, const itps_params* p, const double* V, const double* init_states, itps_system** out);
void itps_destroy(itps_system* s);
/* ITPSystem::step (itpsystem.cpp:407-470) */
int itps_step(itps_system* s);
/* while (!finished) step(); returns the status of the last failing call or ITP_OK */
int itps_run(itps_system* s);
int itps_finished(const itps_system* s);
int itps_error_flag(const itps_system* s);
int itps_total_steps(const itps_system* s);
double itps_eps(const itps_system* s);
size_t itps_num_converged(const itps_system* s);
size_t itps_num_timestep_converged(const itps_system* s);
/* Sorted energies, deviations and slot indices of the latest step (length N each, nullable). */
int itps_sorted(const itps_system* s, double* E, double* sigma, size_t* index);
/* Final states, by slot ([N][sy][sx] complex double). */
int itps_states(itps_system* s, double* out);
/* Wall time from the first step to finish (s), as the reference's total timer (itpsystem.cpp:416-421,533). */
double itps_total_time(const itps_system* s);
/* Number of orthonormalization failures recovered from. */
int itps_recoveries(const itps_system* s);
/* Steps of the mixed path's FP32 phase (included in itps_total_steps). */
int itps_mixed_steps(const itps_system* s);
/* out[4]: FP32-phase steps, FP64 checks made in it, CholeskyQR steps that fell
 * back to the eigendecomposition route, steps whose orthonormalization was
 * skipped. */
int itps_mixed_stats(const itps_system* s, int* out);
/* Abort flag polled at the start of every step (the reference's
 * abort_flagptr, itpsystem.cpp:407-416, raised by its SIGINT handler,
 * itp2d.cpp:35-60): when *flag is non-zero the run finishes with the error
 * flag set. NULL removes it. The pointer must stay valid while steps run. */
int itps_set_abort_flag(itps_system* s, const volatile int* flag);
/* Save flag checked after propagation, orthonormalization and the energies of
 * every step and at its end (the reference's save_flagptr and
 * check_save_flag, itpsystem.cpp:235-241, 437-469, raised by its SIGUSR1
 * handler, itp2d.cpp:57-60): when *flag is non-zero, fn(user) runs (the
 * caller saves the states, e.g. with itps_states and itps_sorted, as the
 * reference's save_states writes them sorted by the last energies,
 * itpsystem.cpp:374-388) and *flag is cleared. NULL flag removes it. */
typedef void (*itps_save_fn)(void* user);
int itps_set_save_flag(itps_system* s, volatile int* flag, itps_save_fn fn, void* user);
/* The reference program's signal handlers (itp2d.cpp:40-66): SIGINT raises
 * *abort_flag (a second SIGINT ends the process at once) and SIGUSR1 raises
 * *save_flag, each with the reference's note on stderr. For command-line
 * drivers; the flags must stay valid for the life of the process. */
int itps_install_signal_flags(volatile int* abort_flag, volatile int* save_flag);
/* Progress printout on stdout/stderr as the reference's verbosity levels
 * (0 quiet, the default; 1: start, epsilon changes and final message;
 * 2: per-step lines, phase ratios and final energies; 3: energies every
 * step), itpsystem.cpp:158-200,405-470,541-589. */
int itps_set_verbosity(itps_system* s, int level);
/* Final-convergence flag of each sorted position (N ints). */
int itps_converged(const itps_system* s, int* flags);
/* Rayleigh-Ritz pass on the current states (itp_rayleigh_ritz): replaces the
 * sorted energies and deviations by the Ritz values and residual norms, and the
 * convergence flags by the error bound sigma^2 / gap <= accuracy |E| (Ritz
 * values within 1e-7 relative form a cluster: summed squared residuals, gap to
 * the other Ritz values). These are the exact discrete eigenvalues of the
 * Hermitian part of H in the states' span, free of the split propagator's
 * O(eps^(2h)) bias that the loop's own energies carry. No counterpart in the
 * reference. */
int itps_rayleigh_ritz(itps_system* s);

/* Step history for result files (the data the reference writes to its
 * datafile during the run, itpsystem.cpp:113-114, 221-226, 399-403):
 * itps_history_steps gives the number of rows (the last step that computed
 * energies); row k holds the sorted energies and deviations of step k + 1
 * (NaN for a step that computed none, e.g. one ended by a recovery), N
 * values each. itps_time_step_history returns the number of time-step
 * changes and fills up to cap (first step using the value, value) pairs; the
 * first pair is the initial time step at step 1. */
size_t itps_history_steps(const itps_system* s);
int itps_history_row(const itps_system* s, size_t step, double* E, double* sigma);
size_t itps_time_step_history(const itps_system* s, int* steps, double* eps, size_t cap);

#ifdef __cplusplus
}
#endif

#endif /* ITP2D_SYSTEM_H */
