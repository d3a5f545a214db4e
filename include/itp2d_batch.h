/*
 * itp2d_batch.h -- C ABI for many independent ITP problems on one GPU.
 *
 * The reference solves one system per process (ITPSystem, itpsystem.hpp:55-156,
 * driven by itp2d.cpp); a batch of instances (different grids, state counts,
 * field strengths and potentials) is the same loop run once per instance.
 * Here the instances of a batch are shared out to worker host threads, each
 * owning one itp_ctx (one HIP stream), so the instances' kernels overlap on
 * the GPU. Every instance runs exactly the loop of itp2d_system.h, with the
 * same results as when run alone: the kernels are deterministic and do not
 * depend on what else runs.
 *
 * The block-Jacobi eigensolver is a cooperative launch whose workgroups wait
 * for each other; the library lets one such launch run at a time per process,
 * so concurrent instances cannot hold parts of two of them resident.
 */
#ifndef ITP2D_BATCH_H
#define ITP2D_BATCH_H

#include <stddef.h>

#include "itp2d_system.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct itp_batch itp_batch;

/* A batch on GPU `device` run by `workers` host threads (1..64), each with its
 * own context and stream. */
int itp_batch_create(int device, int workers, itp_batch** out);
void itp_batch_destroy(itp_batch* b);
/* Launch-shape autotuning (itp_ctx_set_autotune) in every worker context of
 * the following runs (default off). */
int itp_batch_set_autotune(itp_batch* b, int on);
/* Adds an instance: p (its eps_values included) and V ([sy][sx], NULL for the
 * null potential) are copied. *index receives the instance's position. */
int itp_batch_add(itp_batch* b, const itps_params* p, const double* V, size_t* index);
size_t itp_batch_size(const itp_batch* b);
/* Runs every instance to the end of its loop (itps_run). Instances are
 * started largest first (N sx sy). Returns ITP_OK, or the first non-OK status
 * of an instance (each instance's own status is in itp_batch_result); the
 * other instances still run. */
int itp_batch_run(itp_batch* b);
/* Wall time of the last itp_batch_run (s). */
double itp_batch_wall_time(const itp_batch* b);
/* Outcome of instance i: status of its run, ITP iterations, finally
 * converged states, error flag and its own wall time (first step to finish).
 * Any pointer may be NULL. */
int itp_batch_result(const itp_batch* b, size_t i, int* status, int* steps, size_t* converged, int* error_flag,
                     double* time_s);
/* itps_mixed_stats of instance i (zeros on the FP64 path). */
int itp_batch_mixed_stats(const itp_batch* b, size_t i, int* out);
/* Sorted energies, deviations and final-convergence flags of instance i
 * (N values each, nullable), as itps_sorted / itps_converged. */
int itp_batch_sorted(const itp_batch* b, size_t i, double* E, double* sigma, int* converged_flags);

#ifdef __cplusplus
}
#endif

#endif /* ITP2D_BATCH_H */
