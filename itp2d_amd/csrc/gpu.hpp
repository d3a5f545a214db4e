// Launchers of the gfx950 kernels (kernels.hip), called by the C ABI layer.
#pragma once
#include <hip/hip_runtime.h>
#include "line_ops.hpp"
#include "potentials.hpp"

namespace itp {

template <class R>
hipError_t launch_sweep(const SweepArgs<R>& a, int nblk, int nfields, hipStream_t st);
// one register-engine sweep of line length N (reg_sweep_kernels.hip, one translation unit
// per precision and length); the caller has checked the geometry
// Ascending eigenvalues lam[N] of the Hermitian np x np row-major A (device,
// both triangles, left unchanged) by Householder tridiagonalization and
// bisection (tridiag.hip), in scratch of hermitian_eigenvalues_scratch bytes;
// the uint32 at hermitian_eigenvalues_barrier_offset + 4 is non-zero after
// the launch if a grid barrier gave up (spin_limit polls).
size_t hermitian_eigenvalues_scratch(int np, int N);
size_t hermitian_eigenvalues_barrier_offset(int np, int N);
hipError_t launch_hermitian_eigenvalues(const cplx<double>* A, int np, int N, void* scratch, double* lam,
                                        unsigned spin_limit, hipStream_t st);
template <class R, int N>
hipError_t launch_reg_sweep_n(const SweepArgs<R>& a, int nblk, int nfields, hipStream_t st);
// p[0..n) = 0 by a kernel (usable inside captured HIP graphs)
hipError_t launch_zero_ints(int* p, int n, hipStream_t st);
// A = I (np x np, row-major complex FP64)
hipError_t launch_identity(cplx<double>* A, int np, hipStream_t st);
hipError_t launch_energy_reduce(const cplx<double>* p1, const cplx<double>* p2, int nblk, int N, double dx,
                                double* E, double* sigma, cplx<double>* mean, cplx<double>* msq, hipStream_t st);
// V (sx x sy, device) = the potential kind q on the cell-centred grid
hipError_t launch_potential(const PotParams& q, int sx, int sy, double dx, double* V, hipStream_t st);
// V += count impurities of a kind: data [count][stride] holds x, y and the
// per-impurity parameters; g0, g1 the kind's shared parameters
hipError_t launch_impurities(int kind, const double* data, int stride, long long count, double g0, double g1, int sx,
                             int sy, double dx, double* V, hipStream_t st);
// n complex elements: cplx<double> -> cplx<float> (to_f32) or back
hipError_t launch_convert(const void* in, void* out, long long n, int to_f32, hipStream_t st);
// N seeded noise states of M elements, each of dx-norm 1, from the splitmix64
// stream at s0 (draws 2 (n M + i) and 2 (n M + i) + 1); part holds
// N * noise_blocks(M) doubles
int noise_blocks(long long M);
template <class R>
hipError_t launch_noise(uint64_t s0, long long M, int N, double dx, double* part, cplx<R>* out, hipStream_t st);
// S = dx^2 Psi^H Phi from the upper 64 x 64 tile pairs, mirrored (Hermitian
// when Phi = Psi or Phi = H Psi with H Hermitian); with full, from all tile
// pairs, and S is the Hermitian part (G + G^H) / 2.
template <class R>
hipError_t launch_gram(const cplx<R>* psi, const cplx<R>* phi, long long M, int N, int nc, cplx<double>* part,
                       int nchunk, long long kchunk, double dx, cplx<double>* S, int np, int* bad, hipStream_t st,
                       int full = 0);
hipError_t launch_gram_reduce(const cplx<double>* part, int nchunk, int nc, int N, int np, double dx, cplx<double>* S,
                              int* bad, hipStream_t st, int full = 0, int acc = 1);
// mixed path (mixed.hip)
hipError_t launch_gram_bf16x3(const cplx<float>* A, const cplx<float>* B, long long M, int N, int nc,
                              cplx<double>* part, int nchunk, long long kchunk, double dx, cplx<double>* S, int np,
                              int* bad, hipStream_t st, int full = 0);
hipError_t launch_rotate_f32(const cplx<float>* in, cplx<float>* out, const cplx<float>* C, int nc, long long M, int N,
                             hipStream_t st);
// the same rotation with split-BF16x3 products on the bf16 matrix cores;
// upper: C is upper triangular (output state r takes inputs c <= r only)
hipError_t launch_rotate_bf16x3(const cplx<float>* in, cplx<float>* out, const cplx<float>* C, int nc, long long M,
                                int N, hipStream_t st, int upper = 0);
template <class R>
hipError_t launch_residual(const cplx<R>* W, const cplx<R>* Psi, const double* E, long long M, int N, double dx,
                           double* sigma, hipStream_t st);
// grid_wgs <= 0 sizes the cooperative grid by the matrix
hipError_t launch_block_jacobi(cplx<double>* A, cplx<double>* V, cplx<double>* Q, cplx<double>* D, int* ident,
                               double* slots, int np, int max_sweeps, int max_inner, double tol, double handoff,
                               double* info, int grid_wgs, unsigned spin_limit, hipStream_t st);
// Refinement of the eigenvector estimate X0 of A0: X0 is overwritten only on
// convergence (iterates go to X1 and to work's fifth np x np slot); work holds
// 5 np^2 complex values. x0_identity: non-null (an identity matrix) when X0
// is the identity, which spares the first iteration's products. part:
// 4 np^2 complex values for split-K products of small matrices (nullptr:
// no split)
hipError_t launch_oa(const cplx<double>* A0, cplx<double>* X0, cplx<double>* X1, cplx<double>* work, double* lam,
                     double* slots, int* state, const double* info, int np, int N, int max_iter, cplx<double>* Aout,
                     hipStream_t st, const cplx<double>* x0_identity = nullptr, cplx<double>* part = nullptr);
// Strong pairs (i < j < N, |a_ij| > tau |a_ii - a_jj|) of the np x np matrix
// A: hdr[0] counts them (the first cap are listed in pairs as (i, j)),
// hdr[1] holds the bits of the largest ratio
hipError_t launch_cluster_pairs(const cplx<double>* A, int np, int N, double tau, int cap, unsigned long long* hdr,
                                int* pairs, hipStream_t st);
// Cluster c (members[offsets[c] .. offsets[c+1]), 2..kMaxCluster indices):
// X[:, members] <- X[:, members] U_c with U_c the eigenvectors of A's
// principal sub-matrix on the members (clusters are disjoint)
hipError_t launch_cluster_solve(const cplx<double>* A, int np, const int* members, const int* offsets, int nclusters,
                                cplx<double>* X, hipStream_t st);
hipError_t launch_eig_finish(const cplx<double>* A, const cplx<double>* V, int np, int N, double* lam_sorted,
                             int* perm, cplx<double>* C, int nc, int scale, hipStream_t st);
// Cholesky S = R^H R of the np x np overlap (chol.hip): R (np x np scratch)
// and W = R^-1 into the nc x nc coefficient matrix; *fail set (and the rest
// skipped) on a pivot below tol times its original diagonal. diag0: np doubles.
hipError_t launch_cholinv(const cplx<double>* S, int np, int N, cplx<double>* R, double* diag0, cplx<double>* W,
                          int nc, int* fail, double tol, hipStream_t st);
// mm[0], mm[1] = largest and smallest R_ii^2 (i < N) of the factor R, then
// mm[2 + i] = R_ii^2
hipError_t launch_chol_diag_range(const cplx<double>* R, int np, int N, double* mm, hipStream_t st);
// out[n] = sum_i |f_n[i]|^2 for N fields of M elements
hipError_t launch_field_norms(const cplx<double>* f, long long M, int N, double* out, hipStream_t st);
template <class R>
hipError_t launch_rotate(const cplx<R>* in, cplx<R>* out, const cplx<double>* C, int nc, long long M, int N,
                         hipStream_t st);

// Runs engine sweeps on the GPU (the Backend of run_program).
struct GpuBackend {
  hipStream_t st = nullptr;
  hipError_t err = hipSuccess;
  template <class R>
  void sweep(const SweepArgs<R>& a, int nblk, int nfields) {
    if (err == hipSuccess) err = launch_sweep<R>(a, nblk, nfields, st);
  }
};

}  // namespace itp
