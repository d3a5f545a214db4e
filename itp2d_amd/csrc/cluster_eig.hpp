// Start vectors for the refinement iteration from small exact
// eigen-decompositions (the reference diagonalises the overlap matrix with
// LAPACK zheev, eigensolver.hpp:80-84; this is one route to the same
// eigenpairs).
//
// Close to convergence the overlap matrix is nearly diagonal: the refinement
// iteration (first-order corrections E_ij = a_ij / (a_jj - a_ii)) converges
// from the identity unless some pairs couple strongly relative to their gap.
// Those pairs are found (rho_ij = |a_ij| / |a_ii - a_jj| > tau), joined into
// clusters (connected components), and each cluster's principal sub-matrix is
// diagonalised exactly by cyclic two-sided Jacobi in one workgroup. The
// cluster rotations, embedded in the identity, are the refinement's start:
// what couples strongly is rotated exactly, what is left is weak.
//
// Shared by the gfx950 kernels and the CPU emulation.
#pragma once

#include "block_jacobi.hpp"
#include "dense_ops.hpp"

namespace itp {

constexpr int kMaxCluster = 48;  // states per cluster solved in one workgroup

// LDS of one cluster solve: M, M', U, U' (kMaxCluster^2 complex each), the
// rotation parameters and the reduction buffer.
ITP_HD size_t cl_lds_bytes(int nthreads) {
  return (size_t)4 * kMaxCluster * kMaxCluster * sizeof(cplx<double>) +
         kMaxCluster / 2 * (sizeof(double) + sizeof(cplx<double>)) + (size_t)(nthreads + 2) * sizeof(double) + 64;
}

// Strong pair test of the upper triangle, row i (i < N): every j in (i, N)
// with |a_ij| > tau |a_ii - a_jj| is reported through push(i, j).
// Returns the largest rho of the row.
template <class Push>
ITP_HD double cl_row_pairs(const cplx<double>* A, int np, int N, int i, int j0, int jstep, double tau, Push&& push) {
  double mx = 0;
  const double aii = A[(long long)i * np + i].x;
  for (int j = j0; j < N; j += jstep) {
    if (j <= i) continue;
    const cplx<double> b = A[(long long)i * np + j];
    const double ab = cabs_fast(b);
    if (ab == 0) continue;
    const double gap = std::fabs(aii - A[(long long)j * np + j].x);
    const double rho = gap > 0 ? ab / gap : 1e300;
    mx = rho > mx ? rho : mx;
    if (rho > tau) push(i, j);
  }
  return mx;
}

// Exact eigen-decomposition U of the principal sub-matrix of A on the k
// indices idx[0..k) (k <= kMaxCluster), applied to the np x np matrix X on
// those columns: X[:, idx] <- X[:, idx] U (with X the identity this writes U
// as the k x k block at rows and columns idx, eigenvector b in column b).
// Cyclic two-sided Jacobi on the size padded to even (a zero pad row and
// column need no rotation), round-robin pairings of dense_ops.hpp; stops
// once every off-diagonal element is below tol sqrt(|a_ii a_jj|) or
// k eps max|a_ii|, or after max_sweeps sweeps. Clusters have disjoint
// columns, so clusters update X concurrently without overlap.
template <class Exec>
ITP_HD void cl_solve(Exec& ex, const cplx<double>* A, int np, const int* idx, int k, cplx<double>* X, double tol,
                     int max_sweeps, void* lds, int nrows) {
  const int k2 = k + (k & 1), m = k2 / 2, kk = k2 * k2;
  cplx<double>* M[2];
  cplx<double>* U[2];
  cplx<double>* p = reinterpret_cast<cplx<double>*>(lds);
  M[0] = p; p += kMaxCluster * kMaxCluster;
  M[1] = p; p += kMaxCluster * kMaxCluster;
  U[0] = p; p += kMaxCluster * kMaxCluster;
  U[1] = p; p += kMaxCluster * kMaxCluster;
  cplx<double>* sj = p; p += kMaxCluster / 2;
  double* cs = reinterpret_cast<double*>(p);
  double* red = cs + kMaxCluster / 2;
  ex.run([&](int tid, int nt) {
    for (int e = tid; e < kk; e += nt) {
      const int a = e / k2, b = e - a * k2;
      M[0][e] = (a < k && b < k) ? A[(long long)idx[a] * np + idx[b]] : mk<double>(0.0, 0.0);
      U[0][e] = mk<double>(a == b ? 1.0 : 0.0, 0.0);
    }
  });
  int nt0 = 0;
  ex.run([&](int tid, int nt) {
    (void)tid;
    nt0 = nt;
  });
  int cur = 0;
  for (int sw = 0; sw <= max_sweeps; sw++) {
    // convergence of the whole block (positions are in their original order
    // between sweeps): every off-diagonal ratio below 1, with the absolute
    // floor k eps max|a_ii|
    const cplx<double>* Mc = M[cur];
    ex.run([&](int tid, int nt) {
      double dmax = 0;
      for (int a = tid; a < k2; a += nt) {
        const double d = std::fabs(Mc[a * k2 + a].x);
        dmax = d > dmax ? d : dmax;
      }
      red[tid] = dmax;
    });
    bj_tree_max(ex, red, red + nt0);
    double floor_ = 0;
    ex.run([&](int tid, int nt) {
      (void)tid;
      (void)nt;
      floor_ = (double)k * 2.220446049250313e-16 * red[nt0];
    });
    ex.run([&](int tid, int nt) {
      double r = 0;
      for (int e = tid; e < kk; e += nt) {
        const int a = e / k2, b = e - a * k2;
        if (b > a) {
          const double q = bj_ratio(Mc[e], Mc[a * k2 + a].x, Mc[b * k2 + b].x, tol, floor_);
          r = q > r ? q : r;
        }
      }
      red[tid] = r;
    });
    bj_tree_max(ex, red, red + nt0);
    bool done = false;
    ex.run([&](int tid, int nt) {
      (void)tid;
      (void)nt;
      done = red[nt0] < 1.0;
    });
    if (done || sw == max_sweeps) break;
    for (int r = 0; r < k2 - 1; r++) {
      const cplx<double>* Ma = M[cur];
      cplx<double>* Mb = M[cur ^ 1];
      const cplx<double>* Ua = U[cur];
      cplx<double>* Ub = U[cur ^ 1];
      ex.run([&](int tid, int nt) {
        (void)nt;
        if (tid < m) {
          const int p0 = 2 * tid, p1 = p0 + 1;
          double c, off;
          cplx<double> s;
          jacobi_rotation(Ma[p0 * k2 + p0].x, Ma[p1 * k2 + p1].x, Ma[p0 * k2 + p1], c, s, off);
          cs[tid] = c;
          sj[tid] = s;
        }
      });
      ex.run([&](int tid, int nt) {
        const int nblk = m * m;
        for (int it = tid; it < nblk + k2 * m; it += nt) {
          if (it < nblk) {
            const int kI = it / m, kJ = it - kI * m;
            if (kJ >= kI) jacobi_block(Ma, Mb, k2, m, kI, kJ, cs, sj);
          } else {
            const int i2 = it - nblk;
            jacobi_vrow(Ua, Ub, k2, m, i2 / m, i2 % m, cs, sj);
          }
        }
      });
      cur ^= 1;
    }
  }
  // X[:, idx] <- X[:, idx] U over the first nrows rows, in chunks of
  // kClRows rows staged in the now free M buffers
  const cplx<double>* Uf = U[cur];
  cplx<double>* Xs = M[0];
  constexpr int kClRows = 64;
  for (int r0 = 0; r0 < nrows; r0 += kClRows) {
    const int rn = nrows - r0 < kClRows ? nrows - r0 : kClRows;
    ex.run([&](int tid, int nt) {
      for (int e = tid; e < rn * k; e += nt) {
        const int r = e / k, a = e - r * k;
        Xs[e] = X[(long long)(r0 + r) * np + idx[a]];
      }
    });
    ex.run([&](int tid, int nt) {
      for (int e = tid; e < rn * k; e += nt) {
        const int r = e / k, b = e - r * k;
        cplx<double> acc = mk<double>(0.0, 0.0);
        for (int a = 0; a < k; a++) acc = acc + Xs[r * k + a] * Uf[a * k2 + b];
        X[(long long)(r0 + r) * np + idx[b]] = acc;
      }
    });
  }
}

}  // namespace itp
