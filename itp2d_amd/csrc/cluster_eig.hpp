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

constexpr int kMaxCluster = 64;  // states per cluster solved in one workgroup

// LDS of one cluster solve: M and U (kMaxCluster^2 complex each, rotated in
// place), the rotation parameters and pairing of a round, the reduction
// buffer.
ITP_HD size_t cl_lds_bytes(int nthreads) {
  return (size_t)2 * kMaxCluster * kMaxCluster * sizeof(cplx<double>) +
         kMaxCluster / 2 * (sizeof(double) + sizeof(cplx<double>) + 2 * sizeof(int)) +
         (size_t)(nthreads + 2) * sizeof(double) + 64;
}

// Pair t of round r of the round-robin (polygon) schedule on k2 indices:
// index k2-1 meets r, and (r + t, r - t) mod (k2 - 1) for t = 1..k2/2-1.
// Every pair meets once in the k2-1 rounds of a sweep.
ITP_HD void cl_pair(int r, int t, int k2, int& p, int& q) {
  const int n1 = k2 - 1;
  if (t == 0) {
    p = r;
    q = n1;
  } else {
    p = (r + t) % n1;
    q = (r - t + n1) % n1;
  }
  if (p > q) {
    const int x = p;
    p = q;
    q = x;
  }
}

// In-place two-sided update of the 2 x 2 block (rows pa, qa; columns pb, qb)
// of the Hermitian M by rotations Ja (rows) and Jb (columns):
// Y = Ja^H M[Pa, Pb] Jb, and its conjugate transpose at the mirror.
ITP_HD void cl_block(cplx<double>* M, int k2, int pa, int qa, double ca, cplx<double> sa, int pb, int qb, double cb,
                     cplx<double> sb, bool diag) {
  const cplx<double> x00 = M[pa * k2 + pb], x01 = M[pa * k2 + qb];
  const cplx<double> x10 = M[qa * k2 + pb], x11 = M[qa * k2 + qb];
  const cplx<double> z00 = x00 * cb - conj(sb) * x01;
  const cplx<double> z01 = sb * x00 + x01 * cb;
  const cplx<double> z10 = x10 * cb - conj(sb) * x11;
  const cplx<double> z11 = sb * x10 + x11 * cb;
  cplx<double> y00 = z00 * ca - sa * z10;
  cplx<double> y01 = z01 * ca - sa * z11;
  cplx<double> y10 = conj(sa) * z00 + z10 * ca;
  cplx<double> y11 = conj(sa) * z01 + z11 * ca;
  if (diag) {  // the pair's own block: real diagonal, the removed element zero
    y00.y = 0;
    y11.y = 0;
    if (sa.x != 0 || sa.y != 0) {
      y01 = mk<double>(0.0, 0.0);
      y10 = y01;
    } else {
      y10 = conj(y01);
    }
    M[pa * k2 + pa] = y00;
    M[pa * k2 + qa] = y01;
    M[qa * k2 + pa] = y10;
    M[qa * k2 + qa] = y11;
    return;
  }
  M[pa * k2 + pb] = y00;
  M[pa * k2 + qb] = y01;
  M[qa * k2 + pb] = y10;
  M[qa * k2 + qb] = y11;
  M[pb * k2 + pa] = conj(y00);
  M[qb * k2 + pa] = conj(y01);
  M[pb * k2 + qa] = conj(y10);
  M[qb * k2 + qa] = conj(y11);
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
// column need no rotation), rotated in place with the polygon schedule of
// cl_pair; stops once every off-diagonal element is below
// tol sqrt(|a_ii a_jj|) or k eps max|a_ii|, or after max_sweeps sweeps.
// Clusters have disjoint columns, so clusters update X concurrently without
// overlap.
template <class Exec>
ITP_HD void cl_solve(Exec& ex, const cplx<double>* A, int np, const int* idx, int k, cplx<double>* X, double tol,
                     int max_sweeps, void* lds, int nrows) {
  const int k2 = k + (k & 1), m = k2 / 2, kk = k2 * k2;
  cplx<double>* M = reinterpret_cast<cplx<double>*>(lds);
  cplx<double>* U = M + kMaxCluster * kMaxCluster;
  cplx<double>* sj = U + kMaxCluster * kMaxCluster;
  double* cs = reinterpret_cast<double*>(sj + kMaxCluster / 2);
  int* pp = reinterpret_cast<int*>(cs + kMaxCluster / 2);
  int* pq = pp + kMaxCluster / 2;
  double* red = reinterpret_cast<double*>(pq + kMaxCluster / 2);
  ex.run([&](int tid, int nt) {
    for (int e = tid; e < kk; e += nt) {
      const int a = e / k2, b = e - a * k2;
      M[e] = (a < k && b < k) ? A[(long long)idx[a] * np + idx[b]] : mk<double>(0.0, 0.0);
      U[e] = mk<double>(a == b ? 1.0 : 0.0, 0.0);
    }
  });
  int nt0 = 0;
  ex.run([&](int tid, int nt) {
    (void)tid;
    nt0 = nt;
  });
  for (int sw = 0; sw <= max_sweeps; sw++) {
    // convergence of the whole block: every off-diagonal ratio below 1, with
    // the absolute floor k eps max|a_ii|
    ex.run([&](int tid, int nt) {
      double dmax = 0;
      for (int a = tid; a < k2; a += nt) {
        const double d = std::fabs(M[a * k2 + a].x);
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
          const double q = bj_ratio(M[e], M[a * k2 + a].x, M[b * k2 + b].x, tol, floor_);
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
      ex.run([&](int tid, int nt) {
        (void)nt;
        if (tid < m) {
          int p, q;
          cl_pair(r, tid, k2, p, q);
          double c, off;
          cplx<double> s;
          jacobi_rotation(M[p * k2 + p].x, M[q * k2 + q].x, M[p * k2 + q], c, s, off);
          cs[tid] = c;
          sj[tid] = s;
          pp[tid] = p;
          pq[tid] = q;
        }
      });
      ex.run([&](int tid, int nt) {
        const int nblk = m * m;
        for (int it = tid; it < nblk + k2 * m; it += nt) {
          if (it < nblk) {
            const int a = it / m, b = it - a * m;
            if (b >= a) cl_block(M, k2, pp[a], pq[a], cs[a], sj[a], pp[b], pq[b], cs[b], sj[b], a == b);
          } else {
            // U[row, (p, q)] J
            const int i2 = it - nblk, row = i2 / m, t = i2 - row * m;
            const int p = pp[t], q = pq[t];
            const cplx<double> up = U[row * k2 + p], uq = U[row * k2 + q];
            U[row * k2 + p] = up * cs[t] - conj(sj[t]) * uq;
            U[row * k2 + q] = sj[t] * up + uq * cs[t];
          }
        }
      });
    }
  }
  // X[:, idx] <- X[:, idx] U over the first nrows rows, in chunks of
  // kClRows rows staged in the now free M buffer
  cplx<double>* Xs = M;
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
        for (int a = 0; a < k; a++) acc = acc + Xs[r * k + a] * U[a * k2 + b];
        X[(long long)(r0 + r) * np + idx[b]] = acc;
      }
    });
  }
}

}  // namespace itp
