// Block two-sided Jacobi for the Hermitian overlap matrix (the reference
// calls LAPACK zheev, eigensolver.hpp:80-84), shared by the gfx950
// cooperative kernel and its CPU emulation.
//
// The np x np matrix (np a multiple of 2*JB) is split into nb = np/JB blocks
// of JB indices. An outer sweep is nb-1 rounds of the round-robin pairing of
// blocks (dense_ops.hpp's rr_next at block granularity). In a round, every
// block pair k (super-block of 2*JB indices at positions 2k, 2k+1) is:
//   phase 1: loaded into LDS and diagonalised by element Jacobi sweeps
//            (jacobi_block / jacobi_vrow on the 2JB x 2JB sub-problem),
//            accumulating the 2JB x 2JB unitary Q_k;
//   phase 2: every super-tile of A becomes Q_K^H A[K, K'] Q_K' and every
//            super-column of V becomes V[:, K] Q_K, written at the next
//            round's block positions (ping-pong buffers).
// After a full outer sweep blocks are back in their original positions, so the
// diagonal of A holds the eigenvalues and V's columns the eigenvectors in the
// original index order.
#pragma once

#include "dense_ops.hpp"

namespace itp {

constexpr int JB = 16;          // indices per block
constexpr int JS = 2 * JB;      // super-block (block pair) size
constexpr int JS2 = JS * JS;

struct BlockJacobiArgs {
  cplx<double>* A[2];
  cplx<double>* V[2];
  cplx<double>* Q;       // [nb/2][JS][JS] rotations of the round
  cplx<double>* D;       // [nb/2][JS][JS] diagonalised sub-blocks of the round
  double* slots;         // [max_sweeps][3]: max relative off-diagonal, max |off|, max |diag| (as atomics)
  double* info;          // [0] sweeps, [1] converged, [2] final buffer
  int np;
  int max_sweeps;
  int max_inner;
  double tol;
};

// LDS for one workgroup: 4 JS x JS complex matrices + rotation parameters + reductions
struct BJLds {
  cplx<double>* m[4];
  double* cs;
  cplx<double>* sj;
  double* red;  // [3 * nthreads]
};

ITP_HD size_t bj_lds_bytes(int nthreads) {
  return (size_t)4 * JS2 * sizeof(cplx<double>) + JB * (sizeof(double) + sizeof(cplx<double>)) +
         (size_t)3 * nthreads * sizeof(double) + 64;
}

ITP_HD BJLds bj_carve(void* base) {
  BJLds s;
  cplx<double>* p = reinterpret_cast<cplx<double>*>(base);
  for (int i = 0; i < 4; i++) { s.m[i] = p; p += JS2; }
  s.sj = p; p += JB;
  s.cs = reinterpret_cast<double*>(p);
  s.red = s.cs + JB;
  return s;
}

// Non-negative doubles order like their bit patterns; the maxima are kept as
// 64-bit integers so atomicMax can combine them.
ITP_HD unsigned long long dbits(double x) {
  union { double d; unsigned long long u; } c;
  c.d = x;
  return c.u;
}
ITP_HD double bitsd(unsigned long long u) {
  union { double d; unsigned long long u; } c;
  c.u = u;
  return c.d;
}

// Phase 1 for block pair k. Writes Q_k and the diagonalised sub-block D_k,
// and returns (in out3) the largest relative and absolute off-diagonal and
// largest diagonal seen in the first inner sweep (the sub-problem as loaded).
template <class Exec>
ITP_HD void bj_inner(Exec& ex, const BlockJacobiArgs& a, const cplx<double>* A, int k, const BJLds& s,
                     double* out3) {
  const int np = a.np;
  const long long r0 = (long long)k * JS;
  cplx<double>* M0 = s.m[0];
  cplx<double>* M1 = s.m[1];
  cplx<double>* Q0 = s.m[2];
  cplx<double>* Q1 = s.m[3];
  ex.run([&](int tid, int nt) {
    for (int e = tid; e < JS2; e += nt) {
      const int i = e / JS, j = e - i * JS;
      M0[e] = A[(r0 + i) * np + r0 + j];
      Q0[e] = mk<double>(i == j ? 1.0 : 0.0, 0.0);
    }
  });
  int cur = 0;
  double first[3] = {0, 0, 0};
  for (int sw = 0; sw < a.max_inner; sw++) {
    for (int r = 0; r < JS - 1; r++) {
      const cplx<double>* Mc = cur ? M1 : M0;
      cplx<double>* Mn = cur ? M0 : M1;
      const cplx<double>* Qc = cur ? Q1 : Q0;
      cplx<double>* Qn = cur ? Q0 : Q1;
      ex.run([&](int tid, int nt) {
        for (int q = tid; q < JB; q += nt) {
          const int p0 = 2 * q, p1 = p0 + 1;
          double c, off;
          cplx<double> sv;
          const double app = Mc[p0 * JS + p0].x, aqq = Mc[p1 * JS + p1].x;
          const cplx<double> b = Mc[p0 * JS + p1];
          jacobi_rotation(app, aqq, b, c, sv, off);
          s.cs[q] = c;
          s.sj[q] = sv;
          const double ab = std::hypot(b.x, b.y);
          const double dd = std::fabs(app) > std::fabs(aqq) ? std::fabs(app) : std::fabs(aqq);
          double* rd = s.red + 3 * tid;
          if (r == 0) { rd[0] = 0; rd[1] = 0; rd[2] = 0; }
          rd[0] = off > rd[0] ? off : rd[0];
          rd[1] = ab > rd[1] ? ab : rd[1];
          rd[2] = dd > rd[2] ? dd : rd[2];
        }
      });
      ex.run([&](int tid, int nt) {
        const int nblk = JB * JB;
        for (int it = tid; it < nblk + JS * JB; it += nt) {
          if (it < nblk) {
            const int kI = it / JB, kJ = it - kI * JB;
            if (kJ >= kI) jacobi_block(Mc, Mn, JS, JB, kI, kJ, s.cs, s.sj);
          } else {
            const int i2 = it - nblk;
            jacobi_vrow(Qc, Qn, JS, JB, i2 / JB, i2 % JB, s.cs, s.sj);
          }
        }
      });
      cur ^= 1;
    }
    // sweep maxima (threads < JB hold them)
    double mx[3] = {0, 0, 0};
    ex.run([&](int tid, int nt) {
      (void)nt;
      if (tid == 0) {
        double m0 = 0, m1 = 0, m2 = 0;
        for (int t = 0; t < JB; t++) {
          const double* rd = s.red + 3 * t;
          m0 = rd[0] > m0 ? rd[0] : m0;
          m1 = rd[1] > m1 ? rd[1] : m1;
          m2 = rd[2] > m2 ? rd[2] : m2;
        }
        s.red[3 * JB] = m0;
        s.red[3 * JB + 1] = m1;
        s.red[3 * JB + 2] = m2;
      }
    });
    mx[0] = s.red[3 * JB];
    mx[1] = s.red[3 * JB + 1];
    mx[2] = s.red[3 * JB + 2];
    if (sw == 0) { first[0] = mx[0]; first[1] = mx[1]; first[2] = mx[2]; }
    const bool done = (mx[0] < a.tol) || (mx[1] <= 1e-15 * mx[2]);
    ex.run([&](int tid, int nt) { (void)tid; (void)nt; });  // all threads read the maxima before reuse
    if (done) break;
  }
  const cplx<double>* Mf = cur ? M1 : M0;
  const cplx<double>* Qf = cur ? Q1 : Q0;
  cplx<double>* Qg = a.Q + (long long)k * JS2;
  cplx<double>* Dg = a.D + (long long)k * JS2;
  ex.run([&](int tid, int nt) {
    for (int e = tid; e < JS2; e += nt) { Qg[e] = Qf[e]; Dg[e] = Mf[e]; }
  });
  out3[0] = first[0];
  out3[1] = first[1];
  out3[2] = first[2];
}

// Global row index after the round for local row r of pair K (2JB rows).
ITP_HD long long bj_newrow(int K, int r, int mo) {
  const int blk = 2 * K + r / JB;
  return (long long)rr_next(blk, mo) * JB + (r % JB);
}

// Phase 2 work item t: A super-tiles (K <= K') first, then V super-columns.
template <class Exec>
ITP_HD void bj_apply(Exec& ex, const BlockJacobiArgs& a, const cplx<double>* A, cplx<double>* An,
                     const cplx<double>* V, cplx<double>* Vn, int t, const BJLds& s) {
  const int np = a.np;
  const int P = np / JS;       // block pairs
  const int mo = np / JB / 2;  // outer round-robin pairs
  const int ntA = P * (P + 1) / 2;
  cplx<double>* X = s.m[0];
  cplx<double>* Qa = s.m[1];
  cplx<double>* Qb = s.m[2];
  cplx<double>* T = s.m[3];
  if (t < ntA) {
    int K = 0, rem = t;
    while (rem >= P - K) { rem -= P - K; K++; }
    const int K2 = K + rem;
    if (K == K2) {
      const cplx<double>* Dg = a.D + (long long)K * JS2;
      ex.run([&](int tid, int nt) {
        for (int e = tid; e < JS2; e += nt) {
          const int i = e / JS, j = e - i * JS;
          An[bj_newrow(K, i, mo) * np + bj_newrow(K, j, mo)] = Dg[e];
        }
      });
      return;
    }
    const cplx<double>* QgA = a.Q + (long long)K * JS2;
    const cplx<double>* QgB = a.Q + (long long)K2 * JS2;
    ex.run([&](int tid, int nt) {
      for (int e = tid; e < JS2; e += nt) {
        const int i = e / JS, j = e - i * JS;
        X[e] = A[((long long)K * JS + i) * np + (long long)K2 * JS + j];
        Qa[e] = QgA[e];
        Qb[e] = QgB[e];
      }
    });
    // T = X Qb
    ex.run([&](int tid, int nt) {
      for (int e = tid; e < JS2; e += nt) {
        const int i = e / JS, j = e - i * JS;
        cplx<double> acc = mk<double>(0.0, 0.0);
        for (int l = 0; l < JS; l++) acc = acc + X[i * JS + l] * Qb[l * JS + j];
        T[e] = acc;
      }
    });
    // Y = Qa^H T, written with its mirror
    ex.run([&](int tid, int nt) {
      for (int e = tid; e < JS2; e += nt) {
        const int i = e / JS, j = e - i * JS;
        cplx<double> acc = mk<double>(0.0, 0.0);
        for (int l = 0; l < JS; l++) acc = acc + cmulc(Qa[l * JS + i], T[l * JS + j]);
        const long long gi = bj_newrow(K, i, mo), gj = bj_newrow(K2, j, mo);
        An[gi * np + gj] = acc;
        An[gj * np + gi] = conj(acc);
      }
    });
    return;
  }
  // V super-column: rows [R*JS, R*JS+JS), pair K
  const int t2 = t - ntA;
  const int R = t2 / P, K = t2 - R * P;
  const cplx<double>* Qg = a.Q + (long long)K * JS2;
  ex.run([&](int tid, int nt) {
    for (int e = tid; e < JS2; e += nt) {
      const int i = e / JS, j = e - i * JS;
      X[e] = V[((long long)R * JS + i) * np + (long long)K * JS + j];
      Qa[e] = Qg[e];
    }
  });
  ex.run([&](int tid, int nt) {
    for (int e = tid; e < JS2; e += nt) {
      const int i = e / JS, j = e - i * JS;
      cplx<double> acc = mk<double>(0.0, 0.0);
      for (int l = 0; l < JS; l++) acc = acc + X[i * JS + l] * Qa[l * JS + j];
      Vn[((long long)R * JS + i) * np + bj_newrow(K, j, mo)] = acc;
    }
  });
}

ITP_HD int bj_items(int np) {
  const int P = np / JS;
  return P * (P + 1) / 2 + P * P;
}

}  // namespace itp
