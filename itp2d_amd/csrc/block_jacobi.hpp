// Block two-sided Jacobi for the Hermitian overlap matrix (the reference
// calls LAPACK zheev, eigensolver.hpp:80-84), shared by the gfx950
// cooperative kernel and its CPU emulation.
//
// The np x np matrix (np a multiple of 2*JB) is split into nb = np/JB blocks
// of JB indices. An outer sweep is nb-1 rounds of the round-robin pairing of
// block positions (dense_ops.hpp's rr_next at block granularity); a table
// pos2blk maps positions to blocks, so the matrix itself never moves. In a
// round, every block pair k (blocks at positions 2k, 2k+1) is:
//   phase 1: gathered into LDS as a 2JB x 2JB sub-problem; if its relative
//            off-diagonal part is already below the tolerance it is left alone
//            (identity rotation), otherwise it is diagonalised by element
//            Jacobi sweeps (jacobi_block / jacobi_vrow), accumulating Q_k;
//   phase 2: every super-tile of A touched by a non-identity rotation becomes
//            Q_K^H A[K, K'] Q_K' (in place, with its mirror) and every
//            super-column of V becomes V[:, K] Q_K.
// Convergence is checked on the whole matrix before the first sweep and after
// each sweep, so an already diagonal matrix costs one check and no sweep.
#pragma once

#include "dense_ops.hpp"

namespace itp {

constexpr int JB = 16;          // indices per block
constexpr int JS = 2 * JB;      // super-block (block pair) size
constexpr int JS2 = JS * JS;
constexpr int kMaxBlocks = 512; // np <= 8192

struct BlockJacobiArgs {
  cplx<double>* A;
  cplx<double>* V;
  cplx<double>* Q;       // [nb/2][JS][JS] rotations of the round
  cplx<double>* D;       // [nb/2][JS][JS] diagonalised sub-blocks of the round
  int* ident;            // [nb/2] 1 if Q_k is the identity
  double* slots;         // [2 (max_sweeps + 1) + 1]: per whole-matrix check, the largest element ratio,
                         // then (second half) the largest coupling-to-gap ratio (atomic max);
                         // the last 8 bytes hold the grid barrier's counter and error flag
  double* info;          // [0] sweeps, [1] 1 converged / 2 handed off / 0 neither / -1 grid barrier timed out
  int np;
  int max_sweeps;
  int max_inner;
  double tol;
  double handoff;        // stop early once every unconverged |a_ij| < handoff |a_ii - a_jj| (0: never)
  unsigned spin_limit;   // polls of a grid barrier before it gives up (0: give up unless last to arrive)
};

// LDS for one workgroup: 4 JS x JS complex matrices + rotation parameters +
// reductions + the position table.
struct BJLds {
  cplx<double>* m[4];
  double* cs;
  cplx<double>* sj;
  double* red;  // [3 * nthreads + 3]
  int* pos2blk; // [kMaxBlocks] block at each position this round
  int* tmpblk;  // [kMaxBlocks]
  double* floor_; // [1] absolute convergence floor of the current sweep
};

ITP_HD size_t bj_lds_bytes(int nthreads) {
  return (size_t)4 * JS2 * sizeof(cplx<double>) + JB * (sizeof(double) + sizeof(cplx<double>)) +
         (size_t)(3 * nthreads + 3) * sizeof(double) + 2 * kMaxBlocks * sizeof(int) + 2 * sizeof(double) + 64;
}

ITP_HD BJLds bj_carve(void* base, int nthreads) {
  BJLds s;
  cplx<double>* p = reinterpret_cast<cplx<double>*>(base);
  for (int i = 0; i < 4; i++) { s.m[i] = p; p += JS2; }
  s.sj = p; p += JB;
  s.cs = reinterpret_cast<double*>(p);
  s.red = s.cs + JB;
  s.pos2blk = reinterpret_cast<int*>(s.red + 3 * nthreads + 3);
  s.tmpblk = s.pos2blk + kMaxBlocks;
  s.floor_ = reinterpret_cast<double*>(s.tmpblk + kMaxBlocks);
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

// Global index of local index i (0..2JB) of the super-block made of blocks b1, b2.
ITP_HD long long bj_gidx(int b1, int b2, int i) {
  return i < JB ? (long long)b1 * JB + i : (long long)b2 * JB + (i - JB);
}

// Convergence of one off-diagonal element: ratio < 1 means converged, with
//   ratio = |a_ij| / max(tol sqrt(|a_ii a_jj|), floor),
// floor = np eps max_k |a_kk| from the last whole-matrix check. The same
// element test decides the whole-matrix check and the skipping of
// sub-problems, so no unconverged element can be left alone.
ITP_HD double bj_ratio(cplx<double> aij, double aii, double ajj, double tol, double floor) {
  const double ab = cabs_fast(aij);
  double den = tol * std::sqrt(std::fabs(aii * ajj));
  den = den > floor ? den : floor;
  return den > 0 ? ab / den : (ab > 0 ? 1e300 : 0.0);
}

// Tree reduction of the per-thread maxima v[0..nt) (nt a power of two; v is
// overwritten) into *out.
template <class Exec>
ITP_HD void bj_tree_max(Exec& ex, double* v, double* out) {
  int half = 0;
  ex.run([&](int tid, int nt) { (void)tid; half = nt / 2; });
  for (; half >= 1; half /= 2) {
    const int h = half;
    ex.run([&](int tid, int nt) {
      (void)nt;
      if (tid < h) v[tid] = v[tid + h] > v[tid] ? v[tid + h] : v[tid];
      if (h == 1 && tid == 0) *out = v[0];
    });
  }
}

// Block reduction of per-thread maxima red[tid] into red[nt].
template <class Exec>
ITP_HD void bj_reduce_max(Exec& ex, const BJLds& s) {
  int nt0 = 0;
  ex.run([&](int tid, int nt) { (void)tid; nt0 = nt; });
  bj_tree_max(ex, s.red, s.red + nt0);
}

// Coupling-to-gap ratio of an unconverged off-diagonal element,
// |a_ij| / |a_ii - a_jj|: small for every pair once the matrix is close
// enough to diagonal for first-order (perturbative) eigenvector corrections.
ITP_HD double bj_gap_ratio(cplx<double> aij, double aii, double ajj) {
  const double ab = cabs_fast(aij);
  const double gap = std::fabs(aii - ajj);
  return gap > 0 ? ab / gap : 1e300;
}

// Phase 1 for block pair k (blocks b1, b2): writes Q_k, D_k and ident[k].
template <class Exec>
ITP_HD void bj_inner(Exec& ex, const BlockJacobiArgs& a, int k, int b1, int b2, const BJLds& s) {
  const int np = a.np;
  cplx<double>* M0 = s.m[0];
  cplx<double>* M1 = s.m[1];
  cplx<double>* Q0 = s.m[2];
  cplx<double>* Q1 = s.m[3];
  const cplx<double>* A = a.A;
  const double fl = s.floor_[0];
  ex.run([&](int tid, int nt) {
    double r = 0;
    for (int e = tid; e < JS2; e += nt) {
      const int i = e / JS, j = e - i * JS;
      const long long gi = bj_gidx(b1, b2, i), gj = bj_gidx(b1, b2, j);
      M0[e] = A[gi * np + gj];
      Q0[e] = mk<double>(i == j ? 1.0 : 0.0, 0.0);
      if (j > i) {
        const double q = bj_ratio(A[gi * np + gj], A[gi * np + gi].x, A[gj * np + gj].x, a.tol, fl);
        r = q > r ? q : r;
      }
    }
    s.red[tid] = r;
  });
  bj_reduce_max(ex, s);
  int cur = 0;
  bool skip = false;
  ex.run([&](int tid, int nt) { (void)tid; skip = s.red[nt] < 1.0; });
  if (!skip) {
    for (int sw = 0; sw < a.max_inner; sw++) {
      for (int r = 0; r < JS - 1; r++) {
        const cplx<double>* Mc = cur ? M1 : M0;
        cplx<double>* Mn = cur ? M0 : M1;
        const cplx<double>* Qc = cur ? Q1 : Q0;
        cplx<double>* Qn = cur ? Q0 : Q1;
        ex.run([&](int tid, int nt) {
          if (tid < JB) {
            const int p0 = 2 * tid, p1 = p0 + 1;
            double c, off;
            cplx<double> sv;
            jacobi_rotation(Mc[p0 * JS + p0].x, Mc[p1 * JS + p1].x, Mc[p0 * JS + p1], c, sv, off);
            s.cs[tid] = c;
            s.sj[tid] = sv;
            const double q = bj_ratio(Mc[p0 * JS + p1], Mc[p0 * JS + p0].x, Mc[p1 * JS + p1].x, a.tol, fl);
            if (r == 0) s.red[tid] = 0;
            s.red[tid] = q > s.red[tid] ? q : s.red[tid];
          }
          (void)nt;
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
      // stop when a whole inner sweep met only converged pairs
      bool done = false;
      ex.run([&](int tid, int nt) {
        (void)nt;
        (void)tid;
        double m = 0;
        for (int t = 0; t < JB; t++) m = s.red[t] > m ? s.red[t] : m;
        done = m < 1.0;
      });
      if (done) break;
    }
  }
  const cplx<double>* Mf = cur ? M1 : M0;
  const cplx<double>* Qf = cur ? Q1 : Q0;
  cplx<double>* Qg = a.Q + (long long)k * JS2;
  cplx<double>* Dg = a.D + (long long)k * JS2;
  int* idg = a.ident + k;
  ex.run([&](int tid, int nt) {
    if (!skip)
      for (int e = tid; e < JS2; e += nt) { Qg[e] = Qf[e]; Dg[e] = Mf[e]; }
    if (tid == 0) *idg = skip ? 1 : 0;
  });
}

// 32 x 32 complex product C = op(A) B from LDS operands (op = conjugate
// transpose if ah). On the GPU: v_mfma_f64_16x16x4f64, one 16 x 16 quarter per
// wavefront (lane l: A[l%16][l/16], B[l/16][l%16], D[l/16+4v][l%16]); in the
// CPU emulation: a plain loop.
template <class Exec>
ITP_HD void bj_cgemm32(Exec& ex, const cplx<double>* A, bool ah, const cplx<double>* B, cplx<double>* C) {
#if defined(__HIP_DEVICE_COMPILE__)
  (void)ex;
  typedef double d4 __attribute__((ext_vector_type(4)));
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wi = w >> 1, wj = w & 1;
  const int li = lane & 15, lk = lane >> 4;
  d4 cr = {0, 0, 0, 0}, ci = {0, 0, 0, 0};
  for (int kk = 0; kk < JS / 4; kk++) {
    const int kx = kk * 4 + lk;
    const cplx<double> av = ah ? conj(A[kx * JS + wi * 16 + li]) : A[(wi * 16 + li) * JS + kx];
    const cplx<double> bv = B[kx * JS + wj * 16 + li];
    cr = __builtin_amdgcn_mfma_f64_16x16x4f64(av.x, bv.x, cr, 0, 0, 0);
    cr = __builtin_amdgcn_mfma_f64_16x16x4f64(-av.y, bv.y, cr, 0, 0, 0);
    ci = __builtin_amdgcn_mfma_f64_16x16x4f64(av.x, bv.y, ci, 0, 0, 0);
    ci = __builtin_amdgcn_mfma_f64_16x16x4f64(av.y, bv.x, ci, 0, 0, 0);
  }
  for (int v = 0; v < 4; v++) C[(wi * 16 + lk + 4 * v) * JS + wj * 16 + li] = mk<double>(cr[v], ci[v]);
  __syncthreads();
#else
  ex.run([&](int tid, int nt) {
    for (int e = tid; e < JS2; e += nt) {
      const int i = e / JS, j = e - i * JS;
      cplx<double> acc = mk<double>(0.0, 0.0);
      for (int l = 0; l < JS; l++) {
        const cplx<double> av = ah ? conj(A[l * JS + i]) : A[i * JS + l];
        acc = acc + av * B[l * JS + j];
      }
      C[e] = acc;
    }
  });
#endif
}

// Phase 2 work item t: A super-tiles (K <= K') first, then V super-columns.
// pos2blk gives the blocks of each pair for this round.
template <class Exec>
ITP_HD void bj_apply(Exec& ex, const BlockJacobiArgs& a, int t, const BJLds& s) {
  const int np = a.np;
  const int P = np / JS;
  const int ntA = P * (P + 1) / 2;
  cplx<double>* X = s.m[0];
  cplx<double>* Qa = s.m[1];
  cplx<double>* Qb = s.m[2];
  cplx<double>* T = s.m[3];
  cplx<double>* A = a.A;
  cplx<double>* V = a.V;
  if (t < ntA) {
    int K = 0, rem = t;
    while (rem >= P - K) { rem -= P - K; K++; }
    const int K2 = K + rem;
    if (a.ident[K] && a.ident[K2]) return;
    const int a1 = s.pos2blk[2 * K], a2 = s.pos2blk[2 * K + 1];
    const int c1 = s.pos2blk[2 * K2], c2 = s.pos2blk[2 * K2 + 1];
    if (K == K2) {
      const cplx<double>* Dg = a.D + (long long)K * JS2;
      ex.run([&](int tid, int nt) {
        for (int e = tid; e < JS2; e += nt) {
          const int i = e / JS, j = e - i * JS;
          A[bj_gidx(a1, a2, i) * np + bj_gidx(a1, a2, j)] = Dg[e];
        }
      });
      return;
    }
    const cplx<double>* QgA = a.Q + (long long)K * JS2;
    const cplx<double>* QgB = a.Q + (long long)K2 * JS2;
    const bool ia = a.ident[K] != 0, ib = a.ident[K2] != 0;
    ex.run([&](int tid, int nt) {
      for (int e = tid; e < JS2; e += nt) {
        const int i = e / JS, j = e - i * JS;
        X[e] = A[bj_gidx(a1, a2, i) * np + bj_gidx(c1, c2, j)];
        Qa[e] = ia ? mk<double>(i == j ? 1.0 : 0.0, 0.0) : QgA[e];
        Qb[e] = ib ? mk<double>(i == j ? 1.0 : 0.0, 0.0) : QgB[e];
      }
    });
    bj_cgemm32(ex, X, false, Qb, T);   // T = X Qb
    bj_cgemm32(ex, Qa, true, T, X);    // X = Qa^H T
    ex.run([&](int tid, int nt) {
      for (int e = tid; e < JS2; e += nt) {
        const int i = e / JS, j = e - i * JS;
        const long long gi = bj_gidx(a1, a2, i), gj = bj_gidx(c1, c2, j);
        A[gi * np + gj] = X[e];
        A[gj * np + gi] = conj(X[e]);
      }
    });
    return;
  }
  // V super-column: rows [R*JS, R*JS+JS), pair K
  const int t2 = t - ntA;
  const int R = t2 / P, K = t2 - R * P;
  if (a.ident[K]) return;
  const int b1 = s.pos2blk[2 * K], b2 = s.pos2blk[2 * K + 1];
  const cplx<double>* Qg = a.Q + (long long)K * JS2;
  ex.run([&](int tid, int nt) {
    for (int e = tid; e < JS2; e += nt) {
      const int i = e / JS, j = e - i * JS;
      X[e] = V[((long long)R * JS + i) * np + bj_gidx(b1, b2, j)];
      Qa[e] = Qg[e];
    }
  });
  bj_cgemm32(ex, X, false, Qa, T);
  ex.run([&](int tid, int nt) {
    for (int e = tid; e < JS2; e += nt) {
      const int i = e / JS, j = e - i * JS;
      V[((long long)R * JS + i) * np + bj_gidx(b1, b2, j)] = T[e];
    }
  });
}

ITP_HD int bj_items(int np) {
  const int P = np / JS;
  return P * (P + 1) / 2 + P * P;
}

// Position table for the next round (all workgroups hold identical copies).
template <class Exec>
ITP_HD void bj_next_positions(Exec& ex, const BJLds& s, int nb) {
  ex.run([&](int tid, int nt) {
    for (int p = tid; p < nb; p += nt) s.tmpblk[rr_next(p, nb / 2)] = s.pos2blk[p];
  });
  ex.run([&](int tid, int nt) {
    for (int p = tid; p < nb; p += nt) s.pos2blk[p] = s.tmpblk[p];
  });
}

template <class Exec>
ITP_HD void bj_reset_positions(Exec& ex, const BJLds& s, int nb) {
  ex.run([&](int tid, int nt) {
    for (int p = tid; p < nb; p += nt) s.pos2blk[p] = p;
  });
}

// Whole-matrix check, elements [e0, e1) of this workgroup's slice: every
// workgroup first derives the same floor from the whole diagonal, then the
// largest element ratio of its slice lands in red[nt] and the largest
// coupling-to-gap ratio of its unconverged elements in red[nt + 1].
template <class Exec>
ITP_HD void bj_check_rows(Exec& ex, const BlockJacobiArgs& a, long long e0, long long e1, const BJLds& s) {
  const int np = a.np;
  const cplx<double>* A = a.A;
  ex.run([&](int tid, int nt) {
    double m = 0;
    for (int i = tid; i < np; i += nt) {
      const double d = std::fabs(A[(long long)i * np + i].x);
      m = d > m ? d : m;
    }
    s.red[tid] = m;
  });
  bj_reduce_max(ex, s);
  ex.run([&](int tid, int nt) {
    if (tid == 0) s.floor_[0] = (double)np * 2.220446049250313e-16 * s.red[nt];
  });
  const double fl = s.floor_[0];
  double* rho = s.red + 0;
  ex.run([&](int tid, int nt) { rho = s.red + nt + 2; });
  ex.run([&](int tid, int nt) {
    double r = 0, g = 0;
    for (long long e = e0 + tid; e < e1; e += nt) {
      const int i = (int)(e / np), j = (int)(e - (long long)i * np);
      if (j > i) {
        const double aii = A[(long long)i * np + i].x, ajj = A[(long long)j * np + j].x;
        const double q = bj_ratio(A[e], aii, ajj, a.tol, fl);
        r = q > r ? q : r;
        if (q >= 1.0) {
          const double gr = bj_gap_ratio(A[e], aii, ajj);
          g = gr > g ? gr : g;
        }
      }
    }
    s.red[tid] = r;
    rho[tid] = g;
  });
  int nt0 = 0;
  ex.run([&](int tid, int nt) { (void)tid; nt0 = nt; });
  bj_tree_max(ex, rho, s.red + nt0 + 1);
  bj_reduce_max(ex, s);
}

}  // namespace itp
