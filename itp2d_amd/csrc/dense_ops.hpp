// Element operations of the parallel two-sided Jacobi eigensolver for the
// N x N Hermitian overlap matrix (the reference calls LAPACK zheev here,
// eigensolver.hpp:80-84, stateset.cpp:156-177). Shared by the gfx950
// cooperative kernel and its CPU emulation.
//
// Storage is by "position": each round pairs positions (2k, 2k+1); after the
// round, rows/columns move to the next round-robin arrangement (position 0
// fixed, the others rotate), written into the other buffer of a ping-pong
// pair. After np-1 rounds (one sweep) every pair has met once and positions
// are back in their original order.
#pragma once

#include <cmath>
#include "line_ops.hpp"

namespace itp {

// New position of the element at `pos` after a round (m = np/2 pairs).
ITP_HD int rr_next(int pos, int m) {
  if (pos == 0 || m == 1) return pos;
  const int k = pos >> 1;
  if (pos & 1) return k == 0 ? 2 : 2 * (k - 1) + 1;
  return (k + 1 < m) ? 2 * (k + 1) : 2 * (m - 1) + 1;
}

// |z| without the overflow scaling of hypot: the eigensolver's matrices have
// entries far inside the FP64 range, and this sits on the critical path of
// every rotation round.
ITP_HD double cabs_fast(cplx<double> z) { return std::sqrt(z.x * z.x + z.y * z.y); }

// Rotation J = [[c, s], [-conj(s), c]] with J^H [[a, b], [conj(b), d]] J diagonal.
// off = |b| / sqrt(|a d|), the relative size of the element being removed.
// For |tau| so large that tau^2 overflows, t = 0: no rotation, the limit.
ITP_HD void jacobi_rotation(double a, double d, cplx<double> b, double& c, cplx<double>& s, double& off) {
  const double ab = cabs_fast(b);
  const double scale = std::sqrt(std::fabs(a * d));
  off = scale > 0 ? ab / scale : ab;
  if (ab == 0 || off < 1e-17) { c = 1; s = mk<double>(0.0, 0.0); return; }
  const double tau = (d - a) / (2 * ab);
  const double t = (tau >= 0 ? 1.0 : -1.0) / (std::fabs(tau) + std::sqrt(1.0 + tau * tau));
  c = 1 / std::sqrt(1 + t * t);
  const double sc = t * c / ab;
  s = mk<double>(b.x * sc, b.y * sc);
}

// Block (kI, kJ), kI <= kJ: Y = J_I^H A[I, J] J_J written at the next round's
// positions, and its conjugate transpose at the mirrored positions.
ITP_HD void jacobi_block(const cplx<double>* A, cplx<double>* An, int np, int m, int kI, int kJ,
                         const double* cs, const cplx<double>* sj) {
  const int pI = 2 * kI, qI = pI + 1, pJ = 2 * kJ, qJ = pJ + 1;
  const cplx<double> x00 = A[(long long)pI * np + pJ], x01 = A[(long long)pI * np + qJ];
  const cplx<double> x10 = A[(long long)qI * np + pJ], x11 = A[(long long)qI * np + qJ];
  const double cI = cs[kI], cJ = cs[kJ];
  const cplx<double> sI = sj[kI], sJ = sj[kJ];
  // Z = X J_J
  const cplx<double> z00 = x00 * cJ - conj(sJ) * x01;
  const cplx<double> z01 = sJ * x00 + x01 * cJ;
  const cplx<double> z10 = x10 * cJ - conj(sJ) * x11;
  const cplx<double> z11 = sJ * x10 + x11 * cJ;
  // Y = J_I^H Z
  cplx<double> y00 = z00 * cI - sI * z10;
  cplx<double> y01 = z01 * cI - sI * z11;
  cplx<double> y10 = conj(sI) * z00 + z10 * cI;
  cplx<double> y11 = conj(sI) * z01 + z11 * cI;
  const int npI = rr_next(pI, m), nqI = rr_next(qI, m), npJ = rr_next(pJ, m), nqJ = rr_next(qJ, m);
  if (kI == kJ) {
    y00.y = 0; y11.y = 0;
    if (sI.x != 0 || sI.y != 0) { y01 = mk<double>(0.0, 0.0); y10 = y01; }
    else { y10 = conj(y01); }
    An[(long long)npI * np + npI] = y00;
    An[(long long)npI * np + nqI] = y01;
    An[(long long)nqI * np + npI] = y10;
    An[(long long)nqI * np + nqI] = y11;
    return;
  }
  An[(long long)npI * np + npJ] = y00;
  An[(long long)npI * np + nqJ] = y01;
  An[(long long)nqI * np + npJ] = y10;
  An[(long long)nqI * np + nqJ] = y11;
  An[(long long)npJ * np + npI] = conj(y00);
  An[(long long)nqJ * np + npI] = conj(y01);
  An[(long long)npJ * np + nqI] = conj(y10);
  An[(long long)nqJ * np + nqI] = conj(y11);
}

// Row c of the eigenvector matrix, pair k: V[c, (p,q)] J_k, moved to the next positions.
ITP_HD void jacobi_vrow(const cplx<double>* V, cplx<double>* Vn, int np, int m, int c, int k, const double* cs,
                        const cplx<double>* sj) {
  const int p = 2 * k, q = p + 1;
  const cplx<double> vp = V[(long long)c * np + p], vq = V[(long long)c * np + q];
  const double cc = cs[k];
  const cplx<double> s = sj[k];
  Vn[(long long)c * np + rr_next(p, m)] = vp * cc - conj(s) * vq;
  Vn[(long long)c * np + rr_next(q, m)] = s * vp + vq * cc;
}

}  // namespace itp
