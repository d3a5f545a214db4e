// Host-side engine: grid tables, multiplier tables with the reference's
// formulas, radix plans, and the sweep sequences that implement the
// operators. Templated on a Backend that owns memory and runs sweeps, so the
// same sequencing drives the gfx950 kernels (gpu backend) and the CPU
// emulation of those kernels used by the tests.
#pragma once

#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "line_ops.hpp"
#include "reg_sweep.hpp"

namespace itp {

enum Boundary { kPeriodic = 0, kDirichlet = 1 };

// Magnetic-field / boundary branch of the kinetic operators
// (expkinetic.cpp:138-197, kinetic.cpp:99-142).
enum Branch { BR_B0P = 0, BR_B0D = 1, BR_BP = 2, BR_BD = 3 };
inline Branch branch_of(double B, int bt) {
  if (B == 0) return bt == kPeriodic ? BR_B0P : BR_B0D;
  return bt == kPeriodic ? BR_BP : BR_BD;
}

// DataLayout (datalayout.cpp:21-36) + Transformer tables (transformer.cpp:66-90).
struct GridInfo {
  int sx = 0, sy = 0;
  long long M = 0;
  double dx = 0, lenx = 0, leny = 0;
  int bt = kPeriodic;
  std::vector<double> posx, posy, fft_kx, fft_ky, dsct_kx, dsct_ky;
  double nFFT = 0, nFFTx = 0, nFFTy = 0, nDSCT = 0, nDSCTx = 0, nDSCTy = 0;
  double kx(int x) const { return bt == kPeriodic ? fft_kx[x] : dsct_kx[x]; }
  double ky(int y) const { return bt == kPeriodic ? fft_ky[y] : dsct_ky[y]; }
  double norm() const { return bt == kPeriodic ? nFFT : nDSCT; }
  double normx() const { return bt == kPeriodic ? nFFTx : nDSCTx; }
  double normy() const { return bt == kPeriodic ? nFFTy : nDSCTy; }
};

inline GridInfo make_grid(int sx, int sy, double dx, int bt) {
  GridInfo g;
  g.sx = sx; g.sy = sy; g.M = (long long)sx * sy; g.dx = dx; g.bt = bt;
  g.lenx = double(sx) * dx;
  g.leny = double(sy) * dx;
  g.posx.resize(sx); g.posy.resize(sy);
  for (int x = 0; x < sx; x++) g.posx[x] = double(2 * x + 1 - sx) * 0.5 * dx;
  for (int y = 0; y < sy; y++) g.posy[y] = double(2 * y + 1 - sy) * 0.5 * dx;
  g.nFFT = 1.0 / double((size_t)sx * sy);
  g.nFFTx = 1.0 / double(sx);
  g.nFFTy = 1.0 / double(sy);
  g.nDSCT = 1.0 / double(4 * (size_t)sx * sy);
  g.nDSCTx = 1.0 / double(2 * sx);
  g.nDSCTy = 1.0 / double(2 * sy);
  const double mx = M_PI / g.lenx, my = M_PI / g.leny;
  g.fft_kx.resize(sx); g.dsct_kx.resize(sx); g.fft_ky.resize(sy); g.dsct_ky.resize(sy);
  for (int x = 0; x < sx; x++) {
    g.fft_kx[x] = double((x < sx / 2) ? x : x - sx) * 2 * mx;
    g.dsct_kx[x] = double(x + 1) * mx;
  }
  for (int y = 0; y < sy; y++) {
    g.fft_ky[y] = double((y < sy / 2) ? y : y - sy) * 2 * my;
    g.dsct_ky[y] = double(y + 1) * my;
  }
  return g;
}

// Radix plan with an even number of stages, radices 2..8 for the power of two
// part plus 3s and 5s. Returns false for lengths it cannot plan.
inline bool make_plan(int n, LinePlan& p) {
  p.n = n;
  if (n < 2) return false;
  int m = n, k = 0, a = 0, b = 0;
  while (m % 2 == 0) { m /= 2; k++; }
  while (m % 3 == 0) { m /= 3; a++; }
  while (m % 5 == 0) { m /= 5; b++; }
  if (m != 1) return false;
  int s2 = (k + 2) / 3;                    // fewest radix<=8 stages for 2^k
  if ((s2 + a + b) % 2) s2++;
  if (s2 > k) return false;
  if (s2 + a + b > kMaxStages || s2 + a + b == 0) return false;
  int st = 0;
  for (int i = 0; i < s2; i++) {
    const int bits = k / s2 + (i < k % s2 ? 1 : 0);
    p.radix[st++] = 1 << bits;
  }
  for (int i = 0; i < a; i++) p.radix[st++] = 3;
  for (int i = 0; i < b; i++) p.radix[st++] = 5;
  p.nst = st;
  for (int i = 0; i < st; i++)
    if (p.radix[i] > 8) return false;
  return p.nst % 2 == 0;
}

inline void twiddle_tables(int n, std::vector<cplx<double>>& tw, std::vector<cplx<double>>& mkt) {
  tw.resize(n); mkt.resize(n);
  for (int t = 0; t < n; t++) {
    const double a = -2.0 * M_PI * double(t) / double(n);
    tw[t] = mk<double>(std::cos(a), std::sin(a));
    const double b = -M_PI * double(t) / double(2 * n);
    mkt[t] = mk<double>(std::cos(b), std::sin(b));
  }
}

// ExpPotential values p*exp(e*c*V) (exppotential.cpp:55-63)
inline void exppot_values(const GridInfo& g, const double* V, double e, double c, double p,
                          std::vector<double>& out) {
  out.resize(g.M);
  for (long long i = 0; i < g.M; i++) out[i] = p * std::exp(e * c * V[i]);
}

// ExpKinetic multipliers (expkinetic.cpp:65-136)
struct ExpKinTabs { std::vector<double> mult, xm, xm2, ym; };
inline void expkin_tables(const GridInfo& g, double B, double e, double coeff, double pref, ExpKinTabs& t) {
  const int sx = g.sx, sy = g.sy;
  const int bt = g.bt;
  if (B == 0) {
    t.mult.resize(g.M);
    const double normfac = g.norm();
    const double A = coeff * e * 0.5;
    const double p = pref * normfac;
    for (int y = 0; y < sy; y++) {
      const double ky = g.ky(y);
      for (int x = 0; x < sx; x++) {
        const double kx = g.kx(x);
        t.mult[(size_t)y * sx + x] = p * std::exp(A * (kx * kx + ky * ky));
      }
    }
    return;
  }
  t.xm.resize(g.M); t.ym.resize(sy);
  if (bt == kDirichlet) t.xm2.resize(g.M);
  const double normfacx = (bt == kPeriodic) ? std::sqrt(g.nFFTx) : g.nDSCTx;
  const double normfacy = g.normy();
  const double z = B * e;
  const double Cy = (std::fabs(z) < 1e-6) ? 1 + std::pow(z, 2) / 6 + std::pow(z, 4) / 120 : std::sinh(z) / z;
  const double Cx = (std::fabs(z) < 1e-6) ? 0.5 - std::pow(z, 2) / 24 + std::pow(z, 4) / 240 : std::tanh(z / 2) / z;
  const double Ax = coeff * e * 0.5 * Cx;
  const double Ay = coeff * e * 0.5 * Cy;
  const double px = normfacx;
  const double py = pref * normfacy;
  for (int y = 0; y < sy; y++) {
    const double ky = g.ky(y);
    const double dy = g.posy[y];
    t.ym[y] = py * std::exp(Ay * ky * ky);
    if (bt == kPeriodic) {
      for (int x = 0; x < sx; x++) {
        const double kx = g.kx(x);
        t.xm[(size_t)y * sx + x] = px * std::exp(Ax * (kx - B * dy) * (kx - B * dy));
      }
    } else {
      const double By = B * dy;
      for (int x = 0; x < sx; x++) {
        const double kx = g.kx(x);
        const double common = px * std::exp(Ax * (kx * kx + By * By));
        const double arg = 2 * Ax * kx * By;
        t.xm[(size_t)y * sx + x] = common * std::cosh(arg);
        t.xm2[(size_t)y * sx + x] = common * std::sinh(arg);
      }
    }
  }
}

// Kinetic tables (kinetic.cpp:25-83)
struct KinTabs { std::vector<double> mxy, mx, mx2, my; };
inline void kin_tables(const GridInfo& g, double B, KinTabs& t) {
  const int sx = g.sx, sy = g.sy;
  const double normfac = g.norm(), normfac_x = g.normx(), normfac_y = g.normy();
  if (B == 0) {
    t.mxy.resize(g.M);
    for (int y = 0; y < sy; y++) {
      const double ky = g.ky(y);
      for (int x = 0; x < sx; x++) {
        const double kx = g.kx(x);
        t.mxy[(size_t)y * sx + x] = 0.5 * (kx * kx + ky * ky) * normfac;
      }
    }
    return;
  }
  t.mx.resize(g.M); t.my.resize(sy);
  if (g.bt == kDirichlet) t.mx2.resize(g.M);
  for (int y = 0; y < sy; y++) {
    const double ky = g.ky(y);
    t.my[y] = 0.5 * ky * ky * normfac_y;
    const double dy = g.posy[y];
    for (int x = 0; x < sx; x++) {
      if (g.bt == kPeriodic) {
        const double kx = g.kx(x) - B * dy;
        t.mx[(size_t)y * sx + x] = 0.5 * kx * kx * normfac_x;
      } else {
        const double kx = g.kx(x);
        t.mx[(size_t)y * sx + x] = 0.5 * (kx * kx + B * B * dy * dy) * normfac_x;
        t.mx2[(size_t)y * sx + x] = B * dy * kx * normfac_x;
      }
    }
  }
}

// MultiProductSplit coefficients, formula (2.12) of Chin's paper
// (multiproductsplit.cpp:49-58)
inline std::vector<double> mps_coefficients(int h) {
  std::vector<double> c(h);
  for (int t = 1; t <= h; t++) {
    int denom = 1;
    for (int n = 1; n <= h; n++) if (n != t) denom *= (t * t - n * n);
    c[t - 1] = std::pow(t, 2 * (h - 1)) / denom;
  }
  return c;
}

// Device-resident (or emulator-resident) copy of one ExpKinetic's tables.
// (the *32 pointers: FP32 copies of the same tables, or null)
struct KinDev {
  const double* mult = nullptr; const double* xm = nullptr; const double* xm2 = nullptr; const double* ym = nullptr;
  const float* mult32 = nullptr; const float* xm32 = nullptr; const float* xm2_32 = nullptr; const float* ym32 = nullptr;
};

// The x-direction halves of one kinetic application, per branch.
//   forward part (first row sweep of an application)
//   inverse part (last row sweep of an application)
inline void fwd_x_ops(Branch br, const KinDev& k, std::vector<LineOp>& ops) {
  switch (br) {
    case BR_B0P: ops.push_back({LOP_FFT_FWD, nullptr, nullptr, 0}); break;
    case BR_B0D: ops.push_back({LOP_DST2, nullptr, nullptr, 0}); break;
    case BR_BP:
      ops.push_back({LOP_FFT_FWD, nullptr, nullptr, 0});
      ops.push_back({LOP_MUL_TAB, k.xm, nullptr, 0, k.xm32, nullptr});
      break;
    case BR_BD:
      ops.push_back({LOP_DST2, nullptr, nullptr, 0});
      ops.push_back({LOP_XSPLIT, k.xm, k.xm2, 0, k.xm32, k.xm2_32});
      ops.push_back({LOP_DST2, nullptr, nullptr, 0});
      break;
  }
}
inline void inv_x_ops(Branch br, const KinDev& k, std::vector<LineOp>& ops) {
  switch (br) {
    case BR_B0P: ops.push_back({LOP_FFT_INV, nullptr, nullptr, 0}); break;
    case BR_B0D: ops.push_back({LOP_DST3, nullptr, nullptr, 0}); break;
    case BR_BP:
      ops.push_back({LOP_MUL_TAB, k.xm, nullptr, 0, k.xm32, nullptr});
      ops.push_back({LOP_FFT_INV, nullptr, nullptr, 0});
      break;
    case BR_BD: ops.push_back({LOP_XSPLIT, k.xm, k.xm2, 0, k.xm32, k.xm2_32}); break;
  }
}
// column sweep: y transform, multiplier, inverse y transform
inline void col_ops(Branch br, const KinDev& k, std::vector<LineOp>& ops) {
  const bool per = (br == BR_B0P || br == BR_BP);
  ops.push_back({per ? LOP_FFT_FWD : LOP_DST2, nullptr, nullptr, 0});
  if (br == BR_B0P || br == BR_B0D) ops.push_back({LOP_MUL_TAB, k.mult, nullptr, 0, k.mult32, nullptr});
  else ops.push_back({LOP_MUL_POS, k.ym, nullptr, 0, k.ym32, nullptr});
  ops.push_back({per ? LOP_FFT_INV : LOP_DST3, nullptr, nullptr, 0});
}

// One sweep to run: lines orientation, ops and buffer bindings.
template <class R>
struct Sweep {
  int orient = 0;
  std::vector<LineOp> ops;
  const cplx<R>* src = nullptr;
  cplx<R>* dst = nullptr;
  const cplx<R>* add1 = nullptr;
  const double* vtab = nullptr;
  const cplx<R>* vsrc = nullptr;
  const cplx<R>* dotsrc = nullptr;
  cplx<double>* partials = nullptr;
};

struct LineGeom {
  LinePlan plan_x, plan_y;
  const cplx<double>* tw_x = nullptr;
  const cplx<double>* mk_x = nullptr;
  const cplx<double>* tw_y = nullptr;
  const cplx<double>* mk_y = nullptr;
  int L_row = 4, L_col = 4;
  bool use_reg = true;  // register-resident engine where the line length allows it
};

inline int pick_L(int lines, int pref) {
  int L = pref;
  while (L > 1 && lines % L) L /= 2;
  return L;
}

// Engine and lines per workgroup of a sweep orientation (0 rows, 1 columns).
inline bool sweep_uses_reg(const GridInfo& g, const LineGeom& lg, int orient) {
  const int n = orient == 0 ? g.sx : g.sy;
  const int lines = orient == 0 ? g.sy : g.sx;
  return lg.use_reg && reg_supported(n) && lines % reg_lines_per_wg(n) == 0;
}
inline int sweep_lines_per_block(const GridInfo& g, const LineGeom& lg, int orient) {
  const int n = orient == 0 ? g.sx : g.sy;
  if (sweep_uses_reg(g, lg, orient)) return reg_lines_per_wg(n);
  return orient == 0 ? pick_L(g.sy, lg.L_row) : pick_L(g.sx, lg.L_col);
}
inline int sweep_blocks(const GridInfo& g, const LineGeom& lg, int orient) {
  const int lines = orient == 0 ? g.sy : g.sx;
  return lines / sweep_lines_per_block(g, lg, orient);
}

// Fill SweepArgs for a sweep over nfields fields.
template <class R>
SweepArgs<R> make_args(const GridInfo& g, const LineGeom& lg, const Sweep<R>& s, long long fstride,
                       int& nblk) {
  SweepArgs<R> a;
  std::memset(&a, 0, sizeof(a));
  a.orient = s.orient;
  a.sx = g.sx; a.sy = g.sy;
  a.fstride = fstride;
  if (s.orient == 0) { a.plan = lg.plan_x; a.tw = lg.tw_x; a.mk = lg.mk_x; }
  else { a.plan = lg.plan_y; a.tw = lg.tw_y; a.mk = lg.mk_y; }
  a.reg = sweep_uses_reg(g, lg, s.orient) ? 1 : 0;
  a.L = sweep_lines_per_block(g, lg, s.orient);
  nblk = sweep_blocks(g, lg, s.orient);
  if ((int)s.ops.size() > kMaxOps) throw std::runtime_error("too many line ops in one sweep");
  a.nops = (int)s.ops.size();
  for (int i = 0; i < a.nops; i++) a.ops[i] = s.ops[i];
  a.src = s.src; a.dst = s.dst; a.add1 = s.add1; a.vtab = s.vtab; a.vsrc = s.vsrc;
  a.dotsrc = s.dotsrc; a.partials = s.partials; a.nblk = nblk;
  return a;
}

// Per-member tables of a MultiProductSplit on the backend.
struct MemberDev {
  KinDev kin;
  const double* vhalf = nullptr;  // exp(-0.5 h V)                (potential_part)
  const double* vfull = nullptr;  // exp(-h V)                    (potential_part_square)
  const double* vpref = nullptr;  // c_t exp(-0.5 h V)            (potential_with_prefactor)
  const float* vhalf32 = nullptr;  // FP32 copies (or null)
  const float* vfull32 = nullptr;
  const float* vpref32 = nullptr;
};

// Sweep sequence of one SecondOrderSplit member with exponent t applied to
// `in`, result written to `out` (out = acc + result if accumulate). The
// operator list is [V_pref, K, (V_sq, K) x (t-1), V_half], applied in reverse
// (secondordersplit.cpp:21-49, operatorproduct.cpp:48-52). Row sweeps at the
// boundary between two kinetic applications fuse the inverse x half of one,
// the potential factor, and the forward x half of the next.
template <class R>
void member_program(Branch br, const MemberDev& m, int t, bool has_v, const cplx<R>* in, cplx<R>* work,
                    cplx<R>* out, bool accumulate, std::vector<Sweep<R>>& prog) {
  // head
  {
    Sweep<R> s;
    s.orient = 0;
    if (has_v) s.ops.push_back({LOP_MUL_TAB, m.vhalf, nullptr, 0, m.vhalf32, nullptr});
    fwd_x_ops(br, m.kin, s.ops);
    s.src = in; s.dst = work;
    prog.push_back(s);
  }
  for (int k = 1; k <= t; k++) {
    Sweep<R> c;
    c.orient = 1;
    col_ops(br, m.kin, c.ops);
    c.src = work; c.dst = work;
    prog.push_back(c);
    Sweep<R> r;
    r.orient = 0;
    inv_x_ops(br, m.kin, r.ops);
    if (k < t) {
      r.ops.push_back({LOP_MUL_TAB, m.vfull, nullptr, 0, m.vfull32, nullptr});
      fwd_x_ops(br, m.kin, r.ops);
      r.src = work; r.dst = work;
    } else {
      if (has_v) r.ops.push_back({LOP_MUL_TAB, m.vpref, nullptr, 0, m.vpref32, nullptr});
      r.src = work; r.dst = out;
      if (accumulate) r.add1 = out;
    }
    prog.push_back(r);
  }
}


// Run a list of sweeps over nfields fields of M elements each.
template <class R, class BE>
void run_program(BE& be, const GridInfo& g, const LineGeom& lg, const std::vector<Sweep<R>>& prog, int nfields) {
  for (const Sweep<R>& s : prog) {
    int nblk = 0;
    SweepArgs<R> a = make_args<R>(g, lg, s, g.M, nblk);
    be.sweep(a, nblk, nfields);
  }
}

// Transformer::transform kinds (transformer.hpp:132-158) as in-place sweeps.
template <class R>
bool transform_program(int kind, cplx<R>* f, std::vector<Sweep<R>>& prog) {
  int rop = LOP_NONE, cop = LOP_NONE;
  switch (kind) {
    case 0: rop = cop = LOP_FFT_FWD; break;    // FFT
    case 1: rop = cop = LOP_FFT_INV; break;    // iFFT
    case 2: rop = LOP_FFT_FWD; break;          // FFTx
    case 3: rop = LOP_FFT_INV; break;          // iFFTx
    case 4: cop = LOP_FFT_FWD; break;          // FFTy
    case 5: cop = LOP_FFT_INV; break;          // iFFTy
    case 6: rop = cop = LOP_DST2; break;       // DST
    case 7: rop = cop = LOP_DST3; break;       // iDST
    case 8: rop = LOP_DST2; break;             // DSTx
    case 9: rop = LOP_DST3; break;             // iDSTx
    case 10: cop = LOP_DST2; break;            // DSTy
    case 11: cop = LOP_DST3; break;            // iDSTy
    case 12: rop = cop = LOP_DCT2; break;      // DCT
    case 13: rop = cop = LOP_DCT3; break;      // iDCT
    case 14: rop = LOP_DCT2; break;            // DCTx
    case 15: rop = LOP_DCT3; break;            // iDCTx
    case 16: cop = LOP_DCT2; break;            // DCTy
    case 17: cop = LOP_DCT3; break;            // iDCTy
    default: return false;
  }
  if (rop != LOP_NONE) {
    Sweep<R> s; s.orient = 0; s.ops.push_back({rop, nullptr, nullptr, 0}); s.src = f; s.dst = f; prog.push_back(s);
  }
  if (cop != LOP_NONE) {
    Sweep<R> s; s.orient = 1; s.ops.push_back({cop, nullptr, nullptr, 0}); s.src = f; s.dst = f; prog.push_back(s);
  }
  return true;
}

// ExpKinetic alone (expkinetic.cpp:138-197): forward x half, column sweep,
// inverse x half; tables already hold coeff and prefactor.
template <class R>
void expkinetic_program(Branch br, const KinDev& k, cplx<R>* f, cplx<R>* work, std::vector<Sweep<R>>& prog) {
  Sweep<R> h; h.orient = 0; fwd_x_ops(br, k, h.ops); h.src = f; h.dst = work; prog.push_back(h);
  Sweep<R> c; c.orient = 1; col_ops(br, k, c.ops); c.src = work; c.dst = work; prog.push_back(c);
  Sweep<R> t; t.orient = 0; inv_x_ops(br, k, t.ops); t.src = work; t.dst = f; prog.push_back(t);
}

// ExpPotential alone (exppotential.hpp:56-65)
template <class R>
void exppotential_program(const double* values, double prefactor, cplx<R>* f, std::vector<Sweep<R>>& prog) {
  Sweep<R> s; s.orient = 0;
  if (values) s.ops.push_back({LOP_MUL_TAB, values, nullptr, 0});
  else if (prefactor != 1.0) s.ops.push_back({LOP_SCALE, nullptr, nullptr, prefactor});
  s.src = f; s.dst = f;
  prog.push_back(s);
}

// Kinetic tables on the backend.
struct KinHDev { const double* mxy = nullptr; const double* mx = nullptr; const double* mx2 = nullptr; const double* my = nullptr; };

// out = H in = T in + V in (kinetic.cpp:99-142, OperatorSum operatorsum.cpp:61-86),
// optionally with partial sums of <dotsrc|H in>. Returns the number of line
// blocks per field of the final sweep (the partials stride).
template <class R>
int hamiltonian_program(const GridInfo& g, const LineGeom& lg, Branch br, const KinHDev& k, const double* V,
                        const cplx<R>* in, cplx<R>* w1, cplx<R>* out, const cplx<R>* dotsrc,
                        cplx<double>* partials, std::vector<Sweep<R>>& prog) {
  if (br == BR_BP || br == BR_BD) {
    Sweep<R> r; r.orient = 0;
    if (br == BR_BP) {
      r.ops.push_back({LOP_FFT_FWD, nullptr, nullptr, 0});
      r.ops.push_back({LOP_MUL_TAB, k.mx, nullptr, 0});
      r.ops.push_back({LOP_FFT_INV, nullptr, nullptr, 0});
    } else {
      r.ops.push_back({LOP_DST2, nullptr, nullptr, 0});
      r.ops.push_back({LOP_XSPLIT, k.mx, k.mx2, 0});
    }
    r.src = in; r.dst = w1;
    prog.push_back(r);
    Sweep<R> c; c.orient = 1;
    const bool per = (br == BR_BP);
    c.ops.push_back({per ? LOP_FFT_FWD : LOP_DST2, nullptr, nullptr, 0});
    c.ops.push_back({LOP_MUL_POS, k.my, nullptr, 0});
    c.ops.push_back({per ? LOP_FFT_INV : LOP_DST3, nullptr, nullptr, 0});
    c.src = in; c.add1 = w1; c.vtab = V; c.vsrc = V ? in : nullptr;
    c.dotsrc = dotsrc; c.partials = partials; c.dst = out;
    prog.push_back(c);
    return sweep_blocks(g, lg, 1);
  }
  const bool per = (br == BR_B0P);
  Sweep<R> r1; r1.orient = 0;
  r1.ops.push_back({per ? LOP_FFT_FWD : LOP_DST2, nullptr, nullptr, 0});
  r1.src = in; r1.dst = w1;
  prog.push_back(r1);
  Sweep<R> c; c.orient = 1;
  c.ops.push_back({per ? LOP_FFT_FWD : LOP_DST2, nullptr, nullptr, 0});
  c.ops.push_back({LOP_MUL_TAB, k.mxy, nullptr, 0});
  c.ops.push_back({per ? LOP_FFT_INV : LOP_DST3, nullptr, nullptr, 0});
  c.src = w1; c.dst = w1;
  prog.push_back(c);
  Sweep<R> r2; r2.orient = 0;
  r2.ops.push_back({per ? LOP_FFT_INV : LOP_DST3, nullptr, nullptr, 0});
  r2.src = w1; r2.vtab = V; r2.vsrc = V ? in : nullptr;
  r2.dotsrc = dotsrc; r2.partials = partials; r2.dst = out;
  prog.push_back(r2);
  return sweep_blocks(g, lg, 0);
}

}  // namespace itp
