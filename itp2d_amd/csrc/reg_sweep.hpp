// Register-resident line engine for the line lengths 32..1024 that are powers
// of two, and 192.
//
// Same sweep contract as sweep_program (line_ops.hpp): a workgroup takes a
// block of whole lines of one field, runs the line-operation program and
// writes the lines back with the accumulate / V*psi / dot-product epilogue.
// The difference is where the lines live: each line is held by P threads, Q
// elements each (n = P*Q), thread t holding positions t + P*q. FFTs are
// Stockham stages of radix 16/8/4/2 computed in registers; only between two
// stages does data go through LDS (one write, one read), so a length-256
// FFT costs one LDS exchange. The Makhoul pre- and post-steps of the
// real-to-real kinds are gathers: a post-step is left pending and folded into
// the next exchange (or the store), so a real-to-real transform costs one
// exchange besides its FFT; a forward one at the start of the program gathers
// its permuted input straight from global memory.
//
// Thread layout: row sweeps put the P threads of a line on consecutive lanes
// (global accesses run along the row); column sweeps put consecutive lines
// on consecutive lanes (global accesses run along a row segment spanning the
// workgroup's columns).
//
// Every step is either a register-only step (ex.local) or a pass followed by
// a workgroup barrier (ex.run); a pass never reads LDS written in the same
// pass, so the CPU emulator may run the threads of a pass one after another.
#pragma once

#include <type_traits>
#include <utility>

#include "line_ops.hpp"

namespace itp {

// Shape of a line: threads per line P, elements per thread Q, Stockham
// radices R0, R1, R2 (NST stages). Every radix divides Q, which makes each
// stage read registers b + r*Q/R of the thread's own positions.
template <int N> struct RegShape;
template <> struct RegShape<32>   { static constexpr int P = 4,  Q = 8,  NST = 2, R0 = 8,  R1 = 4,  R2 = 1; };
template <> struct RegShape<64>   { static constexpr int P = 8,  Q = 8,  NST = 2, R0 = 8,  R1 = 8,  R2 = 1; };
template <> struct RegShape<128>  { static constexpr int P = 8,  Q = 16, NST = 2, R0 = 16, R1 = 8,  R2 = 1; };
template <> struct RegShape<256>  { static constexpr int P = 16, Q = 16, NST = 2, R0 = 16, R1 = 16, R2 = 1; };
template <> struct RegShape<512>  { static constexpr int P = 32, Q = 16, NST = 3, R0 = 16, R1 = 16, R2 = 2; };
template <> struct RegShape<1024> { static constexpr int P = 64, Q = 16, NST = 3, R0 = 16, R1 = 16, R2 = 4; };
// 192 = 12 * 4 * 4: twelve elements per thread, radix 12 (3 x 4) then two radix-4 stages
template <> struct RegShape<192>  { static constexpr int P = 16, Q = 12, NST = 3, R0 = 12, R1 = 4,  R2 = 4; };

constexpr int kRegThreads = 256;

ITP_HD bool reg_supported(int n) {
  return n == 32 || n == 64 || n == 128 || n == 192 || n == 256 || n == 512 || n == 1024;
}
// Lines per workgroup for a line length (kRegThreads / P).
ITP_HD constexpr int reg_lines_per_wg(int n) {
  switch (n) {
    case 32: return 64;
    case 64: return 32;
    case 128: return 32;
    case 192: return 16;
    case 256: return 16;
    case 512: return 8;
    case 1024: return 4;
    default: return 0;
  }
}
// LDS line stride: one pad element per 16 (stage writes of stride 16 fall in
// distinct banks) plus one, so consecutive lines start in distinct banks.
ITP_HD constexpr int reg_ld(int n) { return n + n / 16 + 1; }
// LDS of a sweep: the line buffers (or the dot-product reduction area, if
// larger), then, when that still leaves room for two workgroups per CU, the
// line's twiddle and Makhoul tables in the working precision (otherwise they
// are read from global memory).
template <class R>
ITP_HD constexpr size_t reg_lds_lines_bytes(int n) {
  const size_t lines = (size_t)reg_lines_per_wg(n) * reg_ld(n) * sizeof(cplx<R>);
  const size_t red = (size_t)(kRegThreads + 16) * sizeof(cplx<double>);
  return lines > red ? lines : red;
}
template <class R>
ITP_HD constexpr bool reg_tabs_in_lds(int n) {
  return reg_lds_lines_bytes<R>(n) + (size_t)2 * n * sizeof(cplx<R>) <= 80 * 1024;
}
template <class R>
ITP_HD constexpr size_t reg_lds_bytes(int n) {
  return reg_lds_lines_bytes<R>(n) + (reg_tabs_in_lds<R>(n) ? (size_t)2 * n * sizeof(cplx<R>) : 0);
}
// A twiddle or Makhoul factor in the working precision, from either table kind.
template <class R, class T>
ITP_HD cplx<R> tab_val(const cplx<T>* t, int i) {
  const cplx<T> v = t[i];
  return mk<R>(R(v.x), R(v.y));
}

ITP_HD int lpad(int i) { return i + (i >> 4); }

// Hides a value from the optimiser so that loads indexed by it are not
// hoisted out of the program's op loop (each hoisted table load would pin
// registers for the whole kernel).
ITP_HD int opaque(int x) {
#if defined(__HIP_DEVICE_COMPILE__)
  asm volatile("" : "+v"(x));
#endif
  return x;
}

// Scheduling fence: keeps the compiler from hoisting the loads of later
// elements above it, which bounds the registers in flight in a gather loop.
ITP_HD void sched_fence() {
#if defined(__HIP_DEVICE_COMPILE__)
  __builtin_amdgcn_sched_barrier(0);
#endif
}

template <class R, int Q>
struct RegState {
  cplx<R> v[Q];
  cplx<R> w[Q];
};

template <int N, int ORIENT>
struct RegGeom {
  static constexpr int n = N;
  static constexpr int P = RegShape<N>::P;
  static constexpr int Q = RegShape<N>::Q;
  static constexpr int LPW = kRegThreads / P;
  static constexpr int LD = N + N / 16 + 1;
  ITP_HD static int t_of(int tid) { return ORIENT == 0 ? tid % P : tid / LPW; }
  ITP_HD static int l_of(int tid) { return ORIENT == 0 ? tid / P : tid % LPW; }
};

// ---------------------------------------------------------------------------
// radix 16 in registers: x[m], m = 4 m1 + m2  ->  X[k1 + 4 k2]
// ---------------------------------------------------------------------------
template <class R, bool INV>
ITP_HD cplx<R> w16_mul(cplx<R> a, int e) {
  // a * e^{-+2 pi i e/16} for the exponents a radix-16 split needs
  const R c1 = R(0.92387953251128675613), s1 = R(0.38268343236508977173);
  const R h = R(0.70710678118654752440);
  cplx<R> w;
  switch (e) {
    case 0: return a;
    case 4: return INV ? mul_i(a) : mul_ni(a);
    case 1: w = mk<R>(c1, -s1); break;
    case 2: w = mk<R>(h, -h); break;
    case 3: w = mk<R>(s1, -c1); break;
    case 6: w = mk<R>(-h, -h); break;
    case 9: w = mk<R>(-c1, s1); break;
    default: w = mk<R>(R(0), R(0)); break;  // not reached by dft16
  }
  if (INV) w.y = -w.y;
  return a * w;
}

template <class R, bool INV>
ITP_HD void dft16(cplx<R>* x) {
  cplx<R> y[4][4];
#pragma unroll
  for (int m2 = 0; m2 < 4; m2++) {
    cplx<R> a[4] = {x[m2], x[4 + m2], x[8 + m2], x[12 + m2]};
    dft4<R, INV>(a);
#pragma unroll
    for (int k1 = 0; k1 < 4; k1++) y[m2][k1] = w16_mul<R, INV>(a[k1], m2 * k1);
  }
#pragma unroll
  for (int k1 = 0; k1 < 4; k1++) {
    cplx<R> a[4] = {y[0][k1], y[1][k1], y[2][k1], y[3][k1]};
    dft4<R, INV>(a);
#pragma unroll
    for (int k2 = 0; k2 < 4; k2++) x[k1 + 4 * k2] = a[k2];
  }
}

// ---------------------------------------------------------------------------
// radix 12 in registers: x[m], m = 4 m1 + m2 (m1 < 3, m2 < 4)  ->  X[k1 + 3 k2]
// ---------------------------------------------------------------------------
template <class R, bool INV>
ITP_HD cplx<R> w12_mul(cplx<R> a, int e) {
  // a * e^{-+2 pi i e/12} for the exponents m2 k1 of the 3 x 4 split
  const R s3 = R(0.86602540378443864676);
  cplx<R> w;
  switch (e) {
    case 0: return a;
    case 3: return INV ? mul_i(a) : mul_ni(a);
    case 6: return -a;
    case 1: w = mk<R>(s3, R(-0.5)); break;
    case 2: w = mk<R>(R(0.5), -s3); break;
    case 4: w = mk<R>(R(-0.5), -s3); break;
    default: w = mk<R>(R(0), R(0)); break;  // not reached by dft12
  }
  if (INV) w.y = -w.y;
  return a * w;
}

template <class R, bool INV>
ITP_HD void dft12(cplx<R>* x) {
  cplx<R> y[4][3];
#pragma unroll
  for (int m2 = 0; m2 < 4; m2++) {
    cplx<R> a[3] = {x[m2], x[4 + m2], x[8 + m2]};
    dft3<R, INV>(a);
#pragma unroll
    for (int k1 = 0; k1 < 3; k1++) y[m2][k1] = w12_mul<R, INV>(a[k1], m2 * k1);
  }
#pragma unroll
  for (int k1 = 0; k1 < 3; k1++) {
    cplx<R> a[4] = {y[0][k1], y[1][k1], y[2][k1], y[3][k1]};
    dft4<R, INV>(a);
#pragma unroll
    for (int k2 = 0; k2 < 4; k2++) x[k1 + 3 * k2] = a[k2];
  }
}

template <class R, int RAD, bool INV>
ITP_HD void dft_reg(cplx<R>* v) {
  if constexpr (RAD == 2) dft2<R, INV>(v);
  else if constexpr (RAD == 4) dft4<R, INV>(v);
  else if constexpr (RAD == 8) dft8<R, INV>(v);
  else if constexpr (RAD == 12) dft12<R, INV>(v);
  else dft16<R, INV>(v);
}

// One Stockham stage on the thread's registers A (positions t + P q).
// Butterfly j = t + P b reads positions j + r n/RAD (registers b + r Q/RAD),
// twiddles by w_n^{r (j mod NS) n/(NS RAD)}, and writes positions
// (j/NS) NS RAD + (j mod NS) + r NS: to LDS, or, in the last stage
// (NS RAD = n), back to registers b + r Q/RAD in natural order.
template <class R, int N, int RAD, int NS, bool INV, bool LAST, class TW>
ITP_HD void reg_stage(cplx<R>* A, int t, cplx<R>* line, const TW* tw) {
  constexpr int P = RegShape<N>::P, Q = RegShape<N>::Q;
  constexpr int NB = Q / RAD;
  constexpr int TWS = N / (NS * RAD);
#pragma unroll
  for (int b = 0; b < NB; b++) {
    const int j = t + P * b;
    const int k = opaque(j % NS);
    cplx<R> x[RAD];
#pragma unroll
    for (int r = 0; r < RAD; r++) x[r] = A[b + r * NB];
    if constexpr (NS > 1) {
#pragma unroll
      for (int r = 1; r < RAD; r++) {
        const cplx<R> w = tab_val<R>(tw, r * k * TWS);
        x[r] = x[r] * (INV ? conj(w) : w);
      }
    }
    dft_reg<R, RAD, INV>(x);
    if constexpr (LAST) {
#pragma unroll
      for (int r = 0; r < RAD; r++) A[b + r * NB] = x[r];
    } else {
      const int base = (j / NS) * NS * RAD + k;
#pragma unroll
      for (int r = 0; r < RAD; r++) line[lpad(base + r * NS)] = x[r];
    }
  }
}

template <int WHICH, class R, int Q>
ITP_HD cplx<R>* reg_arr(RegState<R, Q>& st) { return WHICH ? st.w : st.v; }

// Natural-order LDS read back into registers after a scattered stage write.
template <class G, int WHICH, class R, class Exec>
ITP_HD void reg_read_natural(Exec& ex, cplx<R>* lds) {
  ex.run([&](int tid, auto& st) {
    const int t = G::t_of(tid), l = G::l_of(tid);
    const cplx<R>* line = lds + l * G::LD;
    cplx<R>* A = reg_arr<WHICH>(st);
#pragma unroll
    for (int q = 0; q < G::Q; q++) A[q] = line[lpad(t + G::P * q)];
  });
}

// Complex DFT of every line (forward e^{-}, or inverse e^{+}), unnormalised.
template <class G, bool INV, int WHICH, class R, class Exec, class TW>
ITP_HD void reg_fft(Exec& ex, cplx<R>* lds, const TW* tw) {
  using S = RegShape<G::n>;
  constexpr int N = G::n;
  if constexpr (S::NST == 2) {
    ex.run([&](int tid, auto& st) {
      reg_stage<R, N, S::R0, 1, INV, false>(reg_arr<WHICH>(st), G::t_of(tid), lds + G::l_of(tid) * G::LD, tw);
    });
    reg_read_natural<G, WHICH>(ex, lds);
    ex.local([&](int tid, auto& st) {
      reg_stage<R, N, S::R1, S::R0, INV, true>(reg_arr<WHICH>(st), G::t_of(tid), nullptr, tw);
    });
  } else {
    ex.run([&](int tid, auto& st) {
      reg_stage<R, N, S::R0, 1, INV, false>(reg_arr<WHICH>(st), G::t_of(tid), lds + G::l_of(tid) * G::LD, tw);
    });
    reg_read_natural<G, WHICH>(ex, lds);
    ex.run([&](int tid, auto& st) {
      reg_stage<R, N, S::R1, S::R0, INV, false>(reg_arr<WHICH>(st), G::t_of(tid), lds + G::l_of(tid) * G::LD, tw);
    });
    reg_read_natural<G, WHICH>(ex, lds);
    ex.local([&](int tid, auto& st) {
      reg_stage<R, N, S::R2, S::R0 * S::R1, INV, true>(reg_arr<WHICH>(st), G::t_of(tid), nullptr, tw);
    });
  }
}

// Makhoul input permutation of DCT-II / DST-II: position m of the FFT input
// holds x_{j(m)}; DST-II also negates odd j.
ITP_HD int r2f_src(int n, int m) { return (m < n / 2) ? 2 * m : 2 * (n - 1 - m) + 1; }
// Makhoul output permutation of DCT-III / DST-III: FFT output position m
// lands at x_{j(m)} (the same map); DST-III negates odd j.
ITP_HD int r2i_dst(int n, int m) { return (m < n / 2) ? 2 * m : 2 * (n - 1 - m) + 1; }

ITP_HD bool lop_elementwise(int c) { return c == LOP_MUL_TAB || c == LOP_MUL_POS || c == LOP_SCALE; }
ITP_HD bool lop_r2f(int c) { return c == LOP_DCT2 || c == LOP_DST2; }
ITP_HD bool lop_r2i(int c) { return c == LOP_DCT3 || c == LOP_DST3; }

// ---------------------------------------------------------------------------
// Pending post-steps. The last step of a real-to-real transform is a gather
// (a pair combination for DCT-II / DST-II, a permutation for DCT-III /
// DST-III), and so is the first step of the next one: rather than spend an
// LDS exchange on each, the registers keep the FFT output L and a note of how
// the logical line values P(p) derive from it. The next exchange reads P(p)
// straight from L.
//   PK_NONE      P(p) = L[p]
//   PK_R2F_DCT   P(p) = f(p) (w_k L[k] + conj(w_k) L[(n-k) mod n]), k = p    (REDFT10 post)
//   PK_R2F_DST   the same with k = n-1-p                                      (RODFT10 post)
//   PK_PERM_DCT  P(p) = L[m(p)], m(p) = p/2 (p even), n-1-(p-1)/2 (odd)      (REDFT01 post)
//   PK_PERM_DST  P(p) = (-1)^p L[m(p)]                                        (RODFT01 post)
// f(p) is at most one pending elementwise factor; elementwise ops met with
// nothing or a permutation pending are applied to the registers at once.
// The kind is a compile-time tag at every use, so the unrolled element loops
// carry no branches on it.
// ---------------------------------------------------------------------------
enum RegPend { PK_NONE = 0, PK_R2F_DCT, PK_R2F_DST, PK_PERM_DCT, PK_PERM_DST };

template <int K, bool E>
struct PendTag {
  static constexpr int kind = K;
  static constexpr bool elem = E;
};

struct RegPending {
  int kind;
  int el;  // index of the pending elementwise op, or -1
};

ITP_HD int perm_src(int n, int p) { return (p & 1) ? (n - 1 - (p >> 1)) : (p >> 1); }

// Factor of an elementwise op with code `code`, table t and scale s at
// global offset g and line position pos (MUL_TAB: t[g], MUL_POS: t[pos],
// SCALE: s).
template <class R, class TT>
ITP_HD R elem_factor(int code, const TT* t, double s, long long g, int pos) {
  if (code == LOP_MUL_TAB) return R(t[g]);
  if (code == LOP_MUL_POS) return R(t[pos]);
  return R(s);
}

// A sweep program known at compile time: PG::n op codes in PG::c. The
// runtime op list (SweepArgs::ops) then only supplies the tables; the
// kernel's op loop unrolls into straight-line code, and FP32 programs read
// the FP32 copies of the tables (LineOp::f0, f1). RtProg: the op list is
// read at run time.
struct RtProg {
  static constexpr int n = -1;
  static constexpr int c[kMaxOps] = {};
};
template <int... C>
struct CtProg {
  static constexpr int n = sizeof...(C);
  static constexpr int c[kMaxOps] = {C...};
};
// Index of the first op that is not elementwise.
template <class PG>
ITP_HD constexpr int ct_lead() {
  int i = 0;
  while (i < PG::n && (PG::c[i] == LOP_MUL_TAB || PG::c[i] == LOP_MUL_POS || PG::c[i] == LOP_SCALE)) i++;
  return i;
}
// Does the compile-time program match the runtime op list?
template <class PG>
ITP_HD bool ct_matches(const LineOp* ops, int nops) {
  if (PG::n != nops) return false;
  for (int i = 0; i < nops; i++)
    if (ops[i].code != PG::c[i]) return false;
  return true;
}
template <class F, int... I>
ITP_HD void static_for(F&& f, std::integer_sequence<int, I...>) {
  (f(std::integral_constant<int, I>{}), ...);
}

// P(p) for a pending kind known at compile time. fg, fp: the pending factor's
// table strides in (global offset, position); ft its table.
template <class TAG, class R, int N, class MK, class TT>
ITP_HD cplx<R> pend_value(const cplx<R>* line, int p, long long g, const MK* mkt, const TT* ft,
                          long long fsg, long long fsp) {
  constexpr int n = N;
  constexpr int K = TAG::kind;
  cplx<R> v;
  if constexpr (K == PK_R2F_DCT || K == PK_R2F_DST) {
    const int k = K == PK_R2F_DST ? n - 1 - p : p;
    const cplx<R> w = tab_val<R>(mkt, opaque(k));
    v = w * line[lpad(k)] + conj(w) * line[lpad(k == 0 ? 0 : n - k)];
    if constexpr (TAG::elem) v = v * R(ft[g * fsg + p * fsp]);
  } else if constexpr (K == PK_PERM_DCT || K == PK_PERM_DST) {
    v = line[lpad(perm_src(n, p))];
    if (K == PK_PERM_DST && (p & 1)) v = -v;
  } else {
    v = line[lpad(p)];
  }
  return v;
}

// Calls f(PendTag<kind, elem>{}) for the runtime pending state.
template <class F>
ITP_HD void with_pend(const RegPending& pd, F&& f) {
  const bool e = pd.el >= 0;
  switch (pd.kind) {
    case PK_R2F_DCT:
      if (e) f(PendTag<PK_R2F_DCT, true>{});
      else f(PendTag<PK_R2F_DCT, false>{});
      break;
    case PK_R2F_DST:
      if (e) f(PendTag<PK_R2F_DST, true>{});
      else f(PendTag<PK_R2F_DST, false>{});
      break;
    case PK_PERM_DCT: f(PendTag<PK_PERM_DCT, false>{}); break;
    case PK_PERM_DST: f(PendTag<PK_PERM_DST, false>{}); break;
    default: f(PendTag<PK_NONE, false>{}); break;
  }
}

// Write registers A (the pending L) to LDS in natural order, then
// A[q] = f(t, l, position t + P q, line). Two barriers.
template <class G, int WHICH, class R, class Exec, class F>
ITP_HD void reg_exchange(Exec& ex, cplx<R>* lds, F&& f) {
  ex.run([&](int tid, auto& st) {
    const int t = G::t_of(tid), l = G::l_of(tid);
    cplx<R>* line = lds + l * G::LD;
    const cplx<R>* A = reg_arr<WHICH>(st);
#pragma unroll
    for (int q = 0; q < G::Q; q++) line[lpad(t + G::P * q)] = A[q];
  });
  ex.run([&](int tid, auto& st) {
    const int t = G::t_of(tid), l = G::l_of(tid);
    const cplx<R>* line = lds + l * G::LD;
    cplx<R>* A = reg_arr<WHICH>(st);
#pragma unroll
    for (int q = 0; q < G::Q; q++) A[q] = f(t, l, t + G::P * q, line);
  });
}

// The whole sweep for line block blk of field `field`. XS: the program may
// contain the magnetic Dirichlet split (rows only); without it the code path
// and its second register set are compiled out.
template <class R, int N, int ORIENT, bool XS, class PG = RtProg, class Exec>
ITP_HD void reg_sweep_program(Exec& ex, const SweepArgs<R>& a, cplx<R>* lds, int field, int blk) {
  using G = RegGeom<N, ORIENT>;
  constexpr int P = G::P, Q = G::Q, n = N;
  constexpr bool CT = PG::n >= 0;
  // tables of elementwise ops: FP32 copies for compile-time FP32 programs
  constexpr bool F32T = CT && sizeof(R) == 4;
  using TT = std::conditional_t<F32T, float, double>;
  const int nops = CT ? PG::n : a.nops;
  const int sx = a.sx;
  const int line0 = blk * G::LPW;
  const long long fo = (long long)field * a.fstride;
  const LineOp* ops = a.ops;
  // twiddles and Makhoul factors, staged in LDS where they fit (every later
  // read of them follows at least one workgroup barrier)
  constexpr bool TL = reg_tabs_in_lds<R>(N);
  using TabT = std::conditional_t<TL, cplx<R>, cplx<double>>;
  const TabT* ltw;
  const TabT* mkt;
  if constexpr (TL) {
    cplx<R>* t = lds + reg_lds_lines_bytes<R>(N) / sizeof(cplx<R>);
    cplx<R>* m = t + N;
    ex.local([&](int tid, auto&) {
      for (int i = tid; i < N; i += kRegThreads) {
        t[i] = tab_val<R>(a.tw, i);
        m[i] = tab_val<R>(a.mk, i);
      }
    });
    ltw = t;
    mkt = m;
  } else {
    ltw = a.tw;
    mkt = a.mk;
  }
  auto goff = [&](int l, int pos) -> long long {
    return ORIENT == 0 ? (long long)(line0 + l) * sx + pos : (long long)pos * sx + (line0 + l);
  };
  auto code_at = [&](int i) -> int {
    if constexpr (CT) return PG::c[i];
    else return ops[i].code;
  };
  auto tab0 = [&](const LineOp& op) -> const TT* {
    if constexpr (F32T) return op.f0;
    else return op.t0;
  };
  auto tab1 = [&](const LineOp& op) -> const TT* {
    if constexpr (F32T) return op.f1;
    else return op.t1;
  };
  // leading elementwise ops, then possibly a forward r2r whose input
  // permutation is taken from global memory
  int lead = 0;
  if constexpr (CT) lead = ct_lead<PG>();
  else
    while (lead < nops && lop_elementwise(code_at(lead))) lead++;
  const bool fold_in = lead < nops && lop_r2f(code_at(lead));
  const bool in_dst = fold_in && code_at(lead) == LOP_DST2;

  // load
  {
    const cplx<R>* src = a.src + fo;
    auto load = [&](auto fi) {
      constexpr bool FI = decltype(fi)::value;
      ex.local([&](int tid, auto& st) {
        const int t = G::t_of(tid), l = G::l_of(tid);
#pragma unroll
        for (int q = 0; q < Q; q++) {
          const int m = t + P * q;
          const int pos = FI ? r2f_src(n, m) : m;
          cplx<R> v = src[goff(l, pos)];
          if (FI && in_dst && (pos & 1)) v = -v;
          st.v[q] = v;
        }
      });
      for (int i = 0; i < lead; i++) {
        const LineOp& op = ops[i];
        const int code = code_at(i);
        const TT* t0 = tab0(op);
        ex.local([&](int tid, auto& st) {
          const int t = G::t_of(tid), l = G::l_of(tid);
#pragma unroll
          for (int q = 0; q < Q; q++) {
            const int m = t + P * q;
            const int pos = FI ? r2f_src(n, m) : m;
            st.v[q] = st.v[q] * elem_factor<R>(code, t0, op.s, goff(l, pos), pos);
          }
        });
      }
    };
    if (fold_in) load(std::integral_constant<bool, true>{});
    else load(std::integral_constant<bool, false>{});
  }
  RegPending pd = {PK_NONE, -1};
  const int start = fold_in ? lead + 1 : lead;
  if (fold_in) {
    reg_fft<G, false, 0>(ex, lds, ltw);
    pd.kind = in_dst ? PK_R2F_DST : PK_R2F_DCT;
  }
  // the pending factor's table and strides
  auto ftab = [&](const RegPending& pc, const TT*& ft, long long& fsg, long long& fsp) {
    ft = nullptr; fsg = 0; fsp = 0;
    if (pc.el >= 0) {
      const int c = code_at(pc.el);
      ft = tab0(ops[pc.el]);
      fsg = c == LOP_MUL_TAB ? 1 : 0;
      fsp = c == LOP_MUL_POS ? 1 : 0;
    }
  };
  // bring registers to natural order (one exchange if a post-step is pending)
  auto materialize = [&]() {
    if (pd.kind == PK_NONE) return;
    const RegPending pc = pd;
    const TT* ft; long long fsg, fsp;
    ftab(pc, ft, fsg, fsp);
    with_pend(pc, [&](auto tag) {
      reg_exchange<G, 0>(ex, lds, [&](int, int l, int p, const cplx<R>* line) {
        return pend_value<decltype(tag), R, N>(line, p, goff(l, p), mkt, ft, fsg, fsp);
      });
    });
    pd.kind = PK_NONE;
    pd.el = -1;
  };
  // one op of the program (for compile-time programs o is a constant after
  // inlining, so every branch on the op code and the pending state folds)
  auto do_op = [&](int o) __attribute__((always_inline)) {
    const LineOp& op = ops[o];
    const int code = code_at(o);
    if (lop_elementwise(code)) {
      if ((pd.kind == PK_R2F_DCT || pd.kind == PK_R2F_DST) && code != LOP_SCALE) {
        if (pd.el >= 0) materialize();  // at most one pending factor
        else { pd.el = o; return; }
      }
      // nothing or a permutation pending: apply at the logical position now
      const bool perm = pd.kind == PK_PERM_DCT || pd.kind == PK_PERM_DST;
      const TT* t0 = tab0(op);
      ex.local([&](int tid, auto& st) {
        const int t = G::t_of(tid), l = G::l_of(tid);
#pragma unroll
        for (int q = 0; q < Q; q++) {
          const int m = t + P * q;
          const int pos = perm ? r2i_dst(n, m) : m;
          st.v[q] = st.v[q] * elem_factor<R>(code, t0, op.s, goff(l, pos), pos);
        }
      });
      return;
    }
    if (code == LOP_FFT_FWD || code == LOP_FFT_INV) {
      materialize();
      if (code == LOP_FFT_FWD) reg_fft<G, false, 0>(ex, lds, ltw);
      else reg_fft<G, true, 0>(ex, lds, ltw);
      return;
    }
    const RegPending pc = pd;
    const TT* ft; long long fsg, fsp;
    ftab(pc, ft, fsg, fsp);
    if (lop_r2f(code)) {
      // DCT-II / DST-II: FFT input m holds +-P(j(m)) (Makhoul's reordering)
      const bool dst = code == LOP_DST2;
      with_pend(pc, [&](auto tag) {
        using T = decltype(tag);
        reg_exchange<G, 0>(ex, lds, [&](int, int l, int m, const cplx<R>* line) {
          const int j = r2f_src(n, m);
          const cplx<R> v = pend_value<T, R, N>(line, j, goff(l, j), mkt, ft, fsg, fsp);
          return (dst && (j & 1)) ? -v : v;
        });
      });
      reg_fft<G, false, 0>(ex, lds, ltw);
      pd.kind = dst ? PK_R2F_DST : PK_R2F_DCT;
    } else if (lop_r2i(code)) {
      // DCT-III / DST-III: U_0 = X_0, U_k = w_k (X_k + i X_{n-k}); DST-III reads X reversed
      const bool dst = code == LOP_DST3;
      with_pend(pc, [&](auto tag) {
        using T = decltype(tag);
        reg_exchange<G, 0>(ex, lds, [&](int, int l, int k, const cplx<R>* line) {
          const int pa = dst ? n - 1 - k : k;
          const int pb = k == 0 ? pa : (dst ? k - 1 : n - k);
          const cplx<R> xa = pend_value<T, R, N>(line, pa, goff(l, pa), mkt, ft, fsg, fsp);
          const cplx<R> xb = pend_value<T, R, N>(line, pb, goff(l, pb), mkt, ft, fsg, fsp);
          const cplx<R> u = tab_val<R>(mkt, opaque(k)) * (xa + mul_i(xb));
          return k == 0 ? xa : u;
        });
      });
      reg_fft<G, false, 0>(ex, lds, ltw);
      pd.kind = dst ? PK_PERM_DST : PK_PERM_DCT;
    } else if (code == LOP_XSPLIT) {
      if constexpr (XS) {
      // rows: with b = P, v_x = X[y][x] b_x and w_x = i X2[y][x-1] b_{x-1} (w_0 = 0),
      // the result RODFT01(v) + REDFT01(w) (expkinetic.cpp:176-184,
      // kinetic.cpp:128-134). Both pre-steps need only b_{n-1-k} and b_{k-1}:
      //   U^v_k = w_k (v_{n-1-k} + i v_{k-1}),  U^v_0 = v_{n-1}
      //   U^w_k = w_k (w_k + i w_{n-k}),        U^w_0 = 0
      const TT* X = tab0(op);
      const TT* X2 = tab1(op);
      with_pend(pc, [&](auto tag) {
        using T = decltype(tag);
        ex.run([&](int tid, auto& st) {
          const int t = G::t_of(tid), l = G::l_of(tid);
          cplx<R>* line = lds + l * G::LD;
#pragma unroll
          for (int q = 0; q < Q; q++) line[lpad(t + P * q)] = st.v[q];
        });
        ex.run([&](int tid, auto& st) {
          const int t = G::t_of(tid), l = G::l_of(tid);
          const cplx<R>* line = lds + l * G::LD;
          const long long row = (long long)(line0 + l) * sx;
#pragma unroll
          for (int q = 0; q < Q; q++) {
            if (q % 4 == 0) sched_fence();
            const int k = t + P * q;
            const int pa = n - 1 - k;
            const int pb = k == 0 ? pa : k - 1;
            const cplx<R> bpa = pend_value<T, R, N>(line, pa, goff(l, pa), mkt, ft, fsg, fsp);
            const cplx<R> bb = pend_value<T, R, N>(line, pb, goff(l, pb), mkt, ft, fsg, fsp);
            const cplx<R> ba = bpa * R(X[row + pa]);
            const cplx<R> w = tab_val<R>(mkt, opaque(k));
            const cplx<R> uv = w * (ba + mul_i(bb * R(X[row + pb])));
            const cplx<R> uw = w * (mul_i(bb * R(X2[row + pb])) - bpa * R(X2[row + pa]));
            st.v[q] = k == 0 ? ba : uv;
            st.w[q] = k == 0 ? mk<R>(R(0), R(0)) : uw;
          }
        });
      });
      reg_fft<G, false, 0>(ex, lds, ltw);
      reg_fft<G, false, 1>(ex, lds, ltw);
      // combine before the shared output permutation: DST-III's (-1)^j sign
      // falls on the FFT outputs m >= n/2
      ex.local([&](int tid, auto& st) {
        const int t = G::t_of(tid);
#pragma unroll
        for (int q = 0; q < Q; q++) {
          const int m = t + P * q;
          st.v[q] = (m < n / 2) ? st.w[q] + st.v[q] : st.w[q] - st.v[q];
        }
      });
      pd.kind = PK_PERM_DCT;
      }
    }
    pd.el = -1;
  };
  if constexpr (CT) {
    static_for([&](auto oc) __attribute__((always_inline)) {
      constexpr int o = decltype(oc)::value;
      if (o >= start) do_op(o);
    }, std::make_integer_sequence<int, (PG::n > 0 ? PG::n : 0)>{});
  } else {
    for (int o = start; o < nops; o++) do_op(o);
  }
  // a pending pair combination needs one more exchange; a pending
  // permutation is applied by the store itself
  if (pd.kind == PK_R2F_DCT || pd.kind == PK_R2F_DST) materialize();
  // store with epilogue; the dot partial of each thread goes to LDS slot tid
  // (every LDS read of the program ended at a barrier)
  {
    cplx<R>* dst = a.dst ? a.dst + fo : nullptr;
    const cplx<R>* add1 = a.add1 ? a.add1 + fo : nullptr;
    const double* vtab = a.vtab;
    const cplx<R>* vsrc = a.vsrc ? a.vsrc + fo : nullptr;
    const cplx<R>* dsrc = a.dotsrc ? a.dotsrc + fo : nullptr;
    cplx<double>* red = reinterpret_cast<cplx<double>*>(lds);
    auto store = [&](auto tag) {
      constexpr int K = decltype(tag)::kind;
      constexpr bool PERM = K == PK_PERM_DCT || K == PK_PERM_DST;
      ex.run([&](int tid, auto& st) {
        const int t = G::t_of(tid), l = G::l_of(tid);
        cplx<double> acc = mk<double>(0.0, 0.0);
#pragma unroll
        for (int q = 0; q < Q; q++) {
          const int m = t + P * q;
          const int pos = PERM ? r2i_dst(n, m) : m;
          const long long g = goff(l, pos);
          cplx<R> r = st.v[q];
          if (K == PK_PERM_DST && (pos & 1)) r = -r;
          if (add1) r = add1[g] + r;
          if (vtab) r = r + vsrc[g] * R(vtab[g]);
          if (dsrc) {
            const cplx<R> d = dsrc[g];
            acc.x += double(d.x) * double(r.x) + double(d.y) * double(r.y);
            acc.y += double(d.x) * double(r.y) - double(d.y) * double(r.x);
          }
          if (dst) dst[g] = r;
        }
        if (dsrc) red[tid] = acc;
      });
    };
    with_pend(pd, store);
    if (a.dotsrc) {
      ex.run([&](int tid, auto&) {
        constexpr int groups = 16, per = kRegThreads / groups;
        if (tid < groups) {
          cplx<double> s = mk<double>(0.0, 0.0);
          for (int i = tid * per; i < (tid + 1) * per; i++) s = s + red[i];
          red[kRegThreads + tid] = s;
        }
      });
      cplx<double>* part = a.partials;
      const int nblk = a.nblk;
      ex.run([&](int tid, auto&) {
        if (tid == 0) {
          cplx<double> s = mk<double>(0.0, 0.0);
          for (int i = 0; i < 16; i++) s = s + red[kRegThreads + i];
          part[(long long)field * nblk + blk] = s;
        }
      });
    }
  }
}

// Does the op list use the magnetic Dirichlet split?
ITP_HD bool ops_have_xsplit(const LineOp* ops, int nops) {
  for (int i = 0; i < nops; i++)
    if (ops[i].code == LOP_XSPLIT) return true;
  return false;
}

// Calls f(integral_constant<N>, integral_constant<ORIENT>, integral_constant<XS>)
// for a supported line length (XS only with rows); false otherwise.
template <class F>
inline bool reg_dispatch(int n, int orient, bool xs, F&& f) {
  auto go = [&](auto nc) {
    if (orient == 0) {
      if (xs) f(nc, std::integral_constant<int, 0>{}, std::integral_constant<bool, true>{});
      else f(nc, std::integral_constant<int, 0>{}, std::integral_constant<bool, false>{});
    } else {
      if (xs) return false;
      f(nc, std::integral_constant<int, 1>{}, std::integral_constant<bool, false>{});
    }
    return true;
  };
  switch (n) {
    case 32: return go(std::integral_constant<int, 32>{});
    case 64: return go(std::integral_constant<int, 64>{});
    case 128: return go(std::integral_constant<int, 128>{});
    case 192: return go(std::integral_constant<int, 192>{});
    case 256: return go(std::integral_constant<int, 256>{});
    case 512: return go(std::integral_constant<int, 512>{});
    case 1024: return go(std::integral_constant<int, 1024>{});
    default: return false;
  }
}

}  // namespace itp
