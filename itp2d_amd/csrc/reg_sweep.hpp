// Register-resident line engine for power-of-two line lengths (32..1024).
//
// Same sweep contract as sweep_program (line_ops.hpp): a workgroup takes a
// block of whole lines of one field, runs the line-operation program and
// writes the lines back with the accumulate / V*psi / dot-product epilogue.
// The difference is where the lines live: each line is held by P threads, Q
// elements each (n = P*Q), thread t holding positions t + P*q. FFTs are
// Stockham stages of radix 16/8/4/2 computed in registers; only between two
// stages does data go through LDS (one write, one read), so a length-256
// FFT costs one LDS exchange. The Makhoul pre- and post-permutations of the
// real-to-real kinds are LDS exchanges as well, except that a real-to-real
// forward transform at the start of the program gathers its permuted input
// straight from global memory, and an inverse one at the end scatters its
// output straight to global memory.
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

constexpr int kRegThreads = 256;

ITP_HD bool reg_supported(int n) {
  return n == 32 || n == 64 || n == 128 || n == 256 || n == 512 || n == 1024;
}
// Lines per workgroup for a line length (kRegThreads / P).
ITP_HD int reg_lines_per_wg(int n) {
  switch (n) {
    case 32: return 64;
    case 64: return 32;
    case 128: return 32;
    case 256: return 16;
    case 512: return 8;
    case 1024: return 4;
    default: return 0;
  }
}
// LDS line stride: one pad element per 16 (stage writes of stride 16 fall in
// distinct banks) plus one, so consecutive lines start in distinct banks.
ITP_HD int reg_ld(int n) { return n + n / 16 + 1; }
template <class R>
ITP_HD size_t reg_lds_bytes(int n) {
  const size_t lines = (size_t)reg_lines_per_wg(n) * reg_ld(n) * sizeof(cplx<R>);
  const size_t red = (size_t)(kRegThreads + 16) * sizeof(cplx<double>);
  return lines > red ? lines : red;
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

template <class R, int RAD, bool INV>
ITP_HD void dft_reg(cplx<R>* v) {
  if constexpr (RAD == 2) dft2<R, INV>(v);
  else if constexpr (RAD == 4) dft4<R, INV>(v);
  else if constexpr (RAD == 8) dft8<R, INV>(v);
  else dft16<R, INV>(v);
}

// One Stockham stage on the thread's registers A (positions t + P q).
// Butterfly j = t + P b reads positions j + r n/RAD (registers b + r Q/RAD),
// twiddles by w_n^{r (j mod NS) n/(NS RAD)}, and writes positions
// (j/NS) NS RAD + (j mod NS) + r NS: to LDS, or, in the last stage
// (NS RAD = n), back to registers b + r Q/RAD in natural order.
template <class R, int N, int RAD, int NS, bool INV, bool LAST>
ITP_HD void reg_stage(cplx<R>* A, int t, cplx<R>* line, const cplx<double>* tw) {
  constexpr int P = RegShape<N>::P, Q = RegShape<N>::Q;
  constexpr int NB = Q / RAD;
  constexpr int TWS = N / (NS * RAD);
#pragma unroll
  for (int b = 0; b < NB; b++) {
    const int j = t + P * b;
    const int k = opaque(j & (NS - 1));
    cplx<R> x[RAD];
#pragma unroll
    for (int r = 0; r < RAD; r++) x[r] = A[b + r * NB];
    if constexpr (NS > 1) {
#pragma unroll
      for (int r = 1; r < RAD; r++) {
        const cplx<double> w = tw[r * k * TWS];
        x[r] = x[r] * mk<R>(R(w.x), INV ? R(-w.y) : R(w.y));
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
template <class G, bool INV, int WHICH, class R, class Exec>
ITP_HD void reg_fft(Exec& ex, cplx<R>* lds, const cplx<double>* tw) {
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

// Write registers A to LDS in natural order, then A[q] = f(t, l, pos, line).
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

// Makhoul input permutation of DCT-II / DST-II: position m of the FFT input
// holds x_{j(m)}; DST-II also negates odd j.
ITP_HD int r2f_src(int n, int m) { return (m < n / 2) ? 2 * m : 2 * (n - 1 - m) + 1; }
// Makhoul output permutation of DCT-III / DST-III: FFT output position m
// lands at x_{j(m)} (the same map); DST-III negates odd j.
ITP_HD int r2i_dst(int n, int m) { return (m < n / 2) ? 2 * m : 2 * (n - 1 - m) + 1; }

ITP_HD bool lop_elementwise(int c) { return c == LOP_MUL_TAB || c == LOP_MUL_POS || c == LOP_SCALE; }
ITP_HD bool lop_r2f(int c) { return c == LOP_DCT2 || c == LOP_DST2; }
ITP_HD bool lop_r2i(int c) { return c == LOP_DCT3 || c == LOP_DST3; }

// Post step of DCT-II / DST-II: Y_k = w_k V_k + conj(w_k) V_{(n-k) mod n},
// DST-II output reversed.
template <class G, bool DST, int WHICH, class R, class Exec>
ITP_HD void reg_r2f_post(Exec& ex, cplx<R>* lds, const cplx<double>* mkt) {
  constexpr int n = G::n;
  reg_exchange<G, WHICH>(ex, lds, [&](int, int, int p, const cplx<R>* line) {
    const int k = DST ? n - 1 - p : p;
    const cplx<double> wd = mkt[opaque(k)];
    const cplx<R> w = mk<R>(R(wd.x), R(wd.y));
    return w * line[lpad(k)] + conj(w) * line[lpad((n - k) & (n - 1))];
  });
}

// Pre step of DCT-III / DST-III: U_0 = X_0, U_k = w_k (X_k + i X_{n-k});
// DST-III reads X reversed.
template <class G, bool DST, int WHICH, class R, class Exec>
ITP_HD void reg_r2i_pre(Exec& ex, cplx<R>* lds, const cplx<double>* mkt) {
  constexpr int n = G::n;
  reg_exchange<G, WHICH>(ex, lds, [&](int, int, int k, const cplx<R>* line) {
    if (k == 0) return line[lpad(DST ? n - 1 : 0)];
    const cplx<R> xk = line[lpad(DST ? n - 1 - k : k)];
    const cplx<R> xnk = line[lpad(DST ? k - 1 : n - k)];
    const cplx<double> wd = mkt[opaque(k)];
    return mk<R>(R(wd.x), R(wd.y)) * (xk + mul_i(xnk));
  });
}

template <class G, int CODE, int WHICH, class R, class Exec>
ITP_HD void reg_r2r(Exec& ex, cplx<R>* lds, const cplx<double>* tw, const cplx<double>* mkt) {
  constexpr int n = G::n;
  constexpr bool DST = (CODE == LOP_DST2 || CODE == LOP_DST3);
  if constexpr (CODE == LOP_DCT2 || CODE == LOP_DST2) {
    reg_exchange<G, WHICH>(ex, lds, [&](int, int, int m, const cplx<R>* line) {
      const int j = r2f_src(n, m);
      const cplx<R> v = line[lpad(j)];
      return (DST && (j & 1)) ? -v : v;
    });
    reg_fft<G, false, WHICH>(ex, lds, tw);
    reg_r2f_post<G, DST, WHICH>(ex, lds, mkt);
  } else {
    reg_r2i_pre<G, DST, WHICH>(ex, lds, mkt);
    reg_fft<G, false, WHICH>(ex, lds, tw);
    reg_exchange<G, WHICH>(ex, lds, [&](int, int, int j, const cplx<R>* line) {
      const int m = (j & 1) ? (n - 1 - (j >> 1)) : (j >> 1);
      const cplx<R> v = line[lpad(m)];
      return (DST && (j & 1)) ? -v : v;
    });
  }
}

template <class G, int WHICH, class R, class Exec>
ITP_HD void reg_r2r_dispatch(Exec& ex, int code, cplx<R>* lds, const cplx<double>* tw, const cplx<double>* mkt) {
  switch (code) {
    case LOP_DST2: reg_r2r<G, LOP_DST2, WHICH>(ex, lds, tw, mkt); break;
    case LOP_DST3: reg_r2r<G, LOP_DST3, WHICH>(ex, lds, tw, mkt); break;
    case LOP_DCT2: reg_r2r<G, LOP_DCT2, WHICH>(ex, lds, tw, mkt); break;
    default: reg_r2r<G, LOP_DCT3, WHICH>(ex, lds, tw, mkt); break;
  }
}

// Elementwise op at global line `line` and position `pos`.
template <class R>
ITP_HD cplx<R> reg_elem(const LineOp& op, cplx<R> v, long long g, int pos) {
  if (op.code == LOP_MUL_TAB) return v * R(op.t0[g]);
  if (op.code == LOP_MUL_POS) return v * R(op.t0[pos]);
  return v * R(op.s);
}

// The whole sweep for line block blk of field `field`.
template <class R, int N, int ORIENT, class Exec>
ITP_HD void reg_sweep_program(Exec& ex, const SweepArgs<R>& a, cplx<R>* lds, int field, int blk) {
  using G = RegGeom<N, ORIENT>;
  constexpr int P = G::P, Q = G::Q, n = N;
  const int nops = a.nops;
  const int sx = a.sx;
  const int line0 = blk * G::LPW;
  const long long fo = (long long)field * a.fstride;
  auto goff = [&](int l, int pos) -> long long {
    return ORIENT == 0 ? (long long)(line0 + l) * sx + pos : (long long)pos * sx + (line0 + l);
  };
  // leading elementwise ops, then possibly a forward r2r whose input
  // permutation is taken from global memory
  int lead = 0;
  while (lead < nops && lop_elementwise(a.ops[lead].code)) lead++;
  const bool fold_in = lead < nops && lop_r2f(a.ops[lead].code);
  const bool in_dst = fold_in && a.ops[lead].code == LOP_DST2;
  const int start = fold_in ? lead + 1 : lead;
  // trailing elementwise ops (not already taken by the load), and possibly an
  // inverse r2r before them whose output permutation goes straight to global
  // memory
  int trail = nops;
  while (trail > start && lop_elementwise(a.ops[trail - 1].code)) trail--;
  const bool fold_out = trail > start && lop_r2i(a.ops[trail - 1].code);
  const bool out_dst = fold_out && a.ops[trail - 1].code == LOP_DST3;

  // load
  {
    const cplx<R>* src = a.src + fo;
    ex.local([&](int tid, auto& st) {
      const int t = G::t_of(tid), l = G::l_of(tid);
#pragma unroll
      for (int q = 0; q < Q; q++) {
        const int m = t + P * q;
        const int pos = fold_in ? r2f_src(n, m) : m;
        const long long g = goff(l, pos);
        cplx<R> v = src[g];
        for (int i = 0; i < lead; i++) v = reg_elem<R>(a.ops[i], v, g, pos);
        if (in_dst && (pos & 1)) v = -v;
        st.v[q] = v;
      }
    });
  }
  if (fold_in) {
    reg_fft<G, false, 0>(ex, lds, a.tw);
    if (in_dst) reg_r2f_post<G, true, 0>(ex, lds, a.mk);
    else reg_r2f_post<G, false, 0>(ex, lds, a.mk);
  }
  const int stop = fold_out ? trail - 1 : trail;
  for (int o = start; o < stop; o++) {
    const LineOp op = a.ops[o];
    switch (op.code) {
      case LOP_FFT_FWD: reg_fft<G, false, 0>(ex, lds, a.tw); break;
      case LOP_FFT_INV: reg_fft<G, true, 0>(ex, lds, a.tw); break;
      case LOP_DST2: case LOP_DST3: case LOP_DCT2: case LOP_DCT3:
        reg_r2r_dispatch<G, 0>(ex, op.code, lds, a.tw, a.mk);
        break;
      case LOP_MUL_TAB: case LOP_MUL_POS: case LOP_SCALE:
        ex.local([&](int tid, auto& st) {
          const int t = G::t_of(tid), l = G::l_of(tid);
#pragma unroll
          for (int q = 0; q < Q; q++) {
            const int pos = t + P * q;
            st.v[q] = reg_elem<R>(op, st.v[q], goff(l, pos), pos);
          }
        });
        break;
      case LOP_XSPLIT: {
        // rows: w[x] = i X2[y][x-1] v[x-1] (w[0] = 0), v *= X, then
        // v = RODFT01(v) + REDFT01(w) (expkinetic.cpp:176-184, kinetic.cpp:128-134)
        const double* X = op.t0;
        const double* X2 = op.t1;
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
            const int x = t + P * q;
            cplx<R> wv = mk<R>(R(0), R(0));
            if (x > 0) wv = mul_i(line[lpad(x - 1)] * R(X2[row + x - 1]));
            st.w[q] = wv;
            st.v[q] = st.v[q] * R(X[row + x]);
          }
        });
        reg_r2r<G, LOP_DST3, 0>(ex, lds, a.tw, a.mk);
        reg_r2r<G, LOP_DCT3, 1>(ex, lds, a.tw, a.mk);
        ex.local([&](int, auto& st) {
#pragma unroll
          for (int q = 0; q < Q; q++) st.v[q] = st.v[q] + st.w[q];
        });
      } break;
      default: break;
    }
  }
  if (fold_out) {
    if (out_dst) reg_r2i_pre<G, true, 0>(ex, lds, a.mk);
    else reg_r2i_pre<G, false, 0>(ex, lds, a.mk);
    reg_fft<G, false, 0>(ex, lds, a.tw);
  }
  // store with epilogue; the dot partial of each thread goes to LDS slot tid
  // (every LDS read of the program ended at a barrier)
  {
    cplx<R>* dst = a.dst ? a.dst + fo : nullptr;
    const cplx<R>* add1 = a.add1 ? a.add1 + fo : nullptr;
    const double* vtab = a.vtab;
    const cplx<R>* vsrc = a.vsrc ? a.vsrc + fo : nullptr;
    const cplx<R>* dsrc = a.dotsrc ? a.dotsrc + fo : nullptr;
    cplx<double>* red = reinterpret_cast<cplx<double>*>(lds);
    ex.run([&](int tid, auto& st) {
      const int t = G::t_of(tid), l = G::l_of(tid);
      cplx<double> acc = mk<double>(0.0, 0.0);
#pragma unroll
      for (int q = 0; q < Q; q++) {
        const int m = t + P * q;
        const int pos = fold_out ? r2i_dst(n, m) : m;
        const long long g = goff(l, pos);
        cplx<R> r = st.v[q];
        if (out_dst && (pos & 1)) r = -r;
        for (int i = trail; i < nops; i++) r = reg_elem<R>(a.ops[i], r, g, pos);
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

// Calls f(integral_constant<N>, integral_constant<ORIENT>) for a supported
// line length; false otherwise.
template <class F>
inline bool reg_dispatch(int n, int orient, F&& f) {
  auto go = [&](auto nc) {
    if (orient == 0) f(nc, std::integral_constant<int, 0>{});
    else f(nc, std::integral_constant<int, 1>{});
    return true;
  };
  switch (n) {
    case 32: return go(std::integral_constant<int, 32>{});
    case 64: return go(std::integral_constant<int, 64>{});
    case 128: return go(std::integral_constant<int, 128>{});
    case 256: return go(std::integral_constant<int, 256>{});
    case 512: return go(std::integral_constant<int, 512>{});
    case 1024: return go(std::integral_constant<int, 1024>{});
    default: return false;
  }
}

}  // namespace itp
