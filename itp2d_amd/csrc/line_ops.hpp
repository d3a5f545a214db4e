// Line-transform engine shared by the gfx950 sweep kernels and the CPU
// emulation used by the tests.
//
// A "sweep" loads L lines (rows along x, or columns along y) of one field into
// LDS, runs a short program of line operations on them (complex FFTs,
// RODFT10/RODFT01/REDFT01 real-to-real transforms, pointwise tables, the
// magnetic Dirichlet sine/cosine split) and writes the lines back, optionally
// accumulating, adding V*psi and forming <a|result> partial sums.
//
// Every step is a "pass": a function of (tid, nthreads) that touches LDS and
// global memory and is followed by a workgroup barrier. On the GPU a pass is
// run by all threads of the workgroup; the CPU emulator runs the same pass for
// tid = 0..nthreads-1 in turn, which is equivalent because no pass reads what
// another thread writes in the same pass.
//
// Transform conventions (FFTW's, which the reference uses,
// transformer.cpp:97-151):
//   FFT forward X_k = sum_j x_j e^{-2 pi i jk/n}, inverse with +, unnormalised
//   RODFT10 (DST-II)  y_k = 2 sum_j x_j sin(pi (2j+1)(k+1) / 2n)
//   RODFT01 (DST-III) y_k = (-1)^k x_{n-1} + 2 sum_{j<n-1} x_j sin(pi (j+1)(2k+1) / 2n)
//   REDFT01 (DCT-III) y_k = x_0 + 2 sum_{j>=1} x_j cos(pi j (2k+1) / 2n)
// applied to complex data (the reference applies them to Re and Im
// separately, transformer.cpp:116-144; the matrices are real, so this is the
// same). The real-to-real kinds are computed with one length-n complex FFT
// and O(n) pre/post passes (Makhoul's even/odd reordering, written here for
// complex input):
//   DCT-II:  v_m = x_{2m} (m < n/2), x_{2(n-1-m)+1} (m >= n/2); V = FFT(v);
//            Y_k = w_k V_k + conj(w_k) V_{(n-k) mod n},  w_k = e^{-i pi k / 2n}
//   DST-II:  DCT-II of (-1)^j x_j, output reversed
//   DCT-III: U_0 = X_0, U_k = w_k (X_k + i X_{n-k}); u = FFT(U);
//            x_{2m} = u_m, x_{2m+1} = u_{n-1-m}
//   DST-III: DCT-III of the reversed input, output times (-1)^j
#pragma once

#include <cmath>
#include <cstddef>
#include <cstdint>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define ITP_HD __host__ __device__ __forceinline__
#else
#define ITP_HD inline
#endif

namespace itp {

template <class R>
struct alignas(2 * sizeof(R)) cplx {
  R x, y;
};

template <class R> ITP_HD cplx<R> mk(R a, R b) { cplx<R> c; c.x = a; c.y = b; return c; }

// Workgroups are dealt to the 8 XCDs of the device round-robin in launch
// order. For kernels whose tiles share input slabs (the rotation's grid-point
// slabs, the overlap's grid chunks), launch-order block b maps to
// (inner, outer) so that, within each full group of 8 slabs, every XCD gets
// one whole slab (outer) and runs all tiles of it (inner) back to back,
// reusing the slab from its own L2. The slabs past the last full group are
// spread over all XCDs in plain order, so a few slabs never leave XCDs idle.
// The grid is exactly ninner * nouter blocks.
constexpr int kXcdCount = 8;
ITP_HD bool xcd_tile(long long b, int ninner, long long nouter, int& inner, long long& outer) {
  const long long full = nouter / kXcdCount * kXcdCount * ninner;
  if (b < full) {
    const long long xcd = b % kXcdCount, slot = b / kXcdCount;
    inner = (int)(slot % ninner);
    outer = (slot / ninner) * kXcdCount + xcd;
  } else {
    const long long r = b - full;
    inner = (int)(r % ninner);
    outer = nouter / kXcdCount * kXcdCount + r / ninner;
  }
  return outer < nouter;
}
inline long long xcd_grid(int ninner, long long nouter) { return nouter * ninner; }
template <class R> ITP_HD cplx<R> operator+(cplx<R> a, cplx<R> b) { return mk<R>(a.x + b.x, a.y + b.y); }
template <class R> ITP_HD cplx<R> operator-(cplx<R> a, cplx<R> b) { return mk<R>(a.x - b.x, a.y - b.y); }
template <class R> ITP_HD cplx<R> operator-(cplx<R> a) { return mk<R>(-a.x, -a.y); }
template <class R> ITP_HD cplx<R> operator*(cplx<R> a, cplx<R> b) {
  return mk<R>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <class R> ITP_HD cplx<R> operator*(cplx<R> a, R s) { return mk<R>(a.x * s, a.y * s); }
template <class R> ITP_HD cplx<R> conj(cplx<R> a) { return mk<R>(a.x, -a.y); }
// i * a
template <class R> ITP_HD cplx<R> mul_i(cplx<R> a) { return mk<R>(-a.y, a.x); }
// -i * a
template <class R> ITP_HD cplx<R> mul_ni(cplx<R> a) { return mk<R>(a.y, -a.x); }
// conj(a) * b
template <class R> ITP_HD cplx<R> cmulc(cplx<R> a, cplx<R> b) {
  return mk<R>(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}

// Line operation codes.
enum LineOpCode : int {
  LOP_NONE = 0,
  LOP_FFT_FWD,   // complex DFT, e^{-}
  LOP_FFT_INV,   // complex DFT, e^{+}
  LOP_DST2,      // RODFT10
  LOP_DST3,      // RODFT01
  LOP_DCT2,      // REDFT10
  LOP_DCT3,      // REDFT01
  LOP_MUL_TAB,   // buf *= t0[y*sx + x] (global coordinates of the element)
  LOP_MUL_POS,   // buf *= t0[position along the line]
  LOP_SCALE,     // buf *= s
  LOP_XSPLIT     // magnetic Dirichlet x-factor: buf <- RODFT01(t0 o buf) + REDFT01(shift(i t1 o buf))
};

struct LineOp {
  int code;
  const double* t0;
  const double* t1;
  double s;
  // FP32 copies of t0 and t1, read by compile-time FP32 sweep programs
  // (reg_sweep.hpp); null where no copy exists (aggregate initialisation
  // without them leaves them null)
  const float* f0;
  const float* f1;
};

constexpr int kMaxOps = 8;
constexpr int kMaxStages = 8;

// Radix plan for a length-n line DFT. nst is even so that a Stockham
// ping-pong ends in the buffer it started in.
struct LinePlan {
  int n;
  int nst;
  int radix[kMaxStages];
};

// Per-sweep arguments (by value into the kernel).
template <class R>
struct SweepArgs {
  int orient;               // 0: lines are rows (along x), 1: lines are columns (along y)
  int sx, sy;
  int L;                    // lines per workgroup
  int nops;
  long long fstride;        // elements between fields
  LinePlan plan;            // for the line length (sx for rows, sy for cols)
  LineOp ops[kMaxOps];
  const cplx<double>* tw;   // [n] e^{-2 pi i t/n}
  const cplx<double>* mk;   // [n] e^{-i pi k / 2n}
  const cplx<R>* src;
  cplx<R>* dst;             // may be null
  const cplx<R>* add1;      // result = add1 + result (may be null)
  const double* vtab;       // result = result + vtab o vsrc (may be null)
  const cplx<R>* vsrc;
  const cplx<R>* dotsrc;    // partial[field][block] = sum conj(dotsrc) * result (may be null)
  cplx<double>* partials;
  int nblk;                 // line blocks per field (partials stride)
  int reg;                  // 1: register-resident engine (reg_sweep.hpp), L = its lines per workgroup
};

// LDS layout for a sweep: three line buffers of L*(n+1) values, the two
// twiddle tables, and a reduction area.
template <class R>
struct SweepLds {
  cplx<R>* buf[3];
  cplx<R>* tw;
  cplx<R>* mk;
  cplx<double>* red;  // [nthreads + 16]: per-thread sums, then 16 group sums
  int ld;
};

template <class R>
ITP_HD size_t sweep_lds_bytes(int L, int n, int nthreads) {
  return (size_t)3 * L * (n + 1) * sizeof(cplx<R>) + (size_t)2 * n * sizeof(cplx<R>) +
         (size_t)(nthreads + 16) * sizeof(cplx<double>);
}

template <class R>
ITP_HD SweepLds<R> carve_lds(void* base, int L, int n, int nthreads) {
  SweepLds<R> s;
  s.ld = n + 1;
  cplx<R>* p = reinterpret_cast<cplx<R>*>(base);
  for (int b = 0; b < 3; b++) { s.buf[b] = p; p += (size_t)L * s.ld; }
  s.tw = p; p += n;
  s.mk = p; p += n;
  s.red = reinterpret_cast<cplx<double>*>(p);
  (void)nthreads;
  return s;
}

// ---------------------------------------------------------------------------
// DFT kernels of small radix, in registers. INV selects e^{+}.
// ---------------------------------------------------------------------------
template <class R, bool INV>
ITP_HD void dft2(cplx<R>* v) {
  cplx<R> a = v[0], b = v[1];
  v[0] = a + b; v[1] = a - b;
}

template <class R, bool INV>
ITP_HD void dft3(cplx<R>* v) {
  const R s3 = R(0.86602540378443864676);  // sqrt(3)/2
  cplx<R> t1 = v[1] + v[2];
  cplx<R> t2 = v[0] - t1 * R(0.5);
  cplx<R> t3 = (v[1] - v[2]) * s3;
  v[0] = v[0] + t1;
  if (!INV) { v[1] = t2 + mul_ni(t3); v[2] = t2 + mul_i(t3); }
  else      { v[1] = t2 + mul_i(t3);  v[2] = t2 + mul_ni(t3); }
}

template <class R, bool INV>
ITP_HD void dft4(cplx<R>* v) {
  cplx<R> t0 = v[0] + v[2], t1 = v[0] - v[2], t2 = v[1] + v[3], t3 = v[1] - v[3];
  v[0] = t0 + t2;
  v[2] = t0 - t2;
  if (!INV) { v[1] = t1 + mul_ni(t3); v[3] = t1 + mul_i(t3); }
  else      { v[1] = t1 + mul_i(t3);  v[3] = t1 + mul_ni(t3); }
}

template <class R, bool INV>
ITP_HD void dft5(cplx<R>* v) {
  const R c1 = R(0.30901699437494742410), c2 = R(-0.80901699437494742410);
  const R s1 = R(0.95105651629515357212), s2 = R(0.58778525229247312917);
  cplx<R> t1 = v[1] + v[4], t2 = v[2] + v[3], t3 = v[1] - v[4], t4 = v[2] - v[3];
  cplx<R> b1 = v[0] + t1 * c1 + t2 * c2;
  cplx<R> b2 = v[0] + t1 * c2 + t2 * c1;
  cplx<R> d1 = t3 * s1 + t4 * s2;
  cplx<R> d2 = t3 * s2 - t4 * s1;
  v[0] = v[0] + t1 + t2;
  if (!INV) {
    v[1] = b1 + mul_ni(d1); v[4] = b1 + mul_i(d1);
    v[2] = b2 + mul_ni(d2); v[3] = b2 + mul_i(d2);
  } else {
    v[1] = b1 + mul_i(d1); v[4] = b1 + mul_ni(d1);
    v[2] = b2 + mul_i(d2); v[3] = b2 + mul_ni(d2);
  }
}

template <class R, bool INV>
ITP_HD void dft8(cplx<R>* v) {
  const R h = R(0.70710678118654752440);
  cplx<R> e[4] = {v[0], v[2], v[4], v[6]};
  cplx<R> o[4] = {v[1], v[3], v[5], v[7]};
  dft4<R, INV>(e);
  dft4<R, INV>(o);
  // o_k *= w8^k, w8 = e^{-i pi/4} (forward) or e^{+i pi/4} (inverse)
  if (!INV) {
    o[1] = mk<R>((o[1].x + o[1].y) * h, (o[1].y - o[1].x) * h);
    o[2] = mul_ni(o[2]);
    o[3] = mk<R>((o[3].y - o[3].x) * h, -(o[3].x + o[3].y) * h);
  } else {
    o[1] = mk<R>((o[1].x - o[1].y) * h, (o[1].x + o[1].y) * h);
    o[2] = mul_i(o[2]);
    o[3] = mk<R>(-(o[3].x + o[3].y) * h, (o[3].x - o[3].y) * h);
  }
  for (int k = 0; k < 4; k++) { v[k] = e[k] + o[k]; v[k + 4] = e[k] - o[k]; }
}

template <class R, int RAD, bool INV>
ITP_HD void dft_small(cplx<R>* v) {
  if (RAD == 2) dft2<R, INV>(v);
  else if (RAD == 3) dft3<R, INV>(v);
  else if (RAD == 4) dft4<R, INV>(v);
  else if (RAD == 5) dft5<R, INV>(v);
  else dft8<R, INV>(v);
}

// One Stockham radix-RAD stage on L lines: src -> dst. Ns is the size of the
// sub-transforms already formed. Element j of the stage reads
// src[j + r n/RAD], twiddles by e^{-+2 pi i r (j mod Ns) / (Ns RAD)}, and writes
// dst[(j/Ns) Ns RAD + (j mod Ns) + r Ns].
template <class R, int RAD, bool INV>
ITP_HD void stockham_stage(int tid, int nt, const cplx<R>* src, cplx<R>* dst, int n, int L, int ld, int Ns,
                           const cplx<R>* tw) {
  const int nb = n / RAD;
  const int twstride = n / (Ns * RAD);
  for (int item = tid; item < L * nb; item += nt) {
    const int line = item / nb, j = item - line * nb;
    const cplx<R>* s = src + line * ld;
    cplx<R>* d = dst + line * ld;
    const int k = j % Ns;
    cplx<R> v[RAD];
    v[0] = s[j];
    for (int r = 1; r < RAD; r++) {
      cplx<R> w = tw[r * k * twstride];
      if (INV) w = conj(w);
      v[r] = s[j + r * nb] * w;
    }
    dft_small<R, RAD, INV>(v);
    const int base = (j / Ns) * Ns * RAD + k;
    for (int r = 0; r < RAD; r++) d[base + r * Ns] = v[r];
  }
}

template <class R, bool INV>
ITP_HD void stockham_dispatch(int tid, int nt, int rad, const cplx<R>* src, cplx<R>* dst, int n, int L, int ld,
                              int Ns, const cplx<R>* tw) {
  switch (rad) {
    case 2: stockham_stage<R, 2, INV>(tid, nt, src, dst, n, L, ld, Ns, tw); break;
    case 3: stockham_stage<R, 3, INV>(tid, nt, src, dst, n, L, ld, Ns, tw); break;
    case 4: stockham_stage<R, 4, INV>(tid, nt, src, dst, n, L, ld, Ns, tw); break;
    case 5: stockham_stage<R, 5, INV>(tid, nt, src, dst, n, L, ld, Ns, tw); break;
    default: stockham_stage<R, 8, INV>(tid, nt, src, dst, n, L, ld, Ns, tw); break;
  }
}

// Full line DFT as a sequence of passes; starts in `a`, ping-pongs with `b`,
// ends in `a` (plan.nst is even).
template <class R, class Exec>
ITP_HD void line_fft(Exec& ex, const LinePlan& plan, cplx<R>* a, cplx<R>* b, int L, int ld, const cplx<R>* tw,
                     bool inv) {
  int Ns = 1;
  cplx<R>* src = a;
  cplx<R>* dst = b;
  for (int s = 0; s < plan.nst; s++) {
    const int rad = plan.radix[s];
    const int n = plan.n;
    const int ns = Ns;
    cplx<R>* cs = src;
    cplx<R>* cd = dst;
    if (inv) ex.run([&](int tid, int nt) { stockham_dispatch<R, true>(tid, nt, rad, cs, cd, n, L, ld, ns, tw); });
    else     ex.run([&](int tid, int nt) { stockham_dispatch<R, false>(tid, nt, rad, cs, cd, n, L, ld, ns, tw); });
    Ns *= rad;
    cplx<R>* t = src; src = dst; dst = t;
  }
}

// Real-to-real transform of the lines in `a` (result in `a`), scratch `b`.
template <class R, class Exec>
ITP_HD void line_r2r(Exec& ex, int code, const LinePlan& plan, cplx<R>* a, cplx<R>* b, int L, int ld,
                     const cplx<R>* tw, const cplx<R>* mkt) {
  const int n = plan.n;
  const int half = n / 2;
  if (code == LOP_DCT2 || code == LOP_DST2) {
    const bool dst = (code == LOP_DST2);
    // pre: b[m] = x'[2m] (m < n/2), x'[2(n-1-m)+1] (m >= n/2), x'_j = (-1)^j x_j for DST
    ex.run([&](int tid, int nt) {
      for (int e = tid; e < L * n; e += nt) {
        const int l = e / n, m = e - l * n;
        const int j = (m < half) ? 2 * m : 2 * (n - 1 - m) + 1;
        cplx<R> v = a[l * ld + j];
        if (dst && (j & 1)) v = -v;
        b[l * ld + m] = v;
      }
    });
    line_fft<R>(ex, plan, b, a, L, ld, tw, false);
    // post: Y_k = w_k V_k + conj(w_k) V_{(n-k) mod n}; DST output reversed
    ex.run([&](int tid, int nt) {
      for (int e = tid; e < L * n; e += nt) {
        const int l = e / n, k = e - l * n;
        const cplx<R> w = mkt[k];
        const cplx<R> v1 = b[l * ld + k];
        const cplx<R> v2 = b[l * ld + ((n - k) % n)];
        const cplx<R> y = w * v1 + conj(w) * v2;
        a[l * ld + (dst ? (n - 1 - k) : k)] = y;
      }
    });
  } else {  // DCT3 / DST3
    const bool dst = (code == LOP_DST3);
    // pre: U_0 = X_0, U_k = w_k (X_k + i X_{n-k}); DST reads X reversed
    ex.run([&](int tid, int nt) {
      for (int e = tid; e < L * n; e += nt) {
        const int l = e / n, k = e - l * n;
        const cplx<R>* x = a + l * ld;
        cplx<R> u;
        if (k == 0) {
          u = dst ? x[n - 1] : x[0];
        } else {
          const cplx<R> xk = dst ? x[n - 1 - k] : x[k];
          const cplx<R> xnk = dst ? x[k - 1] : x[n - k];
          u = mkt[k] * (xk + mul_i(xnk));
        }
        b[l * ld + k] = u;
      }
    });
    line_fft<R>(ex, plan, b, a, L, ld, tw, false);
    // post: x_{2m} = u_m, x_{2m+1} = u_{n-1-m}; DST multiplies by (-1)^j
    ex.run([&](int tid, int nt) {
      for (int e = tid; e < L * n; e += nt) {
        const int l = e / n, j = e - l * n;
        const int m = (j & 1) ? (n - 1 - (j >> 1)) : (j >> 1);
        cplx<R> v = b[l * ld + m];
        if (dst && (j & 1)) v = -v;
        a[l * ld + j] = v;
      }
    });
  }
}

// Row index of line l and the global element offset of (line l, position p).
ITP_HD long long elem_offset(int orient, int sx, int line, int pos) {
  return orient == 0 ? (long long)line * sx + pos : (long long)pos * sx + line;
}

// Run one line operation on buf[0] (buf[1], buf[2] are scratch).
template <class R, class Exec>
ITP_HD void line_op(Exec& ex, const SweepArgs<R>& a, const SweepLds<R>& s, int line0, const LineOp& op) {
  const int n = a.plan.n, L = a.L, ld = s.ld;
  cplx<R>* b0 = s.buf[0];
  cplx<R>* b1 = s.buf[1];
  cplx<R>* b2 = s.buf[2];
  switch (op.code) {
    case LOP_FFT_FWD: line_fft<R>(ex, a.plan, b0, b2, L, ld, s.tw, false); break;
    case LOP_FFT_INV: line_fft<R>(ex, a.plan, b0, b2, L, ld, s.tw, true); break;
    case LOP_DST2: case LOP_DST3: case LOP_DCT2: case LOP_DCT3:
      line_r2r<R>(ex, op.code, a.plan, b0, b2, L, ld, s.tw, s.mk);
      break;
    case LOP_MUL_TAB: {
      const double* t = op.t0;
      const int orient = a.orient, sx = a.sx;
      ex.run([&](int tid, int nt) {
        for (int e = tid; e < L * n; e += nt) {
          const int l = e / n, p = e - l * n;
          b0[l * ld + p] = b0[l * ld + p] * R(t[elem_offset(orient, sx, line0 + l, p)]);
        }
      });
    } break;
    case LOP_MUL_POS: {
      const double* t = op.t0;
      ex.run([&](int tid, int nt) {
        for (int e = tid; e < L * n; e += nt) {
          const int l = e / n, p = e - l * n;
          b0[l * ld + p] = b0[l * ld + p] * R(t[p]);
        }
      });
    } break;
    case LOP_SCALE: {
      const R sc = R(op.s);
      ex.run([&](int tid, int nt) {
        for (int e = tid; e < L * n; e += nt) {
          const int l = e / n, p = e - l * n;
          b0[l * ld + p] = b0[l * ld + p] * sc;
        }
      });
    } break;
    case LOP_XSPLIT: {
      // rows only: b1[x] = i t1[y][x-1] b0[x-1] (b1[0] = 0, the top mode is
      // dropped, state.hpp:194-201); then b0 *= t0; b0 = RODFT01(b0) + REDFT01(b1)
      // (expkinetic.cpp:176-184; kinetic.cpp:128-134)
      const double* X = op.t0;
      const double* X2 = op.t1;
      const int sx = a.sx;
      ex.run([&](int tid, int nt) {
        for (int e = tid; e < L * n; e += nt) {
          const int l = e / n, x = e - l * n;
          const long long row = (long long)(line0 + l) * sx;
          cplx<R> v = mk<R>(R(0), R(0));
          if (x > 0) v = mul_i(b0[l * ld + x - 1] * R(X2[row + x - 1]));
          b1[l * ld + x] = v;
        }
      });
      ex.run([&](int tid, int nt) {
        for (int e = tid; e < L * n; e += nt) {
          const int l = e / n, x = e - l * n;
          const long long row = (long long)(line0 + l) * sx;
          b0[l * ld + x] = b0[l * ld + x] * R(X[row + x]);
        }
      });
      line_r2r<R>(ex, LOP_DST3, a.plan, b0, b2, L, ld, s.tw, s.mk);
      line_r2r<R>(ex, LOP_DCT3, a.plan, b1, b2, L, ld, s.tw, s.mk);
      ex.run([&](int tid, int nt) {
        for (int e = tid; e < L * n; e += nt) {
          const int l = e / n, x = e - l * n;
          b0[l * ld + x] = b0[l * ld + x] + b1[l * ld + x];
        }
      });
    } break;
    default: break;
  }
}

// The whole sweep for lines [line0, line0+L) of field `field`, block index blk.
template <class R, class Exec>
ITP_HD void sweep_program(Exec& ex, const SweepArgs<R>& a, const SweepLds<R>& s, int field, int blk) {
  const int n = a.plan.n, L = a.L, ld = s.ld;
  const int line0 = blk * L;
  const int orient = a.orient, sx = a.sx;
  const long long fo = (long long)field * a.fstride;
  cplx<R>* b0 = s.buf[0];
  // load lines and twiddles. Row lines: consecutive threads walk along the
  // line; column lines: consecutive threads walk across the L lines so each
  // row segment of L values is one contiguous access.
  {
    const cplx<R>* src = a.src + fo;
    const cplx<double>* gtw = a.tw;
    const cplx<double>* gmk = a.mk;
    cplx<R>* tw = s.tw;
    cplx<R>* mkt = s.mk;
    ex.run([&](int tid, int nt) {
      for (int e = tid; e < L * n; e += nt) {
        int l, p;
        if (orient == 0) { l = e / n; p = e - l * n; } else { p = e / L; l = e - p * L; }
        b0[l * ld + p] = src[elem_offset(orient, sx, line0 + l, p)];
      }
      for (int k = tid; k < n; k += nt) {
        tw[k] = mk<R>(R(gtw[k].x), R(gtw[k].y));
        mkt[k] = mk<R>(R(gmk[k].x), R(gmk[k].y));
      }
    });
  }
  for (int o = 0; o < a.nops; o++) line_op<R>(ex, a, s, line0, a.ops[o]);
  // store with epilogue
  {
    cplx<R>* dst = a.dst ? a.dst + fo : nullptr;
    const cplx<R>* add1 = a.add1 ? a.add1 + fo : nullptr;
    const double* vtab = a.vtab;
    const cplx<R>* vsrc = a.vsrc ? a.vsrc + fo : nullptr;
    const cplx<R>* dsrc = a.dotsrc ? a.dotsrc + fo : nullptr;
    cplx<double>* red = s.red;
    ex.run([&](int tid, int nt) {
      cplx<double> acc = mk<double>(0.0, 0.0);
      for (int e = tid; e < L * n; e += nt) {
        int l, p;
        if (orient == 0) { l = e / n; p = e - l * n; } else { p = e / L; l = e - p * L; }
        const long long g = elem_offset(orient, sx, line0 + l, p);
        cplx<R> r = b0[l * ld + p];
        if (add1) r = add1[g] + r;
        if (vtab) r = r + vsrc[g] * R(vtab[g]);
        if (dsrc) {
          const cplx<R> d = dsrc[g];
          acc.x += double(d.x) * double(r.x) + double(d.y) * double(r.y);
          acc.y += double(d.x) * double(r.y) - double(d.y) * double(r.x);
        }
        if (dst) dst[g] = r;
      }
      red[tid] = acc;
    });
    if (a.dotsrc) {
      // two-level reduction in a fixed order: 16 partial sums, then one
      ex.run([&](int tid, int nt) {
        const int groups = 16;
        const int per = (nt + groups - 1) / groups;
        if (tid < groups) {
          cplx<double> sacc = mk<double>(0.0, 0.0);
          for (int i = tid * per; i < (tid + 1) * per && i < nt; i++) sacc = sacc + red[i];
          red[nt + tid] = sacc;
        }
      });
      cplx<double>* part = a.partials;
      const int nblk = a.nblk;
      ex.run([&](int tid, int nt) {
        if (tid == 0) {
          cplx<double> sacc = mk<double>(0.0, 0.0);
          for (int i = 0; i < 16; i++) sacc = sacc + red[nt + i];
          part[(long long)field * nblk + blk] = sacc;
        }
      });
    }
  }
}

// Real part of the principal complex square root of (a + i b).
ITP_HD double csqrt_re(double a, double b) {
  if (a == 0 && b == 0) return 0;
  const double r = std::hypot(a, b);
  if (a >= 0) return std::sqrt(0.5 * (r + a));
  return std::fabs(b) / std::sqrt(2.0 * (r - a));
}

// E and sigma of one state from its two partial-sum rows
// (operators.cpp:40-50: mean = <psi|H psi>, sd = sqrt(<psi|H H psi> - mean^2)).
ITP_HD void energy_from_partials(const cplx<double>* p1, const cplx<double>* p2, int nblk1, int nblk2, double dx,
                                 double* E, double* sigma, cplx<double>* mean_out, cplx<double>* msq_out) {
  cplx<double> s1 = mk<double>(0.0, 0.0), s2 = mk<double>(0.0, 0.0);
  for (int b = 0; b < nblk1; b++) s1 = s1 + p1[b];
  for (int b = 0; b < nblk2; b++) s2 = s2 + p2[b];
  const cplx<double> mean = (s1 * dx) * dx;
  const cplx<double> msq = (s2 * dx) * dx;
  const cplx<double> m2 = mean * mean;
  *E = mean.x;
  *sigma = csqrt_re(msq.x - m2.x, msq.y - m2.y);
  if (mean_out) *mean_out = mean;
  if (msq_out) *msq_out = msq;
}

}  // namespace itp
