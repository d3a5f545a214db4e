// FP64 CPU oracle for the itp2d ITP loop. TEST INFRASTRUCTURE ONLY (see the
// header): the product never links or calls this file.
//
// Every block cites the reference lines (paths under src/) it restates.
#include "itp2d_oracle.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <tuple>
#include <vector>
#include <omp.h>

namespace {

typedef std::complex<double> cd;
const double PI = M_PI;

// ---------------------------------------------------------------------------
// 1D complex DFT, unnormalised, X[k] = sum_j x[j] exp(sign*2*pi*i*j*k/n).
// Recursive mixed-radix decimation in time; twiddles come from one table of
// exact angles per length. FFTW_FORWARD is sign = -1 (transformer.cpp:97-106).
// ---------------------------------------------------------------------------
struct FFTPlan {
  size_t n;
  std::vector<cd> tw;        // exp(-2 pi i t / n)
  std::vector<size_t> factors;
};

const FFTPlan& plan_for(size_t n) {
  static std::mutex mu;
  static std::map<size_t, std::unique_ptr<FFTPlan>> cache;
  std::lock_guard<std::mutex> lk(mu);
  auto it = cache.find(n);
  if (it != cache.end()) return *it->second;
  std::unique_ptr<FFTPlan> p(new FFTPlan);
  p->n = n;
  p->tw.resize(n);
  for (size_t t = 0; t < n; t++) p->tw[t] = std::polar(1.0, -2.0 * PI * double(t) / double(n));
  size_t m = n;
  for (size_t f : {4, 2, 3, 5}) {
    while (m % f == 0 && !(f == 4 && m == 4 && false)) { p->factors.push_back(f); m /= f; }
  }
  for (size_t f = 7; m > 1; f += 2)
    while (m % f == 0) { p->factors.push_back(f); m /= f; }
  const FFTPlan& ref = *p;
  cache[n] = std::move(p);
  return ref;
}

// out[k] = sum_{j<n} in[j*is] w_n^{jk}, w_n = exp(sign 2 pi i / n); n = P.n / twstep
void fft_rec(const cd* in, size_t is, cd* out, size_t n, size_t twstep, const FFTPlan& P,
             size_t fi, int sign) {
  if (n == 1) { out[0] = in[0]; return; }
  const size_t p = P.factors[fi], m = n / p;
  for (size_t r = 0; r < p; r++) fft_rec(in + r * is, is * p, out + r * m, m, twstep * p, P, fi + 1, sign);
  cd t[16];
  std::vector<cd> big;
  cd* tp = t;
  if (p > 16) { big.resize(p); tp = big.data(); }
  const size_t pstep = P.n / p;
  for (size_t k = 0; k < m; k++) {
    for (size_t r = 0; r < p; r++) {
      cd w = P.tw[(r * k * twstep) % P.n];
      if (sign > 0) w = std::conj(w);
      tp[r] = out[r * m + k] * w;
    }
    for (size_t q = 0; q < p; q++) {
      cd s = 0;
      for (size_t r = 0; r < p; r++) {
        cd w = P.tw[((r * q) % p) * pstep];
        if (sign > 0) w = std::conj(w);
        s += tp[r] * w;
      }
      out[q * m + k] = s;
    }
  }
}

void fft_inplace(cd* x, size_t n, size_t stride, int sign) {
  if (n <= 1) return;
  const FFTPlan& P = plan_for(n);
  std::vector<cd> in(n), out(n);
  for (size_t j = 0; j < n; j++) in[j] = x[j * stride];
  fft_rec(in.data(), 1, out.data(), n, 1, P, 0, sign);
  for (size_t j = 0; j < n; j++) x[j * stride] = out[j];
}

// ---------------------------------------------------------------------------
// Real-to-real transforms with FFTW's definitions (RODFT10 = DST-II,
// RODFT01 = DST-III, REDFT10 = DCT-II, REDFT01 = DCT-III), applied to complex
// data. The reference runs them on the real and imaginary parts separately
// (transformer.cpp:116-144); these are real matrices, so applying them to the
// complex vector is the same thing. Computed through a length-2n complex DFT.
// ---------------------------------------------------------------------------
void r2r_inplace(cd* x, size_t n, size_t stride, int kind) {
  const size_t n2 = 2 * n;
  std::vector<cd> z(n2);
  const cd I(0, 1);
  switch (kind) {
    case ORC_REDFT10:  // y_k = 2 sum x_j cos(pi (2j+1) k / 2n)
      for (size_t j = 0; j < n; j++) { z[j] = x[j * stride]; z[n2 - 1 - j] = x[j * stride]; }
      fft_inplace(z.data(), n2, 1, -1);
      for (size_t k = 0; k < n; k++) x[k * stride] = std::polar(1.0, -PI * double(k) / double(n2)) * z[k];
      break;
    case ORC_RODFT10:  // y_k = 2 sum x_j sin(pi (2j+1)(k+1) / 2n)
      for (size_t j = 0; j < n; j++) { z[j] = x[j * stride]; z[n2 - 1 - j] = -x[j * stride]; }
      fft_inplace(z.data(), n2, 1, -1);
      for (size_t k = 0; k < n; k++)
        x[k * stride] = I * std::polar(1.0, -PI * double(k + 1) / double(n2)) * z[k + 1];
      break;
    case ORC_REDFT01:  // y_j = x_0 + 2 sum_{k>=1} x_k cos(pi k (2j+1) / 2n)
      z[0] = x[0];
      for (size_t k = 1; k < n; k++) {
        z[k] = x[k * stride] * std::polar(1.0, PI * double(k) / double(n2));
        z[n2 - k] = x[k * stride] * std::polar(1.0, -PI * double(k) / double(n2));
      }
      z[n] = 0;
      fft_inplace(z.data(), n2, 1, +1);
      for (size_t j = 0; j < n; j++) x[j * stride] = z[j];
      break;
    case ORC_RODFT01:  // y_j = (-1)^j x_{n-1} + 2 sum_{k<n-1} x_k sin(pi (k+1)(2j+1) / 2n)
      z[0] = 0;
      for (size_t p = 1; p < n; p++) {
        const cd v = x[(p - 1) * stride];
        z[p] = -I * v * std::polar(1.0, PI * double(p) / double(n2));
        z[n2 - p] = I * v * std::polar(1.0, -PI * double(p) / double(n2));
      }
      z[n] = x[(n - 1) * stride];
      fft_inplace(z.data(), n2, 1, +1);
      for (size_t j = 0; j < n; j++) x[j * stride] = z[j];
      break;
    default:
      throw std::invalid_argument("r2r kind");
  }
}

void r2r_direct(cd* x, size_t n, size_t stride, int kind) {
  std::vector<cd> in(n), out(n, 0.0);
  for (size_t j = 0; j < n; j++) in[j] = x[j * stride];
  const double N = double(n);
  for (size_t k = 0; k < n; k++) {
    cd s = 0;
    for (size_t j = 0; j < n; j++) {
      double c = 0;
      switch (kind) {
        case ORC_REDFT10: c = 2 * std::cos(PI * (j + 0.5) * k / N); break;
        case ORC_RODFT10: c = 2 * std::sin(PI * (j + 0.5) * (k + 1.0) / N); break;
        case ORC_REDFT01: c = (j == 0) ? 1.0 : 2 * std::cos(PI * j * (k + 0.5) / N); break;
        case ORC_RODFT01:
          c = (j == n - 1) ? ((k % 2) ? -1.0 : 1.0) : 2 * std::sin(PI * (j + 1.0) * (k + 0.5) / N);
          break;
      }
      s += c * in[j];
    }
    out[k] = s;
  }
  for (size_t k = 0; k < n; k++) x[k * stride] = out[k];
}

// ---------------------------------------------------------------------------
// Grid (datalayout.cpp:21-36) and Transformer tables (transformer.cpp:66-90).
// ---------------------------------------------------------------------------
struct Grid {
  size_t sx, sy, M;
  double dx, lenx, leny;
  std::vector<double> posx, posy, fft_kx, fft_ky, dsct_kx, dsct_ky;
  double nFFT, nFFTx, nFFTy, nDSCT, nDSCTx, nDSCTy;
  Grid(size_t sx_, size_t sy_, double dx_) : sx(sx_), sy(sy_), M(sx_ * sy_), dx(dx_) {
    lenx = double(sx) * dx;
    leny = double(sy) * dx;
    const int isx = int(sx), isy = int(sy);
    posx.resize(sx); posy.resize(sy);
    for (int x = 0; x < isx; x++) posx[x] = double(2 * x + 1 - isx) * 0.5 * dx;
    for (int y = 0; y < isy; y++) posy[y] = double(2 * y + 1 - isy) * 0.5 * dx;
    nFFT = 1.0 / double(sx * sy);
    nFFTx = 1.0 / double(sx);
    nFFTy = 1.0 / double(sy);
    nDSCT = 1.0 / double(4 * sx * sy);
    nDSCTx = 1.0 / double(2 * sx);
    nDSCTy = 1.0 / double(2 * sy);
    const double mx = PI / lenx, my = PI / leny;
    fft_kx.resize(sx); dsct_kx.resize(sx); fft_ky.resize(sy); dsct_ky.resize(sy);
    for (int x = 0; x < isx; x++) {
      fft_kx[x] = double((x < isx / 2) ? x : x - isx) * 2 * mx;
      dsct_kx[x] = double(x + 1) * mx;
    }
    for (int y = 0; y < isy; y++) {
      fft_ky[y] = double((y < isy / 2) ? y : y - isy) * 2 * my;
      dsct_ky[y] = double(y + 1) * my;
    }
  }
  double kx(size_t x, int bt) const { return bt == ORC_PERIODIC ? fft_kx[x] : dsct_kx[x]; }
  double ky(size_t y, int bt) const { return bt == ORC_PERIODIC ? fft_ky[y] : dsct_ky[y]; }
  double norm(int bt) const { return bt == ORC_PERIODIC ? nFFT : nDSCT; }
  double normx(int bt) const { return bt == ORC_PERIODIC ? nFFTx : nDSCTx; }
  double normy(int bt) const { return bt == ORC_PERIODIC ? nFFTy : nDSCTy; }
};

// Transformer::transform (transformer.hpp:132-158) on one [sy][sx] field.
void rows(cd* d, size_t sx, size_t sy, int sign, int r2r) {
  for (size_t y = 0; y < sy; y++) {
    if (r2r < 0) fft_inplace(d + y * sx, sx, 1, sign);
    else r2r_inplace(d + y * sx, sx, 1, r2r);
  }
}
void cols(cd* d, size_t sx, size_t sy, int sign, int r2r) {
  for (size_t x = 0; x < sx; x++) {
    if (r2r < 0) fft_inplace(d + x, sy, sx, sign);
    else r2r_inplace(d + x, sy, sx, r2r);
  }
}
void transform(cd* d, size_t sx, size_t sy, int kind) {
  switch (kind) {
    case ORC_FFT: rows(d, sx, sy, -1, -1); cols(d, sx, sy, -1, -1); break;
    case ORC_iFFT: rows(d, sx, sy, +1, -1); cols(d, sx, sy, +1, -1); break;
    case ORC_FFTx: rows(d, sx, sy, -1, -1); break;
    case ORC_iFFTx: rows(d, sx, sy, +1, -1); break;
    case ORC_FFTy: cols(d, sx, sy, -1, -1); break;
    case ORC_iFFTy: cols(d, sx, sy, +1, -1); break;
    case ORC_DST: rows(d, sx, sy, 0, ORC_RODFT10); cols(d, sx, sy, 0, ORC_RODFT10); break;
    case ORC_iDST: rows(d, sx, sy, 0, ORC_RODFT01); cols(d, sx, sy, 0, ORC_RODFT01); break;
    case ORC_DSTx: rows(d, sx, sy, 0, ORC_RODFT10); break;
    case ORC_iDSTx: rows(d, sx, sy, 0, ORC_RODFT01); break;
    case ORC_DSTy: cols(d, sx, sy, 0, ORC_RODFT10); break;
    case ORC_iDSTy: cols(d, sx, sy, 0, ORC_RODFT01); break;
    case ORC_DCT: rows(d, sx, sy, 0, ORC_REDFT10); cols(d, sx, sy, 0, ORC_REDFT10); break;
    case ORC_iDCT: rows(d, sx, sy, 0, ORC_REDFT01); cols(d, sx, sy, 0, ORC_REDFT01); break;
    case ORC_DCTx: rows(d, sx, sy, 0, ORC_REDFT10); break;
    case ORC_iDCTx: rows(d, sx, sy, 0, ORC_REDFT01); break;
    case ORC_DCTy: cols(d, sx, sy, 0, ORC_REDFT10); break;
    case ORC_iDCTy: cols(d, sx, sy, 0, ORC_REDFT01); break;
    default: throw std::invalid_argument("transform kind");
  }
}

// ---------------------------------------------------------------------------
// State pointwise helpers (state.hpp:180-211) and BLAS-1 (state.hpp:130-240).
// ---------------------------------------------------------------------------
inline void pw_mul(cd* s, const double* v, size_t M) { for (size_t i = 0; i < M; i++) s[i] *= v[i]; }
inline void pw_mul_y(cd* s, const double* v, size_t sx, size_t sy) {
  for (size_t y = 0; y < sy; y++) for (size_t x = 0; x < sx; x++) s[y * sx + x] *= v[y];
}
inline void pw_mul_imag_shiftx(cd* s, const double* v, size_t sx, size_t sy) {
  for (size_t x = sx - 1; x > 0; x--)
    for (size_t y = 0; y < sy; y++) s[y * sx + x] = s[y * sx + x - 1] * cd(0, v[y * sx + x - 1]);
  for (size_t y = 0; y < sy; y++) s[y * sx] = 0;
}
inline void axpy(cd* y, const cd* x, size_t M) { for (size_t i = 0; i < M; i++) y[i] += x[i]; }
inline cd dot(const cd* a, const cd* b, size_t M, double dx) {
  cd s = 0;
  for (size_t i = 0; i < M; i++) s += std::conj(a[i]) * b[i];
  return s * dx * dx;
}
inline double norm2(const cd* a, size_t M, double dx) {
  double s = 0;
  for (size_t i = 0; i < M; i++) s += std::norm(a[i]);
  return std::sqrt(s) * dx;
}

// ---------------------------------------------------------------------------
// Operators.
// ---------------------------------------------------------------------------
struct Op {
  virtual ~Op() {}
  virtual void operate(cd* s, cd* const* ws) const = 0;
  virtual size_t workspace() const = 0;
  virtual void set_time_step(double) {}
};

// ExpPotential: p*exp(c*e*V) (exppotential.cpp:55-63; operate exppotential.hpp:56-65)
struct ExpPotential : Op {
  const Grid& g; const double* V; double e, c, p; std::vector<double> values;
  ExpPotential(const Grid& g_, const double* V_, double e_, double c_, double p_)
      : g(g_), V(V_), e(e_), c(c_), p(p_) { recalc(); }
  void recalc() {
    if (!V) return;
    values.resize(g.M);
    for (size_t i = 0; i < g.M; i++) values[i] = p * std::exp(e * c * V[i]);
  }
  void set_time_step(double ne) override { e = ne; recalc(); }
  void operate(cd* s, cd* const*) const override {
    if (!V) { if (p != 1.0) for (size_t i = 0; i < g.M; i++) s[i] *= p; }
    else pw_mul(s, values.data(), g.M);
  }
  size_t workspace() const override { return 0; }
};

// ExpKinetic: p*exp(-e*c*T), exact for any B (expkinetic.cpp:33-197)
struct ExpKinetic : Op {
  const Grid& g; int bt; double B, coeff, pref, e;
  std::vector<double> mult, xm, xm2, ym;
  ExpKinetic(const Grid& g_, double e_, double B_, int bt_, double c_ = -1.0, double p_ = 1.0)
      : g(g_), bt(bt_), B(B_), coeff(c_), pref(p_), e(e_) { calc(); }
  void set_time_step(double ne) override { e = ne; calc(); }
  void calc() {  // expkinetic.cpp:65-136
    const size_t sx = g.sx, sy = g.sy;
    if (B == 0) {
      mult.resize(g.M);
      const double normfac = g.norm(bt);
      const double A = coeff * e * 0.5;
      const double p = pref * normfac;
      for (size_t y = 0; y < sy; y++) {
        const double ky = g.ky(y, bt);
        for (size_t x = 0; x < sx; x++) {
          const double kx = g.kx(x, bt);
          mult[y * sx + x] = p * std::exp(A * (kx * kx + ky * ky));
        }
      }
    } else {
      xm.resize(g.M); ym.resize(sy);
      if (bt == ORC_DIRICHLET) xm2.resize(g.M);
      const double normfacx = (bt == ORC_PERIODIC) ? std::sqrt(g.nFFTx) : g.nDSCTx;
      const double normfacy = g.normy(bt);
      const double z = B * e;
      const double Cy = (std::fabs(z) < 1e-6) ? 1 + std::pow(z, 2) / 6 + std::pow(z, 4) / 120 : std::sinh(z) / z;
      const double Cx = (std::fabs(z) < 1e-6) ? 0.5 - std::pow(z, 2) / 24 + std::pow(z, 4) / 240 : std::tanh(z / 2) / z;
      const double Ax = coeff * e * 0.5 * Cx;
      const double Ay = coeff * e * 0.5 * Cy;
      const double px = normfacx;
      const double py = pref * normfacy;
      for (size_t y = 0; y < sy; y++) {
        const double ky = g.ky(y, bt);
        const double dy = g.posy[y];
        ym[y] = py * std::exp(Ay * ky * ky);
        if (bt == ORC_PERIODIC) {
          for (size_t x = 0; x < sx; x++) {
            const double kx = g.kx(x, bt);
            xm[y * sx + x] = px * std::exp(Ax * (kx - B * dy) * (kx - B * dy));
          }
        } else {
          const double By = B * dy;
          for (size_t x = 0; x < sx; x++) {
            const double kx = g.kx(x, bt);
            const double common = px * std::exp(Ax * (kx * kx + By * By));
            const double arg = 2 * Ax * kx * By;
            xm[y * sx + x] = common * std::cosh(arg);
            xm2[y * sx + x] = common * std::sinh(arg);
          }
        }
      }
    }
  }
  size_t workspace() const override { return (B == 0 || bt == ORC_PERIODIC) ? 0 : 1; }
  void operate(cd* s, cd* const* ws) const override {  // expkinetic.cpp:138-197
    const size_t sx = g.sx, sy = g.sy, M = g.M;
    if (B == 0) {
      if (bt == ORC_PERIODIC) { transform(s, sx, sy, ORC_FFT); pw_mul(s, mult.data(), M); transform(s, sx, sy, ORC_iFFT); }
      else { transform(s, sx, sy, ORC_DST); pw_mul(s, mult.data(), M); transform(s, sx, sy, ORC_iDST); }
    } else if (bt == ORC_PERIODIC) {
      transform(s, sx, sy, ORC_FFTx);
      pw_mul(s, xm.data(), M);
      transform(s, sx, sy, ORC_FFTy);
      pw_mul_y(s, ym.data(), sx, sy);
      transform(s, sx, sy, ORC_iFFTy);
      pw_mul(s, xm.data(), M);
      transform(s, sx, sy, ORC_iFFTx);
    } else {
      cd* temp = ws[0];
      transform(s, sx, sy, ORC_DSTx);
      std::memcpy(temp, s, M * sizeof(cd));
      pw_mul(s, xm.data(), M);
      pw_mul_imag_shiftx(temp, xm2.data(), sx, sy);
      transform(s, sx, sy, ORC_iDSTx);
      transform(temp, sx, sy, ORC_iDCTx);
      axpy(s, temp, M);
      transform(s, sx, sy, ORC_DST);
      pw_mul_y(s, ym.data(), sx, sy);
      transform(s, sx, sy, ORC_iDSTy);
      std::memcpy(temp, s, M * sizeof(cd));
      pw_mul(s, xm.data(), M);
      pw_mul_imag_shiftx(temp, xm2.data(), sx, sy);
      transform(s, sx, sy, ORC_iDSTx);
      transform(temp, sx, sy, ORC_iDCTx);
      axpy(s, temp, M);
    }
  }
};

// Kinetic T (kinetic.cpp:25-142)
struct Kinetic : Op {
  const Grid& g; int bt; double B;
  std::vector<double> mxy, mx, my, mx2;
  Kinetic(const Grid& g_, double B_, int bt_) : g(g_), bt(bt_), B(B_) {
    const size_t sx = g.sx, sy = g.sy;
    const double normfac = g.norm(bt), normfac_x = g.normx(bt), normfac_y = g.normy(bt);
    if (B == 0) {
      mxy.resize(g.M);
      for (size_t y = 0; y < sy; y++) {
        const double ky = g.ky(y, bt);
        for (size_t x = 0; x < sx; x++) {
          const double kx = g.kx(x, bt);
          mxy[y * sx + x] = 0.5 * (kx * kx + ky * ky) * normfac;
        }
      }
    } else {
      mx.resize(g.M); my.resize(sy);
      if (bt == ORC_DIRICHLET) mx2.resize(g.M);
      for (size_t y = 0; y < sy; y++) {
        const double ky = g.ky(y, bt);
        my[y] = 0.5 * ky * ky * normfac_y;
        const double dy = g.posy[y];
        for (size_t x = 0; x < sx; x++) {
          if (bt == ORC_PERIODIC) {
            const double kx = g.kx(x, bt) - B * dy;
            mx[y * sx + x] = 0.5 * kx * kx * normfac_x;
          } else {
            const double kx = g.kx(x, bt);
            mx[y * sx + x] = 0.5 * (kx * kx + B * B * dy * dy) * normfac_x;
            mx2[y * sx + x] = B * dy * kx * normfac_x;
          }
        }
      }
    }
  }
  size_t workspace() const override { return B == 0 ? 0 : (bt == ORC_PERIODIC ? 1 : 2); }
  void operate(cd* s, cd* const* ws) const override {
    const size_t sx = g.sx, sy = g.sy, M = g.M;
    if (B == 0) {
      if (bt == ORC_PERIODIC) { transform(s, sx, sy, ORC_FFT); pw_mul(s, mxy.data(), M); transform(s, sx, sy, ORC_iFFT); }
      else { transform(s, sx, sy, ORC_DST); pw_mul(s, mxy.data(), M); transform(s, sx, sy, ORC_iDST); }
      return;
    }
    cd* temp = ws[0];
    std::memcpy(temp, s, M * sizeof(cd));
    if (bt == ORC_PERIODIC) {
      transform(s, sx, sy, ORC_FFTx);
      pw_mul(s, mx.data(), M);
      transform(s, sx, sy, ORC_iFFTx);
      transform(temp, sx, sy, ORC_FFTy);
      pw_mul_y(temp, my.data(), sx, sy);
      transform(temp, sx, sy, ORC_iFFTy);
    } else {
      cd* temp2 = ws[1];
      transform(s, sx, sy, ORC_DSTx);
      std::memcpy(temp2, s, M * sizeof(cd));
      pw_mul(s, mx.data(), M);
      transform(s, sx, sy, ORC_iDSTx);
      pw_mul_imag_shiftx(temp2, mx2.data(), sx, sy);
      transform(temp2, sx, sy, ORC_iDCTx);
      axpy(s, temp2, M);
      transform(temp, sx, sy, ORC_DSTy);
      pw_mul_y(temp, my.data(), sx, sy);
      transform(temp, sx, sy, ORC_iDSTy);
    }
    axpy(s, temp, M);
  }
};

// Potential (potential.hpp:52-58)
struct PotentialOp : Op {
  const Grid& g; const double* V;
  PotentialOp(const Grid& g_, const double* V_) : g(g_), V(V_) {}
  void operate(cd* s, cd* const*) const override {
    if (!V) std::fill(s, s + g.M, cd(0)); else pw_mul(s, V, g.M);
  }
  size_t workspace() const override { return 0; }
};

// OperatorProduct: components applied in reverse insertion order (operatorproduct.cpp:48-52)
struct Product : Op {
  std::vector<const Op*> comps;
  void operate(cd* s, cd* const* ws) const override {
    for (auto it = comps.rbegin(); it != comps.rend(); ++it) (*it)->operate(s, ws);
  }
  size_t workspace() const override {
    size_t w = 0; for (auto c : comps) w = std::max(w, c->workspace()); return w;
  }
};

// OperatorSum (operatorsum.cpp:40-86)
struct Sum : Op {
  const Grid& g; std::vector<const Op*> comps;
  explicit Sum(const Grid& g_) : g(g_) {}
  size_t workspace() const override {
    if (comps.empty()) return 0;
    if (comps.size() == 1) return comps[0]->workspace();
    size_t w = 0; for (auto c : comps) w = std::max(w, c->workspace()); return w + 2;
  }
  void operate(cd* s, cd* const* ws) const override {
    if (comps.empty()) { std::fill(s, s + g.M, cd(0)); return; }
    if (comps.size() == 1) { comps[0]->operate(s, ws); return; }
    cd* orig = ws[0];
    cd* inter = ws[1];
    std::memcpy(orig, s, g.M * sizeof(cd));
    cd* const* slice = ws + 2;
    comps[0]->operate(s, slice);
    for (size_t k = 1; k < comps.size(); k++) {
      std::memcpy(inter, orig, g.M * sizeof(cd));
      comps[k]->operate(inter, slice);
      axpy(s, inter, g.M);
    }
  }
};

// SecondOrderSplit (secondordersplit.cpp:21-69)
struct SecondOrderSplit : Product {
  std::unique_ptr<ExpKinetic> kin;
  std::unique_ptr<ExpPotential> pot, pot_pref, pot_sq;
  SecondOrderSplit(const Grid& g, const double* V, double e, double B, int bt, double prefactor, int exponent) {
    if (V) {
      kin.reset(new ExpKinetic(g, e, B, bt, -1.0));
      pot.reset(new ExpPotential(g, V, e, -0.5, 1.0));
      if (prefactor != 1.0) pot_pref.reset(new ExpPotential(g, V, e, -0.5, prefactor));
      if (exponent >= 2) pot_sq.reset(new ExpPotential(g, V, e, -1.0, 1.0));
      comps.push_back(pot_pref ? pot_pref.get() : pot.get());
      comps.push_back(kin.get());
      for (int t = 2; t <= exponent; t++) { comps.push_back(pot_sq.get()); comps.push_back(kin.get()); }
      comps.push_back(pot.get());
    } else {
      kin.reset(new ExpKinetic(g, e, B, bt, -1.0 * exponent, prefactor));
      comps.push_back(kin.get());
    }
  }
  void set_time_step(double e) override {
    kin->set_time_step(e);
    if (pot) pot->set_time_step(e);
    if (pot_sq) pot_sq->set_time_step(e);
    if (pot_pref) pot_pref->set_time_step(e);
  }
};

// MultiProductSplit (multiproductsplit.cpp:23-58)
int mps_coefficients(int h, std::vector<double>& c) {
  c.resize(h);
  for (int t = 1; t <= h; t++) {
    int denom = 1;
    for (int n = 1; n <= h; n++) if (n != t) denom *= (t * t - n * n);
    c[t - 1] = std::pow(t, 2 * (h - 1)) / denom;
  }
  return h;
}

struct MultiProductSplit : Sum {
  int halforder;
  std::vector<double> coeffs;
  std::vector<std::unique_ptr<SecondOrderSplit>> members;
  MultiProductSplit(const Grid& g, int h, const double* V, double e, int bt, double B)
      : Sum(g), halforder(V ? h : 1) {
    mps_coefficients(halforder, coeffs);
    for (int t = 0; t < halforder; t++) {
      members.emplace_back(new SecondOrderSplit(g, V, e / (t + 1), B, bt, coeffs[t], t + 1));
      comps.push_back(members.back().get());
    }
  }
  void set_time_step(double e) override {
    for (size_t t = 0; t < members.size(); t++) members[t]->set_time_step(e / double(t + 1));
  }
};

// Hamiltonian H = kin + pot (itpsystem.cpp:117-119)
struct Hamiltonian {
  Kinetic kin; PotentialOp pot; Sum H;
  Hamiltonian(const Grid& g, double B, int bt, const double* V) : kin(g, B, bt), pot(g, V), H(g) {
    H.comps.push_back(&kin);
    if (V) H.comps.push_back(&pot);
  }
};

// Per-thread workspace: `count` fields of M values, handed out as a pointer array.
struct Workspace {
  std::vector<std::vector<cd>> store;
  std::vector<cd*> ptrs;
  Workspace(size_t count, size_t M) : store(count, std::vector<cd>(M)), ptrs(count) {
    for (size_t i = 0; i < count; i++) ptrs[i] = store[i].data();
  }
  cd* const* get() const { return ptrs.data(); }
};
Workspace make_ws(size_t count, size_t M) { return Workspace(count, M); }

// Operator::mean_and_standard_deviation (operators.cpp:40-50)
std::pair<cd, cd> mean_and_sd(const Op& H, const cd* state, size_t M, double dx,
                              cd* const* ws) {
  cd* const* slice = ws + 1;
  cd* temp = ws[0];
  std::memcpy(temp, state, M * sizeof(cd));
  H.operate(temp, slice);
  const cd mean = dot(state, temp, M, dx);
  H.operate(temp, slice);
  const cd meansqr = dot(state, temp, M, dx);
  return {mean, std::sqrt(meansqr - mean * mean)};
}

// ---------------------------------------------------------------------------
// Hermitian eigensolver: cyclic two-sided Jacobi with round-robin ordering,
// so each round's rotations are disjoint and can be applied in parallel.
// Result sorted ascending, eigenvectors in columns, like zheev('V','U')
// (eigensolver.hpp:57-84).
// ---------------------------------------------------------------------------
int jacobi_eigh(size_t n, std::vector<cd>& a /* row-major full */, std::vector<double>& w,
                std::vector<cd>& v /* row-major, columns = eigenvectors */) {
  for (size_t i = 0; i < n * n; i++)
    if (!std::isfinite(a[i].real()) || !std::isfinite(a[i].imag())) return ORC_E_EIGENSOLVER;
  const size_t np = n + (n & 1);
  std::vector<cd> A(np * np, 0.0), V(np * np, 0.0);
  for (size_t i = 0; i < n; i++) for (size_t j = 0; j < n; j++) A[i * np + j] = a[i * n + j];
  for (size_t i = 0; i < np; i++) { V[i * np + i] = 1.0; A[i * np + i] = A[i * np + i].real(); }
  const size_t m = np / 2;
  std::vector<size_t> pos(np);  // player at each position
  for (size_t i = 0; i < np; i++) pos[i] = i;
  std::vector<double> cs(m);
  std::vector<cd> sj(m);
  std::vector<size_t> P(m), Q(m);
  const double tol = 1e-14;
  bool converged = false;
  for (int sweep = 0; sweep < 60 && !converged; sweep++) {
    double maxoff = 0, maxb = 0, maxd = 0;
    for (size_t round = 0; round + 1 < np; round++) {
      for (size_t k = 0; k < m; k++) {
        size_t p = pos[2 * k], q = pos[2 * k + 1];
        if (p > q) std::swap(p, q);
        P[k] = p; Q[k] = q;
        const double app = A[p * np + p].real(), aqq = A[q * np + q].real();
        const cd b = A[p * np + q];
        const double ab = std::abs(b);
        const double scale = std::sqrt(std::fabs(app * aqq));
        const double off = scale > 0 ? ab / scale : ab;
        maxoff = std::max(maxoff, off);
        maxb = std::max(maxb, ab);
        maxd = std::max(maxd, std::max(std::fabs(app), std::fabs(aqq)));
        if (ab == 0 || off < 1e-17) { cs[k] = 1; sj[k] = 0; continue; }
        const double tau = (aqq - app) / (2 * ab);
        const double t = (tau >= 0 ? 1.0 : -1.0) / (std::fabs(tau) + std::hypot(1.0, tau));
        const double c = 1 / std::sqrt(1 + t * t);
        cs[k] = c;
        sj[k] = (t * c) * (b / ab);
      }
      // A <- A J (columns), V <- V J
#pragma omp parallel for schedule(static)
      for (size_t k = 0; k < m; k++) {
        const size_t p = P[k], q = Q[k];
        const double c = cs[k]; const cd s = sj[k];
        if (s == cd(0)) continue;
        for (size_t i = 0; i < np; i++) {
          const cd aip = A[i * np + p], aiq = A[i * np + q];
          A[i * np + p] = c * aip - std::conj(s) * aiq;
          A[i * np + q] = s * aip + c * aiq;
          const cd vip = V[i * np + p], viq = V[i * np + q];
          V[i * np + p] = c * vip - std::conj(s) * viq;
          V[i * np + q] = s * vip + c * viq;
        }
      }
      // A <- J^H A (rows)
#pragma omp parallel for schedule(static)
      for (size_t k = 0; k < m; k++) {
        const size_t p = P[k], q = Q[k];
        const double c = cs[k]; const cd s = sj[k];
        if (s == cd(0)) continue;
        for (size_t j = 0; j < np; j++) {
          const cd apj = A[p * np + j], aqj = A[q * np + j];
          A[p * np + j] = c * apj - s * aqj;
          A[q * np + j] = std::conj(s) * apj + c * aqj;
        }
        A[p * np + p] = A[p * np + p].real();
        A[q * np + q] = A[q * np + q].real();
      }
      // round-robin: position 0 fixed, the rest rotate
      std::vector<size_t> np_pos(np);
      np_pos[0] = pos[0];
      for (size_t k = 0; k < m; k++) {
        const size_t top = 2 * k, bot = 2 * k + 1;
        // top k (k>=1) -> top k+1, or bottom m-1 for the last
        if (k >= 1) { if (k + 1 < m) np_pos[2 * (k + 1)] = pos[top]; else np_pos[2 * (m - 1) + 1] = pos[top]; }
        // bottom 0 -> top 1 ; bottom k (k>=1) -> bottom k-1
        if (k == 0) { if (m > 1) np_pos[2] = pos[bot]; else np_pos[1] = pos[bot]; }
        else np_pos[2 * (k - 1) + 1] = pos[bot];
      }
      pos.swap(np_pos);
    }
    // relative off-diagonal below tol, or all off-diagonals at the rounding
    // level of the largest diagonal element
    if (maxoff < tol || maxb <= 1e-15 * maxd) converged = true;
  }
  if (!converged) return ORC_E_EIGENSOLVER;
  // drop the padding player (its row/column stayed zero and untouched)
  std::vector<size_t> idx;
  for (size_t i = 0; i < n; i++) idx.push_back(i);
  std::vector<double> lam(n);
  for (size_t i = 0; i < n; i++) lam[i] = A[i * np + i].real();
  std::stable_sort(idx.begin(), idx.end(), [&](size_t x, size_t y) { return lam[x] < lam[y]; });
  w.resize(n); v.assign(n * n, 0.0);
  for (size_t r = 0; r < n; r++) {
    w[r] = lam[idx[r]];
    for (size_t i = 0; i < n; i++) v[i * n + r] = V[i * np + idx[r]];
  }
  return ORC_OK;
}

// ---------------------------------------------------------------------------
// StateSet::orthonormalize (stateset.cpp:136-232), out-of-place rotation
// (the HighMem branch; the Default branch computes the same product).
// ---------------------------------------------------------------------------
void overlap(const cd* st, size_t N, size_t M, double dx, std::vector<cd>& S /* col-major */) {
  S.assign(N * N, 0.0);
#pragma omp parallel for schedule(dynamic)
  for (size_t i = 0; i < N; i++)
    for (size_t j = i; j < N; j++) S[N * j + i] = dot(st + i * M, st + j * M, M, dx);
}

int orthonormalize(cd* st, size_t N, size_t M, double dx, size_t* bad_n, double* bad_l) {
  if (N == 1) {
    const double nrm = norm2(st, M, dx);
    for (size_t i = 0; i < M; i++) st[i] *= 1.0 / nrm;
    return ORC_OK;
  }
  std::vector<cd> S;
  overlap(st, N, M, dx, S);
  // full Hermitian, row-major, from the upper triangle (i <= j)
  std::vector<cd> a(N * N);
  for (size_t i = 0; i < N; i++)
    for (size_t j = 0; j < N; j++)
      a[i * N + j] = (i <= j) ? S[N * j + i] : std::conj(S[N * i + j]);
  std::vector<double> w; std::vector<cd> v;
  int rc = jacobi_eigh(N, a, w, v);
  if (rc) return rc;
  std::vector<cd> C(N * N);  // C[c*N + r] = v_r[c] / sqrt(lambda_r)
  for (size_t n = 0; n < N; n++) {
    const double ev = w[n];
    if (ev <= 0) { if (bad_n) *bad_n = n; if (bad_l) *bad_l = ev; return ORC_E_NONPOSITIVE_EIG; }
    if (std::fpclassify(ev) != FP_NORMAL) { if (bad_n) *bad_n = n; if (bad_l) *bad_l = ev; return ORC_E_NONNORMAL_EIG; }
    const double sc = 1 / std::sqrt(ev);
    for (size_t c = 0; c < N; c++) C[c * N + n] = v[c * N + n] * sc;
  }
  std::vector<cd> out(N * M);
#pragma omp parallel for schedule(static)
  for (size_t t = 0; t < M; t++) {
    for (size_t r = 0; r < N; r++) {
      cd s = 0;
      for (size_t c = 0; c < N; c++) s += C[c * N + r] * st[c * M + t];
      out[r * M + t] = s;
    }
  }
  std::memcpy(st, out.data(), N * M * sizeof(cd));
  return ORC_OK;
}

// splitmix64 stream for seeded uniform noise
inline uint64_t splitmix64(uint64_t& s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
inline double uniform01(uint64_t& s) { return double(splitmix64(s) >> 11) * (1.0 / 9007199254740992.0); }

void init_states(uint64_t& rng, size_t N, size_t M, double dx, cd* out) {
  for (size_t n = 0; n < N; n++) {
    cd* s = out + n * M;
    for (size_t i = 0; i < M; i++) {
      const double re = 2 * uniform01(rng) - 1;
      const double im = 2 * uniform01(rng) - 1;
      s[i] = cd(re, im);
    }
    const double nrm = norm2(s, M, dx);
    const double f = 1.0 / nrm;
    for (size_t i = 0; i < M; i++) s[i] *= f;
  }
}

// ---------------------------------------------------------------------------
// ITP driver (itpsystem.cpp:23-139, 202-511) and convergence tests
// (convergence.cpp:75-200).
// ---------------------------------------------------------------------------
struct System {
  orc_params p;
  std::vector<double> eps_list_storage;
  Grid g;
  std::vector<double> V;
  bool has_V;
  std::unique_ptr<MultiProductSplit> T;
  std::unique_ptr<Hamiltonian> H;
  std::vector<cd> states;
  uint64_t rng;
  double eps;
  std::deque<double> eps_values;
  bool exhausting;
  int step_counter = 0, total_step_counter = 0;
  bool finished = false, error_flag = false;
  bool all_ts = false, all_final = false;
  std::vector<std::tuple<double, double, size_t>> esn;
  std::vector<std::vector<double>> energies, stdevs;
  std::vector<char> ts_conv, fin_conv;
  size_t n_ts = 0, n_fin = 0;
  std::chrono::steady_clock::time_point t0;
  double total_time = 0;

  System(const orc_params& pp, const double* Vin, const double* init)
      : p(pp), g(pp.sx, pp.sy, pp.lenx / double(pp.sx)), has_V(Vin != nullptr) {
    if (p.n_eps) eps_list_storage.assign(p.eps_values, p.eps_values + p.n_eps);
    p.eps_values = nullptr;
    if (has_V) V.assign(Vin, Vin + g.M);
    for (double e : eps_list_storage) eps_values.push_back(e);
    exhausting = p.exhaust_eps != 0;
    if (!eps_values.empty()) { eps = eps_values.front(); eps_values.pop_front(); }
    else eps = 0.5;  // Parameters::default_initial_eps
    H.reset(new Hamiltonian(g, p.B, p.boundary, has_V ? V.data() : nullptr));
    T.reset(new MultiProductSplit(g, p.halforder, has_V ? V.data() : nullptr, eps, p.boundary, p.B));
    states.resize(p.N * g.M);
    rng = p.seed;
    if (init) std::memcpy(static_cast<void*>(states.data()), init, p.N * g.M * sizeof(cd));
    else init_states(rng, p.N, g.M, g.dx, states.data());
    ts_conv.assign(p.N, 0); fin_conv.assign(p.N, 0);
  }
  void set_ts(size_t n, bool v) { if (bool(ts_conv[n]) == v) return; ts_conv[n] = v; n_ts += v ? 1 : size_t(-1); }
  void set_fin(size_t n, bool v) { if (bool(fin_conv[n]) == v) return; fin_conv[n] = v; n_fin += v ? 1 : size_t(-1); }

  bool test(const orc_test& t, size_t n) const {
    switch (t.kind) {
      case ORC_TEST_NONE: return false;
      case ORC_TEST_ONESTEP: return step_counter == 1 && ts_conv[std::get<2>(esn[n])];
      case ORC_TEST_RELCHANGE:
      case ORC_TEST_ABSCHANGE: {
        if (energies.size() < 2) return false;
        const double a = energies[energies.size() - 1][n], b = energies[energies.size() - 2][n];
        if (t.kind == ORC_TEST_RELCHANGE) return std::fabs((a - b) / b) < t.p0;
        return std::fabs(a - b) < t.p0;
      }
      case ORC_TEST_ABSSTDEV: {
        if (stdevs.size() < 2) return false;
        const size_t last = stdevs.size() - 1;
        const double a = stdevs[last][n], b = stdevs[last - 1][n];
        return (a < t.p0) || (std::fabs(a - b) < t.p1);
      }
      case ORC_TEST_RELSTDEV: {
        if (energies.size() < 2) return false;
        const size_t last = stdevs.size() - 1;
        const double a = stdevs[last][n] / energies[last][n];
        const double b = stdevs[last - 1][n] / energies[last - 1][n];
        return (a < t.p0) || (std::fabs(a - b) < t.p1);
      }
    }
    return false;
  }

  void finish() {
    if (finished) return;
    total_time = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    finished = true;
  }

  void change_time_step() {  // itpsystem.cpp:202-233
    if (!eps_values.empty()) {
      eps = eps_values.front(); eps_values.pop_front();
      if (eps_values.empty()) exhausting = false;
    } else eps /= p.eps_divisor;
    if (eps < p.min_time_step) { error_flag = true; finish(); return; }
    T->set_time_step(eps);
    step_counter = 0;
    all_ts = false;
    for (size_t n = 0; n < p.N; n++) set_ts(n, false);
  }

  void propagate() {  // itpsystem.cpp:243-256
    const size_t ws = std::max(T->workspace(), H->H.workspace() + 1);
#pragma omp parallel
    {
      auto w = make_ws(ws, g.M);
#pragma omp for schedule(dynamic)
      for (size_t n = 0; n < p.N; n++) T->operate(states.data() + n * g.M, w.get());
    }
  }

  void orthonormalize_step() {  // itpsystem.cpp:258-289
    size_t bn = 0; double bl = 0;
    int rc = orthonormalize(states.data(), p.N, g.M, g.dx, &bn, &bl);
    if (rc == ORC_E_NONPOSITIVE_EIG) {
      if (p.recover) {
        change_time_step();
        init_states(rng, p.N, g.M, g.dx, states.data());
      } else { error_flag = true; finish(); }
    } else if (rc != ORC_OK) { error_flag = true; finish(); }
  }

  void calculate_energies() {  // itpsystem.cpp:473-511
    if (total_step_counter == 0) return;
    esn.assign(p.N, std::make_tuple(0.0, 0.0, size_t(0)));
    const size_t ws = std::max(T->workspace(), H->H.workspace() + 1);
#pragma omp parallel
    {
      auto w = make_ws(ws, g.M);
#pragma omp for schedule(dynamic)
      for (size_t n = 0; n < p.N; n++) {
        auto r = mean_and_sd(H->H, states.data() + n * g.M, g.M, g.dx, w.get());
        esn[n] = std::make_tuple(r.first.real(), r.second.real(), n);
      }
    }
    std::sort(esn.begin(), esn.end());
    if (energies.size() < size_t(total_step_counter)) energies.resize(total_step_counter);
    if (stdevs.size() < size_t(total_step_counter)) stdevs.resize(total_step_counter);
    auto& ne = energies[total_step_counter - 1];
    auto& nd = stdevs[total_step_counter - 1];
    ne.clear(); nd.clear();
    for (auto& t : esn) { ne.push_back(std::get<0>(t)); nd.push_back(std::get<1>(t)); }
  }

  void check_ts() {  // itpsystem.cpp:291-332
    for (size_t n = 0; n < p.N; n++) if (test(p.timestep_test, n)) set_ts(std::get<2>(esn[n]), true);
    bool flag = true;
    for (size_t n = p.ignore_lowest; n < p.ignore_lowest + p.n_wanted; n++)
      if (!ts_conv[std::get<2>(esn[n])]) flag = false;
    all_ts = flag;
  }
  void check_final() {  // itpsystem.cpp:334-372
    for (size_t n = 0; n < p.N; n++) set_fin(std::get<2>(esn[n]), test(p.final_test, n));
    bool flag = true;
    for (size_t n = p.ignore_lowest; n < p.ignore_lowest + p.n_wanted; n++)
      if (!fin_conv[std::get<2>(esn[n])]) flag = false;
    all_final = flag;
  }

  void step() {  // itpsystem.cpp:407-470
    if (total_step_counter == 0) t0 = std::chrono::steady_clock::now();
    if (total_step_counter >= p.max_steps) { error_flag = true; finish(); return; }
    step_counter++; total_step_counter++;
    propagate();
    orthonormalize_step();
    calculate_energies();
    check_ts();
    if (all_ts) {
      check_final();
      if (all_final) { finish(); return; }
      change_time_step();
    } else if (exhausting) change_time_step();
  }
};

template <class F> int guard(F f) {
  try { return f(); } catch (const std::exception&) { return ORC_E_INVALID; }
}

template <class MakeOp>
int apply_each(size_t sx, size_t sy, double* states, size_t nstates, MakeOp make) {
  return guard([&] {
    const size_t M = sx * sy;
    auto op = make();
    const size_t ws = op->workspace() + 2;
#pragma omp parallel
    {
      auto w = make_ws(ws, M);
#pragma omp for
      for (size_t n = 0; n < nstates; n++) op->operate(reinterpret_cast<cd*>(states) + n * M, w.get());
    }
    return 0;
  });
}

}  // namespace

extern "C" {

int orc_num_threads(void) { return omp_get_max_threads(); }
void orc_set_num_threads(int n) { omp_set_num_threads(n); }

int orc_fft1d(double* data, size_t n, size_t stride, int sign) {
  return guard([&] { fft_inplace(reinterpret_cast<cd*>(data), n, stride, sign); return 0; });
}
int orc_r2r1d(double* data, size_t n, size_t stride, int kind) {
  return guard([&] { r2r_inplace(reinterpret_cast<cd*>(data), n, stride, kind); return 0; });
}
int orc_r2r1d_direct(double* data, size_t n, size_t stride, int kind) {
  return guard([&] { r2r_direct(reinterpret_cast<cd*>(data), n, stride, kind); return 0; });
}
int orc_transform(size_t sx, size_t sy, double* data, int kind) {
  return guard([&] { transform(reinterpret_cast<cd*>(data), sx, sy, kind); return 0; });
}
int orc_grid_tables(size_t sx, size_t sy, double dx, double* posx, double* posy, double* fft_kx,
                    double* fft_ky, double* dsct_kx, double* dsct_ky, double* norms) {
  Grid g(sx, sy, dx);
  std::copy(g.posx.begin(), g.posx.end(), posx);
  std::copy(g.posy.begin(), g.posy.end(), posy);
  std::copy(g.fft_kx.begin(), g.fft_kx.end(), fft_kx);
  std::copy(g.fft_ky.begin(), g.fft_ky.end(), fft_ky);
  std::copy(g.dsct_kx.begin(), g.dsct_kx.end(), dsct_kx);
  std::copy(g.dsct_ky.begin(), g.dsct_ky.end(), dsct_ky);
  norms[0] = g.nFFT; norms[1] = g.nFFTx; norms[2] = g.nFFTy;
  norms[3] = g.nDSCT; norms[4] = g.nDSCTx; norms[5] = g.nDSCTy;
  return 0;
}

int orc_apply_exppotential(size_t sx, size_t sy, double dx, const double* V, double e, double coeff,
                           double prefactor, double* states, size_t nstates) {
  Grid g(sx, sy, dx);
  return apply_each(sx, sy, states, nstates, [&] { return std::unique_ptr<Op>(new ExpPotential(g, V, e, coeff, prefactor)); });
}
int orc_apply_expkinetic(size_t sx, size_t sy, double dx, int boundary, double B, double e, double coeff,
                         double prefactor, double* states, size_t nstates) {
  Grid g(sx, sy, dx);
  return apply_each(sx, sy, states, nstates, [&] { return std::unique_ptr<Op>(new ExpKinetic(g, e, B, boundary, coeff, prefactor)); });
}
int orc_apply_kinetic(size_t sx, size_t sy, double dx, int boundary, double B, double* states, size_t nstates) {
  Grid g(sx, sy, dx);
  return apply_each(sx, sy, states, nstates, [&] { return std::unique_ptr<Op>(new Kinetic(g, B, boundary)); });
}
int orc_apply_hamiltonian(size_t sx, size_t sy, double dx, int boundary, double B, const double* V,
                          double* states, size_t nstates) {
  return guard([&] {
    Grid g(sx, sy, dx);
    Hamiltonian H(g, B, boundary, V);
    const size_t M = g.M;
    auto w = make_ws(H.H.workspace() + 2, M);
    for (size_t n = 0; n < nstates; n++) H.H.operate(reinterpret_cast<cd*>(states) + n * M, w.get());
    return 0;
  });
}
int orc_apply_secondorder(size_t sx, size_t sy, double dx, int boundary, double B, const double* V, double e,
                          double prefactor, int exponent, double* states, size_t nstates) {
  Grid g(sx, sy, dx);
  return apply_each(sx, sy, states, nstates, [&] {
    return std::unique_ptr<Op>(new SecondOrderSplit(g, V, e, B, boundary, prefactor, exponent));
  });
}
int orc_propagate(size_t sx, size_t sy, double dx, int boundary, double B, const double* V, int halforder,
                  double eps, double* states, size_t nstates) {
  Grid g(sx, sy, dx);
  return apply_each(sx, sy, states, nstates, [&] {
    return std::unique_ptr<Op>(new MultiProductSplit(g, halforder, V, eps, boundary, B));
  });
}
int orc_mps_coefficients(int halforder, double* out) {
  std::vector<double> c;
  mps_coefficients(halforder, c);
  std::copy(c.begin(), c.end(), out);
  return halforder;
}

int orc_eigh(size_t n, double* A, double* w) {
  return guard([&] {
    cd* a = reinterpret_cast<cd*>(A);
    std::vector<cd> full(n * n);
    for (size_t i = 0; i < n; i++)
      for (size_t j = 0; j < n; j++)
        full[i * n + j] = (i <= j) ? a[n * j + i] : std::conj(a[n * i + j]);
    std::vector<double> ww; std::vector<cd> v;
    int rc = jacobi_eigh(n, full, ww, v);
    if (rc) return rc;
    for (size_t r = 0; r < n; r++) {
      w[r] = ww[r];
      for (size_t c = 0; c < n; c++) a[r * n + c] = v[c * n + r];  // column r = eigenvector r
    }
    return 0;
  });
}

int orc_orthonormalize(size_t sx, size_t sy, double dx, double* states, size_t N, size_t* bad_n,
                       double* bad_lambda) {
  return guard([&] { return orthonormalize(reinterpret_cast<cd*>(states), N, sx * sy, dx, bad_n, bad_lambda); });
}
int orc_overlap(size_t sx, size_t sy, double dx, const double* states, size_t N, double* S) {
  return guard([&] {
    std::vector<cd> s;
    overlap(reinterpret_cast<const cd*>(states), N, sx * sy, dx, s);
    cd* o = reinterpret_cast<cd*>(S);
    for (size_t i = 0; i < N; i++)
      for (size_t j = 0; j < N; j++) o[N * j + i] = (i <= j) ? s[N * j + i] : std::conj(s[N * i + j]);
    return 0;
  });
}
int orc_energies(size_t sx, size_t sy, double dx, int boundary, double B, const double* V, const double* states,
                 size_t N, double* E, double* sigma, double* mean_c, double* meansq_c) {
  return guard([&] {
    Grid g(sx, sy, dx);
    Hamiltonian H(g, B, boundary, V);
    const size_t M = g.M;
#pragma omp parallel
    {
      auto w = make_ws(H.H.workspace() + 2, M);
#pragma omp for schedule(dynamic)
      for (size_t n = 0; n < N; n++) {
        const cd* st = reinterpret_cast<const cd*>(states) + n * M;
        cd* const* slice = w.get() + 1;
        cd* temp = w.get()[0];
        std::memcpy(temp, st, M * sizeof(cd));
        H.H.operate(temp, slice);
        const cd mean = dot(st, temp, M, dx);
        H.H.operate(temp, slice);
        const cd msq = dot(st, temp, M, dx);
        const cd sd = std::sqrt(msq - mean * mean);
        E[n] = mean.real();
        sigma[n] = sd.real();
        if (mean_c) { mean_c[2 * n] = mean.real(); mean_c[2 * n + 1] = mean.imag(); }
        if (meansq_c) { meansq_c[2 * n] = msq.real(); meansq_c[2 * n + 1] = msq.imag(); }
      }
    }
    return 0;
  });
}
double orc_how_orthonormal(size_t sx, size_t sy, double dx, const double* states, size_t N) {
  const size_t M = sx * sy;
  const cd* st = reinterpret_cast<const cd*>(states);
  double mx = 0;
  for (size_t n = 0; n < N; n++)
    for (size_t k = 0; k <= n; k++) {
      const cd z = dot(st + n * M, st + k * M, M, dx);
      mx = std::max(mx, std::fabs(z.imag()));
      mx = std::max(mx, std::fabs(n != k ? z.real() : z.real() - 1));
    }
  return mx;
}
int orc_init_states(uint64_t* rng_state, size_t sx, size_t sy, double dx, size_t N, double* out) {
  init_states(*rng_state, N, sx * sy, dx, reinterpret_cast<cd*>(out));
  return 0;
}

void* orc_system_create(const orc_params* p, const double* V, const double* init) {
  try { return new System(*p, V, init); } catch (...) { return nullptr; }
}
void orc_system_destroy(void* s) { delete static_cast<System*>(s); }
int orc_system_step(void* s) { static_cast<System*>(s)->step(); return 0; }
int orc_system_run(void* s) {
  System* sys = static_cast<System*>(s);
  while (!sys->finished) sys->step();
  return sys->total_step_counter;
}
int orc_system_finished(void* s) { return static_cast<System*>(s)->finished; }
int orc_system_error_flag(void* s) { return static_cast<System*>(s)->error_flag; }
int orc_system_total_steps(void* s) { return static_cast<System*>(s)->total_step_counter; }
double orc_system_eps(void* s) { return static_cast<System*>(s)->eps; }
size_t orc_system_num_converged(void* s) { return static_cast<System*>(s)->n_fin; }
size_t orc_system_num_timestep_converged(void* s) { return static_cast<System*>(s)->n_ts; }
int orc_system_sorted(void* s, double* E, double* sigma, size_t* index) {
  System* sys = static_cast<System*>(s);
  for (size_t n = 0; n < sys->esn.size(); n++) {
    if (E) E[n] = std::get<0>(sys->esn[n]);
    if (sigma) sigma[n] = std::get<1>(sys->esn[n]);
    if (index) index[n] = std::get<2>(sys->esn[n]);
  }
  return int(sys->esn.size());
}
// State pointwise and BLAS-1 operations on one field (state.hpp:180-240):
// kind 0 pointwise_multiply, 1 pointwise_multiply_imaginary_shiftx,
// 2 pointwise_multiply_y, 3 pointwise_multiply_and_add (with addstate).
int orc_pointwise(int kind, size_t sx, size_t sy, double* state, const double* values, const double* addstate) {
  cd* s = reinterpret_cast<cd*>(state);
  const size_t M = sx * sy;
  switch (kind) {
    case 0: pw_mul(s, values, M); return 0;
    case 1: pw_mul_imag_shiftx(s, values, sx, sy); return 0;
    case 2: pw_mul_y(s, values, sx, sy); return 0;
    case 3: {
      if (!addstate) return 1;
      const cd* a = reinterpret_cast<const cd*>(addstate);
      pw_mul(s, values, M);
      axpy(s, a, M);
      return 0;
    }
    default: return 1;
  }
}
// State::dot (out[0], out[1]) and State::norm (return value)
double orc_dot_norm(size_t M, double dx, const double* a, const double* b, double* out) {
  const cd* x = reinterpret_cast<const cd*>(a);
  const cd z = dot(x, reinterpret_cast<const cd*>(b), M, dx);
  out[0] = z.real();
  out[1] = z.imag();
  return norm2(x, M, dx);
}
int orc_system_converged(void* s, int* flags) {
  System* sys = static_cast<System*>(s);
  for (size_t n = 0; n < sys->esn.size(); n++) flags[n] = sys->fin_conv[std::get<2>(sys->esn[n])] ? 1 : 0;
  return int(sys->esn.size());
}
int orc_system_states(void* s, double* out) {
  System* sys = static_cast<System*>(s);
  std::memcpy(out, sys->states.data(), sys->states.size() * sizeof(cd));
  return 0;
}
double orc_system_total_time(void* s) { return static_cast<System*>(s)->total_time; }
int orc_system_history_len(void* s) { return int(static_cast<System*>(s)->energies.size()); }

}  // extern "C"
