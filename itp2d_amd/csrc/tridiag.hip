// Eigenvalues of a Hermitian matrix through Householder tridiagonalization
// and Sturm-sequence bisection, without eigenvectors: the check of the
// overlap eigenvalues (stateset.cpp:164-173, NonPositiveEigenvalue and
// NonNormalEigenvalue) at sizes where the full eigendecomposition is
// expensive. The reference computes them with LAPACK zheev
// (eigensolver.hpp:80-84); this is the reduction of zhetrd/zhetd2 (LAPACK's
// published algorithm: reflectors H = I - tau v v^H from zlarfg, with real
// beta, so the tridiagonal matrix is real) followed by bisection on the
// Sturm counts of T - x I (as dstebz), written for the GPU:
//
//  tridiag_kernel  one cooperative launch. Workgroup 0 forms each reflector
//                  and the vector w = p - tau/2 (p^H v) v; all workgroups
//                  then make one pass over the trailing matrix that applies
//                  A <- A - v w^H - w v^H and, fused into the same pass, the
//                  product p' = tau' A v' for the next reflector (whose
//                  column workgroup 0 formed from the not yet updated row
//                  plus the pending update). Two grid barriers per column;
//                  the matrix is kept in full (both triangles), row-major,
//                  so every row is one contiguous read and write.
//  bisect_kernel   one thread per eigenvalue: bisection of the Sturm count
//                  (LDL^T of T - x I, pivots bounded away from zero by
//                  pivmin) from the Gershgorin interval to a width of a few
//                  ulp, so eigenvalues come out in ascending order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>

#include "gpu.hpp"
#include "grid_sync.hpp"

namespace itp {

namespace {

constexpr int kTdThreads = 512;
// rows of a pass handled together by one wavefront (the vectors' entries are
// read once for all of them)
constexpr int kRows = 2;

struct TridiagArgs {
  cplx<double>* A;    // np x np row-major Hermitian (both triangles), destroyed
  int np, N;
  cplx<double>* vec;  // [2][3][N]: v, w, p of the reflector of each parity
  cplx<double>* tau;  // [2]
  double* d;          // [N] diagonal of T
  double* e;          // [N] off-diagonal of T (e[k] couples k and k+1)
  unsigned* bar;      // grid barrier counter and error flag
  unsigned spin_limit;
};

__global__ void td_zero_kernel(unsigned long long* p, long long n) {
  for (long long i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0ull;
}

__device__ __forceinline__ cplx<double> cdiv(cplx<double> a, cplx<double> b) {
  const double s = 1.0 / (b.x * b.x + b.y * b.y);
  return mk<double>((a.x * b.x + a.y * b.y) * s, (a.y * b.x - a.x * b.y) * s);
}

__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Sum over the workgroup (all threads get the result).
__device__ double block_sum(double v, double* red) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[w] = v;
  __syncthreads();
  double s = 0;
  for (int i = 0; i < kTdThreads / 64; i++) s += red[i];
  return s;
}

// zlarfg on x[0..n) in place (n >= 1): x[0] <- 1, x[1..n) <- the reflector's
// vector; returns tau and beta (real). Workgroup 0 only.
__device__ void reflector(cplx<double>* x, int n, cplx<double>& tau, double& beta, double* red) {
  const cplx<double> alpha = x[0];
  double s = 0;
  for (int r = 1 + (int)threadIdx.x; r < n; r += kTdThreads) s += x[r].x * x[r].x + x[r].y * x[r].y;
  const double xnorm2 = block_sum(s, red);
  if (xnorm2 == 0.0 && alpha.y == 0.0) {
    tau = mk<double>(0.0, 0.0);
    beta = alpha.x;
  } else {
    const double b = -std::copysign(std::sqrt(alpha.x * alpha.x + alpha.y * alpha.y + xnorm2), alpha.x);
    tau = mk<double>((b - alpha.x) / b, -alpha.y / b);
    const cplx<double> scale = cdiv(mk<double>(1.0, 0.0), mk<double>(alpha.x - b, alpha.y));
    for (int r = 1 + (int)threadIdx.x; r < n; r += kTdThreads) x[r] = x[r] * scale;
    beta = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) x[0] = mk<double>(1.0, 0.0);
  __syncthreads();
}

__global__ __launch_bounds__(kTdThreads) void tridiag_kernel(TridiagArgs a) {
  __shared__ double red[kTdThreads / 64];
  unsigned epoch = 0;
  const int N = a.N, np = a.np;
  cplx<double>* A = a.A;
  auto V = [&](int q) { return a.vec + (long long)(3 * q + 0) * N; };
  auto W = [&](int q) { return a.vec + (long long)(3 * q + 1) * N; };
  auto Pv = [&](int q) { return a.vec + (long long)(3 * q + 2) * N; };
  const int lane = threadIdx.x & 63;
  const int gwave = (int)blockIdx.x * (kTdThreads / 64) + ((int)threadIdx.x >> 6);
  const int nwaves = (int)gridDim.x * (kTdThreads / 64);
  auto barrier = [&]() -> bool { return grid_barrier(a.bar, epoch, a.bar + 1, a.spin_limit); };
  if (N == 1) {
    if (blockIdx.x == 0 && threadIdx.x == 0) a.d[0] = A[0].x;
    return;
  }
  // reflector of column 0 (from row 0: the matrix is Hermitian)
  if (blockIdx.x == 0) {
    cplx<double>* v = V(0);
    for (int r = threadIdx.x; r < N - 1; r += kTdThreads) v[r] = conj(A[r + 1]);
    __syncthreads();
    cplx<double> tau;
    double beta;
    reflector(v, N - 1, tau, beta, red);
    if (threadIdx.x == 0) {
      a.tau[0] = tau;
      a.d[0] = A[0].x;
      a.e[0] = beta;
    }
  }
  if (barrier()) return;
  // p_0 = tau_0 A[1:, 1:] v_0
  {
    const int m = N - 1;
    const cplx<double> t0 = a.tau[0];
    const cplx<double>* v = V(0);
    cplx<double>* p = Pv(0);
    for (int r0 = gwave * kRows; r0 < m; r0 += nwaves * kRows) {
      double sx[kRows], sy[kRows];
      const cplx<double>* row[kRows];
#pragma unroll
      for (int i = 0; i < kRows; i++) {
        sx[i] = sy[i] = 0;
        row[i] = A + (long long)(1 + std::min(r0 + i, m - 1)) * np + 1;
      }
#pragma unroll 4
      for (int c = lane; c < m; c += 64) {
        const cplx<double> vc = v[c];
#pragma unroll
        for (int i = 0; i < kRows; i++) {
          const cplx<double> b = row[i][c] * vc;
          sx[i] += b.x;
          sy[i] += b.y;
        }
      }
#pragma unroll
      for (int i = 0; i < kRows; i++) {
        const double x = wave_sum(sx[i]), y = wave_sum(sy[i]);
        if (lane == 0 && r0 + i < m) p[r0 + i] = t0 * mk<double>(x, y);
      }
    }
  }
  if (barrier()) return;
  for (int j = 0; j <= N - 2; j++) {
    const int q = j & 1, q1 = q ^ 1;
    const int m = N - 1 - j;  // length of v_j, w_j, p_j
    const bool next = j + 1 <= N - 2;
    if (blockIdx.x == 0) {
      const cplx<double>* v = V(q);
      const cplx<double>* p = Pv(q);
      cplx<double>* w = W(q);
      const cplx<double> tj = a.tau[q];
      // alpha = -tau/2 (p^H v), w = p + alpha v  (zhetd2)
      double sx = 0, sy = 0;
      for (int r = threadIdx.x; r < m; r += kTdThreads) {
        const cplx<double> c = cmulc(p[r], v[r]);
        sx += c.x;
        sy += c.y;
      }
      sx = block_sum(sx, red);
      sy = block_sum(sy, red);
      const cplx<double> alpha = tj * mk<double>(-0.5 * sx, -0.5 * sy);
      for (int r = threadIdx.x; r < m; r += kTdThreads) w[r] = p[r] + alpha * v[r];
      __syncthreads();
      const cplx<double> v0 = v[0], w0 = w[0];
      if (threadIdx.x == 0) {
        const cplx<double> a11 = A[(long long)(j + 1) * np + (j + 1)];
        a.d[j + 1] = a11.x - 2.0 * (v0.x * w0.x + v0.y * w0.y);
      }
      if (next) {
        // column j+1 below the diagonal after this update, from row j+1
        cplx<double>* x = V(q1);
        const cplx<double>* row = A + (long long)(j + 1) * np + (j + 2);
        for (int r = threadIdx.x; r < m - 1; r += kTdThreads)
          x[r] = conj(row[r]) - v[1 + r] * conj(w0) - w[1 + r] * conj(v0);
        __syncthreads();
        cplx<double> tau;
        double beta;
        reflector(x, m - 1, tau, beta, red);
        if (threadIdx.x == 0) {
          a.tau[q1] = tau;
          a.e[j + 1] = beta;
        }
      }
    }
    if (!next) break;
    if (barrier()) return;
    // trailing block B = A[j+1:, j+1:]: rows and columns 1.. updated, and
    // p_{j+1} = tau_{j+1} B[1:, 1:] v_{j+1} from the updated values
    {
      const cplx<double>* v = V(q);
      const cplx<double>* w = W(q);
      const cplx<double>* vn = V(q1);
      cplx<double>* pn = Pv(q1);
      const cplx<double> tn = a.tau[q1];
      for (int r0 = 1 + gwave * kRows; r0 < m; r0 += nwaves * kRows) {
        // rows r0 .. r0 + kRows - 1 of B (a row past the end repeats the last
        // one's reads and writes nothing)
        double sx[kRows], sy[kRows];
        cplx<double>* row[kRows];
        cplx<double> vr[kRows], wr[kRows];
        bool live[kRows];
#pragma unroll
        for (int i = 0; i < kRows; i++) {
          const int r = std::min(r0 + i, m - 1);
          live[i] = r0 + i < m;
          sx[i] = sy[i] = 0;
          row[i] = A + (long long)(j + 1 + r) * np + (j + 1);
          vr[i] = v[r];
          wr[i] = w[r];
        }
#pragma unroll 4
        for (int c = 1 + lane; c < m; c += 64) {
          const cplx<double> wc = conj(w[c]), vc = conj(v[c]), vnc = vn[c - 1];
#pragma unroll
          for (int i = 0; i < kRows; i++) {
            const cplx<double> b = row[i][c] - vr[i] * wc - wr[i] * vc;
            if (live[i]) row[i][c] = b;
            const cplx<double> t = b * vnc;
            sx[i] += t.x;
            sy[i] += t.y;
          }
        }
#pragma unroll
        for (int i = 0; i < kRows; i++) {
          const double x = wave_sum(sx[i]), y = wave_sum(sy[i]);
          if (lane == 0 && live[i]) pn[r0 + i - 1] = tn * mk<double>(x, y);
        }
      }
    }
    if (barrier()) return;
  }
}

// Sturm count: eigenvalues of T below x (LDL^T pivots of T - x I).
__device__ __forceinline__ int sturm_count(const double* d, const double* e2, int n, double x, double pivmin) {
  int cnt = 0;
  double q = d[0] - x;
  if (std::fabs(q) < pivmin) q = -pivmin;
  cnt += q < 0;
  for (int i = 1; i < n; i++) {
    q = d[i] - x - e2[i - 1] / q;
    if (std::fabs(q) < pivmin) q = -pivmin;
    cnt += q < 0;
  }
  return cnt;
}

__global__ __launch_bounds__(256) void bisect_kernel(const double* d, const double* e, int n, double* lam) {
  extern __shared__ double sh[];
  double* sd = sh;
  double* se2 = sh + n;
  __shared__ double bounds[3];
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    sd[i] = d[i];
    se2[i] = i < n - 1 ? e[i] * e[i] : 0.0;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    // Gershgorin interval and the pivot bound (dstebz)
    double lo = DBL_MAX, hi = -DBL_MAX, emax = 0;
    for (int i = 0; i < n; i++) {
      const double r = (i > 0 ? std::sqrt(se2[i - 1]) : 0.0) + (i < n - 1 ? std::sqrt(se2[i]) : 0.0);
      lo = std::fmin(lo, sd[i] - r);
      hi = std::fmax(hi, sd[i] + r);
      emax = std::fmax(emax, se2[i]);
    }
    const double pad = 2.0 * DBL_EPSILON * std::fmax(std::fabs(lo), std::fabs(hi)) * n + 2.2250738585072014e-308;
    bounds[0] = lo - pad;
    bounds[1] = hi + pad;
    bounds[2] = 2.2250738585072014e-308 * std::fmax(1.0, emax);
  }
  __syncthreads();
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  double lo = bounds[0], hi = bounds[1];
  const double pivmin = bounds[2];
  // smallest x with more than k eigenvalues below it
  for (int it = 0; it < 200; it++) {
    const double mid = 0.5 * (lo + hi);
    if (mid <= lo || mid >= hi) break;
    if (hi - lo <= 2.0 * DBL_EPSILON * std::fmax(std::fabs(lo), std::fabs(hi)) + pivmin) break;
    if (sturm_count(sd, se2, n, mid, pivmin) > k) hi = mid;
    else lo = mid;
  }
  // an interval around zero (width at the pivot bound) is an exact zero
  lam[k] = (lo <= 0.0 && hi >= 0.0) ? 0.0 : 0.5 * (lo + hi);
}

}  // namespace

size_t hermitian_eigenvalues_scratch(int np, int N) {
  // work copy, vectors, tau, d, e, then the barrier counter and error flag
  return (size_t)np * np * 16 + (size_t)6 * N * 16 + 2 * 16 + (size_t)2 * N * 8 + 64;
}

hipError_t launch_hermitian_eigenvalues(const cplx<double>* A, int np, int N, void* scratch, double* lam,
                                        unsigned spin_limit, hipStream_t st) {
  if (N < 1 || np < N) return hipErrorInvalidValue;
  if ((size_t)2 * N * sizeof(double) > 64 * 1024) return hipErrorInvalidValue;  // bisection stages T in LDS
  char* p = (char*)scratch;
  TridiagArgs a;
  a.A = (cplx<double>*)p;
  p += (size_t)np * np * 16;
  a.vec = (cplx<double>*)p;
  p += (size_t)6 * N * 16;
  a.tau = (cplx<double>*)p;
  p += 2 * 16;
  a.d = (double*)p;
  p += (size_t)N * 8;
  a.e = (double*)p;
  p += (size_t)N * 8;
  a.bar = (unsigned*)p;
  a.np = np;
  a.N = N;
  a.spin_limit = spin_limit;
  hipError_t e = hipMemcpyAsync(a.A, A, (size_t)np * np * 16, hipMemcpyDeviceToDevice, st);
  if (e != hipSuccess) return e;
  // barrier words, and e (its last entry is never written)
  hipLaunchKernelGGL(td_zero_kernel, dim3(1), dim3(64), 0, st, reinterpret_cast<unsigned long long*>(a.bar), 1LL);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL(td_zero_kernel, dim3(1), dim3(256), 0, st, reinterpret_cast<unsigned long long*>(a.e), (long long)N);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  int per_cu = 0, dev = 0, ncu = 0;
  if ((e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, tridiag_kernel, kTdThreads, 0)) != hipSuccess) return e;
  if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    return hipErrorInvalidValue;
  // one workgroup per CU: the passes are bandwidth-bound, and fewer
  // workgroups make the grid barriers cheaper
  const int grid = std::max(1, std::min(ncu, per_cu * ncu));
  void* args[] = {&a};
  e = hipLaunchCooperativeKernel((const void*)tridiag_kernel, dim3(grid), dim3(kTdThreads), args, 0, st);
  if (e != hipSuccess) return e;
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL(bisect_kernel, dim3((N + 255) / 256), dim3(256), (size_t)2 * N * sizeof(double), st, a.d, a.e, N,
                     lam);
  return hipGetLastError();
}

size_t hermitian_eigenvalues_barrier_offset(int np, int N) {
  return (size_t)np * np * 16 + (size_t)6 * N * 16 + 2 * 16 + (size_t)2 * N * 8;
}

}  // namespace itp
