// Cholesky factor and its inverse of the N x N overlap matrix, for the
// Cholesky orthonormalization Psi <- Psi R^-1 (S = R^H R, R upper
// triangular). No counterpart in the reference, which orthonormalizes through
// the eigendecomposition of S (stateset.cpp:136-232); the mixed-precision path
// uses this between its Rayleigh-Ritz steps, as the CholeskyQR factorization.
//
// Blocked right-looking factorization on 32 x 32 tiles of the np x np
// row-major matrix (np a multiple of 32), in 2 launches per block column k:
//   A_k: panel  R_kj = R_kk^-H S'_kj       (j > k)
//        and    W_Jk = W'_Jk W_kk          (J < k)
//   B_k: update S'_ij -= R_ki^H R_kj       (k < i <= j)
//        and    W'_Ji -= W_Jk R_ki         (J <= k < i)
//        and, by the workgroup of tile (k+1, k+1), that tile's factorization
//        R_kk and inverse W_kk in LDS.
// W = R^-1 builds up with the factorization: W'_Jk accumulates
// -sum_{J<=l<k} W_Jl R_lk, so W_Jk = -(sum W_Jl R_lk) W_kk as W R = I
// requires, and every tile of a launch is independent (no serial chains).
// Tile products run on v_mfma_f64_16x16x4f64 (bj_cgemm32's lane layout).
// Rows and columns at or beyond N are the identity. A pivot below tol times
// the row's original diagonal (or not finite) sets *fail; later launches then
// do nothing and the caller falls back to the eigendecomposition.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "block_jacobi.hpp"
#include "gpu.hpp"

namespace itp {

namespace {

constexpr int T = JS;  // tile side (32)
constexpr int T2 = T * T;
typedef double d4 __attribute__((ext_vector_type(4)));

struct CholArgs {
  const cplx<double>* S;  // np x np overlap (upper triangle read)
  cplx<double>* R;        // np x np: the factor (upper tiles)
  cplx<double>* W;        // nc x nc: R^-1 (upper), zero elsewhere
  double* diag0;          // np: original diagonal
  int* fail;
  int np, nc, N, nt;
  double tol;
};

__device__ __forceinline__ long long tile_off(int ti, int tj, int ld) { return (long long)ti * T * ld + (long long)tj * T; }

// Tile copies between global memory and LDS by the kernels' 256 threads;
// each thread's four loads are issued together (and the tiles a kernel loads
// one after another overlap, the loads being independent of the LDS stores).
constexpr int kTileThreads = 256;
__device__ __forceinline__ void load_tile(const cplx<double>* G, int ld, int ti, int tj, cplx<double>* L) {
  const cplx<double>* g = G + tile_off(ti, tj, ld);
  cplx<double> v[T2 / kTileThreads];
#pragma unroll
  for (int r = 0; r < T2 / kTileThreads; r++) {
    const int e = threadIdx.x + kTileThreads * r;
    v[r] = g[(long long)(e / T) * ld + (e % T)];
  }
#pragma unroll
  for (int r = 0; r < T2 / kTileThreads; r++) L[threadIdx.x + kTileThreads * r] = v[r];
}

__device__ __forceinline__ void store_tile(cplx<double>* G, int ld, int ti, int tj, const cplx<double>* L) {
  cplx<double>* g = G + tile_off(ti, tj, ld);
#pragma unroll
  for (int r = 0; r < T2 / kTileThreads; r++) {
    const int e = threadIdx.x + kTileThreads * r;
    g[(long long)(e / T) * ld + (e % T)] = L[e];
  }
}

// acc += op(A) B for LDS tiles (op = conjugate transpose if ah); each
// wavefront holds one 16 x 16 quarter (wi, wj) of the result: register v of
// lane l is element (wi*16 + l/16 + 4v, wj*16 + l%16).
__device__ __forceinline__ void mma32(const cplx<double>* A, bool ah, const cplx<double>* B, d4& cr, d4& ci) {
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wi = w >> 1, wj = w & 1;
  const int li = lane & 15, lk = lane >> 4;
#pragma unroll 2
  for (int kk = 0; kk < T / 4; kk++) {
    const int kx = kk * 4 + lk;
    const cplx<double> av = ah ? conj(A[kx * T + wi * 16 + li]) : A[(wi * 16 + li) * T + kx];
    const cplx<double> bv = B[kx * T + wj * 16 + li];
    cr = __builtin_amdgcn_mfma_f64_16x16x4f64(av.x, bv.x, cr, 0, 0, 0);
    cr = __builtin_amdgcn_mfma_f64_16x16x4f64(-av.y, bv.y, cr, 0, 0, 0);
    ci = __builtin_amdgcn_mfma_f64_16x16x4f64(av.x, bv.y, ci, 0, 0, 0);
    ci = __builtin_amdgcn_mfma_f64_16x16x4f64(av.y, bv.x, ci, 0, 0, 0);
  }
}

// LDS tile C = s * acc
__device__ __forceinline__ void acc_store(const d4& cr, const d4& ci, double s, cplx<double>* C) {
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wi = w >> 1, wj = w & 1;
  const int li = lane & 15, lk = lane >> 4;
  for (int v = 0; v < 4; v++) C[(wi * 16 + lk + 4 * v) * T + wj * 16 + li] = mk<double>(s * cr[v], s * ci[v]);
}

// Factor the LDS tile a (updated S_kk, upper triangle valid) in place into
// R_kk and write W_kk = R_kk^-1 into inv (lin: LDS workspace). The
// elimination runs on [S_kk | I]: the same row operations that turn S_kk into
// R_kk turn I into L^-1 = R^-H, so the inverse needs no serial triangular
// solve. Row r of the tile is global index k*T + r; the pivot test uses the
// original diagonal. The elimination keeps the tile in registers (thread t:
// column t % 32 of rows t / 32 + 8 s) and passes only the pivot row through
// LDS, double-buffered: one barrier per column. Needs 256 threads.
__device__ __forceinline__ cplx<double> pick4(cplx<double> v0, cplx<double> v1, cplx<double> v2, cplx<double> v3,
                                               int s) {
  return s == 0 ? v0 : s == 1 ? v1 : s == 2 ? v2 : v3;
}
__device__ void factor_diag(cplx<double>* a, cplx<double>* lin, cplx<double>* inv, const CholArgs& p, int k) {
  __shared__ double piv[T];
  __shared__ double ref0[T];  // the rows' original diagonal (read once, not per column)
  __shared__ cplx<double> rowA[2][T], rowL[2][T];
  __shared__ int bad;
  const int tid = threadIdx.x;
  const int col = tid & (T - 1), grp = tid >> 5;
  cplx<double> ra[4], rl[4];
#pragma unroll
  for (int s = 0; s < 4; s++) {
    const int r = grp + 8 * s;
    ra[s] = a[r * T + col];
    rl[s] = mk<double>(r == col ? 1.0 : 0.0, 0.0);
  }
  if (tid == 0) bad = 0;
  if (tid < T) ref0[tid] = p.diag0[k * T + tid];
  if (grp == 0) {
    rowA[0][col] = ra[0];
    rowL[0][col] = rl[0];
  }
  __syncthreads();
  for (int c = 0; c < T; c++) {
    const int b = c & 1;
    // every LDS read of the column first and unconditionally, so that they
    // are in flight together (one wait instead of one per row)
    const double d = rowA[b][c].x;
    const double ref = ref0[c];
    const cplx<double> sc = rowA[b][col], lc = rowL[b][col];
    cplx<double> aci[4];
#pragma unroll
    for (int s = 0; s < 4; s++) aci[s] = rowA[b][grp + 8 * s];
    const bool ok = std::isfinite(d) & (d > p.tol * ref) & (d > 0);
    // the pivot's square root waits for the end (off the column chain)
    if (tid == 0) {
      piv[c] = ok ? d : -1.0;
      if (!ok) bad = 1;
    }
    const double rd = 1.0 / d;
    const double invd = ok ? rd : 0.0;
    // rows i > c: S part a_ij -= conj(a_ci) a_cj / d (j >= i), identity part
    // lin_ij -= conj(a_ci) lin_cj / d (j <= c); this thread's j is col
#pragma unroll
    for (int s = 0; s < 4; s++) {
      const int i = grp + 8 * s;
      if (i > c) {
        const cplx<double> m = conj(aci[s]) * invd;
        if (col >= i) ra[s] = ra[s] - m * sc;
        else if (col <= c) rl[s] = rl[s] - m * lc;
      }
    }
    // row c + 1 is final now: the next column's pivot row
    if (c + 1 < T && ((c + 1) & 7) == grp) {
      rowA[b ^ 1][col] = pick4(ra[0], ra[1], ra[2], ra[3], (c + 1) >> 3);
      rowL[b ^ 1][col] = pick4(rl[0], rl[1], rl[2], rl[3], (c + 1) >> 3);
    }
    __syncthreads();
  }
#pragma unroll
  for (int s = 0; s < 4; s++) {
    const int r = grp + 8 * s;
    a[r * T + col] = ra[s];
    lin[r * T + col] = rl[s];
  }
  if (tid < T) {
    const double sq = piv[tid] > 0 ? std::sqrt(piv[tid]) : 1.0;
    piv[tid] = sq;
    ref0[tid] = 1.0 / sq;  // ref0 is free now: the reciprocal square roots
  }
  __syncthreads();
  // R: rows scaled by 1 / sqrt(pivot), lower triangle zero; W = (L^-1)^H with
  // L^-1 the scaled identity part
  for (int e = tid; e < T2; e += blockDim.x) {
    const int i = e / T, j = e % T;
    if (j < i) a[e] = mk<double>(0.0, 0.0);
    else if (j == i) a[e] = mk<double>(piv[i], 0.0);
    else a[e] = a[e] * ref0[i];
    inv[e] = j < i ? mk<double>(0.0, 0.0) : conj(lin[j * T + i]) * ref0[j];
  }
  __syncthreads();
  if (tid == 0 && bad) atomicOr(p.fail, 1);
}

// R = upper(S) with the identity past N, the original diagonal, W = 0 and the
// failure flag cleared (a kernel write, so the sequence is graph-safe).
__global__ __launch_bounds__(256) void chol_init_kernel(CholArgs p) {
  const long long tot = (long long)p.nc * p.nc;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < tot;
       e += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(e / p.nc), j = (int)(e % p.nc);
    p.W[e] = mk<double>(0.0, 0.0);
    if (i >= p.np || j >= p.np) continue;
    cplx<double> v = mk<double>(0.0, 0.0);
    if (i < p.N && j < p.N && j >= i) v = p.S[(long long)i * p.np + j];
    if (i == j) {
      v = i < p.N ? mk<double>(v.x, 0.0) : mk<double>(1.0, 0.0);
      p.diag0[i] = v.x;
    }
    p.R[(long long)i * p.np + j] = v;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) *p.fail = 0;
}

// B_k: trailing update for block column k (k = -1: none), the update of
// W's accumulators, and the factorization of tile (k+1, k+1).
__global__ __launch_bounds__(256) void chol_update_kernel(CholArgs p, int k) {
  if (__hip_atomic_load(p.fail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
  extern __shared__ __align__(16) char lds[];
  cplx<double>* A = reinterpret_cast<cplx<double>*>(lds);
  cplx<double>* B = A + T2;
  cplx<double>* C = B + T2;
  const int nt = p.nt;
  const int ntrail = (nt - k - 1) * (nt - k) / 2;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wi = w >> 1, wj = w & 1, li = lane & 15, lk = lane >> 4;
  if ((int)blockIdx.x >= ntrail) {
    // W'_Ji -= W_Jk R_ki, J <= k < i
    const int t = (int)blockIdx.x - ntrail;
    const int J = t / (nt - k - 1), i = k + 1 + t % (nt - k - 1);
    load_tile(p.W, p.nc, J, k, A);
    load_tile(p.R, p.np, k, i, B);
    load_tile(p.W, p.nc, J, i, C);
    __syncthreads();
    d4 cr = {0, 0, 0, 0}, ci = {0, 0, 0, 0};
    mma32(A, false, B, cr, ci);
    for (int v = 0; v < 4; v++) {
      const int e = (wi * 16 + lk + 4 * v) * T + wj * 16 + li;
      C[e] = C[e] - mk<double>(cr[v], ci[v]);
    }
    __syncthreads();
    store_tile(p.W, p.nc, J, i, C);
    return;
  }
  // tile t -> (i, j), k < i <= j < nt, row by row
  int t = blockIdx.x, i = k + 1;
  while (t >= nt - i) { t -= nt - i; i++; }
  const int j = i + t;
  load_tile(p.R, p.np, i, j, C);
  if (k >= 0) {
    load_tile(p.R, p.np, k, i, A);
    load_tile(p.R, p.np, k, j, B);
    __syncthreads();
    d4 cr = {0, 0, 0, 0}, ci = {0, 0, 0, 0};
    mma32(A, true, B, cr, ci);
    for (int v = 0; v < 4; v++) {
      const int e = (wi * 16 + lk + 4 * v) * T + wj * 16 + li;
      C[e] = C[e] - mk<double>(cr[v], ci[v]);
    }
  }
  __syncthreads();
  if (i == k + 1 && j == k + 1) {
    factor_diag(C, B, A, p, i);
    store_tile(p.R, p.np, i, i, C);
    store_tile(p.W, p.nc, i, i, A);
  } else {
    store_tile(p.R, p.np, i, j, C);
  }
}

// A_k: panel row k of R and block column k of W.
__global__ __launch_bounds__(256) void chol_panel_kernel(CholArgs p, int k) {
  if (__hip_atomic_load(p.fail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
  extern __shared__ __align__(16) char lds[];
  cplx<double>* A = reinterpret_cast<cplx<double>*>(lds);
  cplx<double>* B = A + T2;
  cplx<double>* C = B + T2;
  const int npanel = p.nt - k - 1;
  const int b = blockIdx.x;
  load_tile(p.W, p.nc, k, k, C);  // W_kk = R_kk^-1
  d4 cr = {0, 0, 0, 0}, ci = {0, 0, 0, 0};
  if (b < npanel) {
    // R_kj = W_kk^H S'_kj
    const int j = k + 1 + b;
    load_tile(p.R, p.np, k, j, B);
    __syncthreads();
    mma32(C, true, B, cr, ci);
    __syncthreads();
    acc_store(cr, ci, 1.0, A);
    __syncthreads();
    store_tile(p.R, p.np, k, j, A);
    return;
  }
  // W_Jk = W'_Jk W_kk
  const int J = b - npanel;  // 0 <= J < k
  load_tile(p.W, p.nc, J, k, B);
  __syncthreads();
  mma32(B, false, C, cr, ci);
  __syncthreads();
  acc_store(cr, ci, 1.0, A);
  __syncthreads();
  store_tile(p.W, p.nc, J, k, A);
}

// out[n] = sum |f_n|^2 over the M elements of field n (one workgroup per field)
__global__ __launch_bounds__(256) void field_norms_kernel(const cplx<double>* f, long long M, double* out) {
  __shared__ double red[256];
  const cplx<double>* g = f + (long long)blockIdx.x * M;
  double s = 0;
  for (long long i = threadIdx.x; i < M; i += blockDim.x) s += g[i].x * g[i].x + g[i].y * g[i].y;
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}

// mm[0] = max, mm[1] = min of R_ii^2 over i < N, mm[2 + i] = R_ii^2 (one workgroup)
__global__ __launch_bounds__(256) void chol_diag_range_kernel(const cplx<double>* R, int np, int N, double* mm) {
  __shared__ double hi[256], lo[256];
  double h = 0, l = 1e300;
  for (int i = threadIdx.x; i < N; i += blockDim.x) {
    const double r = R[(long long)i * np + i].x;
    h = fmax(h, r * r);
    l = fmin(l, r * r);
    mm[2 + i] = r * r;
  }
  hi[threadIdx.x] = h;
  lo[threadIdx.x] = l;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      hi[threadIdx.x] = fmax(hi[threadIdx.x], hi[threadIdx.x + w]);
      lo[threadIdx.x] = fmin(lo[threadIdx.x], lo[threadIdx.x + w]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) { mm[0] = hi[0]; mm[1] = lo[0]; }
}

}  // namespace

hipError_t launch_chol_diag_range(const cplx<double>* R, int np, int N, double* mm, hipStream_t st) {
  hipLaunchKernelGGL(chol_diag_range_kernel, dim3(1), dim3(256), 0, st, R, np, N, mm);
  return hipGetLastError();
}

hipError_t launch_field_norms(const cplx<double>* f, long long M, int N, double* out, hipStream_t st) {
  if (N < 1) return hipSuccess;
  hipLaunchKernelGGL(field_norms_kernel, dim3(N), dim3(256), 0, st, f, M, out);
  return hipGetLastError();
}

hipError_t launch_cholinv(const cplx<double>* S, int np, int N, cplx<double>* R, double* diag0, cplx<double>* W,
                          int nc, int* fail, double tol, hipStream_t st) {
  if (np % T || nc < np || N > np || N < 1) return hipErrorInvalidValue;
  CholArgs p;
  p.S = S; p.R = R; p.W = W; p.diag0 = diag0; p.fail = fail;
  p.np = np; p.nc = nc; p.N = N; p.nt = np / T; p.tol = tol;
  const long long tot = (long long)nc * nc;
  const int ib = (int)std::min<long long>(4096, (tot + 255) / 256);
  hipLaunchKernelGGL(chol_init_kernel, dim3(ib), dim3(256), 0, st, p);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const size_t lds = 3 * T2 * sizeof(cplx<double>);
  const int nt = p.nt;
  // B_{-1}: factor tile (0, 0)
  hipLaunchKernelGGL(chol_update_kernel, dim3(1), dim3(256), lds, st, p, -1);
  for (int k = 0; k < nt; k++) {
    const int na = (nt - k - 1) + k;
    if (na > 0) hipLaunchKernelGGL(chol_panel_kernel, dim3(na), dim3(256), lds, st, p, k);
    const int nb = (nt - k - 1) * (nt - k) / 2 + (k + 1) * (nt - k - 1);
    if (nb > 0) hipLaunchKernelGGL(chol_update_kernel, dim3(nb), dim3(256), lds, st, p, k);
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  return hipSuccess;
}

}  // namespace itp
