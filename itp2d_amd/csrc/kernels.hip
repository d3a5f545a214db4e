// gfx950 kernels of the ITP loop.
//
//  sweep_kernel      one workgroup = L lines of one field; the line program of
//                    line_ops.hpp (LDS-resident FFT / DST / DCT passes fused
//                    with the multiplier tables and the epilogue)
//  energy_reduce     E and sigma per state from the sweep partial sums
//  gram_kernel       overlap S = Psi^H Phi on FP64 MFMA (v_mfma_f64_16x16x4f64),
//                    64x64 tiles of state pairs, split over the grid points
//  gram_reduce       sums the split-K partials, scales by dx^2, fills the
//                    Hermitian matrix for the eigensolver
//  block_jacobi      cooperative block two-sided Jacobi on the overlap matrix
//  zgemm / oa_*      refinement of the Jacobi eigenvectors by n x n MFMA products
//  eig_sort / eig_coeffs  ascending eigenvalues, C = U Lambda^(-1/2)
//  rotate_kernel     Psi' = Psi C on FP64 MFMA
#include <hip/hip_runtime.h>
#include <hip/hip_cooperative_groups.h>

#include "block_jacobi.hpp"
#include "cluster_eig.hpp"
#include "dense_ops.hpp"
#include "gpu.hpp"
#include "grid_sync.hpp"
#include "line_ops.hpp"
#include "potentials.hpp"
#include "reg_sweep.hpp"

namespace itp {

namespace cg = cooperative_groups;
typedef double d4 __attribute__((ext_vector_type(4)));

struct GpuExec {
  template <class F>
  __device__ void run(F&& f) {
    f((int)threadIdx.x, (int)blockDim.x);
    __syncthreads();
  }
};

template <class R>
__global__ __launch_bounds__(256) void sweep_kernel(SweepArgs<R> a) {
  extern __shared__ __align__(16) char lds_raw[];
  SweepLds<R> s = carve_lds<R>(lds_raw, a.L, a.plan.n, (int)blockDim.x);
  GpuExec ex;
  sweep_program<R>(ex, a, s, (int)blockIdx.y, (int)blockIdx.x);
}

template <class R>
hipError_t launch_reg_sweep(const SweepArgs<R>& a, int nblk, int nfields, hipStream_t st) {
  const int n = a.plan.n;
  if (!reg_supported(n)) return hipErrorInvalidValue;
  const size_t lds = reg_lds_bytes<R>(n);
  if (lds > 160 * 1024 || a.L != reg_lines_per_wg(n)) return hipErrorInvalidValue;
  const int lines = a.orient == 0 ? a.sy : a.sx;
  if ((a.orient == 0 ? a.sx : a.sy) != n || nblk * a.L != lines) return hipErrorInvalidValue;
  // the magnetic Dirichlet split runs along rows only
  if (a.orient != 0 && ops_have_xsplit(a.ops, a.nops)) return hipErrorInvalidValue;
  switch (n) {
    case 32: return launch_reg_sweep_n<R, 32>(a, nblk, nfields, st);
    case 64: return launch_reg_sweep_n<R, 64>(a, nblk, nfields, st);
    case 128: return launch_reg_sweep_n<R, 128>(a, nblk, nfields, st);
    case 192: return launch_reg_sweep_n<R, 192>(a, nblk, nfields, st);
    case 256: return launch_reg_sweep_n<R, 256>(a, nblk, nfields, st);
    case 512: return launch_reg_sweep_n<R, 512>(a, nblk, nfields, st);
    case 1024: return launch_reg_sweep_n<R, 1024>(a, nblk, nfields, st);
    default: return hipErrorInvalidValue;
  }
}

__global__ void zero_ints_kernel(int* p, int n) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0;
}
hipError_t launch_zero_ints(int* p, int n, hipStream_t st) {
  hipLaunchKernelGGL(zero_ints_kernel, dim3(1), dim3(64), 0, st, p, n);
  return hipGetLastError();
}

__global__ __launch_bounds__(256) void identity_kernel(cplx<double>* A, int np) {
  const long long n2 = (long long)np * np;
  for (long long e = blockIdx.x * 256LL + threadIdx.x; e < n2; e += (long long)gridDim.x * 256)
    A[e] = mk<double>(e / np == e % np ? 1.0 : 0.0, 0.0);
}
hipError_t launch_identity(cplx<double>* A, int np, hipStream_t st) {
  if (np < 1) return hipErrorInvalidValue;
  const long long n2 = (long long)np * np;
  const unsigned blocks = (unsigned)std::min<long long>((n2 + 255) / 256, 4096);
  hipLaunchKernelGGL(identity_kernel, dim3(blocks), dim3(256), 0, st, A, np);
  return hipGetLastError();
}

template <class R>
hipError_t launch_sweep(const SweepArgs<R>& a, int nblk, int nfields, hipStream_t st) {
  if (nblk <= 0 || nfields <= 0) return hipSuccess;
  if (nfields > 65535) return hipErrorInvalidValue;
  if (a.reg) return launch_reg_sweep<R>(a, nblk, nfields, st);
  const size_t lds = sweep_lds_bytes<R>(a.L, a.plan.n, 256);
  if (lds > 160 * 1024) return hipErrorInvalidValue;
  // allow the full 160 KiB of LDS per workgroup (once, thread-safe static initialisation)
  static const hipError_t attr_set =
      hipFuncSetAttribute((const void*)sweep_kernel<R>, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
  (void)attr_set;
  if (nblk <= 0 || nfields <= 0) return hipSuccess;
  if (nfields > 65535) return hipErrorInvalidValue;
  hipLaunchKernelGGL(sweep_kernel<R>, dim3(nblk, nfields), dim3(256), lds, st, a);
  return hipGetLastError();
}
template hipError_t launch_sweep<double>(const SweepArgs<double>&, int, int, hipStream_t);
template hipError_t launch_sweep<float>(const SweepArgs<float>&, int, int, hipStream_t);

__global__ void energy_reduce_kernel(const cplx<double>* p1, const cplx<double>* p2, int nblk, int N, double dx,
                                     double* E, double* sigma, cplx<double>* mean, cplx<double>* msq) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  energy_from_partials(p1 + (long long)n * nblk, p2 + (long long)n * nblk, nblk, nblk, dx, E + n, sigma + n,
                       mean + n, msq + n);
}

hipError_t launch_energy_reduce(const cplx<double>* p1, const cplx<double>* p2, int nblk, int N, double dx,
                                double* E, double* sigma, cplx<double>* mean, cplx<double>* msq, hipStream_t st) {
  hipLaunchKernelGGL(energy_reduce_kernel, dim3((N + 127) / 128), dim3(128), 0, st, p1, p2, nblk, N, dx, E, sigma,
                     mean, msq);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Precision conversion for FP32 state sets.
// ---------------------------------------------------------------------------
__global__ void f64_to_f32_kernel(const cplx<double>* in, cplx<float>* out, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    out[i] = mk<float>(float(in[i].x), float(in[i].y));
}
__global__ void f32_to_f64_kernel(const cplx<float>* in, cplx<double>* out, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    out[i] = mk<double>(double(in[i].x), double(in[i].y));
}
hipError_t launch_convert(const void* in, void* out, long long n, int to_f32, hipStream_t st) {
  const int blocks = (int)std::min<long long>((n + 255) / 256, 4096);
  if (n <= 0) return hipSuccess;
  if (to_f32) hipLaunchKernelGGL(f64_to_f32_kernel, dim3(blocks), dim3(256), 0, st, (const cplx<double>*)in, (cplx<float>*)out, n);
  else hipLaunchKernelGGL(f32_to_f64_kernel, dim3(blocks), dim3(256), 0, st, (const cplx<float>*)in, (cplx<double>*)out, n);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Potentials (potentials.hpp): V of a potential kind on the cell-centred
// grid, and impurities added point by point in impurity order (the order of
// the reference's loop over impurities, noise.cpp:80-110).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void potential_kernel(PotParams q, int sx, int sy, double dx, double* V) {
  const long long M = (long long)sx * sy;
  for (long long i = blockIdx.x * 256ll + threadIdx.x; i < M; i += (long long)gridDim.x * 256) {
    const long long y = i / sx, x = i - y * sx;
    V[i] = pot_value(q, grid_pos(x, sx, dx), grid_pos(y, sy, dx));
  }
}

__global__ __launch_bounds__(256) void impurity_kernel(int kind, const double* data, int stride, long long count,
                                                       double g0, double g1, int sx, int sy, double dx, double* V) {
  const long long M = (long long)sx * sy;
  const double g[2] = {g0, g1};
  for (long long i = blockIdx.x * 256ll + threadIdx.x; i < M; i += (long long)gridDim.x * 256) {
    const long long y = i / sx, x = i - y * sx;
    const double px = grid_pos(x, sx, dx), py = grid_pos(y, sy, dx);
    double v = V[i];
    for (long long n = 0; n < count; n++) {
      const double* d = data + n * stride;
      v += imp_value(kind, d + 2, g, d[0], d[1], px, py);
    }
    V[i] = v;
  }
}

hipError_t launch_potential(const PotParams& q, int sx, int sy, double dx, double* V, hipStream_t st) {
  const long long M = (long long)sx * sy;
  const int blocks = (int)std::min<long long>((M + 255) / 256, 8192);
  hipLaunchKernelGGL(potential_kernel, dim3(blocks), dim3(256), 0, st, q, sx, sy, dx, V);
  return hipGetLastError();
}

hipError_t launch_impurities(int kind, const double* data, int stride, long long count, double g0, double g1, int sx,
                             int sy, double dx, double* V, hipStream_t st) {
  const long long M = (long long)sx * sy;
  if (count <= 0) return hipSuccess;
  const int blocks = (int)std::min<long long>((M + 255) / 256, 8192);
  hipLaunchKernelGGL(impurity_kernel, dim3(blocks), dim3(256), 0, st, kind, data, stride, count, g0, g1, sx, sy, dx, V);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Seeded uniform-noise initial states on the device.
//
// The host generator draws x_k = splitmix64 output number k + 1 of the stream
// starting at s0, i.e. mix(s0 + (k + 1) * gamma); each draw is independent of
// the others, so state n, element i takes draws 2 (n M + i) and 2 (n M + i) + 1
// directly. Values are 2 u - 1 with u = (x >> 11) 2^-53 (exact in FP64), each
// state scaled to dx-norm 1 in FP64 and then rounded to R. Pass 1 writes
// per-block sums of squares (fixed order), pass 2 sums them per state in
// block order and writes the scaled values.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double noise_value(uint64_t s0, unsigned long long k) {
  uint64_t z = s0 + (k + 1) * 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  z ^= z >> 31;
  return 2.0 * (double(z >> 11) * (1.0 / 9007199254740992.0)) - 1.0;
}

__global__ __launch_bounds__(256) void noise_norm_kernel(uint64_t s0, long long M, double* part) {
  const int n = blockIdx.y, nb = gridDim.x;
  const unsigned long long base = 2ull * (unsigned long long)n * (unsigned long long)M;
  double acc = 0.0;
  for (long long i = blockIdx.x * 256ll + threadIdx.x; i < M; i += 256ll * nb) {
    const double re = noise_value(s0, base + 2ull * i), im = noise_value(s0, base + 2ull * i + 1);
    acc += re * re + im * im;
  }
  __shared__ double red[256];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[(long long)n * nb + blockIdx.x] = red[0];
}

template <class R>
__global__ __launch_bounds__(256) void noise_write_kernel(uint64_t s0, long long M, double dx, const double* part,
                                                          cplx<R>* out) {
  const int n = blockIdx.y, nb = gridDim.x;
  double acc = 0.0;
  for (int b = 0; b < nb; b++) acc += part[(long long)n * nb + b];
  const double f = 1.0 / (sqrt(acc) * dx);
  const unsigned long long base = 2ull * (unsigned long long)n * (unsigned long long)M;
  cplx<R>* o = out + (long long)n * M;
  for (long long i = blockIdx.x * 256ll + threadIdx.x; i < M; i += 256ll * nb) {
    const double re = noise_value(s0, base + 2ull * i), im = noise_value(s0, base + 2ull * i + 1);
    o[i] = mk<R>(R(re * f), R(im * f));
  }
}

int noise_blocks(long long M) { return (int)std::max<long long>(1, std::min<long long>(64, M / 4096)); }

template <class R>
hipError_t launch_noise(uint64_t s0, long long M, int N, double dx, double* part, cplx<R>* out, hipStream_t st) {
  if (N <= 0 || M <= 0) return hipSuccess;
  const dim3 grid(noise_blocks(M), N);
  hipLaunchKernelGGL(noise_norm_kernel, grid, dim3(256), 0, st, s0, M, part);
  hipLaunchKernelGGL(noise_write_kernel<R>, grid, dim3(256), 0, st, s0, M, dx, (const double*)part, out);
  return hipGetLastError();
}
template hipError_t launch_noise<double>(uint64_t, long long, int, double, double*, cplx<double>*, hipStream_t);
template hipError_t launch_noise<float>(uint64_t, long long, int, double, double*, cplx<float>*, hipStream_t);

// ---------------------------------------------------------------------------
// Overlap matrix on FP64 MFMA.
//
// v_mfma_f64_16x16x4f64: D(16x16) += A(16x4) B(4x16). Lane l holds
// A[l%16][l/16], B[l/16][l%16] and D[l/16 + 4v][l%16], v = 0..3 (measured on
// gfx950 with indicator inputs; the f32 16x16 shapes use 4(l/16)+v instead).
// S_ij = sum_t conj(psi_i(t)) psi_j(t): with a = psi_i, b = psi_j,
// Re += a.x b.x + a.y b.y, Im += a.x b.y - a.y b.x.
// ---------------------------------------------------------------------------
template <class R>
__global__ __launch_bounds__(256) void gram_kernel(const cplx<R>* psi, const cplx<R>* phi, long long M, int N, int nb,
                                                   long long kchunk, cplx<double>* part, int nc, int full) {
  // block order by xcd_tile: every XCD takes whole grid chunks and all tile
  // pairs of a chunk, so the chunk's state slabs are reused from its L2
  const int npairs = full ? nb * nb : nb * (nb + 1) / 2;
  int pair;
  long long chunk;
  if (!xcd_tile(blockIdx.x, npairs, (M + kchunk - 1) / kchunk, pair, chunk)) return;
  int bi = 0, bj = 0;
  if (full) {
    bi = pair / nb;
    bj = pair - bi * nb;
  } else {
    int rem = pair;
    while (rem >= nb - bi) { rem -= nb - bi; bi++; }
    bj = bi + rem;
  }
  __shared__ cplx<double> As[64][17];
  __shared__ cplx<double> Bs[64][17];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, wi = wave >> 1, wj = wave & 1;
  const int li = lane & 15, lk = lane >> 4;
  d4 accr[2][2], acci[2][2];
  for (int a = 0; a < 2; a++)
    for (int b = 0; b < 2; b++)
      for (int v = 0; v < 4; v++) { accr[a][b][v] = 0; acci[a][b][v] = 0; }
  const long long k0 = chunk * kchunk;
  const long long k1 = (k0 + kchunk < M) ? k0 + kchunk : M;
  // thread's staging slots e = tid + 256 r of the 64 x 16 tiles; the next
  // step's values are fetched into registers while this step's products run
  cplx<double> pa[4], pb[4];
  auto fetch = [&](long long t0) {
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int e = tid + 256 * r;
      const int s = e >> 4, k = e & 15;
      const long long t = t0 + k;
      const int ia = bi * 64 + s, ib = bj * 64 + s;
      cplx<double> va = mk<double>(0.0, 0.0), vb = va;
      if (t < k1) {
        if (ia < N) { const cplx<R> x = psi[(long long)ia * M + t]; va = mk<double>(double(x.x), double(x.y)); }
        if (ib < N) { const cplx<R> x = phi[(long long)ib * M + t]; vb = mk<double>(double(x.x), double(x.y)); }
      }
      pa[r] = va;
      pb[r] = vb;
    }
  };
  if (k0 < k1) fetch(k0);
  for (long long t0 = k0; t0 < k1; t0 += 16) {
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int e = tid + 256 * r;
      As[e >> 4][e & 15] = pa[r];
      Bs[e >> 4][e & 15] = pb[r];
    }
    __syncthreads();
    if (t0 + 16 < k1) fetch(t0 + 16);
#pragma unroll
    for (int kk = 0; kk < 4; kk++) {
      cplx<double> a[2], b[2];
      for (int mi = 0; mi < 2; mi++) a[mi] = As[wi * 32 + mi * 16 + li][kk * 4 + lk];
      for (int nj = 0; nj < 2; nj++) b[nj] = Bs[wj * 32 + nj * 16 + li][kk * 4 + lk];
      for (int mi = 0; mi < 2; mi++)
        for (int nj = 0; nj < 2; nj++) {
          accr[mi][nj] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[mi].x, b[nj].x, accr[mi][nj], 0, 0, 0);
          accr[mi][nj] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[mi].y, b[nj].y, accr[mi][nj], 0, 0, 0);
          acci[mi][nj] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[mi].x, b[nj].y, acci[mi][nj], 0, 0, 0);
          acci[mi][nj] = __builtin_amdgcn_mfma_f64_16x16x4f64(-a[mi].y, b[nj].x, acci[mi][nj], 0, 0, 0);
        }
    }
    __syncthreads();
  }
  cplx<double>* P = part + chunk * nc * nc;
  for (int mi = 0; mi < 2; mi++)
    for (int nj = 0; nj < 2; nj++)
      for (int v = 0; v < 4; v++) {
        const int i = bi * 64 + wi * 32 + mi * 16 + lk + 4 * v;
        const int j = bj * 64 + wj * 32 + nj * 16 + li;
        P[(long long)i * nc + j] = mk<double>(accr[mi][nj][v], acci[mi][nj][v]);
      }
}

// ACC partial sums over the chunks (ACC = 1: chunk order, the FP64 overlap's;
// ACC = 4: four interleaved sums, more loads in flight, for the split-BF16x3
// overlap)
template <int ACC>
__global__ void gram_reduce_kernel(const cplx<double>* part, int nchunk, int nc, int N, int np, double dx,
                                   cplx<double>* S, int* bad, int full) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y;
  if (j >= N || i > j) return;
  cplx<double> s = mk<double>(0.0, 0.0);
  if constexpr (ACC == 1) {
    for (int z = 0; z < nchunk; z++) s = s + part[((long long)z * nc + i) * nc + j];
  } else {
    cplx<double> t[ACC];
#pragma unroll
    for (int a = 0; a < ACC; a++) t[a] = mk<double>(0.0, 0.0);
    int z = 0;
    for (; z + ACC <= nchunk; z += ACC)
#pragma unroll
      for (int a = 0; a < ACC; a++) t[a] = t[a] + part[((long long)(z + a) * nc + i) * nc + j];
    for (; z < nchunk; z++) t[0] = t[0] + part[((long long)z * nc + i) * nc + j];
#pragma unroll
    for (int a = 0; a < ACC; a++) s = s + t[a];
  }
  if (full) {  // Hermitian part (G + G^H) / 2 of a matrix computed in full
    cplx<double> t = mk<double>(0.0, 0.0);
    for (int z = 0; z < nchunk; z++) t = t + part[((long long)z * nc + j) * nc + i];
    s = (s + conj(t)) * 0.5;
  }
  s = (s * dx) * dx;
  if (!(std::isfinite(s.x) && std::isfinite(s.y))) atomicOr(bad, 1);
  if (i == j) {
    S[(long long)i * np + i] = mk<double>(s.x, 0.0);
  } else {
    S[(long long)i * np + j] = s;
    S[(long long)j * np + i] = conj(s);
  }
}

template <class R>
hipError_t launch_gram(const cplx<R>* psi, const cplx<R>* phi, long long M, int N, int nc, cplx<double>* part,
                       int nchunk, long long kchunk, double dx, cplx<double>* S, int np, int* bad, hipStream_t st,
                       int full) {
  const int nb = nc / 64;
  const int npairs = full ? nb * nb : nb * (nb + 1) / 2;
  if (nchunk != (M + kchunk - 1) / kchunk) return hipErrorInvalidValue;
  const long long blocks = xcd_grid(npairs, nchunk);
  if (blocks > 0x7fffffffLL) return hipErrorInvalidValue;
  hipLaunchKernelGGL(gram_kernel<R>, dim3((unsigned)blocks), dim3(256), 0, st, psi, phi, M, N, nb, kchunk, part, nc,
                     full);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return launch_gram_reduce(part, nchunk, nc, N, np, dx, S, bad, st, full);
}

hipError_t launch_gram_reduce(const cplx<double>* part, int nchunk, int nc, int N, int np, double dx, cplx<double>* S,
                              int* bad, hipStream_t st, int full, int acc) {
  if (acc == 4)
    hipLaunchKernelGGL(gram_reduce_kernel<4>, dim3((N + 127) / 128, N), dim3(128), 0, st, part, nchunk, nc, N, np, dx,
                       S, bad, full);
  else
    hipLaunchKernelGGL(gram_reduce_kernel<1>, dim3((N + 127) / 128, N), dim3(128), 0, st, part, nchunk, nc, N, np, dx,
                       S, bad, full);
  return hipGetLastError();
}
template hipError_t launch_gram<double>(const cplx<double>*, const cplx<double>*, long long, int, int, cplx<double>*,
                                        int, long long, double, cplx<double>*, int, int*, hipStream_t, int);
template hipError_t launch_gram<float>(const cplx<float>*, const cplx<float>*, long long, int, int, cplx<double>*, int,
                                       long long, double, cplx<double>*, int, int*, hipStream_t, int);

// ---------------------------------------------------------------------------
// Block Jacobi (block_jacobi.hpp), one cooperative launch for all sweeps:
// per round, phase 1 (one workgroup per block pair, sub-problem in LDS) and
// phase 2 (super-tile updates of A and V), each followed by a grid barrier.
// The convergence maxima of a sweep are combined with atomicMax into that
// sweep's own slot, which every workgroup reads after the sweep's last
// barrier, so all workgroups take the same decision.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void block_jacobi_kernel(BlockJacobiArgs a) {
  unsigned* bar = reinterpret_cast<unsigned*>(a.slots + 2 * (a.max_sweeps + 1));
  unsigned* bar_err = bar + 1;
  unsigned epoch = 0;
  extern __shared__ __align__(16) char bjl[];
  const int nt = (int)blockDim.x;
  BJLds s = bj_carve(bjl, nt);
  GpuExec ex;
  const int np = a.np, nb = np / JB, P = np / JS;
  const int items = bj_items(np);
  const long long tot = (long long)np * np;
  const long long per = (tot + gridDim.x - 1) / gridDim.x;
  const long long e0 = (long long)blockIdx.x * per;
  const long long e1 = e0 + per < tot ? e0 + per : tot;
  int sweeps = 0;
  bool conv = false, handed = false, failed = false;
  for (int chk = 0;; chk++) {
    // whole-matrix convergence check
    unsigned long long* slot = reinterpret_cast<unsigned long long*>(a.slots + chk);
    unsigned long long* gslot = reinterpret_cast<unsigned long long*>(a.slots + a.max_sweeps + 1 + chk);
    bj_check_rows(ex, a, e0, e1, s);
    if (threadIdx.x == 0) {
      atomicMax(slot, dbits(s.red[nt]));
      atomicMax(gslot, dbits(s.red[nt + 1]));
    }
    if ((failed = grid_barrier(bar, epoch, bar_err, a.spin_limit))) break;
    const double worst = bitsd(__hip_atomic_load(slot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    const double rho = bitsd(__hip_atomic_load(gslot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    if (worst < 1.0) { conv = true; break; }
    if (a.handoff > 0 && rho < a.handoff) { handed = true; break; }
    if (chk == a.max_sweeps) break;
    // one sweep of nb-1 rounds
    bj_reset_positions(ex, s, nb);
    for (int r = 0; r < nb - 1 && !failed; r++) {
      for (int k = blockIdx.x; k < P; k += gridDim.x) bj_inner(ex, a, k, s.pos2blk[2 * k], s.pos2blk[2 * k + 1], s);
      if ((failed = grid_barrier(bar, epoch, bar_err, a.spin_limit))) break;
      for (int t = blockIdx.x; t < items; t += gridDim.x) bj_apply(ex, a, t, s);
      if ((failed = grid_barrier(bar, epoch, bar_err, a.spin_limit))) break;
      bj_next_positions(ex, s, nb);
    }
    if (failed) break;
    sweeps++;
  }
  // the host also reads the error flag itself after the launch has drained:
  // a workgroup that timed out after workgroup 0 wrote these words still
  // marks the call as failed
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    a.info[0] = sweeps;
    a.info[1] = failed ? -1.0 : (conv ? 1.0 : (handed ? 2.0 : 0.0));
    a.info[2] = 0;
  }
}

__global__ void fill_u64_kernel(unsigned long long* p, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) p[i] = 0ull;
}

hipError_t launch_block_jacobi(cplx<double>* A, cplx<double>* V, cplx<double>* Q, cplx<double>* D, int* ident,
                               double* slots, int np, int max_sweeps, int max_inner, double tol, double handoff,
                               double* info, int grid_wgs, unsigned spin_limit, hipStream_t st) {
  if (np % JS || np / JB > kMaxBlocks) return hipErrorInvalidValue;
  BlockJacobiArgs a;
  a.A = A; a.V = V; a.Q = Q; a.D = D; a.ident = ident; a.slots = slots; a.info = info;
  a.np = np; a.max_sweeps = max_sweeps; a.max_inner = max_inner; a.tol = tol; a.handoff = handoff;
  a.spin_limit = spin_limit;
  // the two maxima per check, then the barrier counter and its error flag
  hipLaunchKernelGGL(fill_u64_kernel, dim3(1), dim3(256), 0, st, reinterpret_cast<unsigned long long*>(slots),
                     (long long)(2 * (max_sweeps + 1) + 1));
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const size_t lds = bj_lds_bytes(256);
  int per_cu = 0;
  e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, block_jacobi_kernel, 256, lds);
  if (e != hipSuccess) return e;
  if (per_cu < 1) return hipErrorInvalidValue;
  int dev = 0, ncu = 0;
  if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    return hipErrorInvalidValue;
  // grid_wgs <= 0: one workgroup per 16 super-tile updates of a round, at
  // least 64 (below that the grid barrier is cheap and the sub-problems of a
  // round are the critical path), at most what is co-resident; from one
  // workgroup per CU on, whole workgroups per CU (with a partial second row,
  // every round waits for the CUs holding two)
  int want = grid_wgs > 0 ? grid_wgs : std::max(64, bj_items(np) / 16);
  if (grid_wgs <= 0 && ncu > 0 && want >= ncu) want = ncu * (want / ncu);
  int grid = std::min(want, per_cu * ncu);
  grid = std::max(1, std::min(grid, bj_items(np)));
  void* args[] = {&a};
  e = hipLaunchCooperativeKernel((const void*)block_jacobi_kernel, dim3(grid), dim3(256), args, (unsigned)lds, st);
  if (e != hipSuccess) return e;
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Refinement of an approximate eigenvector matrix X of the Hermitian A
// (Ogita and Aishima, "Iterative refinement for symmetric eigenvalue
// decomposition", Japan J. Indust. Appl. Math. 35 (2018), here for complex
// Hermitian A): with S = X^H A X, G = X^H X, R = I - G,
//   lambda_i = S_ii / G_ii,
//   E_ij = (S_ij + lambda_j R_ij) / (lambda_j - lambda_i)   (i != j),
//   E_ii = R_ii / 2,
//   X <- X + X E,
// which converges quadratically once every |S_ij| is small against its gap.
// Every iteration is four n x n complex products on FP64 MFMA. The control
// words live on the device, so the iterations queue without host round trips:
//   state[0]: 0 running, 1 done, 2 failed, 3 not needed; state[1]: X buffer
//   holding the result.
// ---------------------------------------------------------------------------

// C = op(A) B (+ Cadd), n x n complex, row-major with leading dimension n
// (n a multiple of 32). op(A) = A^H if ah. 32 x 32 output tile per workgroup,
// one 16 x 16 quarter per wavefront.
__global__ __launch_bounds__(256) void zgemm_kernel(const cplx<double>* A, int ah, const cplx<double>* B,
                                                    cplx<double>* C, const cplx<double>* Cadd, int n,
                                                    const int* state, int herm, cplx<double>* part) {
  if (state && state[0] != 0) return;
  // herm: C is Hermitian (A^H B with B = A, or with B = M A for Hermitian M):
  // the upper tiles only, each written with its mirror
  if (herm && blockIdx.y > blockIdx.x) return;
  __shared__ cplx<double> As[32][17];
  __shared__ cplx<double> Bs[16][33];
  const int i0 = blockIdx.y * 32, j0 = blockIdx.x * 32;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wi = w >> 1, wj = w & 1;
  const int li = lane & 15, lk = lane >> 4;
  // split-K: slice blockIdx.z of gridDim.z equal K ranges (multiples of 16),
  // its partial product to part[z] (summed by zgemm_reduce_kernel)
  const int kn = n / (int)gridDim.z, kb = (int)blockIdx.z * kn, ke = kb + kn;
  d4 cr = {0, 0, 0, 0}, ci = {0, 0, 0, 0};
  cplx<double> pa[2], pb[2];
  auto fetch = [&](int k0) {
#pragma unroll
    for (int q = 0; q < 2; q++) {
      const int e = tid + 256 * q;
      const int r = e >> 4, kk = e & 15;  // A tile: 32 rows of op(A) x 16 k
      pa[q] = ah ? conj(A[(long long)(k0 + kk) * n + i0 + r]) : A[(long long)(i0 + r) * n + k0 + kk];
      const int kb2 = e >> 5, cidx = e & 31;  // B tile: 16 k x 32 columns
      pb[q] = B[(long long)(k0 + kb2) * n + j0 + cidx];
    }
  };
  fetch(kb);
  for (int k0 = kb; k0 < ke; k0 += 16) {
#pragma unroll
    for (int q = 0; q < 2; q++) {
      const int e = tid + 256 * q;
      As[e >> 4][e & 15] = pa[q];
      Bs[e >> 5][e & 31] = pb[q];
    }
    __syncthreads();
    if (k0 + 16 < ke) fetch(k0 + 16);
#pragma unroll
    for (int ks = 0; ks < 4; ks++) {
      const cplx<double> av = As[wi * 16 + li][ks * 4 + lk];
      const cplx<double> bv = Bs[ks * 4 + lk][wj * 16 + li];
      cr = __builtin_amdgcn_mfma_f64_16x16x4f64(av.x, bv.x, cr, 0, 0, 0);
      cr = __builtin_amdgcn_mfma_f64_16x16x4f64(-av.y, bv.y, cr, 0, 0, 0);
      ci = __builtin_amdgcn_mfma_f64_16x16x4f64(av.x, bv.y, ci, 0, 0, 0);
      ci = __builtin_amdgcn_mfma_f64_16x16x4f64(av.y, bv.x, ci, 0, 0, 0);
    }
    __syncthreads();
  }
  cplx<double>* out = gridDim.z > 1 ? part + (long long)blockIdx.z * n * n : C;
  for (int v = 0; v < 4; v++) {
    const long long o = (long long)(i0 + wi * 16 + lk + 4 * v) * n + j0 + wj * 16 + li;
    cplx<double> r = mk<double>(cr[v], ci[v]);
    if (gridDim.z > 1) {
      out[o] = r;
      continue;
    }
    if (Cadd) r = Cadd[o] + r;
    C[o] = r;
    if (herm && blockIdx.y != blockIdx.x)
      C[(long long)(j0 + wj * 16 + li) * n + i0 + wi * 16 + lk + 4 * v] = conj(r);
  }
}

// C = sum of the ksplit partial products (in slice order) [+ Cadd]; with herm
// only the upper 32 x 32 tiles were computed and each is written with its
// mirror.
__global__ __launch_bounds__(256) void zgemm_reduce_kernel(const cplx<double>* part, int ksplit, cplx<double>* C,
                                                           const cplx<double>* Cadd, int n, const int* state,
                                                           int herm) {
  if (state && state[0] != 0) return;
  const long long n2 = (long long)n * n;
  for (long long e = blockIdx.x * 256ll + threadIdx.x; e < n2; e += (long long)gridDim.x * 256) {
    const int i = (int)(e / n), j = (int)(e - (long long)i * n);
    if (herm && (i >> 5) > (j >> 5)) continue;
    cplx<double> r = part[e];
    for (int z = 1; z < ksplit; z++) r = r + part[(long long)z * n2 + e];
    if (Cadd) r = Cadd[e] + r;
    C[e] = r;
    if (herm && (i >> 5) != (j >> 5)) C[(long long)j * n + i] = conj(r);
  }
}

// E from S and G; indices >= N (padding) stay fixed. lam[i] = S_ii / G_ii.
// The largest |E_ij| (i != j) goes to *slot (atomic max of the bit pattern).
constexpr int kOARows = 8;  // rows of E per workgroup of oa_update_kernel

__global__ __launch_bounds__(256) void oa_update_kernel(const cplx<double>* S, const cplx<double>* G, cplx<double>* E,
                                                        double* lam, int np, int N, unsigned long long* slot,
                                                        const int* state) {
  if (state[0] != 0) return;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  double m = 0;
  for (int r = 0; r < kOARows; r++) {
    const int i = blockIdx.y * kOARows + r;
    if (i >= np || j >= np) continue;
    const long long o = (long long)i * np + j;
    cplx<double> e = mk<double>(0.0, 0.0);
    if (i < N && j < N) {
      const double li = S[(long long)i * np + i].x / G[(long long)i * np + i].x;
      if (i == j) {
        e = mk<double>(0.5 * (1.0 - G[o].x), 0.0);
        lam[i] = li;
      } else {
        const double lj = S[(long long)j * np + j].x / G[(long long)j * np + j].x;
        // S and G are Hermitian, but inside the diagonal 32 x 32 tiles of the
        // products both triangles are computed and differ by rounding; their
        // Hermitian part keeps that difference out of E's Hermitian part
        // (which must be exactly R/2), where divided by a small gap it would
        // stall the iteration at eps |A| / gap. Elsewhere the lower triangle
        // is the exact mirror, and the Hermitian part is the element itself.
        cplx<double> sij = S[o], gij = G[o];
        if ((i >> 5) == (j >> 5)) {
          const long long ot = (long long)j * np + i;
          sij = (sij + conj(S[ot])) * 0.5;
          gij = (gij + conj(G[ot])) * 0.5;
        }
        const cplx<double> num = sij - gij * lj;  // S_ij + lambda_j R_ij with R_ij = -G_ij
        const double gap = lj - li;
        e = mk<double>(num.x / gap, num.y / gap);
        const double a = cabs_fast(e);
        const double mm = (a <= 1e300) ? a : 1e300;  // a zero gap or a NaN fails the iteration
        m = mm > m ? mm : m;
      }
    }
    E[o] = e;
  }
  // workgroup maximum, then one atomic per workgroup
  for (int off = 32; off >= 1; off >>= 1) {
    const double t = __shfl_xor(m, off);
    m = t > m ? t : m;
  }
  __shared__ double wm[4];
  if ((threadIdx.x & 63) == 0) wm[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double b = wm[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); w++) b = wm[w] > b ? wm[w] : b;
    if (b > 0) atomicMax(slot, dbits(b));
  }
}

// After iteration k: done once the correction fell below 1e-8 (the update
// just applied leaves errors of its square), failed if it is not small.
__global__ void oa_decide_kernel(const unsigned long long* slot, int k, int* state) {
  if (state[0] != 0) return;
  const double m = bitsd(*slot);
  if (!(m < 0.5)) { state[0] = 2; return; }
  if (m < 1e-8) { state[0] = 1; state[1] = (k & 1) ? 2 : 1; }
}

// Starts the refinement if the Jacobi stage handed off (info[1] == 2).
__global__ void oa_init_kernel(const double* info, int* state) {
  state[0] = info[1] == 2.0 ? 0 : 3;
  state[1] = 0;
}

// Result into the places the eigenvalue sort reads: X -> V (from the buffer
// the last iterate went to), lambda -> diagonal of A. Without convergence V
// and A keep the Jacobi stage's eigenvector estimate and rotated matrix.
__global__ void oa_finalize_kernel(const int* state, const cplx<double>* X1, const cplx<double>* X2, cplx<double>* V,
                                   cplx<double>* A, const double* lam, int np, int N) {
  if (state[0] != 1) return;
  const long long n2 = (long long)np * np;
  const cplx<double>* X = state[1] == 1 ? X1 : X2;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < n2; e += (long long)gridDim.x * blockDim.x) {
    V[e] = X[e];
    const int i = (int)(e / np), j = (int)(e - (long long)i * np);
    if (i == j && i < N) A[e] = mk<double>(lam[i], 0.0);
  }
}

hipError_t launch_oa(const cplx<double>* A0, cplx<double>* X0, cplx<double>* X1, cplx<double>* work, double* lam,
                     double* slots, int* state, const double* info, int np, int N, int max_iter, cplx<double>* Aout,
                     hipStream_t st, const cplx<double>* x0_identity, cplx<double>* part) {
  if (np % 32) return hipErrorInvalidValue;
  const long long n2 = (long long)np * np;
  cplx<double>* Y = work;
  cplx<double>* S = work + n2;
  cplx<double>* G = work + 2 * n2;
  cplx<double>* E = work + 3 * n2;
  cplx<double>* X2 = work + 4 * n2;
  hipLaunchKernelGGL(oa_init_kernel, dim3(1), dim3(1), 0, st, info, state);
  hipLaunchKernelGGL(fill_u64_kernel, dim3(1), dim3(64), 0, st, reinterpret_cast<unsigned long long*>(slots),
                     (long long)max_iter);
  // split-K for small matrices (few 32 x 32 tiles): 4 slices when the tiles
  // alone cannot fill the device and part is given
  int ks = 1;
  if (part && (np / 32) * (np / 32) < 1024) ks = np % 64 == 0 ? (np % 128 == 0 ? 4 : 2) : 1;
  const dim3 gg(np / 32, np / 32, ks);
  const dim3 ug((np + 255) / 256, (np + kOARows - 1) / kOARows);
  const int rb = (int)std::min<long long>(((long long)np * np + 255) / 256, 2048);
  auto gemm = [&](const cplx<double>* a, int ah, const cplx<double>* b, cplx<double>* c, const cplx<double>* cadd,
                  int herm) {
    hipLaunchKernelGGL(zgemm_kernel, gg, dim3(256), 0, st, a, ah, b, c, cadd, np, state, herm, part);
    if (ks > 1)
      hipLaunchKernelGGL(zgemm_reduce_kernel, dim3(rb), dim3(256), 0, st, (const cplx<double>*)part, ks, c, cadd, np,
                         state, herm);
  };
  for (int k = 0; k < max_iter; k++) {
    // X0 is only read (iterates alternate between X1 and X2), so a failed
    // refinement leaves the starting estimate in place
    cplx<double>* X = k == 0 ? X0 : ((k & 1) ? X1 : X2);
    cplx<double>* Xn = (k & 1) ? X2 : X1;
    unsigned long long* slot = reinterpret_cast<unsigned long long*>(slots) + k;
    if (k == 0 && x0_identity) {
      // X = I: S = A0 and G = I, no products needed
      hipLaunchKernelGGL(oa_update_kernel, ug, dim3(256), 0, st, A0, x0_identity, E, lam, np, N, slot, state);
    } else {
      gemm(A0, 0, X, Y, nullptr, 0);
      gemm(X, 1, Y, S, nullptr, 1);
      gemm(X, 1, X, G, nullptr, 1);
      hipLaunchKernelGGL(oa_update_kernel, ug, dim3(256), 0, st, S, G, E, lam, np, N, slot, state);
    }
    gemm(X, 0, E, Xn, X, 0);
    hipLaunchKernelGGL(oa_decide_kernel, dim3(1), dim3(1), 0, st, slot, k, state);
  }
  hipLaunchKernelGGL(oa_finalize_kernel, dim3(256), dim3(256), 0, st, state, X1, X2, X0, Aout, lam, np, N);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Cluster start for the refinement (cluster_eig.hpp): strong pairs of the
// upper triangle, then one workgroup per cluster.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void cl_pairs_kernel(const cplx<double>* A, int np, int N, double tau, int cap,
                                                       unsigned long long* hdr, int* pairs) {
  const int i = blockIdx.x;  // grid = N rows: the whole block takes the same path
  const double mx = cl_row_pairs(A, np, N, i, (int)threadIdx.x, (int)blockDim.x, tau, [&](int a, int b) {
    const unsigned long long slot = atomicAdd(&hdr[0], 1ull);
    if (slot < (unsigned long long)cap) {
      pairs[2 * slot] = a;
      pairs[2 * slot + 1] = b;
    }
  });
  __shared__ double red[256];
  red[threadIdx.x] = mx;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] = red[threadIdx.x + w] > red[threadIdx.x] ? red[threadIdx.x + w] : red[threadIdx.x];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicMax(&hdr[1], dbits(red[0]));
}

hipError_t launch_cluster_pairs(const cplx<double>* A, int np, int N, double tau, int cap, unsigned long long* hdr,
                                int* pairs, hipStream_t st) {
  hipLaunchKernelGGL(fill_u64_kernel, dim3(1), dim3(64), 0, st, hdr, 2ll);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || N < 2) return e;
  hipLaunchKernelGGL(cl_pairs_kernel, dim3(N), dim3(256), 0, st, A, np, N, tau, cap, hdr, pairs);
  return hipGetLastError();
}

__global__ __launch_bounds__(256) void cl_solve_kernel(const cplx<double>* A, int np, const int* members,
                                                       const int* offsets, cplx<double>* X, double tol,
                                                       int max_sweeps) {
  extern __shared__ __align__(16) char cll[];
  const int c = blockIdx.x;
  const int o0 = offsets[c], k = offsets[c + 1] - o0;
  if (k < 1 || k > kMaxCluster) return;  // the host builds 2..kMaxCluster; the whole block agrees
  GpuExec ex;
  cl_solve(ex, A, np, members + o0, k, X, tol, max_sweeps, cll, np);
}

hipError_t launch_cluster_solve(const cplx<double>* A, int np, const int* members, const int* offsets, int nclusters,
                                cplx<double>* X, hipStream_t st) {
  if (nclusters <= 0) return hipSuccess;
  const size_t lds = cl_lds_bytes(256);
  static const hipError_t attr =
      hipFuncSetAttribute((const void*)cl_solve_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (attr != hipSuccess) return attr;
  hipLaunchKernelGGL(cl_solve_kernel, dim3(nclusters), dim3(256), lds, st, A, np, members, offsets, X, 1e-15, 30);
  return hipGetLastError();
}

// Ascending order of the eigenvalues (ties by index, like a stable sort):
// each thread ranks one eigenvalue against the diagonal, which every
// workgroup stages through LDS in chunks (read as broadcasts).
__global__ __launch_bounds__(256) void eig_sort_kernel(const cplx<double>* A, int np, int N, double* lam_sorted,
                                                       int* perm) {
  constexpr int kChunk = 2048;
  __shared__ double d[kChunk];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const double li = i < N ? A[(long long)i * np + i].x : 0.0;
  int rank = 0;
  for (int j0 = 0; j0 < N; j0 += kChunk) {
    const int jn = N - j0 < kChunk ? N - j0 : kChunk;
    __syncthreads();
    for (int j = threadIdx.x; j < jn; j += blockDim.x) d[j] = A[(long long)(j0 + j) * np + (j0 + j)].x;
    __syncthreads();
    if (i < N)
      for (int j = 0; j < jn; j++) {
        const double lj = d[j];
        rank += (lj < li) || (lj == li && j0 + j < i);
      }
  }
  if (i < N && rank < N) {
    lam_sorted[rank] = li;
    perm[rank] = i;
  }
}

// C[c][r] = v_{perm r}[c] / sqrt(lambda_r)  (stateset.cpp:175 then the row-major read at :205-210)
__global__ void eig_coeff_kernel(const cplx<double>* V, int np, int N, const double* lam_sorted, const int* perm,
                                 cplx<double>* C, int nc, int scale) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  const int c = blockIdx.y;
  if (r >= N || c >= N) return;
  const double sc = scale ? 1.0 / std::sqrt(lam_sorted[r]) : 1.0;
  int p = perm[r];
  if (p < 0 || p >= N) p = 0;  // ranks broken by a NaN: the host rejects the result anyway
  C[(long long)c * nc + r] = V[(long long)c * np + p] * sc;
}

__global__ void fill_kernel(double* p, long long n, double v) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) p[i] = v;
}

hipError_t launch_eig_finish(const cplx<double>* A, const cplx<double>* V, int np, int N, double* lam_sorted, int* perm,
                             cplx<double>* C, int nc, int scale, hipStream_t st) {
  hipLaunchKernelGGL(fill_kernel, dim3((N + 255) / 256), dim3(256), 0, st, lam_sorted, (long long)N, (double)NAN);
  hipError_t e = hipMemsetAsync(perm, 0xff, (size_t)N * sizeof(int), st);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(eig_sort_kernel, dim3((N + 255) / 256), dim3(256), 0, st, A, np, N, lam_sorted, perm);
  hipLaunchKernelGGL(eig_coeff_kernel, dim3((N + 127) / 128, N), dim3(128), 0, st, V, np, N, lam_sorted, perm, C, nc, scale);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Rotation Psi_out[r][t] = sum_c C[c][r] Psi_in[c][t] on FP64 MFMA.
// Tile: 64 output states x 64 grid points per workgroup, K = input states,
// staged 16 at a time through LDS with the next step's operands prefetched
// into registers. Workgroups are dealt to the 8 XCDs round-robin in launch
// order; the block order below gives each XCD whole grid-point slabs and
// runs all output-state tiles of a slab back to back there, so the slab
// (N x 64 values) is read from HBM once and reused from that XCD's L2.
// ---------------------------------------------------------------------------
template <class R>
__global__ __launch_bounds__(256) void rotate_kernel(const cplx<R>* in, cplx<R>* out, const cplx<double>* C, int nc,
                                                     long long M, int N) {
  __shared__ cplx<double> Cs[16][65];
  __shared__ cplx<double> Ps[16][65];
  int rt;
  long long tt;
  if (!xcd_tile(blockIdx.x, nc / 64, (M + 63) / 64, rt, tt)) return;
  const long long t0 = tt * 64;
  const int r0 = rt * 64;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, wr = wave >> 1, wt = wave & 1;
  const int li = lane & 15, lk = lane >> 4;
  d4 accr[2][2], acci[2][2];
  for (int a = 0; a < 2; a++)
    for (int b = 0; b < 2; b++)
      for (int v = 0; v < 4; v++) { accr[a][b][v] = 0; acci[a][b][v] = 0; }
  // thread's staging slots: element e = tid + 256 q of the 16 x 64 tiles
  cplx<double> pc[4], pp[4];
  auto fetch = [&](int c0) {
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int e = tid + 256 * q;
      const int kc = e >> 6, x = e & 63;
      const int c = c0 + kc;
      // C is zero-padded to nc x nc; rows c >= N contribute nothing
      pc[q] = (c < N) ? C[(long long)c * nc + r0 + x] : mk<double>(0.0, 0.0);
      cplx<double> v = mk<double>(0.0, 0.0);
      if (c < N && t0 + x < M) { const cplx<R> p = in[(long long)c * M + t0 + x]; v = mk<double>(double(p.x), double(p.y)); }
      pp[q] = v;
    }
  };
  fetch(0);
  for (int c0 = 0; c0 < N; c0 += 16) {
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int e = tid + 256 * q;
      Cs[e >> 6][e & 63] = pc[q];
      Ps[e >> 6][e & 63] = pp[q];
    }
    __syncthreads();
    if (c0 + 16 < N) fetch(c0 + 16);
#pragma unroll
    for (int kk = 0; kk < 4; kk++) {
      cplx<double> a[2], b[2];
      for (int mi = 0; mi < 2; mi++) a[mi] = Cs[kk * 4 + lk][wr * 32 + mi * 16 + li];
      for (int nj = 0; nj < 2; nj++) b[nj] = Ps[kk * 4 + lk][wt * 32 + nj * 16 + li];
      for (int mi = 0; mi < 2; mi++)
        for (int nj = 0; nj < 2; nj++) {
          accr[mi][nj] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[mi].x, b[nj].x, accr[mi][nj], 0, 0, 0);
          accr[mi][nj] = __builtin_amdgcn_mfma_f64_16x16x4f64(-a[mi].y, b[nj].y, accr[mi][nj], 0, 0, 0);
          acci[mi][nj] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[mi].x, b[nj].y, acci[mi][nj], 0, 0, 0);
          acci[mi][nj] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[mi].y, b[nj].x, acci[mi][nj], 0, 0, 0);
        }
    }
    __syncthreads();
  }
  for (int mi = 0; mi < 2; mi++)
    for (int nj = 0; nj < 2; nj++)
      for (int v = 0; v < 4; v++) {
        const int r = r0 + wr * 32 + mi * 16 + lk + 4 * v;
        const long long t = t0 + wt * 32 + nj * 16 + li;
        if (r < N && t < M) out[(long long)r * M + t] = mk<R>(R(accr[mi][nj][v]), R(acci[mi][nj][v]));
      }
}

template <class R>
hipError_t launch_rotate(const cplx<R>* in, cplx<R>* out, const cplx<double>* C, int nc, long long M, int N,
                         hipStream_t st) {
  const long long blocks = xcd_grid(nc / 64, (M + 63) / 64);
  if (nc % 64 || blocks > 0x7fffffffLL) return hipErrorInvalidValue;
  hipLaunchKernelGGL(rotate_kernel<R>, dim3((unsigned)blocks), dim3(256), 0, st, in, out, C, nc, M, N);
  return hipGetLastError();
}
template hipError_t launch_rotate<double>(const cplx<double>*, cplx<double>*, const cplx<double>*, int, long long, int,
                                          hipStream_t);
template hipError_t launch_rotate<float>(const cplx<float>*, cplx<float>*, const cplx<double>*, int, long long, int,
                                         hipStream_t);

}  // namespace itp
