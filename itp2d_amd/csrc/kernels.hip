// gfx950 kernels of the ITP loop.
//
//  sweep_kernel      one workgroup = L lines of one field; the line program of
//                    line_ops.hpp (LDS-resident FFT / DST / DCT passes fused
//                    with the multiplier tables and the epilogue)
//  energy_reduce     E and sigma per state from the sweep partial sums
//  gram_kernel       overlap S = Psi^H Psi on FP64 MFMA (v_mfma_f64_16x16x4f64),
//                    64x64 tiles of state pairs, split over the grid points
//  gram_reduce       sums the split-K partials, scales by dx^2, fills the
//                    Hermitian matrix for the eigensolver
//  jacobi_kernel     cooperative parallel two-sided Jacobi on the overlap matrix
//  eig_sort / eig_coeffs  ascending eigenvalues, C = U Lambda^(-1/2)
//  rotate_kernel     Psi' = Psi C on FP64 MFMA
#include <hip/hip_runtime.h>
#include <hip/hip_cooperative_groups.h>

#include "block_jacobi.hpp"
#include "dense_ops.hpp"
#include "gpu.hpp"
#include "line_ops.hpp"
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

// Per-thread state of the register engine lives in the exec object, so the
// same pass code runs per thread on the GPU and per emulated thread on the CPU.
// The thread index is passed through opaque() in every pass: addresses and
// table loads derived from it then stay inside their pass instead of being
// hoisted out of the op loop, which would pin registers for the whole kernel.
template <class S>
struct GpuRegExec {
  S st;
  template <class F>
  __device__ void run(F&& f) {
    f(opaque((int)threadIdx.x), st);
    __syncthreads();
  }
  template <class F>
  __device__ void local(F&& f) {
    f(opaque((int)threadIdx.x), st);
  }
};

template <class R, int N, int ORIENT>
__global__ __launch_bounds__(256) void reg_sweep_kernel(SweepArgs<R> a) {
  extern __shared__ __align__(16) char lds_raw[];
  GpuRegExec<RegState<R, RegShape<N>::Q>> ex;
  reg_sweep_program<R, N, ORIENT>(ex, a, reinterpret_cast<cplx<R>*>(lds_raw), (int)blockIdx.y, (int)blockIdx.x);
}

template <class R>
hipError_t launch_reg_sweep(const SweepArgs<R>& a, int nblk, int nfields, hipStream_t st) {
  const int n = a.plan.n;
  const size_t lds = reg_lds_bytes<R>(n);
  if (lds > 160 * 1024 || a.L != reg_lines_per_wg(n)) return hipErrorInvalidValue;
  const int lines = a.orient == 0 ? a.sy : a.sx;
  if ((a.orient == 0 ? a.sx : a.sy) != n || nblk * a.L != lines) return hipErrorInvalidValue;
  hipError_t err = hipSuccess;
  const bool ok = reg_dispatch(n, a.orient, [&](auto nc, auto oc) {
    constexpr int N = decltype(nc)::value, O = decltype(oc)::value;
    static bool attr_set = false;
    if (!attr_set) {
      (void)hipFuncSetAttribute((const void*)reg_sweep_kernel<R, N, O>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
      attr_set = true;
    }
    hipLaunchKernelGGL((reg_sweep_kernel<R, N, O>), dim3(nblk, nfields), dim3(kRegThreads), lds, st, a);
    err = hipGetLastError();
  });
  return ok ? err : hipErrorInvalidValue;
}

template <class R>
hipError_t launch_sweep(const SweepArgs<R>& a, int nblk, int nfields, hipStream_t st) {
  if (nblk <= 0 || nfields <= 0) return hipSuccess;
  if (nfields > 65535) return hipErrorInvalidValue;
  if (a.reg) return launch_reg_sweep<R>(a, nblk, nfields, st);
  const size_t lds = sweep_lds_bytes<R>(a.L, a.plan.n, 256);
  if (lds > 160 * 1024) return hipErrorInvalidValue;
  static bool attr_set = false;
  if (!attr_set) {
    // allow the full 160 KiB of LDS per workgroup
    (void)hipFuncSetAttribute((const void*)sweep_kernel<R>, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    attr_set = true;
  }
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
// Overlap matrix on FP64 MFMA.
//
// v_mfma_f64_16x16x4f64: D(16x16) += A(16x4) B(4x16). Lane l holds
// A[l%16][l/16], B[l/16][l%16] and D[l/16 + 4v][l%16], v = 0..3 (measured on
// gfx950 with indicator inputs; the f32 16x16 shapes use 4(l/16)+v instead).
// S_ij = sum_t conj(psi_i(t)) psi_j(t): with a = psi_i, b = psi_j,
// Re += a.x b.x + a.y b.y, Im += a.x b.y - a.y b.x.
// ---------------------------------------------------------------------------
template <class R>
__global__ __launch_bounds__(256) void gram_kernel(const cplx<R>* psi, long long M, int N, int nb, long long kchunk,
                                                   cplx<double>* part, int nc) {
  int bi = 0, rem = blockIdx.x;
  while (rem >= nb - bi) { rem -= nb - bi; bi++; }
  const int bj = bi + rem;
  __shared__ cplx<double> As[64][17];
  __shared__ cplx<double> Bs[64][17];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, wi = wave >> 1, wj = wave & 1;
  const int li = lane & 15, lk = lane >> 4;
  d4 accr[2][2], acci[2][2];
  for (int a = 0; a < 2; a++)
    for (int b = 0; b < 2; b++)
      for (int v = 0; v < 4; v++) { accr[a][b][v] = 0; acci[a][b][v] = 0; }
  const long long k0 = (long long)blockIdx.y * kchunk;
  const long long k1 = (k0 + kchunk < M) ? k0 + kchunk : M;
  for (long long t0 = k0; t0 < k1; t0 += 16) {
    for (int r = 0; r < 4; r++) {
      const int e = tid + 256 * r;
      const int s = e >> 4, k = e & 15;
      const long long t = t0 + k;
      const int ia = bi * 64 + s, ib = bj * 64 + s;
      cplx<double> va = mk<double>(0.0, 0.0), vb = va;
      if (t < k1) {
        if (ia < N) { const cplx<R> x = psi[(long long)ia * M + t]; va = mk<double>(double(x.x), double(x.y)); }
        if (ib < N) { const cplx<R> x = psi[(long long)ib * M + t]; vb = mk<double>(double(x.x), double(x.y)); }
      }
      As[s][k] = va;
      Bs[s][k] = vb;
    }
    __syncthreads();
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
  cplx<double>* P = part + (long long)blockIdx.y * nc * nc;
  for (int mi = 0; mi < 2; mi++)
    for (int nj = 0; nj < 2; nj++)
      for (int v = 0; v < 4; v++) {
        const int i = bi * 64 + wi * 32 + mi * 16 + lk + 4 * v;
        const int j = bj * 64 + wj * 32 + nj * 16 + li;
        P[(long long)i * nc + j] = mk<double>(accr[mi][nj][v], acci[mi][nj][v]);
      }
}

__global__ void gram_reduce_kernel(const cplx<double>* part, int nchunk, int nc, int N, int np, double dx,
                                   cplx<double>* S, int* bad) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y;
  if (j >= N || i > j) return;
  cplx<double> s = mk<double>(0.0, 0.0);
  for (int z = 0; z < nchunk; z++) s = s + part[((long long)z * nc + i) * nc + j];
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
hipError_t launch_gram(const cplx<R>* psi, long long M, int N, int nc, cplx<double>* part, int nchunk,
                       long long kchunk, double dx, cplx<double>* S, int np, int* bad, hipStream_t st) {
  const int nb = nc / 64;
  const int npairs = nb * (nb + 1) / 2;
  hipLaunchKernelGGL(gram_kernel<R>, dim3(npairs, nchunk), dim3(256), 0, st, psi, M, N, nb, kchunk, part, nc);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(gram_reduce_kernel, dim3((N + 127) / 128, N), dim3(128), 0, st, part, nchunk, nc, N, np, dx, S,
                     bad);
  return hipGetLastError();
}
template hipError_t launch_gram<double>(const cplx<double>*, long long, int, int, cplx<double>*, int, long long,
                                        double, cplx<double>*, int, int*, hipStream_t);
template hipError_t launch_gram<float>(const cplx<float>*, long long, int, int, cplx<double>*, int, long long,
                                       double, cplx<double>*, int, int*, hipStream_t);

// ---------------------------------------------------------------------------
// Jacobi eigensolver, one cooperative launch for all sweeps. Every workgroup
// computes all rotation parameters of a round itself (they only read the
// 2x2 diagonal blocks), so the only grid-wide barrier is the one per round.
// The convergence measure (largest relative off-diagonal element met in a
// sweep) is therefore identical in every workgroup, and all stop together.
// ---------------------------------------------------------------------------
struct JacobiArgs {
  cplx<double>* A[2];
  cplx<double>* V[2];
  int np;
  int max_sweeps;
  double tol;
  double* info;  // [0] sweeps, [1] converged, [2] final buffer, [3] last sweep's largest relative
                 // off-diagonal, [4] stop flag written by block 0
};

__global__ __launch_bounds__(256) void jacobi_kernel(JacobiArgs a) {
  cg::grid_group grid = cg::this_grid();
  extern __shared__ __align__(16) char jl[];
  const int np = a.np, m = np / 2;
  cplx<double>* sj = reinterpret_cast<cplx<double>*>(jl);
  double* cs = reinterpret_cast<double*>(sj + m);
  double* red = cs + m;  // [3 * blockDim + 3]
  const int nt = (int)blockDim.x;
  const long long gtid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long gthreads = (long long)gridDim.x * blockDim.x;
  const long long nblock_items = (long long)m * m;
  const long long nitems = nblock_items + (long long)np * m;
  int cur = 0, sweeps = 0;
  bool conv = false;
  double wg_off = 0;
  for (int sw = 0; sw < a.max_sweeps; sw++) {
    double thr_off = 0, thr_b = 0, thr_d = 0;
    for (int r = 0; r < np - 1; r++) {
      const cplx<double>* A = a.A[cur];
      cplx<double>* An = a.A[cur ^ 1];
      const cplx<double>* V = a.V[cur];
      cplx<double>* Vn = a.V[cur ^ 1];
      for (int k = threadIdx.x; k < m; k += nt) {
        const int p = 2 * k, q = p + 1;
        double c, off;
        cplx<double> s;
        const double app = A[(long long)p * np + p].x, aqq = A[(long long)q * np + q].x;
        const cplx<double> b = A[(long long)p * np + q];
        jacobi_rotation(app, aqq, b, c, s, off);
        cs[k] = c;
        sj[k] = s;
        thr_off = off > thr_off ? off : thr_off;
        const double ab = std::hypot(b.x, b.y);
        thr_b = ab > thr_b ? ab : thr_b;
        const double dd = std::fabs(app) > std::fabs(aqq) ? std::fabs(app) : std::fabs(aqq);
        thr_d = dd > thr_d ? dd : thr_d;
      }
      __syncthreads();
      for (long long it = gtid; it < nitems; it += gthreads) {
        if (it < nblock_items) {
          const int kI = (int)(it / m), kJ = (int)(it - (long long)kI * m);
          if (kJ >= kI) jacobi_block(A, An, np, m, kI, kJ, cs, sj);
        } else {
          const long long i2 = it - nblock_items;
          const int c = (int)(i2 / m), k = (int)(i2 - (long long)c * m);
          jacobi_vrow(V, Vn, np, m, c, k, cs, sj);
        }
      }
      grid.sync();
      cur ^= 1;
    }
    sweeps++;
    // block 0 decides whether the sweep met the tolerance and broadcasts it
    if (blockIdx.x == 0) {
      red[threadIdx.x] = thr_off;
      red[nt + threadIdx.x] = thr_b;
      red[2 * nt + threadIdx.x] = thr_d;
      __syncthreads();
      if (threadIdx.x == 0) {
        double mo = 0, mb = 0, md = 0;
        for (int t = 0; t < nt; t++) {
          mo = red[t] > mo ? red[t] : mo;
          mb = red[nt + t] > mb ? red[nt + t] : mb;
          md = red[2 * nt + t] > md ? red[2 * nt + t] : md;
        }
        // relative off-diagonal below tol, or all off-diagonals at the
        // rounding level of the largest diagonal element
        const bool done = (mo < a.tol) || (mb <= 1e-15 * md);
        a.info[4] = done ? 1.0 : 0.0;
        a.info[3] = mo;
      }
    }
    grid.sync();
    conv = __hip_atomic_load(&a.info[4], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0.0;
    if (conv) break;
    grid.sync();  // everyone has read the flag before block 0 may overwrite it
  }
  // leave the result in buffer 0 (the diagonal of A and all of V)
  if (cur == 1) {
    const long long tot = (long long)np * np;
    for (long long i = gtid; i < tot; i += gthreads) a.V[0][i] = a.V[1][i];
    for (long long i = gtid; i < np; i += gthreads) a.A[0][i * np + i] = a.A[1][i * np + i];
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    a.info[0] = sweeps;
    a.info[1] = conv ? 1.0 : 0.0;
    a.info[2] = cur;
  }
  (void)wg_off;
}

size_t jacobi_lds_bytes(int np) { return (size_t)(np / 2) * (sizeof(cplx<double>) + sizeof(double)) + (3 * 256 + 3) * sizeof(double); }

hipError_t launch_jacobi(cplx<double>* A0, cplx<double>* A1, cplx<double>* V0, cplx<double>* V1, int np,
                         int max_sweeps, double tol, double* info, int num_cus, hipStream_t st) {
  JacobiArgs a;
  a.A[0] = A0; a.A[1] = A1; a.V[0] = V0; a.V[1] = V1;
  a.np = np; a.max_sweeps = max_sweeps; a.tol = tol; a.info = info;
  const size_t lds = jacobi_lds_bytes(np);
  if (lds > 160 * 1024) return hipErrorInvalidValue;
  (void)hipFuncSetAttribute((const void*)jacobi_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
  int per_cu = 0;
  hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, jacobi_kernel, 256, lds);
  if (e != hipSuccess) return e;
  if (per_cu < 1) return hipErrorInvalidValue;
  const long long m = np / 2;
  const long long items = m * m + (long long)np * m;
  long long want = (items + 255) / 256;
  long long cap = (long long)per_cu * num_cus;
  if (cap > (long long)num_cus) cap = (long long)num_cus;  // barrier cost grows with workgroups
  int grid = (int)std::max<long long>(1, std::min(want, cap));
  void* args[] = {&a};
  e = hipLaunchCooperativeKernel((const void*)jacobi_kernel, dim3(grid), dim3(256), args, (unsigned)lds, st);
  if (e != hipSuccess) return e;
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Block Jacobi (block_jacobi.hpp), one cooperative launch for all sweeps:
// per round, phase 1 (one workgroup per block pair, sub-problem in LDS) and
// phase 2 (super-tile updates of A and V), each followed by a grid barrier.
// The convergence maxima of a sweep are combined with atomicMax into that
// sweep's own slot, which every workgroup reads after the sweep's last
// barrier, so all workgroups take the same decision.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void block_jacobi_kernel(BlockJacobiArgs a) {
  cg::grid_group grid = cg::this_grid();
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
  bool conv = false;
  for (int chk = 0;; chk++) {
    // whole-matrix convergence check
    unsigned long long* slot = reinterpret_cast<unsigned long long*>(a.slots + chk);
    bj_check_rows(ex, a, e0, e1, s);
    if (threadIdx.x == 0) atomicMax(slot, dbits(s.red[nt]));
    grid.sync();
    const double worst = bitsd(__hip_atomic_load(slot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    if (worst < 1.0) { conv = true; break; }
    if (chk == a.max_sweeps) break;
    // one sweep of nb-1 rounds
    bj_reset_positions(ex, s, nb);
    for (int r = 0; r < nb - 1; r++) {
      for (int k = blockIdx.x; k < P; k += gridDim.x) bj_inner(ex, a, k, s.pos2blk[2 * k], s.pos2blk[2 * k + 1], s);
      grid.sync();
      for (int t = blockIdx.x; t < items; t += gridDim.x) bj_apply(ex, a, t, s);
      grid.sync();
      bj_next_positions(ex, s, nb);
    }
    sweeps++;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    a.info[0] = sweeps;
    a.info[1] = conv ? 1.0 : 0.0;
    a.info[2] = 0;
  }
}

__global__ void fill_u64_kernel(unsigned long long* p, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) p[i] = 0ull;
}

hipError_t launch_block_jacobi(cplx<double>* A, cplx<double>* V, cplx<double>* Q, cplx<double>* D, int* ident,
                               double* slots, int np, int max_sweeps, int max_inner, double tol, double* info,
                               int grid_wgs, hipStream_t st) {
  if (np % JS || np / JB > kMaxBlocks) return hipErrorInvalidValue;
  BlockJacobiArgs a;
  a.A = A; a.V = V; a.Q = Q; a.D = D; a.ident = ident; a.slots = slots; a.info = info;
  a.np = np; a.max_sweeps = max_sweeps; a.max_inner = max_inner; a.tol = tol;
  hipLaunchKernelGGL(fill_u64_kernel, dim3(1), dim3(256), 0, st, reinterpret_cast<unsigned long long*>(slots),
                     (long long)(max_sweeps + 1));
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
  int grid = std::min(grid_wgs, per_cu * ncu);
  grid = std::max(1, std::min(grid, bj_items(np)));
  void* args[] = {&a};
  e = hipLaunchCooperativeKernel((const void*)block_jacobi_kernel, dim3(grid), dim3(256), args, (unsigned)lds, st);
  if (e != hipSuccess) return e;
  return hipGetLastError();
}

// Ascending order of the eigenvalues (ties by index, like a stable sort).
__global__ void eig_sort_kernel(const cplx<double>* A, int np, int N, double* lam_sorted, int* perm) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const double li = A[(long long)i * np + i].x;
  int rank = 0;
  for (int j = 0; j < N; j++) {
    const double lj = A[(long long)j * np + j].x;
    rank += (lj < li) || (lj == li && j < i);
  }
  if (rank < N) { lam_sorted[rank] = li; perm[rank] = i; }
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
  hipLaunchKernelGGL(eig_sort_kernel, dim3((N + 127) / 128), dim3(128), 0, st, A, np, N, lam_sorted, perm);
  hipLaunchKernelGGL(eig_coeff_kernel, dim3((N + 127) / 128, N), dim3(128), 0, st, V, np, N, lam_sorted, perm, C, nc, scale);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Rotation Psi_out[r][t] = sum_c C[c][r] Psi_in[c][t] on FP64 MFMA.
// Tile: 64 output states x 64 grid points per workgroup, K = input states.
// ---------------------------------------------------------------------------
template <class R>
__global__ __launch_bounds__(256) void rotate_kernel(const cplx<R>* in, cplx<R>* out, const cplx<double>* C, int nc,
                                                     long long M, int N) {
  __shared__ cplx<double> Cs[16][65];
  __shared__ cplx<double> Ps[16][65];
  const long long t0 = (long long)blockIdx.x * 64;
  const int r0 = blockIdx.y * 64;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, wr = wave >> 1, wt = wave & 1;
  const int li = lane & 15, lk = lane >> 4;
  d4 accr[2][2], acci[2][2];
  for (int a = 0; a < 2; a++)
    for (int b = 0; b < 2; b++)
      for (int v = 0; v < 4; v++) { accr[a][b][v] = 0; acci[a][b][v] = 0; }
  for (int c0 = 0; c0 < N; c0 += 16) {
    for (int q = 0; q < 4; q++) {
      const int e = tid + 256 * q;
      const int kc = e >> 6, x = e & 63;
      const int c = c0 + kc;
      // C is zero-padded to nc x nc; rows c >= N contribute nothing
      Cs[kc][x] = (c < N) ? C[(long long)c * nc + r0 + x] : mk<double>(0.0, 0.0);
      cplx<double> v = mk<double>(0.0, 0.0);
      if (c < N && t0 + x < M) { const cplx<R> p = in[(long long)c * M + t0 + x]; v = mk<double>(double(p.x), double(p.y)); }
      Ps[kc][x] = v;
    }
    __syncthreads();
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
  const long long tiles = (M + 63) / 64;
  if (tiles > 0x7fffffffLL) return hipErrorInvalidValue;
  hipLaunchKernelGGL(rotate_kernel<R>, dim3((unsigned)tiles, nc / 64), dim3(256), 0, st, in, out, C, nc, M, N);
  return hipGetLastError();
}
template hipError_t launch_rotate<double>(const cplx<double>*, cplx<double>*, const cplx<double>*, int, long long, int,
                                          hipStream_t);
template hipError_t launch_rotate<float>(const cplx<float>*, cplx<float>*, const cplx<double>*, int, long long, int,
                                         hipStream_t);

}  // namespace itp
