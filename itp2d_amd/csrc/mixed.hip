// gfx950 kernels of the mixed-precision path (complex FP32 fields).
//
//  gram_bf16x3_kernel  G = A^H B for 64 x 64 tiles of state pairs (I <= J, or
//                      all pairs for a non-Hermitian product), split over the
//                      grid points. Each FP32 value is split
//                      into three bfloat16 parts x = x1 + x2 + x3; the six
//                      products x_p y_q with p + q <= 4 (1-based) run on
//                      v_mfma_f32_16x16x32_bf16 and carry FP32 accuracy.
//                      The complex product is two real dot products over the
//                      interleaved (re, im) sequence: Re = sum a.b,
//                      Im = sum a.b' with b' = (b_im, -b_re) per element.
//                      MFMA partial sums are FP32 over one 64-real step, then
//                      added in FP64; the chunk partials are reduced in FP64.
//  rotate_f32_kernel   Psi_out = Psi_in C on v_mfma_f32_16x16x4_f32.
//  residual_kernel     sigma_n = ||W_n - E_n Psi_n|| dx (FP64 sums).
//
// Lane layouts (measured on gfx950): for both v_mfma_f32_16x16x4_f32 (KL = 1)
// and v_mfma_f32_16x16x32_bf16 (KL = 8 elements per lane), A[i][k] comes from
// lane i + 16 (k / KL), element k % KL; B[k][j] from lane j + 16 (k / KL),
// element k % KL; D[i][j] lands in lane j + 16 (i / 4), register i % 4.
#include <hip/hip_runtime.h>

#include "gpu.hpp"
#include "line_ops.hpp"

namespace itp {

typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef double d4m __attribute__((ext_vector_type(4)));
typedef __bf16 bf4v __attribute__((ext_vector_type(4)));
typedef __bf16 bf2v __attribute__((ext_vector_type(2)));

constexpr int kG3Rows = 64;     // states per tile side
constexpr int kG3KT = 64;       // reals per K step
constexpr int kG3Pad = 8;       // bf16 row padding in LDS
constexpr int kG3Sub = 64;      // reals per FP32 partial sum (one K step: 12 MFMA roundings)

// x = x1 + x2 + x3 in bfloat16 (round to nearest each time)
__device__ __forceinline__ void split3(float x, __bf16& h, __bf16& m, __bf16& l) {
  h = (__bf16)x;
  const float r1 = x - (float)h;
  m = (__bf16)r1;
  l = (__bf16)(r1 - (float)m);
}
// the same for two values at once: each rounding is one packed conversion
typedef float f2v __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void split3x2(f2v x, bf2v& h, bf2v& m, bf2v& l) {
  h = __builtin_convertvector(x, bf2v);
  const f2v r1 = x - __builtin_convertvector(h, f2v);
  m = __builtin_convertvector(r1, bf2v);
  l = __builtin_convertvector(r1 - __builtin_convertvector(m, f2v), bf2v);
}
__device__ __forceinline__ bf4v cat2(bf2v a, bf2v b) { return __builtin_shufflevector(a, b, 0, 1, 2, 3); }

// (b0, b1, b2, ...) -> (b1, -b0, b3, -b2, ...): the operand of the imaginary part
__device__ __forceinline__ bf8 swapneg(bf8 b) {
  bf8 r;
#pragma unroll
  for (int c = 0; c < 4; c++) {
    r[2 * c] = b[2 * c + 1];
    r[2 * c + 1] = -b[2 * c];
  }
  return r;
}

// Eight wavefronts per workgroup: wavefront (wi, wj) = (w / 4, w % 4) holds
// rows wi*32 .. +31 of the I tile against columns wj*16 .. +15 of the J tile
// (two 16 x 16 MFMA tiles), so that a wavefront's registers allow four
// wavefronts per SIMD with two workgroups (the LDS limit) per CU.
constexpr int kG3Threads = 512;
__global__ __launch_bounds__(kG3Threads) void gram_bf16x3_kernel(const float* A, const float* B, long long M, int N,
                                                                 int nb, long long kchunk, cplx<double>* part, int nc,
                                                                 int full) {
  __shared__ __align__(16) __bf16 As[3][kG3Rows][kG3KT + kG3Pad];
  __shared__ __align__(16) __bf16 Bs[3][kG3Rows][kG3KT + kG3Pad];
  // tile pair (all of them if full, else I <= J) and grid chunk in the block
  // order of xcd_tile: every XCD takes whole chunks and all pairs of a chunk
  const int npairs = full ? nb * nb : nb * (nb + 1) / 2;
  int pair;
  long long chunk;
  if (!xcd_tile(blockIdx.x, npairs, (M + kchunk - 1) / kchunk, pair, chunk)) return;
  int I = 0, J = 0;
  if (full) {
    I = pair / nb;
    J = pair - I * nb;
  } else {
    int rem = pair;
    while (rem >= nb - I) { rem -= nb - I; I++; }
    J = I + rem;
  }
  const long long K2 = 2 * M;  // reals per field
  const long long k_begin = (long long)chunk * 2 * kchunk;
  const long long k_end = k_begin + 2 * kchunk < K2 ? k_begin + 2 * kchunk : K2;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wi = w >> 2, wj = w & 3;
  const int i0 = I * kG3Rows, j0 = J * kG3Rows;
  f4 re[2], im[2];
  d4m dre[2], dim[2];
#pragma unroll
  for (int x = 0; x < 2; x++) {
    re[x] = f4{0, 0, 0, 0};
    im[x] = f4{0, 0, 0, 0};
    dre[x] = d4m{0, 0, 0, 0};
    dim[x] = d4m{0, 0, 0, 0};
  }
  int sub = 0;
  // the next step's reals are loaded while this step's products run (the
  // loads of step k0 + 64 are issued before its MFMAs); item j of a thread is
  // row (tid + 512 j) / 16, reals 4 ((tid + 512 j) % 16) .. + 3 of the step
  constexpr int kItems = kG3Rows * kG3KT / 4 / kG3Threads;
  float4 nxa[kItems], nxb[kItems];
  auto load_step = [&](long long k0) {
#pragma unroll
    for (int j = 0; j < kItems; j++) {
      const int idx = tid + kG3Threads * j;
      const int row = idx >> 4, c4 = idx & 15;
      const long long kk = k0 + 4 * c4;
      nxa[j] = make_float4(0.f, 0.f, 0.f, 0.f);
      nxb[j] = make_float4(0.f, 0.f, 0.f, 0.f);
      if (kk < k_end) {  // k_end - kk is a multiple of 4 (2 M and the chunk bounds are)
        if (i0 + row < N) nxa[j] = *reinterpret_cast<const float4*>(A + (long long)(i0 + row) * K2 + kk);
        if (j0 + row < N) nxb[j] = *reinterpret_cast<const float4*>(B + (long long)(j0 + row) * K2 + kk);
      }
    }
  };
  load_step(k_begin);
  for (long long k0 = k_begin; k0 < k_end; k0 += kG3KT) {
    // stage 64 rows x 64 reals of A and of B as three bf16 parts
#pragma unroll
    for (int j = 0; j < kItems; j++) {
      const int idx = tid + kG3Threads * j;
      const int row = idx >> 4, c4 = idx & 15;
      const float4 va = nxa[j], vb = nxb[j];
      bf2v h0, m0, l0, h1, m1, l1;
      split3x2(f2v{va.x, va.y}, h0, m0, l0);
      split3x2(f2v{va.z, va.w}, h1, m1, l1);
      *reinterpret_cast<bf4v*>(&As[0][row][4 * c4]) = cat2(h0, h1);
      *reinterpret_cast<bf4v*>(&As[1][row][4 * c4]) = cat2(m0, m1);
      *reinterpret_cast<bf4v*>(&As[2][row][4 * c4]) = cat2(l0, l1);
      split3x2(f2v{vb.x, vb.y}, h0, m0, l0);
      split3x2(f2v{vb.z, vb.w}, h1, m1, l1);
      *reinterpret_cast<bf4v*>(&Bs[0][row][4 * c4]) = cat2(h0, h1);
      *reinterpret_cast<bf4v*>(&Bs[1][row][4 * c4]) = cat2(m0, m1);
      *reinterpret_cast<bf4v*>(&Bs[2][row][4 * c4]) = cat2(l0, l1);
    }
    __syncthreads();
    if (k0 + kG3KT < k_end) load_step(k0 + kG3KT);
#pragma unroll
    for (int ks = 0; ks < kG3KT / 32; ks++) {
      const int kcol = ks * 32 + 8 * (lane >> 4);
      bf8 a[3][2], b[3], bp[3];
#pragma unroll
      for (int s = 0; s < 3; s++) {
#pragma unroll
        for (int x = 0; x < 2; x++)
          a[s][x] = *reinterpret_cast<const bf8*>(&As[s][wi * 32 + x * 16 + (lane & 15)][kcol]);
        b[s] = *reinterpret_cast<const bf8*>(&Bs[s][wj * 16 + (lane & 15)][kcol]);
        bp[s] = swapneg(b[s]);
      }
#pragma unroll
      for (int x = 0; x < 2; x++) {
        // split products with p + q <= 4 (1-based): (1,1) (1,2) (2,1) (2,2) (1,3) (3,1)
        re[x] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[2][x], b[0], re[x], 0, 0, 0);
        re[x] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0][x], b[2], re[x], 0, 0, 0);
        re[x] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[1][x], b[1], re[x], 0, 0, 0);
        re[x] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[1][x], b[0], re[x], 0, 0, 0);
        re[x] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0][x], b[1], re[x], 0, 0, 0);
        re[x] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0][x], b[0], re[x], 0, 0, 0);
        im[x] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[2][x], bp[0], im[x], 0, 0, 0);
        im[x] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0][x], bp[2], im[x], 0, 0, 0);
        im[x] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[1][x], bp[1], im[x], 0, 0, 0);
        im[x] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[1][x], bp[0], im[x], 0, 0, 0);
        im[x] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0][x], bp[1], im[x], 0, 0, 0);
        im[x] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0][x], bp[0], im[x], 0, 0, 0);
      }
    }
    __syncthreads();
    // FP32 partial sums span at most kG3Sub reals, then move to FP64
    if (++sub == kG3Sub / kG3KT) {
      sub = 0;
#pragma unroll
      for (int x = 0; x < 2; x++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
          dre[x][v] += (double)re[x][v];
          dim[x][v] += (double)im[x][v];
          re[x][v] = 0.f;
          im[x][v] = 0.f;
        }
    }
  }
#pragma unroll
  for (int x = 0; x < 2; x++)
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const int i = i0 + wi * 32 + x * 16 + 4 * (lane >> 4) + v;
      const int j = j0 + wj * 16 + (lane & 15);
      part[((long long)chunk * nc + i) * nc + j] = mk<double>(dre[x][v] + (double)re[x][v], dim[x][v] + (double)im[x][v]);
    }
}

hipError_t launch_gram_bf16x3(const cplx<float>* A, const cplx<float>* B, long long M, int N, int nc,
                              cplx<double>* part, int nchunk, long long kchunk, double dx, cplx<double>* S, int np,
                              int* bad, hipStream_t st, int full) {
  if (nc % kG3Rows || (2 * kchunk) % kG3Sub || M % 2) return hipErrorInvalidValue;
  const int nb = nc / kG3Rows;
  const int npairs = full ? nb * nb : nb * (nb + 1) / 2;
  if (nchunk != (M + kchunk - 1) / kchunk) return hipErrorInvalidValue;
  const long long blocks = xcd_grid(npairs, nchunk);
  if (blocks > 0x7fffffffLL) return hipErrorInvalidValue;
  hipLaunchKernelGGL(gram_bf16x3_kernel, dim3((unsigned)blocks), dim3(kG3Threads), 0, st, reinterpret_cast<const float*>(A),
                     reinterpret_cast<const float*>(B), M, N, nb, kchunk, part, nc, full);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return launch_gram_reduce(part, nchunk, nc, N, np, dx, S, bad, st, full, 4);
}

// ---------------------------------------------------------------------------
// Psi_out[r][t] = sum_c C[c][r] Psi_in[c][t] for complex FP32 fields and an
// FP32 copy of the coefficients, on v_mfma_f32_16x16x4_f32: 64 output states
// x 64 grid points per workgroup, one 32 x 32 quarter per wavefront.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void rotate_f32_kernel(const cplx<float>* in, cplx<float>* out, const cplx<float>* C,
                                                         int nc, long long M, int N) {
  __shared__ cplx<float> Cs[16][65];
  __shared__ cplx<float> Ps[16][65];
  int rt;
  long long tt;
  // block order by xcd_tile: each XCD takes whole grid-point slabs (read from
  // HBM once) and all output-state tiles of a slab
  if (!xcd_tile(blockIdx.x, nc / 64, (M + 63) / 64, rt, tt)) return;
  const long long t0 = tt * 64;
  const int r0 = rt * 64;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, wr = wave >> 1, wt = wave & 1;
  const int li = lane & 15, lk = lane >> 4;
  f4 accr[2][2], acci[2][2];
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++) {
      accr[a][b] = f4{0, 0, 0, 0};
      acci[a][b] = f4{0, 0, 0, 0};
    }
  for (int c0 = 0; c0 < N; c0 += 16) {
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int e = tid + 256 * q;
      const int kc = e >> 6, x = e & 63;
      const int c = c0 + kc;
      Cs[kc][x] = (c < N) ? C[(long long)c * nc + r0 + x] : mk<float>(0.f, 0.f);
      Ps[kc][x] = (c < N && t0 + x < M) ? in[(long long)c * M + t0 + x] : mk<float>(0.f, 0.f);
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < 4; kk++) {
      cplx<float> a[2], b[2];
#pragma unroll
      for (int mi = 0; mi < 2; mi++) a[mi] = Cs[kk * 4 + lk][wr * 32 + mi * 16 + li];
#pragma unroll
      for (int nj = 0; nj < 2; nj++) b[nj] = Ps[kk * 4 + lk][wt * 32 + nj * 16 + li];
#pragma unroll
      for (int mi = 0; mi < 2; mi++)
#pragma unroll
        for (int nj = 0; nj < 2; nj++) {
          accr[mi][nj] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[mi].x, b[nj].x, accr[mi][nj], 0, 0, 0);
          accr[mi][nj] = __builtin_amdgcn_mfma_f32_16x16x4f32(-a[mi].y, b[nj].y, accr[mi][nj], 0, 0, 0);
          acci[mi][nj] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[mi].x, b[nj].y, acci[mi][nj], 0, 0, 0);
          acci[mi][nj] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[mi].y, b[nj].x, acci[mi][nj], 0, 0, 0);
        }
    }
    __syncthreads();
  }
#pragma unroll
  for (int mi = 0; mi < 2; mi++)
#pragma unroll
    for (int nj = 0; nj < 2; nj++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int r = r0 + wr * 32 + mi * 16 + 4 * lk + v;
        const long long t = t0 + wt * 32 + nj * 16 + li;
        if (r < N && t < M) out[(long long)r * M + t] = mk<float>(accr[mi][nj][v], acci[mi][nj][v]);
      }
}

// ---------------------------------------------------------------------------
// Psi_out[r][t] = sum_c C[c][r] Psi_in[c][t] for complex FP32 fields and an
// FP32 copy of the coefficients, on v_mfma_f32_16x16x32_bf16 with the
// split-BF16x3 products of the overlap (p + q <= 4): 64 output states x 64
// grid points per workgroup, K = 16 input states (32 interleaved reals) per
// step. A[r][2c + q] holds C[c][r] (re, im); the states are staged as
// (re, im) pairs and turned in registers into (re, -im) for the real part and
// (im, re) for the imaginary part (negating a bfloat16 is exact). FP32
// accumulation over all input states, as on the FP32 MFMA path.
// ---------------------------------------------------------------------------
constexpr int kR3Pad = 8;  // bf16 row padding in LDS

// (b0, b1, b2, b3, ...) -> (b0, -b1, b2, -b3, ...): the real part's operand
__device__ __forceinline__ bf8 negodd(bf8 b) {
  bf8 r;
#pragma unroll
  for (int c = 0; c < 4; c++) {
    r[2 * c] = b[2 * c];
    r[2 * c + 1] = -b[2 * c + 1];
  }
  return r;
}
// (b0, b1, b2, b3, ...) -> (b1, b0, b3, b2, ...): the imaginary part's operand
__device__ __forceinline__ bf8 swappairs(bf8 b) {
  bf8 r;
#pragma unroll
  for (int c = 0; c < 4; c++) {
    r[2 * c] = b[2 * c + 1];
    r[2 * c + 1] = b[2 * c];
  }
  return r;
}

// Eight wavefronts per workgroup: wavefront (wr, wt) = (w / 4, w % 4) holds
// output states wr*32 .. +31 against grid points wt*16 .. +15 (two 16 x 16
// MFMA tiles). The register bound of six wavefronts per SIMD (at most 85
// VGPRs; 78 without spills) lets three workgroups share a CU (30 KiB of LDS
// each); at 88 VGPRs only two fit, which measured slower.
constexpr int kR3Threads = 512;
__global__ __launch_bounds__(kR3Threads) __attribute__((amdgpu_waves_per_eu(6))) void rotate_bf16x3_kernel(
    const cplx<float>* in, cplx<float>* out, const cplx<float>* C, int nc, long long M, int N, int upper) {
  __shared__ __align__(16) __bf16 As[3][64][32 + kR3Pad];  // [part][output state r][interleaved input k]
  __shared__ __align__(16) __bf16 Bs[3][64][32 + kR3Pad];  // [part][grid point t][interleaved input k]
  int rt;
  long long tt;
  // block order by xcd_tile: each XCD takes whole grid-point slabs and all
  // output-state tiles of a slab
  if (!xcd_tile(blockIdx.x, nc / 64, (M + 63) / 64, rt, tt)) return;
  const long long t0 = tt * 64;
  const int r0 = rt * 64;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, wr = wave >> 2, wt = wave & 3;
  f4 re[2], im[2];
#pragma unroll
  for (int x = 0; x < 2; x++) {
    re[x] = f4{0, 0, 0, 0};
    im[x] = f4{0, 0, 0, 0};
  }
  // upper triangular C: output states r0..r0+63 take inputs c < r0 + 64 only
  const int cend = upper ? (N < r0 + 64 ? N : r0 + 64) : N;
  // the next step's coefficients and states are loaded while this step's
  // products run (the loads of step c0 + 16 are issued before its MFMAs)
  cplx<float> nxa[2], nxp[2];
  // thread (w, x) = (tid / 64, tid % 64) stages inputs cc = 2 w + q of row x:
  // global loads coalesced along x, and one 8-byte LDS store per split part
  auto load_step = [&](int c0) {
#pragma unroll
    for (int q = 0; q < 2; q++) {
      const int cc = 2 * (tid >> 6) + q, x = tid & 63;
      const int c = c0 + cc;
      // C is zero-padded to nc x nc; states c >= N contribute nothing
      const cplx<float> cv = C[(long long)(c < nc ? c : nc - 1) * nc + r0 + x];
      nxa[q] = (c < N) ? cv : mk<float>(0.f, 0.f);
      nxp[q] = (c < N && t0 + x < M) ? in[(long long)c * M + t0 + x] : mk<float>(0.f, 0.f);
    }
  };
  load_step(0);
  for (int c0 = 0; c0 < cend; c0 += 16) {
    {
      // the (re, im) pairs of inputs 2 w, 2 w + 1, per split part
      bf2v a[2][3], p[2][3];
#pragma unroll
      for (int q = 0; q < 2; q++) {
        split3x2(f2v{nxa[q].x, nxa[q].y}, a[q][0], a[q][1], a[q][2]);
        split3x2(f2v{nxp[q].x, nxp[q].y}, p[q][0], p[q][1], p[q][2]);
      }
      const int x = tid & 63, k4 = 4 * (tid >> 6);
#pragma unroll
      for (int sp = 0; sp < 3; sp++) {
        *reinterpret_cast<bf4v*>(&As[sp][x][k4]) = cat2(a[0][sp], a[1][sp]);
        *reinterpret_cast<bf4v*>(&Bs[sp][x][k4]) = cat2(p[0][sp], p[1][sp]);
      }
    }
    __syncthreads();
    if (c0 + 16 < cend) load_step(c0 + 16);
    const int kcol = 8 * (lane >> 4);
    bf8 a[3][2], br[3], bi[3];
#pragma unroll
    for (int sp = 0; sp < 3; sp++) {
#pragma unroll
      for (int x = 0; x < 2; x++)
        a[sp][x] = *reinterpret_cast<const bf8*>(&As[sp][wr * 32 + x * 16 + (lane & 15)][kcol]);
      const bf8 b = *reinterpret_cast<const bf8*>(&Bs[sp][wt * 16 + (lane & 15)][kcol]);
      br[sp] = negodd(b);
      bi[sp] = swappairs(b);
    }
#pragma unroll
    for (int x = 0; x < 2; x++) {
      // split products with p + q <= 4 (1-based), smallest first
      re[x] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[2][x], br[0], re[x], 0, 0, 0);
      re[x] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0][x], br[2], re[x], 0, 0, 0);
      re[x] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[1][x], br[1], re[x], 0, 0, 0);
      re[x] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[1][x], br[0], re[x], 0, 0, 0);
      re[x] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0][x], br[1], re[x], 0, 0, 0);
      re[x] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0][x], br[0], re[x], 0, 0, 0);
      im[x] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[2][x], bi[0], im[x], 0, 0, 0);
      im[x] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0][x], bi[2], im[x], 0, 0, 0);
      im[x] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[1][x], bi[1], im[x], 0, 0, 0);
      im[x] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[1][x], bi[0], im[x], 0, 0, 0);
      im[x] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0][x], bi[1], im[x], 0, 0, 0);
      im[x] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0][x], bi[0], im[x], 0, 0, 0);
    }
    __syncthreads();
  }
#pragma unroll
  for (int x = 0; x < 2; x++)
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const int r = r0 + wr * 32 + x * 16 + 4 * (lane >> 4) + v;
      const long long t = t0 + wt * 16 + (lane & 15);
      if (r < N && t < M) out[(long long)r * M + t] = mk<float>(re[x][v], im[x][v]);
    }
}

hipError_t launch_rotate_bf16x3(const cplx<float>* in, cplx<float>* out, const cplx<float>* C, int nc, long long M,
                                int N, hipStream_t st, int upper) {
  const long long blocks = xcd_grid(nc / 64, (M + 63) / 64);
  if (blocks > 0x7fffffffLL || nc % 64) return hipErrorInvalidValue;
  hipLaunchKernelGGL(rotate_bf16x3_kernel, dim3((unsigned)blocks), dim3(kR3Threads), 0, st, in, out, C, nc, M, N, upper);
  return hipGetLastError();
}

hipError_t launch_rotate_f32(const cplx<float>* in, cplx<float>* out, const cplx<float>* C, int nc, long long M, int N,
                             hipStream_t st) {
  const long long blocks = xcd_grid(nc / 64, (M + 63) / 64);
  if (blocks > 0x7fffffffLL || nc % 64) return hipErrorInvalidValue;
  hipLaunchKernelGGL(rotate_f32_kernel, dim3((unsigned)blocks), dim3(256), 0, st, in, out, C, nc, M, N);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// sigma_n = dx ||W_n - E_n Psi_n||: one workgroup per state, FP64 sums.
// ---------------------------------------------------------------------------
template <class R>
__global__ __launch_bounds__(256) void residual_kernel(const cplx<R>* W, const cplx<R>* Psi, const double* E,
                                                       long long M, double dx, double* sigma) {
  __shared__ double red[256];
  const int n = blockIdx.x;
  const double e = E[n];
  const cplx<R>* w = W + (long long)n * M;
  const cplx<R>* p = Psi + (long long)n * M;
  double acc = 0;
  for (long long i = threadIdx.x; i < M; i += 256) {
    const double rx = double(w[i].x) - e * double(p[i].x);
    const double ry = double(w[i].y) - e * double(p[i].y);
    acc += rx * rx + ry * ry;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int h = 128; h >= 1; h >>= 1) {
    if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) sigma[n] = std::sqrt(red[0]) * dx;
}

template <class R>
hipError_t launch_residual(const cplx<R>* W, const cplx<R>* Psi, const double* E, long long M, int N, double dx,
                           double* sigma, hipStream_t st) {
  hipLaunchKernelGGL(residual_kernel<R>, dim3(N), dim3(256), 0, st, W, Psi, E, M, dx, sigma);
  return hipGetLastError();
}
template hipError_t launch_residual<double>(const cplx<double>*, const cplx<double>*, const double*, long long, int,
                                            double, double*, hipStream_t);
template hipError_t launch_residual<float>(const cplx<float>*, const cplx<float>*, const double*, long long, int,
                                           double, double*, hipStream_t);

}  // namespace itp
