// The register engine's sweep kernels (reg_sweep.hpp) for one working
// precision and one line length, fixed at compile time by ITP_SWEEP_REAL
// (float or double) and ITP_SWEEP_N. The build compiles this file once per
// pair, so the instantiations compile in parallel; launch_reg_sweep in
// kernels.hip picks the launcher for a sweep's line length.
//
// Besides the kernel that reads its op list at run time, the sweeps of the
// split-operator members (engine.hpp member_program: head, column, inner and
// last sweeps of each kinetic branch, with a potential) have kernels compiled
// for their op lists: their op loop is straight-line code, and in FP32 they
// read the FP32 copies of the tables.
#include <hip/hip_runtime.h>

#include "gpu.hpp"
#include "reg_sweep.hpp"

#if !defined(ITP_SWEEP_REAL) || !defined(ITP_SWEEP_N)
#error "compile with -DITP_SWEEP_REAL=<float|double> -DITP_SWEEP_N=<line length>"
#endif

namespace itp {

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

// At least two waves per SIMD: the line exchanges are latency-bound, and the
// LDS of two workgroups per CU allows it.
template <class R, int N, int ORIENT, bool XS, class PG>
__global__ __launch_bounds__(256, 2) void reg_sweep_kernel(SweepArgs<R> a) {
  extern __shared__ __align__(16) char lds_raw[];
  GpuRegExec<RegState<R, RegShape<N>::Q>> ex;
  int field = (int)blockIdx.y, blk = (int)blockIdx.x;
  if constexpr (ORIENT == 1 && reg_lines_per_wg(N) * sizeof(cplx<R>) < 128) {
    // A workgroup's columns are narrower than a 128-byte line: workgroups are
    // dealt to the XCDs in turn, so neighbouring column groups would fetch the
    // same lines into different L2s (4x the read traffic for 4-column groups).
    // Each XCD takes whole fields instead, their column groups one after
    // another, so the groups that share lines share them in one L2.
    const long long b = (long long)blockIdx.y * gridDim.x + blockIdx.x;
    long long outer = 0;
    if (!xcd_tile(b, (int)gridDim.x, (long long)gridDim.y, blk, outer)) return;
    field = (int)outer;
  }
  reg_sweep_program<R, N, ORIENT, XS, PG>(ex, a, reinterpret_cast<cplx<R>*>(lds_raw), field, blk);
}

namespace {

constexpr int T_ = LOP_MUL_TAB, P_ = LOP_MUL_POS, F_ = LOP_FFT_FWD, I_ = LOP_FFT_INV, D2 = LOP_DST2,
              D3 = LOP_DST3, X_ = LOP_XSPLIT;

// member sweeps with a potential (engine.hpp fwd_x_ops, inv_x_ops, col_ops,
// member_program), per kinetic branch: B = 0 periodic, B != 0 periodic,
// B = 0 Dirichlet, B != 0 Dirichlet
using RowPrograms = std::tuple<
    CtProg<T_, F_>, CtProg<I_, T_, F_>, CtProg<I_, T_>,                      // B = 0 periodic
    CtProg<T_, F_, T_>, CtProg<T_, I_, T_, F_, T_>, CtProg<T_, I_, T_>,       // B != 0 periodic
    CtProg<T_, D2>, CtProg<D3, T_, D2>, CtProg<D3, T_>>;                      // B = 0 Dirichlet
// (the FP64 last row sweep, split then table, spills registers when compiled
// for its op list: the runtime kernel runs it)
using XsRowPrograms = std::conditional_t<
    sizeof(ITP_SWEEP_REAL) == 8, std::tuple<CtProg<T_, D2, X_, D2>, CtProg<X_, T_, D2, X_, D2>>,
    std::tuple<CtProg<T_, D2, X_, D2>, CtProg<X_, T_, D2, X_, D2>, CtProg<X_, T_>>>;
using ColPrograms = std::tuple<CtProg<F_, T_, I_>, CtProg<F_, P_, I_>, CtProg<D2, T_, D3>, CtProg<D2, P_, D3>>;

template <class R, int N, int ORIENT, bool XS, class PG>
hipError_t launch_one(const SweepArgs<R>& a, int nblk, int nfields, hipStream_t st) {
  // once per kernel (thread-safe static initialisation)
  static const hipError_t attr_set = hipFuncSetAttribute((const void*)reg_sweep_kernel<R, N, ORIENT, XS, PG>,
                                                         hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
  (void)attr_set;
  hipLaunchKernelGGL((reg_sweep_kernel<R, N, ORIENT, XS, PG>), dim3(nblk, nfields), dim3(kRegThreads),
                     reg_lds_bytes<R>(N), st, a);
  return hipGetLastError();
}

// Launches the first compile-time program of the tuple that matches the op
// list; false if none does.
template <class R, int N, int ORIENT, bool XS, class... PGs>
bool launch_ct(std::tuple<PGs...>*, const SweepArgs<R>& a, int nblk, int nfields, hipStream_t st, hipError_t& err) {
  bool done = false;
  auto one = [&](auto* pg) {
    using PG = std::remove_pointer_t<decltype(pg)>;
    if (done || !ct_matches<PG>(a.ops, a.nops)) return;
    err = launch_one<R, N, ORIENT, XS, PG>(a, nblk, nfields, st);
    done = true;
  };
  (one((PGs*)nullptr), ...);
  return done;
}

// Compile-time FP32 programs read the FP32 table copies: every table an op
// uses must have one.
bool have_f32_tables(const LineOp* ops, int nops) {
  for (int i = 0; i < nops; i++) {
    const LineOp& o = ops[i];
    if ((o.code == LOP_MUL_TAB || o.code == LOP_MUL_POS) && !o.f0) return false;
    if (o.code == LOP_XSPLIT && (!o.f0 || !o.f1)) return false;
  }
  return true;
}

}  // namespace

template <class R, int N>
hipError_t launch_reg_sweep_n(const SweepArgs<R>& a, int nblk, int nfields, hipStream_t st) {
  const bool xs = ops_have_xsplit(a.ops, a.nops);
  constexpr bool kCt = N >= 64;  // shorter lines appear only in tests
  if (kCt && (sizeof(R) == 8 || have_f32_tables(a.ops, a.nops))) {
    hipError_t err = hipSuccess;
    if (a.orient != 0) {
      if (launch_ct<R, N, 1, false>((ColPrograms*)nullptr, a, nblk, nfields, st, err)) return err;
    } else if (xs) {
      if (launch_ct<R, N, 0, true>((XsRowPrograms*)nullptr, a, nblk, nfields, st, err)) return err;
    } else {
      if (launch_ct<R, N, 0, false>((RowPrograms*)nullptr, a, nblk, nfields, st, err)) return err;
    }
  }
  if (a.orient != 0) return launch_one<R, N, 1, false, RtProg>(a, nblk, nfields, st);
  if (xs) return launch_one<R, N, 0, true, RtProg>(a, nblk, nfields, st);
  return launch_one<R, N, 0, false, RtProg>(a, nblk, nfields, st);
}

template hipError_t launch_reg_sweep_n<ITP_SWEEP_REAL, ITP_SWEEP_N>(const SweepArgs<ITP_SWEEP_REAL>&, int, int,
                                                                    hipStream_t);

}  // namespace itp
