// The register engine's sweep kernels (reg_sweep.hpp) for one working
// precision and one line length, fixed at compile time by ITP_SWEEP_REAL
// (float or double) and ITP_SWEEP_N. The build compiles this file once per
// pair, so the instantiations compile in parallel; launch_reg_sweep in
// kernels.hip picks the launcher for a sweep's line length.
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
template <class R, int N, int ORIENT, bool XS>
__global__ __launch_bounds__(256, 2) void reg_sweep_kernel(SweepArgs<R> a) {
  extern __shared__ __align__(16) char lds_raw[];
  GpuRegExec<RegState<R, RegShape<N>::Q>> ex;
  reg_sweep_program<R, N, ORIENT, XS>(ex, a, reinterpret_cast<cplx<R>*>(lds_raw), (int)blockIdx.y, (int)blockIdx.x);
}

template <class R, int N>
hipError_t launch_reg_sweep_n(const SweepArgs<R>& a, int nblk, int nfields, hipStream_t st) {
  const size_t lds = reg_lds_bytes<R>(N);
  auto go = [&](auto kern) {
    // once per instantiation (thread-safe static initialisation)
    static const hipError_t attr_set =
        hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    (void)attr_set;
    hipLaunchKernelGGL(kern, dim3(nblk, nfields), dim3(kRegThreads), lds, st, a);
    return hipGetLastError();
  };
  if (a.orient != 0) return go(reg_sweep_kernel<R, N, 1, false>);
  if (ops_have_xsplit(a.ops, a.nops)) return go(reg_sweep_kernel<R, N, 0, true>);
  return go(reg_sweep_kernel<R, N, 0, false>);
}

template hipError_t launch_reg_sweep_n<ITP_SWEEP_REAL, ITP_SWEEP_N>(const SweepArgs<ITP_SWEEP_REAL>&, int, int,
                                                                    hipStream_t);

}  // namespace itp
