// Linear device load for gfx950: the standalone launch of the linear
// devices (resistors, capacitors, diodes, sources; device code in
// xg_linear_dev.h).
#include "xg_linear_dev.h"

namespace xg {

namespace {

__global__ void __launch_bounds__(256)
    linear_eval_kernel(LinArgs a, SolverStateDev ss, LoadPtrs p, double time,
                       double bp_tol, const double *dstep) {
  if (p.gate && *p.gate) return;
  lin::linear_eval_one(a, ss, p, time, bp_tol, dstep, blockIdx.x * 256 + threadIdx.x);
}

}  // namespace

hipError_t launch_linear_eval(const LinArgs &a, const SolverStateDev &ss,
                              const LoadPtrs &p, double time, double bp_tol,
                              const double *dstep, hipStream_t s) {
  if (a.n_slots == 0) return hipSuccess;
  hipLaunchKernelGGL(linear_eval_kernel, dim3((a.n_slots + 255) / 256),
                     dim3(256), 0, s, a, ss, p, time, bp_tol, dstep);
  return hipGetLastError();
}

}  // namespace xg
