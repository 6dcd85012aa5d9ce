// Ordered segmented reduction: folds the per-statement stamp values of the
// slot buffer into output vectors / CSR values without atomics, in the
// reference's accumulation order, so the result is bitwise reproducible.
//
// For every touched output entry k of a program:
//   acc = out[k];  for each source s of k in order:  acc = acc + slot[s];
//   out[k] = acc;
// which is the sequence of "*ptr += v" / "vec[li] += v" statements the
// reference executes on a zeroed matrix (N_DEV_MOSFET1.C:4441-4722,
// N_DEV_MOSFET_B3.C:9654-9928; vector loads likewise).
//
// Several programs (e.g. F, Q, dFdxdVp, dQdxdVp) run in one launch: the grid
// is the concatenation of their work items.
#include "xg_internal.h"

namespace xg {

namespace {

constexpr int kMaxProgs = 4;
constexpr int kBlock = 256;

struct ReduceBatch {
  int n;
  int32_t start[kMaxProgs + 1];
  ReduceDev prog[kMaxProgs];
  double *out[kMaxProgs];
};

__global__ void __launch_bounds__(kBlock)
    reduce_kernel(ReduceBatch b, const double *__restrict__ slots) {
  const int w = blockIdx.x * kBlock + threadIdx.x;
  if (w >= b.start[b.n]) return;
  int k = 0;
#pragma unroll
  for (int j = 1; j < kMaxProgs; ++j)
    if (j < b.n && w >= b.start[j]) k = j;
  const ReduceDev &p = b.prog[k];
  const int i = w - b.start[k];
  const int32_t t = p.tgt[i];
  const int32_t beg = p.ptr[i], end = p.ptr[i + 1];
  double *out = b.out[k];
  double acc = out[t];
  for (int32_t j = beg; j < end; ++j) {
    const int32_t s = p.src[j];
    acc = acc + (s >= 0 ? slots[s] : 0.0);
  }
  out[t] = acc;
}

__global__ void count_kernel(const int32_t *flags, int32_t n, int64_t *out) {
  __shared__ int64_t part[256];
  int64_t c = 0;
  for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += gridDim.x * blockDim.x)
    c += flags[i] != 0;
  part[threadIdx.x] = c;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) part[threadIdx.x] += part[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = part[0];
}

}  // namespace

hipError_t launch_reduce(int n_out, const ReduceDev *progs, double *const *outs,
                         const double *slots, hipStream_t s) {
  if (n_out <= 0 || n_out > kMaxProgs) return hipErrorInvalidValue;
  ReduceBatch b{};
  b.n = n_out;
  b.start[0] = 0;
  for (int k = 0; k < n_out; ++k) {
    b.prog[k] = progs[k];
    b.out[k] = outs[k];
    if (!outs[k] && progs[k].n_work) return hipErrorInvalidValue;
    b.start[k + 1] = b.start[k] + progs[k].n_work;
  }
  for (int k = n_out; k < kMaxProgs; ++k) b.start[k + 1] = b.start[n_out];
  const int total = b.start[n_out];
  if (total == 0) return hipSuccess;
  const int grid = (total + kBlock - 1) / kBlock;
  hipLaunchKernelGGL(reduce_kernel, dim3(grid), dim3(kBlock), 0, s, b, slots);
  return hipGetLastError();
}

// Writes one partial count per block (64 blocks); the caller sums them.
hipError_t launch_count_flags(const int32_t *flags, int32_t n, int64_t *out,
                              hipStream_t s) {
  hipLaunchKernelGGL(count_kernel, dim3(64), dim3(256), 0, s, flags, n, out);
  return hipGetLastError();
}

}  // namespace xg
