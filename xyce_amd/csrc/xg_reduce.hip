// Ordered segmented reduction: folds the per-statement stamp values of the
// slot buffer into output vectors / CSR values without atomics, in a fixed
// order, so results are bitwise reproducible run to run.
//
// Every output entry k owns a segment of sources (slot indices) listed in
// the reference's accumulation order (device type, instance, statement).
// Segments are cut into chunks of at most kChunk sources.
//
//  * A segment of one chunk is folded by one thread starting from out[k]:
//      acc = out[k]; acc = acc + s0; acc = acc + s1; ...; out[k] = acc
//    which is exactly the sequence of "*ptr += v" / "vec[li] += v"
//    statements the reference executes (N_DEV_MOSFET1.C:4441-4722,
//    N_DEV_MOSFET_B3.C:9654-9928), so these entries are bit-identical to
//    the reference's order.
//  * A longer segment (a supply rail or a heavily shared node) is folded in
//    two passes: each chunk is summed in order into a partial, then one
//    wavefront adds the partials to out[k] in chunk order (lane runs in
//    lane order).  The order is fixed; the result differs from the strictly
//    sequential sum by rounding only.
//
// Sources are gathered eight at a time so that a thread keeps eight loads
// in flight; the additions themselves stay in order.
#include "xg_internal.h"

namespace xg {

namespace {

constexpr int kMaxProgs = 6;
constexpr int kBlock = 256;
constexpr int kWaves = kBlock / 64;

struct ReduceBatch {
  int n;
  int32_t cstart[kMaxProgs + 1];  // chunk work items, prefix over programs
  int32_t mstart[kMaxProgs + 1];  // multi-chunk segments, prefix
  ReduceDev prog[kMaxProgs];
  double *out[kMaxProgs];
  double *partial;                // scratch for chunk partial sums
};

__device__ __forceinline__ int find_prog(const int32_t *start, int n, int w) {
  int k = 0;
#pragma unroll
  for (int j = 1; j < kMaxProgs; ++j)
    if (j < n && w >= start[j]) k = j;
  return k;
}

__device__ __forceinline__ double fold(double acc, const int32_t *__restrict__ src,
                                       int32_t beg, int32_t end,
                                       const double *__restrict__ slots) {
  int32_t j = beg;
  for (; j + 8 <= end; j += 8) {
    int32_t s[8];
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) s[u] = src[j + u];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = s[u] >= 0 ? slots[s[u]] : 0.0;
#pragma unroll
    for (int u = 0; u < 8; ++u) acc = acc + v[u];
  }
  for (; j < end; ++j) {
    const int32_t s = src[j];
    acc = acc + (s >= 0 ? slots[s] : 0.0);
  }
  return acc;
}

// Pass 1: one thread per chunk.
__global__ void __launch_bounds__(kBlock)
    chunk_kernel(ReduceBatch b, const double *__restrict__ slots) {
  const int w = blockIdx.x * kBlock + threadIdx.x;
  if (w >= b.cstart[b.n]) return;
  const int k = find_prog(b.cstart, b.n, w);
  const ReduceDev &p = b.prog[k];
  const int i = w - b.cstart[k];
  const int32_t beg = p.cptr[i], end = p.cptr[i + 1];
  const int32_t dst = p.cdst[i];
  if (dst >= 0) {
    // whole segment in this chunk: exact sequential fold into the output
    double *out = b.out[k];
    out[dst] = fold(out[dst], p.src, beg, end, slots);
  } else {
    // -0.0 is the additive identity, so the chunk sum starts exactly
    b.partial[p.pbase + i] = fold(-0.0, p.src, beg, end, slots);
  }
}

// Pass 2: one wavefront per multi-chunk segment.
__global__ void __launch_bounds__(kBlock)
    combine_kernel(ReduceBatch b) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int w = blockIdx.x * kWaves + wave;
  if (w >= b.mstart[b.n]) return;
  const int k = find_prog(b.mstart, b.n, w);
  const ReduceDev &p = b.prog[k];
  const int i = w - b.mstart[k];
  const int32_t t = p.mtgt[i];
  const int32_t c0 = p.mchunk[2 * i], c1 = p.mchunk[2 * i + 1];
  const int32_t len = c1 - c0;
  const int32_t run = (len + 63) / 64;
  const int32_t rb = c0 + lane * run;
  const int32_t re = min(c1, rb + run);
  double acc = -0.0;
  for (int32_t c = rb; c < re; ++c) acc = acc + b.partial[p.pbase + c];
  double *out = b.out[k];
  double r = out[t];
  for (int l = 0; l < 64; ++l) r = r + __shfl(acc, l, 64);
  if (lane == 0) out[t] = r;
}

// Per-instance flag words: bit 0 limited, bit 1 or bit 9 non-finite.
// Block b writes its two counts to out[b] and out[64 + b].
__global__ void __launch_bounds__(256)
    count_kernel(const int32_t *flags, int32_t n, int64_t *out) {
  __shared__ int64_t cnt[2][256];
  int64_t c0 = 0, c1 = 0;
  for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += gridDim.x * blockDim.x) {
    c0 += flags[i] & 1;
    c1 += (flags[i] & 0x202) != 0;
  }
  cnt[0][threadIdx.x] = c0;
  cnt[1][threadIdx.x] = c1;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      cnt[0][threadIdx.x] += cnt[0][threadIdx.x + s];
      cnt[1][threadIdx.x] += cnt[1][threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[blockIdx.x] = cnt[0][0];
    out[64 + blockIdx.x] = cnt[1][0];
  }
}

}  // namespace

hipError_t launch_reduce(int n_out, const ReduceDev *progs, double *const *outs,
                         const double *slots, double *partial, hipStream_t s) {
  if (n_out <= 0 || n_out > kMaxProgs) return hipErrorInvalidValue;
  ReduceBatch b{};
  b.n = n_out;
  b.partial = partial;
  for (int k = 0; k < n_out; ++k) {
    b.prog[k] = progs[k];
    b.out[k] = outs[k];
    if (!outs[k] && progs[k].n_chunk) return hipErrorInvalidValue;
    b.cstart[k + 1] = b.cstart[k] + progs[k].n_chunk;
    b.mstart[k + 1] = b.mstart[k] + progs[k].n_multi;
  }
  for (int k = n_out; k < kMaxProgs; ++k) {
    b.cstart[k + 1] = b.cstart[n_out];
    b.mstart[k + 1] = b.mstart[n_out];
  }
  const int n_chunks = b.cstart[n_out], n_multi = b.mstart[n_out];
  if (n_chunks == 0) return hipSuccess;
  hipLaunchKernelGGL(chunk_kernel, dim3((n_chunks + kBlock - 1) / kBlock),
                     dim3(kBlock), 0, s, b, slots);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || n_multi == 0) return e;
  hipLaunchKernelGGL(combine_kernel, dim3((n_multi + kWaves - 1) / kWaves),
                     dim3(kBlock), 0, s, b);
  return hipGetLastError();
}

// Writes two partial counts per block (64 blocks: limited in out[0..63],
// non-finite in out[64..127]); the caller sums them.
hipError_t launch_count_flags(const int32_t *flags, int32_t n, int64_t *out,
                              hipStream_t s) {
  hipLaunchKernelGGL(count_kernel, dim3(64), dim3(256), 0, s, flags, n, out);
  return hipGetLastError();
}

}  // namespace xg
