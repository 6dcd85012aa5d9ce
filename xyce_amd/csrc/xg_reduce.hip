// Ordered segmented reduction: folds the per-statement stamp values of the
// slot buffer into output vectors / CSR values without atomics, in a fixed
// order, so results are bitwise reproducible run to run.
//
// Every output entry k owns a segment of sources (slot indices) listed in
// the reference's accumulation order (device type, instance, statement).
//
//  * A segment of at most kShort sources is a fixed-width record of four
//    (one 16-byte load) or eight sources (two); one thread loads the
//    record, gathers its values and folds
//      acc = out[k]; acc = acc + s0; acc = acc + s1; ...; out[k] = acc
//    which is exactly the sequence of "*ptr += v" / "vec[li] += v"
//    statements the reference executes (N_DEV_MOSFET1.C:4441-4722,
//    N_DEV_MOSFET_B3.C:9654-9928); padding adds -0.0, the exact identity.
//  * A segment of at most kChunk sources is folded the same way by one
//    thread from the contiguous source list.
//  * A longer segment (a supply rail or a heavily shared node) is folded in
//    two passes: each chunk is summed in order into a partial, then one
//    wavefront adds the partials to out[k] in chunk order (lane runs in
//    lane order).  The order is fixed; the result differs from the strictly
//    sequential sum by rounding only.
//
// In the overwrite mode the fold starts from +0.0 instead of out[k] (the
// same bits as zeroing first) and entries no source reaches are set to
// +0.0, so the caller needs no zeroing pass.
#include "xg_internal.h"

namespace xg {

namespace {

constexpr int kMaxProgs = 6;
constexpr int kBlock = 256;
constexpr int kWaves = kBlock / 64;

struct ReduceBatch {
  int n;
  int overwrite;
  // work-item prefixes over the programs: short records, chunks, zero
  // entries, multi-chunk segments
  int32_t s4start[kMaxProgs + 1];
  int32_t sstart[kMaxProgs + 1];
  int32_t cstart[kMaxProgs + 1];
  int32_t zstart[kMaxProgs + 1];
  int32_t mstart[kMaxProgs + 1];
  ReduceDev prog[kMaxProgs];
  double *out[kMaxProgs];
  double *partial;  // scratch for chunk partial sums
  const int32_t *gate;  // stop word: nonzero skips the launch
};

__device__ __forceinline__ int find_prog(const int32_t *start, int n, int w) {
  int k = 0;
#pragma unroll
  for (int j = 1; j < kMaxProgs; ++j)
    if (j < n && w >= start[j]) k = j;
  return k;
}

// Value of one source index: a slot, "+ 0.0" (-1) or the padding identity.
__device__ __forceinline__ double source_value(int32_t s,
                                               const double *__restrict__ slots) {
  return s >= 0 ? slots[s] : (s == -1 ? 0.0 : -0.0);
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
    for (int u = 0; u < 8; ++u) v[u] = source_value(s[u], slots);
#pragma unroll
    for (int u = 0; u < 8; ++u) acc = acc + v[u];
  }
  for (; j < end; ++j) acc = acc + source_value(src[j], slots);
  return acc;
}

// One launch for every short record, every chunk and (overwrite mode) every
// zero entry of all programs, in that order of thread ids, so each wave
// does one kind of work.
__global__ void __launch_bounds__(kBlock)
    fold_kernel(ReduceBatch b, const double *__restrict__ slots) {
  if (b.gate && *b.gate) return;
  int w = blockIdx.x * kBlock + threadIdx.x;
  const int n_short4 = b.s4start[b.n];
  const int n_short = b.sstart[b.n], n_chunk = b.cstart[b.n];
  if (w < n_short4) {
    const int k = find_prog(b.s4start, b.n, w);
    const ReduceDev &p = b.prog[k];
    const int i = w - b.s4start[k];
    const int4 r0 = reinterpret_cast<const int4 *>(p.ssrc4)[i];
    const int32_t dst = p.sdst4[i];
    double *out = b.out[k];
    const double v0 = source_value(r0.x, slots), v1 = source_value(r0.y, slots),
                 v2 = source_value(r0.z, slots), v3 = source_value(r0.w, slots);
    double acc = b.overwrite ? 0.0 : out[dst];
    acc = acc + v0;
    acc = acc + v1;
    acc = acc + v2;
    acc = acc + v3;
    out[dst] = acc;
    return;
  }
  w -= n_short4;
  if (w < n_short) {
    const int k = find_prog(b.sstart, b.n, w);
    const ReduceDev &p = b.prog[k];
    const int i = w - b.sstart[k];
    const int4 *rec = reinterpret_cast<const int4 *>(p.ssrc) + 2 * (size_t)i;
    const int4 r0 = rec[0], r1 = rec[1];
    const int32_t dst = p.sdst[i];
    double *out = b.out[k];
    const double v0 = source_value(r0.x, slots), v1 = source_value(r0.y, slots),
                 v2 = source_value(r0.z, slots), v3 = source_value(r0.w, slots),
                 v4 = source_value(r1.x, slots), v5 = source_value(r1.y, slots),
                 v6 = source_value(r1.z, slots), v7 = source_value(r1.w, slots);
    double acc = b.overwrite ? 0.0 : out[dst];
    acc = acc + v0;
    acc = acc + v1;
    acc = acc + v2;
    acc = acc + v3;
    acc = acc + v4;
    acc = acc + v5;
    acc = acc + v6;
    acc = acc + v7;
    out[dst] = acc;
    return;
  }
  w -= n_short;
  if (w < n_chunk) {
    const int k = find_prog(b.cstart, b.n, w);
    const ReduceDev &p = b.prog[k];
    const int i = w - b.cstart[k];
    const int32_t beg = p.cptr[i], end = p.cptr[i + 1];
    const int32_t dst = p.cdst[i];
    if (dst >= 0) {
      // whole segment in this chunk: sequential fold into the output
      double *out = b.out[k];
      out[dst] = fold(b.overwrite ? 0.0 : out[dst], p.src, beg, end, slots);
    } else {
      // -0.0 is the additive identity, so the chunk sum starts exactly
      b.partial[p.pbase + i] = fold(-0.0, p.src, beg, end, slots);
    }
    return;
  }
  w -= n_chunk;
  if (b.overwrite && w < b.zstart[b.n]) {
    const int k = find_prog(b.zstart, b.n, w);
    b.out[k][b.prog[k].zidx[w - b.zstart[k]]] = 0.0;
  }
}

// Second pass: one wavefront per multi-chunk segment.
__global__ void __launch_bounds__(kBlock)
    combine_kernel(ReduceBatch b) {
  if (b.gate && *b.gate) return;
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
  double r = b.overwrite ? 0.0 : out[t];
  // lane runs in lane order; each lane's run sum read with readlane (the
  // reads do not wait on the additions, so only the adds are serial)
  const long long bits = __double_as_longlong(acc);
  const int lo = (int)bits, hi = (int)(bits >> 32);
#pragma unroll
  for (int l = 0; l < 64; ++l) {
    const unsigned a = (unsigned)__builtin_amdgcn_readlane(lo, l);
    const long long h = (long long)__builtin_amdgcn_readlane(hi, l);
    r = r + __longlong_as_double((long long)((unsigned long long)h << 32 | a));
  }
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

// count_kernel and comp_count_kernel in one pass (the native Newton loop;
// cnt zeroed beforehand).  Grid: 64 blocks.
__global__ void __launch_bounds__(256)
    iteration_count_kernel(const int32_t *flags, const int32_t *slot_comp, int32_t n,
                           int32_t *cnt, int64_t *out, const int32_t *gate) {
  __shared__ int64_t c2[2][256];
  if (gate && *gate) return;
  int64_t c0 = 0, c1 = 0;
  for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += gridDim.x * blockDim.x) {
    const int32_t f = flags[i];
    c0 += f & 1;
    c1 += (f & 0x202) != 0;
    if (f & 1) {
      const int32_t c = slot_comp[i];
      if (c >= 0) atomicAdd(cnt + c, 1);
    }
  }
  c2[0][threadIdx.x] = c0;
  c2[1][threadIdx.x] = c1;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      c2[0][threadIdx.x] += c2[0][threadIdx.x + s];
      c2[1][threadIdx.x] += c2[1][threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[blockIdx.x] = c2[0][0];
    out[64 + blockIdx.x] = c2[1][0];
  }
}

__global__ void zero_i32_kernel(int32_t *p, int32_t n) {
  const int32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) p[i] = 0;
}

// Limited instances per component (a count, so the order of the atomic
// increments does not matter).
__global__ void comp_count_kernel(const int32_t *flags, const int32_t *slot_comp,
                                  int32_t n, int32_t *cnt) {
  const int32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const int32_t c = slot_comp[i];
  if (c >= 0 && (flags[i] & 1)) atomicAdd(cnt + c, 1);
}

}  // namespace

hipError_t launch_component_limited(const int32_t *flags, const int32_t *slot_comp,
                                    int32_t n, int32_t *cnt, int32_t n_comp,
                                    hipStream_t s) {
  if (n_comp > 0)
    hipLaunchKernelGGL(zero_i32_kernel, dim3((n_comp + 255) / 256), dim3(256), 0,
                       s, cnt, n_comp);
  if (n > 0)
    hipLaunchKernelGGL(comp_count_kernel, dim3((n + 255) / 256), dim3(256), 0, s,
                       flags, slot_comp, n, cnt);
  return hipGetLastError();
}

hipError_t launch_reduce(int n_out, const ReduceDev *progs, double *const *outs,
                         const double *slots, double *partial, bool overwrite,
                         hipStream_t s, const int32_t *gate) {
  if (n_out <= 0 || n_out > kMaxProgs) return hipErrorInvalidValue;
  ReduceBatch b{};
  b.n = n_out;
  b.gate = gate;
  b.overwrite = overwrite ? 1 : 0;
  b.partial = partial;
  for (int k = 0; k < n_out; ++k) {
    b.prog[k] = progs[k];
    b.out[k] = outs[k];
    const int32_t items = progs[k].n_short4 + progs[k].n_short + progs[k].n_chunk +
                          (overwrite ? progs[k].n_zero : 0);
    if (!outs[k] && items) return hipErrorInvalidValue;
    b.s4start[k + 1] = b.s4start[k] + progs[k].n_short4;
    b.sstart[k + 1] = b.sstart[k] + progs[k].n_short;
    b.cstart[k + 1] = b.cstart[k] + progs[k].n_chunk;
    b.zstart[k + 1] = b.zstart[k] + progs[k].n_zero;
    b.mstart[k + 1] = b.mstart[k] + progs[k].n_multi;
  }
  for (int k = n_out; k < kMaxProgs; ++k) {
    b.s4start[k + 1] = b.s4start[n_out];
    b.sstart[k + 1] = b.sstart[n_out];
    b.cstart[k + 1] = b.cstart[n_out];
    b.zstart[k + 1] = b.zstart[n_out];
    b.mstart[k + 1] = b.mstart[n_out];
  }
  const int64_t items = (int64_t)b.s4start[n_out] + b.sstart[n_out] + b.cstart[n_out] +
                        (overwrite ? b.zstart[n_out] : 0);
  const int n_multi = b.mstart[n_out];
  if (items == 0) return hipSuccess;
  hipLaunchKernelGGL(fold_kernel, dim3((unsigned)((items + kBlock - 1) / kBlock)),
                     dim3(kBlock), 0, s, b, slots);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || n_multi == 0) return e;
  hipLaunchKernelGGL(combine_kernel, dim3((n_multi + kWaves - 1) / kWaves),
                     dim3(kBlock), 0, s, b);
  return hipGetLastError();
}

// Writes two partial counts per block (64 blocks: limited in out[0..63],
// non-finite in out[64..127]); the caller sums them.
hipError_t launch_iteration_counts(const int32_t *flags, const int32_t *slot_comp,
                                   int32_t n, int32_t *comp_cnt, int64_t *out,
                                   hipStream_t s, const int32_t *gate) {
  hipLaunchKernelGGL(iteration_count_kernel, dim3(64), dim3(256), 0, s, flags, slot_comp,
                     n, comp_cnt, out, gate);
  return hipGetLastError();
}

hipError_t launch_count_flags(const int32_t *flags, int32_t n, int64_t *out,
                              hipStream_t s) {
  hipLaunchKernelGGL(count_kernel, dim3(64), dim3(256), 0, s, flags, n, out);
  return hipGetLastError();
}

}  // namespace xg
