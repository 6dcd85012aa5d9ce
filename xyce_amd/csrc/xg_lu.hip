// Static-pivot sparse LU of each independent sub-circuit on gfx950, with
// the Newton update and its norms: the refactor-and-solve that replaces
// N_LAS_AmesosSolver::solve (KLU, N_LAS_AmesosSolver.C:231-470) for every
// Newton iteration after setup.
//
// One workgroup per component.  The component's filled factor pattern and
// its solution vector sit in LDS when they fit (else in a global-memory
// block of their own).  The task lists are level-scheduled: every task of a
// level is independent of the others, so the workgroup runs a level's tasks
// in parallel, one task per thread, and a barrier separates levels.  A task
// is a dot product of earlier values, summed by one thread in ascending
// order, so results do not depend on the launch or on other components.
//
// Each factorisation is checked against KLU's threshold pivoting rule
// (klu defaults: tol = 0.001): every multiplier |l_ik| <= 1/tol, every
// pivot nonzero and finite.  A component that fails is flagged; the dense
// partial-pivoting kernel launched right after refactors just those.
#include "xg_internal.h"

namespace xg {

namespace {

constexpr int kBlock = 256;
constexpr int kLuVal = 16384;  // 128 KB of factor values in LDS
constexpr int kLuVec = 2048;   //  16 KB of solution vector in LDS

__device__ __forceinline__ double blk_max(double v, double *sv) {
  const int t = threadIdx.x;
  sv[t] = v;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if (t < s) sv[t] = fmax(sv[t], sv[t + s]);
    __syncthreads();
  }
  const double r = sv[0];
  __syncthreads();
  return r;
}

__device__ __forceinline__ double blk_sum(double v, double *sv) {
  const int t = threadIdx.x;
  sv[t] = v;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if (t < s) sv[t] = sv[t] + sv[t + s];
    __syncthreads();
  }
  const double r = sv[0];
  __syncthreads();
  return r;
}

template <bool kLds>
__global__ void __launch_bounds__(kBlock) sparse_lu_kernel(SparseLuArgs a) {
  __shared__ double lval[kLds ? kLuVal : 1];
  __shared__ double lvec[kLds ? kLuVec : 1];
  __shared__ double sv[kBlock];
  const int c = blockIdx.x;
  const int t = threadIdx.x;
  const int beg = a.cstart[c], n = a.cstart[c + 1] - beg;
  const int64_t nv = a.voff[c + 1] - a.voff[c];
  if ((n <= kLuVec && nv <= kLuVal) != kLds) return;  // the other form
  double *stats = a.stats + 4 * (size_t)c;
  if (a.active && !a.active[c]) {
    if (t == 0) {
      stats[0] = 0.0;
      stats[1] = 0.0;
      stats[2] = 0.0;
      stats[3] = 0.0;
      a.fallback[c] = 0;
    }
    return;
  }
  if (!a.ok[c]) {  // structurally singular under matching: dense kernel
    if (t == 0) a.fallback[c] = 1;
    return;
  }
  double *val = kLds ? lval : a.val_work + a.voff[c];
  double *vec = kLds ? lvec : a.vec_work + beg;
  const int32_t *prow = a.prow + beg, *qcol = a.qcol + beg;
  for (int64_t e = t; e < nv; e += kBlock) val[e] = 0.0;
  __syncthreads();
  double rmax = 0.0;
  for (int i = t; i < n; i += kBlock) {
    const int g = prow[i];
    for (int k = a.rowptr[g]; k < a.rowptr[g + 1]; ++k) {
      const int p = a.lu_pos[k];
      if (p >= 0) val[p] = a.J[k];
    }
    const double r = a.rhs[g];
    vec[i] = r;
    rmax = fmax(rmax, fabs(r));
  }
  const double max_rhs = blk_max(rmax, sv);  // includes the barrier
  // ---- factor
  const int32_t *lv = a.lv_comp + 3 * (size_t)c;
  int bad = 0;
  for (int L = lv[0]; L < lv[1]; ++L) {
    for (int k = a.lv_ptr[L] + t; k < a.lv_ptr[L + 1]; k += kBlock) {
      const int d = a.t_dst[k], dv = a.t_div[k];
      double s = val[d];
      for (int o = a.t_op[k]; o < a.t_op[k + 1]; ++o) {
        const int2 q = a.ops[o];
        s = s - val[q.x] * val[q.y];
      }
      if (dv >= 0) {
        s = s / val[dv];
        bad |= !(fabs(s) <= a.inv_tau);  // also catches NaN
      } else if (dv == -2) {
        bad |= !(fabs(s) > 0.0) || isinf(s);
      }
      val[d] = s;
    }
    __syncthreads();
  }
  if (__syncthreads_or(bad)) {
    if (t == 0) a.fallback[c] = 1;
    return;
  }
  // ---- L y = P b (unit diagonal), then U x' = y, in place
  for (int L = lv[1]; L < lv[2]; ++L) {
    for (int k = a.lv_ptr[L] + t; k < a.lv_ptr[L + 1]; k += kBlock) {
      const int d = a.t_dst[k];
      double s = vec[d];
      for (int o = a.t_op[k]; o < a.t_op[k + 1]; ++o) {
        const int2 q = a.ops[o];
        s = s - val[q.x] * vec[q.y];
      }
      vec[d] = s;
    }
    __syncthreads();
  }
  for (int L = lv[2]; L < lv[3]; ++L) {
    for (int k = a.lv_ptr[L] + t; k < a.lv_ptr[L + 1]; k += kBlock) {
      const int d = a.t_dst[k];
      double s = vec[d];
      for (int o = a.t_op[k]; o < a.t_op[k + 1]; ++o) {
        const int2 q = a.ops[o];
        s = s - val[q.x] * vec[q.y];
      }
      vec[d] = s / val[a.t_div[k]];
    }
    __syncthreads();
  }
  // ---- Newton update x += dx and the weighted RMS norm of dx
  // (N_NLS_DampedNewton.C:311-335)
  double acc = 0.0;
  for (int i = t; i < n; i += kBlock) {
    const int g = qcol[i];
    const double dx = vec[i];
    const double w = a.reltol * fmax(fabs(a.wa[g]), fabs(a.wb[g])) + a.abstol;
    const double q = dx / w;
    acc = acc + q * q;
    a.x[g] = a.x[g] + dx;
  }
  const double ss = blk_sum(acc, sv);
  if (t == 0) {
    stats[0] = max_rhs;
    stats[1] = sqrt(ss / n);
    stats[2] = 0.0;
    stats[3] = (double)n;
    a.fallback[c] = 0;
  }
}

}  // namespace

bool sparse_lu_fits_lds(int64_t n, int64_t fnnz) {
  return n <= kLuVec && fnnz <= kLuVal;
}

hipError_t launch_sparse_lu(const SparseLuArgs &a, hipStream_t s) {
  if (a.n_comp == 0) return hipSuccess;
  hipLaunchKernelGGL(sparse_lu_kernel<true>, dim3(a.n_comp), dim3(kBlock), 0, s, a);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || !a.any_global) return e;
  hipLaunchKernelGGL(sparse_lu_kernel<false>, dim3(a.n_comp), dim3(kBlock), 0, s, a);
  return hipGetLastError();
}

}  // namespace xg
