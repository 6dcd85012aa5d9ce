// One Newton iteration's linear algebra on gfx950: assembly of the
// residual and Jacobian of the time-discretised DAE, and a dense LU solve
// per independent sub-circuit (connected component of the circuit graph
// without ground), with the Newton update and the convergence norms.
//
// Assembly follows N_TIA_OneStep::obtainResidual / obtainJacobian
// (N_TIA_OneStep.C:215-330), statement for statement, for the transient;
// the DC operating point uses the F part alone (RHS = -F + dFdxdVp,
// J = dF/dx).  The solve replaces N_LAS_AmesosSolver (KLU): LU with
// partial pivoting (the largest magnitude in the column, first on ties, as
// LAPACK getrf), one workgroup per component; components of at most
// kLdsN unknowns factor in LDS, larger ones in a global-memory workspace.
#include "xg_internal.h"

namespace xg {

namespace {

constexpr int kBlock = 256;
constexpr int kLdsN = 128;  // 128 x 128 doubles = 128 KB of the 160 KB LDS

__global__ void __launch_bounds__(kBlock)
    assemble_kernel(TiaAssembleArgs a) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  double inv_dt = a.inv_dt;
  if (a.dstep && a.transient) {  // step size and order from device memory
    inv_dt = a.dstep[2];
    a.order = (int32_t)a.dstep[3];
  }
  if (i < a.n1) {
    double r;
    if (a.transient) {
      // RHS = Q - qn0; order 2: RHS/h + F/2 + qHistory[2]/2, else RHS/h + F
      r = 1.0 * a.Q[i] + -1.0 * a.qn0[i];
      if (a.order == 2) {
        r = inv_dt * r + 0.5 * a.F[i];
        r = 1.0 * r + 0.5 * a.fprev[i];
      } else {
        r = inv_dt * r + 1.0 * a.F[i];
      }
      r = -r;
      if (a.limiter) {
        const double ql = a.QL[i] * inv_dt;  // scale(-alphas / h), alphas = -1
        r = r + 1.0 * ql;
        r = r + (a.order == 2 ? 0.5 : 1.0) * a.FL[i];
      }
    } else {
      r = -a.F[i];
      if (a.limiter) r = r + 1.0 * a.FL[i];
    }
    a.rhs[i] = i == a.n1 - 1 ? 0.0 : r;  // the ground slot carries nothing
  }
  if (i < a.nnz) {
    if (a.transient) {
      // J = (-alphas / h) dQ/dx + fscalar dF/dx
      a.J[i] = inv_dt * a.dQdx[i] + (a.order == 2 ? 0.5 : 1.0) * a.dFdx[i];
    } else {
      a.J[i] = a.dFdx[i];
    }
  }
}

// ONESTEP predictor (N_TIA_OneStep.C:115-133): xn0 = xHistory[0], then
// xn0 = beta_i xHistory[i] + 1.0 xn0 for i = 1..order, each product and sum
// rounded separately (the linearCombo sequence), so the result does not
// depend on how many circuits share the vector.
__global__ void __launch_bounds__(kBlock)
    predict_kernel(int64_t n1, int32_t order, double beta1, double beta2,
                   const double *xh0, const double *xh1, const double *xh2,
                   double *out) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n1) return;
  double v = xh0[i];
  v = beta1 * xh1[i] + 1.0 * v;
  if (order == 2) v = beta2 * xh2[i] + 1.0 * v;
  out[i] = v;
}

// Block-wide max of (|v|, index) with the smallest index on ties.
__device__ __forceinline__ void argmax_abs(double v, int idx, double *sv, int *si,
                                           double &best, int &bidx) {
  const int t = threadIdx.x;
  sv[t] = v;
  si[t] = idx;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if (t < s) {
      const double a = sv[t], b = sv[t + s];
      const int ia = si[t], ib = si[t + s];
      if (b > a || (b == a && ib < ia)) {
        sv[t] = b;
        si[t] = ib;
      }
    }
    __syncthreads();
  }
  best = sv[0];
  bidx = si[0];
  __syncthreads();
}

__device__ __forceinline__ double block_sum(double v, double *sv) {
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

__device__ __forceinline__ double block_max(double v, double *sv) {
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

// One component per workgroup.  A is n x n row-major (LDS or workspace).
template <bool kLds>
__global__ void __launch_bounds__(kBlock)
    dense_solve_kernel(DenseSolveArgs a) {
  __shared__ double lds_a[kLds ? kLdsN * kLdsN : 1];
  __shared__ double sv[kBlock];
  __shared__ int si[kBlock];
  __shared__ double sb[kLds ? kLdsN : 1];
  const int c = blockIdx.x;
  const int t = threadIdx.x;
  const int beg = a.cstart[c], n = a.cstart[c + 1] - beg;
  double *stats = a.stats + 4 * (size_t)c;
  if ((kLds && n > kLdsN) || (!kLds && n <= kLdsN)) return;  // other variant
  if (a.active && !a.active[c]) {
    if (t == 0 && !a.only_flagged) {
      stats[0] = 0.0;
      stats[1] = 0.0;
      stats[2] = 0.0;
      stats[3] = 0.0;
    }
    return;
  }
  if (!kLds && a.aoff[c] < 0 && !a.norm_only) {
    // no dense form for a component this large
    if (t == 0) {
      stats[0] = 0.0;
      stats[1] = 0.0;
      stats[2] = 1.0;
      stats[3] = (double)n;
    }
    return;
  }
  double *A = kLds ? lds_a : a.work + (a.aoff[c] < 0 ? 0 : a.aoff[c]);
  double *b = kLds ? sb : a.bwork + beg;
  const int32_t *nodes = a.nodes + beg;
  // gather the component's rows of J (and its RHS) into dense form
  if (!a.norm_only)
    for (int64_t e = t; e < (int64_t)n * n; e += kBlock) A[e] = 0.0;
  __syncthreads();
  double rmax = 0.0;
  for (int r = t; r < n; r += kBlock) {
    const int g = nodes[r];
    if (!a.norm_only) {
      for (int k = a.rowptr[g]; k < a.rowptr[g + 1]; ++k) {
        const int col = a.local[a.colidx[k]];
        if (col >= 0) A[(int64_t)r * n + col] = a.J[k];
      }
      b[r] = a.rhs[g];
    }
    rmax = fmax(rmax, fabs(a.rhs[g]));
  }
  const double max_rhs = block_max(rmax, sv);
  if (a.norm_only) {
    if (t == 0) stats[0] = max_rhs;
    return;
  }
  // LU with partial pivoting, row swaps applied to b as they happen
  int singular = 0;
  for (int k = 0; k < n; ++k) {
    double best = -1.0;
    int bidx = k;
    for (int i = k + t; i < n; i += kBlock) {
      const double v = fabs(A[(int64_t)i * n + k]);
      if (v > best) {
        best = v;
        bidx = i;
      }
    }
    argmax_abs(best, bidx, sv, si, best, bidx);
    if (!(best > 0.0)) {
      singular = 1;
      break;
    }
    if (bidx != k) {
      for (int j = t; j < n; j += kBlock) {
        const double tmp = A[(int64_t)k * n + j];
        A[(int64_t)k * n + j] = A[(int64_t)bidx * n + j];
        A[(int64_t)bidx * n + j] = tmp;
      }
      if (t == 0) {
        const double tmp = b[k];
        b[k] = b[bidx];
        b[bidx] = tmp;
      }
    }
    __syncthreads();
    const double piv = A[(int64_t)k * n + k];
    for (int i = k + 1 + t; i < n; i += kBlock) A[(int64_t)i * n + k] /= piv;
    __syncthreads();
    const int m = n - k - 1;
    for (int64_t e = t; e < (int64_t)m * m; e += kBlock) {
      const int i = k + 1 + (int)(e / m), j = k + 1 + (int)(e % m);
      A[(int64_t)i * n + j] -= A[(int64_t)i * n + k] * A[(int64_t)k * n + j];
    }
    __syncthreads();
  }
  if (singular) {
    if (t == 0) {
      stats[0] = max_rhs;
      stats[1] = 0.0;
      stats[2] = 1.0;
      stats[3] = 0.0;
    }
    return;
  }
  // L y = P b (unit lower), then U x = y; one row at a time, the inner
  // product spread over the workgroup
  for (int r = 0; r < n; ++r) {
    double s = 0.0;
    for (int j = t; j < r; j += kBlock) s = s + A[(int64_t)r * n + j] * b[j];
    s = block_sum(s, sv);
    if (t == 0) b[r] = b[r] - s;
    __syncthreads();
  }
  for (int r = n - 1; r >= 0; --r) {
    double s = 0.0;
    for (int j = r + 1 + t; j < n; j += kBlock) s = s + A[(int64_t)r * n + j] * b[j];
    s = block_sum(s, sv);
    if (t == 0) b[r] = (b[r] - s) / A[(int64_t)r * n + r];
    __syncthreads();
  }
  // Newton update x += dx and the weighted RMS norm of dx
  // (N_NLS_DampedNewton.C:311-335 weights, wRMSNorm)
  double acc = 0.0;
  for (int r = t; r < n; r += kBlock) {
    const int g = nodes[r];
    const double dx = b[r];
    const double w = a.reltol * fmax(fabs(a.wa[g]), fabs(a.wb[g])) + a.abstol;
    const double q = dx / w;
    acc = acc + q * q;
    a.x[g] = a.x[g] + dx;
  }
  const double ss = block_sum(acc, sv);
  if (t == 0) {
    stats[0] = max_rhs;
    stats[1] = sqrt(ss / n);
    stats[2] = 0.0;
    stats[3] = (double)n;
  }
}

}  // namespace

hipError_t launch_tia_assemble(const TiaAssembleArgs &a, hipStream_t s) {
  const int64_t items = a.n1 > a.nnz ? a.n1 : a.nnz;
  if (items == 0) return hipSuccess;
  hipLaunchKernelGGL(assemble_kernel, dim3((unsigned)((items + kBlock - 1) / kBlock)),
                     dim3(kBlock), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_dense_solve(const DenseSolveArgs &a, hipStream_t s) {
  if (a.n_comp == 0) return hipSuccess;
  hipLaunchKernelGGL(dense_solve_kernel<true>, dim3(a.n_comp), dim3(kBlock), 0,
                     s, a);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || !a.any_large) return e;
  hipLaunchKernelGGL(dense_solve_kernel<false>, dim3(a.n_comp), dim3(kBlock), 0,
                     s, a);
  return hipGetLastError();
}

int dense_lds_max() { return kLdsN; }

hipError_t launch_predict(int64_t n1, int32_t order, double beta1, double beta2,
                          const double *xh0, const double *xh1, const double *xh2,
                          double *out, hipStream_t s) {
  if (n1 == 0) return hipSuccess;
  hipLaunchKernelGGL(predict_kernel, dim3((unsigned)((n1 + kBlock - 1) / kBlock)),
                     dim3(kBlock), 0, s, n1, order, beta1, beta2, xh0, xh1, xh2, out);
  return hipGetLastError();
}

}  // namespace xg
