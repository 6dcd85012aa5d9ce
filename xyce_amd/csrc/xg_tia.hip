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
// kLdsN unknowns factor in LDS, larger ones in a global-memory workspace,
// both as a blocked elimination (column panels factored in LDS, rank-16
// updates below them) and blocked triangular solves.
#include "xg_internal.h"

namespace xg {

namespace {

constexpr int kBlock = 256;
// The dense LU's workgroup (one per component and CU for its LDS): four
// waves, each with the whole register file of its SIMD (the unrolled panel
// solve keeps a column's sixteen U values and the panel rows in registers;
// more waves per SIMD would spill them).
constexpr int kDenseBlock = 256;
constexpr int kLdsN = 128;  // 128 x 128 doubles = 128 KB of the 160 KB LDS

__global__ void __launch_bounds__(kBlock)
    assemble_kernel(TiaAssembleArgs a) {
  if (a.gate && *a.gate) return;
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  double inv_dt = a.inv_dt;
  if (a.dstep && a.transient) {  // step size and order from device memory
    inv_dt = a.dstep[2];
    a.order = (int32_t)a.dstep[3];
  }
  double nrm = 0.0;     // |rhs| of this unknown (NaN: 0, as fmax skips it)
  int32_t comp = -1;    // its component (-1: none, ground, out of range)
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
    if (a.norms && i < a.n1 - 1) {
      comp = a.node_comp[i];
      nrm = r == r ? fabs(r) : 0.0;
    }
  }
  if (a.norms) {
    // per-component max |rhs|: non-negative doubles order as their bit
    // patterns, so an integer atomic max is exact.  A wave whose unknowns
    // all lie in one component (components are contiguous runs of nodes)
    // folds its max first and makes one atomic.
    int32_t lo = comp >= 0 ? comp : 0x7fffffff, hi = comp;
    double wm = nrm;
    for (int o = 32; o > 0; o >>= 1) {
      lo = min(lo, __shfl_xor(lo, o, 64));
      hi = max(hi, __shfl_xor(hi, o, 64));
      wm = fmax(wm, __shfl_xor(wm, o, 64));
    }
    unsigned long long *dst = nullptr;
    double v = 0.0;
    if (lo == hi) {  // every unknown of the wave in one component
      if ((threadIdx.x & 63) == 0) {
        dst = reinterpret_cast<unsigned long long *>(a.norms + 4 * (size_t)lo);
        v = wm;
      }
    } else if (comp >= 0) {
      dst = reinterpret_cast<unsigned long long *>(a.norms + 4 * (size_t)comp);
      v = nrm;
    }
    if (dst && v > 0.0) atomicMax(dst, (unsigned long long)__double_as_longlong(v));
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

// Sum over the 64 lanes of a wave (every lane gets the same value).
__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v = v + __shfl_xor(v, o, 64);
  return v;
}

// Block-wide max of (|v|, index) with the smallest index on ties; callers
// synchronise before the next call (sv / si are reused).
__device__ __forceinline__ void block_argmax(double v, int idx, double *sv, int *si,
                                             double &best, int &bidx) {
  for (int o = 32; o > 0; o >>= 1) {
    const double ov = __shfl_xor(v, o, 64);
    const int oi = __shfl_xor(idx, o, 64);
    if (ov > v || (ov == v && oi < idx)) {
      v = ov;
      idx = oi;
    }
  }
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    sv[w] = v;
    si[w] = idx;
  }
  __syncthreads();
  best = sv[0];
  bidx = si[0];
  for (int k = 1; k < kDenseBlock / 64; ++k)
    if (sv[k] > best || (sv[k] == best && si[k] < bidx)) {
      best = sv[k];
      bidx = si[k];
    }
}

__device__ __forceinline__ double block_sum(double v, double *sv) {
  v = wave_sum(v);
  if ((threadIdx.x & 63) == 0) sv[threadIdx.x >> 6] = v;
  __syncthreads();
  double r = sv[0];
  for (int k = 1; k < kDenseBlock / 64; ++k) r = r + sv[k];
  __syncthreads();
  return r;
}

__device__ __forceinline__ double block_max(double v, double *sv) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  if ((threadIdx.x & 63) == 0) sv[threadIdx.x >> 6] = v;
  __syncthreads();
  double r = sv[0];
  for (int k = 1; k < kDenseBlock / 64; ++k) r = fmax(r, sv[k]);
  __syncthreads();
  return r;
}

// Blocked right-looking LU with partial pivoting of the n x n row-major A
// (LDS or global), row swaps applied to b as they happen.  Panels of NB
// columns factor in the LDS buffer P (rows kb..n-1, stride NB + 1); the
// rows below the
// panel then take the panel's NB rank-1 updates in one pass per element,
// in the order of the unblocked elimination (each element's updates in
// ascending pivot order, so the factors are those of the column-by-column
// algorithm).  Returns false when a pivot column has no nonzero magnitude.
template <int NB>
__device__ __forceinline__ bool dense_factor(double *A, int n, double *b, double *P, double *sv,
                             int *si, int *spiv) {
  const int t = threadIdx.x;
  // panel row stride NB + 1 doubles: the lanes of a column access (pivot
  // search, scaling) fall in different LDS banks
  constexpr int PS = NB + 1;
  for (int kb = 0; kb < n; kb += NB) {
    const int nb = min(NB, n - kb);
    const int m = n - kb;
    for (int e = t; e < m * NB; e += kDenseBlock) {
      const int r = e / NB, c = e % NB;
      P[r * PS + c] = c < nb ? A[(int64_t)(kb + r) * n + kb + c] : 0.0;
    }
    __syncthreads();
    for (int c = 0; c < nb; ++c) {
      double v = -1.0;
      int idx = 0x7fffffff;
      for (int r = c + t; r < m; r += kDenseBlock) {
        const double a = fabs(P[r * PS + c]);
        if (a > v) {
          v = a;
          idx = r;
        }
      }
      double best;
      int br;
      block_argmax(v, idx, sv, si, best, br);
      if (!(best > 0.0)) return false;  // the same on every thread
      if (t == 0) spiv[c] = kb + br;
      if (br != c) {
        for (int cc = t; cc < NB; cc += kDenseBlock) {
          const double tmp = P[c * PS + cc];
          P[c * PS + cc] = P[br * PS + cc];
          P[br * PS + cc] = tmp;
        }
        if (t == 0) {
          const double tmp = b[kb + c];
          b[kb + c] = b[kb + br];
          b[kb + br] = tmp;
        }
      }
      __syncthreads();
      const double piv = P[c * PS + c];
      for (int r = c + 1 + t; r < m; r += kDenseBlock) {
        const double l = P[r * PS + c] / piv;
        P[r * PS + c] = l;
        for (int cc = c + 1; cc < nb; ++cc) P[r * PS + cc] -= l * P[c * PS + cc];
      }
      __syncthreads();
    }
    // the factored panel back into A; the panel's row swaps on the other
    // columns (left of the panel too, so L ends in the final row order)
    for (int e = t; e < m * NB; e += kDenseBlock) {
      const int r = e / NB, c = e % NB;
      if (c < nb) A[(int64_t)(kb + r) * n + kb + c] = P[r * PS + c];
    }
    for (int j = t; j < n; j += kDenseBlock) {
      if (j >= kb && j < kb + nb) continue;
      for (int c = 0; c < nb; ++c) {
        const int pr = spiv[c];
        if (pr != kb + c) {
          const double tmp = A[(int64_t)(kb + c) * n + j];
          A[(int64_t)(kb + c) * n + j] = A[(int64_t)pr * n + j];
          A[(int64_t)pr * n + j] = tmp;
        }
      }
    }
    __syncthreads();
    // U12 (rows kb..kb+NB-1 right of the panel) and the trailing update;
    // a column is solved in registers, then swept down its rows.  Columns
    // remain only when the panel is full (nb == NB).
    const int j0 = kb + nb, ncols = n - j0;
    // solve column j's U12 part into u (rows kb..kb+NB-1)
    auto solve_u = [&](int j, double *u) {
#pragma unroll
      for (int c = 0; c < NB; ++c) u[c] = A[(int64_t)(kb + c) * n + j];
#pragma unroll
      for (int c = 0; c < NB; ++c)
#pragma unroll
        for (int i = c + 1; i < NB; ++i) u[i] -= P[i * PS + c] * u[c];
    };
    // rows j0+g, j0+g+G, ... of column j take the panel's updates, four
    // rows' loads issued before their stores
    auto sweep = [&](int j, int g, int G, const double *u) {
      for (int i = j0 + g; i < n; i += 4 * G) {
        double x[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int ii = i + q * G;
          x[q] = ii < n ? A[(int64_t)ii * n + j] : 0.0;
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int ii = i + q * G;
          if (ii < n) {
            const double *l = P + (ii - kb) * PS;
            double v = x[q];
#pragma unroll
            for (int c = 0; c < NB; ++c) v -= l[c] * u[c];
            A[(int64_t)ii * n + j] = v;
          }
        }
      }
    };
    if (ncols >= kDenseBlock) {
      // a column per thread
      for (int j = j0 + t; j < n; j += kDenseBlock) {
        double u[NB];
        solve_u(j, u);
#pragma unroll
        for (int c = 0; c < NB; ++c) A[(int64_t)(kb + c) * n + j] = u[c];
        sweep(j, 0, 1, u);
      }
    } else if (ncols > 0) {
      // G threads per column, each sweeping every G-th row; all of them
      // read the column's U12 rows before the first one writes them
      const int G = min(kDenseBlock / ncols, 16);
      const bool mine = t < ncols * G;
      const int j = j0 + t % ncols, g = t / ncols;
      double u[NB];
      if (mine) solve_u(j, u);
      __syncthreads();
      if (mine) {
        if (g == 0)
#pragma unroll
          for (int c = 0; c < NB; ++c) A[(int64_t)(kb + c) * n + j] = u[c];
        sweep(j, g, G, u);
      }
    }
    __syncthreads();
  }
  return true;
}

// L y = b (unit lower) then U x = y on the factored A, b in LDS, blocks of
// 64 rows: the block's products with the solved part are wave dot
// products, the 64 x 64 diagonal block is solved by one wave.  D is an
// LDS buffer of 64 x 65 doubles for the diagonal block (rows padded by one
// so a wave's column reads fall in different banks; unused when A is
// itself in LDS).
template <bool kLds>
__device__ __forceinline__ void dense_solve_tri(const double *A, int n, double *b, double *D,
                                double *ss) {
  const int t = threadIdx.x, w = t >> 6, lane = t & 63;
  constexpr int kWaves = kDenseBlock / 64;
  auto diag = [&](int r0, int l, int i) -> double {
    return kLds ? A[(int64_t)(r0 + l) * n + r0 + i] : D[l * 65 + i];
  };
  for (int r0 = 0; r0 < n; r0 += 64) {
    const int R = min(64, n - r0);
    for (int rr = w; rr < R; rr += kWaves) {
      const double *row = A + (int64_t)(r0 + rr) * n;
      double s = 0.0;
      for (int j = lane; j < r0; j += 64) s = s + row[j] * b[j];
      s = wave_sum(s);
      if (lane == 0) ss[rr] = s;
    }
    if (!kLds)
      for (int e = t; e < R * 64; e += kDenseBlock) {
        const int l = e >> 6, i = e & 63;
        if (i < R) D[l * 65 + i] = A[(int64_t)(r0 + l) * n + r0 + i];
      }
    __syncthreads();
    if (w == 0) {
      double bl = lane < R ? b[r0 + lane] - ss[lane] : 0.0;
      for (int i = 0; i < R; ++i) {
        const double yi = __shfl(bl, i, 64);
        if (lane > i && lane < R) bl -= diag(r0, lane, i) * yi;
      }
      if (lane < R) b[r0 + lane] = bl;
    }
    __syncthreads();
  }
  for (int r0 = ((n - 1) / 64) * 64; r0 >= 0; r0 -= 64) {
    const int R = min(64, n - r0);
    for (int rr = w; rr < R; rr += kWaves) {
      const double *row = A + (int64_t)(r0 + rr) * n;
      double s = 0.0;
      for (int j = r0 + R + lane; j < n; j += 64) s = s + row[j] * b[j];
      s = wave_sum(s);
      if (lane == 0) ss[rr] = s;
    }
    if (!kLds)
      for (int e = t; e < R * 64; e += kDenseBlock) {
        const int l = e >> 6, i = e & 63;
        if (i < R) D[l * 65 + i] = A[(int64_t)(r0 + l) * n + r0 + i];
      }
    __syncthreads();
    if (w == 0) {
      double bl = lane < R ? b[r0 + lane] - ss[lane] : 0.0;
      for (int i = R - 1; i >= 0; --i) {
        if (lane == i) bl = bl / diag(r0, i, i);
        const double xi = __shfl(bl, i, 64);
        if (lane < i) bl -= diag(r0, lane, i) * xi;
      }
      if (lane < R) b[r0 + lane] = bl;
    }
    __syncthreads();
  }
}

// Panel buffer of the global-memory form: n x 17 doubles (16-column panels)
// up to n = 1024, n x 9 up to 1934, n x 5 beyond (kDenseMax = 2048).
constexpr int kPanelDoubles = 17408;

// One component per workgroup.  A is n x n row-major (LDS or workspace).
template <bool kLds>
__global__ void __launch_bounds__(kDenseBlock)
    dense_solve_kernel(DenseSolveArgs a) {
  __shared__ double lds_a[kLds ? kLdsN * kLdsN : 1];
  __shared__ double panel[kLds ? kLdsN * 17 : kPanelDoubles];
  __shared__ double sb[kLds ? kLdsN : kDenseMax];
  __shared__ double ss[64];
  __shared__ double sv[kDenseBlock / 64];
  __shared__ int si[kDenseBlock / 64];
  __shared__ int spiv[16];
  const int c = blockIdx.x;
  const int t = threadIdx.x;
  if (a.gate && *a.gate) return;
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
  double *b = sb;
  const int32_t *nodes = a.nodes + beg;
  // gather the component's rows of J (and its RHS) into dense form
  if (!a.norm_only)
    for (int64_t e = t; e < (int64_t)n * n; e += kDenseBlock) A[e] = 0.0;
  __syncthreads();
  double rmax = 0.0;
  for (int r = t; r < n; r += kDenseBlock) {
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
  const bool ok = (kLds || n * 17 <= kPanelDoubles)
                      ? dense_factor<16>(A, n, b, panel, sv, si, spiv)
                      : n * 9 <= kPanelDoubles ? dense_factor<8>(A, n, b, panel, sv, si, spiv)
                                               : dense_factor<4>(A, n, b, panel, sv, si, spiv);
  if (!ok) {
    if (t == 0) {
      stats[0] = max_rhs;
      stats[1] = 0.0;
      stats[2] = 1.0;
      stats[3] = 0.0;
    }
    return;
  }
  dense_solve_tri<kLds>(A, n, b, panel, ss);
  // Newton update x += dx and the weighted RMS norm of dx
  // (N_NLS_DampedNewton.C:311-335 weights, wRMSNorm)
  double acc = 0.0;
  for (int r = t; r < n; r += kDenseBlock) {
    const int g = nodes[r];
    const double dx = b[r];
    const double w = a.reltol * fmax(fabs(a.wa[g]), fabs(a.wb[g])) + a.abstol;
    const double q = dx / w;
    acc = acc + q * q;
    a.x[g] = a.x[g] + dx;
  }
  const double ss2 = block_sum(acc, sv);
  if (t == 0) {
    stats[0] = max_rhs;
    stats[1] = sqrt(ss2 / n);
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
  hipLaunchKernelGGL(dense_solve_kernel<true>, dim3(a.n_comp), dim3(kDenseBlock), 0,
                     s, a);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || !a.any_large) return e;
  hipLaunchKernelGGL(dense_solve_kernel<false>, dim3(a.n_comp), dim3(kDenseBlock), 0,
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
