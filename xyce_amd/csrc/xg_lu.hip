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
// Static pivots are accepted when the solve is backward stable: every
// pivot nonzero and every factor finite, and the normwise backward error of
// the solution, ||J dx - rhs|| / (||J|| ||dx|| + ||rhs||) (infinity norms,
// the residual taken with the unfactored values), at most 1e-12 -- the
// bound partial pivoting gives in practice (static pivoting with a
// backward-error check, as in Li and Demmel's GESP for SuperLU_DIST).  A
// component that fails is flagged, its iterate left as it was, and the
// dense partial-pivoting kernel launched right after refactors just those.
#include <cfloat>

#include "xg_internal.h"

namespace xg {

namespace {

constexpr int kBlock = 256;
constexpr int kLdsBytes = 144 * 1024;  // factors, vector and task lists

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

__host__ __device__ constexpr int64_t align16(int64_t b) { return (b + 15) & ~(int64_t)15; }

// The task lists of one component, with op ranges relative to `ops` and
// level pointers relative to `tasks`.
struct Lists {
  const int4 *tasks;
  const int2 *ops;
  const int32_t *lvp, *lvm, *lvl;  // level -> first task, first mid, first long
  int obase;  // subtracted from the ops fields of tasks (0 once staged)
};

// Finishes a task value: division by the pivot (factor and backward
// solve) and the factor's acceptance checks.
template <int kPhase>
__device__ __forceinline__ void finish_task(const int4 &tk, double s, double *val,
                                            double *vec, int &bad) {
  if (kPhase == 0) {
    if (tk.y == kLuFwdTask) {  // a forward-substitution row (y in vec)
      vec[tk.x] = s;
      return;
    }
    if (tk.y >= 0) {
      s = s / val[tk.y];
      bad |= !(fabs(s) <= DBL_MAX);  // infinite or NaN
    } else if (tk.y == -2) {
      bad |= !(fabs(s) > 0.0) || isinf(s);
    }
    val[tk.x] = s;
  } else {
    vec[tk.x] = kPhase == 2 ? s / val[tk.y] : s;
  }
}

// One level.  Short tasks (at most four terms): one thread each, the terms'
// operands loaded together, then summed in list order.  Mid tasks: eight
// lanes each, long tasks (rows of supply rails, which touch most of the
// circuit): a wavefront each; lane-strided partial sums folded by a fixed
// butterfly.  The association is fixed by the plan, so results do not
// depend on the launch.
template <int kPhase>
__device__ __forceinline__ void run_level(const Lists &ls, int L, double *val,
                                          double *vec, int &bad) {
  const int k0 = ls.lvp[L], km = ls.lvm[L], kl = ls.lvl[L], k1 = ls.lvp[L + 1];
  const double *src0 = kPhase == 0 ? val : vec;
  for (int k = k0 + (int)threadIdx.x; k < km; k += kBlock) {
    const int4 tk = ls.tasks[k];
    const double *src = kPhase == 0 && tk.y == kLuFwdTask ? vec : src0;
    const int o0 = tk.z - ls.obase, m = tk.w - tk.z;
    double a[4], b[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      if (j < m) {
        const int2 q = ls.ops[o0 + j];
        a[j] = val[q.x];
        b[j] = src[q.y];
      }
    }
    double s = src[tk.x];
#pragma unroll
    for (int j = 0; j < 4; ++j)
      if (j < m) s = s - a[j] * b[j];
    finish_task<kPhase>(tk, s, val, vec, bad);
  }
  {
    const int grp = threadIdx.x >> 3, sub = threadIdx.x & 7;
    for (int k = km + grp; k < kl; k += kBlock / 8) {
      const int4 tk = ls.tasks[k];
      const double *src = kPhase == 0 && tk.y == kLuFwdTask ? vec : src0;
      const int o0 = tk.z - ls.obase, o1 = tk.w - ls.obase;
      double p = 0.0;
      for (int o = o0 + sub; o < o1; o += 8) {
        const int2 q = ls.ops[o];
        p = p + val[q.x] * src[q.y];
      }
      for (int off = 4; off >= 1; off >>= 1) p = p + __shfl_xor(p, off);
      if (sub == 0) finish_task<kPhase>(tk, src[tk.x] - p, val, vec, bad);
    }
  }
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int k = kl + wave; k < k1; k += kBlock / 64) {
    const int4 tk = ls.tasks[k];
    const double *src = kPhase == 0 && tk.y == kLuFwdTask ? vec : src0;
    const int o0 = tk.z - ls.obase, o1 = tk.w - ls.obase;
    double p = 0.0;
    for (int o = o0 + lane; o < o1; o += 64) {
      const int2 q = ls.ops[o];
      p = p + val[q.x] * src[q.y];
    }
    for (int off = 32; off >= 1; off >>= 1) p = p + __shfl_xor(p, off);
    if (lane == 0) finish_task<kPhase>(tk, src[tk.x] - p, val, vec, bad);
  }
}

// Lays the Jacobian out per component in entry order and the right-hand
// side in position order, so that each component's workgroup reads its
// inputs contiguously.
__global__ void __launch_bounds__(kBlock)
    permute_kernel(int32_t n_ent, int32_t n_nodes, const int32_t *ent_k,
                   const int32_t *prow, const double *J, const double *rhs,
                   double *Jp, double *bp, const int32_t *gate) {
  if (gate && *gate) return;
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n_ent) Jp[i] = J[ent_k[i]];
  if (i < n_nodes) bp[i] = rhs[prow[i]];
}

template <bool kLds>
__global__ void __launch_bounds__(kBlock) sparse_lu_kernel(SparseLuArgs a) {
  __shared__ __align__(16) char lds[kLds ? kLdsBytes : 16];
  __shared__ double sv[kBlock];
  const int c = blockIdx.x;
  const int t = threadIdx.x;
  if (a.gate && *a.gate) return;
  if ((a.in_lds[c] != 0) != kLds) return;  // the other form
  const int beg = a.cstart[c], n = a.cstart[c + 1] - beg;
  const int64_t nv = a.voff[c + 1] - a.voff[c];
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
  if (a.front && a.front[c]) {  // the dense-front kernels solve it
    if (t == 0) a.fallback[c] = 0;
    return;
  }
  if (!a.ok[c]) {  // structurally singular under matching: dense kernel
    if (t == 0) a.fallback[c] = 1;
    return;
  }
  const int32_t *lv = a.lv_comp + 3 * (size_t)c;
  const int L0 = lv[0], nlev = lv[3] - lv[0];
  const int tbase = a.lv_ptr[L0], ntask = a.lv_ptr[lv[3]] - tbase;
  const int obase = ntask ? a.tasks[tbase].z : 0;
  const int nops = ntask ? a.tasks[tbase + ntask - 1].w - obase : 0;
  double *val, *vec;
  Lists ls;
  if (kLds) {
    // stage: [values][vector][tasks][ops][level pointers]
    char *p = lds;
    val = (double *)p;
    p += align16(nv * 8);
    vec = (double *)p;
    p += align16((int64_t)n * 8);
    int4 *tk = (int4 *)p;
    p += (int64_t)ntask * 16;
    int2 *op = (int2 *)p;
    p += align16((int64_t)nops * 8);
    int32_t *lvp = (int32_t *)p;
    int32_t *lvm = lvp + nlev + 1;
    int32_t *lvl = lvm + nlev;
    for (int k = t; k < ntask; k += kBlock) tk[k] = a.tasks[tbase + k];
    for (int k = t; k < nops; k += kBlock) op[k] = a.ops[obase + k];
    for (int k = t; k <= nlev; k += kBlock) lvp[k] = a.lv_ptr[L0 + k] - tbase;
    for (int k = t; k < nlev; k += kBlock) {
      lvm[k] = a.lv_mid[L0 + k] - tbase;
      lvl[k] = a.lv_long[L0 + k] - tbase;
    }
    ls = Lists{tk, op, lvp, lvm, lvl, obase};
  } else {
    val = a.val_work + a.voff[c];
    vec = a.vec_work + beg;
    // global lists: level pointers stay absolute, ops absolute
    ls = Lists{a.tasks, a.ops, a.lv_ptr + L0, a.lv_mid + L0, a.lv_long + L0, 0};
  }
  const int32_t *qcol = a.qcol + beg;
  // the component's Jacobian entries and right-hand side, already laid out
  // contiguously by the permute kernel
  const int eb = a.ent_off[c], ne = a.ent_off[c + 1] - eb;
  const double *jp = a.Jp + eb, *bp = a.bp + beg;
  const int32_t *epos = a.ent_pos + eb, *ecol = a.ent_col + eb, *ers = a.ent_rs + beg + c;
  for (int64_t e = t; e < nv; e += kBlock) val[e] = 0.0;
  __syncthreads();
#pragma unroll 4
  for (int j = t; j < ne; j += kBlock) val[epos[j]] = jp[j];
  double rmax = 0.0;
  for (int i = t; i < n; i += kBlock) {
    const double r = bp[i];
    vec[i] = r;
    rmax = fmax(rmax, fabs(r));
  }
  const double max_rhs = blk_max(rmax, sv);  // includes the barriers
  // ---- factor
  const int f1 = lv[1] - L0, s1 = lv[2] - L0;
  int bad = 0;
  for (int L = 0; L < f1; ++L) {
    run_level<0>(ls, L, val, vec, bad);
    __syncthreads();
  }
  if (__syncthreads_or(bad)) {
    if (t == 0) a.fallback[c] = 1;
    return;
  }
  // ---- L y = P b (unit diagonal), then U x' = y, in place
  for (int L = f1; L < s1; ++L) {
    run_level<1>(ls, L, val, vec, bad);
    __syncthreads();
  }
  for (int L = s1; L < nlev; ++L) {
    run_level<2>(ls, L, val, vec, bad);
    __syncthreads();
  }
  // ---- backward error of the solution against the unfactored rows
  double rm = 0.0, am = 0.0, xm = 0.0;
  for (int i = t; i < n; i += kBlock) {
    double r = bp[i], as = 0.0;
    const int j0 = ers[i], j1 = ers[i + 1];
#pragma unroll 4
    for (int j = j0; j < j1; ++j) {
      const double v = jp[j];
      r = r - v * vec[ecol[j]];
      as = as + fabs(v);
    }
    rm = fmax(rm, fabs(r));
    am = fmax(am, as);
    xm = fmax(xm, fabs(vec[i]));
  }
  rm = blk_max(rm, sv);
  am = blk_max(am, sv);
  xm = blk_max(xm, sv);
  if (!(rm <= a.max_berr * (am * xm + max_rhs))) {  // also NaN / infinite
    if (t == 0) a.fallback[c] = 1;
    return;
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

int64_t sparse_lu_lds_bytes(int64_t n, int64_t fnnz, int64_t tasks, int64_t ops,
                            int64_t levels) {
  return align16(fnnz * 8) + align16(n * 8) + tasks * 16 + align16(ops * 8) +
         (3 * levels + 1) * 4;
}

int64_t sparse_lu_lds_budget() { return kLdsBytes; }

hipError_t launch_sparse_lu(const SparseLuArgs &a, hipStream_t s) {
  if (a.n_comp == 0) return hipSuccess;
  const int64_t nperm = a.n_ent > a.n_nodes ? a.n_ent : a.n_nodes;
  if (nperm > 0)
    hipLaunchKernelGGL(permute_kernel, dim3((unsigned)((nperm + kBlock - 1) / kBlock)),
                       dim3(kBlock), 0, s, a.n_ent, a.n_nodes, a.ent_k, a.prow, a.J,
                       a.rhs, a.Jp, a.bp, a.gate);
  hipLaunchKernelGGL(sparse_lu_kernel<true>, dim3(a.n_comp), dim3(kBlock), 0, s, a);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || !a.any_global) return e;
  hipLaunchKernelGGL(sparse_lu_kernel<false>, dim3(a.n_comp), dim3(kBlock), 0, s, a);
  return hipGetLastError();
}

}  // namespace xg
