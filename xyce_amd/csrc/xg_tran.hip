// Native transient driver: the DC operating point and ONESTEP time steps
// with the whole Newton loop on the device, the convergence decision
// included, so the host only reads a few words per iteration.
//
// Per Newton iteration (one hipGraph per solver-state variant, captured
// on first use and replayed):
//   xg_lu_solve            linear solve, x += dx, wRMS(dx) per component
//   xg_load_all            device load at the new iterate
//   xg_tia_assemble        residual and Jacobian (N_TIA_OneStep.C:215-330)
//   xg_rhs_norms           max |rhs| per component
//   xg_component_limited   per-component limiting count
//   count kernel           non-finite device values
//   newton_ctl_kernel      per-component convergence (N_NLS_DampedNewton.C:
//                          1787-1830), active mask, done / failure flags
// Step-dependent values (time, 1/dt, order) reach the captured kernels
// through device memory (xg_set_step_source), so one graph serves every
// step.  The control flow -- retry with half the step on failure -- is the
// host's, as in the Python driver it replaces.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <vector>

#include "xg_internal.h"

struct xg_tran {
  xg_handle *h = nullptr;
  xg_tran_buffers b{};
  hipStream_t stream = nullptr, old_stream = nullptr;
  int32_t n_comp = 0;
  int64_t n1 = 0, n_state = 0, n_store = 0, nnz = 0;
  double t = 0.0, bp_tol = 0.0;
  int64_t n_steps = 0, newton_total = 0, step_failures = 0;
  int32_t use_graphs = 1;
  // graphs by (variant, iterations): variant 0 DC iteration, 1 first-step
  // iteration, 2 later-step iteration, 3 / 4 a whole first / later step's
  // first load, its first iterations and (when they converged) completion.
  // Iterations after convergence inside one graph skip every launch (the
  // stop word), so a graph of K iterations costs ~K only when all are used.
  static constexpr int kMaxK = 4;
  hipGraphExec_t exec[5][kMaxK + 1] = {};
  hipGraph_t graph[5][kMaxK + 1] = {};
  // iterations the next step's graph holds: those of the last step
  // (adapted per step; no counterpart in the reference)
  int32_t k_step = 1;
  int64_t host_reads = 0;  // control-word reads (host round trips)
  int32_t *ctl_host = nullptr;  // pinned copy of the control words
  // fused fixed steps (xg_tran_run): graphs of kRunSteps whole steps by
  // iterations per step, their device descriptor and its pinned copies in
  // and out
  hipGraphExec_t run_exec[kMaxK + 1] = {};
  hipGraph_t run_graph[kMaxK + 1] = {};
  void *run_dev = nullptr, *run_in = nullptr, *run_out = nullptr;
  std::string err;
  // variable-step control (xg_tran_adaptive_*): the scalar state of the
  // reference's N_TIA_StepErrorControl and N_TIA_OneStep, and the device
  // buffers of the error estimate
  struct Adaptive {
    bool on = false;
    double tstop = 0.0, tstep = 0.0, reltol = 1e-3, abstol = 1e-6;
    std::vector<double> bps;  // 0, the PULSE corners, tstop
    double cur = 0.0, stop = 0.0, h = 0.0, last_h = 0.0, h_sum = 0.0;
    double saved_h = 1e-10, min_h = 0.0, max_h = 0.0, bp_tol = 0.0, nxt = 0.0;
    double est = 0.0, psi[3] = {0.0, 0.0, 0.0};
    int order = 1, n_steps = 0, nef = 0;
    int64_t step_number = 0;
    bool beginning = true, attempt_ok = true;
    int64_t accepted = 0, rejected_error = 0, rejected_newton = 0;
  } ad;
  double *qw = nullptr;        // n1: Q at the predictor, for the Q weights
  double *lte_part = nullptr;  // 3 x kLteBlocks partial sums
  double *lte_out = nullptr;   // 3: sum (x - xpred)^2, sum ((Q - qn0)/w)^2, max |x_n|
  double *lte_host = nullptr;  // pinned copy of lte_out
  // gmin stepping (xg_tran_dcop_continuation): the conductance (device
  // scalar, read by the captured DC iteration), the voltage unknowns and
  // their diagonal CSR positions, the last converged solution
  double *gmin_dev = nullptr;
  int32_t *gv_ids = nullptr, *gv_pos = nullptr;
  int32_t n_gv = 0;
  double *x_prev = nullptr;
  int64_t gmin_steps = 0;
  // eviction (xg_tran_set_eviction): components taken out of the solve
  // after their Newton solve fails at the smallest step; their unknowns
  // stay at the last accepted values
  int32_t *disabled = nullptr;     // n_comp
  int32_t *comp_of_row = nullptr;  // n1 (ground: -1)
  int32_t evict = 0;
  double min_dt = 1e-18;
  int64_t n_evicted = 0;
};

namespace xg {

namespace {

constexpr int kBlock = 256;

struct NewtonParams {
  double reltol, abstol, rhstol, dxtol, small_update;
  int32_t maxit;
};
// N_NLS_NLParams.C:108-125 and DampedNewton.C (small update 1e-6)
constexpr NewtonParams kDC{1e-3, 1e-12, 1e-6, 1.0, 1e-6, 200};
constexpr NewtonParams kTran{1e-2, 1e-6, 1e-2, 0.33, 1e-6, 20};

// ctl words: [0] iterations, [1] done, [2] failure (1 singular Jacobian,
// 2 non-finite device values, 3 no convergence), [3] unconverged count
// [4] stop: done or failed -- every launch of an iteration checks it
enum { kIter = 0, kDone = 1, kFail = 2, kLeft = 3, kStop = 4, kCtlWords = 8 };
// iterations per graph of the DC and the continuing-iteration loops
constexpr int kDcIters = 4, kTranIters = 2;

// Fused fixed steps: the words the host sets per group of steps (time,
// step, probe output) and those the graph returns (steps done, iterations
// per step, halt), in device memory so that one captured graph serves
// every group.  Halt: a step of the group did not converge within the
// graph's iterations, or failed; every later launch of the group does
// nothing and the host takes that step over.
constexpr int kRunSteps = 8;
struct RunDesc {
  double t, dt, inv_dt;
  const int32_t *ids;
  double *out;
  int64_t s0;  // steps of the whole run before this group
  int32_t every, n_ids, halt, done;
  int32_t iters[kRunSteps];
};

// Step start: predictor (the rounding sequence of the predictor kernel),
// iterate = predictor, step parameters, control words and masks reset.
__global__ void __launch_bounds__(kBlock)
    tran_begin_kernel(int64_t n1, int32_t n_comp, int32_t order, double beta1,
                      double beta2, const double *x_curr, const double *xh1,
                      const double *xh2, double *xpred, double *x, double *step,
                      double t, double bp_tol, double inv_dt, int32_t *conv,
                      int32_t *active, int32_t *ctl, const int32_t *disabled,
                      const RunDesc *run) {
  if (run) {  // a fused step: its time from the group's descriptor
    if (run->halt) return;
    t = run->t + run->dt;
    inv_dt = run->inv_dt;
  }
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i < n1) {
    double v = x_curr[i];
    v = beta1 * xh1[i] + 1.0 * v;
    if (order == 2) v = beta2 * xh2[i] + 1.0 * v;
    xpred[i] = v;
    x[i] = v;
  }
  if (i < n_comp) {
    const int32_t off = disabled[i];
    conv[i] = off;
    active[i] = 1 - off;
  }
  if (i == 0) {
    step[0] = t;
    step[1] = bp_tol;
    step[2] = inv_dt;
    step[3] = (double)order;
    for (int k = 0; k < kCtlWords; ++k) ctl[k] = 0;
  }
}

// DC start: iterate, current state and store zero; masks and words reset.
__global__ void __launch_bounds__(kBlock)
    dc_begin_kernel(int64_t n1, int64_t n_state, int64_t n_store, int32_t n_comp,
                    double *x, double *x_curr, double *sta_c, double *sto_c,
                    double *step, double bp_tol, int32_t *conv, int32_t *active,
                    int32_t *ctl, const int32_t *disabled) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i < n1) {
    x[i] = 0.0;
    x_curr[i] = 0.0;
  }
  if (i < n_state) sta_c[i] = 0.0;
  if (i < n_store) sto_c[i] = 0.0;
  if (i < n_comp) {
    const int32_t off = disabled[i];
    conv[i] = off;
    active[i] = 1 - off;
  }
  if (i == 0) {
    step[0] = 0.0;
    step[1] = bp_tol;
    step[2] = 0.0;
    step[3] = 1.0;
    for (int k = 0; k < kCtlWords; ++k) ctl[k] = 0;
  }
}

// The Newton convergence test per component after an iteration's reload:
//   devices converged (no limiting) or iteration limit reached, and
//   (max|rhs| <= RHSTol and wRMS(dx) <= deltaXTol, or wRMS(dx) <= 1e-6);
// a converged component stays converged and leaves the active mask.
__global__ void __launch_bounds__(kBlock)
    newton_ctl_kernel(int32_t n_comp, const double *stats, double *nstats, int32_t *lim,
                      const int64_t *counts, NewtonParams prm, int32_t *conv,
                      int32_t *active, int32_t *ctl, const int32_t *disabled) {
  __shared__ int32_t s_left, s_sing;
  const int t = threadIdx.x;
  if (ctl[kStop]) return;  // an iteration after convergence or failure
  if (t == 0) {
    s_left = 0;
    s_sing = 0;
  }
  __syncthreads();
  const int32_t it = ctl[kIter] + 1;
  int left = 0, sing = 0;
  for (int c = t; c < n_comp; c += kBlock) {
    if (disabled[c]) {  // evicted: out of the solve and the test
      nstats[4 * (size_t)c] = 0.0;
      lim[c] = 0;
      conv[c] = 1;
      active[c] = 0;
      continue;
    }
    const double upd = stats[4 * (size_t)c + 1];
    sing |= stats[4 * (size_t)c + 2] != 0.0;
    const double max_rhs = nstats[4 * (size_t)c];
    const bool dev_ok = lim[c] == 0 || it >= prm.maxit;
    nstats[4 * (size_t)c] = 0.0;  // the accumulators of the next iteration
    lim[c] = 0;
    const bool now = dev_ok && ((max_rhs <= prm.rhstol && upd <= prm.dxtol) ||
                                upd <= prm.small_update);
    const int32_t cv = conv[c] | (now ? 1 : 0);
    conv[c] = cv;
    active[c] = cv ? 0 : 1;
    left += cv ? 0 : 1;
  }
  if (left) atomicAdd(&s_left, left);
  if (sing) atomicOr(&s_sing, 1);
  __syncthreads();
  if (t == 0) {
    int64_t nonfinite = 0;
    for (int k = 0; k < 64; ++k) nonfinite += counts[64 + k];
    ctl[kIter] = it;
    ctl[kLeft] = s_left;
    if (s_sing)
      ctl[kFail] = 1;
    else if (nonfinite)
      ctl[kFail] = 2;
    else if (s_left == 0)
      ctl[kDone] = 1;
    else if (it >= prm.maxit)
      ctl[kFail] = 3;
    if (ctl[kFail] || ctl[kDone]) ctl[kStop] = 1;
  }
}

// Accepted step (N_TIA_OneStep::completeStep history and state rotation).
// With ctl given, only when the Newton solve has converged (the fused step
// graph runs it right after the first iteration).
__global__ void __launch_bounds__(kBlock)
    tran_complete_kernel(int64_t n1, int64_t n_state, int64_t n_store, const double *x,
                         double *x_curr, double *xh1, double *xh2, const double *Q,
                         const double *F, double *qn0, double *fprev,
                         const double *sta_n, double *sta_c, const double *sto_n,
                         double *sto_c, int32_t dc, const int32_t *ctl) {
  if (ctl && !ctl[kDone]) return;
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i < n1) {
    if (dc) {
      xh2[i] = 0.0;
      xh1[i] = 0.0;
    } else {
      xh2[i] = xh1[i];
      xh1[i] = x[i] - x_curr[i];
    }
    x_curr[i] = x[i];
    qn0[i] = Q[i];
    fprev[i] = F[i];
  }
  if (i < n_state) sta_c[i] = sta_n[i];
  if (i < n_store) sto_c[i] = sto_n[i];
}

// gmin stepping: a conductance g from every voltage unknown to ground, in
// F (the residual is -F) and on the dF/dx diagonal after each DC load
// (N_NLS_NOX_AugmentLinSys_GStepping.C augmentResidual / augmentJacobian).
// g = 0 (no continuation, or its last step) leaves the load untouched.
__global__ void __launch_bounds__(kBlock)
    gmin_augment_kernel(int32_t n_v, const int32_t *ids, const int32_t *pos,
                        const double *g_ptr, const double *x, double *F, double *JF,
                        const int32_t *gate) {
  if (gate && *gate) return;
  const double g = *g_ptr;
  if (g == 0.0) return;
  const int32_t k = blockIdx.x * kBlock + threadIdx.x;
  if (k >= n_v) return;
  const int32_t i = ids[k];
  F[i] = F[i] + g * x[i];
  JF[pos[k]] = JF[pos[k]] + g;
}

// Error-estimate sums of an attempted step over the n unknowns, in a fixed
// order (per block a tree over the threads' strided sums, then the blocks
// in a second kernel): sum (x - x_pred)^2, sum ((Q - Q_n) / (reltol |Q_w|
// + abstol))^2 and max |x_n| (N_TIA_DataStore.C:1174-1330, 1377-1420,
// 1782-1800).
constexpr int kLteBlocks = 256;

__device__ __forceinline__ void tree3(double *a, double *b, double *m) {
  const int t = threadIdx.x;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if (t < s) {
      a[t] = a[t] + a[t + s];
      b[t] = b[t] + b[t + s];
      m[t] = fmax(m[t], m[t + s]);
    }
    __syncthreads();
  }
}

__global__ void __launch_bounds__(kBlock)
    lte_partials_kernel(int64_t n, const double *x, const double *xpred,
                        const double *x_curr, const double *Q, const double *qn0,
                        const double *qw, double reltol, double abstol, double *part) {
  __shared__ double a[kBlock], b[kBlock], m[kBlock];
  double sa = 0.0, sb = 0.0, sm = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock) {
    const double d = x[i] - xpred[i];
    sa = sa + d * d;
    const double e = (Q[i] - qn0[i]) / (reltol * fabs(qw[i]) + abstol);
    sb = sb + e * e;
    sm = fmax(sm, fabs(x_curr[i]));
  }
  a[threadIdx.x] = sa;
  b[threadIdx.x] = sb;
  m[threadIdx.x] = sm;
  tree3(a, b, m);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = a[0];
    part[kLteBlocks + blockIdx.x] = b[0];
    part[2 * kLteBlocks + blockIdx.x] = m[0];
  }
}

__global__ void __launch_bounds__(kBlock) lte_final_kernel(const double *part, double *out) {
  __shared__ double a[kBlock], b[kBlock], m[kBlock];
  static_assert(kLteBlocks == kBlock, "one thread per block partial");
  a[threadIdx.x] = part[threadIdx.x];
  b[threadIdx.x] = part[kLteBlocks + threadIdx.x];
  m[threadIdx.x] = part[2 * kLteBlocks + threadIdx.x];
  tree3(a, b, m);
  if (threadIdx.x == 0) {
    out[0] = a[0];
    out[1] = b[0];
    out[2] = m[0];
  }
}

// Evicted components: their unknowns back at the accepted values and their
// history cleared, so the predictor keeps them there.
__global__ void __launch_bounds__(kBlock)
    freeze_kernel(int64_t n1, const int32_t *comp_of_row, const int32_t *disabled,
                  const double *x_curr, double *x, double *xh1, double *xh2) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n1) return;
  const int32_t c = comp_of_row[i];
  if (c < 0 || !disabled[c]) return;
  x[i] = x_curr[i];
  xh1[i] = 0.0;
  xh2[i] = 0.0;
}

// End of fused step j: a converged step advances the group's time (the
// host's t + dt, in the same order) and counts as done; otherwise the
// group halts, its remaining launches stopped by the stop word.
__global__ void run_end_kernel(RunDesc *run, int32_t *ctl, int32_t j) {
  if (threadIdx.x != 0 || run->halt) return;
  run->iters[j] = ctl[kIter];
  if (ctl[kDone]) {
    run->t = run->t + run->dt;
    run->done = j + 1;
  } else {
    run->halt = 1;
    ctl[kStop] = 1;
  }
}

// The probes of fused step j when its step number is a multiple of
// `every`: row s / every - 1 of the run's output.
constexpr int kRunProbeBlocks = 4;
__global__ void __launch_bounds__(kBlock)
    run_probe_kernel(const RunDesc *run, const double *x, int64_t n_nodes, int32_t j) {
  if (run->halt) return;
  const int64_t s = run->s0 + j + 1;
  if (s % run->every) return;
  double *out = run->out + (s / run->every - 1) * (int64_t)run->n_ids;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < run->n_ids;
       i += kRunProbeBlocks * kBlock) {
    const int32_t id = run->ids[i];
    out[i] = (id >= 0 && id < n_nodes) ? x[id] : __builtin_nan("");
  }
}

// A step halted inside a fused group continues with the ordinary
// iteration graphs: its stop word cleared unless it failed.
__global__ void run_resume_kernel(int32_t *ctl) {
  if (threadIdx.x == 0 && !ctl[kFail] && !ctl[kDone]) ctl[kStop] = 0;
}

// ids outside 0..n_nodes-1 (ground, a typo mapped to nothing) read no
// memory: their output is NaN, so a bad probe is visible, never another
// node's waveform.
__global__ void probe_kernel(const double *x, int64_t n_nodes, const int32_t *ids,
                             int32_t n, double *out) {
  const int i = blockIdx.x * 64 + threadIdx.x;
  if (i < n) {
    const int32_t id = ids[i];
    out[i] = (id >= 0 && id < n_nodes) ? x[id] : __builtin_nan("");
  }
}

unsigned grid_for(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

}  // namespace

}  // namespace xg

namespace {

int tfail(xg_tran *T, int code, const std::string &m) {
  T->err = m;
  return code;
}

void set_state(xg_handle *h, int dc, int init_jct, int init_tran, int it,
               int transient) {
  xg_solver_state s{};
  s.dcopFlag = dc;
  s.tranopFlag = dc;
  s.initJctFlag = init_jct;
  s.initTranFlag = init_tran;
  s.newtonIter = it;
  s.transientFlag = transient;
  xg_set_solver_state(h, &s);
}

int enqueue_gmin(xg_tran *T, const int32_t *gate) {
  if (!T->n_gv) return XG_OK;
  hipLaunchKernelGGL(xg::gmin_augment_kernel, dim3(xg::grid_for(T->n_gv)), dim3(xg::kBlock),
                     0, T->stream, T->n_gv, (const int32_t *)T->gv_ids,
                     (const int32_t *)T->gv_pos, (const double *)T->gmin_dev,
                     (const double *)T->b.x, T->b.F, T->b.JF, gate);
  return hipGetLastError() == hipSuccess ? XG_OK : XG_E_HIP;
}

// The first load of a Newton solve and its assembly (init_jct < 0: the DC
// default, junction initialisation on a DC solve).
int enqueue_first_load(xg_tran *T, int dc, int first, int init_jct = -1) {
  const xg_tran_buffers &b = T->b;
  set_state(T->h, dc, init_jct < 0 ? dc : init_jct, first, 0, !dc);
  int rc = xg_load_all(T->h, b.x, b.sta_c, b.sto_c, b.sta_n, b.sto_n, b.F, b.Q, b.FL,
                       b.QL, b.JF, b.JQ);
  if (rc) return rc;
  if (dc && (rc = enqueue_gmin(T, nullptr))) return rc;
  return xg_tia_assemble(T->h, !dc, 1, 0.0, b.qn0, b.fprev, b.F, b.Q, b.FL, b.QL, b.JF,
                         b.JQ, b.rhs, b.J);
}

// One Newton iteration after the first load; variant 0 DC, 1 first step,
// 2 later steps.
int enqueue_iteration(xg_tran *T, int variant) {
  const xg_tran_buffers &b = T->b;
  const bool dc = variant == 0;
  const xg::NewtonParams &prm = dc ? xg::kDC : xg::kTran;
  int rc = xg_lu_solve(T->h, b.J, b.rhs, b.x, b.x_curr, dc ? b.x : b.xpred, prm.reltol,
                       prm.abstol, b.active, b.stats);
  if (rc) return rc;
  set_state(T->h, dc, 0, variant == 1, 1, !dc);
  rc = xg_load_all(T->h, b.x, b.sta_c, b.sto_c, b.sta_n, b.sto_n, b.F, b.Q, b.FL, b.QL,
                   b.JF, b.JQ);
  if (rc) return rc;
  if (dc && (rc = enqueue_gmin(T, b.ctl + xg::kStop))) return rc;
  // assembly with the per-component max |rhs| (xg_rhs_norms), then one
  // pass for the per-component limiting counts (xg_component_limited) and
  // the non-finite count; newton_ctl_kernel clears both accumulators for
  // the next iteration
  rc = xg::handle_assemble_norms(T->h, !dc, b.qn0, b.fprev, b.F, b.Q, b.FL, b.QL, b.JF,
                                 b.JQ, b.rhs, b.J, b.nstats);
  if (rc) return rc;
  hipError_t e = xg::handle_iteration_counts(T->h, b.lim, b.counts);
  if (e != hipSuccess) return XG_E_HIP;
  hipLaunchKernelGGL(xg::newton_ctl_kernel, dim3(1), dim3(xg::kBlock), 0, T->stream,
                     T->n_comp, (const double *)b.stats, b.nstats, b.lim,
                     (const int64_t *)b.counts, prm, b.conv, b.active, b.ctl,
                     (const int32_t *)T->disabled);
  return hipGetLastError() == hipSuccess ? XG_OK : XG_E_HIP;
}

int enqueue_complete(xg_tran *T, int dc, const int32_t *ctl) {
  const xg_tran_buffers &b = T->b;
  const int64_t n = std::max(T->n1, std::max(T->n_state, T->n_store));
  hipLaunchKernelGGL(xg::tran_complete_kernel, dim3(xg::grid_for(n)), dim3(xg::kBlock), 0,
                     T->stream, T->n1, T->n_state, T->n_store, (const double *)b.x,
                     b.x_curr, b.xh1, b.xh2, (const double *)b.Q, (const double *)b.F,
                     b.qn0, b.fprev, (const double *)b.sta_n, b.sta_c,
                     (const double *)b.sto_n, b.sto_c, dc, ctl);
  return hipGetLastError() == hipSuccess ? XG_OK : XG_E_HIP;
}

// Variants 0-2: K Newton iterations; 3 / 4: a first / later step's first
// load and assembly, its first K iterations, and the completion if they
// converged (one launch for a step that converges within K).  Every launch
// checks the stop word, so the iterations after convergence do nothing.
int enqueue_body(xg_tran *T, int variant, int K) {
  if (variant < 3) {
    for (int k = 0; k < K; ++k)
      if (int rc = enqueue_iteration(T, variant)) return rc;
    return XG_OK;
  }
  const int first = variant == 3;
  int rc = enqueue_first_load(T, 0, first);
  if (rc) return rc;
  for (int k = 0; k < K; ++k)
    if ((rc = enqueue_iteration(T, first ? 1 : 2))) return rc;
  return enqueue_complete(T, 0, T->b.ctl);
}

int enqueue_variant(xg_tran *T, int variant, int K) {
  xg::handle_set_gate(T->h, T->b.ctl + xg::kStop);
  const int rc = enqueue_body(T, variant, K);
  xg::handle_set_gate(T->h, nullptr);
  return rc;
}

int capture_variant(xg_tran *T, int variant, int K) {
  if (T->exec[variant][K]) return XG_OK;
  if (hipStreamBeginCapture(T->stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
    return tfail(T, XG_E_HIP, "graph capture begin");
  const int rc = enqueue_variant(T, variant, K);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(T->stream, &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) return tfail(T, XG_E_HIP, "graph capture end");
  T->graph[variant][K] = g;
  if (hipGraphInstantiate(&T->exec[variant][K], g, nullptr, nullptr, 0) != hipSuccess)
    return tfail(T, XG_E_HIP, "graph instantiate");
  return XG_OK;
}

int launch_iteration(xg_tran *T, int variant, int K) {
  if (!T->use_graphs) return enqueue_variant(T, variant, K);
  const int rc = capture_variant(T, variant, K);
  if (rc) return rc;
  return hipGraphLaunch(T->exec[variant][K], T->stream) == hipSuccess ? XG_OK : XG_E_HIP;
}

// kRunSteps whole later steps of a fixed step (order 2), K iterations
// each: the launches of a step's graph (variant 4) between a fused step
// start and the step's end and probe kernels.
int enqueue_run(xg_tran *T, int K) {
  const xg_tran_buffers &b = T->b;
  const int64_t n = std::max(T->n1, (int64_t)T->n_comp);
  auto *run = (xg::RunDesc *)T->run_dev;
  for (int j = 0; j < xg::kRunSteps; ++j) {
    hipLaunchKernelGGL(xg::tran_begin_kernel, dim3(xg::grid_for(n)), dim3(xg::kBlock), 0,
                       T->stream, T->n1, T->n_comp, 2, 1.5, -0.5, (const double *)b.x_curr,
                       (const double *)b.xh1, (const double *)b.xh2, b.xpred, b.x, b.step,
                       0.0, T->bp_tol, 0.0, b.conv, b.active, b.ctl,
                       (const int32_t *)T->disabled, (const xg::RunDesc *)run);
    int rc = enqueue_variant(T, 4, K);
    if (rc) return rc;
    hipLaunchKernelGGL(xg::run_end_kernel, dim3(1), dim3(64), 0, T->stream, run, b.ctl, j);
    hipLaunchKernelGGL(xg::run_probe_kernel, dim3(xg::kRunProbeBlocks), dim3(xg::kBlock), 0,
                       T->stream, (const xg::RunDesc *)run, (const double *)b.x_curr,
                       T->n1 - 1, j);
    if (hipGetLastError() != hipSuccess) return XG_E_HIP;
  }
  return XG_OK;
}

int capture_run(xg_tran *T, int K) {
  if (T->run_exec[K]) return XG_OK;
  if (hipStreamBeginCapture(T->stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
    return tfail(T, XG_E_HIP, "graph capture begin");
  const int rc = enqueue_run(T, K);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(T->stream, &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) return tfail(T, XG_E_HIP, "graph capture end");
  T->run_graph[K] = g;
  if (hipGraphInstantiate(&T->run_exec[K], g, nullptr, nullptr, 0) != hipSuccess)
    return tfail(T, XG_E_HIP, "graph instantiate");
  return XG_OK;
}

int read_ctl(xg_tran *T) {
  T->host_reads++;
  if (hipMemcpyAsync(T->ctl_host, T->b.ctl, sizeof(int32_t) * xg::kCtlWords,
                     hipMemcpyDeviceToHost, T->stream) != hipSuccess ||
      hipStreamSynchronize(T->stream) != hipSuccess)
    return tfail(T, XG_E_HIP, "control word copy");
  return XG_OK;
}

// Newton iterations until convergence: returns 0 converged, >0 the failure
// word, <0 an error code.  `first_variant`, if not negative, is launched
// first with `first_k` iterations (the fused step graph); then graphs of
// `k` iterations follow, the control words read once per graph.
int newton_loop(xg_tran *T, int variant, int k, int32_t *iters, int first_variant = -1,
                int first_k = 1) {
  for (int j = 0;; ++j) {
    const bool fused = j == 0 && first_variant >= 0;
    int rc = launch_iteration(T, fused ? first_variant : variant, fused ? first_k : k);
    if (rc) return rc < 0 ? rc : XG_E_HIP;
    rc = read_ctl(T);
    if (rc) return rc;
    *iters = T->ctl_host[xg::kIter];
    if (T->ctl_host[xg::kFail]) return T->ctl_host[xg::kFail];
    if (T->ctl_host[xg::kDone]) return 0;
  }
}

int complete(xg_tran *T, int dc) { return enqueue_complete(T, dc, nullptr); }

// Marks the components whose last Newton solve did not converge as evicted
// and freezes their unknowns; returns how many (< 0 on a HIP error).
int evict_unconverged(xg_tran *T) {
  const int32_t nc = T->n_comp;
  if (nc <= 0) return 0;
  std::vector<int32_t> conv(nc), off(nc);
  if (hipMemcpyAsync(conv.data(), T->b.conv, sizeof(int32_t) * nc, hipMemcpyDeviceToHost,
                     T->stream) != hipSuccess ||
      hipMemcpyAsync(off.data(), T->disabled, sizeof(int32_t) * nc, hipMemcpyDeviceToHost,
                     T->stream) != hipSuccess ||
      hipStreamSynchronize(T->stream) != hipSuccess)
    return -1;
  int k = 0;
  for (int32_t c = 0; c < nc; ++c)
    if (!off[c] && !conv[c]) {
      off[c] = 1;
      ++k;
    }
  if (k == 0) return 0;
  if (hipMemcpyAsync(T->disabled, off.data(), sizeof(int32_t) * nc, hipMemcpyHostToDevice,
                     T->stream) != hipSuccess)
    return -1;
  const xg_tran_buffers &b = T->b;
  hipLaunchKernelGGL(xg::freeze_kernel, dim3(xg::grid_for(T->n1)), dim3(xg::kBlock), 0,
                     T->stream, T->n1, (const int32_t *)T->comp_of_row,
                     (const int32_t *)T->disabled, (const double *)b.x_curr, b.x, b.xh1, b.xh2);
  if (hipGetLastError() != hipSuccess || hipStreamSynchronize(T->stream) != hipSuccess)
    return -1;
  T->n_evicted += k;
  return k;
}

// ---- variable-step ONESTEP: the step-size and order control of
// N_TIA_StepErrorControl / N_TIA_OneStep around one device Newton solve per
// attempt.  Scalar host logic; the attempt, the error sums and the history
// update run on the device.

using Ad = xg_tran::Adaptive;

// updateStopTime (N_TIA_StepErrorControl.C:520-620): the first breakpoint
// after the current time; one reached at the start of an integration is
// passed over.
void ad_update_stop(Ad &a) {
  const double old = a.stop;
  auto it = std::upper_bound(a.bps.begin(), a.bps.end(), a.cur);
  a.stop = std::min(a.tstop, it != a.bps.end() ? *it : a.tstop);
  if (std::fabs(a.stop - old) < a.bp_tol && a.beginning && a.stop != a.tstop) {
    auto nx = std::upper_bound(a.bps.begin(), a.bps.end(), a.stop);
    a.stop = nx != a.bps.end() ? *nx : a.tstop;
  }
}

// N_TIA_OneStep::initialize (N_TIA_OneStep.C:1482-1640) at the start of the
// integration and after each breakpoint: first step min(TSTEP, 0.005 times
// the time to the stop), later 0.1 min(saved step, time to stop); order 1
// and a constant history (xh1 = xh2 = 0).
int ad_initialize(xg_tran *T) {
  Ad &a = T->ad;
  const double to_stop = a.stop - a.cur;
  if (a.cur == 0.0) {
    a.h = std::min(0.005 * std::fabs(to_stop), a.tstep);
  } else {
    a.h = std::min(a.h, 0.1 * std::min(a.saved_h, std::fabs(to_stop)));
  }
  a.h_sum = 2.0 * a.h;
  a.last_h = a.h;
  a.nxt = a.cur + a.h;
  a.n_steps = 0;
  a.order = 1;
  a.psi[0] = a.h;
  const size_t bytes = sizeof(double) * (size_t)T->n1;
  if (hipMemsetAsync(T->b.xh1, 0, bytes, T->stream) != hipSuccess ||
      hipMemsetAsync(T->b.xh2, 0, bytes, T->stream) != hipSuccess)
    return tfail(T, XG_E_HIP, "history reset");
  return XG_OK;
}

// Error estimate of a converged attempt: ck sqrt((|dx|_w^2 + |dQ|_w^2) / 2)
// with the weight reltol max|x_n| + abstol for every unknown.
int ad_estimate(xg_tran *T, double ck, double *est) {
  const xg_tran_buffers &b = T->b;
  const Ad &a = T->ad;
  const int64_t n = T->n1 - 1;
  hipLaunchKernelGGL(xg::lte_partials_kernel, dim3(xg::kLteBlocks), dim3(xg::kBlock), 0,
                     T->stream, n, (const double *)b.x, (const double *)b.xpred,
                     (const double *)b.x_curr, (const double *)b.Q, (const double *)b.qn0,
                     (const double *)T->qw, a.reltol, a.abstol, T->lte_part);
  hipLaunchKernelGGL(xg::lte_final_kernel, dim3(1), dim3(xg::kBlock), 0, T->stream,
                     (const double *)T->lte_part, T->lte_out);
  T->host_reads++;
  if (hipGetLastError() != hipSuccess ||
      hipMemcpyAsync(T->lte_host, T->lte_out, 3 * sizeof(double), hipMemcpyDeviceToHost,
                     T->stream) != hipSuccess ||
      hipStreamSynchronize(T->stream) != hipSuccess)
    return tfail(T, XG_E_HIP, "error estimate");
  const double wx = a.reltol * T->lte_host[2] + a.abstol;
  const double ex2 = n ? T->lte_host[0] / (double)n / (wx * wx) : 0.0;
  const double eq2 = n ? T->lte_host[1] / (double)n : 0.0;
  *est = ck * std::sqrt(0.5 * ex2 + 0.5 * eq2);
  return XG_OK;
}

// One attempt of the step from a.cur to a.nxt with order a.order.  Returns
// XG_OK with *newton_ok set, or an error code.
int ad_attempt(xg_tran *T, double beta1, double beta2, bool *newton_ok) {
  const xg_tran_buffers &b = T->b;
  Ad &a = T->ad;
  const bool first = a.step_number <= 0;
  if (hipMemcpyAsync(T->qw, b.Q, sizeof(double) * (size_t)T->n1, hipMemcpyDeviceToDevice,
                     T->stream) != hipSuccess)
    return tfail(T, XG_E_HIP, "Q weights");
  const int64_t n = std::max(T->n1, (int64_t)T->n_comp);
  hipLaunchKernelGGL(xg::tran_begin_kernel, dim3(xg::grid_for(n)), dim3(xg::kBlock), 0,
                     T->stream, T->n1, T->n_comp, a.order, beta1, beta2,
                     (const double *)b.x_curr, (const double *)b.xh1, (const double *)b.xh2,
                     b.xpred, b.x, b.step, a.nxt, a.bp_tol, 1.0 / a.h, b.conv, b.active,
                     b.ctl, (const int32_t *)T->disabled, nullptr);
  if (hipGetLastError() != hipSuccess) return tfail(T, XG_E_HIP, "step begin");
  int rc = enqueue_first_load(T, 0, first);
  if (rc) return rc;
  int32_t iters = 0;
  rc = newton_loop(T, first ? 1 : 2, xg::kTranIters, &iters);
  T->newton_total += iters;
  if (rc < 0) return rc;
  *newton_ok = rc == 0;
  return XG_OK;
}

// completeStep (N_TIA_OneStep.C:1945-2190) after an accepted attempt, with
// the minimum step and breakpoint tolerance of processSuccessfulStep
// (N_ANP_Transient.C:1039-1130).
void ad_complete(Ad &a) {
  a.min_h = a.cur * 10.0 * std::numeric_limits<double>::epsilon();
  a.bp_tol = 2.0 * a.min_h;
  a.n_steps++;
  a.nef = 0;
  a.cur = a.nxt;
  double new_h = a.h;
  a.last_h = a.h;
  double rr = std::pow(0.5 / (a.est + 0.0001), 1.0 / (a.order + 1.0));
  if (a.n_steps >= 2 && a.order == 1) {
    a.order = 2;
    rr = std::pow(0.5 / (a.est + 0.0001), 1.0 / (a.order + 1.0));
    if (rr <= 1.05) a.order = 1;
  }
  if (rr >= 2.0)
    new_h = 2.0 * a.h;
  else if (rr <= 1.0)
    new_h = std::max(0.25, std::min(0.9, rr)) * a.h;
  new_h = std::min(std::max(new_h, a.min_h), a.max_h);
  if (a.stop - a.cur >= a.min_h) {
    double npt = a.cur + new_h;
    if (npt > a.stop) {
      a.saved_h = new_h;
      npt = a.stop;
      new_h = a.stop - a.cur;
    }
    a.nxt = npt;
    a.h_sum = new_h + a.last_h;
  }
  a.h = new_h;
  a.step_number++;
  a.accepted++;
  a.beginning = std::fabs(a.cur - a.stop) <= a.bp_tol && std::fabs(a.cur - a.tstop) > a.bp_tol;
}

// rejectStep (N_TIA_OneStep.C:1695-1930).
int ad_reject(xg_tran *T, bool newton_ok) {
  Ad &a = T->ad;
  a.nef++;
  for (int i = 1; i <= a.order; ++i) a.psi[i - 1] = a.psi[i];
  if (a.nef >= 15) return tfail(T, XG_E_STATE, "maximum number of local error test failures");
  double new_h;
  if (!newton_ok) {
    new_h = a.h / 8;
    a.rejected_newton++;
  } else if (a.nef == 1) {
    const double rr = std::pow(0.5 / (a.est + 0.0001), 1.0 / (a.order + 1.0));
    new_h = std::max(0.25, std::min(0.9, rr)) * a.h;
    a.rejected_error++;
  } else {
    new_h = 0.25 * a.h;
    a.rejected_error++;
  }
  a.order = 1;
  new_h = std::min(std::max(new_h, a.min_h), a.max_h);
  double npt = a.cur + new_h;
  if (npt > a.stop) {
    npt = a.stop;
    new_h = a.stop - a.cur;
  }
  a.nxt = npt;
  a.h_sum = new_h + a.last_h;
  a.h = new_h;
  T->step_failures++;
  if (a.h <= a.min_h) return tfail(T, XG_E_STATE, "time step too small");
  return XG_OK;
}

}  // namespace

extern "C" {

int xg_tran_create(xg_handle *h, const xg_tran_buffers *b, double bp_tol,
                   xg_tran **out) {
  if (!h || !b || !out) return XG_E_ARG;
  const void *need[] = {b->x,   b->x_curr, b->xh1,  b->xh2,  b->xpred,  b->F,
                        b->Q,   b->FL,     b->QL,   b->JF,   b->JQ,     b->J,
                        b->rhs, b->qn0,    b->fprev, b->stats, b->nstats, b->lim,
                        b->active, b->conv, b->counts, b->ctl, b->step};
  for (const void *p : need)
    if (!p) return XG_E_ARG;
  auto *T = new xg_tran;
  T->h = h;
  T->b = *b;
  T->bp_tol = bp_tol;
  T->n_comp = xg::handle_n_comp(h);
  T->n1 = xg::handle_n_nodes(h) + 1;
  xg::handle_sizes(h, &T->n_state, &T->n_store, &T->nnz);
  if ((T->n_state && (!b->sta_c || !b->sta_n)) || (T->n_store && (!b->sto_c || !b->sto_n))) {
    delete T;
    return XG_E_ARG;
  }
  T->old_stream = xg::handle_stream(h);
  if (hipStreamCreateWithFlags(&T->stream, hipStreamNonBlocking) != hipSuccess ||
      hipHostMalloc((void **)&T->ctl_host, sizeof(int32_t) * xg::kCtlWords) != hipSuccess) {
    if (T->stream) (void)hipStreamDestroy(T->stream);
    delete T;
    return XG_E_HIP;
  }
  // work queued by the caller on the handle's stream comes first
  (void)hipStreamSynchronize(T->old_stream);
  xg_set_stream(h, T->stream);
  xg_set_step_source(h, b->step);
  xg_set_active_mask(h, b->active);
  // the per-iteration accumulators start at zero (newton_ctl_kernel clears
  // them after every iteration)
  if (hipMemsetAsync(b->lim, 0, sizeof(int32_t) * (size_t)T->n_comp, T->stream) !=
          hipSuccess ||
      hipMemsetAsync(b->nstats, 0, sizeof(double) * 4 * (size_t)T->n_comp, T->stream) !=
          hipSuccess ||
      hipStreamSynchronize(T->stream) != hipSuccess) {
    xg_tran_destroy(T);
    return XG_E_HIP;
  }
  {
    // eviction flags (none evicted) and each row's component
    std::vector<int32_t> comp((size_t)T->n1, -1);
    int32_t nc = 0;
    if (T->n1 > 1) xg_components(h, &nc, comp.data());
    std::vector<int32_t> zeros((size_t)std::max(T->n_comp, 1), 0);
    if (hipMalloc((void **)&T->disabled, sizeof(int32_t) * zeros.size()) != hipSuccess ||
        hipMemcpy(T->disabled, zeros.data(), sizeof(int32_t) * zeros.size(),
                  hipMemcpyHostToDevice) != hipSuccess ||
        hipMalloc((void **)&T->comp_of_row, sizeof(int32_t) * comp.size()) != hipSuccess ||
        hipMemcpy(T->comp_of_row, comp.data(), sizeof(int32_t) * comp.size(),
                  hipMemcpyHostToDevice) != hipSuccess) {
      xg_tran_destroy(T);
      return XG_E_NOMEM;
    }
  }
  {
    std::vector<int32_t> ids, pos;
    xg::handle_voltage_diagonal(h, ids, pos);
    T->n_gv = (int32_t)ids.size();
    const double zero = 0.0;
    if (hipMalloc((void **)&T->gmin_dev, sizeof(double)) != hipSuccess ||
        hipMemcpy(T->gmin_dev, &zero, sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
        (T->n_gv &&
         (hipMalloc((void **)&T->gv_ids, sizeof(int32_t) * ids.size()) != hipSuccess ||
          hipMalloc((void **)&T->gv_pos, sizeof(int32_t) * pos.size()) != hipSuccess ||
          hipMemcpy(T->gv_ids, ids.data(), sizeof(int32_t) * ids.size(),
                    hipMemcpyHostToDevice) != hipSuccess ||
          hipMemcpy(T->gv_pos, pos.data(), sizeof(int32_t) * pos.size(),
                    hipMemcpyHostToDevice) != hipSuccess))) {
      xg_tran_destroy(T);
      return XG_E_NOMEM;
    }
  }
  // every graph is captured here, so that the run itself makes no capture:
  // a capture must not overlap another thread's device allocations or
  // synchronous copies, and creation is where a caller can serialise those
  for (int v = 0; v < 5; ++v)
    for (int K = 1; K <= xg_tran::kMaxK; ++K) {
      const bool need = v >= 3 || K == (v == 0 ? xg::kDcIters : xg::kTranIters);
      if (!need) continue;
      const int rc = capture_variant(T, v, K);
      if (rc) {
        xg_tran_destroy(T);
        return rc;
      }
    }
  *out = T;
  return XG_OK;
}

int xg_tran_set_graphs(xg_tran *T, int32_t on) {
  if (!T) return XG_E_ARG;
  T->use_graphs = on != 0;
  return XG_OK;
}

int xg_tran_dcop(xg_tran *T) {
  if (!T) return XG_E_ARG;
  const xg_tran_buffers &b = T->b;
  const int64_t n = std::max(T->n1, std::max(T->n_state, std::max(T->n_store,
                                                                  (int64_t)T->n_comp)));
  hipLaunchKernelGGL(xg::dc_begin_kernel, dim3(xg::grid_for(n)), dim3(xg::kBlock), 0,
                     T->stream, T->n1, T->n_state, T->n_store, T->n_comp, b.x, b.x_curr,
                     b.sta_c, b.sto_c, b.step, T->bp_tol, b.conv, b.active, b.ctl,
                     (const int32_t *)T->disabled);
  if (hipGetLastError() != hipSuccess) return tfail(T, XG_E_HIP, "dc begin");
  int rc = enqueue_first_load(T, 1, 0);
  if (rc) return rc;
  int32_t iters = 0;
  rc = newton_loop(T, 0, xg::kDcIters, &iters);
  T->newton_total += iters;
  if (rc) return rc;
  rc = complete(T, 1);
  if (rc) return rc;
  T->t = 0.0;
  T->n_steps = 0;
  return XG_OK;
}

// gmin stepping (.options nonlin continuation=gmin, N_NLS_NOX_Interface.C:
// 894-960 with N_NLS_LOCA_StepSizeControl.C): natural continuation in p
// from 4 to -4, conductance 10^p - 10^-4 on every voltage unknown, first
// step -2, grown by 1 + 0.01 ((20 - its) / 20)^2 after a converged solve,
// halved after a failed one (from the last converged solution), the last
// clipped onto -4 (conductance 0); at most 400 solves; junction
// initialisation on the first solve only.
static int set_gmin(xg_tran *T, double g) {
  if (hipMemcpyAsync(T->gmin_dev, &g, sizeof(double), hipMemcpyHostToDevice, T->stream) !=
          hipSuccess ||
      hipStreamSynchronize(T->stream) != hipSuccess)
    return tfail(T, XG_E_HIP, "gmin value");
  return XG_OK;
}

static int dc_begin(xg_tran *T) {
  const xg_tran_buffers &b = T->b;
  const int64_t n = std::max(T->n1, std::max(T->n_state, std::max(T->n_store,
                                                                  (int64_t)T->n_comp)));
  hipLaunchKernelGGL(xg::dc_begin_kernel, dim3(xg::grid_for(n)), dim3(xg::kBlock), 0,
                     T->stream, T->n1, T->n_state, T->n_store, T->n_comp, b.x, b.x_curr,
                     b.sta_c, b.sto_c, b.step, T->bp_tol, b.conv, b.active, b.ctl,
                     (const int32_t *)T->disabled);
  return hipGetLastError() == hipSuccess ? XG_OK : tfail(T, XG_E_HIP, "dc begin");
}

// One DC Newton solve from the iterate in b.x (control words and masks
// reset); 0 converged, > 0 the failure word, < 0 an error.
static int dc_solve(xg_tran *T, int init_jct, int32_t *iters) {
  const xg_tran_buffers &b = T->b;
  const int64_t n = std::max(T->n1, (int64_t)T->n_comp);
  // masks and control words as at a step start; the predictor of
  // tran_begin_kernel with beta = 0 leaves the iterate where it is
  hipLaunchKernelGGL(xg::tran_begin_kernel, dim3(xg::grid_for(n)), dim3(xg::kBlock), 0,
                     T->stream, T->n1, T->n_comp, 1, 0.0, 0.0, (const double *)b.x,
                     (const double *)b.xh1, (const double *)b.xh2, b.xpred, b.x, b.step, 0.0,
                     T->bp_tol, 0.0, b.conv, b.active, b.ctl, (const int32_t *)T->disabled,
                     nullptr);
  if (hipGetLastError() != hipSuccess) return tfail(T, XG_E_HIP, "dc restart");
  int rc = enqueue_first_load(T, 1, 0, init_jct);
  if (rc) return rc;
  *iters = 0;
  rc = newton_loop(T, 0, xg::kDcIters, iters);
  T->newton_total += *iters;
  return rc;
}

static int gmin_stepping(xg_tran *T) {
  const size_t bytes = sizeof(double) * (size_t)T->n1;
  if (!T->x_prev && hipMalloc((void **)&T->x_prev, bytes) != hipSuccess)
    return tfail(T, XG_E_NOMEM, "gmin stepping buffer");
  if (int rc = dc_begin(T)) return rc;
  if (hipMemcpyAsync(T->x_prev, T->b.x, bytes, hipMemcpyDeviceToDevice, T->stream) !=
      hipSuccess)
    return tfail(T, XG_E_HIP, "gmin stepping copy");
  const double p_end = -4.0;
  double p = 4.0, step = -2.0;
  bool first = true;
  int rc = 0;
  for (int n = 1;; ++n) {
    const double p_try = first ? p : std::max(p + step, p_end);
    if ((rc = set_gmin(T, std::pow(10.0, p_try) - std::pow(10.0, p_end) + 0.0))) break;
    if (hipMemcpyAsync(T->b.x, T->x_prev, bytes, hipMemcpyDeviceToDevice, T->stream) !=
        hipSuccess) {
      rc = tfail(T, XG_E_HIP, "gmin stepping copy");
      break;
    }
    int32_t iters = 0;
    rc = dc_solve(T, first ? 1 : 0, &iters);
    if (rc < 0) break;
    T->gmin_steps = n;
    if (first && rc > 0) break;
    if (rc == 0) {
      p = p_try;
      if (hipMemcpyAsync(T->x_prev, T->b.x, bytes, hipMemcpyDeviceToDevice, T->stream) !=
          hipSuccess) {
        rc = tfail(T, XG_E_HIP, "gmin stepping copy");
        break;
      }
      if (!first) {
        const double f = (20.0 - iters) / 20.0;
        step *= 1.0 + 0.01 * f * f;
        step = std::max(step, -1e12);
      }
      first = false;
      if (p == p_end) break;
    } else {
      step *= 0.5;
      if (std::fabs(step) < 1e-6) break;
    }
    if (n >= 400) {
      rc = 3;
      break;
    }
  }
  const int rc2 = set_gmin(T, 0.0);
  return rc ? rc : rc2;
}

int xg_tran_dcop_continuation(xg_tran *T, int32_t mode) {
  if (!T || mode < 0 || mode > 2) return XG_E_ARG;
  int rc;
  if (mode == 0) return xg_tran_dcop(T);
  if (mode == 2) {
    rc = xg_tran_dcop(T);
    if (rc <= 0) return rc;
  }
  rc = gmin_stepping(T);
  if (rc) return rc;
  rc = complete(T, 1);
  if (rc) return rc;
  T->t = 0.0;
  T->n_steps = 0;
  return XG_OK;
}

int xg_tran_gmin_steps(xg_tran *T, int64_t *out) {
  if (!T || !out) return XG_E_ARG;
  *out = T->gmin_steps;
  return XG_OK;
}

int xg_tran_uic(xg_tran *T) {
  if (!T) return XG_E_ARG;
  const xg_tran_buffers &b = T->b;
  const int64_t n = std::max(T->n1, std::max(T->n_state, std::max(T->n_store,
                                                                  (int64_t)T->n_comp)));
  hipLaunchKernelGGL(xg::dc_begin_kernel, dim3(xg::grid_for(n)), dim3(xg::kBlock), 0,
                     T->stream, T->n1, T->n_state, T->n_store, T->n_comp, b.x, b.x_curr,
                     b.sta_c, b.sto_c, b.step, T->bp_tol, b.conv, b.active, b.ctl,
                     (const int32_t *)T->disabled);
  if (hipGetLastError() != hipSuccess) return tfail(T, XG_E_HIP, "uic begin");
  // one load at x = 0 with the operating-point flags but no junction
  // initialisation, to start the charge and F history there
  set_state(T->h, 1, 0, 0, 1, 0);
  const int rc = xg_load_all(T->h, b.x, b.sta_c, b.sto_c, b.sta_n, b.sto_n, b.F, b.Q,
                             b.FL, b.QL, b.JF, b.JQ);
  if (rc) return rc;
  const int rc2 = complete(T, 1);
  if (rc2) return rc2;
  T->t = 0.0;
  T->n_steps = 0;
  return XG_OK;
}

// After a step's Newton loop (rc as newton_loop's, iters its iterations):
// the halving of a failed step, or the completion (unless the step's graph
// completed it) and the time advance of a converged one.
static int step_end(xg_tran *T, double dt, double t_next, int rc, int32_t iters,
                    bool completed) {
  T->newton_total += iters;
  if (rc < 0) return rc;
  if (rc > 0) {  // no convergence: two half steps
    if (dt / 2 < T->min_dt) {
      // eviction: the components that did not converge leave the solve,
      // at their last accepted values, and the step is taken again
      if (T->evict && evict_unconverged(T) > 0) return xg_tran_step(T, dt);
      return tfail(T, XG_E_STATE, "time step too small");
    }
    T->step_failures++;
    rc = xg_tran_step(T, dt / 2);
    if (rc) return rc;
    return xg_tran_step(T, dt / 2);
  }
  T->k_step = std::max(1, std::min<int>(iters, xg_tran::kMaxK));
  if (!completed) {  // converged after the step's graph: complete now
    rc = complete(T, 0);
    if (rc) return rc;
  }
  T->t = t_next;
  T->n_steps++;
  return XG_OK;
}

int xg_tran_step(xg_tran *T, double dt) {
  if (!T || !(dt > 0.0)) return XG_E_ARG;
  const int order = T->n_steps >= 2 ? 2 : 1;
  const int first = T->n_steps == 0;
  const double t_next = T->t + dt;
  // constant-step ONESTEP predictor coefficients (N_TIA_OneStep.C:1380-1405)
  const double b1 = order == 2 ? 1.5 : 1.0, b2 = order == 2 ? -0.5 : 0.0;
  const xg_tran_buffers &b = T->b;
  const int64_t n = std::max(T->n1, (int64_t)T->n_comp);
  hipLaunchKernelGGL(xg::tran_begin_kernel, dim3(xg::grid_for(n)), dim3(xg::kBlock), 0,
                     T->stream, T->n1, T->n_comp, order, b1, b2, (const double *)b.x_curr,
                     (const double *)b.xh1, (const double *)b.xh2, b.xpred, b.x, b.step,
                     t_next, T->bp_tol, 1.0 / dt, b.conv, b.active, b.ctl,
                     (const int32_t *)T->disabled, nullptr);
  if (hipGetLastError() != hipSuccess) return tfail(T, XG_E_HIP, "step begin");
  // first load, the first k_step iterations and (if converged) completion
  // in one launch; k_step follows the iterations the last step took
  int32_t iters = 0;
  const int k_first = T->use_graphs ? T->k_step : 1;
  int rc = newton_loop(T, first ? 1 : 2, xg::kTranIters, &iters, first ? 3 : 4, k_first);
  return step_end(T, dt, t_next, rc, iters, iters <= k_first);
}

int xg_tran_prepare_run(xg_tran *T) {
  if (!T) return XG_E_ARG;
  if (T->run_dev) return XG_OK;
  // dense-front solves (components too large for the task lists) take
  // milliseconds per iteration and hundreds of launches: their steps stay
  // one per launch
  if (xg::handle_n_fronts(T->h) > 0) return XG_OK;
  if (hipMalloc(&T->run_dev, sizeof(xg::RunDesc)) != hipSuccess ||
      hipHostMalloc(&T->run_in, sizeof(xg::RunDesc)) != hipSuccess ||
      hipHostMalloc(&T->run_out, sizeof(xg::RunDesc)) != hipSuccess)
    return XG_E_NOMEM;
  std::memset(T->run_in, 0, sizeof(xg::RunDesc));
  for (int K = 1; K <= xg_tran::kMaxK; ++K)
    if (int rc = capture_run(T, K)) return rc;
  return XG_OK;
}

int xg_tran_run(xg_tran *T, double dt, int64_t n_steps, const int32_t *ids, int32_t n_ids,
                int32_t every, int64_t first_step, double *out, double *times) {
  if (!T || !(dt > 0.0) || n_steps < 0 || n_ids < 0 || every < 1 || first_step < 0 ||
      (n_ids > 0 && (!ids || !out)))
    return XG_E_ARG;
  const int32_t *ctl = T->ctl_host;
  int64_t s = 0;
  while (s < n_steps) {
    const int K = T->k_step;
    if (T->use_graphs && T->run_dev && T->n_steps >= 2 && n_steps - s >= xg::kRunSteps &&
        T->run_exec[K]) {
      auto *in = (xg::RunDesc *)T->run_in;
      auto *res = (const xg::RunDesc *)T->run_out;
      in->t = T->t;
      in->dt = dt;
      in->inv_dt = 1.0 / dt;
      in->ids = ids;
      in->out = out;
      in->s0 = first_step + s;
      in->every = every;
      in->n_ids = n_ids;
      in->halt = 0;
      in->done = 0;
      T->host_reads++;
      if (hipMemcpyAsync(T->run_dev, in, sizeof(xg::RunDesc), hipMemcpyHostToDevice,
                         T->stream) != hipSuccess ||
          hipGraphLaunch(T->run_exec[K], T->stream) != hipSuccess ||
          hipMemcpyAsync(T->run_out, T->run_dev, sizeof(xg::RunDesc), hipMemcpyDeviceToHost,
                         T->stream) != hipSuccess ||
          hipMemcpyAsync(T->ctl_host, T->b.ctl, sizeof(int32_t) * xg::kCtlWords,
                         hipMemcpyDeviceToHost, T->stream) != hipSuccess ||
          hipStreamSynchronize(T->stream) != hipSuccess)
        return tfail(T, XG_E_HIP, "fused steps");
      for (int j = 0; j < res->done; ++j) {
        T->t = T->t + dt;
        T->n_steps++;
        T->newton_total += res->iters[j];
        T->k_step = std::max(1, std::min<int>(res->iters[j], xg_tran::kMaxK));
        ++s;
        if (times && (first_step + s) % every == 0) times[(first_step + s) / every - 1] = T->t;
      }
      if (!res->halt) continue;
      // the halted step: more iterations, or the failure's halving
      int32_t iters = res->iters[res->done];
      int rc = ctl[xg::kFail];
      if (!rc) {
        hipLaunchKernelGGL(xg::run_resume_kernel, dim3(1), dim3(64), 0, T->stream, T->b.ctl);
        if (hipGetLastError() != hipSuccess) return tfail(T, XG_E_HIP, "fused steps");
        rc = newton_loop(T, 2, xg::kTranIters, &iters);
      }
      if ((rc = step_end(T, dt, T->t + dt, rc, iters, false))) return rc;
    } else if (int rc = xg_tran_step(T, dt)) {
      return rc;
    }
    ++s;
    const int64_t g = first_step + s;
    if (g % every) continue;
    if (times) times[g / every - 1] = T->t;
    if (n_ids &&
        xg_tran_probe(T, ids, n_ids, out + (g / every - 1) * (int64_t)n_ids) != XG_OK)
      return tfail(T, XG_E_HIP, "probe");
  }
  return XG_OK;
}

int xg_tran_adaptive_begin(xg_tran *T, double tstop, double tstep, const double *bps,
                           int32_t n_bp, double reltol, double abstol) {
  if (!T || !(tstop > 0.0) || !(tstep > 0.0) || n_bp < 0 || (n_bp && !bps) ||
      !(reltol > 0.0) || !(abstol > 0.0))
    return XG_E_ARG;
  if (T->t != 0.0 || T->n_steps != 0)
    return tfail(T, XG_E_STATE, "adaptive run must start from xg_tran_dcop / xg_tran_uic");
  if (!T->qw) {
    if (hipMalloc((void **)&T->qw, sizeof(double) * (size_t)T->n1) != hipSuccess ||
        hipMalloc((void **)&T->lte_part, sizeof(double) * 3 * xg::kLteBlocks) != hipSuccess ||
        hipMalloc((void **)&T->lte_out, sizeof(double) * 4) != hipSuccess ||
        hipHostMalloc((void **)&T->lte_host, sizeof(double) * 4) != hipSuccess)
      return tfail(T, XG_E_NOMEM, "adaptive buffers");
  }
  Ad a;
  a.on = true;
  a.tstop = tstop;
  a.tstep = tstep;
  a.reltol = reltol;
  a.abstol = abstol;
  a.bps.push_back(0.0);
  for (int32_t i = 0; i < n_bp; ++i)
    if (bps[i] > 0.0 && bps[i] < tstop) a.bps.push_back(bps[i]);
  a.bps.push_back(tstop);
  std::sort(a.bps.begin(), a.bps.end());
  a.bps.erase(std::unique(a.bps.begin(), a.bps.end()), a.bps.end());
  a.h = tstep;
  a.last_h = tstep;
  a.h_sum = 2.0 * tstep;
  a.max_h = 0.1 * tstop;
  T->ad = a;
  return XG_OK;
}

int xg_tran_adaptive_step(xg_tran *T, double *t_out) {
  if (!T || !T->ad.on) return XG_E_ARG;
  Ad &a = T->ad;
  for (;;) {
    if (std::fabs(a.cur - a.tstop) < 1e-10 * a.tstop) break;
    ad_update_stop(a);
    if (a.beginning && a.attempt_ok)
      if (int rc = ad_initialize(T)) return rc;
    // updateCoeffs (N_TIA_OneStep.C:1355-1410)
    if (a.order == 2) a.psi[2] = a.psi[1];
    a.psi[1] = a.psi[0];
    a.psi[0] = a.h;
    double beta1, beta2 = 0.0, ck;
    if (a.order == 2) {
      const double t2 = a.psi[1];
      beta1 = a.h / t2 + (a.h / t2) * (a.h / t2) / 2;
      beta2 = -1.0 * a.h * a.h / t2 / a.psi[2] / 2;
      ck = (a.h / a.h_sum) / 3.0;
    } else {
      beta1 = a.h / a.psi[1];
      ck = a.h / a.h_sum;
    }
    bool newton_ok = false;
    if (int rc = ad_attempt(T, beta1, beta2, &newton_ok)) return rc;
    a.attempt_ok = newton_ok;
    if (newton_ok && a.step_number >= 1 && !a.beginning) {
      if (int rc = ad_estimate(T, ck, &a.est)) return rc;
      a.attempt_ok = a.est <= 1.0;
    }
    if (a.attempt_ok) {
      if (int rc = complete(T, 0)) return rc;
      ad_complete(a);
      T->t = a.cur;
      T->n_steps++;
      break;
    }
    if (int rc = ad_reject(T, newton_ok)) return rc;
  }
  if (t_out) *t_out = a.cur;
  return XG_OK;
}

int xg_tran_adaptive_info(xg_tran *T, double *h_est, int64_t *counts) {
  if (!T) return XG_E_ARG;
  const Ad &a = T->ad;
  if (h_est) {
    h_est[0] = a.h;
    h_est[1] = a.est;
    h_est[2] = a.stop;
  }
  if (counts) {
    counts[0] = a.accepted;
    counts[1] = a.rejected_error;
    counts[2] = a.rejected_newton;
    counts[3] = a.order;
  }
  return XG_OK;
}

int xg_tran_set_eviction(xg_tran *T, int32_t on, double min_dt) {
  if (!T || !(min_dt > 0.0)) return XG_E_ARG;
  T->evict = on != 0;
  T->min_dt = min_dt;
  return XG_OK;
}

int xg_tran_evicted(xg_tran *T, int32_t *flags, int64_t *n) {
  if (!T) return XG_E_ARG;
  if (n) *n = T->n_evicted;
  if (flags && T->n_comp > 0 &&
      (hipMemcpyAsync(flags, T->disabled, sizeof(int32_t) * T->n_comp, hipMemcpyDeviceToHost,
                      T->stream) != hipSuccess ||
       hipStreamSynchronize(T->stream) != hipSuccess))
    return XG_E_HIP;
  return XG_OK;
}

int xg_tran_probe(xg_tran *T, const int32_t *ids, int32_t n, double *out) {
  if (!T || (n > 0 && (!ids || !out)) || n < 0) return XG_E_ARG;
  if (n == 0) return XG_OK;
  hipLaunchKernelGGL(xg::probe_kernel, dim3((n + 63) / 64), dim3(64), 0, T->stream,
                     (const double *)T->b.x_curr, T->n1 - 1, ids, n, out);
  return hipGetLastError() == hipSuccess ? XG_OK : XG_E_HIP;
}

int xg_tran_info(xg_tran *T, double *t, int64_t *counters) {
  if (!T) return XG_E_ARG;
  if (t) *t = T->t;
  if (counters) {
    counters[0] = T->n_steps;
    counters[1] = T->newton_total;
    counters[2] = T->step_failures;
    counters[3] = T->ctl_host ? T->ctl_host[xg::kFail] : 0;
  }
  return XG_OK;
}

int xg_tran_counters(xg_tran *T, int64_t *out) {
  if (!T || !out) return XG_E_ARG;
  out[0] = T->host_reads;
  out[1] = T->n_steps;
  out[2] = T->newton_total;
  out[3] = T->k_step;
  return XG_OK;
}

int xg_tran_restore(xg_tran *T, double time, int64_t n_steps, int64_t newton_total,
                    int64_t step_failures) {
  if (!T || n_steps < 0) return XG_E_ARG;
  T->t = time;
  T->n_steps = n_steps;
  T->newton_total = newton_total;
  T->step_failures = step_failures;
  return XG_OK;
}

int xg_tran_sync(xg_tran *T) {
  if (!T) return XG_E_ARG;
  return hipStreamSynchronize(T->stream) == hipSuccess ? XG_OK : XG_E_HIP;
}

int xg_tran_destroy(xg_tran *T) {
  if (!T) return XG_OK;
  (void)hipStreamSynchronize(T->stream);
  for (int v = 0; v < 5; ++v)
    for (int K = 0; K <= xg_tran::kMaxK; ++K) {
      if (T->exec[v][K]) (void)hipGraphExecDestroy(T->exec[v][K]);
      if (T->graph[v][K]) (void)hipGraphDestroy(T->graph[v][K]);
    }
  for (int K = 0; K <= xg_tran::kMaxK; ++K) {
    if (T->run_exec[K]) (void)hipGraphExecDestroy(T->run_exec[K]);
    if (T->run_graph[K]) (void)hipGraphDestroy(T->run_graph[K]);
  }
  if (T->run_dev) (void)hipFree(T->run_dev);
  if (T->run_in) (void)hipHostFree(T->run_in);
  if (T->run_out) (void)hipHostFree(T->run_out);
  xg_set_step_source(T->h, nullptr);
  xg_set_active_mask(T->h, nullptr);
  xg_set_stream(T->h, T->old_stream);
  if (T->ctl_host) (void)hipHostFree(T->ctl_host);
  if (T->lte_host) (void)hipHostFree(T->lte_host);
  if (T->qw) (void)hipFree(T->qw);
  if (T->lte_part) (void)hipFree(T->lte_part);
  if (T->lte_out) (void)hipFree(T->lte_out);
  if (T->gmin_dev) (void)hipFree(T->gmin_dev);
  if (T->gv_ids) (void)hipFree(T->gv_ids);
  if (T->gv_pos) (void)hipFree(T->gv_pos);
  if (T->x_prev) (void)hipFree(T->x_prev);
  if (T->disabled) (void)hipFree(T->disabled);
  if (T->comp_of_row) (void)hipFree(T->comp_of_row);
  (void)hipStreamDestroy(T->stream);
  delete T;
  return XG_OK;
}

}  // extern "C"
