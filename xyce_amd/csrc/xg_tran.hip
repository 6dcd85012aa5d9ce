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
#include <cstring>
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
  std::string err;
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

// Step start: predictor (the rounding sequence of the predictor kernel),
// iterate = predictor, step parameters, control words and masks reset.
__global__ void __launch_bounds__(kBlock)
    tran_begin_kernel(int64_t n1, int32_t n_comp, int32_t order, double beta1,
                      double beta2, const double *x_curr, const double *xh1,
                      const double *xh2, double *xpred, double *x, double *step,
                      double t, double bp_tol, double inv_dt, int32_t *conv,
                      int32_t *active, int32_t *ctl) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i < n1) {
    double v = x_curr[i];
    v = beta1 * xh1[i] + 1.0 * v;
    if (order == 2) v = beta2 * xh2[i] + 1.0 * v;
    xpred[i] = v;
    x[i] = v;
  }
  if (i < n_comp) {
    conv[i] = 0;
    active[i] = 1;
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
                    int32_t *ctl) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i < n1) {
    x[i] = 0.0;
    x_curr[i] = 0.0;
  }
  if (i < n_state) sta_c[i] = 0.0;
  if (i < n_store) sto_c[i] = 0.0;
  if (i < n_comp) {
    conv[i] = 0;
    active[i] = 1;
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
                      int32_t *active, int32_t *ctl) {
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

// The first load of a Newton solve and its assembly.
int enqueue_first_load(xg_tran *T, int dc, int first) {
  const xg_tran_buffers &b = T->b;
  set_state(T->h, dc, dc, first, 0, !dc);
  int rc = xg_load_all(T->h, b.x, b.sta_c, b.sto_c, b.sta_n, b.sto_n, b.F, b.Q, b.FL,
                       b.QL, b.JF, b.JQ);
  if (rc) return rc;
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
                     (const int64_t *)b.counts, prm, b.conv, b.active, b.ctl);
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
                     b.sta_c, b.sto_c, b.step, T->bp_tol, b.conv, b.active, b.ctl);
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

int xg_tran_uic(xg_tran *T) {
  if (!T) return XG_E_ARG;
  const xg_tran_buffers &b = T->b;
  const int64_t n = std::max(T->n1, std::max(T->n_state, std::max(T->n_store,
                                                                  (int64_t)T->n_comp)));
  hipLaunchKernelGGL(xg::dc_begin_kernel, dim3(xg::grid_for(n)), dim3(xg::kBlock), 0,
                     T->stream, T->n1, T->n_state, T->n_store, T->n_comp, b.x, b.x_curr,
                     b.sta_c, b.sto_c, b.step, T->bp_tol, b.conv, b.active, b.ctl);
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
                     t_next, T->bp_tol, 1.0 / dt, b.conv, b.active, b.ctl);
  if (hipGetLastError() != hipSuccess) return tfail(T, XG_E_HIP, "step begin");
  // first load, the first k_step iterations and (if converged) completion
  // in one launch; k_step follows the iterations the last step took
  int32_t iters = 0;
  const int k_first = T->use_graphs ? T->k_step : 1;
  int rc = newton_loop(T, first ? 1 : 2, xg::kTranIters, &iters, first ? 3 : 4, k_first);
  T->newton_total += iters;
  if (rc < 0) return rc;
  if (rc > 0) {  // no convergence: two half steps
    if (dt < 1e-18) return tfail(T, XG_E_STATE, "time step too small");
    T->step_failures++;
    rc = xg_tran_step(T, dt / 2);
    if (rc) return rc;
    return xg_tran_step(T, dt / 2);
  }
  T->k_step = std::max(1, std::min<int>(iters, xg_tran::kMaxK));
  if (iters > k_first) {  // converged after the fused graph: complete now
    rc = complete(T, 0);
    if (rc) return rc;
  }
  T->t = t_next;
  T->n_steps++;
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
  xg_set_step_source(T->h, nullptr);
  xg_set_active_mask(T->h, nullptr);
  xg_set_stream(T->h, T->old_stream);
  if (T->ctl_host) (void)hipHostFree(T->ctl_host);
  (void)hipStreamDestroy(T->stream);
  delete T;
  return XG_OK;
}

}  // extern "C"
