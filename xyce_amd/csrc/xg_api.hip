// C ABI of libxg_load.so: circuit setup, device-resident handles and the
// per-Newton load entry points.  See include/xg_load.h for the contract.
#include <cstdio>
#include <cstring>
#include <algorithm>
#include <memory>
#include <new>
#include <stdexcept>
#include <cmath>
#include <map>
#include <mutex>
#include <tuple>
#include <string>

#include "xg_bsim3.h"
#include "xg_internal.h"
#include "xg_setup.h"

struct xg_circuit {
  xg::Circuit c;
};

namespace xg {

struct DevBuf {
  void *p = nullptr;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
};

struct Handle {
  Plan plan;
  hipStream_t stream = nullptr;
  SolverStateDev ss{};
  std::vector<std::unique_ptr<DevBuf>> bufs;
  // MOSFET1 device data
  Mos1Args m1{};
  // BSIM3 device data
  Bsim3Dev b3{};
  // linear devices and the time their sources are evaluated at
  LinArgs lin{};
  double time = 0.0, bp_tol = 0.0;
  const double *dstep = nullptr;  // device {time, bp_tol, 1/dt, order}
  const int32_t *active = nullptr;  // component activity (xg_set_active_mask)
  // per-component dense solve, and the static-pivot sparse LU
  DenseSolveArgs ds{};
  SparseLuArgs lu{};
  FrontArgs fr{};
  std::vector<int32_t> fr_lv_ptr, fr_lv_maxm;  // host copies for the launches
  // fronts of more than fr.big_m rows by level (host, offsets, device)
  std::vector<int32_t> fr_big, fr_big_ptr;
  int32_t *fr_big_dev = nullptr;
  std::vector<int64_t> fr_xr_off;  // receiving rows per front (host copy)
  int32_t *d_rowptr = nullptr, *d_colidx = nullptr;
  int32_t *slot_comp = nullptr;
  int32_t *node_comp = nullptr;  // node -> component (ground excluded)
  // reduction programs
  ReduceDev red_vec[4]{}, red_mat[2]{};
  double *slots = nullptr;
  double *partial = nullptr;  // chunk partial sums of the reduction
  int64_t *count_buf = nullptr;
  int32_t *limited_all = nullptr;  // m1 then b3 eval slots
  int32_t n_limited_slots = 0;
  // the *_host variants: device mirrors of the caller's host vectors and
  // one pinned staging buffer, allocated by their first call
  double *pinned = nullptr;
  double *hx = nullptr, *hsta_c = nullptr, *hsto_c = nullptr, *hsta_n = nullptr,
         *hsto_n = nullptr, *hvec = nullptr, *hmat = nullptr;
  int64_t pinned_len = 0;
  bool plan_only = false;
  bool overwrite = false;  // XG_OUT_OVERWRITE
  // stop word of the transient driver's multi-iteration graphs: set only
  // while the driver enqueues (its launches then skip once it is nonzero)
  const int32_t *gate = nullptr;
  // autotuned launch choices (result-neutral): linear devices in trailing
  // workgroups of the BSIM3 launch or their own launch
  bool fuse_linear = true;
  int32_t tune_source = 0;   // 0 not tuned, 1 timed at upload, 2 cached
  double tune_us[6] = {};    // per candidate, NaN where not timed
  int device = -1;         // the GPU the handle's buffers live on

  template <class T>
  T *alloc(size_t n) {
    auto b = std::make_unique<DevBuf>();
    if (n == 0) n = 1;
    if (hipMalloc(&b->p, n * sizeof(T)) != hipSuccess) return nullptr;
    T *r = (T *)b->p;
    bufs.push_back(std::move(b));
    return r;
  }
  template <class T>
  T *upload(const std::vector<T> &v) {
    T *d = alloc<T>(v.size());
    if (!d) return nullptr;
    if (!v.empty() &&
        hipMemcpy(d, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice) !=
            hipSuccess)
      return nullptr;
    return d;
  }
  ~Handle() {
    if (pinned) (void)hipHostFree(pinned);
  }
};

}  // namespace xg

struct xg_handle {
  xg::Handle h;
};

namespace {
thread_local std::string g_err;

int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

int hip_fail(hipError_t e, const char *where) {
  g_err = std::string(where) + ": " + hipGetErrorString(e);
  return XG_E_HIP;
}
}  // namespace

extern "C" {

const char *xg_last_error(void) { return g_err.c_str(); }

int xg_circuit_new(xg_circuit **c, int32_t n_nodes, int32_t n_state,
                   int32_t n_store) {
  if (!c || n_nodes < 0 || n_state < 0 || n_store < 0)
    return fail(XG_E_ARG, "xg_circuit_new: bad argument");
  auto *p = new (std::nothrow) xg_circuit;
  if (!p) return fail(XG_E_NOMEM, "xg_circuit_new: out of memory");
  p->c.n_nodes = n_nodes;
  p->c.n_state = n_state;
  p->c.n_store = n_store;
  *c = p;
  return XG_OK;
}

void xg_circuit_free(xg_circuit *c) { delete c; }

int xg_circuit_set_lu_mode(xg_circuit *c, int32_t mode) {
  if (!c || (mode != XG_LU_AUTO && mode != XG_LU_FRONTS && mode != XG_LU_TASKS))
    return fail(XG_E_ARG, "xg_circuit_set_lu_mode: bad argument");
  c->c.opt.lu_mode = mode;
  return XG_OK;
}

int xg_circuit_set_options(xg_circuit *c, double gmin, int32_t limiter,
                           double temp, double tnom) {
  if (!c || !(temp > 0) || !(tnom > 0) || !(gmin >= 0))
    return fail(XG_E_ARG, "xg_circuit_set_options: bad argument");
  if (!c->c.mos1_models.empty() || !c->c.b3_models.empty())
    return fail(XG_E_STATE, "xg_circuit_set_options: set options before models");
  c->c.opt.gmin = gmin;
  c->c.opt.limiter = limiter ? 1 : 0;
  c->c.opt.temp = temp;
  c->c.opt.tnom = tnom;
  return XG_OK;
}

int xg_add_mos1_model(xg_circuit *c, int32_t dtype, int32_t n,
                      const char *const *keys, const double *vals,
                      int32_t *model_id) {
  if (!c || (dtype != 1 && dtype != -1) || !model_id)
    return fail(XG_E_ARG, "xg_add_mos1_model: bad argument");
  xg::Mos1Model m;
  m.dtype = dtype;
  if (xg::parse_params(n, keys, vals, m.p))
    return fail(XG_E_ARG, "xg_add_mos1_model: bad parameter list");
  std::string err;
  int rc = xg::mos1_model_process(m, c->c.opt, err);
  if (rc) return fail(rc, err);
  *model_id = (int32_t)c->c.mos1_models.size();
  c->c.mos1_models.push_back(std::move(m));
  return XG_OK;
}

// Solution ids must address the circuit's nodes (or be negative = ground),
// and the instance's state / store slots must lie inside the vectors: the
// kernels index x, state and store with these values unchecked.
static int check_ids(const xg::Circuit &c, const int32_t lids[6],
                     int32_t sta_base, int32_t n_sta, int32_t sto_base,
                     int32_t n_sto, const char *who) {
  for (int k = 0; k < 6; ++k)
    if (lids[k] >= c.n_nodes)
      return fail(XG_E_ARG, std::string(who) + ": node id out of range");
  if (sta_base < 0 || (int64_t)sta_base + n_sta > c.n_state ||
      sto_base < 0 || (int64_t)sto_base + n_sto > c.n_store)
    return fail(XG_E_ARG, std::string(who) + ": state/store slots out of range");
  return XG_OK;
}

int xg_add_mos1_instance(xg_circuit *c, int32_t model_id, int32_t n,
                         const char *const *keys, const double *vals,
                         const int32_t lids[6], int32_t sta_base,
                         int32_t sto_base, int32_t *inst_id) {
  if (!c || !lids || !inst_id || model_id < 0 ||
      model_id >= (int32_t)c->c.mos1_models.size())
    return fail(XG_E_ARG, "xg_add_mos1_instance: bad argument");
  if (int rc = check_ids(c->c, lids, sta_base, 8, sto_base, 5,
                         "xg_add_mos1_instance"))
    return rc;
  xg::Mos1Inst in;
  in.model = model_id;
  if (xg::parse_params(n, keys, vals, in.p))
    return fail(XG_E_ARG, "xg_add_mos1_instance: bad parameter list");
  for (int k = 0; k < 6; ++k) in.lids[k] = lids[k] < 0 ? -1 : lids[k];
  in.sta_base = sta_base;
  in.sto_base = sto_base;
  std::string err;
  int rc = xg::mos1_instance_setup(c->c.mos1_models[model_id], in, c->c.opt,
                                   err);
  if (rc) return fail(rc, err);
  *inst_id = (int32_t)c->c.mos1.size();
  c->c.mos1.push_back(std::move(in));
  return XG_OK;
}

int xg_add_bsim3_model(xg_circuit *c, int32_t dtype, int32_t n,
                       const char *const *keys, const double *vals,
                       int32_t *model_id) {
  if (!c || (dtype != 1 && dtype != -1) || !model_id)
    return fail(XG_E_ARG, "xg_add_bsim3_model: bad argument");
  xg::ParamSet p;
  if (xg::parse_params(n, keys, vals, p))
    return fail(XG_E_ARG, "xg_add_bsim3_model: bad parameter list");
  std::string err;
  int rc = xg::bsim3_add_model(c->c, dtype, p, model_id, err);
  return rc ? fail(rc, err) : XG_OK;
}

int xg_add_bsim3_instance(xg_circuit *c, int32_t model_id, int32_t n,
                          const char *const *keys, const double *vals,
                          const int32_t lids[6], int32_t sta_base,
                          int32_t sto_base, int32_t *inst_id) {
  if (!c || !lids || !inst_id)
    return fail(XG_E_ARG, "xg_add_bsim3_instance: bad argument");
  if (int rc = check_ids(c->c, lids, sta_base, 12, sto_base, 5,
                         "xg_add_bsim3_instance"))
    return rc;
  xg::ParamSet p;
  if (xg::parse_params(n, keys, vals, p))
    return fail(XG_E_ARG, "xg_add_bsim3_instance: bad parameter list");
  std::string err;
  int rc = xg::bsim3_add_instance(c->c, model_id, p, lids, sta_base, sto_base,
                                  inst_id, err);
  return rc ? fail(rc, err) : XG_OK;
}

// n instances per call, as that many single calls in array order
typedef int (*AddOne)(xg_circuit *, int32_t, int32_t, const char *const *, const double *,
                      const int32_t *, int32_t, int32_t, int32_t *);
static int add_instances(AddOne one, xg_circuit *c, const int32_t *model_ids, int32_t count,
                         int32_t n_keys, const char *const *keys, const double *vals,
                         const int32_t *lids, const int32_t *sta_base, const int32_t *sto_base,
                         int32_t *inst_ids, const char *what) {
  if (!c || count < 0 || n_keys < 0 ||
      (count > 0 && (!model_ids || !lids || !sta_base || !sto_base || !inst_ids ||
                     (n_keys > 0 && (!keys || !vals)))))
    return fail(XG_E_ARG, std::string(what) + ": bad argument");
  for (int32_t i = 0; i < count; ++i)
    if (int rc = one(c, model_ids[i], n_keys, keys, vals + (size_t)i * n_keys,
                     lids + 6 * (size_t)i, sta_base[i], sto_base[i], inst_ids + i))
      return rc;
  return XG_OK;
}

int xg_add_mos1_instances(xg_circuit *c, const int32_t *model_ids, int32_t count, int32_t n_keys,
                          const char *const *keys, const double *vals, const int32_t *lids,
                          const int32_t *sta_base, const int32_t *sto_base,
                          int32_t *inst_ids) {
  return add_instances(xg_add_mos1_instance, c, model_ids, count, n_keys, keys, vals, lids,
                       sta_base, sto_base, inst_ids, "xg_add_mos1_instances");
}

int xg_add_bsim3_instances(xg_circuit *c, const int32_t *model_ids, int32_t count, int32_t n_keys,
                           const char *const *keys, const double *vals, const int32_t *lids,
                           const int32_t *sta_base, const int32_t *sto_base,
                           int32_t *inst_ids) {
  return add_instances(xg_add_bsim3_instance, c, model_ids, count, n_keys, keys, vals, lids,
                       sta_base, sto_base, inst_ids, "xg_add_bsim3_instances");
}

static int add_lin(xg_circuit *c, int32_t type, int32_t pos, int32_t neg,
                   int32_t br, double val, const double *pulse, int32_t sta,
                   const char *who) {
  if (!c) return fail(XG_E_ARG, std::string(who) + ": bad argument");
  const int32_t n = c->c.n_nodes;
  if (pos >= n || neg >= n || br >= n || (type == xg::LIN_V && br < 0))
    return fail(XG_E_ARG, std::string(who) + ": node id out of range");
  if (type == xg::LIN_C && (sta < 0 || sta >= c->c.n_state))
    return fail(XG_E_ARG, std::string(who) + ": state slot out of range");
  if (type == xg::LIN_R && val == 0.0)
    return fail(XG_E_PARAM, std::string(who) + ": zero resistance");
  xg::LinDev d{};
  d.type = type;
  d.lids[0] = pos < 0 ? -1 : pos;
  d.lids[1] = neg < 0 ? -1 : neg;
  d.lids[2] = br < 0 ? -1 : br;
  d.val = val;
  d.sta = sta;
  if (pulse) {
    d.pulse[0] = 1.0;
    for (int k = 0; k < 7; ++k) d.pulse[k + 1] = pulse[k];
  }
  c->c.lin.push_back(d);
  return XG_OK;
}

int xg_add_resistor(xg_circuit *c, int32_t pos, int32_t neg, double r) {
  return add_lin(c, xg::LIN_R, pos, neg, -1, r, nullptr, 0, "xg_add_resistor");
}

int xg_add_capacitor(xg_circuit *c, int32_t pos, int32_t neg, double cap,
                     int32_t sta) {
  return add_lin(c, xg::LIN_C, pos, neg, -1, cap, nullptr, sta,
                 "xg_add_capacitor");
}

int xg_add_resistors(xg_circuit *c, int32_t n, const int32_t *pos,
                     const int32_t *neg, const double *r) {
  if (!c || n < 0 || (n > 0 && (!pos || !neg || !r)))
    return fail(XG_E_ARG, "xg_add_resistors: bad argument");
  if (c->c.lin.size() + (size_t)n > c->c.lin.capacity())
    c->c.lin.reserve(std::max(c->c.lin.size() + (size_t)n, 2 * c->c.lin.capacity()));
  for (int32_t i = 0; i < n; ++i)
    if (int rc = add_lin(c, xg::LIN_R, pos[i], neg[i], -1, r[i], nullptr, 0,
                         "xg_add_resistors")) {
      g_err += " (device " + std::to_string(i) + ")";
      return rc;
    }
  return XG_OK;
}

int xg_add_capacitors(xg_circuit *c, int32_t n, const int32_t *pos,
                      const int32_t *neg, const double *cap, const int32_t *sta) {
  if (!c || n < 0 || (n > 0 && (!pos || !neg || !cap || !sta)))
    return fail(XG_E_ARG, "xg_add_capacitors: bad argument");
  if (c->c.lin.size() + (size_t)n > c->c.lin.capacity())
    c->c.lin.reserve(std::max(c->c.lin.size() + (size_t)n, 2 * c->c.lin.capacity()));
  for (int32_t i = 0; i < n; ++i)
    if (int rc = add_lin(c, xg::LIN_C, pos[i], neg[i], -1, cap[i], nullptr, sta[i],
                         "xg_add_capacitors")) {
      g_err += " (device " + std::to_string(i) + ")";
      return rc;
    }
  return XG_OK;
}

int xg_add_vsource(xg_circuit *c, int32_t pos, int32_t neg, int32_t branch,
                   double dc, const double *pulse) {
  return add_lin(c, xg::LIN_V, pos, neg, branch, dc, pulse, 0, "xg_add_vsource");
}

int xg_add_diode(xg_circuit *c, int32_t pos, int32_t neg, int32_t n_model,
                 const char *const *mkeys, const double *mvals, int32_t n_inst,
                 const char *const *ikeys, const double *ivals, int32_t sto) {
  if (!c) return fail(XG_E_ARG, "xg_add_diode: bad argument");
  if (sto < 0 || sto >= c->c.n_store)
    return fail(XG_E_ARG, "xg_add_diode: store slot out of range");
  xg::ParamSet mp, ip;
  if (xg::parse_params(n_model, mkeys, mvals, mp) ||
      xg::parse_params(n_inst, ikeys, ivals, ip))
    return fail(XG_E_ARG, "xg_add_diode: bad parameter list");
  double f[xg::DIO_N];
  std::string err;
  int rc = xg::diode_setup(mp, ip, c->c.opt, f, err);
  if (rc) return fail(rc, "xg_add_diode: " + err);
  rc = add_lin(c, xg::LIN_D, pos, neg, -1, 0.0, nullptr, 0, "xg_add_diode");
  if (rc) return rc;
  xg::LinDev &d = c->c.lin.back();
  d.sta = sto;
  for (int k = 0; k < xg::DIO_N; ++k) d.dpar[k] = f[k];
  return XG_OK;
}

int xg_components(const xg_handle *H, int32_t *n_comp, int32_t *comp_of_node) {
  if (!H || !n_comp) return fail(XG_E_ARG, "xg_components: bad argument");
  const xg::Plan &P = H->h.plan;
  *n_comp = (int32_t)P.comp_start.size() - 1;
  if (comp_of_node)
    for (int32_t k = 0; k < *n_comp; ++k)
      for (int32_t j = P.comp_start[k]; j < P.comp_start[k + 1]; ++j)
        comp_of_node[P.comp_nodes[j]] = k;
  return XG_OK;
}

// Assembly with the per-component max |rhs| folded in when norms is given.
static int tia_assemble(xg_handle *H, int32_t transient, int32_t order, double dt,
                        const double *qn0, const double *fprev, const double *F,
                        const double *Q, const double *dFdxdVp, const double *dQdxdVp,
                        const double *dFdx, const double *dQdx, double *rhs, double *J,
                        double *norms) {
  if (H && H->h.plan_only)
    return fail(XG_E_STATE, "handle was finalized without a device");
  if (!H || !F || !dFdxdVp || !dFdx || !rhs || !J ||
      (transient && (!qn0 || !Q || !dQdxdVp || !dQdx ||
                     (!(dt > 0.0) && !H->h.dstep) ||
                     ((order == 2 || H->h.dstep) && !fprev))) ||
      (order != 1 && order != 2))
    return fail(XG_E_ARG, "xg_tia_assemble: bad argument");
  xg::Handle &h = H->h;
  xg::TiaAssembleArgs a{};
  a.n1 = h.plan.n_nodes + 1;
  a.nnz = (int64_t)h.plan.colidx.size();
  a.transient = transient != 0;
  a.order = order;
  a.limiter = h.plan.limiter;
  a.inv_dt = transient ? 1.0 / dt : 0.0;
  a.dstep = h.dstep;
  a.gate = h.gate;
  a.qn0 = qn0;
  a.fprev = fprev ? fprev : F;
  a.F = F;
  a.Q = Q ? Q : F;
  a.FL = dFdxdVp;
  a.QL = dQdxdVp ? dQdxdVp : dFdxdVp;
  a.dFdx = dFdx;
  a.dQdx = dQdx ? dQdx : dFdx;
  a.rhs = rhs;
  a.J = J;
  a.node_comp = h.node_comp;
  a.norms = norms;
  hipError_t e = xg::launch_tia_assemble(a, h.stream);
  return e == hipSuccess ? XG_OK : hip_fail(e, "assemble");
}

int xg_tia_assemble(xg_handle *H, int32_t transient, int32_t order, double dt,
                    const double *qn0, const double *fprev, const double *F,
                    const double *Q, const double *dFdxdVp,
                    const double *dQdxdVp, const double *dFdx,
                    const double *dQdx, double *rhs, double *J) {
  return tia_assemble(H, transient, order, dt, qn0, fprev, F, Q, dFdxdVp, dQdxdVp, dFdx,
                      dQdx, rhs, J, nullptr);
}

int xg_dense_solve(xg_handle *H, const double *J, const double *rhs, double *x,
                   const double *wa, const double *wb, double reltol,
                   double abstol, const int32_t *active, double *stats) {
  if (H && H->h.plan_only)
    return fail(XG_E_STATE, "handle was finalized without a device");
  if (!H || !J || !rhs || !x || !wa || !wb || !stats)
    return fail(XG_E_ARG, "xg_dense_solve: bad argument");
  xg::Handle &h = H->h;
  xg::DenseSolveArgs a = h.ds;
  a.J = J;
  a.rhs = rhs;
  a.x = x;
  a.wa = wa;
  a.wb = wb;
  a.reltol = reltol;
  a.abstol = abstol;
  a.active = active;
  a.stats = stats;
  hipError_t e = xg::launch_dense_solve(a, h.stream);
  return e == hipSuccess ? XG_OK : hip_fail(e, "dense solve");
}

int xg_lu_solve(xg_handle *H, const double *J, const double *rhs, double *x,
                const double *wa, const double *wb, double reltol, double abstol,
                const int32_t *active, double *stats) {
  if (H && H->h.plan_only)
    return fail(XG_E_STATE, "handle was finalized without a device");
  if (!H || !J || !rhs || !x || !wa || !wb || !stats)
    return fail(XG_E_ARG, "xg_lu_solve: bad argument");
  xg::Handle &h = H->h;
  xg::SparseLuArgs u = h.lu;
  u.J = J;
  u.rhs = rhs;
  u.x = x;
  u.wa = wa;
  u.wb = wb;
  u.reltol = reltol;
  u.abstol = abstol;
  u.active = active;
  u.stats = stats;
  u.gate = h.gate;
  hipError_t e = xg::launch_sparse_lu(u, h.stream);
  if (e != hipSuccess) return hip_fail(e, "sparse LU");
  if (h.fr.n_fronts > 0) {
    xg::FrontArgs f = h.fr;
    f.active = active;
    f.gate = h.gate;
    f.lv_ptr_host = h.fr_lv_ptr.data();
    f.lv_maxm_host = h.fr_lv_maxm.data();
    const xg::LuPlan &L = h.plan.lu;
    f.lv_host = L.fr_lv.data();
    f.m_host = L.fr_m.data();
    f.k_host = L.fr_k.data();
    f.ch_ptr_host = L.fr_ch_ptr.data();
    f.ch_host = L.fr_ch.data();
    f.big = h.fr_big_dev;
    f.big_host = h.fr_big.data();
    f.big_ptr_host = h.fr_big_ptr.data();
    f.xr_off_host = h.fr_xr_off.data();
    e = xg::launch_fronts(f, u, h.stream);
    if (e != hipSuccess) return hip_fail(e, "front LU");
  }
  // components whose static pivots fail the threshold test: dense LU
  // with partial pivoting, in the same stream, no host round trip
  xg::DenseSolveArgs a = h.ds;
  a.J = J;
  a.rhs = rhs;
  a.x = x;
  a.wa = wa;
  a.wb = wb;
  a.reltol = reltol;
  a.abstol = abstol;
  a.active = h.lu.fallback;
  a.only_flagged = 1;
  a.stats = stats;
  a.gate = h.gate;
  e = xg::launch_dense_solve(a, h.stream);
  return e == hipSuccess ? XG_OK : hip_fail(e, "pivoting LU");
}

int xg_lu_info(xg_handle *H, int64_t *info) {
  if (!H || !info) return fail(XG_E_ARG, "xg_lu_info: bad argument");
  xg::Handle &h = H->h;
  const xg::LuPlan &L = h.plan.lu;
  int64_t n_global = 0;
  const int32_t nc = (int32_t)h.plan.comp_start.size() - 1;
  for (int32_t k = 0; k < nc; ++k) n_global += L.in_lds[k] ? 0 : 1;
  info[0] = L.fnnz;
  info[1] = L.factor_ops;
  info[2] = L.max_factor_levels;
  info[3] = L.max_solve_levels;
  info[4] = L.n_static_fail;
  info[5] = n_global;
  info[6] = -1;
  if (!h.plan_only && nc > 0) {
    std::vector<int32_t> f(nc);
    if (hipStreamSynchronize(h.stream) != hipSuccess ||
        hipMemcpy(f.data(), h.lu.fallback, sizeof(int32_t) * nc, hipMemcpyDeviceToHost) !=
            hipSuccess)
      return fail(XG_E_HIP, "xg_lu_info: copy failed");
    info[6] = 0;
    for (int32_t v : f) info[6] += v != 0;
  } else if (nc == 0) {
    info[6] = 0;
  }
  info[7] = (int64_t)L.t_dst.size();
  return XG_OK;
}

// The fronts of more than big_m rows grouped by level, on the host and (once
// uploaded) on the device.
static int front_big_lists(xg::Handle &h) {
  const xg::LuPlan &L = h.plan.lu;
  h.fr_big.clear();
  h.fr_big_ptr.assign(1, 0);
  for (size_t l = 0; l + 1 < L.fr_lv_ptr.size(); ++l) {
    for (int32_t q = L.fr_lv_ptr[l]; q < L.fr_lv_ptr[l + 1]; ++q)
      if (L.fr_m[L.fr_lv[q]] > h.fr.big_m) h.fr_big.push_back(L.fr_lv[q]);
    h.fr_big_ptr.push_back((int32_t)h.fr_big.size());
  }
  if (h.fr_big_dev && !h.fr_big.empty() &&
      hipMemcpy(h.fr_big_dev, h.fr_big.data(), sizeof(int32_t) * h.fr_big.size(),
                hipMemcpyHostToDevice) != hipSuccess)
    return XG_E_HIP;
  return XG_OK;
}

int xg_lu_cache_stats(int64_t *out) {
  if (!out) return fail(XG_E_ARG, "xg_lu_cache_stats: bad argument");
  xg::lu_cache_stats(out);
  return XG_OK;
}

int xg_lu_cache_set_limit(int64_t bytes) {
  if (bytes < 0) return fail(XG_E_ARG, "xg_lu_cache_set_limit: negative limit");
  xg::lu_cache_set_limit(bytes);
  return XG_OK;
}

int xg_lu_front_set_big(xg_handle *H, int32_t big_m) {
  if (!H || big_m < 1) return fail(XG_E_ARG, "xg_lu_front_set_big");
  H->h.fr.big_m = big_m;
  if (front_big_lists(H->h)) return fail(XG_E_HIP, "xg_lu_front_set_big: upload");
  return XG_OK;
}

int xg_lu_front_info(const xg_handle *H, int64_t *info) {
  if (!H || !info) return fail(XG_E_ARG, "xg_lu_front_info: bad argument");
  const xg::LuPlan &L = H->h.plan.lu;
  info[0] = L.n_front_comp;
  info[1] = (int64_t)L.fr_m.size();
  info[2] = L.fr_lv_ptr.empty() ? 0 : (int64_t)L.fr_lv_ptr.size() - 1;
  info[3] = L.fr_fo.empty() ? 0 : L.fr_fo.back();
  info[4] = L.fr_max_m;
  info[5] = L.fr_flops;
  return XG_OK;
}

int xg_lu_front_export(const xg_handle *H, int64_t *sizes, int32_t *const *i32,
                       int64_t *const *i64) {
  if (!H || !sizes) return fail(XG_E_ARG, "xg_lu_front_export: bad argument");
  const xg::Plan &P = H->h.plan;
  const xg::LuPlan &L = P.lu;
  const int64_t nf = (int64_t)L.fr_m.size();
  sizes[0] = nf;
  sizes[1] = nf ? L.fr_ro.back() : 0;
  sizes[2] = (int64_t)L.fr_ch.size();
  sizes[3] = L.fr_lv_ptr.empty() ? 0 : (int64_t)L.fr_lv_ptr.size() - 1;
  sizes[4] = (int64_t)L.fr_ent_k.size();
  sizes[5] = (int64_t)P.comp_start.size() - 1;
  if (!i32 && !i64) return XG_OK;
  if (i32) {
    const std::vector<int32_t> *src[11] = {&L.fr_comp,  &L.fr_m,     &L.fr_k,
                                           &L.fr_rows,  &L.fr_cbmap, &L.fr_ch_ptr,
                                           &L.fr_ch,    &L.fr_lv_ptr, &L.fr_lv,
                                           &L.fr_ent_k, &L.front_comp};
    for (int k = 0; k < 11; ++k)
      if (i32[k] && !src[k]->empty())
        std::memcpy(i32[k], src[k]->data(), src[k]->size() * sizeof(int32_t));
  }
  if (i64) {
    const std::vector<int64_t> *src[3] = {&L.fr_fo, &L.fr_ro, &L.fr_ent_fpos};
    for (int k = 0; k < 3; ++k)
      if (i64[k] && !src[k]->empty())
        std::memcpy(i64[k], src[k]->data(), src[k]->size() * sizeof(int64_t));
  }
  return XG_OK;
}

int xg_lu_export(const xg_handle *H, int64_t *sizes, int32_t *const *arrays,
                 int64_t *voff) {
  if (!H || !sizes) return fail(XG_E_ARG, "xg_lu_export: bad argument");
  const xg::Plan &P = H->h.plan;
  const xg::LuPlan &L = P.lu;
  sizes[0] = P.n_nodes;
  sizes[1] = (int64_t)P.colidx.size();
  sizes[2] = (int64_t)P.comp_start.size() - 1;
  sizes[3] = (int64_t)L.lv_ptr.size();
  sizes[4] = (int64_t)L.t_dst.size();
  sizes[5] = (int64_t)L.ops.size() / 2;
  if (!arrays) return XG_OK;
  const std::vector<int32_t> *src[9] = {&L.prow, &L.qcol, &L.lu_pos, &L.ok, &L.lv_comp,
                                        &L.lv_ptr, &L.t_dst, &L.t_div, &L.t_op};
  for (int k = 0; k < 9; ++k)
    if (arrays[k]) std::memcpy(arrays[k], src[k]->data(), src[k]->size() * sizeof(int32_t));
  if (arrays[9]) std::memcpy(arrays[9], L.ops.data(), L.ops.size() * sizeof(int32_t));
  if (voff) std::memcpy(voff, L.voff.data(), L.voff.size() * sizeof(int64_t));
  return XG_OK;
}

int xg_tia_predict(xg_handle *H, int32_t order, double beta1, double beta2,
                   const double *xh0, const double *xh1, const double *xh2,
                   double *xpred) {
  if (!H || !xh0 || !xh1 || !xpred || (order == 2 && !xh2) ||
      (order != 1 && order != 2))
    return fail(XG_E_ARG, "xg_tia_predict: bad argument");
  xg::Handle &h = H->h;
  hipError_t e = xg::launch_predict(h.plan.n_nodes + 1, order, beta1, beta2, xh0,
                                    xh1, xh2 ? xh2 : xh1, xpred, h.stream);
  return e == hipSuccess ? XG_OK : hip_fail(e, "predict");
}

int xg_component_limited(xg_handle *H, int32_t *counts) {
  if (H && H->h.plan_only)
    return fail(XG_E_STATE, "handle was finalized without a device");
  if (!H || !counts) return fail(XG_E_ARG, "xg_component_limited: bad argument");
  xg::Handle &h = H->h;
  hipError_t e = xg::launch_component_limited(h.limited_all, h.slot_comp,
                                              h.n_limited_slots, counts,
                                              h.ds.n_comp, h.stream);
  return e == hipSuccess ? XG_OK : hip_fail(e, "component limited");
}

int xg_rhs_norms(xg_handle *H, const double *rhs, double *stats) {
  if (H && H->h.plan_only)
    return fail(XG_E_STATE, "handle was finalized without a device");
  if (!H || !rhs || !stats) return fail(XG_E_ARG, "xg_rhs_norms: bad argument");
  xg::Handle &h = H->h;
  xg::DenseSolveArgs a = h.ds;
  a.rhs = rhs;
  a.stats = stats;
  a.norm_only = 1;
  a.active = nullptr;
  hipError_t e = xg::launch_dense_solve(a, h.stream);
  return e == hipSuccess ? XG_OK : hip_fail(e, "rhs norms");
}

}  // extern "C"

namespace xg {
hipError_t handle_count_flags(xg_handle *H, int64_t *out128) {
  Handle &h = H->h;
  return launch_count_flags(h.limited_all, h.n_limited_slots, out128, h.stream);
}
int handle_assemble_norms(xg_handle *H, int32_t transient, const double *qn0,
                          const double *fprev, const double *F, const double *Q,
                          const double *FL, const double *QL, const double *JF,
                          const double *JQ, double *rhs, double *J, double *norms) {
  return tia_assemble(H, transient, 1, 0.0, qn0, fprev, F, Q, FL, QL, JF, JQ, rhs, J,
                      norms);
}
hipError_t handle_iteration_counts(xg_handle *H, int32_t *comp_cnt, int64_t *out128) {
  Handle &h = H->h;
  return launch_iteration_counts(h.limited_all, h.slot_comp, h.n_limited_slots, comp_cnt,
                                 out128, h.stream, h.gate);
}
void handle_set_gate(xg_handle *H, const int32_t *gate) { H->h.gate = gate; }
hipStream_t handle_stream(const xg_handle *H) { return H->h.stream; }
int32_t handle_n_comp(const xg_handle *H) {
  return (int32_t)H->h.plan.comp_start.size() - 1;
}
int64_t handle_n_nodes(const xg_handle *H) { return H->h.plan.n_nodes; }
int32_t handle_n_fronts(const xg_handle *H) { return H->h.fr.n_fronts; }

void handle_voltage_diagonal(const xg_handle *H, std::vector<int32_t> &ids,
                             std::vector<int32_t> &pos) {
  const xg::Plan &P = H->h.plan;
  const int32_t n = P.n_nodes, S = P.lin_slots;
  std::vector<char> branch(n, 0);
  for (int32_t s = 0; s < S; ++s)
    if (P.lin_type[s] == xg::LIN_V) {
      const int32_t b = P.lin_lid[(size_t)2 * S + s];
      if (b >= 0 && b < n) branch[b] = 1;
    }
  ids.clear();
  pos.clear();
  for (int32_t i = 0; i < n; ++i) {
    if (branch[i]) continue;
    const int32_t *b = P.colidx.data() + P.rowptr[i];
    const int32_t *e = P.colidx.data() + P.rowptr[i + 1];
    const int32_t *it = std::lower_bound(b, e, i);
    if (it == e || *it != i) continue;
    ids.push_back(i);
    pos.push_back((int32_t)(it - P.colidx.data()));
  }
}
void handle_sizes(const xg_handle *H, int64_t *n_state, int64_t *n_store, int64_t *nnz) {
  *n_state = H->h.plan.n_state;
  *n_store = H->h.plan.n_store;
  *nnz = (int64_t)H->h.plan.colidx.size();
}
}  // namespace xg

extern "C" {

int xg_instance_state(xg_handle *H, int32_t mode, double *buf, int64_t *n) {
  if (!H || !n || mode < 0 || mode > 2 || (mode && !buf))
    return fail(XG_E_ARG, "xg_instance_state: bad argument");
  xg::Handle &h = H->h;
  const int64_t m = h.plan.lin_slots;
  if (mode == 0) {
    *n = m;
    return XG_OK;
  }
  if (*n != m) return fail(XG_E_ARG, "xg_instance_state: size mismatch");
  if (h.plan_only) return fail(XG_E_STATE, "handle was finalized without a device");
  if (m == 0) return XG_OK;
  hipError_t e = hipStreamSynchronize(h.stream);
  if (e == hipSuccess)
    e = mode == 1 ? hipMemcpy(buf, h.lin.dvd, sizeof(double) * m, hipMemcpyDeviceToHost)
                  : hipMemcpy(h.lin.dvd, buf, sizeof(double) * m, hipMemcpyHostToDevice);
  return e == hipSuccess ? XG_OK : hip_fail(e, "instance state copy");
}

int xg_set_active_mask(xg_handle *H, const int32_t *active) {
  if (!H) return fail(XG_E_ARG, "xg_set_active_mask: bad argument");
  H->h.active = active;
  return XG_OK;
}

int xg_set_step_source(xg_handle *H, const double *dstep) {
  if (!H) return fail(XG_E_ARG, "xg_set_step_source: bad argument");
  H->h.dstep = dstep;
  return XG_OK;
}

int xg_set_time(xg_handle *H, double time, double bp_tol) {
  if (!H) return fail(XG_E_ARG, "xg_set_time: bad argument");
  H->h.time = time;
  H->h.bp_tol = bp_tol;
  return XG_OK;
}

int xg_add_foreign_stamp(xg_circuit *c, int32_t n_dev_nodes,
                         const int32_t *gids, int32_t n_entries,
                         const int32_t *rows, const int32_t *cols) {
  if (!c || n_dev_nodes <= 0 || n_entries < 0 || !gids ||
      (n_entries && (!rows || !cols)))
    return fail(XG_E_ARG, "xg_add_foreign_stamp: bad argument");
  xg::ForeignStamp fs;
  for (int32_t k = 0; k < n_dev_nodes; ++k) {
    if (gids[k] >= c->c.n_nodes)
      return fail(XG_E_ARG, "xg_add_foreign_stamp: node id out of range");
    fs.gids.push_back(gids[k] < 0 ? -1 : gids[k]);
  }
  for (int32_t e = 0; e < n_entries; ++e) {
    if (rows[e] < 0 || rows[e] >= n_dev_nodes || cols[e] < 0 ||
        cols[e] >= n_dev_nodes)
      return fail(XG_E_ARG, "xg_add_foreign_stamp: entry out of range");
    fs.rows.push_back(rows[e]);
    fs.cols.push_back(cols[e]);
  }
  c->c.foreign.push_back(std::move(fs));
  return XG_OK;
}

int xg_circuit_set_pattern(xg_circuit *c, const int32_t *rowptr,
                           const int32_t *colidx) {
  if (!c || !rowptr || !colidx)
    return fail(XG_E_ARG, "xg_circuit_set_pattern: bad argument");
  const int32_t nr = c->c.n_nodes + 1;
  if (rowptr[0] != 0) return fail(XG_E_ARG, "rowptr[0] != 0");
  for (int32_t r = 0; r < nr; ++r) {
    if (rowptr[r + 1] < rowptr[r]) return fail(XG_E_ARG, "rowptr not monotone");
    for (int32_t j = rowptr[r]; j < rowptr[r + 1]; ++j) {
      if (colidx[j] < 0 || colidx[j] >= nr)
        return fail(XG_E_ARG, "column index out of range");
      if (j > rowptr[r] && colidx[j] <= colidx[j - 1])
        return fail(XG_E_ARG, "row columns not strictly increasing");
    }
  }
  c->c.rowptr.assign(rowptr, rowptr + nr + 1);
  c->c.colidx.assign(colidx, colidx + rowptr[nr]);
  c->c.have_pattern = true;
  return XG_OK;
}

}  // extern "C"

// Device allocation and upload of a built plan (the device half of
// xg_finalize).  Returns false if an allocation or copy failed.
static bool upload_handle(xg_handle *H) {
  if (hipGetDevice(&H->h.device) != hipSuccess) return false;
  xg::Handle &h = H->h;
  int rc = 0;
  const xg::Plan &P = h.plan;
  bool ok = true;
  auto chk = [&](const void *p) { ok = ok && p != nullptr; };
  // MOSFET1
  h.m1.n_slots = P.m1_slots;
  chk(h.m1.inst_of = h.upload(P.m1_inst_of));
  chk(h.m1.wave_card = h.upload(P.m1_wave_card));
  chk(h.m1.cards = h.upload(P.m1_cards));
  chk(h.m1.fields = h.upload(P.m1_fields));
  chk(h.m1.lid4 = h.upload(P.m1_lid4));
  chk(h.m1.lid_int = h.upload(P.m1_lid_int));
  chk(h.m1.sta_base = h.upload(P.m1_sta));
  chk(h.m1.sto_base = h.upload(P.m1_sto));
  chk(h.m1.flags = h.upload(P.m1_flags));
  // BSIM3
  if (ok) {
    rc = xg::bsim3_upload(P, h.b3, [&](const std::vector<double> &v) {
      return h.upload(v);
    }, [&](const std::vector<int32_t> &v) { return h.upload(v); });
    ok = rc == 0;
  }
  // slots, flags
  chk(h.slots = h.alloc<double>((size_t)P.n_slot_values));
  h.n_limited_slots = P.m1_slots + xg::bsim3_eval_slots(P) + P.lin_slots;
  chk(h.limited_all = h.alloc<int32_t>((size_t)h.n_limited_slots));
  chk(h.count_buf = h.alloc<int64_t>(128));
  // circuit CSR and components for the Newton solve
  chk(h.d_rowptr = h.upload(P.rowptr));
  chk(h.slot_comp = h.upload(P.slot_comp));
  {
    std::vector<int32_t> nc((size_t)P.n_nodes, -1);
    for (size_t k = 0; k + 1 < P.comp_start.size(); ++k)
      for (int32_t j = P.comp_start[k]; j < P.comp_start[k + 1]; ++j)
        nc[P.comp_nodes[j]] = (int32_t)k;
    chk(h.node_comp = h.upload(nc));
  }
  chk(h.d_colidx = h.upload(P.colidx));
  {
    auto &d = h.ds;
    d.n_comp = (int32_t)P.comp_start.size() - 1;
    d.any_large = P.comp_work > 0;
    int32_t *cs, *cn, *cl;
    int64_t *ao;
    chk(cs = h.upload(P.comp_start));
    chk(cn = h.upload(P.comp_nodes));
    chk(cl = h.upload(P.comp_local));
    chk(ao = h.upload(P.comp_aoff));
    d.cstart = cs;
    d.nodes = cn;
    d.local = cl;
    d.aoff = ao;
    d.rowptr = h.d_rowptr;
    d.colidx = h.d_colidx;
    chk(d.work = h.alloc<double>((size_t)(P.comp_work > 0 ? P.comp_work : 1)));
    chk(d.bwork = h.alloc<double>((size_t)P.n_nodes + 1));
    // sparse LU task lists; factor blocks of components beyond LDS
    auto &u = h.lu;
    const xg::LuPlan &L = P.lu;
    u.n_comp = d.n_comp;
    int64_t gwork = 0;
    std::vector<int64_t> voff(L.voff);
    for (int32_t k = 0; k < u.n_comp; ++k)
      if (!L.in_lds[k]) u.any_global = 1;
    if (u.any_global) gwork = L.voff[u.n_comp];
    u.cstart = cs;
    chk(u.prow = h.upload(L.prow));
    chk(u.qcol = h.upload(L.qcol));
    chk(u.lu_pos = h.upload(L.lu_pos));
    chk(u.voff = h.upload(voff));
    chk(u.ok = h.upload(L.ok));
    chk(u.lv_comp = h.upload(L.lv_comp));
    chk(u.lv_ptr = h.upload(L.lv_ptr));
    chk(u.lv_mid = h.upload(L.lv_mid));
    chk(u.lv_long = h.upload(L.lv_long));
    chk(u.in_lds = h.upload(L.in_lds));
    {
      std::vector<int32_t> t4(4 * L.t_dst.size());
      for (size_t k = 0; k < L.t_dst.size(); ++k) {
        t4[4 * k] = L.t_dst[k];
        t4[4 * k + 1] = L.t_div[k];
        t4[4 * k + 2] = L.t_op[k];
        t4[4 * k + 3] = L.t_op[k + 1];
      }
      int32_t *o;
      chk(o = h.upload(t4));
      u.tasks = (const int4 *)o;
    }
    {
      int32_t *o;
      chk(o = h.upload(L.ops));
      u.ops = (const int2 *)o;
    }
    chk(u.ent_off = h.upload(L.ent_off));
    chk(u.ent_k = h.upload(L.ent_k));
    chk(u.ent_pos = h.upload(L.ent_pos));
    chk(u.ent_col = h.upload(L.ent_col));
    chk(u.ent_rs = h.upload(L.ent_rs));
    u.n_ent = (int32_t)L.ent_k.size();
    u.n_nodes = P.n_nodes;
    chk(u.Jp = h.alloc<double>((size_t)(u.n_ent > 0 ? u.n_ent : 1)));
    chk(u.bp = h.alloc<double>((size_t)(P.n_nodes > 0 ? P.n_nodes : 1)));
    u.max_berr = 1e-12;
    chk(u.val_work = h.alloc<double>((size_t)(gwork > 0 ? gwork : 1)));
    u.vec_work = d.bwork;
    chk(u.fallback = h.alloc<int32_t>((size_t)(u.n_comp > 0 ? u.n_comp : 1)));
    // dense fronts
    xg::FrontArgs &f = h.fr;
    f.n_fronts = (int32_t)L.fr_m.size();
    f.n_comp = u.n_comp;
    if (f.n_fronts > 0) {
      h.fr_lv_ptr = L.fr_lv_ptr;
      h.fr_lv_maxm = L.fr_lv_maxm;
      f.n_levels = (int32_t)L.fr_lv_maxm.size();
      f.n_vals = L.fr_fo.back();
      f.n_ent = (int64_t)L.fr_ent_k.size();
      chk(f.comp = h.upload(L.fr_comp));
      chk(f.m = h.upload(L.fr_m));
      chk(f.k = h.upload(L.fr_k));
      chk(f.rows = h.upload(L.fr_rows));
      chk(f.cbmap = h.upload(L.fr_cbmap));
      chk(f.ch_ptr = h.upload(L.fr_ch_ptr));
      chk(f.ch = h.upload(L.fr_ch));
      chk(f.lv = h.upload(L.fr_lv));
      chk(f.ent_k = h.upload(L.fr_ent_k));
      chk(f.front_comp = h.upload(L.front_comp));
      chk(f.fo = h.upload(L.fr_fo));
      chk(f.ro = h.upload(L.fr_ro));
      chk(f.ent_fpos = h.upload(L.fr_ent_fpos));
      chk(f.vals = h.alloc<double>((size_t)f.n_vals));
      chk(f.wrhs = h.alloc<double>((size_t)L.fr_ro.back()));
      chk(f.v = h.alloc<double>((size_t)(P.n_nodes > 0 ? P.n_nodes : 1)));
      chk(f.piv = h.alloc<int32_t>((size_t)L.fr_ro.back()));
      chk(f.bad = h.alloc<int32_t>((size_t)(u.n_comp > 0 ? u.n_comp : 1)));
      chk(h.fr_big_dev = h.alloc<int32_t>((size_t)f.n_fronts));
      if (ok && front_big_lists(h)) ok = false;
      // the children's contribution rows grouped by the row they land on
      // (children in order within each row), for the many-workgroup form
      {
        const int32_t nf = f.n_fronts;
        std::vector<int64_t> xr_off(nf + 1, 0), ext_ptr(1, 0);
        std::vector<int32_t> xr_row, ext_ch, ext_i, cnt, at;
        for (int32_t s = 0; s < nf; ++s) {
          const int32_t m = L.fr_m[s];
          cnt.assign(m, 0);
          for (int32_t c = L.fr_ch_ptr[s]; c < L.fr_ch_ptr[s + 1]; ++c) {
            const int32_t ch = L.fr_ch[c], kc = L.fr_k[ch], w = L.fr_m[ch] - kc;
            for (int32_t i = 0; i < w; ++i) cnt[L.fr_cbmap[L.fr_ro[ch] + kc + i]]++;
          }
          at.assign(m, 0);
          int64_t run = ext_ptr.back();
          const int64_t base = run;
          for (int32_t r = 0; r < m; ++r) {
            if (!cnt[r]) continue;
            xr_row.push_back(r);
            at[r] = (int32_t)(run - base);
            run += cnt[r];
            ext_ptr.push_back(run);
          }
          ext_ch.resize((size_t)run);
          ext_i.resize((size_t)run);
          for (int32_t c = L.fr_ch_ptr[s]; c < L.fr_ch_ptr[s + 1]; ++c) {
            const int32_t ch = L.fr_ch[c], kc = L.fr_k[ch], w = L.fr_m[ch] - kc;
            for (int32_t i = 0; i < w; ++i) {
              const int64_t e = base + at[L.fr_cbmap[L.fr_ro[ch] + kc + i]]++;
              ext_ch[(size_t)e] = ch;
              ext_i[(size_t)e] = i;
            }
          }
          xr_off[s + 1] = (int64_t)xr_row.size();
        }
        h.fr_xr_off = xr_off;
        chk(f.xr_off = h.upload(xr_off));
        chk(f.xr_row = h.upload(xr_row));
        chk(f.ext_ptr = h.upload(ext_ptr));
        chk(f.ext_ch = h.upload(ext_ch));
        chk(f.ext_i = h.upload(ext_i));
      }
      // rows of more than kLongRow entries of the front components
      {
        std::vector<int32_t> lptr(1, 0), lrows;
        for (int32_t c = 0; c < u.n_comp; ++c) {
          if (L.front_comp[c]) {
            const int32_t beg = P.comp_start[c], n = P.comp_start[c + 1] - beg;
            const int32_t *rs = L.ent_rs.data() + beg + c;
            for (int32_t i = 0; i < n; ++i)
              if (rs[i + 1] - rs[i] > xg::kLongRow) lrows.push_back(i);
          }
          lptr.push_back((int32_t)lrows.size());
        }
        chk(f.long_ptr = h.upload(lptr));
        chk(f.long_rows = h.upload(lrows));
      }
      // chunked finish of the largest front components
      {
        std::vector<int32_t> split(u.n_comp > 0 ? u.n_comp : 1, 0), ch_comp, ch_beg, comps,
            cptr(1, 0);
        for (int32_t c = 0; c < u.n_comp; ++c) {
          const int32_t n = P.comp_start[c + 1] - P.comp_start[c];
          if (!L.front_comp[c] || n <= xg::kFinishSplit) continue;
          split[c] = 1;
          comps.push_back(c);
          for (int32_t r = 0; r < n; r += xg::kFinishChunk) {
            ch_comp.push_back(c);
            ch_beg.push_back(r);
          }
          cptr.push_back((int32_t)ch_comp.size());
        }
        if (!comps.empty()) {
          f.n_fin_chunks = (int32_t)ch_comp.size();
          f.n_fin_comps = (int32_t)comps.size();
          chk(f.fin_split = h.upload(split));
          chk(f.fin_chunk_comp = h.upload(ch_comp));
          chk(f.fin_chunk_beg = h.upload(ch_beg));
          chk(f.fin_comps = h.upload(comps));
          chk(f.fin_cptr = h.upload(cptr));
          chk(f.fin_part = h.alloc<double>((size_t)4 * ch_comp.size()));
          chk(f.fin_good = h.alloc<int32_t>((size_t)u.n_comp));
        }
      }
      u.front = f.front_comp;
    }
    if (ok && hipMemset(u.fallback, 0, sizeof(int32_t) * (u.n_comp > 0 ? u.n_comp : 1)) !=
                  hipSuccess)
      ok = false;
  }
  // linear devices
  h.lin.n_slots = P.lin_slots;
  chk(h.lin.type = h.upload(P.lin_type));
  chk(h.lin.lid = h.upload(P.lin_lid));
  chk(h.lin.val = h.upload(P.lin_val));
  chk(h.lin.pulse = h.upload(P.lin_pulse));
  chk(h.lin.sta = h.upload(P.lin_sta));
  chk(h.lin.dpar = h.upload(P.lin_dpar));
  {
    std::vector<double> z((size_t)(P.lin_slots > 0 ? P.lin_slots : 1), 0.0);
    chk(h.lin.dvd = h.upload(z));
  }
  h.lin.gmin = P.gmin;
  h.lin.limiter = P.limiter;
  if (ok) {
    h.lin.slots = h.slots + P.lin_slot_base;
    h.lin.limited = h.limited_all + P.m1_slots + xg::bsim3_eval_slots(P);
    h.m1.slots = h.slots;
    h.m1.limited = h.limited_all;
    xg::bsim3_bind(h.b3, h.slots, h.limited_all + P.m1_slots);
    if (hipMemset(h.limited_all, 0, sizeof(int32_t) * (h.n_limited_slots ? h.n_limited_slots : 1)) != hipSuccess)
      ok = false;
    if (P.n_slot_values &&
        hipMemset(h.slots, 0, sizeof(double) * P.n_slot_values) != hipSuccess)
      ok = false;
  }
  auto up_prog = [&](const xg::ReduceProgram &rp, xg::ReduceDev &rd) {
    rd.n_short4 = (int32_t)rp.sdst4.size();
    rd.n_short = (int32_t)rp.sdst.size();
    rd.n_chunk = (int32_t)rp.cdst.size();
    rd.n_multi = (int32_t)rp.mtgt.size();
    rd.n_zero = (int32_t)rp.zidx.size();
    rd.pbase = rp.pbase;
    chk(rd.ssrc4 = h.upload(rp.ssrc4));
    chk(rd.sdst4 = h.upload(rp.sdst4));
    chk(rd.ssrc = h.upload(rp.ssrc));
    chk(rd.sdst = h.upload(rp.sdst));
    chk(rd.zidx = h.upload(rp.zidx));
    chk(rd.src = h.upload(rp.src));
    chk(rd.cptr = h.upload(rp.cptr));
    chk(rd.cdst = h.upload(rp.cdst));
    chk(rd.mtgt = h.upload(rp.mtgt));
    chk(rd.mchunk = h.upload(rp.mchunk));
  };
  chk(h.partial = h.alloc<double>((size_t)P.n_partials));
  for (int k = 0; k < 4; ++k) up_prog(P.red_vec[k], h.red_vec[k]);
  for (int k = 0; k < 2; ++k) up_prog(P.red_mat[k], h.red_mat[k]);
  (void)rc;
  return ok;
}

extern "C" {

namespace {
// Host setup allocates freely; no exception may cross the C boundary.
int build_plan_nothrow(xg::Circuit &c, xg::Plan &plan, std::string &err) {
  try {
    return xg::build_plan(c, plan, err);
  } catch (const std::bad_alloc &) {
    err = "setup: out of host memory";
    return XG_E_NOMEM;
  } catch (const std::exception &e) {
    err = std::string("setup: ") + e.what();
    return XG_E_ARG;
  }
}
bool upload_handle_nothrow(xg_handle *H) {
  try {
    return upload_handle(H);
  } catch (const std::exception &) {
    return false;
  }
}
}  // namespace

}  // extern "C"

static int eval_types(xg::Handle &h, const xg::Mos1Args &m1, const xg::Bsim3Dev &b3,
                      const xg::LinArgs &lin, int32_t type, const double *x_next,
                      double *sta_curr, const double *sto_curr, double *sta_next,
                      double *sto_next);

// ---- autotuner: the BSIM3 eval kernel form and whether the linear devices
// run in trailing workgroups of its launch are timed on the circuit itself
// at upload and the fastest kept; the choice is cached by circuit shape
// (eval slots per device type and model cards), so further circuits of the
// same shape take it without timing.  Every candidate computes the same
// bits; only the launch configuration differs.
namespace {

struct TuneKey {
  int64_t m1, b3, lin, cards;
  bool operator<(const TuneKey &o) const {
    return std::tie(m1, b3, lin, cards) < std::tie(o.m1, o.b3, o.lin, o.cards);
  }
};
struct TuneChoice {
  int32_t variant;
  bool fuse;
  double us[6];
};
std::mutex g_tune_mu;
std::map<TuneKey, TuneChoice> g_tune_cache;

constexpr int32_t kTuneMinWaves = 8;  // smaller circuits keep the default

int autotune(xg::Handle &h) {
  const xg::Plan &P = h.plan;
  if (h.b3.n_slots < 64 * kTuneMinWaves) return XG_OK;
  const TuneKey key{P.m1_slots, h.b3.n_slots, P.lin_slots, (int64_t)h.b3.n_sdp};
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tune_cache.find(key);
    if (it != g_tune_cache.end()) {
      h.b3.variant = it->second.variant;
      h.fuse_linear = it->second.fuse;
      std::memcpy(h.tune_us, it->second.us, sizeof(h.tune_us));
      h.tune_source = 2;
      return XG_OK;
    }
  }
  const int64_t n1 = P.n_nodes + 1, ns = std::max(P.n_state, 1), nt = std::max(P.n_store, 1);
  std::vector<double> hx(n1);
  for (int64_t i = 0; i < n1; ++i) hx[i] = i + 1 < n1 ? 0.2 + std::fmod(0.618034 * i, 1.0) : 0.0;
  double *x = nullptr, *st = nullptr, *dvd = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = XG_OK;
  const int64_t nd = std::max<int64_t>(h.lin.n_slots, 1);
  if (hipMalloc(&x, sizeof(double) * n1) != hipSuccess ||
      hipMalloc(&st, sizeof(double) * 2 * (ns + nt)) != hipSuccess ||
      hipMalloc(&dvd, sizeof(double) * nd) != hipSuccess ||
      hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
    rc = XG_E_NOMEM;
  }
  const xg::SolverStateDev ss_saved = h.ss;
  const int32_t v_saved = h.b3.variant;
  if (rc == XG_OK &&
      (hipMemcpy(x, hx.data(), sizeof(double) * n1, hipMemcpyHostToDevice) != hipSuccess ||
       hipMemset(st, 0, sizeof(double) * 2 * (ns + nt)) != hipSuccess ||
       (h.lin.n_slots &&
        hipMemcpy(dvd, h.lin.dvd, sizeof(double) * nd, hipMemcpyDeviceToDevice) != hipSuccess)))
    rc = XG_E_HIP;
  double best = -1.0;
  TuneChoice ch{XG_EVAL_AUTO, true, {}};
  for (double &u : ch.us) u = std::nan("");
  if (rc == XG_OK) {
    h.ss = xg::SolverStateDev{};
    h.ss.transient = 1;
    h.ss.newtonIter = 1;
    const int32_t forms[3] = {XG_EVAL_HBM, XG_EVAL_STAGED, XG_EVAL_STAGED_SPLIT};
    for (int f = 0; f < 3 && rc == XG_OK; ++f)
      for (int fuse = 1; fuse >= 0 && rc == XG_OK; --fuse) {
        if (!fuse && h.lin.n_slots == 0) continue;
        h.b3.variant = forms[f];
        h.fuse_linear = fuse != 0;
        auto run = [&]() {
          return eval_types(h, h.m1, h.b3, h.lin, 0, x, st, st + ns, st + ns + nt,
                            st + 2 * ns + nt);
        };
        for (int k = 0; k < 2 && rc == XG_OK; ++k) rc = run();
        if (rc) break;
        const int reps = 5;
        (void)hipEventRecord(e0, h.stream);
        for (int k = 0; k < reps && rc == XG_OK; ++k) rc = run();
        (void)hipEventRecord(e1, h.stream);
        float ms = 0.0f;
        if (rc == XG_OK && (hipEventSynchronize(e1) != hipSuccess ||
                            hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
          rc = XG_E_HIP;
        if (rc) break;
        const double us = 1e3 * ms / reps;
        ch.us[2 * f + (fuse ? 0 : 1)] = us;
        if (best < 0 || us < best) {
          best = us;
          ch.variant = forms[f];
          ch.fuse = fuse != 0;
        }
      }
  }
  // the tuning launches leave no trace: statement values, flags and the
  // diodes' junction voltages back to their state after upload
  if (h.lin.n_slots && dvd)
    (void)hipMemcpy(h.lin.dvd, dvd, sizeof(double) * nd, hipMemcpyDeviceToDevice);
  if (P.n_slot_values) (void)hipMemset(h.slots, 0, sizeof(double) * P.n_slot_values);
  if (h.n_limited_slots)
    (void)hipMemset(h.limited_all, 0, sizeof(int32_t) * h.n_limited_slots);
  (void)hipStreamSynchronize(h.stream);
  h.ss = ss_saved;
  if (x) (void)hipFree(x);
  if (st) (void)hipFree(st);
  if (dvd) (void)hipFree(dvd);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (rc != XG_OK || best < 0) {
    h.b3.variant = v_saved;
    h.fuse_linear = true;
    return rc;
  }
  h.b3.variant = ch.variant;
  h.fuse_linear = ch.fuse;
  std::memcpy(h.tune_us, ch.us, sizeof(h.tune_us));
  h.tune_source = 1;
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tune_cache[key] = ch;
  return XG_OK;
}

}  // namespace

extern "C" {

int xg_autotune_info(const xg_handle *H, int32_t *choice, double *us) {
  if (!H) return fail(XG_E_ARG, "xg_autotune_info: bad argument");
  const xg::Handle &h = H->h;
  if (choice) {
    choice[0] = h.b3.variant;
    choice[1] = h.fuse_linear ? 1 : 0;
    choice[2] = h.tune_source;
  }
  if (us) std::memcpy(us, h.tune_us, sizeof(h.tune_us));
  return XG_OK;
}

int xg_finalize(xg_circuit *c, xg_handle **out) {
  if (!c || !out) return fail(XG_E_ARG, "xg_finalize: bad argument");
  auto *H = new (std::nothrow) xg_handle;
  if (!H) return fail(XG_E_NOMEM, "xg_finalize: out of memory");
  std::string err;
  int rc = build_plan_nothrow(c->c, H->h.plan, err);
  if (rc) {
    delete H;
    return fail(rc, err);
  }
  if (!upload_handle_nothrow(H)) {
    delete H;
    return fail(XG_E_NOMEM, "xg_finalize: device allocation/upload failed");
  }
  if (int rc2 = autotune(H->h)) {
    delete H;
    return fail(rc2, "xg_finalize: autotuning failed");
  }
  *out = H;
  return XG_OK;
}

int xg_upload(xg_handle *H) {
  if (!H) return fail(XG_E_ARG, "xg_upload: bad argument");
  if (!H->h.plan_only) return fail(XG_E_STATE, "xg_upload: handle is already on the device");
  if (!upload_handle_nothrow(H)) return fail(XG_E_NOMEM, "xg_upload: device allocation/upload failed");
  H->h.plan_only = false;
  if (int rc = autotune(H->h)) return fail(rc, "xg_upload: autotuning failed");
  return XG_OK;
}

int xg_finalize_plan(xg_circuit *c, xg_handle **out) {
  if (!c || !out) return fail(XG_E_ARG, "xg_finalize_plan: bad argument");
  auto *H = new (std::nothrow) xg_handle;
  if (!H) return fail(XG_E_NOMEM, "xg_finalize_plan: out of memory");
  std::string err;
  int rc = build_plan_nothrow(c->c, H->h.plan, err);
  if (rc) {
    delete H;
    return fail(rc, err);
  }
  H->h.plan_only = true;
  *out = H;
  return XG_OK;
}

void xg_destroy(xg_handle *h) {
  if (h) {
    // the handle's work is on its stream (a transient driver syncs its own)
    if (!h->h.plan_only) (void)hipStreamSynchronize(h->h.stream);
    delete h;
  }
}

int xg_pattern_size(const xg_handle *h, int32_t *n_rows, int64_t *nnz) {
  if (!h || !n_rows || !nnz) return fail(XG_E_ARG, "xg_pattern_size");
  *n_rows = h->h.plan.n_nodes + 1;
  *nnz = (int64_t)h->h.plan.colidx.size();
  return XG_OK;
}

int xg_get_pattern(const xg_handle *h, int32_t *rowptr, int32_t *colidx) {
  if (!h || !rowptr || !colidx) return fail(XG_E_ARG, "xg_get_pattern");
  const auto &P = h->h.plan;
  std::memcpy(rowptr, P.rowptr.data(), P.rowptr.size() * sizeof(int32_t));
  std::memcpy(colidx, P.colidx.data(), P.colidx.size() * sizeof(int32_t));
  return XG_OK;
}

int xg_get_mos1_offsets(const xg_handle *h, int32_t *off) {
  if (!h || !off) return fail(XG_E_ARG, "xg_get_mos1_offsets");
  const auto &v = h->h.plan.m1_offsets;
  std::memcpy(off, v.data(), v.size() * sizeof(int32_t));
  return XG_OK;
}

int xg_get_bsim3_offsets(const xg_handle *h, int32_t *off) {
  if (!h || !off) return fail(XG_E_ARG, "xg_get_bsim3_offsets");
  return xg::bsim3_get_offsets(h->h.plan, off);
}

int xg_get_mos1_fields(const xg_handle *h, double *out, int32_t *n_fields) {
  if (!h || !n_fields) return fail(XG_E_ARG, "xg_get_mos1_fields");
  *n_fields = xg::M1F_N;
  if (out) {
    const auto &v = h->h.plan.m1_fields_ref;
    std::memcpy(out, v.data(), v.size() * sizeof(double));
  }
  return XG_OK;
}

int xg_get_bsim3_fields(const xg_handle *h, double *out, int32_t *n_fields) {
  if (!h || !n_fields) return fail(XG_E_ARG, "xg_get_bsim3_fields");
  return xg::bsim3_get_fields(h->h.plan, out, n_fields);
}

int xg_set_solver_state(xg_handle *H, const xg_solver_state *s) {
  if (!H || !s) return fail(XG_E_ARG, "xg_set_solver_state");
  auto &d = H->h.ss;
  d.dcop = s->dcopFlag;
  d.initJct = s->initJctFlag;
  d.initFix = s->initFixFlag;
  d.initTran = s->initTranFlag;
  d.newtonIter = s->newtonIter;
  d.tranop = s->tranopFlag;
  d.acop = s->acopFlag;
  d.transient = s->transientFlag;
  return XG_OK;
}

int xg_set_output_mode(xg_handle *H, int32_t mode) {
  if (!H || (mode != XG_OUT_ACCUMULATE && mode != XG_OUT_OVERWRITE))
    return fail(XG_E_ARG, "xg_set_output_mode: bad argument");
  H->h.overwrite = mode == XG_OUT_OVERWRITE;
  return XG_OK;
}

int xg_set_eval_variant(xg_handle *H, int32_t variant) {
  if (!H || variant < XG_EVAL_AUTO || variant > XG_EVAL_STAGED_SPLIT)
    return fail(XG_E_ARG, "xg_set_eval_variant: bad argument");
  H->h.b3.variant = variant;
  return XG_OK;
}

int xg_set_stream(xg_handle *H, void *stream) {
  if (!H) return fail(XG_E_ARG, "xg_set_stream");
  H->h.stream = (hipStream_t)stream;
  return XG_OK;
}

// The eval launches of one device type (0 = every type, in load order),
// with the device argument sets given (the handle's own, or copies bound to
// other statement-value memory).
static int eval_types(xg::Handle &h, const xg::Mos1Args &m1, const xg::Bsim3Dev &b3,
                      const xg::LinArgs &lin, int32_t type, const double *x_next,
                      double *sta_curr, const double *sto_curr, double *sta_next,
                      double *sto_next) {
  xg::LoadPtrs p{x_next, sta_curr, sto_curr, sta_next, sto_next};
  p.active = h.active;
  p.gate = h.gate;
  hipError_t e = hipSuccess;
  if (type == 0 || type == XG_DEV_MOS1) {
    p.comp = h.slot_comp;
    e = xg::launch_mos1_eval(m1, h.ss, p, h.stream);
    if (e != hipSuccess) return hip_fail(e, "mos1 eval");
  }
  // every type: the linear devices ride in trailing workgroups of the
  // BSIM3 launch when there are both
  const bool fused = type == 0 && b3.n_slots > 0 && lin.n_slots > 0 && h.fuse_linear;
  if (type == 0 || type == XG_DEV_BSIM3) {
    p.comp = h.slot_comp + h.plan.m1_slots;
    xg::LinTail tail;
    if (fused) {
      tail.a = lin;
      tail.p = p;
      tail.p.comp = h.slot_comp + h.plan.m1_slots + xg::bsim3_eval_slots(h.plan);
      tail.time = h.time;
      tail.bp_tol = h.bp_tol;
      tail.dstep = h.dstep;
    }
    e = xg::launch_bsim3_eval(b3, h.ss, p, h.stream, fused ? &tail : nullptr);
    if (e != hipSuccess) return hip_fail(e, "bsim3 eval");
  }
  if ((type == 0 && !fused) || type == XG_DEV_LINEAR) {
    p.comp = h.slot_comp + h.plan.m1_slots + xg::bsim3_eval_slots(h.plan);
    e = xg::launch_linear_eval(lin, h.ss, p, h.time, h.bp_tol, h.dstep, h.stream);
    if (e != hipSuccess) return hip_fail(e, "linear eval");
  }
  return XG_OK;
}

static int check_update_args(xg_handle *H, const double *x_next, const double *sta_curr,
                             const double *sto_curr, const double *sta_next,
                             const double *sto_next) {
  if (H && H->h.plan_only)
    return fail(XG_E_STATE, "handle was finalized without a device");
  if (!H || !x_next) return fail(XG_E_ARG, "xg_update_state: null pointer");
  const xg::Handle &h = H->h;
  // state / store pointers may be null when the circuit has none
  if ((h.plan.n_state > 0 && (!sta_curr || !sta_next)) ||
      (h.plan.n_store > 0 && (!sto_curr || !sto_next)))
    return fail(XG_E_ARG, "xg_update_state: null pointer");
  return XG_OK;
}

static int update_state_types(xg_handle *H, int32_t type, const double *x_next,
                              double *sta_curr, const double *sto_curr,
                              double *sta_next, double *sto_next) {
  if (int rc = check_update_args(H, x_next, sta_curr, sto_curr, sta_next, sto_next))
    return rc;
  if (type < 0 || type > XG_DEV_LINEAR)
    return fail(XG_E_ARG, "xg_update_state_type: unknown device type");
  xg::Handle &h = H->h;
  return eval_types(h, h.m1, h.b3, h.lin, type, x_next, sta_curr, sto_curr, sta_next,
                    sto_next);
}

int xg_update_state(xg_handle *H, const double *x_next, double *sta_curr,
                    const double *sto_curr, double *sta_next,
                    double *sto_next) {
  return update_state_types(H, 0, x_next, sta_curr, sto_curr, sta_next, sto_next);
}

int xg_update_state_type(xg_handle *H, int32_t type, const double *x_next,
                         double *sta_curr, const double *sto_curr,
                         double *sta_next, double *sto_next) {
  if (type == 0) return fail(XG_E_ARG, "xg_update_state_type: unknown device type");
  return update_state_types(H, type, x_next, sta_curr, sto_curr, sta_next, sto_next);
}

int xg_load_vectors(xg_handle *H, double *F, double *Q, double *dFdxdVp,
                    double *dQdxdVp) {
  if (H && H->h.plan_only)
    return fail(XG_E_STATE, "handle was finalized without a device");
  if (!H || !F || !Q || !dFdxdVp || !dQdxdVp)
    return fail(XG_E_ARG, "xg_load_vectors: null pointer");
  xg::Handle &h = H->h;
  double *outs[4] = {F, Q, dFdxdVp, dQdxdVp};
  hipError_t e = xg::launch_reduce(4, h.red_vec, outs, h.slots, h.partial, h.overwrite, h.stream, h.gate);
  if (e != hipSuccess) return hip_fail(e, "vector reduction");
  return XG_OK;
}

int xg_load_matrices(xg_handle *H, double *dFdx, double *dQdx) {
  if (H && H->h.plan_only)
    return fail(XG_E_STATE, "handle was finalized without a device");
  if (!H || !dFdx || !dQdx)
    return fail(XG_E_ARG, "xg_load_matrices: null pointer");
  xg::Handle &h = H->h;
  double *outs[2] = {dFdx, dQdx};
  hipError_t e = xg::launch_reduce(2, h.red_mat, outs, h.slots, h.partial, h.overwrite, h.stream, h.gate);
  if (e != hipSuccess) return hip_fail(e, "matrix reduction");
  return XG_OK;
}

int xg_load_all(xg_handle *H, const double *x_next, double *sta_curr,
                const double *sto_curr, double *sta_next, double *sto_next,
                double *F, double *Q, double *dFdxdVp, double *dQdxdVp,
                double *dFdx, double *dQdx) {
  int rc = xg_update_state(H, x_next, sta_curr, sto_curr, sta_next, sto_next);
  if (rc) return rc;
  if (!F || !Q || !dFdxdVp || !dQdxdVp || !dFdx || !dQdx)
    return fail(XG_E_ARG, "xg_load_all: null pointer");
  // vector and matrix folds in one reduction pass
  xg::Handle &h = H->h;
  xg::ReduceDev progs[6] = {h.red_vec[0], h.red_vec[1], h.red_vec[2],
                            h.red_vec[3], h.red_mat[0], h.red_mat[1]};
  double *outs[6] = {F, Q, dFdxdVp, dQdxdVp, dFdx, dQdx};
  hipError_t e = xg::launch_reduce(6, progs, outs, h.slots, h.partial, h.overwrite, h.stream, h.gate);
  if (e != hipSuccess) return hip_fail(e, "reduction");
  return XG_OK;
}

}  // extern "C"

// ---- xg_load_all_noalloc: the statement values and reduction partials of
// the load live in one library-wide device array instead of the handle's
// buffers, so the call allocates nothing.  Each GPU holds its own copy of
// the array (at its own address, found with hipGetSymbolAddress on that
// device's first call); every call on a GPU shares it, so calls that
// overlap on one GPU race.
namespace {

constexpr int64_t kScratchDoubles = XG_NOALLOC_SCRATCH_BYTES / sizeof(double);
__device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_dev[kMaxDevices] = {};

// Device address of `scratch` on the current device (which must be
// `device`), looked up once per device.
int scratch_address(int device, double **out) {
  if (device < 0 || device >= kMaxDevices)
    return fail(XG_E_ARG, "xg_load_all_noalloc: device id out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_dev[device]) {
    void *ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(e, "xg_load_all_noalloc: hipGetSymbolAddress");
    if (!ptr) return fail(XG_E_HIP, "xg_load_all_noalloc: null scratch address");
    g_scratch_dev[device] = (double *)ptr;
  }
  *out = g_scratch_dev[device];
  return XG_OK;
}

__global__ void zero_doubles_kernel(double *p, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    p[i] = 0.0;
}

}  // namespace

extern "C" {

int xg_load_all_noalloc(xg_handle *H, const double *x_next, double *sta_curr,
                        const double *sto_curr, double *sta_next, double *sto_next,
                        double *F, double *Q, double *dFdxdVp, double *dQdxdVp,
                        double *dFdx, double *dQdx) {
  if (!H) return fail(XG_E_ARG, "xg_load_all_noalloc: null handle");
  xg::Handle &h = H->h;
  // statement values, then the chunk partials at a 256-byte boundary
  const int64_t n_val = h.plan.n_slot_values;
  const int64_t pbase = (n_val + 31) / 32 * 32;
  const int64_t need = pbase + (int64_t)h.plan.n_partials;
  if (need > kScratchDoubles)
    return fail(XG_E_ARG, "xg_load_all_noalloc: the circuit needs " + std::to_string(need * 8) +
                              " bytes of scratch; the limit is " +
                              std::to_string(kScratchDoubles * 8) + " (use xg_load_all)");
  if (int rc = check_update_args(H, x_next, sta_curr, sto_curr, sta_next, sto_next)) return rc;
  if (!F || !Q || !dFdxdVp || !dQdxdVp || !dFdx || !dQdx)
    return fail(XG_E_ARG, "xg_load_all_noalloc: null pointer");
  if (h.active)
    return fail(XG_E_STATE,
                "xg_load_all_noalloc: an active mask keeps statement values across calls, "
                "which the shared scratch does not; clear it or use xg_load_all");
  int cur = -1;
  hipError_t e = hipGetDevice(&cur);
  if (e != hipSuccess) return hip_fail(e, "xg_load_all_noalloc: hipGetDevice");
  if (cur != h.device)
    return fail(XG_E_STATE, "xg_load_all_noalloc: the handle's GPU is not the current device");
  double *base = nullptr;
  if (int rc = scratch_address(h.device, &base)) return rc;
  double *vals = base, *partial = base + pbase;
  // Statements an instance never writes (absent drain'/source' rows,
  // padding) read as +0.0, as in the handle's buffer, which starts zeroed;
  // the array holds the last call's values, so clear it (a fill kernel,
  // not a memset node, so the call may be graph-captured).
  if (n_val > 0) {
    const int64_t blocks = std::min<int64_t>((n_val + 255) / 256, 4096);
    hipLaunchKernelGGL(zero_doubles_kernel, dim3((unsigned)blocks), dim3(256), 0, h.stream,
                       vals, n_val);
    if ((e = hipGetLastError()) != hipSuccess) return hip_fail(e, "scratch clear");
  }
  xg::Mos1Args m1 = h.m1;
  m1.slots = vals;
  xg::LinArgs lin = h.lin;
  lin.slots = vals + h.plan.lin_slot_base;
  xg::Bsim3Dev b3 = h.b3;
  xg::bsim3_bind(b3, vals, h.limited_all + h.plan.m1_slots);
  if (int rc = eval_types(h, m1, b3, lin, 0, x_next, sta_curr, sto_curr, sta_next, sto_next))
    return rc;
  xg::ReduceDev progs[6] = {h.red_vec[0], h.red_vec[1], h.red_vec[2],
                            h.red_vec[3], h.red_mat[0], h.red_mat[1]};
  double *outs[6] = {F, Q, dFdxdVp, dQdxdVp, dFdx, dQdx};
  e = xg::launch_reduce(6, progs, outs, vals, partial, h.overwrite, h.stream);
  if (e != hipSuccess) return hip_fail(e, "reduction");
  return XG_OK;
}

// ---- host-pointer variants: the caller's vectors stay in host memory
// (the reference's N_LAS vectors); each call stages them through one
// pinned buffer to device mirrors and back, and returns when the results
// are in the caller's arrays.
static int host_mirrors(xg_handle *H) {
  xg::Handle &h = H->h;
  if (h.hx) return XG_OK;
  const xg::Plan &P = h.plan;
  const int64_t n1 = P.n_nodes + 1, nnz = (int64_t)P.colidx.size();
  const int64_t ns = std::max<int64_t>(P.n_state, 1), nt = std::max<int64_t>(P.n_store, 1);
  h.pinned_len = std::max<int64_t>({n1 + 2 * ns + 2 * nt, 4 * n1, 2 * nnz});
  if (hipHostMalloc((void **)&h.pinned, sizeof(double) * h.pinned_len) != hipSuccess) {
    h.pinned = nullptr;
    return fail(XG_E_NOMEM, "host-pointer loads: pinned staging allocation failed");
  }
  if (!(h.hx = h.alloc<double>(n1)) || !(h.hsta_c = h.alloc<double>(ns)) ||
      !(h.hsto_c = h.alloc<double>(nt)) || !(h.hsta_n = h.alloc<double>(ns)) ||
      !(h.hsto_n = h.alloc<double>(nt)) || !(h.hvec = h.alloc<double>(4 * n1)) ||
      !(h.hmat = h.alloc<double>(2 * std::max<int64_t>(nnz, 1)))) {
    h.hx = nullptr;
    return fail(XG_E_NOMEM, "host-pointer loads: device mirrors allocation failed");
  }
  return XG_OK;
}

// Copies n doubles host -> device (through the staging buffer at off).
static hipError_t stage_in(xg::Handle &h, double *dev, const double *host, int64_t n,
                           int64_t off) {
  if (n <= 0 || !host) return hipSuccess;
  std::memcpy(h.pinned + off, host, sizeof(double) * n);
  return hipMemcpyAsync(dev, h.pinned + off, sizeof(double) * n, hipMemcpyHostToDevice,
                        h.stream);
}

static hipError_t stage_out(xg::Handle &h, const double *dev, int64_t n, int64_t off) {
  if (n <= 0) return hipSuccess;
  return hipMemcpyAsync(h.pinned + off, dev, sizeof(double) * n, hipMemcpyDeviceToHost,
                        h.stream);
}

int xg_update_state_host(xg_handle *H, const double *x_next, double *sta_curr,
                         const double *sto_curr, double *sta_next, double *sto_next) {
  if (int rc = check_update_args(H, x_next, sta_curr, sto_curr, sta_next, sto_next)) return rc;
  if (int rc = host_mirrors(H)) return rc;
  xg::Handle &h = H->h;
  const int64_t n1 = h.plan.n_nodes + 1, ns = h.plan.n_state, nt = h.plan.n_store;
  // in: x, current state and store, and the next store (the previous Newton
  // iteration's voltages, read by the limiters); sta_next is written whole
  hipError_t e = stage_in(h, h.hx, x_next, n1, 0);
  if (e == hipSuccess) e = stage_in(h, h.hsta_c, sta_curr, ns, n1);
  if (e == hipSuccess) e = stage_in(h, h.hsto_c, sto_curr, nt, n1 + ns);
  if (e == hipSuccess) e = stage_in(h, h.hsto_n, sto_next, nt, n1 + ns + nt);
  if (e == hipSuccess) e = stage_in(h, h.hsta_n, sta_next, ns, n1 + ns + 2 * nt);
  if (e != hipSuccess) return hip_fail(e, "xg_update_state_host: copy in");
  if (int rc = eval_types(h, h.m1, h.b3, h.lin, 0, h.hx, h.hsta_c, h.hsto_c, h.hsta_n,
                          h.hsto_n))
    return rc;
  // out: next state and store, and the current state (written at initTran)
  e = stage_out(h, h.hsta_n, ns, 0);
  if (e == hipSuccess) e = stage_out(h, h.hsto_n, nt, ns);
  if (e == hipSuccess) e = stage_out(h, h.hsta_c, ns, ns + nt);
  if (e == hipSuccess) e = hipStreamSynchronize(h.stream);
  if (e != hipSuccess) return hip_fail(e, "xg_update_state_host: copy out");
  if (ns > 0) {
    std::memcpy(sta_next, h.pinned, sizeof(double) * ns);
    std::memcpy(sta_curr, h.pinned + ns + nt, sizeof(double) * ns);
  }
  if (nt > 0) std::memcpy(sto_next, h.pinned + ns, sizeof(double) * nt);
  return XG_OK;
}

// The reductions of the last xg_update_state_host into the caller's host
// arrays (added to what they hold, or written, per the output mode).
static int reduce_host(xg_handle *H, int nprog, double *const *outs, int64_t len,
                       const char *who) {
  if (H && H->h.plan_only)
    return fail(XG_E_STATE, "handle was finalized without a device");
  if (!H) return fail(XG_E_ARG, std::string(who) + ": null handle");
  for (int k = 0; k < nprog; ++k)
    if (!outs[k]) return fail(XG_E_ARG, std::string(who) + ": null pointer");
  xg::Handle &h = H->h;
  if (!h.hx) return fail(XG_E_STATE, std::string(who) + ": call xg_update_state_host first");
  double *dev = nprog == 4 ? h.hvec : h.hmat;
  hipError_t e = hipSuccess;
  if (!h.overwrite)
    for (int k = 0; k < nprog && e == hipSuccess; ++k)
      e = stage_in(h, dev + k * len, outs[k], len, k * len);
  if (e != hipSuccess) return hip_fail(e, (std::string(who) + ": copy in").c_str());
  double *douts[6];
  for (int k = 0; k < nprog; ++k) douts[k] = dev + k * len;
  e = nprog == 4 ? xg::launch_reduce(4, h.red_vec, douts, h.slots, h.partial, h.overwrite,
                                     h.stream)
                 : xg::launch_reduce(2, h.red_mat, douts, h.slots, h.partial, h.overwrite,
                                     h.stream);
  if (e == hipSuccess) e = stage_out(h, dev, nprog * len, 0);
  if (e == hipSuccess) e = hipStreamSynchronize(h.stream);
  if (e != hipSuccess) return hip_fail(e, (std::string(who) + ": reduction / copy out").c_str());
  for (int k = 0; k < nprog; ++k)
    std::memcpy(outs[k], h.pinned + k * len, sizeof(double) * len);
  return XG_OK;
}

int xg_load_vectors_host(xg_handle *H, double *F, double *Q, double *dFdxdVp,
                         double *dQdxdVp) {
  double *outs[4] = {F, Q, dFdxdVp, dQdxdVp};
  return reduce_host(H, 4, outs, H ? (int64_t)H->h.plan.n_nodes + 1 : 0,
                     "xg_load_vectors_host");
}

int xg_load_matrices_host(xg_handle *H, double *dFdx, double *dQdx) {
  double *outs[2] = {dFdx, dQdx};
  return reduce_host(H, 2, outs, H ? (int64_t)H->h.plan.colidx.size() : 0,
                     "xg_load_matrices_host");
}

int xg_get_stats(xg_handle *H, xg_stats *st) {
  if (H && H->h.plan_only)
    return fail(XG_E_STATE, "handle was finalized without a device");
  if (!H || !st) return fail(XG_E_ARG, "xg_get_stats");
  xg::Handle &h = H->h;
  st->n_limited = 0;
  st->n_nonfinite = 0;
  if (h.n_limited_slots == 0) return XG_OK;
  hipError_t e = xg::launch_count_flags(h.limited_all, h.n_limited_slots,
                                        h.count_buf, h.stream);
  if (e != hipSuccess) return hip_fail(e, "count");
  int64_t part[128];
  e = hipMemcpyAsync(part, h.count_buf, sizeof(part), hipMemcpyDeviceToHost,
                     h.stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h.stream);
  if (e != hipSuccess) return hip_fail(e, "count copy");
  for (int k = 0; k < 64; ++k) {
    st->n_limited += part[k];
    st->n_nonfinite += part[64 + k];
  }
  return XG_OK;
}

int xg_sync(xg_handle *H) {
  if (H && H->h.plan_only)
    return fail(XG_E_STATE, "handle was finalized without a device");
  if (!H) return fail(XG_E_ARG, "xg_sync");
  hipError_t e = hipStreamSynchronize(H->h.stream);
  return e == hipSuccess ? XG_OK : hip_fail(e, "xg_sync");
}

int xg_sizes(const xg_handle *H, int32_t *n_nodes, int32_t *n_state,
             int32_t *n_store, int32_t *n_mos1, int32_t *n_bsim3) {
  if (!H) return fail(XG_E_ARG, "xg_sizes");
  const auto &P = H->h.plan;
  if (n_nodes) *n_nodes = P.n_nodes;
  if (n_state) *n_state = P.n_state;
  if (n_store) *n_store = P.n_store;
  if (n_mos1) *n_mos1 = P.m1_n;
  if (n_bsim3) *n_bsim3 = xg::bsim3_count(P);
  return XG_OK;
}

}  // extern "C"
