// BSIM3 host-side setup: parameter resolution, model processing, the
// size-dependent-parameter records, temperature, cards, stamps and the
// structure-of-arrays plan the eval kernel reads.
//
// The formulas are the reference's, operation for operation (results are
// compared to 1e-12):
//   Model::Model / processParams     N_DEV_MOSFET_B3.C:8481-9100
//   Instance::processParams           N_DEV_MOSFET_B3.C:1845-1880
//   Instance::updateTemperature       N_DEV_MOSFET_B3.C:3611-4518
// including the SizeDependParam cache keyed by (model, L, W) whose first
// instance's temperature wins.  What is this library's own: one record per
// cache entry shared by every instance that hits it, cards deduplicated by
// content, and the per-instance junction geometry folded at setup.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <numeric>
#include <string>
#include <unordered_map>

#include "xg_bsim3.h"
#include "xg_bsim3_layout.h"
#include "xg_setup.h"

namespace xg {

namespace {
constexpr double kQ = 1.6021918e-19;
constexpr double kBoltz = 1.3806226e-23;
constexpr double kRefTemp = 300.15;
constexpr double kKoverQ = kBoltz / kQ;
constexpr double kEg0 = 1.16, kAlphaEg = 7.02e-4, kBetaEg = 1108.0;
constexpr double kNi0 = 1.45e10;
constexpr double kEpsOx = 3.453133e-11, kEpsSi = 1.03594e-10;
constexpr double kExpThreshold = 34.0, kMinExp = 1.713908431e-15;
const double kVt0 = kBoltz * (27.0 + 273.15) / kQ;

enum : int {
#define XG_B3P(N, f, d) P_##N,
#include "b3_params.def"
#undef XG_B3P
  P_COUNT
};
const char *kNames[P_COUNT] = {
#define XG_B3P(N, f, d) #N,
#include "b3_params.def"
#undef XG_B3P
};
const double kDefaults[P_COUNT] = {
#define XG_B3P(N, f, d) (d),
#include "b3_params.def"
#undef XG_B3P
};
}  // namespace

struct Bsim3Model {
  int dtype = 1;
  double v[P_COUNT];
  bool g[P_COUNT];
  double cox, vcrit, factor1, Vtm0, Eg0, ni;
  double &at(int k) { return v[k]; }
  double at(int k) const { return v[k]; }
};

// One SizeDependParam cache entry.
struct B3Sdp {
  double L, W, temp;
  double r[B3R_N];
};

struct Bsim3Inst {
  int model;
  ParamSet p;
  std::array<int32_t, 6> lids;
  int32_t sta, sto;
  double l, w, ad, as, pd, ps, nrd, nrs, m, temp;
  bool off;
  double gd, gs;
  int sdp;   // index into Bsim3Plan::sdps
  int card;
  double geo[B3I_N];
};

struct Bsim3Plan {
  int32_t n = 0, slots = 0;
  std::vector<B3Sdp> sdps;
  std::vector<std::vector<int>> sdp_of_model;  // model -> sdp indices
  std::vector<int32_t> inst_of, wave_card, lid4, lid_int, sta, sto, flags,
      sdp_idx;
  std::vector<double> cards, sdp_soa, geo_soa, rec_blk;
  std::vector<int32_t> offsets;
  std::vector<double> fields_ref;  // per instance: B3R_N sdp + B3I_N geo
  int64_t slot_base = 0;
};

void bsim3_free(Circuit &c) {
  for (auto *m : c.b3_models) delete m;
  for (auto *i : c.b3) delete i;
  c.b3_models.clear();
  c.b3.clear();
}
void bsim3_plan_free(Plan &plan) {
  delete plan.b3;
  plan.b3 = nullptr;
}

// Model::Model + processParams.
static int b3_model_process(Bsim3Model &m, const Options &opt,
                            std::string &err) {
  auto &v = m.v;
  auto G = [&](int k) { return m.g[k]; };
  if (G(P_TNOM)) v[P_TNOM] += 273.15;
  else v[P_TNOM] = opt.tnom;
  if (v[P_CAPMOD] != 0.0 && v[P_CAPMOD] != 1.0 && v[P_CAPMOD] != 2.0 &&
      v[P_CAPMOD] != 3.0) {
    err = "BSIM3 capMod must be 0, 1, 2 or 3";
    return XG_E_PARAM;
  }
  if (!G(P_VTH0)) v[P_VTH0] = (m.dtype == 1) ? 0.7 : -0.7;
  if (!G(P_UC)) v[P_UC] = (v[P_MOBMOD] == 3) ? -0.0465 : -0.0465e-9;
  if (!G(P_UC1)) v[P_UC1] = (v[P_MOBMOD] == 3) ? -0.056 : -0.056e-9;
  if (!G(P_U0)) v[P_U0] = (m.dtype == 1) ? 0.067 : 0.025;
  const double tox = v[P_TOX];
  m.cox = 3.453133e-11 / tox;
  if (!G(P_TOXM)) v[P_TOXM] = tox;
  if (!G(P_DSUB)) v[P_DSUB] = v[P_DROUT];
  if (!G(P_LLC)) v[P_LLC] = v[P_LL];
  if (!G(P_LWC)) v[P_LWC] = v[P_LW];
  if (!G(P_LWLC)) v[P_LWLC] = v[P_LWL];
  if (!G(P_WLC)) v[P_WLC] = v[P_WL];
  // The reference tests WWL here (not WWLC) before its WWLC default; both
  // statements are kept so that a WWLC given without WWL is overridden the
  // same way (N_DEV_MOSFET_B3.C:8494-8495).
  if (!G(P_WWL)) v[P_WWLC] = v[P_WWL];
  if (!G(P_WWLC)) v[P_WWLC] = v[P_WWL];
  if (!G(P_DWC)) v[P_DWC] = v[P_WINT];
  if (!G(P_DLC)) v[P_DLC] = v[P_LINT];
  if (!G(P_CF)) {
    const double C1 = 2.0 * kEpsOx;
    const double C5 = M_PI;
    const double C2 = 1.0 + (0.4e-6 / tox);
    const double C3 = std::log(C2);
    v[P_CF] = C1 * C3 / C5;
  }
  if (!G(P_CGDO)) {
    if (G(P_DLC) && (v[P_DLC] > 0.0)) v[P_CGDO] = v[P_DLC] * m.cox - v[P_CGDL];
    else v[P_CGDO] = 0.6 * v[P_XJ] * m.cox;
  }
  if (!G(P_CGSO)) {
    if (G(P_DLC) && (v[P_DLC] > 0.0)) v[P_CGSO] = v[P_DLC] * m.cox - v[P_CGSL];
    else v[P_CGSO] = 0.6 * v[P_XJ] * m.cox;
  }
  if (!G(P_CGBO)) v[P_CGBO] = 2.0 * v[P_DWC] * m.cox;
  if (!G(P_CJSWG)) v[P_CJSWG] = v[P_CJSW];
  if (!G(P_PBSWG)) v[P_PBSWG] = v[P_PBSW];
  if (!G(P_MJSWG)) v[P_MJSWG] = v[P_MJSW];
  if (v[P_PB] < 0.1) v[P_PB] = 0.1;
  if (v[P_PBSW] < 0.1) v[P_PBSW] = 0.1;
  if (v[P_PBSWG] < 0.1) v[P_PBSWG] = 0.1;
  m.vcrit = kVt0 * std::log(kVt0 / (std::sqrt(2.0) * 1.0e-14));
  m.factor1 = std::sqrt(kEpsSi / kEpsOx * tox);
  const double tnom = v[P_TNOM];
  m.Vtm0 = kKoverQ * tnom;
  m.Eg0 = kEg0 - kAlphaEg * tnom * tnom / (tnom + kBetaEg);
  m.ni = kNi0 * (tnom / kRefTemp) * std::sqrt(tnom / kRefTemp) *
         std::exp(21.5565981 - m.Eg0 / (2.0 * m.Vtm0));
  return XG_OK;
}

int bsim3_add_model(Circuit &c, int dtype, const ParamSet &p, int32_t *id,
                    std::string &err) {
  auto *m = new Bsim3Model;
  m->dtype = dtype;
  for (int k = 0; k < P_COUNT; ++k) {
    m->v[k] = kDefaults[k];
    m->g[k] = false;
  }
  // name -> parameter index, built once (thread-safe static init)
  static const std::unordered_map<std::string, int> index = [] {
    std::unordered_map<std::string, int> m;
    for (int k = 0; k < P_COUNT; ++k) m.emplace(kNames[k], k);
    return m;
  }();
  for (auto &kv : p.v) {
    if (kv.first == "LEVEL") continue;
    const auto it = index.find(kv.first);
    const int k = it == index.end() ? P_COUNT : it->second;
    if (k == P_COUNT) {
      err = "unknown BSIM3 model parameter " + kv.first;
      delete m;
      return XG_E_ARG;
    }
    m->v[k] = kv.second;
    m->g[k] = true;
  }
  int rc = b3_model_process(*m, c.opt, err);
  if (rc) {
    delete m;
    return rc;
  }
  *id = (int32_t)c.b3_models.size();
  c.b3_models.push_back(m);
  return XG_OK;
}

int bsim3_add_instance(Circuit &c, int model, const ParamSet &p,
                       const int32_t lids[6], int32_t sta, int32_t sto,
                       int32_t *id, std::string &err) {
  if (model < 0 || model >= (int)c.b3_models.size()) {
    err = "bad BSIM3 model id";
    return XG_E_ARG;
  }
  static const char *keys[] = {"L",  "W",  "AD", "AS",  "NRD", "NRS",
                               "PD", "PS", "M",  "OFF", "TEMP"};
  for (auto &kv : p.v) {
    bool ok = false;
    for (const char *k : keys) ok = ok || kv.first == k;
    if (!ok) {
      if (kv.first == "IC1" || kv.first == "IC2" || kv.first == "IC3" ||
          kv.first == "NQSMOD") {
        err = "BSIM3 " + kv.first + " (extra unknowns) is not supported";
        return XG_E_PARAM;
      }
      err = "unknown BSIM3 instance parameter " + kv.first;
      return XG_E_ARG;
    }
  }
  const Bsim3Model &m = *c.b3_models[model];
  auto *in = new Bsim3Inst;
  in->model = model;
  in->p = p;
  for (int k = 0; k < 6; ++k) in->lids[k] = lids[k] < 0 ? -1 : lids[k];
  in->sta = sta;
  in->sto = sto;
  in->l = p.given("L") ? p.get("L", 0) : m.at(P_L);
  in->w = p.given("W") ? p.get("W", 0) : m.at(P_W);
  in->ad = p.get("AD", 0.0);
  in->as = p.get("AS", 0.0);
  in->pd = p.get("PD", 0.0);
  in->ps = p.get("PS", 0.0);
  in->nrd = p.get("NRD", 1.0);
  in->nrs = p.get("NRS", 1.0);
  in->m = p.get("M", 1.0);
  in->off = p.get("OFF", 0.0) != 0.0;
  in->temp = p.given("TEMP") ? p.get("TEMP", 0) + 273.15 : c.opt.temp;
  // Instance::processParams (sheet resistance)
  double gd = m.at(P_RSH) * in->nrd;
  in->gd = gd > 0.0 ? 1.0 / gd : 0.0;
  double gs = m.at(P_RSH) * in->nrs;
  in->gs = gs > 0.0 ? 1.0 / gs : 0.0;
  in->sdp = -1;
  in->card = -1;
  *id = (int32_t)c.b3.size();
  c.b3.push_back(in);
  return XG_OK;
}

// The size-dependent record of one cache entry (updateTemperature's
// "new SizeDependParam" branch).
static void b3_sdp_compute(const Bsim3Model &mm, double l, double w,
                           double temp, double *r) {
  const auto &m = mm.v;
  const double Tnom = m[P_TNOM];
  const double TRatio = temp / Tnom;
  const double Vtm0 = mm.Vtm0, ni = mm.ni, tox = m[P_TOX], cox = mm.cox;
  double T0, T1, T2, T3, T4, T5, tmp, tmp1, tmp2, tmp3;
  T0 = std::pow(l, m[P_LLN]);
  T1 = std::pow(w, m[P_LWN]);
  tmp1 = m[P_LL] / T0 + m[P_LW] / T1 + m[P_LWL] / (T0 * T1);
  const double dl = m[P_LINT] + tmp1;
  tmp2 = m[P_LLC] / T0 + m[P_LWC] / T1 + m[P_LWLC] / (T0 * T1);
  const double dlc = m[P_DLC] + tmp2;
  T2 = std::pow(l, m[P_WLN]);
  T3 = std::pow(w, m[P_WWN]);
  tmp1 = m[P_WL] / T2 + m[P_WW] / T3 + m[P_WWL] / (T2 * T3);
  const double dw = m[P_WINT] + tmp1;
  tmp2 = m[P_WLC] / T2 + m[P_WWC] / T3 + m[P_WWLC] / (T2 * T3);
  const double dwc = m[P_DWC] + tmp2;
  const double leff = l - 2.0 * dl;
  const double weff = w - 2.0 * dw;
  const double leffCV = l - 2.0 * dlc;
  const double weffCV = w - 2.0 * dwc;
  double Inv_L, Inv_W, Inv_LW;
  if (m[P_BINUNIT] == 1) {
    Inv_L = 1.0e-6 / leff;
    Inv_W = 1.0e-6 / weff;
    Inv_LW = 1.0e-12 / (leff * weff);
  } else {
    Inv_L = 1.0 / leff;
    Inv_W = 1.0 / weff;
    Inv_LW = 1.0 / (leff * weff);
  }
  // Binned parameter X = X + LX/L + WX/W + PX/(LW).
#define BIN(X) \
  (m[P_##X] + m[P_L##X] * Inv_L + m[P_W##X] * Inv_W + m[P_P##X] * Inv_LW)
  double cdsc = BIN(CDSC), cdscb = BIN(CDSCB), cdscd = BIN(CDSCD),
         cit = BIN(CIT), nfactor = BIN(NFACTOR), xj = BIN(XJ),
         vsat = BIN(VSAT), at = BIN(AT), a0 = BIN(A0), ags = BIN(AGS),
         a1 = BIN(A1), a2 = BIN(A2), keta = BIN(KETA), nsub = BIN(NSUB),
         npeak = BIN(NCH), ngate = BIN(NGATE), gamma1 = BIN(GAMMA1),
         gamma2 = BIN(GAMMA2), vbx = BIN(VBX), vbm = BIN(VBM), xt = BIN(XT),
         vfb = BIN(VFB), k1 = BIN(K1), kt1 = BIN(KT1), kt1l = BIN(KT1L),
         k2 = BIN(K2), kt2 = BIN(KT2), k3 = BIN(K3), k3b = BIN(K3B),
         w0 = BIN(W0), nlx = BIN(NLX), dvt0 = BIN(DVT0), dvt1 = BIN(DVT1),
         dvt2 = BIN(DVT2), dvt0w = BIN(DVT0W), dvt1w = BIN(DVT1W),
         dvt2w = BIN(DVT2W), drout = BIN(DROUT), dsub = BIN(DSUB),
         vth0 = BIN(VTH0), ua = BIN(UA), ua1 = BIN(UA1), ub = BIN(UB),
         ub1 = BIN(UB1), uc = BIN(UC), uc1 = BIN(UC1), u0 = BIN(U0),
         ute = BIN(UTE), voff = BIN(VOFF), delta = BIN(DELTA),
         rdsw = BIN(RDSW), prwg = BIN(PRWG), prwb = BIN(PRWB), prt = BIN(PRT),
         eta0 = BIN(ETA0), etab = BIN(ETAB), pclm = BIN(PCLM),
         pdibl1 = BIN(PDIBLC1), pdibl2 = BIN(PDIBLC2), pdiblb = BIN(PDIBLCB),
         pscbe1 = BIN(PSCBE1), pscbe2 = BIN(PSCBE2), pvag = BIN(PVAG),
         wr = BIN(WR), dwg = BIN(DWG), dwb = BIN(DWB), b0 = BIN(B0),
         b1 = BIN(B1), alpha0 = BIN(ALPHA0), alpha1 = BIN(ALPHA1),
         beta0 = BIN(BETA0), elm = BIN(ELM), cgsl = BIN(CGSL),
         cgdl = BIN(CGDL), ckappa = BIN(CKAPPA), cf = BIN(CF), clc = BIN(CLC),
         cle = BIN(CLE), acde = BIN(ACDE), moin = BIN(MOIN), noff = BIN(NOFF),
         voffcv = BIN(VOFFCV), vfbcv = BIN(VFBCV);
#undef BIN
  const double abulkCVfactor = 1.0 + std::pow((clc / leffCV), cle);
  T0 = (TRatio - 1.0);
  ua = ua + ua1 * T0;
  ub = ub + ub1 * T0;
  uc = uc + uc1 * T0;
  if (u0 > 1.0) u0 = u0 / 1.0e4;
  const double u0temp = u0 * std::pow(TRatio, ute);
  const double vsattemp = vsat - at * T0;
  const double rds0 = (rdsw + prt * T0) / std::pow(weff * 1E6, wr);
  const double cgdo = (m[P_CGDO] + cf) * weffCV;
  const double cgso = (m[P_CGSO] + cf) * weffCV;
  const double cgbo = m[P_CGBO] * leffCV;
  (void)elm;
  if (!mm.g[P_NCH] && mm.g[P_GAMMA1]) {
    T0 = gamma1 * cox;
    npeak = 3.021E22 * T0 * T0;
  }
  const double phi = 2.0 * Vtm0 * std::log(npeak / ni);
  const double sqrtPhi = std::sqrt(phi);
  const double phis3 = sqrtPhi * phi;
  const double Xdep0 = std::sqrt(2.0 * kEpsSi / (kQ * npeak * 1.0e6)) * sqrtPhi;
  const double litl = std::sqrt(3.0 * xj * tox);
  const double vbi = Vtm0 * std::log(1.0e20 * npeak / (ni * ni));
  const double cdep0 = std::sqrt(kQ * kEpsSi * npeak * 1.0e6 / 2.0 / phi);
  const double ldeb = std::sqrt(kEpsSi * Vtm0 / (kQ * npeak * 1.0e6)) / 3.0;
  acde *= std::pow((npeak / 2.0e16), -0.25);
  if (mm.g[P_K1] || mm.g[P_K2]) {
    if (!mm.g[P_K1]) k1 = 0.53;
    if (!mm.g[P_K2]) k2 = -0.0186;
  } else {
    if (!mm.g[P_VBX]) vbx = phi - 7.7348e-4 * npeak * xt * xt;
    if (vbx > 0.0) vbx = -vbx;
    if (vbm > 0.0) vbm = -vbm;
    if (!mm.g[P_GAMMA1]) gamma1 = 5.753e-12 * std::sqrt(npeak) / cox;
    if (!mm.g[P_GAMMA2]) gamma2 = 5.753e-12 * std::sqrt(nsub) / cox;
    T0 = gamma1 - gamma2;
    T1 = std::sqrt(phi - vbx) - sqrtPhi;
    T2 = std::sqrt(phi * (phi - vbm)) - phi;
    k2 = T0 * T1 / (2.0 * T2 + vbm);
    k1 = gamma2 - 2.0 * k2 * std::sqrt(phi - vbm);
  }
  double vbsc;
  if (k2 < 0.0) {
    T0 = 0.5 * k1 / k2;
    vbsc = 0.9 * (phi - T0 * T0);
    if (vbsc > -3.0) vbsc = -3.0;
    else if (vbsc < -30.0) vbsc = -30.0;
  } else {
    vbsc = -30.0;
  }
  if (vbsc > vbm) vbsc = vbm;
  if (!mm.g[P_VFB]) {
    if (mm.g[P_VTH0]) vfb = mm.dtype * vth0 - phi - k1 * sqrtPhi;
    else vfb = -1.0;
  }
  if (!mm.g[P_VTH0]) vth0 = mm.dtype * (vfb + phi + k1 * sqrtPhi);
  const double k1ox = k1 * tox / m[P_TOXM];
  const double k2ox = k2 * tox / m[P_TOXM];
  T1 = std::sqrt(kEpsSi / kEpsOx * tox * Xdep0);
  T0 = std::exp(-0.5 * dsub * leff / T1);
  const double theta0vb0 = (T0 + 2.0 * T0 * T0);
  T0 = std::exp(-0.5 * drout * leff / T1);
  T2 = (T0 + 2.0 * T0 * T0);
  const double thetaRout = pdibl1 * T2 + pdibl2;
  tmp = std::sqrt(Xdep0);
  tmp1 = vbi - phi;
  tmp2 = mm.factor1 * tmp;
  T0 = -0.5 * dvt1w * weff * leff / tmp2;
  T1 = T0 > -kExpThreshold ? std::exp(T0) : kMinExp;
  T2 = T1 * (1.0 + 2.0 * T1);
  T0 = dvt0w * T2;
  T2 = T0 * tmp1;
  T0 = -0.5 * dvt1 * leff / tmp2;
  T1 = T0 > -kExpThreshold ? std::exp(T0) : kMinExp;
  T3 = T1 * (1.0 + 2.0 * T1);
  T3 = dvt0 * T3 * tmp1;
  T4 = tox * phi / (weff + w0);
  T0 = std::sqrt(1.0 + nlx / leff);
  T5 = k1ox * (T0 - 1.0) * sqrtPhi + (kt1 + kt1l / leff) * (TRatio - 1.0);
  tmp3 = mm.dtype * vth0 - T2 - T3 + k3 * T4 + T5;
  const double vfbzb = tmp3 - phi - k1 * sqrtPhi;

  r[B3R_LEFF] = leff; r[B3R_WEFF] = weff; r[B3R_LEFFCV] = leffCV;
  r[B3R_WEFFCV] = weffCV; r[B3R_VTH0] = vth0; r[B3R_VBSC] = vbsc;
  r[B3R_PHI] = phi; r[B3R_PHIS3] = phis3; r[B3R_XDEP0] = Xdep0;
  r[B3R_SQRTPHI] = sqrtPhi; r[B3R_VBI] = vbi; r[B3R_DVT2] = dvt2;
  r[B3R_DVT2W] = dvt2w; r[B3R_DVT1] = dvt1; r[B3R_DVT0] = dvt0;
  r[B3R_DVT1W] = dvt1w; r[B3R_DVT0W] = dvt0w; r[B3R_NLX] = nlx;
  r[B3R_K1OX] = k1ox; r[B3R_KT1] = kt1; r[B3R_KT1L] = kt1l; r[B3R_KT2] = kt2;
  r[B3R_W0] = w0; r[B3R_ETA0] = eta0; r[B3R_ETAB] = etab;
  r[B3R_THETA0VB0] = theta0vb0; r[B3R_K1] = k1; r[B3R_K2OX] = k2ox;
  r[B3R_K3] = k3; r[B3R_K3B] = k3b; r[B3R_NFACTOR] = nfactor;
  r[B3R_CDSC] = cdsc; r[B3R_CDSCB] = cdscb; r[B3R_CDSCD] = cdscd;
  r[B3R_CIT] = cit; r[B3R_VFB] = vfb; r[B3R_NGATE] = ngate;
  r[B3R_VOFF] = voff; r[B3R_CDEP0] = cdep0; r[B3R_DWG] = dwg;
  r[B3R_DWB] = dwb; r[B3R_PRWG] = prwg; r[B3R_PRWB] = prwb;
  r[B3R_RDS0] = rds0; r[B3R_XJ] = xj; r[B3R_A0] = a0; r[B3R_B1] = b1;
  r[B3R_B0] = b0; r[B3R_AGS] = ags; r[B3R_KETA] = keta; r[B3R_UA] = ua;
  r[B3R_UC] = uc; r[B3R_UB] = ub; r[B3R_U0TEMP] = u0temp;
  r[B3R_VSATTEMP] = vsattemp; r[B3R_A1] = a1; r[B3R_A2] = a2;
  r[B3R_DELTA] = delta; r[B3R_PCLM] = pclm; r[B3R_LITL] = litl;
  r[B3R_THETAROUT] = thetaRout; r[B3R_PDIBLB] = pdiblb; r[B3R_PVAG] = pvag;
  r[B3R_PSCBE1] = pscbe1; r[B3R_PSCBE2] = pscbe2; r[B3R_ALPHA0] = alpha0;
  r[B3R_ALPHA1] = alpha1; r[B3R_BETA0] = beta0;
  r[B3R_VFBQ] = (m[P_CAPMOD] == 0.0) ? vfbcv : vfbzb;
  r[B3R_ACDE] = acde; r[B3R_LDEB] = ldeb; r[B3R_MOIN] = moin;
  r[B3R_ABULKCVF] = abulkCVfactor; r[B3R_NOFF] = noff;
  r[B3R_VOFFCV] = voffcv; r[B3R_CGBO] = cgbo; r[B3R_CGDO] = cgdo;
  r[B3R_CGSO] = cgso; r[B3R_CGDL] = cgdl; r[B3R_CGSL] = cgsl;
  r[B3R_CKAPPA] = ckappa; r[B3R_TOX] = tox; r[B3R_COX] = cox;
  r[B3R_FACTOR1] = mm.factor1;
}

// Temperature-level quantities of one instance (updateTemperature before
// the size-dependent branch) and its junction geometry.
struct B3Temp {
  double vtm, jsat, jswsat, cj, cjsw, cjswg, phib, phibsw, phibswg;
};

static B3Temp b3_temperature(const Bsim3Model &mm, double temp) {
  const auto &m = mm.v;
  B3Temp t{};
  const double Tnom = m[P_TNOM];
  t.vtm = kKoverQ * temp;
  const double Eg = kEg0 - kAlphaEg * temp * temp / (temp + kBetaEg);
  if (temp != Tnom) {
    double T0 = mm.Eg0 / mm.Vtm0 - Eg / t.vtm + m[P_XTI] * std::log(temp / Tnom);
    double T1 = std::exp(T0 / m[P_NJ]);
    t.jsat = m[P_JS] * T1;
    t.jswsat = m[P_JSW] * T1;
  } else {
    t.jsat = m[P_JS];
    t.jswsat = m[P_JSW];
  }
  if (t.jsat < 0.0) t.jsat = 0.0;
  if (t.jswsat < 0.0) t.jswsat = 0.0;
  const double delTemp = temp - Tnom;
  // A temperature coefficient driving a capacitance below zero leaves it at
  // the value of a freshly built instance, zero (N_DEV_MOSFET_B3.C:3664-3700).
  double T0 = m[P_TCJ] * delTemp;
  t.cj = T0 >= -1.0 ? m[P_CJ] * (1.0 + T0) : 0.0;
  T0 = m[P_TCJSW] * delTemp;
  t.cjsw = T0 >= -1.0 ? m[P_CJSW] * (1.0 + T0) : 0.0;
  T0 = m[P_TCJSWG] * delTemp;
  t.cjswg = T0 >= -1.0 ? m[P_CJSWG] * (1.0 + T0) : 0.0;
  t.phib = m[P_PB] - m[P_TPB] * delTemp;
  if (t.phib < 0.01) t.phib = 0.01;
  t.phibsw = m[P_PBSW] - m[P_TPBSW] * delTemp;
  if (t.phibsw <= 0.01) t.phibsw = 0.01;
  t.phibswg = m[P_PBSWG] - m[P_TPBSWG] * delTemp;
  if (t.phibswg <= 0.01) t.phibswg = 0.01;
  return t;
}

static void b3_card(const Bsim3Model &mm, double temp, const Options &opt,
                    double *c) {
  const auto &m = mm.v;
  const B3Temp t = b3_temperature(mm, temp);
  std::fill(c, c + B3C_N, 0.0);
  c[B3C_DTYPE] = mm.dtype;
  c[B3C_MOBMOD] = m[P_MOBMOD];
  c[B3C_XPART] = m[P_XPART];
  c[B3C_VCRIT] = mm.vcrit;
  c[B3C_NVTM] = t.vtm * m[P_NJ];
  c[B3C_VTM] = t.vtm;
  c[B3C_IJTH] = m[P_IJTH];
  c[B3C_TEMPRATIO] = temp / m[P_TNOM] - 1.0;
  c[B3C_PHIB] = t.phib;
  c[B3C_PHIBSW] = t.phibsw;
  c[B3C_PHIBSWG] = t.phibswg;
  c[B3C_MJ] = m[P_MJ];
  c[B3C_MJSW] = m[P_MJSW];
  c[B3C_MJSWG] = m[P_MJSWG];
  c[B3C_GMIN] = opt.gmin;
  c[B3C_LIMITER] = opt.limiter ? 1.0 : 0.0;
  c[B3C_CAPMOD] = m[P_CAPMOD];
}

int bsim3_prepare(Circuit &c, std::string &err) {
  if (c.b3.empty()) return XG_OK;
  for (auto *in : c.b3) {
    for (int k = 0; k < 6; ++k)
      if (in->lids[k] >= c.n_nodes) {
        err = "BSIM3 node id out of range";
        return XG_E_ARG;
      }
    if (in->sta < 0 || in->sta + 12 > c.n_state || in->sto < 0 ||
        in->sto + 5 > c.n_store) {
      err = "BSIM3 state/store range out of bounds";
      return XG_E_ARG;
    }
  }
  return XG_OK;
}

static inline int32_t b3_node(const Bsim3Inst &in, int local) {
  if (local == 4 && in.gd == 0.0) local = 0;
  if (local == 5 && in.gs == 0.0) local = 2;
  return in.lids[local];
}

// The stamp and its 22 offsets are MOSFET1's (N_DEV_MOSFET_B3.C:2378-2418,
// 3150-3330).
static const int kOffRC[22][2] = {
    {0, 0}, {0, 4}, {1, 1}, {1, 3}, {1, 4}, {1, 5}, {2, 2}, {2, 5},
    {3, 1}, {3, 3}, {3, 4}, {3, 5}, {4, 0}, {4, 1}, {4, 3}, {4, 4},
    {4, 5}, {5, 1}, {5, 2}, {5, 3}, {5, 4}, {5, 5}};
// dF/dx statements (N_DEV_MOSFET_B3.C:9660-9760) -> offset index.
static const int kJF[18] = {0,  1,  6,  7,  8,  9,  10, 11, 12,
                            14, 13, 15, 16, 17, 19, 18, 20, 21};
// dQ/dx statements (N_DEV_MOSFET_B3.C:9880-9915) -> offset index.
static const int kJQ[16] = {2,  3,  4,  5,  8,  9,  10, 11,
                            14, 13, 15, 16, 17, 19, 20, 21};

void bsim3_pattern_entries(const Circuit &c,
                           std::vector<std::array<int32_t, 2>> &entries) {
  const int32_t G = c.n_nodes;
  for (auto *in : c.b3)
    for (int k = 0; k < 22; ++k) {
      int32_t r = b3_node(*in, kOffRC[k][0]), cc = b3_node(*in, kOffRC[k][1]);
      entries.push_back({r < 0 ? G : r, cc < 0 ? G : cc});
    }
}

int bsim3_build(Circuit &c, Plan &plan, int64_t slot_base,
                std::vector<Contrib> *vec, std::vector<Contrib> *mat,
                std::string &err) {
  if (c.b3.empty()) return XG_OK;
  auto *bp = new Bsim3Plan;
  plan.b3 = bp;
  const int32_t n = (int32_t)c.b3.size();
  const int32_t G = c.n_nodes;
  auto L = [G](int32_t g) { return g < 0 ? G : g; };
  bp->n = n;
  bp->slot_base = slot_base;

  // size-dependent cache, in instance order (first instance's temperature)
  bp->sdp_of_model.assign(c.b3_models.size(), {});
  std::vector<std::vector<double>> card_rows;
  std::map<std::vector<double>, int> card_index;
  for (auto *in : c.b3) {
    const Bsim3Model &m = *c.b3_models[in->model];
    int found = -1;
    for (int s : bp->sdp_of_model[in->model])
      if (bp->sdps[s].L == in->l && bp->sdps[s].W == in->w) found = s;
    if (found < 0) {
      B3Sdp rec;
      rec.L = in->l;
      rec.W = in->w;
      rec.temp = in->temp;
      b3_sdp_compute(m, in->l, in->w, in->temp, rec.r);
      found = (int)bp->sdps.size();
      bp->sdps.push_back(rec);
      bp->sdp_of_model[in->model].push_back(found);
    }
    in->sdp = found;
    const double *r = bp->sdps[found].r;
    std::vector<double> card(B3C_N);
    b3_card(m, in->temp, c.opt, card.data());
    auto it = card_index.find(card);
    if (it == card_index.end()) {
      it = card_index.emplace(card, (int)card_rows.size()).first;
      card_rows.push_back(card);
    }
    in->card = it->second;
    // per-instance junction geometry (updateTemperature tail and the
    // junction section of updateIntermediateVars)
    const B3Temp t = b3_temperature(m, in->temp);
    const double Nvtm = t.vtm * m.at(P_NJ);
    double *g = in->geo;
    std::fill(g, g + B3I_N, 0.0);
    double ss = (in->as <= 0.0 && in->ps <= 0.0) ? 1.0e-14
                                                  : in->as * t.jsat + in->ps * t.jswsat;
    double ds = (in->ad <= 0.0 && in->pd <= 0.0) ? 1.0e-14
                                                  : in->ad * t.jsat + in->pd * t.jswsat;
    g[B3I_SSATCUR] = ss;
    g[B3I_DSATCUR] = ds;
    if (ss > 0.0 && m.at(P_IJTH) > 0.0) {
      g[B3I_VJSM] = Nvtm * std::log(m.at(P_IJTH) / ss + 1.0);
      g[B3I_ISEVJSM] = ss * std::exp(g[B3I_VJSM] / Nvtm);
    }
    if (ds > 0.0 && m.at(P_IJTH) > 0.0) {
      g[B3I_VJDM] = Nvtm * std::log(m.at(P_IJTH) / ds + 1.0);
      g[B3I_ISEVJDM] = ds * std::exp(g[B3I_VJDM] / Nvtm);
    }
    const double weff = r[B3R_WEFF];
    g[B3I_CZBD] = t.cj * in->ad;
    g[B3I_CZBS] = t.cj * in->as;
    if (in->pd < weff) {
      g[B3I_CZBDSWG] = t.cjswg * in->pd;
      g[B3I_CZBDSW] = 0.0;
    } else {
      g[B3I_CZBDSW] = t.cjsw * (in->pd - weff);
      g[B3I_CZBDSWG] = t.cjswg * weff;
    }
    if (in->ps < weff) {
      g[B3I_CZBSSW] = 0.0;
      g[B3I_CZBSSWG] = t.cjswg * in->ps;
    } else {
      g[B3I_CZBSSW] = t.cjsw * (in->ps - weff);
      g[B3I_CZBSSWG] = t.cjswg * weff;
    }
    g[B3I_GD] = in->gd;
    g[B3I_GS] = in->gs;
    g[B3I_M] = in->m;
  }

  // evaluation order: by card, then by size-dependent record; card groups
  // padded to whole wavefronts
  std::vector<int32_t> order(n), slot_of(n, -1);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) {
    if (c.b3[a]->card != c.b3[b]->card) return c.b3[a]->card < c.b3[b]->card;
    return c.b3[a]->sdp < c.b3[b]->sdp;
  });
  int32_t s = 0;
  for (size_t k = 0; k < order.size();) {
    const int card = c.b3[order[k]]->card;
    size_t e = k;
    while (e < order.size() && c.b3[order[e]]->card == card) ++e;
    for (size_t j = k; j < e; ++j) {
      slot_of[order[j]] = s++;
      bp->inst_of.push_back(order[j]);
    }
    while (s % 64) {
      bp->inst_of.push_back(-1);
      ++s;
    }
    k = e;
  }
  const int32_t S = s;
  bp->slots = S;
  for (int32_t w = 0; w < S / 64; ++w)
    bp->wave_card.push_back(c.b3[bp->inst_of[w * 64]]->card);
  bp->cards.clear();
  for (auto &row : card_rows) bp->cards.insert(bp->cards.end(), row.begin(), row.end());
  const int32_t nsdp = (int32_t)bp->sdps.size();
  bp->sdp_soa.assign((size_t)B3R_N * nsdp, 0.0);
  for (int32_t k = 0; k < nsdp; ++k)
    for (int f = 0; f < B3R_N; ++f) bp->sdp_soa[(size_t)f * nsdp + k] = bp->sdps[k].r[f];
  bp->geo_soa.assign((size_t)B3I_N * S, 0.0);
  bp->lid4.assign((size_t)4 * S, G);
  bp->lid_int.assign((size_t)2 * S, G);
  bp->sta.assign(S, 0);
  bp->sto.assign(S, 0);
  bp->flags.assign(S, 0);
  bp->sdp_idx.assign(S, 0);
  bp->fields_ref.assign((size_t)n * (B3R_N + B3I_N), 0.0);
  for (int32_t i = 0; i < n; ++i) {
    const Bsim3Inst &in = *c.b3[i];
    const int32_t sl = slot_of[i];
    for (int f = 0; f < B3I_N; ++f) bp->geo_soa[(size_t)f * S + sl] = in.geo[f];
    for (int k = 0; k < 4; ++k) bp->lid4[(size_t)k * S + sl] = L(in.lids[k]);
    bp->lid_int[sl] = L(b3_node(in, 4));
    bp->lid_int[(size_t)S + sl] = L(b3_node(in, 5));
    bp->sta[sl] = in.sta;
    bp->sto[sl] = in.sto;
    int32_t fl = 0;
    if (in.off) fl |= XF_OFF;
    if (in.gd != 0.0) fl |= XF_HAS_DP;
    if (in.gs != 0.0) fl |= XF_HAS_SP;
    bp->flags[sl] = fl;
    bp->sdp_idx[sl] = in.sdp;
    double *fr = &bp->fields_ref[(size_t)i * (B3R_N + B3I_N)];
    std::memcpy(fr, bp->sdps[in.sdp].r, sizeof(double) * B3R_N);
    std::memcpy(fr + B3R_N, in.geo, sizeof(double) * B3I_N);
  }
  // wave-blocked copy of the same fields: block w holds field f of lane l
  // at ((w * kRec) + f) * 64 + l
  {
    constexpr int kRec = B3R_N + B3I_N;
    bp->rec_blk.assign((size_t)S * kRec, 0.0);
    for (int32_t i = 0; i < n; ++i) {
      const int32_t sl = slot_of[i];
      const double *fr = &bp->fields_ref[(size_t)i * kRec];
      const size_t base = (size_t)(sl / 64) * kRec * 64 + (size_t)(sl % 64);
      for (int f = 0; f < kRec; ++f) bp->rec_blk[base + (size_t)f * 64] = fr[f];
    }
  }

  // offsets and reduction contributions in reference order
  auto csr_pos = [&](int32_t r, int32_t col, int32_t &pos) -> bool {
    const int32_t *b = plan.colidx.data() + plan.rowptr[r];
    const int32_t *e = plan.colidx.data() + plan.rowptr[r + 1];
    const int32_t *it = std::lower_bound(b, e, col);
    if (it == e || *it != col) return false;
    pos = (int32_t)(it - plan.colidx.data());
    return true;
  };
  auto SL = [&](int slot, int32_t sl) -> int32_t {
    return (int32_t)(slot_base + (int64_t)slot * S + sl);
  };
  bp->offsets.assign((size_t)n * 22, 0);
  for (int32_t i = 0; i < n; ++i) {
    const Bsim3Inst &in = *c.b3[i];
    const int32_t sl = slot_of[i];
    const bool dp = in.gd != 0.0, sp = in.gs != 0.0;
    const int32_t d = L(in.lids[0]), g = L(in.lids[1]), so = L(in.lids[2]),
                  b = L(in.lids[3]);
    const int32_t dpr = L(b3_node(in, 4)), spr = L(b3_node(in, 5));
    if (dp) vec[0].push_back({d, SL(B3S_F_D, sl)});
    if (sp) vec[0].push_back({so, SL(B3S_F_S, sl)});
    vec[0].push_back({b, SL(B3S_F_B, sl)});
    vec[0].push_back({dpr, SL(B3S_F_DP, sl)});
    vec[0].push_back({spr, SL(B3S_F_SP, sl)});
    vec[2].push_back({b, SL(B3S_FL_B, sl)});
    vec[2].push_back({dpr, SL(B3S_FL_DP, sl)});
    vec[2].push_back({spr, SL(B3S_FL_SP, sl)});
    vec[1].push_back({g, SL(B3S_Q_G, sl)});
    vec[1].push_back({b, SL(B3S_Q_B, sl)});
    vec[1].push_back({dpr, SL(B3S_Q_DP, sl)});
    vec[1].push_back({spr, SL(B3S_Q_SP, sl)});
    vec[3].push_back({g, SL(B3S_QL_G, sl)});
    vec[3].push_back({b, SL(B3S_QL_B, sl)});
    vec[3].push_back({dpr, SL(B3S_QL_DP, sl)});
    vec[3].push_back({spr, SL(B3S_QL_SP, sl)});
  }
  for (int32_t i = 0; i < n; ++i) {
    const Bsim3Inst &in = *c.b3[i];
    const int32_t sl = slot_of[i];
    int32_t pos[22];
    for (int k = 0; k < 22; ++k) {
      const int32_t r = L(b3_node(in, kOffRC[k][0]));
      const int32_t col = L(b3_node(in, kOffRC[k][1]));
      if (!csr_pos(r, col, pos[k])) {
        err = "supplied CSR pattern lacks a BSIM3 stamp entry";
        return XG_E_PATTERN;
      }
      bp->offsets[(size_t)i * 22 + k] = pos[k] - plan.rowptr[r];
    }
    const bool dp = in.gd != 0.0, sp = in.gs != 0.0;
    for (int e = 0; e < 18; ++e) {
      const bool dterm = (e == 0 || e == 1 || e == 8);
      const bool sterm = (e == 2 || e == 3 || e == 15);
      int32_t src = SL(B3S_JF0 + e, sl);
      if ((dterm && !dp) || (sterm && !sp)) {
        const bool plus = (e == 0 || e == 2);
        if (!plus) continue;
        src = -1;
      }
      mat[0].push_back({pos[kJF[e]], src});
    }
    for (int e = 0; e < 16; ++e) mat[1].push_back({pos[kJQ[e]], SL(B3S_JQ0 + e, sl)});
  }
  return XG_OK;
}

void bsim3_slot_nodes(const Plan &plan, int32_t ground, std::vector<int32_t> &out) {
  if (!plan.b3) return;
  const Bsim3Plan &b = *plan.b3;
  const int32_t S = b.slots;
  for (int32_t s = 0; s < S; ++s) {
    int32_t node = -1;
    if (b.inst_of[s] >= 0)
      for (int k = 0; k < 4 && node < 0; ++k)
        if (b.lid4[(size_t)k * S + s] != ground) node = b.lid4[(size_t)k * S + s];
    out.push_back(node);
  }
}

int64_t bsim3_slot_count(const Plan &plan) {
  return plan.b3 ? (int64_t)B3S_N * plan.b3->slots : 0;
}

int bsim3_upload(const Plan &P, Bsim3Dev &d, UpD upd, UpI upi) {
  if (!P.b3) return 0;
  const Bsim3Plan &b = *P.b3;
  d.n_slots = b.slots;
  d.n_sdp = (int32_t)b.sdps.size();
  d.slot_base = b.slot_base;
  bool ok = true;
  auto chk = [&](const void *p) { ok = ok && p != nullptr; };
  chk(d.inst_of = upi(b.inst_of));
  chk(d.wave_card = upi(b.wave_card));
  chk(d.cards = upd(b.cards));
  chk(d.sdp = upd(b.sdp_soa));
  chk(d.geo = upd(b.geo_soa));
  chk(d.lid4 = upi(b.lid4));
  chk(d.lid_int = upi(b.lid_int));
  chk(d.sta_base = upi(b.sta));
  chk(d.sto_base = upi(b.sto));
  chk(d.flags = upi(b.flags));
  chk(d.sdp_idx = upi(b.sdp_idx));
  chk(d.rec = upd(b.rec_blk));
  return ok ? 0 : -1;
}

void bsim3_bind(Bsim3Dev &d, double *slots, int32_t *limited) {
  d.slots = slots ? slots + d.slot_base : nullptr;
  d.limited = limited;
}
int32_t bsim3_eval_slots(const Plan &P) { return P.b3 ? P.b3->slots : 0; }
int32_t bsim3_count(const Plan &P) { return P.b3 ? P.b3->n : 0; }
int bsim3_get_offsets(const Plan &P, int32_t *off) {
  if (P.b3) std::memcpy(off, P.b3->offsets.data(), P.b3->offsets.size() * sizeof(int32_t));
  return XG_OK;
}
int bsim3_get_fields(const Plan &P, double *out, int32_t *n_fields) {
  *n_fields = B3R_N + B3I_N;
  if (out && P.b3)
    std::memcpy(out, P.b3->fields_ref.data(), P.b3->fields_ref.size() * sizeof(double));
  return XG_OK;
}

}  // namespace xg
