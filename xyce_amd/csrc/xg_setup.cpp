// Host-side setup for the device-load library: parameter resolution, MOSFET1
// temperature/geometry precompute, CSR pattern, stamp offsets, evaluation
// order and the ordered-reduction programs.
//
// The temperature and geometry formulas must be the reference's (results are
// compared to 1e-12), so they follow
//   MOSFET1::Model::processParams       N_DEV_MOSFET1.C:3786-3868
//   MOSFET1::Instance::Instance         N_DEV_MOSFET1.C:1170-1260
//   MOSFET1::Instance::updateTemperature N_DEV_MOSFET1.C:3437-3680
// operation for operation.  The data layout they produce (one card per
// model/temperature, one array per instance field) is this library's own.
#include "xg_setup.h"

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <functional>
#include <cctype>
#include <cmath>
#include <cstring>
#include <numeric>

namespace xg {

namespace {
constexpr double kQ = 1.6021918e-19;
constexpr double kBoltz = 1.3806226e-23;
constexpr double kRefTemp = 300.15;
constexpr double kKoverQ = kBoltz / kQ;
}  // namespace

// Diode level 1 with RS = 0 (the merged stamp): model defaults of
// N_DEV_Diode.C:97-247, Model::processParams :1772-1793,
// Instance::updateTemperature :1332-1470 (breakdown voltage by the
// default, non-"ORIGINAL" Picard iteration).  TNOM and TEMP are Celsius.
int diode_setup(const ParamSet &mp, const ParamSet &ip, const Options &opt,
                double *f, std::string &err) {
  static const char *kModel[] = {"LEVEL", "IS", "JS", "RS", "N", "TT", "CJO", "CJ",
                                 "CJ0", "VJ", "M", "EG", "XTI", "FC", "BV", "VB",
                                 "IBV", "IRF", "TNOM", "KF", "AF"};
  static const char *kInst[] = {"AREA", "IC", "TEMP", "OFF", "LAMBERTW"};
  for (auto &kv : mp.v) {
    bool ok = false;
    for (const char *k : kModel) ok = ok || kv.first == k;
    if (!ok) {
      err = "diode model parameter " + kv.first + " is not modelled here";
      return XG_E_PARAM;
    }
  }
  for (auto &kv : ip.v) {
    bool ok = false;
    for (const char *k : kInst) ok = ok || kv.first == k;
    if (!ok) {
      err = "diode instance parameter " + kv.first + " is not modelled here";
      return XG_E_PARAM;
    }
  }
  if (mp.get("LEVEL", 1.0) != 1.0 || mp.get("RS", 0.0) != 0.0 ||
      ip.get("LAMBERTW", 0.0) != 0.0) {
    err = "diode: only level 1, RS = 0, LAMBERTW = 0 are modelled";
    return XG_E_PARAM;
  }
  const double IS = mp.given("IS") ? mp.get("IS", 0) : mp.get("JS", 1.0e-14);
  const double N = mp.get("N", 1.0);
  const double TT = mp.get("TT", 0.0);
  const double CJO = mp.given("CJO") ? mp.get("CJO", 0)
                     : mp.given("CJ") ? mp.get("CJ", 0) : mp.get("CJ0", 0.0);
  const double VJ = mp.get("VJ", 1.0);
  double M = mp.get("M", 0.5), EG = mp.get("EG", 1.11), FC = mp.get("FC", 0.5);
  const double XTI = mp.get("XTI", 3.0);
  const bool bv_given = mp.given("BV") || mp.given("VB");
  const double BV = mp.given("BV") ? mp.get("BV", 0) : mp.get("VB", 1e99);
  const double IBV = mp.get("IBV", 1.0e-3), IRF = mp.get("IRF", 1.0);
  const double TNOM = mp.given("TNOM") ? mp.get("TNOM", 0) + 273.15 : opt.tnom;
  const double Temp = ip.given("TEMP") ? ip.get("TEMP", 0) + 273.15 : opt.temp;
  const double Area = ip.get("AREA", 1.0);
  // Model::processParams
  if (M > 0.9) M = 0.9;
  if (EG < 0.1) EG = 0.1;
  if (FC > 0.95) FC = 0.95;
  const double xfc_m = std::log(1.0 - FC);
  const double F2 = std::exp((1.0 + M) * xfc_m);
  const double F3 = 1.0 - FC * (1.0 + M);
  // Instance::updateTemperature
  const double Eg0 = 1.16, alphaEg = 7.02e-4, betaEg = 1108.0, Eg300 = 1.1150877;
  const double vtnom = kKoverQ * TNOM;
  const double xfc = std::log(1.0 - FC);
  const double vt = kKoverQ * Temp;
  const double fact2 = Temp / kRefTemp;
  const double egfet = Eg0 - (alphaEg * Temp * Temp) / (Temp + betaEg);
  const double arg = -egfet / (2.0 * kBoltz * Temp) +
                     Eg300 / (kBoltz * (kRefTemp + kRefTemp));
  const double pbfact = -2.0 * vt * (1.5 * std::log(fact2) + kQ * arg);
  const double egfet1 = Eg0 - (alphaEg * TNOM * TNOM) / (TNOM + betaEg);
  const double arg1 = -egfet1 / (2.0 * kBoltz * TNOM) + Eg300 / (2.0 * kBoltz * kRefTemp);
  const double fact1 = TNOM / kRefTemp;
  const double pbfact1 = -2.0 * vtnom * (1.5 * std::log(fact1) + kQ * arg1);
  const double pbo = (VJ - pbfact1) / fact1;
  const double gmaold = (VJ - pbo) / pbo;
  double tJctCap = CJO / (1.0 + M * (4.0e-4 * (TNOM - kRefTemp) - gmaold));
  const double tJctPot = pbfact + fact2 * pbo;
  const double gmanew = (tJctPot - pbo) / pbo;
  tJctCap *= 1.0 + M * (4.0e-4 * (Temp - kRefTemp) - gmanew);
  const double tSatCur = IS * std::exp(((Temp / TNOM) - 1.0) * EG / (N * vt) +
                                       XTI / N * std::log(Temp / TNOM));
  const double tF1 = tJctPot * (1.0 - std::exp((1.0 - M) * xfc)) / (1.0 - M);
  const double tDepCap = FC * tJctPot;
  const double vte = N * vt;
  const double tVcrit = vte * std::log(vte / (std::sqrt(2.0) * tSatCur));
  const double tIRF = IRF * std::pow(fact2, 1.6);
  double tBrkdwnV = 0.0;
  if (bv_given) {
    const double e = std::exp(1.0);
    const double cbv = IBV;
    double a1 = 3.0 * vte / (e * BV);
    a1 = a1 * a1 * a1;
    const double lo = tSatCur * tIRF * (1 - a1);
    const double hi = tSatCur * tIRF * (1 - 1.0 / (e * e * e)) *
                      std::exp(-1.0 * (3.0 * vte - BV) / vte);
    if (cbv >= hi) {
      tBrkdwnV = 3.0 * vte;
    } else if (cbv <= lo) {
      tBrkdwnV = BV;
    } else {
      double xbv = BV;
      for (int it = 0; it < 8; ++it) {
        double a2 = 3.0 * vte / (e * xbv);
        a2 = a2 * a2 * a2;
        xbv = BV - vte * std::log(cbv / (tSatCur * tIRF)) + vte * std::log(1 - a2);
      }
      tBrkdwnV = xbv;
    }
  }
  f[DIO_ISAT] = tSatCur * Area;
  f[DIO_VTE] = vte;
  f[DIO_VCRIT] = tVcrit;
  f[DIO_IRF] = tIRF;
  f[DIO_BRKV] = tBrkdwnV;
  f[DIO_BVGIVEN] = bv_given ? 1.0 : 0.0;
  f[DIO_BV] = BV;
  f[DIO_TJCAP] = tJctCap;
  f[DIO_CZERO] = tJctCap * Area;
  f[DIO_TJPOT] = tJctPot;
  f[DIO_TDEPCAP] = tDepCap;
  f[DIO_TF1] = tF1;
  f[DIO_M] = M;
  f[DIO_F2] = F2;
  f[DIO_F3] = F3;
  f[DIO_TT] = TT;
  f[DIO_GSPR] = 0.0;  // tCOND * AREA with COND = 0 for RS = 0
  f[DIO_OFF] = ip.get("OFF", 0.0) != 0.0 ? 1.0 : 0.0;
  f[DIO_HASIC] = ip.given("IC") ? 1.0 : 0.0;
  f[DIO_IC] = ip.get("IC", 0.0);
  return XG_OK;
}

Circuit::~Circuit() { bsim3_free(*this); }
Plan::~Plan() { bsim3_plan_free(*this); }

int parse_params(int32_t n, const char *const *keys, const double *vals,
                 ParamSet &out) {
  if (n < 0 || (n > 0 && (!keys || !vals))) return XG_E_ARG;
  for (int32_t i = 0; i < n; ++i) {
    if (!keys[i]) return XG_E_ARG;
    std::string k(keys[i]);
    for (auto &ch : k) ch = (char)std::toupper((unsigned char)ch);
    out.v[k] = vals[i];
  }
  return XG_OK;
}

// ---------------------------------------------------------------------------
// MOSFET1 model: defaults of the parameter table (N_DEV_MOSFET1.C:117-263),
// constructor fix-ups (N_DEV_MOSFET1.C:3950-3980), processParams.
static const char *kMos1ModelKeys[] = {
    "L", "W", "VTO", "KP", "GAMMA", "PHI", "LAMBDA", "RD", "RS", "CBD", "CBS",
    "IS", "PB", "CGSO", "CGDO", "CGBO", "RSH", "CJ", "MJ", "CJSW", "MJSW",
    "JS", "TOX", "LD", "UO", "U0", "FC", "NSUB", "NSS", "TNOM", "KF", "AF",
    "TPG", "LEVEL"};

int mos1_model_process(Mos1Model &m, const Options &opt, std::string &err) {
  for (auto &kv : m.p.v) {
    bool ok = false;
    for (const char *k : kMos1ModelKeys) ok = ok || kv.first == k;
    if (!ok) {
      err = "unknown MOSFET1 model parameter " + kv.first;
      return XG_E_ARG;
    }
  }
  const ParamSet &p = m.p;
  m.model_l = p.get("L", opt.tnom > 0 ? 1e-4 : 1e-4);
  m.model_w = p.get("W", 1e-4);
  m.vt0 = p.get("VTO", 0.0);
  m.kp = p.get("KP", 2e-5);
  m.gamma = p.get("GAMMA", 0.0);
  m.phi = p.get("PHI", 0.6);
  m.lambda = p.get("LAMBDA", 0.0);
  m.rd = p.get("RD", 0.0);
  m.rs = p.get("RS", 0.0);
  m.capBD = p.get("CBD", 0.0);
  m.capBS = p.get("CBS", 0.0);
  m.jctSatCur = p.get("IS", 1e-14);
  m.pb = p.get("PB", 0.8);
  m.cgso = p.get("CGSO", 0.0);
  m.cgdo = p.get("CGDO", 0.0);
  m.cgbo = p.get("CGBO", 0.0);
  m.rsh = p.get("RSH", 0.0);
  m.cj = p.get("CJ", 0.0);
  m.mj = p.get("MJ", 0.5);
  m.cjsw = p.get("CJSW", 0.0);
  m.mjsw = p.get("MJSW", 0.5);
  m.jctSatCurDensity = p.get("JS", 0.0);
  m.tox = p.get("TOX", 1e-7);
  m.latDiff = p.get("LD", 0.0);
  m.uo = p.get("UO", 600.0);
  m.fc = p.get("FC", 0.5);
  m.nsub = p.get("NSUB", 0.0);
  m.nss = p.get("NSS", 0.0);
  // TNOM / TEMP are given in Celsius (N_DEV_DeviceEntity.C:863-866); the
  // circuit options carry Kelvin.
  m.tnom = p.given("TNOM") ? p.get("TNOM", 0.0) + 273.15 : opt.tnom;
  m.tpg = (int)p.get("TPG", 0.0);
  m.capBDGiven = p.given("CBD") || m.capBD != 0;
  m.capBSGiven = p.given("CBS") || m.capBS != 0;
  m.cjGiven = p.given("CJ");
  m.cjswGiven = p.given("CJSW");
  m.rshGiven = p.given("RSH");
  if (p.given("U0")) {
    if (p.given("UO")) {
      err = "both UO and U0 given";
      return XG_E_PARAM;
    }
    m.uo = p.get("U0", 600.0);
  }
  // processParams
  m.fact1 = m.tnom / kRefTemp;
  m.vtnom = m.tnom * kKoverQ;
  const double kt1 = kBoltz * m.tnom;
  m.egfet1 = 1.16 - (7.02e-4 * m.tnom * m.tnom) / (m.tnom + 1108);
  const double arg1 =
      -m.egfet1 / (kt1 + kt1) + 1.1150877 / (kBoltz * (kRefTemp + kRefTemp));
  m.pbfact1 = -2 * m.vtnom * (1.5 * std::log(m.fact1) + kQ * arg1);
  if (!p.given("TOX") || m.tox == 0) {
    m.oxideCapFactor = 0;
  } else {
    m.oxideCapFactor = 3.9 * 8.854214871e-12 / m.tox;
    if (!p.given("KP")) {
      if (!p.given("UO") && !p.given("U0")) m.uo = 600;
      m.kp = m.uo * m.oxideCapFactor * 1e-4;
    }
    if (p.given("NSUB")) {
      if (m.nsub * 1e6 > 1.45e16) {
        if (!p.given("PHI")) {
          m.phi = 2 * m.vtnom * std::log(m.nsub * 1e6 / 1.45e16);
          m.phi = std::max(0.1, m.phi);
        }
        const double fermis = m.dtype * .5 * m.phi;
        double wkfng = 3.2;
        if (!p.given("TPG")) m.tpg = 1;
        if (m.tpg != 0) {
          const double fermig = m.dtype * m.tpg * .5 * m.egfet1;
          wkfng = 3.25 + .5 * m.egfet1 - fermig;
        }
        const double wkfngs = wkfng - (3.25 + .5 * m.egfet1 + fermis);
        if (!p.given("GAMMA"))
          m.gamma = std::sqrt(2 * 11.70 * 8.854214871e-12 * kQ * m.nsub * 1e6) /
                    m.oxideCapFactor;
        if (!p.given("VTO")) {
          const double nss = p.given("NSS") ? m.nss : 0.0;
          const double vfb = wkfngs - nss * 1e4 * kQ / m.oxideCapFactor;
          m.vt0 = vfb + m.dtype * (m.gamma * std::sqrt(m.phi) + m.phi);
        }
      } else {
        err = "MOSFET1 NSUB < Ni";
        return XG_E_PARAM;
      }
    }
  }
  return XG_OK;
}

// Temperature-only quantities of one (model, temp) pair.
struct Mos1Temp {
  double vt, tKP, tPhi, tVbi, tVto, tSatCur, tSatCurDens, tCbd, tCbs, tCj,
      tCjsw, tBulkPot, tDepCap;
};

static Mos1Temp mos1_temperature(const Mos1Model &m, double temp) {
  Mos1Temp t{};
  const double tnom = m.tnom;
  const double ratio = temp / tnom;
  t.vt = temp * kKoverQ;
  const double fact2 = temp / kRefTemp;
  const double kt = temp * kBoltz;
  const double egfet = 1.16 - (7.02e-4 * temp * temp) / (temp + 1108);
  const double arg =
      -egfet / (kt + kt) + 1.1150877 / (kBoltz * (kRefTemp + kRefTemp));
  const double pbfact = -2 * t.vt * (1.5 * std::log(fact2) + kQ * arg);
  const double ratio4 = ratio * std::sqrt(ratio);
  t.tKP = m.kp / ratio4;
  const double phio = (m.phi - m.pbfact1) / m.fact1;
  t.tPhi = fact2 * phio + pbfact;
  t.tVbi = m.vt0 - m.dtype * (m.gamma * std::sqrt(m.phi)) +
           .5 * (m.egfet1 - egfet) + m.dtype * .5 * (t.tPhi - m.phi);
  t.tVto = t.tVbi + m.dtype * m.gamma * std::sqrt(t.tPhi);
  t.tSatCur = m.jctSatCur * std::exp(-egfet / t.vt + m.egfet1 / m.vtnom);
  t.tSatCurDens =
      m.jctSatCurDensity * std::exp(-egfet / t.vt + m.egfet1 / m.vtnom);
  const double pbo = (m.pb - m.pbfact1) / m.fact1;
  const double gmaold = (m.pb - pbo) / pbo;
  double capfact = 1 / (1 + m.mj * (4e-4 * (m.tnom - kRefTemp) - gmaold));
  t.tCbd = m.capBD * capfact;
  t.tCbs = m.capBS * capfact;
  t.tCj = m.cj * capfact;
  capfact = 1 / (1 + m.mjsw * (4e-4 * (m.tnom - kRefTemp) - gmaold));
  t.tCjsw = m.cjsw * capfact;
  t.tBulkPot = fact2 * pbo + pbfact;
  const double gmanew = (t.tBulkPot - pbo) / pbo;
  capfact = (1 + m.mj * (4e-4 * (temp - kRefTemp) - gmanew));
  t.tCbd *= capfact;
  t.tCbs *= capfact;
  t.tCj *= capfact;
  capfact = (1 + m.mjsw * (4e-4 * (temp - kRefTemp) - gmanew));
  t.tCjsw *= capfact;
  t.tDepCap = m.fc * t.tBulkPot;
  return t;
}

// Junction-capacitance polynomial coefficients f2/f3/f4 for one side.
static void mos1_junction(const Mos1Model &m, const Mos1Temp &t, double czb,
                          double czbsw, double &f2, double &f3, double &f4) {
  const double arg = 1 - m.fc;
  const double sarg = std::exp((-m.mj) * std::log(arg));
  const double sargsw = std::exp((-m.mjsw) * std::log(arg));
  f2 = czb * (1 - m.fc * (1 + m.mj)) * sarg / arg +
       czbsw * (1 - m.fc * (1 + m.mjsw)) * sargsw / arg;
  f3 = czb * m.mj * sarg / arg / t.tBulkPot +
       czbsw * m.mjsw * sargsw / arg / t.tBulkPot;
  f4 = czb * t.tBulkPot * (1 - arg * sarg) / (1 - m.mj) +
       czbsw * t.tBulkPot * (1 - arg * sargsw) / (1 - m.mjsw) -
       f3 / 2 * (t.tDepCap * t.tDepCap) - t.tDepCap * f2;
}

static const char *kMos1InstKeys[] = {"L",  "W",   "AD",  "AS",  "NRD",
                                      "NRS", "PD", "PS",  "M",   "OFF",
                                      "IC1", "IC2", "IC3", "TEMP"};

int mos1_instance_setup(const Mos1Model &m, Mos1Inst &in, const Options &opt,
                        std::string &err) {
  for (auto &kv : in.p.v) {
    bool ok = false;
    for (const char *k : kMos1InstKeys) ok = ok || kv.first == k;
    if (!ok) {
      err = "unknown MOSFET1 instance parameter " + kv.first;
      return XG_E_ARG;
    }
  }
  const ParamSet &p = in.p;
  const double l = p.given("L") ? p.get("L", 0) : m.model_l;
  const double w = p.given("W") ? p.get("W", 0) : m.model_w;
  const double ad = p.get("AD", 0.0), as = p.get("AS", 0.0);
  const double nrd = p.get("NRD", 1.0), nrs = p.get("NRS", 1.0);
  const double pd = p.get("PD", 0.0), ps = p.get("PS", 0.0);
  const double np = p.get("M", 1.0);
  in.temp = p.given("TEMP") ? p.get("TEMP", 0) + 273.15 : opt.temp;
  // series resistances (N_DEV_MOSFET1.C:1184-1225)
  double gd = 0, gs = 0;
  if (m.rd != 0) gd = 1 / m.rd;
  else if (m.rshGiven && m.rsh != 0) gd = 1 / (m.rsh * nrd);
  if (m.rs != 0) gs = 1 / m.rs;
  else if (m.rshGiven && m.rsh != 0) gs = 1 / (m.rsh * nrs);
  in.gd = gd;
  in.gs = gs;
  if (l - 2 * m.latDiff <= 0) {
    err = "MOSFET1 effective channel length less than zero";
    return XG_E_PARAM;
  }
  const double leff = l - 2 * m.latDiff;
  const Mos1Temp t = mos1_temperature(m, in.temp);
  auto &f = in.f;
  f.fill(0.0);
  f[M1F_BETA] = t.tKP * w / leff;
  if ((t.tSatCurDens == 0) || (ad == 0) || (as == 0)) {
    f[M1F_DSATCUR] = t.tSatCur;
    f[M1F_SSATCUR] = t.tSatCur;
  } else {
    f[M1F_DSATCUR] = t.tSatCurDens * ad;
    f[M1F_SSATCUR] = t.tSatCurDens * as;
  }
  if ((m.jctSatCurDensity == 0) || (ad == 0) || (as == 0)) {
    f[M1F_SVCRIT] = f[M1F_DVCRIT] =
        t.vt * std::log(t.vt / (std::sqrt(2.0) * m.jctSatCur));
  } else {
    f[M1F_DVCRIT] =
        t.vt * std::log(t.vt / (std::sqrt(2.0) * m.jctSatCurDensity * ad));
    f[M1F_SVCRIT] =
        t.vt * std::log(t.vt / (std::sqrt(2.0) * m.jctSatCurDensity * as));
  }
  const double czbd = m.capBDGiven ? t.tCbd : (m.cjGiven ? t.tCj * ad : 0);
  const double czbdsw = m.cjswGiven ? t.tCjsw * pd : 0;
  const double czbs = m.capBSGiven ? t.tCbs : (m.cjGiven ? t.tCj * as : 0);
  const double czbssw = m.cjswGiven ? t.tCjsw * ps : 0;
  f[M1F_CBD] = czbd;
  f[M1F_CBDSW] = czbdsw;
  f[M1F_CBS] = czbs;
  f[M1F_CBSSW] = czbssw;
  mos1_junction(m, t, czbd, czbdsw, f[M1F_F2D], f[M1F_F3D], f[M1F_F4D]);
  mos1_junction(m, t, czbs, czbssw, f[M1F_F2S], f[M1F_F3S], f[M1F_F4S]);
  f[M1F_OXIDECAP] = m.oxideCapFactor * leff * w;
  f[M1F_GSO] = m.cgso * w;
  f[M1F_GDO] = m.cgdo * w;
  f[M1F_GBO] = m.cgbo * leff;
  f[M1F_GD] = gd;
  f[M1F_GS] = gs;
  f[M1F_M] = np;
  f[M1F_ICVDS] = p.get("IC1", 0.0);
  f[M1F_ICVGS] = p.get("IC2", 0.0);
  f[M1F_ICVBS] = p.get("IC3", 0.0);
  in.flags = 0;
  if (p.get("OFF", 0.0) != 0.0) in.flags |= XF_OFF;
  if (p.given("IC1") || p.given("IC2") || p.given("IC3")) in.flags |= XF_IC;
  if (gd != 0.0) in.flags |= XF_HAS_DP;
  if (gs != 0.0) in.flags |= XF_HAS_SP;
  return XG_OK;
}

static void mos1_card(const Mos1Model &m, double temp, const Options &opt,
                      double *card) {
  const Mos1Temp t = mos1_temperature(m, temp);
  std::fill(card, card + M1C_N, 0.0);
  card[M1C_DTYPE] = m.dtype;
  card[M1C_GAMMA] = m.gamma;
  card[M1C_LAMBDA] = m.lambda;
  card[M1C_MJ] = m.mj;
  card[M1C_MJSW] = m.mjsw;
  card[M1C_TPHI] = t.tPhi;
  card[M1C_TVBI] = t.tVbi;
  card[M1C_TVTO] = t.tVto;
  card[M1C_VT] = t.vt;
  card[M1C_TDEPCAP] = t.tDepCap;
  card[M1C_TBULKPOT] = t.tBulkPot;
  card[M1C_GMIN] = opt.gmin;
  card[M1C_LIMITER] = opt.limiter ? 1.0 : 0.0;
  card[M1C_MJ_EQ_HALF] = (m.mj == .5) ? 1.0 : 0.0;
  card[M1C_MJSW_EQ_HALF] = (m.mjsw == .5) ? 1.0 : 0.0;
  card[M1C_MJ_EQ_MJSW] = (m.mj == m.mjsw) ? 1.0 : 0.0;
}

// ---------------------------------------------------------------------------
// Stamps.  A stamp entry is (row node, col node) in device-local numbering.
// MOSFET1 local nodes: 0=d 1=g 2=s 3=b 4=d' 5=s'; an absent d'/s' collapses
// onto d/s, which is what jacStampMap's merged stamps express
// (N_DEV_MOSFET1.C:1127-1166, N_DEV_DeviceInstance.C:663-883).
//
// The 22 offsets registerJacLIDs assigns (N_DEV_MOSFET1.C:1650-1673):
static const int kMos1OffsetRC[22][2] = {
    {0, 0}, {0, 4}, {1, 1}, {1, 3}, {1, 4}, {1, 5}, {2, 2}, {2, 5},
    {3, 1}, {3, 3}, {3, 4}, {3, 5}, {4, 0}, {4, 1}, {4, 3}, {4, 4},
    {4, 5}, {5, 1}, {5, 2}, {5, 3}, {5, 4}, {5, 5}};
// dF/dx statements (N_DEV_MOSFET1.C:4488-4620, old-Meyer): offset index.
static const int kMos1JF[17] = {0, 1, 6, 7, 9, 10, 11, 12, 13,
                                14, 15, 16, 17, 18, 19, 20, 21};
// dQ/dx statements (N_DEV_MOSFET1.C:4690-4716): offset index.
static const int kMos1JQ[14] = {2, 3, 4, 5, 8, 9, 10, 11, 13, 14, 15, 17, 19, 21};

static inline int32_t mos1_node(const Mos1Inst &in, int local) {
  if (local == 4 && !(in.flags & XF_HAS_DP)) local = 0;
  if (local == 5 && !(in.flags & XF_HAS_SP)) local = 2;
  return in.lids[local];
}

// ---------------------------------------------------------------------------
// Reduction-program assembly.

// Groups the contributions by output entry, keeping the reference's order
// within each entry.  Entries flagged in `skip` (the ground row and column:
// the reference writes them into the ground slot that N_LAS discards before
// the solve, N_LAS_Builder.C:424-434) get no summing work item; they and the
// entries without contributions are listed for the overwrite mode.
static void finish_program(const std::vector<Contrib> &cs, int32_t n_out,
                           const std::vector<char> &skip,
                           ReduceProgram &prog) {
  std::vector<int32_t> cnt(n_out + 1, 0);
  for (auto &c : cs) cnt[c.out + 1]++;
  for (int32_t i = 0; i < n_out; ++i) cnt[i + 1] += cnt[i];
  std::vector<int32_t> pos(cnt.begin(), cnt.end() - 1);
  std::vector<int32_t> srcs(cs.size());
  for (auto &c : cs) srcs[pos[c.out]++] = c.src;
  prog.ssrc4.clear();
  prog.sdst4.clear();
  prog.ssrc.clear();
  prog.sdst.clear();
  prog.src.clear();
  prog.cptr.assign(1, 0);
  prog.cdst.clear();
  prog.mtgt.clear();
  prog.mchunk.clear();
  prog.zidx.clear();
  for (int32_t k = 0; k < n_out; ++k) {
    const int32_t len = cnt[k + 1] - cnt[k];
    if (len == 0 || skip[k]) {
      prog.zidx.push_back(k);
      continue;
    }
    if (len <= 4) {
      for (int32_t q = 0; q < 4; ++q)
        prog.ssrc4.push_back(q < len ? srcs[cnt[k] + q] : kPad);
      prog.sdst4.push_back(k);
      continue;
    }
    if (len <= kShort) {
      for (int32_t q = 0; q < kShort; ++q)
        prog.ssrc.push_back(q < len ? srcs[cnt[k] + q] : kPad);
      prog.sdst.push_back(k);
      continue;
    }
    const int32_t first_chunk = (int32_t)prog.cdst.size();
    for (int32_t j = cnt[k]; j < cnt[k + 1]; j += kChunk) {
      const int32_t e = std::min(cnt[k + 1], j + kChunk);
      for (int32_t q = j; q < e; ++q) prog.src.push_back(srcs[q]);
      prog.cptr.push_back((int32_t)prog.src.size());
      prog.cdst.push_back(len <= kChunk ? k : -1);
    }
    if (len > kChunk) {
      prog.mtgt.push_back(k);
      prog.mchunk.push_back(first_chunk);
      prog.mchunk.push_back((int32_t)prog.cdst.size());
    }
  }
}

int build_plan(Circuit &c, Plan &plan, std::string &err) {
  const auto t_start = std::chrono::steady_clock::now();
  const int32_t n = c.n_nodes;
  const int32_t G = n;  // ground slot / row / column
  auto L = [G](int32_t g) { return g < 0 ? G : g; };
  plan.n_nodes = n;
  plan.n_state = c.n_state;
  plan.n_store = c.n_store;

  // ---- instance setup (MOSFET1) and cards
  std::vector<std::pair<int, double>> card_keys;
  for (auto &in : c.mos1) {
    int rc = mos1_instance_setup(c.mos1_models[in.model], in, c.opt, err);
    if (rc) return rc;
    int ci = -1;
    for (size_t k = 0; k < card_keys.size(); ++k)
      if (card_keys[k].first == in.model && card_keys[k].second == in.temp)
        ci = (int)k;
    if (ci < 0) {
      ci = (int)card_keys.size();
      card_keys.push_back({in.model, in.temp});
    }
    in.card = ci;
    for (int k = 0; k < 6; ++k)
      if (in.lids[k] >= n) {
        err = "MOSFET1 node id out of range";
        return XG_E_ARG;
      }
    if (in.sta_base < 0 || in.sta_base + 8 > c.n_state || in.sto_base < 0 ||
        in.sto_base + 5 > c.n_store) {
      err = "MOSFET1 state/store range out of bounds";
      return XG_E_ARG;
    }
  }

  // ---- CSR pattern: union of every stamp, rows sorted, ground last
  //      (N_TOP_TopoLSUtil.C:2375-2430, N_LAS_Builder.C:336-338, 403-445).
  std::vector<std::array<int32_t, 2>> entries;
  for (auto &in : c.mos1)
    for (int k = 0; k < 22; ++k)
      entries.push_back({L(mos1_node(in, kMos1OffsetRC[k][0])),
                         L(mos1_node(in, kMos1OffsetRC[k][1]))});
  {
    int rc = bsim3_prepare(c, err);
    if (rc) return rc;
    bsim3_pattern_entries(c, entries);
  }
  for (auto &fs : c.foreign)
    for (size_t e = 0; e < fs.rows.size(); ++e)
      entries.push_back({L(fs.gids[fs.rows[e]]), L(fs.gids[fs.cols[e]])});
  for (auto &ld : c.lin) {
    const int32_t p = L(ld.lids[0]), q = L(ld.lids[1]), b = L(ld.lids[2]);
    if (ld.type == LIN_V) {
      // N_DEV_Vsrc.C jacStamp: (pos, br), (neg, br), (br, pos), (br, neg)
      entries.push_back({p, b});
      entries.push_back({q, b});
      entries.push_back({b, p});
      entries.push_back({b, q});
    } else {
      // N_DEV_Resistor.C / N_DEV_Capacitor.C, and N_DEV_Diode.C with RS = 0
      // (internal node merged into the positive one): the full 2 x 2 block
      entries.push_back({p, p});
      entries.push_back({p, q});
      entries.push_back({q, p});
      entries.push_back({q, q});
    }
  }

  if (c.have_pattern) {
    plan.rowptr = c.rowptr;
    plan.colidx = c.colidx;
  } else {
    std::vector<int32_t> cnt(n + 2, 0);
    std::sort(entries.begin(), entries.end());
    entries.erase(std::unique(entries.begin(), entries.end()), entries.end());
    plan.rowptr.assign(n + 2, 0);
    plan.colidx.resize(entries.size());
    for (size_t e = 0; e < entries.size(); ++e) {
      plan.rowptr[entries[e][0] + 1]++;
      plan.colidx[e] = entries[e][1];
    }
    for (int32_t r = 0; r <= n; ++r) plan.rowptr[r + 1] += plan.rowptr[r];
  }
  auto csr_pos = [&](int32_t r, int32_t col, int32_t &pos) -> bool {
    const int32_t *b = plan.colidx.data() + plan.rowptr[r];
    const int32_t *e = plan.colidx.data() + plan.rowptr[r + 1];
    const int32_t *it = std::lower_bound(b, e, col);
    if (it == e || *it != col) return false;
    pos = (int32_t)(it - plan.colidx.data());
    return true;
  };

  // ---- MOSFET1 evaluation order: grouped by card, each group padded to a
  //      whole wavefront so that every wave reads one card.
  const int32_t n1 = (int32_t)c.mos1.size();
  plan.m1_n = n1;
  std::vector<int32_t> slot_of(n1, -1);
  {
    std::vector<int32_t> order(n1);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) {
      return c.mos1[a].card < c.mos1[b].card;
    });
    int32_t s = 0;
    plan.m1_inst_of.clear();
    plan.m1_wave_card.clear();
    for (size_t k = 0; k < order.size();) {
      const int card = c.mos1[order[k]].card;
      size_t e = k;
      while (e < order.size() && c.mos1[order[e]].card == card) ++e;
      for (size_t j = k; j < e; ++j) {
        slot_of[order[j]] = s++;
        plan.m1_inst_of.push_back(order[j]);
      }
      while (s % 64) {
        plan.m1_inst_of.push_back(-1);
        ++s;
      }
      k = e;
    }
    plan.m1_slots = s;
    for (int32_t w = 0; w < s / 64; ++w)
      plan.m1_wave_card.push_back(
          c.mos1[plan.m1_inst_of[w * 64]].card);
  }
  const int32_t S1 = plan.m1_slots;
  plan.m1_cards.assign(card_keys.size() * M1C_N, 0.0);
  for (size_t k = 0; k < card_keys.size(); ++k)
    mos1_card(c.mos1_models[card_keys[k].first], card_keys[k].second, c.opt,
              &plan.m1_cards[k * M1C_N]);
  plan.m1_fields.assign((size_t)M1F_N * S1, 0.0);
  plan.m1_lid4.assign((size_t)4 * S1, G);
  plan.m1_lid_int.assign((size_t)2 * S1, G);
  plan.m1_sta.assign(S1, 0);
  plan.m1_sto.assign(S1, 0);
  plan.m1_flags.assign(S1, 0);
  plan.m1_fields_ref.assign((size_t)n1 * M1F_N, 0.0);
  for (int32_t i = 0; i < n1; ++i) {
    const Mos1Inst &in = c.mos1[i];
    const int32_t s = slot_of[i];
    for (int f = 0; f < M1F_N; ++f) {
      plan.m1_fields[(size_t)f * S1 + s] = in.f[f];
      plan.m1_fields_ref[(size_t)i * M1F_N + f] = in.f[f];
    }
    for (int k = 0; k < 4; ++k) plan.m1_lid4[(size_t)k * S1 + s] = L(in.lids[k]);
    plan.m1_lid_int[s] = L(mos1_node(in, 4));
    plan.m1_lid_int[(size_t)S1 + s] = L(mos1_node(in, 5));
    plan.m1_sta[s] = in.sta_base;
    plan.m1_sto[s] = in.sto_base;
    plan.m1_flags[s] = in.flags;
  }

  // ---- offsets + reduction contributions, in reference order
  std::vector<Contrib> vec[4], mat[2];
  plan.m1_slot_base = 0;
  plan.m1_offsets.assign((size_t)n1 * 22, 0);
  auto SL = [&](int slot, int32_t s) -> int32_t {
    return (int32_t)(plan.m1_slot_base + (int64_t)slot * S1 + s);
  };
  // loadDAEVectors loops instances once, touching F, Q, dFdxdVp, dQdxdVp.
  for (int32_t i = 0; i < n1; ++i) {
    const Mos1Inst &in = c.mos1[i];
    const int32_t s = slot_of[i];
    const bool dp = in.flags & XF_HAS_DP, sp = in.flags & XF_HAS_SP;
    const int32_t d = L(in.lids[0]), g = L(in.lids[1]), so = L(in.lids[2]),
                  b = L(in.lids[3]);
    const int32_t dpr = L(mos1_node(in, 4)), spr = L(mos1_node(in, 5));
    if (dp) vec[0].push_back({d, SL(M1S_F_D, s)});
    vec[0].push_back({g, SL(M1S_F_G, s)});
    if (sp) vec[0].push_back({so, SL(M1S_F_S, s)});
    vec[0].push_back({b, SL(M1S_F_B, s)});
    vec[0].push_back({dpr, SL(M1S_F_DP, s)});
    vec[0].push_back({spr, SL(M1S_F_SP, s)});
    vec[1].push_back({g, SL(M1S_Q_G, s)});
    vec[1].push_back({b, SL(M1S_Q_B, s)});
    vec[1].push_back({dpr, SL(M1S_Q_DP, s)});
    vec[1].push_back({spr, SL(M1S_Q_SP, s)});
    vec[2].push_back({b, SL(M1S_FL_B, s)});
    vec[2].push_back({dpr, SL(M1S_FL_DP, s)});
    vec[2].push_back({spr, SL(M1S_FL_SP, s)});
    vec[3].push_back({g, SL(M1S_QL_G, s)});
    vec[3].push_back({b, SL(M1S_QL_B, s)});
    vec[3].push_back({dpr, SL(M1S_QL_DP, s)});
    vec[3].push_back({spr, SL(M1S_QL_SP, s)});
  }
  for (int32_t i = 0; i < n1; ++i) {
    const Mos1Inst &in = c.mos1[i];
    const int32_t s = slot_of[i];
    int32_t pos[22];
    for (int k = 0; k < 22; ++k) {
      const int32_t r = L(mos1_node(in, kMos1OffsetRC[k][0]));
      const int32_t col = L(mos1_node(in, kMos1OffsetRC[k][1]));
      if (!csr_pos(r, col, pos[k])) {
        err = "supplied CSR pattern lacks a MOSFET1 stamp entry";
        return XG_E_PATTERN;
      }
      plan.m1_offsets[(size_t)i * 22 + k] = pos[k] - plan.rowptr[r];
    }
    // dF/dx statements 0-3, 7, 13 are conductance terms that are constant
    // zero when the node is collapsed: "+= 0" keeps a "+0.0" step, "-= 0" is
    // the identity and is dropped.
    const bool dp = in.flags & XF_HAS_DP, sp = in.flags & XF_HAS_SP;
    for (int e = 0; e < 17; ++e) {
      const bool dterm = (e == 0 || e == 1 || e == 7);
      const bool sterm = (e == 2 || e == 3 || e == 13);
      int32_t src = SL(M1S_JF0 + e, s);
      if ((dterm && !dp) || (sterm && !sp)) {
        const bool plus = (e == 0 || e == 2);
        if (!plus) continue;
        src = -1;
      }
      mat[0].push_back({pos[kMos1JF[e]], src});
    }
    for (int e = 0; e < 14; ++e)
      mat[1].push_back({pos[kMos1JQ[e]], SL(M1S_JQ0 + e, s)});
  }
  int64_t slot_total = (int64_t)M1S_N * S1;

  // ---- BSIM3 (after MOSFET1, the order the oracle's device manager uses)
  {
    int rc = bsim3_build(c, plan, slot_total, vec, mat, err);
    if (rc) return rc;
  }
  slot_total += bsim3_slot_count(plan);

  // ---- linear devices, after the MOSFETs; resistors, capacitors, diodes,
  //      then sources (one Master loop per device type), each in the order
  //      they were added, each type padded to whole waves
  {
    std::vector<int32_t> order;
    for (int t = LIN_R; t <= LIN_V; ++t) {
      for (int32_t i = 0; i < (int32_t)c.lin.size(); ++i)
        if (c.lin[i].type == t) order.push_back(i);
      while (order.size() % 64) order.push_back(-1);
    }
    // drop the padding of empty groups
    std::vector<int32_t> ord2;
    for (size_t w = 0; w < order.size(); w += 64) {
      bool any = false;
      for (size_t j = w; j < w + 64; ++j) any = any || order[j] >= 0;
      if (any) ord2.insert(ord2.end(), order.begin() + w, order.begin() + w + 64);
    }
    const int32_t SL = (int32_t)ord2.size();
    plan.lin_n = (int32_t)c.lin.size();
    plan.lin_slots = SL;
    plan.lin_slot_base = slot_total;
    plan.lin_type.assign(SL, -1);
    plan.lin_lid.assign((size_t)3 * SL, G);
    plan.lin_sta.assign(SL, 0);
    plan.lin_val.assign(SL, 0.0);
    plan.lin_pulse.assign((size_t)8 * SL, 0.0);
    plan.lin_dpar.assign((size_t)DIO_N * SL, 0.0);
    auto LS = [&](int e, int32_t s) -> int32_t {
      return (int32_t)(slot_total + (int64_t)e * SL + s);
    };
    // vectors first for every device (loadDAEVectors), matrices after
    // (loadDAEMatrices), each in device-type then instance order
    for (int32_t s = 0; s < SL; ++s) {
      const int32_t i = ord2[s];
      if (i < 0) continue;
      const LinDev &ld = c.lin[i];
      const int32_t p = L(ld.lids[0]), q = L(ld.lids[1]), b = L(ld.lids[2]);
      plan.lin_type[s] = ld.type;
      plan.lin_lid[s] = p;
      plan.lin_lid[(size_t)SL + s] = q;
      plan.lin_lid[(size_t)2 * SL + s] = b;
      plan.lin_sta[s] = ld.sta;
      plan.lin_val[s] = ld.type == LIN_R ? 1.0 / ld.val : ld.val;
      for (int k = 0; k < 8; ++k) plan.lin_pulse[(size_t)k * SL + s] = ld.pulse[k];
      if (ld.type == LIN_D)
        for (int k = 0; k < DIO_N; ++k) plan.lin_dpar[(size_t)k * SL + s] = ld.dpar[k];
      if (ld.type == LIN_R) {          // N_DEV_Resistor.C:1118-1131
        vec[0].push_back({p, LS(0, s)});
        vec[0].push_back({q, LS(1, s)});
      } else if (ld.type == LIN_C) {   // N_DEV_Capacitor.C:1881-1960
        vec[1].push_back({p, LS(0, s)});
        vec[1].push_back({q, LS(1, s)});
      } else if (ld.type == LIN_D) {   // N_DEV_Diode.C:2040-2067 (Pri = Pos)
        vec[0].push_back({p, LS(0, s)});
        vec[0].push_back({q, LS(1, s)});
        vec[0].push_back({p, LS(2, s)});
        vec[1].push_back({q, LS(3, s)});
        vec[1].push_back({p, LS(4, s)});
        if (c.opt.limiter) {
          vec[2].push_back({q, LS(5, s)});
          vec[2].push_back({p, LS(6, s)});
          vec[3].push_back({q, LS(7, s)});
          vec[3].push_back({p, LS(8, s)});
        }
      } else {                         // N_DEV_Vsrc.C:1002-1013
        vec[0].push_back({p, LS(0, s)});
        vec[0].push_back({q, LS(1, s)});
        vec[0].push_back({b, LS(2, s)});
        vec[0].push_back({b, LS(3, s)});
      }
    }
    for (int32_t s = 0; s < SL; ++s) {
      const int32_t i = ord2[s];
      if (i < 0) continue;
      const LinDev &ld = c.lin[i];
      const int32_t p = L(ld.lids[0]), q = L(ld.lids[1]), b = L(ld.lids[2]);
      if (ld.type == LIN_D) {
        // N_DEV_Diode.C:2094-2108 with Pri = Pos: dF/dx (P,P) +Gspr,
        // (P,P) -Gspr, (N,N) +Gd, (N,P) -Gd, (P,P) -Gspr, (P,N) -Gd,
        // (P,P) +Gspr+Gd; dQ/dx (N,N) +Cd, (N,P) -Cd, (P,N) -Cd, (P,P) +Cd
        const int32_t tf[7][2] = {{p, p}, {p, p}, {q, q}, {q, p}, {p, p}, {p, q}, {p, p}};
        const int32_t tq[4][2] = {{q, q}, {q, p}, {p, q}, {p, p}};
        for (int k = 0; k < 11; ++k) {
          const int32_t *e = k < 7 ? tf[k] : tq[k - 7];
          int32_t pos;
          if (!csr_pos(e[0], e[1], pos)) {
            err = "supplied CSR pattern lacks a diode stamp entry";
            return XG_E_PATTERN;
          }
          (k < 7 ? mat[0] : mat[1]).push_back({pos, LS(9 + k, s)});
        }
        continue;
      }
      int32_t rc_[4][2];
      if (ld.type == LIN_V) {          // N_DEV_Vsrc.C:1023-1036
        const int32_t t[4][2] = {{p, b}, {q, b}, {b, p}, {b, q}};
        std::memcpy(rc_, t, sizeof(t));
      } else {                         // Resistor.C:1159-1174, Capacitor.C:2044-2050
        const int32_t t[4][2] = {{p, p}, {p, q}, {q, p}, {q, q}};
        std::memcpy(rc_, t, sizeof(t));
      }
      const int e0 = ld.type == LIN_V ? 4 : 2;
      auto &prog = ld.type == LIN_C ? mat[1] : mat[0];
      for (int k = 0; k < 4; ++k) {
        int32_t pos;
        if (!csr_pos(rc_[k][0], rc_[k][1], pos)) {
          err = "supplied CSR pattern lacks a linear-device stamp entry";
          return XG_E_PATTERN;
        }
        prog.push_back({pos, LS(e0 + k, s)});
      }
    }
    slot_total += (int64_t)kLinStatements * SL;
  }
  plan.n_slot_values = slot_total;
  if (slot_total > INT32_MAX) {
    err = "slot buffer exceeds 2^31 values";
    return XG_E_ARG;
  }
  std::vector<char> vskip(n + 1, 0);
  vskip[G] = 1;
  for (int k = 0; k < 4; ++k) finish_program(vec[k], n + 1, vskip, plan.red_vec[k]);
  const int32_t nnz = (int32_t)plan.colidx.size();
  std::vector<char> mskip(nnz, 0);
  for (int32_t r = 0; r <= n; ++r)
    for (int32_t j = plan.rowptr[r]; j < plan.rowptr[r + 1]; ++j)
      mskip[j] = (r == G || plan.colidx[j] == G) ? 1 : 0;
  for (int k = 0; k < 2; ++k) finish_program(mat[k], nnz, mskip, plan.red_mat[k]);
  plan.limiter = c.opt.limiter;
  plan.gmin = c.opt.gmin;
  // ---- connected components without ground (union-find over the pattern)
  {
    std::vector<int32_t> par(n);
    std::iota(par.begin(), par.end(), 0);
    std::function<int32_t(int32_t)> find = [&](int32_t a) {
      while (par[a] != a) a = par[a] = par[par[a]];
      return a;
    };
    for (int32_t r = 0; r < n; ++r)
      for (int32_t j = plan.rowptr[r]; j < plan.rowptr[r + 1]; ++j) {
        const int32_t col = plan.colidx[j];
        if (col >= n) continue;
        const int32_t a = find(r), b = find(col);
        if (a != b) par[std::max(a, b)] = std::min(a, b);
      }
    std::vector<int32_t> id(n, -1);
    int32_t nc = 0;
    for (int32_t v = 0; v < n; ++v) {
      const int32_t root = find(v);
      if (id[root] < 0) id[root] = nc++;
    }
    std::vector<int32_t> cnt(nc + 1, 0);
    for (int32_t v = 0; v < n; ++v) cnt[id[find(v)] + 1]++;
    for (int32_t k = 0; k < nc; ++k) cnt[k + 1] += cnt[k];
    plan.comp_start = cnt;
    plan.comp_nodes.assign(n, 0);
    plan.comp_local.assign(n + 1, -1);
    std::vector<int32_t> at(cnt.begin(), cnt.end() - 1);
    for (int32_t v = 0; v < n; ++v) {
      const int32_t k = id[find(v)];
      plan.comp_local[v] = at[k] - cnt[k];
      plan.comp_nodes[at[k]++] = v;
    }
    plan.comp_aoff.assign(nc, 0);
    int64_t w = 0;
    for (int32_t k = 0; k < nc; ++k) {
      const int64_t m = cnt[k + 1] - cnt[k];
      plan.comp_aoff[k] = m <= kDenseMax ? w : -1;
      if (m > dense_lds_max() && m <= kDenseMax) w += m * m;
    }
    plan.comp_work = w;
    // eval slot -> component (an instance's nodes all lie in one component)
    std::vector<int32_t> node_of;
    for (int32_t s = 0; s < plan.m1_slots; ++s) {
      int32_t node = -1;
      if (plan.m1_inst_of[s] >= 0)
        for (int k = 0; k < 4 && node < 0; ++k)
          if (plan.m1_lid4[(size_t)k * plan.m1_slots + s] != G)
            node = plan.m1_lid4[(size_t)k * plan.m1_slots + s];
      node_of.push_back(node);
    }
    bsim3_slot_nodes(plan, G, node_of);
    // linear-device slots (only diodes flag limiting)
    for (int32_t s = 0; s < plan.lin_slots; ++s) {
      const size_t S = (size_t)plan.lin_slots;
      int32_t node = plan.lin_lid[s] != G ? plan.lin_lid[s] : plan.lin_lid[S + s];
      if (plan.lin_type[s] < 0 || node == G) node = -1;
      node_of.push_back(node);
    }
    plan.slot_comp.assign(node_of.size(), -1);
    for (size_t s = 0; s < node_of.size(); ++s)
      if (node_of[s] >= 0) plan.slot_comp[s] = id[find(node_of[s])];
  }
  int32_t pbase = 0;
  for (auto *prog : {&plan.red_vec[0], &plan.red_vec[1], &plan.red_vec[2],
                     &plan.red_vec[3], &plan.red_mat[0], &plan.red_mat[1]}) {
    prog->pbase = pbase;
    pbase += (int32_t)prog->cdst.size();
  }
  plan.n_partials = pbase;
  plan.lu_mode = c.opt.lu_mode;
  const auto t_lu = std::chrono::steady_clock::now();
  const int rc = lu_analyse(plan, err);
  if (std::getenv("XG_LU_DEBUG")) {
    const auto t_end = std::chrono::steady_clock::now();
    std::fprintf(stderr, "setup: plan %.3f s, LU analysis %.3f s\n",
                 std::chrono::duration<double>(t_lu - t_start).count(),
                 std::chrono::duration<double>(t_end - t_lu).count());
  }
  return rc;
}

}  // namespace xg
