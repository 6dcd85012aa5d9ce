// BSIM3 v3.2 fused load kernel for gfx950 (capMod 3, mobMod 1/2/3).
//
// One thread per instance.  Instances are grouped by card and padded to
// whole 64-lane wavefronts (one card per wave, staged in LDS); within a card
// they are ordered by size-dependent record so that instances sharing a
// record read the same cache lines.  The thread runs updateIntermediateVars
// and Master::updateState, then computes every value that the Master's
// loadDAEVectors and loadDAEMatrices statements add and writes them to the
// slot buffer (one array per statement); the ordered reduction kernel folds
// them into the outputs in the reference's order.
//
// Physics follows, statement for statement (FP64, no FMA contraction):
//   Instance::updateIntermediateVars   N_DEV_MOSFET_B3.C:4526-7273
//   setupCapacitors_oldDAE/newDAE      N_DEV_MOSFET_B3.C:7558-7885
//   auxChargeCalculations              N_DEV_MOSFET_B3.C:7465-7555
//   Master::updateState                N_DEV_MOSFET_B3.C:9222-9296
//   Master::loadDAEVectors             N_DEV_MOSFET_B3.C:9306-9643
//   Master::loadDAEMatrices            N_DEV_MOSFET_B3.C:9654-9928
// nqsMod = 0, so the non-quasi-static terms of the dF/dx statements
// (dxpart*ggt*, T1global*ddxpart_dV*) are identically zero and omitted;
// they could only change the sign of an entry that is exactly zero.
#include "xg_bsim3.h"
#include "xg_bsim3_layout.h"
#include "xg_devmath.h"
#include "xg_linear_dev.h"

namespace xg {

namespace {

constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / 64;
constexpr double kExpThreshold = 34.0;
constexpr double kMinExp = 1.713908431e-15;
constexpr double kMaxExp = 5.834617425e14;
constexpr double kEpsSi = 1.03594e-10;
constexpr double kQ = 1.6021918e-19;
constexpr double kDelta1 = 0.02, kDelta3 = 0.02, kDelta4 = 0.02;

// Depletion charge/capacitance of one bulk junction
// (N_DEV_MOSFET_B3.C:7009-7134).
__device__ __forceinline__ void b3_junction(double v, double czb, double czbsw,
                                            double czbswg, const double *cd,
                                            double &q, double &cap) {
  const double MJ = cd[B3C_MJ], MJSW = cd[B3C_MJSW], MJSWG = cd[B3C_MJSWG];
  const double PhiB = cd[B3C_PHIB], PhiBSW = cd[B3C_PHIBSW],
               PhiBSWG = cd[B3C_PHIBSWG];
  if (v == 0.0) {
    q = 0.0;
    cap = czb + czbsw + czbswg;
  } else if (v < 0.0) {
    double arg, sarg;
    if (czb > 0.0) {
      arg = 1.0 - v / PhiB;
      sarg = (MJ == 0.5) ? 1.0 / sqrt(arg) : dm_exp(-MJ * dm_log(arg));
      q = PhiB * czb * (1.0 - arg * sarg) / (1.0 - MJ);
      cap = czb * sarg;
    } else {
      q = 0.0;
      cap = 0.0;
    }
    if (czbsw > 0.0) {
      arg = 1.0 - v / PhiBSW;
      sarg = (MJSW == 0.5) ? 1.0 / sqrt(arg) : dm_exp(-MJSW * dm_log(arg));
      q += PhiBSW * czbsw * (1.0 - arg * sarg) / (1.0 - MJSW);
      cap += czbsw * sarg;
    }
    if (czbswg > 0.0) {
      arg = 1.0 - v / PhiBSWG;
      sarg = (MJSWG == 0.5) ? 1.0 / sqrt(arg) : dm_exp(-MJSWG * dm_log(arg));
      q += PhiBSWG * czbswg * (1.0 - arg * sarg) / (1.0 - MJSWG);
      cap += czbswg * sarg;
    }
  } else {
    const double T0 = czb + czbsw + czbswg;
    const double T1 = v * (czb * MJ / PhiB + czbsw * MJSW / PhiBSW +
                           czbswg * MJSWG / PhiBSWG);
    q = v * (T0 + 0.5 * T1);
    cap = T0 + T1;
  }
}

// Smooth overlap charge/capacitance of one side, capMod 3
// (N_DEV_MOSFET_B3.C:7217-7238).
__device__ __forceinline__ void b3_overlap(double v, double cgxo, double cgxl,
                                           double weffCV, double ckappa,
                                           double &cap, double &q) {
  const double T0 = v + kDelta1;
  const double T1 = sqrt(T0 * T0 + 4.0 * kDelta1);
  const double T2 = 0.5 * (T0 - T1);
  const double T3 = weffCV * cgxl;
  const double T4 = sqrt(1.0 - 4.0 * T2 / ckappa);
  cap = cgxo + T3 - T3 * (1.0 - 1.0 / T4) * (0.5 - 0.5 * T0 / T1);
  q = (cgxo + T3) * v - T3 * (T2 + 0.5 * ckappa * (T4 - 1.0));
}

// Overlap charge/capacitance of one side for capMod 0 (constant cgxo,
// N_DEV_MOSFET_B3.C:7164-7187) and capMod 1 (7188-7215: bias-dependent
// overlap cgxl over weffCV, square-root form below zero, linear above).
__device__ __forceinline__ void b3_overlap01(int capmod, double v, double cgxo,
                                             double cgxl, double weffCV,
                                             double ckappa, double &cap,
                                             double &q) {
  if (capmod == 0) {
    cap = cgxo;
    q = cgxo * v;
  } else if (v < 0.0) {
    const double r = sqrt(1.0 - 4.0 * v / ckappa);
    cap = cgxo + weffCV * cgxl / r;
    q = cgxo * v - weffCV * 0.5 * cgxl * ckappa * (r - 1.0);
  } else {
    cap = cgxo + weffCV * cgxl;
    q = (weffCV * cgxl + cgxo) * v;
  }
}

// Channel charges and their capacitances, as the charge section hands them
// to the companion step (qsrc is rebuilt there from the others).
struct B3ChanQ {
  double qgate, qdrn, qsrc, qbulk;
  double cggb, cgsb, cgdb, cdgb, cdsb, cddb, cbgb, cbsb, cbdb;
};

// capMod 0 (N_DEV_MOSFET_B3.C:6005-6308): long-channel charges on the
// flat-band voltage vfbcv in accumulation (Arg1 <= 0), depletion
// (Vgst <= 0) and inversion, where saturation and linear forms split the
// channel charge 0/100 (xpart > 0.5), 40/60 (xpart < 0.5) or 50/50.  The
// body bias is Vbs when Vbseff < 0, else phi - Phis.
__device__ __forceinline__ void b3_charge_cap0(
    double xpart, double vfbcv, double phi, double k1ox, double CoxWL,
    double abulkCVfactor, double Vbs, double Vbseff, double Phis,
    double dPhis_dVb, double sqrtPhis, double dsqrtPhis_dVb, double Vgs_eff,
    double dVgs_eff_dVg, double Vds, double Abulk0, double dAbulk0_dVb,
    B3ChanQ &o) {
  double vb, dvb_dVb;
  if (Vbseff < 0.0) {
    vb = Vbs;
    dvb_dVb = 1.0;
  } else {
    vb = phi - Phis;
    dvb_dVb = -dPhis_dVb;
  }
  const double Vth = vfbcv + phi + k1ox * sqrtPhis;
  const double Vgst = Vgs_eff - Vth;
  const double dVth_dVb = k1ox * dsqrtPhis_dVb;
  const double Arg1 = Vgs_eff - vb - vfbcv;
  o.qsrc = 0.0;
  o.cgdb = o.cdgb = o.cddb = o.cdsb = o.cbdb = 0.0;
  if (Arg1 <= 0.0) {
    o.qgate = CoxWL * Arg1;
    o.qbulk = -o.qgate;
    o.qdrn = 0.0;
    o.cggb = CoxWL * dVgs_eff_dVg;
    o.cgsb = CoxWL * (dvb_dVb - dVgs_eff_dVg);
    o.cbgb = -CoxWL * dVgs_eff_dVg;
    o.cbsb = -o.cgsb;
    return;
  }
  if (Vgst <= 0.0) {
    const double h = 0.5 * k1ox;
    const double r = sqrt(h * h + Arg1);
    o.qgate = CoxWL * k1ox * (r - h);
    o.qbulk = -o.qgate;
    o.qdrn = 0.0;
    const double c = CoxWL * h / r;
    o.cggb = c * dVgs_eff_dVg;
    o.cgsb = c * (dvb_dVb - dVgs_eff_dVg);
    o.cbgb = -o.cggb;
    o.cbsb = -o.cgsb;
    return;
  }
  const double c3 = CoxWL / 3.0, c23 = 2.0 * c3;
  const double AbulkCV = Abulk0 * abulkCVfactor;
  const double dAbulkCV_dVb = abulkCVfactor * dAbulk0_dVb;
  const double Vdsat = Vgst / AbulkCV;
  const double dVdsat_dVg = dVgs_eff_dVg / AbulkCV;
  const double dVdsat_dVb = -(Vdsat * dAbulkCV_dVb + dVth_dVb) / AbulkCV;
  const int part = (xpart > 0.5) ? 0 : (xpart < 0.5) ? 1 : 2;
  if (Vdsat <= Vds) {
    o.qgate = CoxWL * (Vgs_eff - vfbcv - phi - Vdsat / 3.0);
    const double qi = -c23 * Vgst;
    o.qbulk = -(o.qgate + qi);
    o.cggb = c3 * (3.0 - dVdsat_dVg) * dVgs_eff_dVg;
    const double gb = -c3 * dVdsat_dVb;
    o.cgsb = -(o.cggb + gb);
    o.cbgb = -(o.cggb - c23 * dVgs_eff_dVg);
    const double bb = -(gb + c23 * dVth_dVb);
    o.cbsb = -(o.cbgb + bb);
    if (part == 0) {
      o.qdrn = 0.0;
    } else if (part == 1) {
      o.qdrn = 0.4 * qi;
      const double k = 0.4 * c23;
      o.cdgb = -k * dVgs_eff_dVg;
      o.cdsb = -(k * dVth_dVb + o.cdgb);
    } else {
      o.qdrn = 0.5 * qi;
      o.cdgb = -c3 * dVgs_eff_dVg;
      o.cdsb = -(c3 * dVth_dVb + o.cdgb);
    }
    return;
  }
  const double Alphaz = Vgst / Vdsat;
  const double T1 = 2.0 * Vdsat - Vds;
  const double T2 = Vds / (3.0 * T1);
  const double T3 = T2 * Vds;
  const double q4 = 0.25 * CoxWL;
  const double T4 = q4 * Alphaz;
  o.qgate = CoxWL * (Vgs_eff - vfbcv - phi - 0.5 * (Vds - T3));
  const double T5 = T3 / T1;
  o.cggb = CoxWL * (1.0 - T5 * dVdsat_dVg) * dVgs_eff_dVg;
  const double gb = -CoxWL * T5 * dVdsat_dVb;
  o.cgdb = CoxWL * (T2 - 0.5 + 0.5 * T5);
  if (part == 0) o.cgsb = -(o.cggb + gb + o.cgdb);
  else o.cgsb = -(o.cggb + o.cgdb + gb);
  const double rV = 1.0 / Vdsat;
  const double dAz_dVg = rV * (1.0 - Alphaz * dVdsat_dVg);
  const double dAz_dVb = -rV * (dVth_dVb + Alphaz * dVdsat_dVb);
  if (part == 0) {
    const double a = 2.0 * Vds - T1 - 3.0 * T3;
    const double b = T3 - T1 - 2.0 * Vds;
    o.qdrn = T4 * a;
    o.qbulk = -(o.qgate + o.qdrn + T4 * b);
    const double qa = q4 * a, qb = q4 * b;
    double k = 2.0 * T4 * (1.0 - 3.0 * T5);
    o.cdgb = (qa * dAz_dVg - k * dVdsat_dVg) * dVgs_eff_dVg;
    const double db = qa * dAz_dVb - k * dVdsat_dVb;
    o.cddb = T4 * (3.0 - 6.0 * T2 - 3.0 * T5);
    o.cdsb = -(o.cdgb + db + o.cddb);
    k = 2.0 * T4 * (1.0 + T5);
    const double bg = (qb * dAz_dVg - k * dVdsat_dVg) * dVgs_eff_dVg;
    const double bbv = qb * dAz_dVb - k * dVdsat_dVb;
    const double bd = T4 * (2.0 * T2 + T5 - 1.0);
    const double bs = -(bg + bbv + bd);
    o.cbgb = -(o.cggb + o.cdgb + bg);
    o.cbdb = -(o.cgdb + o.cddb + bd);
    o.cbsb = -(o.cgsb + o.cdsb + bs);
  } else if (part == 1) {
    const double w = 8.0 * Vdsat * Vdsat - 6.0 * Vdsat * Vds + 1.2 * Vds * Vds;
    const double u = T2 / T1;
    double a = Vds - T1 - u * w;
    o.qdrn = T4 * a;
    a *= q4;
    const double v = u / T1;
    const double k = T4 * (2.0 - 4.0 * v * w + u * (16.0 * Vdsat - 6.0 * Vds));
    o.cdgb = (a * dAz_dVg - k * dVdsat_dVg) * dVgs_eff_dVg;
    const double db = a * dAz_dVb - k * dVdsat_dVb;
    o.cddb = T4 * (2.0 - (1.0 / (3.0 * T1 * T1) + 2.0 * v) * w +
                   u * (6.0 * Vdsat - 2.4 * Vds));
    o.cdsb = -(o.cdgb + db + o.cddb);
    double e = 2.0 * (T1 + T3);
    o.qbulk = -(o.qgate - T4 * e);
    e *= q4;
    const double f = 4.0 * T4 * (1.0 - T5);
    const double bg = (-e * dAz_dVg - o.cdgb - f * dVdsat_dVg) * dVgs_eff_dVg;
    const double bbv = -e * dAz_dVb - db - f * dVdsat_dVb;
    const double bd = -4.0 * T4 * (T2 - 0.5 + 0.5 * T5) - o.cddb;
    const double bs = -(bd + bbv + bg);
    o.cbgb = -(o.cggb + o.cdgb + bg);
    // the Vb term, as N_DEV_MOSFET_B3.C:6238 has it
    o.cbdb = -(o.cgdb + o.cddb + bbv);
    o.cbsb = -(o.cgsb + o.cdsb + bs);
  } else {
    double a = T1 + T3;
    o.qdrn = -T4 * a;
    o.qbulk = -(o.qgate + o.qdrn + o.qdrn);
    a *= q4;
    const double k = T4 * (2.0 * T5 - 2.0);
    o.cdgb = (k * dVdsat_dVg - a * dAz_dVg) * dVgs_eff_dVg;
    const double db = k * dVdsat_dVb - a * dAz_dVb;
    o.cddb = T4 * (1.0 - 2.0 * T2 - T5);
    o.cdsb = -(o.cdgb + db + o.cddb);
    o.cbgb = -(o.cggb + 2.0 * o.cdgb);
    o.cbdb = -(o.cgdb + 2.0 * o.cddb);
    o.cbsb = -(o.cgsb + 2.0 * o.cdsb);
  }
}

// capMod 1 (N_DEV_MOSFET_B3.C:6360-6520, piecewise: body charge, then the
// saturation or linear inversion charge) and capMod 2 (6521-6685, smooth:
// accumulation charge Qac0 on Vfbeff, body charge Qsub0, inversion charge
// on VdseffCV), on the smoothed Vgsteff of the capMod > 0 preamble and
// the flat-band voltage vfbzb.  xpart selects the source share.
__device__ __forceinline__ void b3_charge_cap12(
    int capmod, double xpart, double vfbzb, double k1ox, double CoxWL,
    double abulkCVfactor, double VbseffCV, double dVbseffCV_dVb,
    double dVbseff_dVb, double Vgs_eff, double dVgs_eff_dVg, double Vds,
    double Abulk0, double dAbulk0_dVb, double Vgsteff, double dVgsteff_dVg,
    double dVgsteff_dVd, double dVgsteff_dVb, B3ChanQ &o) {
  const double AbulkCV = Abulk0 * abulkCVfactor;
  const double dAbulkCV_dVb = abulkCVfactor * dAbulk0_dVb;
  const double VdsatCV = Vgsteff / AbulkCV;
  const int part = (xpart > 0.5) ? 0 : (xpart < 0.5) ? 1 : 2;
  double qgate, qbulk, qsrc;
  double Cgg, Cgd, Cgb, Cbg, Cbd, Cbb, Csg, Csd, Csb;
  double Cgg1, Cgd1, Cgb1, Cbg1, Cbd1, Cbb1;
  double T0, T1, T2, T3, T4, T5, T6, T7;
  if (capmod == 1) {
    const double Arg1 = Vgs_eff - VbseffCV - vfbzb - Vgsteff;
    if (Arg1 <= 0.0) {
      qgate = CoxWL * Arg1;
      Cgg = CoxWL * (dVgs_eff_dVg - dVgsteff_dVg);
      Cgd = -CoxWL * dVgsteff_dVd;
      Cgb = -CoxWL * (dVbseffCV_dVb + dVgsteff_dVb);
    } else {
      T0 = 0.5 * k1ox;
      T1 = sqrt(T0 * T0 + Arg1);
      T2 = CoxWL * T0 / T1;
      qgate = CoxWL * k1ox * (T1 - T0);
      Cgg = T2 * (dVgs_eff_dVg - dVgsteff_dVg);
      Cgd = -T2 * dVgsteff_dVd;
      Cgb = -T2 * (dVbseffCV_dVb + dVgsteff_dVb);
    }
    qbulk = -qgate;
    Cbg = -Cgg;
    Cbd = -Cgd;
    Cbb = -Cgb;
    const double c3 = CoxWL / 3.0, c23 = 2.0 * c3;
    if (VdsatCV < Vds) {
      const double dVdsatCV_dVg = 1.0 / AbulkCV;
      const double dVdsatCV_dVb = -VdsatCV * dAbulkCV_dVb / AbulkCV;
      T0 = Vgsteff - VdsatCV / 3.0;
      double dg = 1.0 - dVdsatCV_dVg / 3.0, db = -dVdsatCV_dVb / 3.0;
      qgate += CoxWL * T0;
      Cgg1 = CoxWL * dg;
      Cgb1 = CoxWL * db + Cgg1 * dVgsteff_dVb;
      Cgd1 = Cgg1 * dVgsteff_dVd;
      Cgg1 *= dVgsteff_dVg;
      Cgg += Cgg1;
      Cgb += Cgb1;
      Cgd += Cgd1;
      T0 = VdsatCV - Vgsteff;
      dg = dVdsatCV_dVg - 1.0;
      db = dVdsatCV_dVb;
      qbulk += c3 * T0;
      Cbg1 = c3 * dg;
      Cbb1 = c3 * db + Cbg1 * dVgsteff_dVb;
      Cbd1 = Cbg1 * dVgsteff_dVd;
      Cbg1 *= dVgsteff_dVg;
      Cbg += Cbg1;
      Cbb += Cbb1;
      Cbd += Cbd1;
      T0 = (part == 0) ? -c23 : (part == 1) ? -0.4 * CoxWL : -c3;
      qsrc = T0 * Vgsteff;
      Csg = T0 * dVgsteff_dVg;
      Csb = T0 * dVgsteff_dVb;
      Csd = T0 * dVgsteff_dVd;
    } else {
      T0 = AbulkCV * Vds;
      T1 = 12.0 * (Vgsteff - 0.5 * T0 + 1.e-20);
      T2 = Vds / T1;
      T3 = T0 * T2;
      const double dT3_dVg = -12.0 * T2 * T2 * AbulkCV;
      const double dT3_dVd = 6.0 * T0 * (4.0 * Vgsteff - T0) / T1 / T1 - 0.5;
      const double dT3_dVb = 12.0 * T2 * T2 * dAbulkCV_dVb * Vgsteff;
      qgate += CoxWL * (Vgsteff - 0.5 * Vds + T3);
      Cgg1 = CoxWL * (1.0 + dT3_dVg);
      Cgb1 = CoxWL * dT3_dVb + Cgg1 * dVgsteff_dVb;
      Cgd1 = CoxWL * dT3_dVd + Cgg1 * dVgsteff_dVd;
      Cgg1 *= dVgsteff_dVg;
      Cgg += Cgg1;
      Cgb += Cgb1;
      Cgd += Cgd1;
      qbulk += CoxWL * (1.0 - AbulkCV) * (0.5 * Vds - T3);
      Cbg1 = -CoxWL * ((1.0 - AbulkCV) * dT3_dVg);
      Cbb1 = -CoxWL * ((1.0 - AbulkCV) * dT3_dVb + (0.5 * Vds - T3) * dAbulkCV_dVb) +
             Cbg1 * dVgsteff_dVb;
      Cbd1 = -CoxWL * (1.0 - AbulkCV) * dT3_dVd + Cbg1 * dVgsteff_dVd;
      Cbg1 *= dVgsteff_dVg;
      Cbg += Cbg1;
      Cbb += Cbb1;
      Cbd += Cbd1;
      if (part == 0) {
        T1 = T1 + T1;
        qsrc = -CoxWL * (0.5 * Vgsteff + 0.25 * T0 - T0 * T0 / T1);
        Csg = -CoxWL * (0.5 + 24.0 * T0 * Vds / T1 / T1 * AbulkCV);
        Csb = -CoxWL * (0.25 * Vds * dAbulkCV_dVb -
                        12.0 * T0 * Vds / T1 / T1 * (4.0 * Vgsteff - T0) * dAbulkCV_dVb) +
              Csg * dVgsteff_dVb;
        Csd = -CoxWL * (0.25 * AbulkCV - 12.0 * AbulkCV * T0 / T1 / T1 * (4.0 * Vgsteff - T0)) +
              Csg * dVgsteff_dVd;
        Csg *= dVgsteff_dVg;
      } else if (part == 1) {
        T1 = T1 / 12.0;
        T2 = 0.5 * CoxWL / (T1 * T1);
        T3 = Vgsteff * (2.0 * T0 * T0 / 3.0 + Vgsteff * (Vgsteff - 4.0 * T0 / 3.0)) -
             2.0 * T0 * T0 * T0 / 15.0;
        qsrc = -T2 * T3;
        T4 = 4.0 / 3.0 * Vgsteff * (Vgsteff - T0) + 0.4 * T0 * T0;
        Csg = -2.0 * qsrc / T1 -
              T2 * (Vgsteff * (3.0 * Vgsteff - 8.0 * T0 / 3.0) + 2.0 * T0 * T0 / 3.0);
        Csb = (qsrc / T1 * Vds + T2 * T4 * Vds) * dAbulkCV_dVb + Csg * dVgsteff_dVb;
        Csd = (qsrc / T1 + T2 * T4) * AbulkCV + Csg * dVgsteff_dVd;
        Csg *= dVgsteff_dVg;
      } else {
        qsrc = -0.5 * (qgate + qbulk);
        Csg = -0.5 * (Cgg1 + Cbg1);
        Csb = -0.5 * (Cgb1 + Cbb1);
        Csd = -0.5 * (Cgd1 + Cbd1);
      }
    }
  } else {
    const double V3 = vfbzb - Vgs_eff + VbseffCV - kDelta3;
    if (vfbzb <= 0.0) T0 = sqrt(V3 * V3 - 4.0 * kDelta3 * vfbzb);
    else T0 = sqrt(V3 * V3 + 4.0 * kDelta3 * vfbzb);
    T1 = 0.5 * (1.0 + V3 / T0);
    const double Vfbeff = vfbzb - 0.5 * (V3 + T0);
    const double dVfbeff_dVg = T1 * dVgs_eff_dVg;
    const double dVfbeff_dVb = -T1 * dVbseffCV_dVb;
    const double Qac0 = CoxWL * (Vfbeff - vfbzb);
    const double dQac0_dVg = CoxWL * dVfbeff_dVg;
    const double dQac0_dVb = CoxWL * dVfbeff_dVb;
    T0 = 0.5 * k1ox;
    T3 = Vgs_eff - Vfbeff - VbseffCV - Vgsteff;
    if (k1ox == 0.0) {
      T1 = 0.0;
      T2 = 0.0;
    } else if (T3 < 0.0) {
      T1 = T0 + T3 / k1ox;
      T2 = CoxWL;
    } else {
      T1 = sqrt(T0 * T0 + T3);
      T2 = CoxWL * T0 / T1;
    }
    const double Qsub0 = CoxWL * k1ox * (T1 - T0);
    const double dQsub0_dVg = T2 * (dVgs_eff_dVg - dVfbeff_dVg - dVgsteff_dVg);
    const double dQsub0_dVd = -T2 * dVgsteff_dVd;
    const double dQsub0_dVb = -T2 * (dVfbeff_dVb + dVbseffCV_dVb + dVgsteff_dVb);
    const double V4 = VdsatCV - Vds - kDelta4;
    T0 = sqrt(V4 * V4 + 4.0 * kDelta4 * VdsatCV);
    double VdseffCV = VdsatCV - 0.5 * (V4 + T0);
    T1 = 0.5 * (1.0 + V4 / T0);
    T2 = kDelta4 / T0;
    T3 = (1.0 - T1 - T2) / AbulkCV;
    double dVdseffCV_dVg = T3;
    const double dVdseffCV_dVd = T1;
    double dVdseffCV_dVb = -T3 * VdsatCV * dAbulkCV_dVb;
    if (Vds == 0.0) {
      VdseffCV = 0.0;
      dVdseffCV_dVg = 0.0;
      dVdseffCV_dVb = 0.0;
    }
    T0 = AbulkCV * VdseffCV;
    T1 = 12.0 * (Vgsteff - 0.5 * T0 + 1e-20);
    T2 = VdseffCV / T1;
    T3 = T0 * T2;
    T4 = (1.0 - 12.0 * T2 * T2 * AbulkCV);
    T5 = (6.0 * T0 * (4.0 * Vgsteff - T0) / (T1 * T1) - 0.5);
    T6 = 12.0 * T2 * T2 * Vgsteff;
    qgate = CoxWL * (Vgsteff - 0.5 * VdseffCV + T3);
    Cgg1 = CoxWL * (T4 + T5 * dVdseffCV_dVg);
    Cgd1 = CoxWL * T5 * dVdseffCV_dVd + Cgg1 * dVgsteff_dVd;
    Cgb1 = CoxWL * (T5 * dVdseffCV_dVb + T6 * dAbulkCV_dVb) + Cgg1 * dVgsteff_dVb;
    Cgg1 *= dVgsteff_dVg;
    T7 = 1.0 - AbulkCV;
    qbulk = CoxWL * T7 * (0.5 * VdseffCV - T3);
    T4 = -T7 * (T4 - 1.0);
    T5 = -T7 * T5;
    T6 = -(T7 * T6 + (0.5 * VdseffCV - T3));
    Cbg1 = CoxWL * (T4 + T5 * dVdseffCV_dVg);
    Cbd1 = CoxWL * T5 * dVdseffCV_dVd + Cbg1 * dVgsteff_dVd;
    Cbb1 = CoxWL * (T5 * dVdseffCV_dVb + T6 * dAbulkCV_dVb) + Cbg1 * dVgsteff_dVb;
    Cbg1 *= dVgsteff_dVg;
    if (part == 0) {
      T1 = T1 + T1;
      qsrc = -CoxWL * (0.5 * Vgsteff + 0.25 * T0 - T0 * T0 / T1);
      T7 = (4.0 * Vgsteff - T0) / (T1 * T1);
      T4 = -(0.5 + 24.0 * T0 * T0 / (T1 * T1));
      T5 = -(0.25 * AbulkCV - 12.0 * AbulkCV * T0 * T7);
      T6 = -(0.25 * VdseffCV - 12.0 * T0 * VdseffCV * T7);
      Csg = CoxWL * (T4 + T5 * dVdseffCV_dVg);
      Csd = CoxWL * T5 * dVdseffCV_dVd + Csg * dVgsteff_dVd;
      Csb = CoxWL * (T5 * dVdseffCV_dVb + T6 * dAbulkCV_dVb) + Csg * dVgsteff_dVb;
      Csg *= dVgsteff_dVg;
    } else if (part == 1) {
      T1 = T1 / 12.0;
      T2 = 0.5 * CoxWL / (T1 * T1);
      T3 = Vgsteff * (2.0 * T0 * T0 / 3.0 + Vgsteff * (Vgsteff - 4.0 * T0 / 3.0)) -
           2.0 * T0 * T0 * T0 / 15.0;
      qsrc = -T2 * T3;
      T7 = 4.0 / 3.0 * Vgsteff * (Vgsteff - T0) + 0.4 * T0 * T0;
      T4 = -2.0 * qsrc / T1 -
           T2 * (Vgsteff * (3.0 * Vgsteff - 8.0 * T0 / 3.0) + 2.0 * T0 * T0 / 3.0);
      T5 = (qsrc / T1 + T2 * T7) * AbulkCV;
      T6 = (qsrc / T1 * VdseffCV + T2 * T7 * VdseffCV);
      Csg = (T4 + T5 * dVdseffCV_dVg);
      Csd = T5 * dVdseffCV_dVd + Csg * dVgsteff_dVd;
      Csb = (T5 * dVdseffCV_dVb + T6 * dAbulkCV_dVb) + Csg * dVgsteff_dVb;
      Csg *= dVgsteff_dVg;
    } else {
      qsrc = -0.5 * (qgate + qbulk);
      Csg = -0.5 * (Cgg1 + Cbg1);
      Csb = -0.5 * (Cgb1 + Cbb1);
      Csd = -0.5 * (Cgd1 + Cbd1);
    }
    qgate += Qac0 + Qsub0;
    qbulk -= (Qac0 + Qsub0);
    Cgg = dQac0_dVg + dQsub0_dVg + Cgg1;
    Cgd = dQsub0_dVd + Cgd1;
    Cgb = dQac0_dVb + dQsub0_dVb + Cgb1;
    Cbg = Cbg1 - dQac0_dVg - dQsub0_dVg;
    Cbd = Cbd1 - dQsub0_dVd;
    Cbb = Cbb1 - dQac0_dVb - dQsub0_dVb;
  }
  Cgb *= dVbseff_dVb;
  Cbb *= dVbseff_dVb;
  Csb *= dVbseff_dVb;
  o.qgate = qgate;
  o.qbulk = qbulk;
  o.qsrc = qsrc;
  o.qdrn = -(qgate + qbulk + qsrc);
  o.cggb = Cgg;
  o.cgsb = -(Cgg + Cgd + Cgb);
  o.cgdb = Cgd;
  o.cdgb = -(Cgg + Cbg + Csg);
  o.cdsb = (Cgg + Cgd + Cgb + Cbg + Cbd + Cbb + Csg + Csd + Csb);
  o.cddb = -(Cgd + Cbd + Csd);
  o.cbgb = Cbg;
  o.cbsb = -(Cbg + Cbd + Cbb);
  o.cbdb = Cbd;
}

// What one instance's evaluation reads from HBM before any arithmetic: its
// flags, slot bases, the six terminal voltages and the five previous-store
// values.  Loaded in the kernels' prologues so that every load is in flight
// before the first wait.
struct B3Inputs {
  int32_t flags, sta, sto;
  double Vd, Vg, Vs, Vb, Vdp, Vsp;
  double old[5];  // vbd, vbs, vgs, vds, von of the store the iteration reads
};

__device__ __forceinline__ void b3_inputs(const Bsim3Dev &a,
                                          const SolverStateDev &ss,
                                          const LoadPtrs &p, int slot,
                                          B3Inputs &in) {
  const int S = a.n_slots;
  in.flags = a.flags[slot];
  in.sta = a.sta_base[slot];
  in.sto = a.sto_base[slot];
  // every index is valid for padding slots too (ground node, slot 0), so
  // nothing here waits on another load before issuing
  const int ld = a.lid4[slot], lg = a.lid4[S + slot], ls = a.lid4[2 * S + slot],
            lb = a.lid4[3 * S + slot];
  const int li_d = a.lid_int[slot], li_s = a.lid_int[S + slot];
  const int ldp = (in.flags & XF_HAS_DP) ? li_d : ld;
  const int lsp = (in.flags & XF_HAS_SP) ? li_s : ls;
  in.Vd = p.x[ld];
  in.Vg = p.x[lg];
  in.Vs = p.x[ls];
  in.Vb = p.x[lb];
  in.Vdp = p.x[ldp];
  in.Vsp = p.x[lsp];
  // Master::updateState reads the next store from the second Newton
  // iteration on, the current store at the first (N_DEV_MOSFET_B3.C:4799-4840)
  const double *st = ss.newtonIter == 0 ? p.sto_curr : p.sto_next;
  const bool from_store = !(ss.newtonIter == 0 && ss.dcop);
#pragma unroll
  for (int k = 0; k < 5; ++k) in.old[k] = from_store ? st[in.sto + k] : 0.0;
}

// Which part of an instance's evaluation a wave performs.  kFull: all of
// it.  kDC / kCV: two waves share the instances of one eval wave; both
// redo the limiting and the threshold voltage (cheap), then the kDC wave
// computes the channel and junction currents, their conductances, the
// store values and the F, dFdxdVp and dF/dx statements, and the kCV wave
// the capMod-3, junction and overlap charges and capacitances, the state
// values and the Q, dQdxdVp and dQ/dx statements.  Every value is computed
// by the same expressions as in kFull, so results are identical.
enum B3Role : int { kFull = 0, kDC = 1, kCV = 2 };

// The evaluation of one instance.  Card fields come from LDS (cd); the
// size-dependent record and junction fields through R / GEO:
//   kStaged = false: R -> the record's column of the field-major record
//                    array, GEO -> the slot's column of the geometry array;
//   kStaged = true:  both from this lane's column of the wave's LDS copy.
template <bool kStaged, int kRole>
__device__ __forceinline__ void b3_eval(const Bsim3Dev &a,
                                        const SolverStateDev &ss,
                                        const LoadPtrs &p, int slot,
                                        const double *cd, const double *R,
                                        const double *Gp, const B3Inputs &in) {
  const int S = a.n_slots;
  const double dtype = cd[B3C_DTYPE];
  const double gmin = cd[B3C_GMIN];
  const bool limiter = cd[B3C_LIMITER] != 0.0;
  const size_t rs = kStaged ? 64 : (size_t)a.n_sdp;
  const size_t gs = kStaged ? 64 : (size_t)S;
  auto P = [&](int f) { return R[(size_t)f * rs]; };
  auto GEO = [&](int f) { return Gp[(size_t)f * gs]; };

  const int32_t flags = in.flags;
  const bool OFF = flags & XF_OFF;
  const int sta = in.sta, sto = in.sto;
  const bool charge = ss.tranop || ss.acop || ss.transient;
  constexpr bool doDC = kRole != kCV, doCV = kRole != kDC;

  // ---- voltages and limiting (N_DEV_MOSFET_B3.C:4691-5018)
  const double Vd = in.Vd, Vg = in.Vg, Vs = in.Vs, Vb = in.Vb;
  const double Vsp = in.Vsp, Vdp = in.Vdp;
  const double Vddp = Vd - Vdp, Vssp = Vs - Vsp;
  double vbs = dtype * (Vb - Vsp), vgs = dtype * (Vg - Vsp),
         vds = dtype * (Vdp - Vsp);
  double vbd = vbs - vds, vgd = vgs - vds;
  bool origFlag = true, limitedFlag = false;
  const double vgs_orig = vgs, vds_orig = vds, vbs_orig = vbs,
               vbd_orig = vbd, vgd_orig = vgd;
  if (ss.initJct && !OFF && limiter) {
    vbs = 0.0;
    vgs = dtype * P(B3R_VTH0) + 0.1;
    vds = 0.1;
    origFlag = false;
    vbd = vbs - vds;
    vgd = vgs - vds;
  } else if ((ss.initFix || ss.initJct) && OFF) {
    vbs = vgs = vds = 0;
  }
  double vbs_old, vbd_old, vgs_old, vds_old, von_local;
  if (ss.newtonIter == 0 && ss.dcop) {
    vbs_old = vbs;
    vbd_old = vbd;
    vgs_old = vgs;
    vds_old = vds;
    von_local = 0.0;
  } else {
    vbd_old = in.old[0];
    vbs_old = in.old[1];
    vgs_old = in.old[2];
    vds_old = in.old[3];
    von_local = in.old[4];
  }
  if (limiter && !(ss.initFix && OFF)) {
    const double vgdo = vgs_old - vds_old;
    const double vt0 = 1.3806226e-23 * (27.0 + 273.15) / 1.6021918e-19;
    int Check = 1;
    if (vds_old >= 0.0) {
      vgs = dm_fetlim(vgs, vgs_old, von_local);
      vds = vgs - vgd;
      vds = dm_limvds(vds, vds_old);
      vgd = vgs - vds;
    } else {
      vgd = dm_fetlim(vgd, vgdo, von_local);
      vds = vgs - vgd;
      vds = -dm_limvds(-vds, -vds_old);
      vgs = vgd + vds;
    }
    if (vds >= 0.0) {
      vbs = dm_pnjlim(vbs, vbs_old, vt0, cd[B3C_VCRIT], Check);
      vbd = vbs - vds;
    } else {
      vbd = dm_pnjlim(vbd, vbd_old, vt0, cd[B3C_VCRIT], Check);
      vbs = vbd + vds;
    }
    if (vgs_orig != vgs || vds_orig != vds || vbs_orig != vbs ||
        vbd_orig != vbd || vgd_orig != vgd)
      origFlag = false;
    if (Check == 1) limitedFlag = true;
  }
  vbd = vbs - vds;
  vgd = vgs - vds;
  const double vgb = vgs - vbs;

  // ---- junction diodes (N_DEV_MOSFET_B3.C:5037-5125)
  const double Nvtm = cd[B3C_NVTM];
  const double ijth = cd[B3C_IJTH];
  double gbs = 0.0, cbs = 0.0, gbd = 0.0, cbd = 0.0;
  if constexpr (doDC) {
    const double ss_ = GEO(B3I_SSATCUR);
    if (ss_ <= 0.0) {
      gbs = gmin;
      cbs = gbs * vbs;
    } else if (ijth == 0.0 || vbs < GEO(B3I_VJSM)) {
      const double evbs = dm_exp(vbs / Nvtm);
      gbs = ss_ * evbs / Nvtm + gmin;
      cbs = ss_ * (evbs - 1.0) + gmin * vbs;
    } else {
      const double IsEv = GEO(B3I_ISEVJSM);
      const double T0 = IsEv / Nvtm;
      gbs = T0 + gmin;
      cbs = IsEv - ss_ + T0 * (vbs - GEO(B3I_VJSM)) + gmin * vbs;
    }
    const double ds_ = GEO(B3I_DSATCUR);
    if (ds_ <= 0.0) {
      gbd = gmin;
      cbd = gbd * vbd;
    } else if (ijth == 0.0 || vbd < GEO(B3I_VJDM)) {
      const double evbd = dm_exp(vbd / Nvtm);
      gbd = ds_ * evbd / Nvtm + gmin;
      cbd = ds_ * (evbd - 1.0) + gmin * vbd;
    } else {
      const double IsEv = GEO(B3I_ISEVJDM);
      const double T0 = IsEv / Nvtm;
      gbd = T0 + gmin;
      cbd = IsEv - ds_ + T0 * (vbd - GEO(B3I_VJDM)) + gmin * vbd;
    }
  }
  int mode;
  double Vds, Vgs, Vbs;
  if (vds >= 0.0) {
    mode = 1;
    Vds = vds;
    Vgs = vgs;
    Vbs = vbs;
  } else {
    mode = -1;
    Vds = -vds;
    Vgs = vgd;
    Vbs = vbd;
  }

  // ---- threshold voltage (N_DEV_MOSFET_B3.C:5186-5346)
  const double cox = P(B3R_COX), tox = P(B3R_TOX), factor1 = P(B3R_FACTOR1);
  const double Vtm = cd[B3C_VTM];
  const double phi = P(B3R_PHI), sqrtPhi = P(B3R_SQRTPHI);
  const double vbsc = P(B3R_VBSC);
  double T0, T1, T2, T3, T4, T5, T6, T7, T8, T9, T10;
  T0 = Vbs - vbsc - 0.001;
  T1 = sqrt(T0 * T0 - 0.004 * vbsc);
  double Vbseff = vbsc + 0.5 * (T0 + T1);
  const double dVbseff_dVb = 0.5 * (1.0 + T0 / T1);
  if (Vbseff < Vbs) Vbseff = Vbs;
  double Phis, dPhis_dVb, sqrtPhis, dsqrtPhis_dVb;
  if (Vbseff > 0.0) {
    T0 = phi / (phi + Vbseff);
    Phis = phi * T0;
    dPhis_dVb = -T0 * T0;
    sqrtPhis = P(B3R_PHIS3) / (phi + 0.5 * Vbseff);
    dsqrtPhis_dVb = -0.5 * sqrtPhis * sqrtPhis / P(B3R_PHIS3);
  } else {
    Phis = phi - Vbseff;
    dPhis_dVb = -1.0;
    sqrtPhis = sqrt(Phis);
    dsqrtPhis_dVb = -0.5 / sqrtPhis;
  }
  const double Xdep0 = P(B3R_XDEP0);
  const double Xdep = Xdep0 * sqrtPhis / sqrtPhi;
  const double dXdep_dVb = (Xdep0 / sqrtPhi) * dsqrtPhis_dVb;
  const double Leff = P(B3R_LEFF);
  T3 = sqrt(Xdep);
  const double V0 = P(B3R_VBI) - phi;
  double lt1, dlt1_dVb, ltw, dltw_dVb;
  {
    const double dvt2 = P(B3R_DVT2);
    T0 = dvt2 * Vbseff;
    if (T0 >= -0.5) {
      T1 = 1.0 + T0;
      T2 = dvt2;
    } else {
      T4 = 1.0 / (3.0 + 8.0 * T0);
      T1 = (1.0 + 3.0 * T0) * T4;
      T2 = dvt2 * T4 * T4;
    }
    lt1 = factor1 * T3 * T1;
    dlt1_dVb = factor1 * (0.5 / T3 * T1 * dXdep_dVb + T3 * T2);
    const double dvt2w = P(B3R_DVT2W);
    T0 = dvt2w * Vbseff;
    if (T0 >= -0.5) {
      T1 = 1.0 + T0;
      T2 = dvt2w;
    } else {
      T4 = 1.0 / (3.0 + 8.0 * T0);
      T1 = (1.0 + 3.0 * T0) * T4;
      T2 = dvt2w * T4 * T4;
    }
    ltw = factor1 * T3 * T1;
    dltw_dVb = factor1 * (0.5 / T3 * T1 * dXdep_dVb + T3 * T2);
  }
  double Theta0, dTheta0_dVb;
  T0 = -0.5 * P(B3R_DVT1) * Leff / lt1;
  if (T0 > -kExpThreshold) {
    T1 = dm_exp(T0);
    Theta0 = T1 * (1.0 + 2.0 * T1);
    const double dT1_dVb = -T0 / lt1 * T1 * dlt1_dVb;
    dTheta0_dVb = (1.0 + 4.0 * T1) * dT1_dVb;
  } else {
    T1 = kMinExp;
    Theta0 = T1 * (1.0 + 2.0 * T1);
    dTheta0_dVb = 0.0;
  }
  const double dvt0 = P(B3R_DVT0);
  const double thetavth = dvt0 * Theta0;
  const double Delt_vth = thetavth * V0;
  const double dDelt_vth_dVb = dvt0 * dTheta0_dVb * V0;
  const double weff = P(B3R_WEFF);
  double dT2_dVb;
  T0 = -0.5 * P(B3R_DVT1W) * weff * Leff / ltw;
  if (T0 > -kExpThreshold) {
    T1 = dm_exp(T0);
    T2 = T1 * (1.0 + 2.0 * T1);
    const double dT1_dVb = -T0 / ltw * T1 * dltw_dVb;
    dT2_dVb = (1.0 + 4.0 * T1) * dT1_dVb;
  } else {
    T1 = kMinExp;
    T2 = T1 * (1.0 + 2.0 * T1);
    dT2_dVb = 0.0;
  }
  const double dvt0w = P(B3R_DVT0W);
  T0 = dvt0w * T2;
  T2 = T0 * V0;
  dT2_dVb = dvt0w * dT2_dVb * V0;
  const double TempRatio = cd[B3C_TEMPRATIO];
  const double k1ox = P(B3R_K1OX), kt2 = P(B3R_KT2);
  T0 = sqrt(1.0 + P(B3R_NLX) / Leff);
  T1 = k1ox * (T0 - 1.0) * sqrtPhi +
       (P(B3R_KT1) + P(B3R_KT1L) / Leff + kt2 * Vbseff) * TempRatio;
  double tmp2 = tox * phi / (weff + P(B3R_W0));
  const double etab = P(B3R_ETAB), theta0vb0 = P(B3R_THETA0VB0);
  T3 = P(B3R_ETA0) + etab * Vbseff;
  if (T3 < 1.0e-4) {
    T9 = 1.0 / (3.0 - 2.0e4 * T3);
    T3 = (2.0e-4 - T3) * T9;
    T4 = T9 * T9;
  } else {
    T4 = 1.0;
  }
  const double dDIBL_Sft_dVd = T3 * theta0vb0;
  const double DIBL_Sft = dDIBL_Sft_dVd * Vds;
  const double k2ox = P(B3R_K2OX), k3b = P(B3R_K3B);
  const double Vth = dtype * P(B3R_VTH0) - P(B3R_K1) * sqrtPhi + k1ox * sqrtPhis -
                     k2ox * Vbseff - Delt_vth - T2 + (P(B3R_K3) + k3b * Vbseff) * tmp2 +
                     T1 - DIBL_Sft;
  const double von = Vth;
  const double dVth_dVb = k1ox * dsqrtPhis_dVb - k2ox - dDelt_vth_dVb - dT2_dVb +
                          k3b * tmp2 - etab * Vds * theta0vb0 * T4 + kt2 * TempRatio;
  const double dVth_dVd = -dDIBL_Sft_dVd;
  double n, dn_dVb, dn_dVd;
  {
    const double cdscb = P(B3R_CDSCB), cdscd = P(B3R_CDSCD);
    tmp2 = P(B3R_NFACTOR) * kEpsSi / Xdep;
    const double tmp3 = P(B3R_CDSC) + cdscb * Vbseff + cdscd * Vds;
    const double tmp4 = (tmp2 + tmp3 * Theta0 + P(B3R_CIT)) / cox;
    if (tmp4 >= -0.5) {
      n = 1.0 + tmp4;
      dn_dVb = (-tmp2 / Xdep * dXdep_dVb + tmp3 * dTheta0_dVb + cdscb * Theta0) / cox;
      dn_dVd = cdscd * Theta0 / cox;
    } else {
      T0 = 1.0 / (3.0 + 8.0 * tmp4);
      n = (1.0 + 3.0 * tmp4) * T0;
      T0 *= T0;
      dn_dVb = (-tmp2 / Xdep * dXdep_dVb + tmp3 * dTheta0_dVb + cdscb * Theta0) / cox * T0;
      dn_dVd = cdscd * Theta0 / cox * T0;
    }
  }
  // ---- poly depletion, effective Vgs-Vth (N_DEV_MOSFET_B3.C:5349-5419)
  double Vgs_eff, dVgs_eff_dVg;
  {
    const double ngate = P(B3R_NGATE);
    T0 = P(B3R_VFB) + phi;
    if ((ngate > 1.e18) && (ngate < 1.e25) && (Vgs > T0)) {
      T1 = 1.0e6 * kQ * kEpsSi * ngate / (cox * cox);
      T4 = sqrt(1.0 + 2.0 * (Vgs - T0) / T1);
      T2 = T1 * (T4 - 1.0);
      T3 = 0.5 * T2 * T2 / T1;
      T7 = 1.12 - T3 - 0.05;
      T6 = sqrt(T7 * T7 + 0.224);
      T5 = 1.12 - 0.5 * (T7 + T6);
      Vgs_eff = Vgs - T5;
      dVgs_eff_dVg = 1.0 - (0.5 - 0.5 / T4) * (1.0 + T7 / T6);
    } else {
      Vgs_eff = Vgs;
      dVgs_eff_dVg = 1.0;
    }
  }
  const double Vgst = Vgs_eff - Vth;
  const double voff = P(B3R_VOFF), cdep0 = P(B3R_CDEP0);
  double Vgsteff = 0.0, dVgsteff_dVg = 0.0, dVgsteff_dVd = 0.0,
         dVgsteff_dVb = 0.0;
  if constexpr (doDC) {
    T10 = 2.0 * n * Vtm;
    const double VgstNVt = Vgst / T10;
    const double ExpArg = (2.0 * voff - Vgst) / T10;
    if (VgstNVt > kExpThreshold) {
      Vgsteff = Vgst;
      dVgsteff_dVg = dVgs_eff_dVg;
      dVgsteff_dVd = -dVth_dVd;
      dVgsteff_dVb = -dVth_dVb;
    } else if (ExpArg > kExpThreshold) {
      T0 = (Vgst - voff) / (n * Vtm);
      const double ExpVgst = dm_exp(T0);
      Vgsteff = Vtm * cdep0 / cox * ExpVgst;
      dVgsteff_dVg = Vgsteff / (n * Vtm);
      dVgsteff_dVd = -dVgsteff_dVg * (dVth_dVd + T0 * Vtm * dn_dVd);
      dVgsteff_dVb = -dVgsteff_dVg * (dVth_dVb + T0 * Vtm * dn_dVb);
      dVgsteff_dVg *= dVgs_eff_dVg;
    } else {
      const double ExpVgst = dm_exp(VgstNVt);
      T1 = T10 * dm_log(1.0 + ExpVgst);
      const double dT1_dVg = ExpVgst / (1.0 + ExpVgst);
      const double dT1_dVb = -dT1_dVg * (dVth_dVb + Vgst / n * dn_dVb) + T1 / n * dn_dVb;
      const double dT1_dVd = -dT1_dVg * (dVth_dVd + Vgst / n * dn_dVd) + T1 / n * dn_dVd;
      const double dT2_dVg = -cox / (Vtm * cdep0) * dm_exp(ExpArg);
      T2 = 1.0 - T10 * dT2_dVg;
      const double dT2_dVd = -dT2_dVg * (dVth_dVd - 2.0 * Vtm * ExpArg * dn_dVd) + (T2 - 1.0) / n * dn_dVd;
      const double dT2_dVb2 = -dT2_dVg * (dVth_dVb - 2.0 * Vtm * ExpArg * dn_dVb) + (T2 - 1.0) / n * dn_dVb;
      Vgsteff = T1 / T2;
      T3 = T2 * T2;
      dVgsteff_dVg = (T2 * dT1_dVg - T1 * dT2_dVg) / T3 * dVgs_eff_dVg;
      dVgsteff_dVd = (T2 * dT1_dVd - T1 * dT2_dVd) / T3;
      dVgsteff_dVb = (T2 * dT1_dVb - T1 * dT2_dVb2) / T3;
    }
  }
  // ---- effective width, Rds, Abulk (N_DEV_MOSFET_B3.C:5422-5508)
  T9 = sqrtPhis - sqrtPhi;
  double Weff = weff - 2.0 * (P(B3R_DWG) * Vgsteff + P(B3R_DWB) * T9);
  double dWeff_dVg = -2.0 * P(B3R_DWG);
  double dWeff_dVb = -2.0 * P(B3R_DWB) * dsqrtPhis_dVb;
  if (Weff < 2.0e-8) {
    T0 = 1.0 / (6.0e-8 - 2.0 * Weff);
    Weff = 2.0e-8 * (4.0e-8 - Weff) * T0;
    T0 *= T0 * 4.0e-16;
    dWeff_dVg *= T0;
    dWeff_dVb *= T0;
  }
  double Rds, dRds_dVg, dRds_dVb;
  {
    const double rds0 = P(B3R_RDS0), prwg = P(B3R_PRWG), prwb = P(B3R_PRWB);
    T0 = prwg * Vgsteff + prwb * T9;
    if (T0 >= -0.9) {
      Rds = rds0 * (1.0 + T0);
      dRds_dVg = rds0 * prwg;
      dRds_dVb = rds0 * prwb * dsqrtPhis_dVb;
    } else {
      T1 = 1.0 / (17.0 + 20.0 * T0);
      Rds = rds0 * (0.8 + T0) * T1;
      T1 *= T1;
      dRds_dVg = rds0 * prwg * T1;
      dRds_dVb = rds0 * prwb * dsqrtPhis_dVb * T1;
    }
  }
  double Abulk, Abulk0, dAbulk_dVg, dAbulk_dVb, dAbulk0_dVb;
  {
    T1 = 0.5 * k1ox / sqrtPhis;
    const double dT1_dVb = -T1 / sqrtPhis * dsqrtPhis_dVb;
    T9 = sqrt(P(B3R_XJ) * Xdep);
    const double tmp1 = Leff + 2.0 * T9;
    T5 = Leff / tmp1;
    const double a0 = P(B3R_A0);
    tmp2 = a0 * T5;
    const double tmp3 = weff + P(B3R_B1);
    const double tmp4 = P(B3R_B0) / tmp3;
    T2 = tmp2 + tmp4;
    const double dT2_dVb2 = -T9 / tmp1 / Xdep * dXdep_dVb;
    T6 = T5 * T5;
    T7 = T5 * T6;
    Abulk0 = 1.0 + T1 * T2;
    dAbulk0_dVb = T1 * tmp2 * dT2_dVb2 + T2 * dT1_dVb;
    T8 = P(B3R_AGS) * a0 * T7;
    dAbulk_dVg = -T1 * T8;
    Abulk = Abulk0 + dAbulk_dVg * Vgsteff;
    dAbulk_dVb = dAbulk0_dVb - T8 * Vgsteff * (dT1_dVb + 3.0 * T1 * dT2_dVb2);
    if (Abulk0 < 0.1) {
      T9 = 1.0 / (3.0 - 20.0 * Abulk0);
      Abulk0 = (0.2 - Abulk0) * T9;
      dAbulk0_dVb *= T9 * T9;
    }
    if (Abulk < 0.1) {
      T9 = 1.0 / (3.0 - 20.0 * Abulk);
      Abulk = (0.2 - Abulk) * T9;
      T10 = T9 * T9;
      dAbulk_dVb *= T10;
      dAbulk_dVg *= T10;
    }
    const double keta = P(B3R_KETA);
    T2 = keta * Vbseff;
    double dT0_dVb;
    if (T2 >= -0.9) {
      T0 = 1.0 / (1.0 + T2);
      dT0_dVb = -keta * T0 * T0;
    } else {
      T1 = 1.0 / (0.8 + T2);
      T0 = (17.0 + 20.0 * T2) * T1;
      dT0_dVb = -keta * T1 * T1;
    }
    dAbulk_dVg *= T0;
    dAbulk_dVb = dAbulk_dVb * T0 + Abulk * dT0_dVb;
    dAbulk0_dVb = dAbulk0_dVb * T0 + Abulk0 * dT0_dVb;
    Abulk *= T0;
    Abulk0 *= T0;
  }
  double cdrain = 0.0, gds = 0.0, gm = 0.0, gmbs = 0.0;
  double gbbs = 0.0, gbgs = 0.0, gbds = 0.0, csub = 0.0;
  if constexpr (doDC) {
  // ---- mobility (N_DEV_MOSFET_B3.C:5511-5563)
  double ueff, dueff_dVg, dueff_dVd, dueff_dVb;
  {
    const double mobMod = cd[B3C_MOBMOD];
    const double ua = P(B3R_UA), ub = P(B3R_UB), uc = P(B3R_UC);
    double Denomi, dDenomi_dVg, dDenomi_dVd, dDenomi_dVb;
    if (mobMod == 1.0) {
      T0 = Vgsteff + Vth + Vth;
      T2 = ua + uc * Vbseff;
      T3 = T0 / tox;
      T5 = T3 * (T2 + ub * T3);
      dDenomi_dVg = (T2 + 2.0 * ub * T3) / tox;
      dDenomi_dVd = dDenomi_dVg * 2.0 * dVth_dVd;
      dDenomi_dVb = dDenomi_dVg * 2.0 * dVth_dVb + uc * T3;
    } else if (mobMod == 2.0) {
      T5 = Vgsteff / tox * (ua + uc * Vbseff + ub * Vgsteff / tox);
      dDenomi_dVg = (ua + uc * Vbseff + 2.0 * ub * Vgsteff / tox) / tox;
      dDenomi_dVd = 0.0;
      dDenomi_dVb = Vgsteff * uc / tox;
    } else {
      T0 = Vgsteff + Vth + Vth;
      T2 = 1.0 + uc * Vbseff;
      T3 = T0 / tox;
      T4 = T3 * (ua + ub * T3);
      T5 = T4 * T2;
      dDenomi_dVg = (ua + 2.0 * ub * T3) * T2 / tox;
      dDenomi_dVd = dDenomi_dVg * 2.0 * dVth_dVd;
      dDenomi_dVb = dDenomi_dVg * 2.0 * dVth_dVb + uc * T4;
    }
    if (T5 >= -0.8) {
      Denomi = 1.0 + T5;
    } else {
      T9 = 1.0 / (7.0 + 10.0 * T5);
      Denomi = (0.6 + T5) * T9;
      T9 *= T9;
      dDenomi_dVg *= T9;
      dDenomi_dVd *= T9;
      dDenomi_dVb *= T9;
    }
    ueff = P(B3R_U0TEMP) / Denomi;
    T9 = -ueff / Denomi;
    dueff_dVg = T9 * dDenomi_dVg;
    dueff_dVd = T9 * dDenomi_dVd;
    dueff_dVb = T9 * dDenomi_dVb;
  }
  // ---- saturation voltage (N_DEV_MOSFET_B3.C:5566-5676)
  const double vsattemp = P(B3R_VSATTEMP);
  const double WVCox = Weff * vsattemp * cox;
  const double WVCoxRds = WVCox * Rds;
  const double Esat = 2.0 * vsattemp / ueff;
  const double EsatL = Esat * Leff;
  T0 = -EsatL / ueff;
  const double dEsatL_dVg = T0 * dueff_dVg;
  const double dEsatL_dVd = T0 * dueff_dVd;
  const double dEsatL_dVb = T0 * dueff_dVb;
  double Lambda, dLambda_dVg;
  {
    const double a1 = P(B3R_A1), a2 = P(B3R_A2);
    if (a1 == 0.0) {
      Lambda = a2;
      dLambda_dVg = 0.0;
    } else if (a1 > 0.0) {
      T0 = 1.0 - a2;
      T1 = T0 - a1 * Vgsteff - 0.0001;
      T2 = sqrt(T1 * T1 + 0.0004 * T0);
      Lambda = a2 + T0 - 0.5 * (T1 + T2);
      dLambda_dVg = 0.5 * a1 * (1.0 + T1 / T2);
    } else {
      T1 = a2 + a1 * Vgsteff - 0.0001;
      T2 = sqrt(T1 * T1 + 0.0004 * a2);
      Lambda = 0.5 * (T1 + T2);
      dLambda_dVg = 0.5 * a1 * (1.0 + T1 / T2);
    }
  }
  const double Vgst2Vtm = Vgsteff + 2.0 * Vtm;
  double tmp1, tmp3;
  if (Rds > 0) {
    tmp2 = dRds_dVg / Rds + dWeff_dVg / Weff;
    tmp3 = dRds_dVb / Rds + dWeff_dVb / Weff;
  } else {
    tmp2 = dWeff_dVg / Weff;
    tmp3 = dWeff_dVb / Weff;
  }
  double Vdsat, dVdsat_dVg, dVdsat_dVd, dVdsat_dVb;
  if ((Rds == 0.0) && (Lambda == 1.0)) {
    T0 = 1.0 / (Abulk * EsatL + Vgst2Vtm);
    tmp1 = 0.0;
    T1 = T0 * T0;
    T2 = Vgst2Vtm * T0;
    T3 = EsatL * Vgst2Vtm;
    Vdsat = T3 * T0;
    const double dT0_dVg = -(Abulk * dEsatL_dVg + EsatL * dAbulk_dVg + 1.0) * T1;
    const double dT0_dVd = -(Abulk * dEsatL_dVd) * T1;
    const double dT0_dVb = -(Abulk * dEsatL_dVb + dAbulk_dVb * EsatL) * T1;
    dVdsat_dVg = T3 * dT0_dVg + T2 * dEsatL_dVg + EsatL * T0;
    dVdsat_dVd = T3 * dT0_dVd + T2 * dEsatL_dVd;
    dVdsat_dVb = T3 * dT0_dVb + T2 * dEsatL_dVb;
  } else {
    tmp1 = dLambda_dVg / (Lambda * Lambda);
    T9 = Abulk * WVCoxRds;
    T8 = Abulk * T9;
    T7 = Vgst2Vtm * T9;
    T6 = Vgst2Vtm * WVCoxRds;
    T0 = 2.0 * Abulk * (T9 - 1.0 + 1.0 / Lambda);
    const double dT0_dVg = 2.0 * (T8 * tmp2 - Abulk * tmp1 + (2.0 * T9 + 1.0 / Lambda - 1.0) * dAbulk_dVg);
    const double dT0_dVb = 2.0 * (T8 * (2.0 / Abulk * dAbulk_dVb + tmp3) + (1.0 / Lambda - 1.0) * dAbulk_dVb);
    const double dT0_dVd = 0.0;
    T1 = Vgst2Vtm * (2.0 / Lambda - 1.0) + Abulk * EsatL + 3.0 * T7;
    const double dT1_dVg = (2.0 / Lambda - 1.0) - 2.0 * Vgst2Vtm * tmp1 + Abulk * dEsatL_dVg +
                           EsatL * dAbulk_dVg + 3.0 * (T9 + T7 * tmp2 + T6 * dAbulk_dVg);
    const double dT1_dVb = Abulk * dEsatL_dVb + EsatL * dAbulk_dVb + 3.0 * (T6 * dAbulk_dVb + T7 * tmp3);
    const double dT1_dVd = Abulk * dEsatL_dVd;
    T2 = Vgst2Vtm * (EsatL + 2.0 * T6);
    const double dT2_dVg = EsatL + Vgst2Vtm * dEsatL_dVg + T6 * (4.0 + 2.0 * Vgst2Vtm * tmp2);
    const double dT2_dVb2 = Vgst2Vtm * (dEsatL_dVb + 2.0 * T6 * tmp3);
    const double dT2_dVd = Vgst2Vtm * dEsatL_dVd;
    T3 = sqrt(T1 * T1 - 2.0 * T0 * T2);
    Vdsat = (T1 - T3) / T0;
    dVdsat_dVg = (dT1_dVg - (T1 * dT1_dVg - dT0_dVg * T2 - T0 * dT2_dVg) / T3 - Vdsat * dT0_dVg) / T0;
    dVdsat_dVb = (dT1_dVb - (T1 * dT1_dVb - dT0_dVb * T2 - T0 * dT2_dVb2) / T3 - Vdsat * dT0_dVb) / T0;
    dVdsat_dVd = (dT1_dVd - (T1 * dT1_dVd - T0 * dT2_dVd) / T3) / T0;
    (void)dT0_dVd;
  }
  // ---- effective Vds (N_DEV_MOSFET_B3.C:5679-5731)
  const double delta = P(B3R_DELTA);
  T1 = Vdsat - Vds - delta;
  double dT1_dVg = dVdsat_dVg, dT1_dVd = dVdsat_dVd - 1.0, dT1_dVb = dVdsat_dVb;
  T2 = sqrt(T1 * T1 + 4.0 * delta * Vdsat);
  T0 = T1 / T2;
  T3 = 2.0 * delta / T2;
  double dT2_dVg = T0 * dT1_dVg + T3 * dVdsat_dVg;
  double dT2_dVd = T0 * dT1_dVd + T3 * dVdsat_dVd;
  dT2_dVb = T0 * dT1_dVb + T3 * dVdsat_dVb;
  double Vdseff = Vdsat - 0.5 * (T1 + T2);
  double dVdseff_dVg = dVdsat_dVg - 0.5 * (dT1_dVg + dT2_dVg);
  const double dVdseff_dVd = dVdsat_dVd - 0.5 * (dT1_dVd + dT2_dVd);
  double dVdseff_dVb = dVdsat_dVb - 0.5 * (dT1_dVb + dT2_dVb);
  if (Vds == 0.0) {
    Vdseff = 0.0;
    dVdseff_dVg = 0.0;
    dVdseff_dVb = 0.0;
  }
  double Vasat, dVasat_dVg, dVasat_dVb, dVasat_dVd;
  {
    const double tmp4 = 1.0 - 0.5 * Abulk * Vdsat / Vgst2Vtm;
    T9 = WVCoxRds * Vgsteff;
    T8 = T9 / Vgst2Vtm;
    T0 = EsatL + Vdsat + 2.0 * T9 * tmp4;
    T7 = 2.0 * WVCoxRds * tmp4;
    const double dT0_dVg = dEsatL_dVg + dVdsat_dVg + T7 * (1.0 + tmp2 * Vgsteff) -
                           T8 * (Abulk * dVdsat_dVg - Abulk * Vdsat / Vgst2Vtm + Vdsat * dAbulk_dVg);
    const double dT0_dVb = dEsatL_dVb + dVdsat_dVb + T7 * tmp3 * Vgsteff -
                           T8 * (dAbulk_dVb * Vdsat + Abulk * dVdsat_dVb);
    const double dT0_dVd = dEsatL_dVd + dVdsat_dVd - T8 * Abulk * dVdsat_dVd;
    T9 = WVCoxRds * Abulk;
    T1 = 2.0 / Lambda - 1.0 + T9;
    dT1_dVg = -2.0 * tmp1 + WVCoxRds * (Abulk * tmp2 + dAbulk_dVg);
    dT1_dVb = dAbulk_dVb * WVCoxRds + T9 * tmp3;
    Vasat = T0 / T1;
    dVasat_dVg = (dT0_dVg - Vasat * dT1_dVg) / T1;
    dVasat_dVb = (dT0_dVb - Vasat * dT1_dVb) / T1;
    dVasat_dVd = dT0_dVd / T1;
  }
  if (Vdseff > Vds) Vdseff = Vds;
  const double diffVds = Vds - Vdseff;
  // ---- output resistance (N_DEV_MOSFET_B3.C:5734-5871)
  const double litl = P(B3R_LITL);
  double VACLM, dVACLM_dVg, dVACLM_dVd, dVACLM_dVb;
  {
    const double pclm = P(B3R_PCLM);
    if ((pclm > 0.0) && (diffVds > 1.0e-10)) {
      T0 = 1.0 / (pclm * Abulk * litl);
      const double dT0_dVb = -T0 / Abulk * dAbulk_dVb;
      const double dT0_dVg = -T0 / Abulk * dAbulk_dVg;
      T2 = Vgsteff / EsatL;
      T1 = Leff * (Abulk + T2);
      dT1_dVg = Leff * ((1.0 - T2 * dEsatL_dVg) / EsatL + dAbulk_dVg);
      dT1_dVb = Leff * (dAbulk_dVb - T2 * dEsatL_dVb / EsatL);
      dT1_dVd = -T2 * dEsatL_dVd / Esat;
      T9 = T0 * T1;
      VACLM = T9 * diffVds;
      dVACLM_dVg = T0 * dT1_dVg * diffVds - T9 * dVdseff_dVg + T1 * diffVds * dT0_dVg;
      dVACLM_dVb = (dT0_dVb * T1 + T0 * dT1_dVb) * diffVds - T9 * dVdseff_dVb;
      dVACLM_dVd = T0 * dT1_dVd * diffVds + T9 * (1.0 - dVdseff_dVd);
    } else {
      VACLM = kMaxExp;
      dVACLM_dVd = dVACLM_dVg = dVACLM_dVb = 0.0;
    }
  }
  double VADIBL, dVADIBL_dVg, dVADIBL_dVd, dVADIBL_dVb;
  {
    const double thetaRout = P(B3R_THETAROUT);
    if (thetaRout > 0.0) {
      T8 = Abulk * Vdsat;
      T0 = Vgst2Vtm * T8;
      const double dT0_dVg = Vgst2Vtm * Abulk * dVdsat_dVg + T8 + Vgst2Vtm * Vdsat * dAbulk_dVg;
      const double dT0_dVb = Vgst2Vtm * (dAbulk_dVb * Vdsat + Abulk * dVdsat_dVb);
      const double dT0_dVd = Vgst2Vtm * Abulk * dVdsat_dVd;
      T1 = Vgst2Vtm + T8;
      dT1_dVg = 1.0 + Abulk * dVdsat_dVg + Vdsat * dAbulk_dVg;
      dT1_dVb = Abulk * dVdsat_dVb + dAbulk_dVb * Vdsat;
      dT1_dVd = Abulk * dVdsat_dVd;
      T9 = T1 * T1;
      T2 = thetaRout;
      VADIBL = (Vgst2Vtm - T0 / T1) / T2;
      dVADIBL_dVg = (1.0 - dT0_dVg / T1 + T0 * dT1_dVg / T9) / T2;
      dVADIBL_dVb = (-dT0_dVb / T1 + T0 * dT1_dVb / T9) / T2;
      dVADIBL_dVd = (-dT0_dVd / T1 + T0 * dT1_dVd / T9) / T2;
      const double pdiblb = P(B3R_PDIBLB);
      T7 = pdiblb * Vbseff;
      if (T7 >= -0.9) {
        T3 = 1.0 / (1.0 + T7);
        VADIBL *= T3;
        dVADIBL_dVg *= T3;
        dVADIBL_dVb = (dVADIBL_dVb - VADIBL * pdiblb) * T3;
        dVADIBL_dVd *= T3;
      } else {
        T4 = 1.0 / (0.8 + T7);
        T3 = (17.0 + 20.0 * T7) * T4;
        dVADIBL_dVg *= T3;
        dVADIBL_dVb = dVADIBL_dVb * T3 - VADIBL * pdiblb * T4 * T4;
        dVADIBL_dVd *= T3;
        VADIBL *= T3;
      }
    } else {
      VADIBL = kMaxExp;
      dVADIBL_dVd = dVADIBL_dVg = dVADIBL_dVb = 0.0;
    }
  }
  double dT0_dVg, dT0_dVd, dT0_dVb;
  T8 = P(B3R_PVAG) / EsatL;
  T9 = T8 * Vgsteff;
  if (T9 > -0.9) {
    T0 = 1.0 + T9;
    dT0_dVg = T8 * (1.0 - Vgsteff * dEsatL_dVg / EsatL);
    dT0_dVb = -T9 * dEsatL_dVb / EsatL;
    dT0_dVd = -T9 * dEsatL_dVd / EsatL;
  } else {
    T1 = 1.0 / (17.0 + 20.0 * T9);
    T0 = (0.8 + T9) * T1;
    T1 *= T1;
    dT0_dVg = T8 * (1.0 - Vgsteff * dEsatL_dVg / EsatL) * T1;
    T9 *= T1 / EsatL;
    dT0_dVb = -T9 * dEsatL_dVb;
    dT0_dVd = -T9 * dEsatL_dVd;
  }
  tmp1 = VACLM * VACLM;
  tmp2 = VADIBL * VADIBL;
  tmp3 = VACLM + VADIBL;
  T1 = VACLM * VADIBL / tmp3;
  tmp3 *= tmp3;
  dT1_dVg = (tmp1 * dVADIBL_dVg + tmp2 * dVACLM_dVg) / tmp3;
  dT1_dVd = (tmp1 * dVADIBL_dVd + tmp2 * dVACLM_dVd) / tmp3;
  dT1_dVb = (tmp1 * dVADIBL_dVb + tmp2 * dVACLM_dVb) / tmp3;
  const double Va = Vasat + T0 * T1;
  const double dVa_dVg = dVasat_dVg + T1 * dT0_dVg + T0 * dT1_dVg;
  const double dVa_dVd = dVasat_dVd + T1 * dT0_dVd + T0 * dT1_dVd;
  const double dVa_dVb = dVasat_dVb + T1 * dT0_dVb + T0 * dT1_dVb;
  double VASCBE, dVASCBE_dVg, dVASCBE_dVd, dVASCBE_dVb;
  {
    const double pscbe1 = P(B3R_PSCBE1), pscbe2 = P(B3R_PSCBE2);
    if (pscbe2 > 0.0) {
      if (diffVds > pscbe1 * litl / kExpThreshold) {
        T0 = pscbe1 * litl / diffVds;
        VASCBE = Leff * dm_exp(T0) / pscbe2;
        T1 = T0 * VASCBE / diffVds;
        dVASCBE_dVg = T1 * dVdseff_dVg;
        dVASCBE_dVd = -T1 * (1.0 - dVdseff_dVd);
        dVASCBE_dVb = T1 * dVdseff_dVb;
      } else {
        VASCBE = kMaxExp * Leff / pscbe2;
        dVASCBE_dVg = dVASCBE_dVd = dVASCBE_dVb = 0.0;
      }
    } else {
      VASCBE = kMaxExp;
      dVASCBE_dVg = dVASCBE_dVd = dVASCBE_dVb = 0.0;
    }
  }
  // ---- drain current (N_DEV_MOSFET_B3.C:5874-5940)
  const double CoxWovL = cox * Weff / Leff;
  const double beta = ueff * CoxWovL;
  const double dbeta_dVg = CoxWovL * dueff_dVg + beta * dWeff_dVg / Weff;
  const double dbeta_dVd = CoxWovL * dueff_dVd;
  const double dbeta_dVb = CoxWovL * dueff_dVb + beta * dWeff_dVb / Weff;
  T0 = 1.0 - 0.5 * Abulk * Vdseff / Vgst2Vtm;
  dT0_dVg = -0.5 * (Abulk * dVdseff_dVg - Abulk * Vdseff / Vgst2Vtm + Vdseff * dAbulk_dVg) / Vgst2Vtm;
  dT0_dVd = -0.5 * Abulk * dVdseff_dVd / Vgst2Vtm;
  dT0_dVb = -0.5 * (Abulk * dVdseff_dVb + dAbulk_dVb * Vdseff) / Vgst2Vtm;
  const double fgche1 = Vgsteff * T0;
  const double dfgche1_dVg = Vgsteff * dT0_dVg + T0;
  const double dfgche1_dVd = Vgsteff * dT0_dVd;
  const double dfgche1_dVb = Vgsteff * dT0_dVb;
  T9 = Vdseff / EsatL;
  const double fgche2 = 1.0 + T9;
  const double dfgche2_dVg = (dVdseff_dVg - T9 * dEsatL_dVg) / EsatL;
  const double dfgche2_dVd = (dVdseff_dVd - T9 * dEsatL_dVd) / EsatL;
  const double dfgche2_dVb = (dVdseff_dVb - T9 * dEsatL_dVb) / EsatL;
  const double gche = beta * fgche1 / fgche2;
  const double dgche_dVg = (beta * dfgche1_dVg + fgche1 * dbeta_dVg - gche * dfgche2_dVg) / fgche2;
  const double dgche_dVd = (beta * dfgche1_dVd + fgche1 * dbeta_dVd - gche * dfgche2_dVd) / fgche2;
  const double dgche_dVb = (beta * dfgche1_dVb + fgche1 * dbeta_dVb - gche * dfgche2_dVb) / fgche2;
  T0 = 1.0 + gche * Rds;
  T9 = Vdseff / T0;
  const double Idl = gche * T9;
  const double dIdl_dVg = (gche * dVdseff_dVg + T9 * dgche_dVg) / T0 - Idl * gche / T0 * dRds_dVg;
  const double dIdl_dVd = (gche * dVdseff_dVd + T9 * dgche_dVd) / T0;
  const double dIdl_dVb = (gche * dVdseff_dVb + T9 * dgche_dVb - Idl * dRds_dVb * gche) / T0;
  T9 = diffVds / Va;
  T0 = 1.0 + T9;
  const double Idsa = Idl * T0;
  const double dIdsa_dVg = T0 * dIdl_dVg - Idl * (dVdseff_dVg + T9 * dVa_dVg) / Va;
  const double dIdsa_dVd = T0 * dIdl_dVd + Idl * (1.0 - dVdseff_dVd - T9 * dVa_dVd) / Va;
  const double dIdsa_dVb = T0 * dIdl_dVb - Idl * (dVdseff_dVb + T9 * dVa_dVb) / Va;
  T9 = diffVds / VASCBE;
  T0 = 1.0 + T9;
  const double Ids = Idsa * T0;
  double Gm = T0 * dIdsa_dVg - Idsa * (dVdseff_dVg + T9 * dVASCBE_dVg) / VASCBE;
  double Gds = T0 * dIdsa_dVd + Idsa * (1.0 - dVdseff_dVd - T9 * dVASCBE_dVd) / VASCBE;
  double Gmb = T0 * dIdsa_dVb - Idsa * (dVdseff_dVb + T9 * dVASCBE_dVb) / VASCBE;
  Gds += Gm * dVgsteff_dVd;
  Gmb += Gm * dVgsteff_dVb;
  Gm *= dVgsteff_dVg;
  Gmb *= dVbseff_dVb;
  // ---- substrate current (N_DEV_MOSFET_B3.C:5943-5977)
  double Isub, Gbd, Gbb, Gbg;
  {
    const double beta0 = P(B3R_BETA0);
    const double tmp = P(B3R_ALPHA0) + P(B3R_ALPHA1) * Leff;
    if ((tmp <= 0.0) || (beta0 <= 0.0)) {
      Isub = Gbd = Gbb = Gbg = 0.0;
    } else {
      T2 = tmp / Leff;
      if (diffVds > beta0 / kExpThreshold) {
        T0 = -beta0 / diffVds;
        T1 = T2 * diffVds * dm_exp(T0);
        T3 = T1 / diffVds * (T0 - 1.0);
        dT1_dVg = T3 * dVdseff_dVg;
        dT1_dVd = T3 * (dVdseff_dVd - 1.0);
        dT1_dVb = T3 * dVdseff_dVb;
      } else {
        T3 = T2 * kMinExp;
        T1 = T3 * diffVds;
        dT1_dVg = -T3 * dVdseff_dVg;
        dT1_dVd = T3 * (1.0 - dVdseff_dVd);
        dT1_dVb = -T3 * dVdseff_dVb;
      }
      Isub = T1 * Idsa;
      Gbg = T1 * dIdsa_dVg + Idsa * dT1_dVg;
      Gbd = T1 * dIdsa_dVd + Idsa * dT1_dVd;
      Gbb = T1 * dIdsa_dVb + Idsa * dT1_dVb;
      Gbd += Gbg * dVgsteff_dVd;
      Gbb += Gbg * dVgsteff_dVb;
      Gbg *= dVgsteff_dVg;
      Gbb *= dVbseff_dVb;
    }
  }
  cdrain = Ids;
  gds = Gds;
  gm = Gm;
  gmbs = Gmb;
  gbbs = Gbb;
  gbgs = Gbg;
  gbds = Gbd;
  csub = Isub;
  }  // doDC

  // ---- charges (capMod 0-3) and capacitances (N_DEV_MOSFET_B3.C:5994-7252)
  double qg = 0.0, qd = 0.0, qb = 0.0, qbs = 0.0, qbd = 0.0;
  double CAPcggb = 0.0, CAPcgdb = 0.0, CAPcgsb = 0.0, CAPcdgb = 0.0,
         CAPcddb = 0.0, CAPcdsb = 0.0, CAPcsgb = 0.0, CAPcsdb = 0.0,
         CAPcssb = 0.0, CAPcbgb = 0.0, CAPcbdb = 0.0, CAPcbsb = 0.0;
  if (doCV && charge) {
    const double xpart = cd[B3C_XPART];
    const int capmod = (int)cd[B3C_CAPMOD];
    double qgate = 0.0, qdrn = 0.0, qsrc = 0.0, qbulk = 0.0;
    double cggb = 0.0, cgsb = 0.0, cgdb = 0.0, cdgb = 0.0, cdsb = 0.0,
           cddb = 0.0, cbgb = 0.0, cbsb = 0.0, cbdb = 0.0;
    if (!(xpart < 0) && capmod == 0) {
      B3ChanQ o;
      b3_charge_cap0(xpart, P(B3R_VFBQ), phi, k1ox,
                     cox * P(B3R_WEFFCV) * P(B3R_LEFFCV), P(B3R_ABULKCVF), Vbs,
                     Vbseff, Phis, dPhis_dVb, sqrtPhis, dsqrtPhis_dVb, Vgs_eff,
                     dVgs_eff_dVg, Vds, Abulk0, dAbulk0_dVb, o);
      qgate = o.qgate; qdrn = o.qdrn; qsrc = o.qsrc; qbulk = o.qbulk;
      cggb = o.cggb; cgsb = o.cgsb; cgdb = o.cgdb; cdgb = o.cdgb;
      cdsb = o.cdsb; cddb = o.cddb; cbgb = o.cbgb; cbsb = o.cbsb;
      cbdb = o.cbdb;
    } else if (!(xpart < 0)) {
      double VbseffCV, dVbseffCV_dVb;
      if (Vbseff < 0.0) {
        VbseffCV = Vbseff;
        dVbseffCV_dVb = 1.0;
      } else {
        VbseffCV = phi - Phis;
        dVbseffCV_dVb = -dPhis_dVb;
      }
      const double CoxWL = cox * P(B3R_WEFFCV) * P(B3R_LEFFCV);
      const double noff = n * P(B3R_NOFF);
      const double dnoff_dVd = P(B3R_NOFF) * dn_dVd;
      const double dnoff_dVb = P(B3R_NOFF) * dn_dVb;
      T0 = Vtm * noff;
      const double voffcv = P(B3R_VOFFCV);
      const double VgstNVt = (Vgst - voffcv) / T0;
      double Vgsteffc, dVgsteffc_dVg, dVgsteffc_dVd, dVgsteffc_dVb;
      if (VgstNVt > kExpThreshold) {
        Vgsteffc = Vgst - voffcv;
        dVgsteffc_dVg = dVgs_eff_dVg;
        dVgsteffc_dVd = -dVth_dVd;
        dVgsteffc_dVb = -dVth_dVb;
      } else if (VgstNVt < -kExpThreshold) {
        Vgsteffc = T0 * dm_log(1.0 + kMinExp);
        dVgsteffc_dVg = 0.0;
        dVgsteffc_dVd = Vgsteffc / noff;
        dVgsteffc_dVb = dVgsteffc_dVd * dnoff_dVb;
        dVgsteffc_dVd *= dnoff_dVd;
      } else {
        const double ExpVgst = dm_exp(VgstNVt);
        Vgsteffc = T0 * dm_log(1.0 + ExpVgst);
        dVgsteffc_dVg = ExpVgst / (1.0 + ExpVgst);
        dVgsteffc_dVd = -dVgsteffc_dVg * (dVth_dVd + (Vgst - voffcv) / noff * dnoff_dVd) + Vgsteffc / noff * dnoff_dVd;
        dVgsteffc_dVb = -dVgsteffc_dVg * (dVth_dVb + (Vgst - voffcv) / noff * dnoff_dVb) + Vgsteffc / noff * dnoff_dVb;
        dVgsteffc_dVg *= dVgs_eff_dVg;
      }
      if (capmod != 3) {
        B3ChanQ o;
        b3_charge_cap12(capmod, xpart, P(B3R_VFBQ), k1ox, CoxWL, P(B3R_ABULKCVF),
                        VbseffCV, dVbseffCV_dVb, dVbseff_dVb, Vgs_eff, dVgs_eff_dVg,
                        Vds, Abulk0, dAbulk0_dVb, Vgsteffc, dVgsteffc_dVg,
                        dVgsteffc_dVd, dVgsteffc_dVb, o);
        qgate = o.qgate; qdrn = o.qdrn; qsrc = o.qsrc; qbulk = o.qbulk;
        cggb = o.cggb; cgsb = o.cgsb; cgdb = o.cgdb; cdgb = o.cdgb;
        cdsb = o.cdsb; cddb = o.cddb; cbgb = o.cbgb; cbsb = o.cbsb;
        cbdb = o.cbdb;
      } else {
        const double vfbzb = P(B3R_VFBQ);
        double V3 = vfbzb - Vgs_eff + VbseffCV - kDelta3;
        if (vfbzb <= 0.0) {
          T0 = sqrt(V3 * V3 - 4.0 * kDelta3 * vfbzb);
        } else {
          T0 = sqrt(V3 * V3 + 4.0 * kDelta3 * vfbzb);
        }
        T1 = 0.5 * (1.0 + V3 / T0);
        const double Vfbeff = vfbzb - 0.5 * (V3 + T0);
        const double dVfbeff_dVg = T1 * dVgs_eff_dVg;
        const double dVfbeff_dVb = -T1 * dVbseffCV_dVb;
        const double Cox = cox;
        double Tox = 1.0e8 * tox;
        T0 = (Vgs_eff - VbseffCV - vfbzb) / Tox;
        const double dT0c_dVg = dVgs_eff_dVg / Tox;
        const double dT0c_dVb = -dVbseffCV_dVb / Tox;
        const double acde = P(B3R_ACDE), ldeb = P(B3R_LDEB);
        double tmp = T0 * acde;
        double Tcen, dTcen_dVg, dTcen_dVd, dTcen_dVb;
        if ((-kExpThreshold < tmp) && (tmp < kExpThreshold)) {
          Tcen = ldeb * dm_exp(tmp);
          dTcen_dVg = acde * Tcen;
          dTcen_dVb = dTcen_dVg * dT0c_dVb;
          dTcen_dVg *= dT0c_dVg;
        } else if (tmp <= -kExpThreshold) {
          Tcen = ldeb * kMinExp;
          dTcen_dVg = dTcen_dVb = 0.0;
        } else {
          Tcen = ldeb * kMaxExp;
          dTcen_dVg = dTcen_dVb = 0.0;
        }
        const double LINK = 1.0e-3 * tox;
        V3 = ldeb - Tcen - LINK;
        double V4 = sqrt(V3 * V3 + 4.0 * LINK * ldeb);
        Tcen = ldeb - 0.5 * (V3 + V4);
        T1 = 0.5 * (1.0 + V3 / V4);
        dTcen_dVg *= T1;
        dTcen_dVb *= T1;
        double Ccen = kEpsSi / Tcen;
        T2 = Cox / (Cox + Ccen);
        double Coxeff = T2 * Ccen;
        T3 = -Ccen / Tcen;
        double dCoxeff_dVg = T2 * T2 * T3;
        double dCoxeff_dVb = dCoxeff_dVg * dTcen_dVb;
        dCoxeff_dVg *= dTcen_dVg;
        double CoxWLcen = CoxWL * Coxeff / Cox;
        const double Qac0 = CoxWLcen * (Vfbeff - vfbzb);
        double QovCox = Qac0 / Coxeff;
        const double dQac0_dVg = CoxWLcen * dVfbeff_dVg + QovCox * dCoxeff_dVg;
        const double dQac0_dVb = CoxWLcen * dVfbeff_dVb + QovCox * dCoxeff_dVb;
        T0 = 0.5 * k1ox;
        T3 = Vgs_eff - Vfbeff - VbseffCV - Vgsteffc;
        if (k1ox == 0.0) {
          T1 = 0.0;
          T2 = 0.0;
        } else if (T3 < 0.0) {
          T1 = T0 + T3 / k1ox;
          T2 = CoxWLcen;
        } else {
          T1 = sqrt(T0 * T0 + T3);
          T2 = CoxWLcen * T0 / T1;
        }
        const double Qsub0 = CoxWLcen * k1ox * (T1 - T0);
        QovCox = Qsub0 / Coxeff;
        const double dQsub0_dVg = T2 * (dVgs_eff_dVg - dVfbeff_dVg - dVgsteffc_dVg) + QovCox * dCoxeff_dVg;
        const double dQsub0_dVd = -T2 * dVgsteffc_dVd;
        const double dQsub0_dVb = -T2 * (dVfbeff_dVb + dVbseffCV_dVb + dVgsteffc_dVb) + QovCox * dCoxeff_dVb;
        double Denomi;
        const double moin = P(B3R_MOIN);
        if (k1ox <= 0.0) {
          Denomi = 0.25 * moin * Vtm;
          T0 = 0.5 * sqrtPhi;
        } else {
          Denomi = moin * Vtm * k1ox * k1ox;
          T0 = k1ox * sqrtPhi;
        }
        T1 = 2.0 * T0 + Vgsteffc;
        const double DeltaPhi = Vtm * dm_log(1.0 + T1 * Vgsteffc / Denomi);
        const double dDeltaPhi_dVg = 2.0 * Vtm * (T1 - T0) / (Denomi + T1 * Vgsteffc);
        T3 = 4.0 * (Vth - vfbzb - phi);
        Tox += Tox;
        double dT0c_dVd, dT0c_dVb2;
        if (T3 >= 0.0) {
          T0 = (Vgsteffc + T3) / Tox;
          dT0c_dVd = (dVgsteffc_dVd + 4.0 * dVth_dVd) / Tox;
          dT0c_dVb2 = (dVgsteffc_dVb + 4.0 * dVth_dVb) / Tox;
        } else {
          T0 = (Vgsteffc + 1.0e-20) / Tox;
          dT0c_dVd = dVgsteffc_dVd / Tox;
          dT0c_dVb2 = dVgsteffc_dVb / Tox;
        }
        tmp = dm_exp(0.7 * dm_log(T0));
        T1 = 1.0 + tmp;
        T2 = 0.7 * tmp / (T0 * Tox);
        Tcen = 1.9e-9 / T1;
        dTcen_dVg = -1.9e-9 * T2 / T1 / T1;
        dTcen_dVd = Tox * dTcen_dVg;
        dTcen_dVb = dTcen_dVd * dT0c_dVb2;
        dTcen_dVd *= dT0c_dVd;
        dTcen_dVg *= dVgsteffc_dVg;
        Ccen = kEpsSi / Tcen;
        T0 = Cox / (Cox + Ccen);
        Coxeff = T0 * Ccen;
        T1 = -Ccen / Tcen;
        dCoxeff_dVg = T0 * T0 * T1;
        const double dCoxeff_dVd = dCoxeff_dVg * dTcen_dVd;
        dCoxeff_dVb = dCoxeff_dVg * dTcen_dVb;
        dCoxeff_dVg *= dTcen_dVg;
        CoxWLcen = CoxWL * Coxeff / Cox;
        const double abulkCVfactor = P(B3R_ABULKCVF);
        const double AbulkCV = Abulk0 * abulkCVfactor;
        const double dAbulkCV_dVb = abulkCVfactor * dAbulk0_dVb;
        const double VdsatCV = (Vgsteffc - DeltaPhi) / AbulkCV;
        V4 = VdsatCV - Vds - kDelta4;
        T0 = sqrt(V4 * V4 + 4.0 * kDelta4 * VdsatCV);
        double VdseffCV = VdsatCV - 0.5 * (V4 + T0);
        T1 = 0.5 * (1.0 + V4 / T0);
        T2 = kDelta4 / T0;
        T3 = (1.0 - T1 - T2) / AbulkCV;
        T4 = T3 * (1.0 - dDeltaPhi_dVg);
        double dVdseffCV_dVg = T4;
        const double dVdseffCV_dVd = T1;
        double dVdseffCV_dVb = -T3 * VdsatCV * dAbulkCV_dVb;
        if (Vds == 0.0) {
          VdseffCV = 0.0;
          dVdseffCV_dVg = 0.0;
          dVdseffCV_dVb = 0.0;
        }
        T0 = AbulkCV * VdseffCV;
        T1 = Vgsteffc - DeltaPhi;
        T2 = 12.0 * (T1 - 0.5 * T0 + 1.0e-20);
        T3 = T0 / T2;
        T4 = 1.0 - 12.0 * T3 * T3;
        T5 = AbulkCV * (6.0 * T0 * (4.0 * T1 - T0) / (T2 * T2) - 0.5);
        T6 = T5 * VdseffCV / AbulkCV;
        qgate = CoxWLcen * (T1 - T0 * (0.5 - T3));
        QovCox = qgate / Coxeff;
        double Cgg1 = CoxWLcen * (T4 * (1.0 - dDeltaPhi_dVg) + T5 * dVdseffCV_dVg);
        const double Cgd1 = CoxWLcen * T5 * dVdseffCV_dVd + Cgg1 * dVgsteffc_dVd + QovCox * dCoxeff_dVd;
        const double Cgb1 = CoxWLcen * (T5 * dVdseffCV_dVb + T6 * dAbulkCV_dVb) + Cgg1 * dVgsteffc_dVb + QovCox * dCoxeff_dVb;
        Cgg1 = Cgg1 * dVgsteffc_dVg + QovCox * dCoxeff_dVg;
        T7 = 1.0 - AbulkCV;
        T8 = T2 * T2;
        T9 = 12.0 * T7 * T0 * T0 / (T8 * AbulkCV);
        T10 = T9 * (1.0 - dDeltaPhi_dVg);
        const double T11 = -T7 * T5 / AbulkCV;
        const double T12 = -(T9 * T1 / AbulkCV + VdseffCV * (0.5 - T0 / T2));
        qbulk = CoxWLcen * T7 * (0.5 * VdseffCV - T0 * VdseffCV / T2);
        QovCox = qbulk / Coxeff;
        double Cbg1 = CoxWLcen * (T10 + T11 * dVdseffCV_dVg);
        const double Cbd1 = CoxWLcen * T11 * dVdseffCV_dVd + Cbg1 * dVgsteffc_dVd + QovCox * dCoxeff_dVd;
        const double Cbb1 = CoxWLcen * (T11 * dVdseffCV_dVb + T12 * dAbulkCV_dVb) + Cbg1 * dVgsteffc_dVb + QovCox * dCoxeff_dVb;
        Cbg1 = Cbg1 * dVgsteffc_dVg + QovCox * dCoxeff_dVg;
        double Csg, Csd, Csb;
        if (xpart > 0.5) {
          qsrc = -CoxWLcen * (T1 / 2.0 + T0 / 4.0 - 0.5 * T0 * T0 / T2);
          QovCox = qsrc / Coxeff;
          T2 += T2;
          T3 = T2 * T2;
          T7 = -(0.25 - 12.0 * T0 * (4.0 * T1 - T0) / T3);
          T4 = -(0.5 + 24.0 * T0 * T0 / T3) * (1.0 - dDeltaPhi_dVg);
          T5 = T7 * AbulkCV;
          T6 = T7 * VdseffCV;
          Csg = CoxWLcen * (T4 + T5 * dVdseffCV_dVg);
          Csd = CoxWLcen * T5 * dVdseffCV_dVd + Csg * dVgsteffc_dVd + QovCox * dCoxeff_dVd;
          Csb = CoxWLcen * (T5 * dVdseffCV_dVb + T6 * dAbulkCV_dVb) + Csg * dVgsteffc_dVb + QovCox * dCoxeff_dVb;
          Csg = Csg * dVgsteffc_dVg + QovCox * dCoxeff_dVg;
        } else if (xpart < 0.5) {
          T2 = T2 / 12.0;
          T3 = 0.5 * CoxWLcen / (T2 * T2);
          T4 = T1 * (2.0 * T0 * T0 / 3.0 + T1 * (T1 - 4.0 * T0 / 3.0)) - 2.0 * T0 * T0 * T0 / 15.0;
          qsrc = -T3 * T4;
          QovCox = qsrc / Coxeff;
          T8 = 4.0 / 3.0 * T1 * (T1 - T0) + 0.4 * T0 * T0;
          T5 = -2.0 * qsrc / T2 - T3 * (T1 * (3.0 * T1 - 8.0 * T0 / 3.0) + 2.0 * T0 * T0 / 3.0);
          T6 = AbulkCV * (qsrc / T2 + T3 * T8);
          T7 = T6 * VdseffCV / AbulkCV;
          Csg = T5 * (1.0 - dDeltaPhi_dVg) + T6 * dVdseffCV_dVg;
          Csd = Csg * dVgsteffc_dVd + T6 * dVdseffCV_dVd + QovCox * dCoxeff_dVd;
          Csb = Csg * dVgsteffc_dVb + T6 * dVdseffCV_dVb + T7 * dAbulkCV_dVb + QovCox * dCoxeff_dVb;
          Csg = Csg * dVgsteffc_dVg + QovCox * dCoxeff_dVg;
        } else {
          qsrc = -0.5 * qgate;
          Csg = -0.5 * Cgg1;
          Csd = -0.5 * Cgd1;
          Csb = -0.5 * Cgb1;
        }
        qgate += Qac0 + Qsub0 - qbulk;
        qbulk -= (Qac0 + Qsub0);
        qdrn = -(qgate + qbulk + qsrc);
        const double Cbg = Cbg1 - dQac0_dVg - dQsub0_dVg;
        const double Cbd = Cbd1 - dQsub0_dVd;
        double Cbb = Cbb1 - dQac0_dVb - dQsub0_dVb;
        const double Cgg = Cgg1 - Cbg;
        const double Cgd = Cgd1 - Cbd;
        double Cgb = Cgb1 - Cbb;
        Cgb *= dVbseff_dVb;
        Cbb *= dVbseff_dVb;
        Csb *= dVbseff_dVb;
        cggb = Cgg;
        cgsb = -(Cgg + Cgd + Cgb);
        cgdb = Cgd;
        cdgb = -(Cgg + Cbg + Csg);
        cdsb = (Cgg + Cgd + Cgb + Cbg + Cbd + Cbb + Csg + Csd + Csb);
        cddb = -(Cgd + Cbd + Csd);
        cbgb = Cbg;
        cbsb = -(Cbg + Cbd + Cbb);
        cbdb = Cbd;
      }
    }
    // junctions (N_DEV_MOSFET_B3.C:6981-7134)
    double capbs, capbd;
    b3_junction(vbs, GEO(B3I_CZBS), GEO(B3I_CZBSSW), GEO(B3I_CZBSSWG), cd, qbs, capbs);
    b3_junction(vbd, GEO(B3I_CZBD), GEO(B3I_CZBDSW), GEO(B3I_CZBDSWG), cd, qbd, capbd);
    // overlaps and the companion charges (7217-7252, 7632-7800)
    const double weffCV = P(B3R_WEFFCV), ckappa = P(B3R_CKAPPA);
    double cgdo, qgdo, cgso, qgso;
    if (capmod >= 2) {
      b3_overlap(vgd, P(B3R_CGDO), P(B3R_CGDL), weffCV, ckappa, cgdo, qgdo);
      b3_overlap(vgs, P(B3R_CGSO), P(B3R_CGSL), weffCV, ckappa, cgso, qgso);
    } else {
      b3_overlap01(capmod, vgd, P(B3R_CGDO), P(B3R_CGDL), weffCV, ckappa, cgdo, qgdo);
      b3_overlap01(capmod, vgs, P(B3R_CGSO), P(B3R_CGSL), weffCV, ckappa, cgso, qgso);
    }
    const double cgbo = P(B3R_CGBO);
    const double qgb = cgbo * vgb;
    qgate += qgdo + qgso + qgb;
    qbulk -= qgb;
    if (mode > 0) {
      qdrn -= qgdo;
      qsrc = -(qgate + qbulk + qdrn);
    } else {
      qsrc = qdrn - qgso;
      qdrn = -(qgate + qbulk + qsrc);
    }
    if (mode > 0) {
      CAPcggb = (cggb + cgdo + cgso + cgbo);
      CAPcgdb = (cgdb - cgdo);
      CAPcgsb = (cgsb - cgso);
      CAPcdgb = (cdgb - cgdo);
      CAPcddb = (cddb + capbd + cgdo);
      CAPcdsb = cdsb;
      CAPcsgb = -(cggb + cbgb + cdgb + cgso);
      CAPcsdb = -(cgdb + cbdb + cddb);
      CAPcssb = (capbs + cgso - (cgsb + cbsb + cdsb));
      CAPcbgb = (cbgb - cgbo);
      CAPcbdb = (cbdb - capbd);
      CAPcbsb = (cbsb - capbs);
    } else {
      CAPcggb = (cggb + cgdo + cgso + cgbo);
      CAPcgdb = (cgsb - cgdo);
      CAPcgsb = (cgdb - cgso);
      CAPcdgb = -(cggb + cbgb + cdgb + cgdo);
      CAPcddb = (capbd + cgdo - (cgsb + cbsb + cdsb));
      CAPcdsb = -(cgdb + cbdb + cddb);
      CAPcsgb = (cdgb - cgso);
      CAPcsdb = cdsb;
      CAPcssb = (cddb + capbs + cgso);
      CAPcbgb = (cbgb - cgbo);
      CAPcbdb = (cbsb - capbd);
      CAPcbsb = (cbdb - capbs);
    }
    qg = qgate;
    qd = qdrn - qbd;
    qb = qbulk + qbd + qbs;
    (void)qsrc;
  }
  const double gdr = GEO(B3I_GD), gsr = GEO(B3I_GS), np = GEO(B3I_M);
  const double Idrain = gdr * Vddp;
  const double Isource = gsr * Vssp;

  // ---- Master::updateState
  const bool first_tran = !ss.dcop && ss.initTran && ss.newtonIter == 0;
  if constexpr (doDC) {
    p.sto_next[sto + 1] = vbs;
    p.sto_next[sto + 2] = vgs;
    p.sto_next[sto + 3] = vds;
    p.sto_next[sto + 0] = vbd;
    p.sto_next[sto + 4] = von;
  }
  if constexpr (doCV) {
    p.sta_next[sta + 0] = qb;
    p.sta_next[sta + 1] = qg;
    p.sta_next[sta + 2] = qd;
    p.sta_next[sta + 3] = qbs;
    p.sta_next[sta + 4] = qbd;
    if (first_tran) {
      p.sta_curr[sta + 0] = qb;
      p.sta_curr[sta + 1] = qg;
      p.sta_curr[sta + 2] = qd;
      p.sta_curr[sta + 3] = qbs;
      p.sta_curr[sta + 4] = qbd;
    }
  }
  // ---- Master::loadDAEVectors values
  double *out = a.slots + slot;
  double chk = 0.0;  // running sum of everything written, for the finiteness flag
  auto W = [&](int k, double v) {
    out[(size_t)k * S] = v;
    chk = chk + v;
  };
  const int JF = B3S_JF0, JQ = B3S_JQ0;
  if constexpr (doDC) {
    double Gm_, Gmbs_, FwdSum, RevSum, cdreq, ceqbd, ceqbs, gbbdp, gbbsp, gbdpg,
        gbdpdp, gbdpb, gbdpsp, gbspg, gbspdp, gbspb, gbspsp;
    const double cdv = cdrain;
    if (mode >= 0) {
      Gm_ = gm;
      Gmbs_ = gmbs;
      FwdSum = Gm_ + Gmbs_;
      RevSum = 0.0;
      cdreq = dtype * (cdv);
      ceqbd = -dtype * (csub);
      ceqbs = 0.0;
      gbbdp = -gbds;
      gbbsp = (gbds + gbgs + gbbs);
      gbdpg = gbgs;
      gbdpdp = gbds;
      gbdpb = gbbs;
      gbdpsp = -(gbdpg + gbdpdp + gbdpb);
      gbspg = 0.0;
      gbspdp = 0.0;
      gbspb = 0.0;
      gbspsp = 0.0;
    } else {
      Gm_ = -gm;
      Gmbs_ = -gmbs;
      FwdSum = 0.0;
      RevSum = -(Gm_ + Gmbs_);
      cdreq = -dtype * (cdv);
      ceqbs = -dtype * (csub);
      ceqbd = 0.0;
      gbbsp = -gbds;
      gbbdp = (gbds + gbgs + gbbs);
      gbdpg = 0.0;
      gbdpsp = 0.0;
      gbdpb = 0.0;
      gbdpdp = 0.0;
      gbspg = gbgs;
      gbspsp = gbds;
      gbspb = gbbs;
      gbspdp = -(gbspg + gbspsp + gbspb);
    }
    if (dtype > 0) {
      ceqbs += (cbs);
      ceqbd += (cbd);
    } else {
      ceqbs -= (cbs);
      ceqbd -= (cbd);
    }
    if (flags & XF_HAS_DP) W(B3S_F_D, Idrain * np);
    if (flags & XF_HAS_SP) W(B3S_F_S, Isource * np);
    W(B3S_F_B, (ceqbs + ceqbd) * np);
    W(B3S_F_DP, (-(ceqbd - cdreq) - Idrain) * np);
    W(B3S_F_SP, (-(cdreq + ceqbs) - Isource) * np);
    if (limiter && !origFlag) {
      double cdreq_J = 0.0, ceqbd_J = 0.0, ceqbs_J = 0.0, tmp;
      if (mode >= 0) {
        tmp = dtype * (-gds * (vds - vds_orig) - Gm_ * (vgs - vgs_orig) - Gmbs_ * (vbs - vbs_orig));
        cdreq_J += tmp;
        tmp = -dtype * (-gbds * (vds - vds_orig) - gbgs * (vgs - vgs_orig) - gbbs * (vbs - vbs_orig));
        ceqbd_J += tmp;
      } else {
        tmp = -dtype * (gds * (vds - vds_orig) + Gm_ * (vgd - vgd_orig) + Gmbs_ * (vbd - vbd_orig));
        cdreq_J += tmp;
        tmp = -dtype * (gbds * (vds - vds_orig) - gbgs * (vgd - vgd_orig) - gbbs * (vbd - vbd_orig));
        ceqbd_J += tmp;
      }
      if (dtype > 0) {
        ceqbs_J += (-gbs * (vbs - vbs_orig));
        ceqbd_J += (-gbd * (vbd - vbd_orig));
      } else {
        ceqbs_J -= (-gbs * (vbs - vbs_orig));
        ceqbd_J -= (-gbd * (vbd - vbd_orig));
      }
      W(B3S_FL_B, -(ceqbs_J + ceqbd_J) * np);
      W(B3S_FL_DP, (ceqbd_J - cdreq_J) * np);
      W(B3S_FL_SP, (cdreq_J + ceqbs_J) * np);
    } else {
      W(B3S_FL_B, -0.0);
      W(B3S_FL_DP, -0.0);
      W(B3S_FL_SP, -0.0);
    }
    // ---- Master::loadDAEMatrices dF/dx values ("x -= v" stored as "+ (-v)")
    if (flags & XF_HAS_DP) {
      W(JF + 0, gdr * np);
      W(JF + 1, -(gdr * np));
      W(JF + 8, -(gdr * np));
    }
    if (flags & XF_HAS_SP) {
      W(JF + 2, gsr * np);
      W(JF + 3, -(gsr * np));
      W(JF + 15, -(gsr * np));
    }
    W(JF + 4, (-gbgs) * np);
    W(JF + 5, (gbd + gbs - gbbs) * np);
    W(JF + 6, (-gbd + gbbdp) * np);
    W(JF + 7, (-gbs + gbbsp) * np);
    W(JF + 9, -((gbd - Gmbs_ - gbdpb) * np));
    W(JF + 10, (Gm_ + gbdpg) * np);
    W(JF + 11, (gdr + gds + gbd + RevSum + gbdpdp) * np);
    W(JF + 12, -((gds + FwdSum - gbdpsp) * np));
    W(JF + 13, (-Gm_ + gbspg) * np);
    W(JF + 14, -((gbs + Gmbs_ - gbspb) * np));
    W(JF + 16, -((gds + RevSum - gbspdp) * np));
    W(JF + 17, (gsr + gds + gbs + FwdSum + gbspsp) * np);
    chk = chk + vbs + vgs + vds + vbd + von;
  }
  if constexpr (doCV) {
    const double Qeqqg = dtype > 0 ? qg : -qg;
    const double Qeqqb = dtype > 0 ? qb : -qb;
    const double Qeqqd = dtype > 0 ? qd : -qd;
    W(B3S_Q_G, Qeqqg * np);
    W(B3S_Q_B, (Qeqqb)*np);
    W(B3S_Q_DP, (-(-Qeqqd)) * np);
    W(B3S_Q_SP, (-(+Qeqqg + Qeqqb + Qeqqd)) * np);
    if (limiter && !origFlag) {
      // auxChargeCalculations; with charges off a fresh instance holds zeros
      double Jg = 0.0, Jb = 0.0, Jd = 0.0;
      if (charge) {
        const double vgb_orig = vgs_orig - vbs_orig;
        const double dvgb = vgb - vgb_orig, dvbd = vbd - vbd_orig,
                     dvbs = vbs - vbs_orig;
        Jg = -CAPcggb * dvgb + CAPcgdb * dvbd + CAPcgsb * dvbs;
        Jb = -CAPcbgb * dvgb + CAPcbdb * dvbd + CAPcbsb * dvbs;
        Jd = -CAPcdgb * dvgb + CAPcddb * dvbd + CAPcdsb * dvbs;
      }
      if (dtype < 0) {
        Jg = -Jg;
        Jb = -Jb;
        Jd = -Jd;
      }
      W(B3S_QL_G, -Jg * np);
      W(B3S_QL_B, -(+Jb) * np);
      W(B3S_QL_DP, (-Jd) * np);
      W(B3S_QL_SP, (+Jg + Jb + Jd) * np);
    } else {
      W(B3S_QL_G, -0.0);
      W(B3S_QL_B, -0.0);
      W(B3S_QL_DP, -0.0);
      W(B3S_QL_SP, -0.0);
    }
    // ---- Master::loadDAEMatrices dQ/dx values
    if (first_tran) {
      for (int e = 0; e < 16; ++e) W(JQ + e, -0.0);
    } else {
      W(JQ + 0, (CAPcggb)*np);
      W(JQ + 1, -((CAPcggb + CAPcgdb + CAPcgsb) * np));
      W(JQ + 2, (CAPcgdb)*np);
      W(JQ + 3, (CAPcgsb)*np);
      W(JQ + 4, (CAPcbgb)*np);
      W(JQ + 5, (-CAPcbgb - CAPcbdb - CAPcbsb) * np);
      W(JQ + 6, (CAPcbdb)*np);
      W(JQ + 7, (CAPcbsb)*np);
      W(JQ + 8, -((+CAPcdgb + CAPcddb + CAPcdsb) * np));
      W(JQ + 9, (CAPcdgb)*np);
      W(JQ + 10, (+CAPcddb) * np);
      W(JQ + 11, -((-CAPcdsb) * np));
      W(JQ + 12, (CAPcsgb)*np);
      W(JQ + 13, -((+CAPcsgb + CAPcsdb + CAPcssb) * np));
      W(JQ + 14, -((-CAPcsdb) * np));
      W(JQ + 15, (+CAPcssb) * np);
    }
    chk = chk + qb + qg + qd + qbs + qbd;
  }
  // Bit 0: junction voltages limited (the reference's limitedFlag).  Bit 1
  // (kFull, kDC) or bit 9 (kCV, byte 1 of the word, so that the two waves of
  // a split evaluation store to different bytes): a non-finite statement,
  // state or store value (their sum is finite exactly when every term is:
  // the magnitudes are far from overflow).
  const bool bad = (chk - chk) != 0.0;
  if constexpr (kRole == kFull) {
    a.limited[slot] = (limitedFlag ? 1 : 0) | (bad ? 2 : 0);
  } else if constexpr (kRole == kDC) {
    reinterpret_cast<uint8_t *>(a.limited + slot)[0] =
        (uint8_t)((limitedFlag ? 1 : 0) | (bad ? 2 : 0));
  } else {
    reinterpret_cast<uint8_t *>(a.limited + slot)[1] = (uint8_t)(bad ? 2 : 0);
  }
}

// Many waves per CU: four waves per block, records read from HBM where the
// evaluation uses them (other waves hide the latency).
__global__ void __launch_bounds__(kBlock)
    bsim3_eval_kernel(Bsim3Dev a, SolverStateDev ss, LoadPtrs p, LinTail lt,
                      int32_t b3_blocks) {
  __shared__ double card_lds[kWavesPerBlock][B3C_N];
  if (p.gate && *p.gate) return;
  if ((int)blockIdx.x >= b3_blocks) {  // a trailing linear-device workgroup
    lin::linear_eval_one(lt.a, ss, lt.p, lt.time, lt.bp_tol, lt.dstep,
                         ((int)blockIdx.x - b3_blocks) * kBlock + threadIdx.x);
    return;
  }
  const int bid = (int)blockIdx.x;
  const int tid = threadIdx.x;
  const int wave = tid >> 6;
  const int slot = bid * kBlock + tid;
  const int S = a.n_slots;
  {
    const int nwaves = S >> 6;
    for (int k = tid; k < kWavesPerBlock * B3C_N; k += kBlock) {
      const int w = k / B3C_N, f = k % B3C_N;
      const int gw = bid * kWavesPerBlock + w;
      card_lds[w][f] = gw < nwaves ? a.cards[a.wave_card[gw] * B3C_N + f] : 0.0;
    }
  }
  __syncthreads();
  if (slot >= S) return;
  if (a.inst_of[slot] < 0) return;
  if (skip_slot(p, slot)) return;
  B3Inputs in;
  b3_inputs(a, ss, p, slot, in);
  b3_eval<false, kFull>(a, ss, p, slot, card_lds[wave], a.sdp + a.sdp_idx[slot],
                        a.geo + slot, in);
}

// Few waves per CU (small circuits, where one wave's latency is the kernel's
// time): one wave per block, and the wave's 64 instance records (size-
// dependent and junction fields, kRecN x 64 doubles, 50.7 KB) are copied to
// LDS with every load in flight at once, so the evaluation never waits on
// HBM for a parameter.
// kSplit: two waves per eval wave, wave 0 the kDC part and wave 1 the kCV
// part of the same 64 instances, sharing one LDS copy of the records (each
// wave stages half of the fields).  Both waves read their inputs --
// including the store the previous Newton iteration wrote, which the kDC
// wave overwrites -- before the workgroup barrier, and write only after it,
// so the in-place store update is race-free.
// kGroups: eval waves per workgroup, each with its own LDS copy.  Two keep
// the workgroup count within the 256 CUs up to 512 eval waves, so every wave
// of the launch has a SIMD to itself (a second workgroup on a CU put a
// kDC wave on a busy SIMD and made it half as slow again).
constexpr int kRecN = B3R_N + B3I_N;

template <bool kSplit, int kGroups>
__global__ void __launch_bounds__((kSplit ? 128 : 64) * kGroups)
    bsim3_eval_staged_kernel(Bsim3Dev a, SolverStateDev ss, LoadPtrs p, LinTail lt,
                             int32_t b3_blocks) {
  constexpr int kWaves = kSplit ? 2 : 1;     // waves per eval wave
  constexpr int kThreads = 64 * kWaves * kGroups;
  __shared__ double rec[kGroups][kRecN * 64];
  __shared__ double card[kGroups][B3C_N];
  if (p.gate && *p.gate) return;
  if ((int)blockIdx.x >= b3_blocks) {  // a trailing linear-device workgroup
    lin::linear_eval_one(lt.a, ss, lt.p, lt.time, lt.bp_tol, lt.dstep,
                         ((int)blockIdx.x - b3_blocks) * kThreads + threadIdx.x);
    return;
  }
  const int lane = threadIdx.x & 63;
  const int wave = kSplit ? (int)((threadIdx.x >> 6) % kWaves) : 0;
  const int grp = (int)threadIdx.x / (64 * kWaves);
  constexpr int kPer = (kRecN + kWaves - 1) / kWaves;
  const int w = (int)blockIdx.x * kGroups + grp;
  // the last workgroup's second group may have no eval wave: it loads
  // wave 0's fields (always valid) and evaluates nothing
  const bool valid = w < (a.n_slots >> 6);
  const int wl = valid ? w : 0;
  const int slot = wl * 64 + lane;
  const double *src = a.rec + (size_t)wl * kRecN * 64 + lane;
  double v[kPer];
#pragma unroll
  for (int j = 0; j < kPer; ++j) {
    const int f = j * kWaves + wave;
    v[j] = f < kRecN ? src[f * 64] : 0.0;
  }
  B3Inputs in;
  b3_inputs(a, ss, p, slot, in);
  const bool live = valid && a.inst_of[slot] >= 0;
  double *R0 = rec[grp];
#pragma unroll
  for (int j = 0; j < kPer; ++j) {
    const int f = j * kWaves + wave;
    if (f < kRecN) R0[f * 64 + lane] = v[j];
  }
  const int gt = (int)threadIdx.x % (64 * kWaves);
  if (gt < B3C_N) card[grp][gt] = a.cards[a.wave_card[wl] * B3C_N + gt];
  __syncthreads();
  if (!live || skip_slot(p, slot)) return;
  const double *R = R0 + lane, *Gp = R0 + B3R_N * 64 + lane;
  if constexpr (kSplit) {
    if (wave)
      b3_eval<true, kCV>(a, ss, p, slot, card[grp], R, Gp, in);
    else
      b3_eval<true, kDC>(a, ss, p, slot, card[grp], R, Gp, in);
  } else {
    b3_eval<true, kFull>(a, ss, p, slot, card[grp], R, Gp, in);
  }
}

}  // namespace

// Eval waves up to which the LDS-staged kernels are used: with two eval
// waves (101 KB of LDS) per workgroup, 512 waves fill the 256 CUs once.
#ifndef XG_B3_STAGED_MAX_WAVES
#define XG_B3_STAGED_MAX_WAVES (2 * 256)
#endif

hipError_t launch_bsim3_eval(const Bsim3Dev &d, const SolverStateDev &ss,
                             const LoadPtrs &p, hipStream_t s, const LinTail *tail) {
  if (d.n_slots == 0) return hipSuccess;
  const int waves = d.n_slots / 64;
  int v = d.variant;
  if (v == XG_EVAL_AUTO)
    v = waves <= XG_B3_STAGED_MAX_WAVES ? XG_EVAL_STAGED_SPLIT : XG_EVAL_HBM;
  // split form: two eval waves per workgroup once one per workgroup would
  // need more workgroups than CUs
  const int groups = v == XG_EVAL_STAGED_SPLIT && waves > 256 ? 2 : 1;
  const int b3_blocks = v == XG_EVAL_HBM ? (d.n_slots + kBlock - 1) / kBlock
                                         : (waves + groups - 1) / groups;
  const int block = v == XG_EVAL_STAGED_SPLIT ? 128 * groups
                                              : (v == XG_EVAL_STAGED ? 64 : kBlock);
  LinTail lt{};
  int lin_blocks = 0;
  if (tail && tail->a.n_slots > 0) {
    lt = *tail;
    lin_blocks = (tail->a.n_slots + block - 1) / block;
  }
  const dim3 grid(b3_blocks + lin_blocks);
  if (v == XG_EVAL_STAGED_SPLIT && groups == 2) {
    hipLaunchKernelGGL((bsim3_eval_staged_kernel<true, 2>), grid, dim3(256), 0, s, d, ss,
                       p, lt, b3_blocks);
  } else if (v == XG_EVAL_STAGED_SPLIT) {
    hipLaunchKernelGGL((bsim3_eval_staged_kernel<true, 1>), grid, dim3(128), 0, s, d, ss,
                       p, lt, b3_blocks);
  } else if (v == XG_EVAL_STAGED) {
    hipLaunchKernelGGL((bsim3_eval_staged_kernel<false, 1>), grid, dim3(64), 0, s, d, ss,
                       p, lt, b3_blocks);
  } else {
    hipLaunchKernelGGL(bsim3_eval_kernel, grid, dim3(kBlock), 0, s, d, ss, p, lt,
                       b3_blocks);
  }
  return hipGetLastError();
}

}  // namespace xg
