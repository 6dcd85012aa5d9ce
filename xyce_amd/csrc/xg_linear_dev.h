// Linear device load for gfx950: resistors, capacitors, junction diodes
// and independent voltage sources (DC or PULSE), one thread per device,
// devices grouped by type in whole waves.  Each thread computes the values
// of its device's Master-loop statements and writes them to the slot
// buffer, where the ordered reduction folds them after the MOSFETs' (device
// types load in the order MOSFET1, BSIM3, resistor, capacitor, diode,
// source).  Device code shared by the standalone linear eval kernel and the
// BSIM3 eval kernels, which can run the linear devices in trailing
// workgroups of their own launch.
//
//   resistor   N_DEV_Resistor.C:1073-1174   i0 = G (v+ - v-)
//   capacitor  N_DEV_Capacitor.C:1821-2060  q0 = C (v+ - v-), state q0
//   diode      N_DEV_Diode.C:873-1316, 2007-2108 (level 1, RS = 0)
//   source     N_DEV_Vsrc.C:941-1036        branch current, v+ - v- - V(t)
//   PULSE      N_DEV_SourceData.C:928-1000
#pragma once
#include <cfloat>

#include "xg_devmath.h"
#include "xg_internal.h"

namespace xg {
namespace lin {

// PulseData::updateSource: V(t) of PULSE(V1 V2 TD TR TF PW PER).
__device__ __forceinline__ double pulse_value(const double *pl, size_t st,
                                              double t, double bp_tol) {
  const double V1 = pl[1 * st], V2 = pl[2 * st], TD = pl[3 * st],
               TR = pl[4 * st], TF = pl[5 * st], PW = pl[6 * st],
               PER = pl[7 * st];
  double time = t - TD;
  if (time > PER && PER != 0.0) time -= PER * floor(time / PER);
  if (time <= 0 || (time > (TR + PW + TF) && fabs(time - (TR + PW + TF)) > bp_tol))
    return V1;
  if ((time > TR && fabs(time - TR) > bp_tol) &&
      (time < (TR + PW) || fabs(time - (TR + PW)) < bp_tol))
    return V2;
  if (time > 0 && (time < TR || fabs(time - TR) < bp_tol))
    return TR != 0.0 ? V1 + (V2 - V1) * (time) / TR : V1;
  return TF != 0.0 ? V2 + (V1 - V2) * (time - (TR + PW)) / TF : V2;
}

// One diode: Master::updateState (the store gets the junction voltage of
// the previous evaluation, which the instance keeps), updateIntermediateVars
// (initial junction voltage, pnjlim, exponential / linear-reverse /
// breakdown current, depletion and diffusion charge), then the statement
// values of loadDAEVectors and loadDAEMatrices.
__device__ __forceinline__ void diode_eval(const LinArgs &a, const SolverStateDev &ss,
                                           const LoadPtrs &p, int s, int S, double vpp,
                                           double vn, double *out) {
  auto F = [&](int k) { return a.dpar[(size_t)k * S + s]; };
  auto W = [&](int k, double v) { out[(size_t)k * S] = v; };
  const int sto = a.sta[s];
  const double gmin = a.gmin;
  // Master::updateState: stoVec[li_storevd] = Vd (the last evaluation's)
  p.sto_next[sto] = a.dvd[s];
  const double Vte = F(DIO_VTE), vcrit = F(DIO_VCRIT), BV = F(DIO_BV);
  const double brk = F(DIO_BRKV);
  double Vd = vpp - vn;
  const double Vd_orig = Vd;
  bool orig = true;
  double Vd_old;
  if (ss.newtonIter == 0) {
    if (ss.initJct && a.limiter) {
      if (F(DIO_HASIC) != 0.0) {
        Vd = F(DIO_IC);
      } else if (F(DIO_OFF) != 0.0) {
        Vd = 0.0;
      } else {
        Vd = vcrit;
      }
      orig = false;
    }
    Vd_old = Vd;
    if (!ss.dcop) Vd_old = p.sto_curr[sto];
  } else {
    Vd_old = p.sto_next[sto];
  }
  if (a.limiter) {
    int ichk = 0;
    if (F(DIO_BVGIVEN) != 0.0 && (Vd < fmin(0.0, -BV + 10.0 * Vte))) {
      double Vdtmp = -(BV + Vd);
      Vdtmp = dm_pnjlim(Vdtmp, -(Vd_old + BV), Vte, vcrit, ichk);
      Vd = -(Vdtmp + BV);
    } else {
      Vd = dm_pnjlim(Vd, Vd_old, Vte, vcrit, ichk);
    }
    if (ichk) orig = false;
  }
  // current and conductance (normal exponential form)
  double Isat = F(DIO_ISAT);
  const double IRfactor = Vd >= -3.0 * Vte ? 1.0 : F(DIO_IRF);
  Isat *= IRfactor;
  double Id, Gd;
  if (Vd >= -3.0 * Vte) {
    double arg1 = Vd / Vte;
    arg1 = fmin(100.0, arg1);
    const double evd = dm_exp(arg1);
    Id = Isat * (evd - 1.0) + gmin * Vd;
    Gd = Isat * evd / Vte + gmin;
  } else if (!(brk != 0.0) || (Vd >= -brk)) {
    double arg = 3.0 * Vte / (Vd * dm_exp(1.0));
    arg = arg * arg * arg;
    Id = -Isat * (1.0 + arg) + gmin * Vd;
    Gd = Isat * 3.0 * arg / Vd + gmin;
  } else {
    double arg1 = -(brk + Vd) / Vte;
    arg1 = fmin(100.0, arg1);
    const double evrev = dm_exp(arg1);
    double arg2 = 3.0 * Vte / (dm_exp(1.0) * brk);
    arg2 = arg2 * arg2 * arg2;
    const double Isat_b = Isat * (1 - arg2);
    Id = -Isat_b * evrev + gmin * Vd;
    Gd = Isat_b * evrev / Vte + gmin;
  }
  const double Vc = Vd;
  double Qd = 0.0, Cd = 0.0;
  if (F(DIO_TJCAP) != 0.0) {
    const double Czero = F(DIO_CZERO), pot = F(DIO_TJPOT), dep = F(DIO_TDEPCAP);
    const double M = F(DIO_M), TT = F(DIO_TT);
    if (Vc < dep) {
      const double arg = 1.0 - Vc / pot;
      double arg1 = -M * dm_log(arg);
      arg1 = fmin(100.0, arg1);
      const double sarg = dm_exp(arg1);
      Qd = TT * Id + pot * Czero * (1.0 - arg * sarg) / (1.0 - M);
      Cd = TT * Gd + Czero * sarg;
    } else {
      const double Czof2 = Czero / F(DIO_F2);
      const double MotJctPot = M / pot;
      Qd = TT * Id + Czero * F(DIO_TF1) +
           Czof2 * (F(DIO_F3) * (Vc - dep) + (0.5 * MotJctPot) * (Vc * Vc - dep * dep));
      Cd = TT * Gd + Czof2 * (F(DIO_F3) + MotJctPot * Vc);
    }
  }
  a.dvd[s] = Vd;
  const double Gspr = F(DIO_GSPR);
  // with RS = 0 the internal node is the positive one: Vp - Vpp = 0
  const double Ir = Gspr * (vpp - vpp);
  // loadDAEVectors: x -= v is stored as +(-v)
  W(0, -(-Ir));
  W(1, -Id);
  W(2, -(-Id + Ir));
  W(3, -Qd);
  W(4, -(-Qd));
  if (a.limiter) {
    const double Vd_diff = Vd - Vd_orig;
    const double Cd_Jdxp = -(Cd)*Vd_diff;
    const double Gd_Jdxp = -(Gd)*Vd_diff;
    W(5, Gd_Jdxp);
    W(6, -Gd_Jdxp);
    W(7, Cd_Jdxp);
    W(8, -Cd_Jdxp);
  }
  // loadDAEMatrices
  W(9, Gspr);
  W(10, -Gspr);
  W(11, Gd);
  W(12, -Gd);
  W(13, -Gspr);
  W(14, -Gd);
  W(15, Gspr + Gd);
  W(16, Cd);
  W(17, -Cd);
  W(18, -Cd);
  W(19, Cd);
  const bool bad = !(fabs(Id) <= DBL_MAX && fabs(Gd) <= DBL_MAX && fabs(Qd) <= DBL_MAX &&
                     fabs(Cd) <= DBL_MAX);
  a.limited[s] = (orig ? 0 : 1) | (bad ? 2 : 0);
}

// One linear-device eval slot s (time / bp_tol from dstep when given).
__device__ __forceinline__ void linear_eval_one(const LinArgs &a, const SolverStateDev &ss,
                                                const LoadPtrs &p, double time,
                                                double bp_tol, const double *dstep,
                                                int s) {
  if (dstep) {  // time and breakpoint tolerance from device memory
    time = dstep[0];
    bp_tol = dstep[1];
  }
  const int S = a.n_slots;
  if (s >= S) return;
  const int32_t type = a.type[s];
  if (type < 0 || skip_slot(p, s)) return;
  const double vp = p.x[a.lid[s]], vn = p.x[a.lid[S + s]];
  const double val = a.val[s];
  double *out = a.slots + s;
  auto W = [&](int k, double v) { out[(size_t)k * S] = v; };
  if (type == LIN_R) {
    const double i0 = (vp - vn) * val;
    W(0, i0);
    W(1, -i0);
    W(2, val);
    W(3, -val);
    W(4, -val);
    W(5, val);
  } else if (type == LIN_C) {
    const double q0 = val * (vp - vn);
    p.sta_next[a.sta[s]] = q0;
    W(0, q0);
    W(1, -q0);
    W(2, val);
    W(3, -val);
    W(4, -val);
    W(5, val);
  } else if (type == LIN_D) {
    diode_eval(a, ss, p, s, S, vp, vn, out);
  } else {
    const double ibr = p.x[a.lid[2 * S + s]];
    const bool use_pulse = (ss.tranop || ss.transient) && a.pulse[s] != 0.0;
    const double src = use_pulse ? pulse_value(a.pulse + s, (size_t)S, time, bp_tol)
                                 : val;
    W(0, ibr);
    W(1, -ibr);
    W(2, vp - vn);
    W(3, -src);
    W(4, 1.0);
    W(5, -1.0);
    W(6, 1.0);
    W(7, -1.0);
  }
}

}  // namespace lin
}  // namespace xg
