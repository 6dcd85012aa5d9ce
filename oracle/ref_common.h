/*
 * ORACLE — test infrastructure only.  Nothing in the product links this.
 *
 * Shared pieces of the CPU restatement of Xyce's device-load path:
 * solver-state flags, device options, physical constants and the SPICE3
 * limiters / Meyer capacitance helpers.  Each item cites the reference
 * file:line it restates.  Plain C, FP64 throughout, same operation order as
 * the reference so that results can be compared to the last few ulps.
 */
#ifndef XG_ORACLE_REF_COMMON_H
#define XG_ORACLE_REF_COMMON_H

#include <math.h>

/* Physical constants: src/DeviceModelPKG/Core/include/N_DEV_Const.h:50-106 */
#define RC_ROOT2      sqrt(2.0)
#define RC_Q          (1.6021918e-19)
#define RC_BOLTZ      (1.3806226e-23)
#define RC_REFTEMP    (300.15)
#define RC_KoverQ     (RC_BOLTZ / RC_Q)
#define RC_MAX_EXP_ARG 100.0
#define RC_EXP_THRESHOLD 34.0
#define RC_MAX_EXP    5.834617425e14
#define RC_MIN_EXP    1.713908431e-15
#define RC_EPSOX      (3.453133e-11)
#define RC_EPSSI      (1.03594e-10)
#define RC_DELTA_1    0.02
#define RC_DELTA_3    0.02
#define RC_DELTA_4    0.02

static inline double rc_min(double a, double b) { return a < b ? a : b; }
static inline double rc_max(double a, double b) { return a > b ? a : b; }

/* Subset of N_DEV_SolverState.h:75-200 that selects branches in the loads. */
typedef struct ref_solver_state {
  int dcopFlag;
  int initJctFlag;
  int initFixFlag;
  int initTranFlag;
  int newtonIter;
  int tranopFlag;
  int acopFlag;
  int transientFlag;
  int inputOPFlag;
  int locaEnabledFlag;
} ref_solver_state;

/* Subset of N_DEV_DeviceOptions.C:82-147 (defaults gmin=1e-12, limiter on). */
typedef struct ref_dev_options {
  double gmin;
  int voltageLimiterFlag;
} ref_dev_options;

/* DeviceSupport::limvds, N_DEV_DeviceSupport.C:186-203 */
static inline double ref_limvds(double vnew, double vold) {
  if (vold >= 3.5) {
    if (vnew > vold) vnew = rc_min(vnew, (3.0 * vold) + 2.0);
    else if (vnew < 3.5) vnew = rc_max(vnew, 2.0);
  } else {
    if (vnew > vold) vnew = rc_min(vnew, 4.0);
    else vnew = rc_max(vnew, -0.5);
  }
  return vnew;
}

/* DeviceSupport::pnjlim, N_DEV_DeviceSupport.C:214-251 */
static inline double ref_pnjlim(double vnew, double vold, double vt,
                                double vcrit, int *icheck) {
  double arg;
  if ((vnew > vcrit) && (fabs(vnew - vold) > (vt + vt))) {
    if (vold > 0) {
      arg = 1 + (vnew - vold) / vt;
      if (arg > 0) vnew = vold + vt * log(arg);
      else vnew = vcrit;
    } else {
      vnew = vt * log(vnew / vt);
    }
    *icheck = 1;
  } else {
    *icheck = 0;
  }
  return vnew;
}

/* DeviceSupport::fetlim, N_DEV_DeviceSupport.C:337-421 */
static inline double ref_fetlim(double vnew, double vold, double vto) {
  double vtsthi = fabs(2 * (vold - vto)) + 2;
  double vtstlo = vtsthi / 2 + 2;
  double vtox = vto + 3.5;
  double delv = vnew - vold;
  double vtemp;
  if (vold >= vto) {
    if (vold >= vtox) {
      if (delv <= 0) {
        if (vnew >= vtox) {
          if (-delv > vtstlo) vnew = vold - vtstlo;
        } else {
          vnew = rc_max(vnew, vto + 2.0);
        }
      } else {
        if (delv >= vtsthi) vnew = vold + vtsthi;
      }
    } else {
      if (delv <= 0) vnew = rc_max(vnew, vto - 0.5);
      else vnew = rc_min(vnew, vto + 4.0);
    }
  } else {
    if (delv <= 0) {
      if (-delv > vtsthi) vnew = vold - vtsthi;
    } else {
      vtemp = vto + 0.5;
      if (vnew <= vtemp) {
        if (delv > vtstlo) vnew = vold + vtstlo;
      } else {
        vnew = vtemp;
      }
    }
  }
  return vnew;
}

/* DeviceSupport::qmeyer, N_DEV_DeviceSupport.C:545-614 */
static inline void ref_qmeyer(double vgs, double vgd, double vgb, double von,
                              double vdsat, double *capgs, double *capgd,
                              double *capgb, double phi, double cox) {
  double vds, vddif, vddif1, vddif2, vgst;
  (void)vgb;
  vgst = vgs - von;
  if (vgst <= -phi) {
    *capgb = cox / 2;
    *capgs = 0;
    *capgd = 0;
  } else if (vgst <= -phi / 2) {
    *capgb = -vgst * cox / (2 * phi);
    *capgs = 0;
    *capgd = 0;
  } else if (vgst <= 0) {
    *capgb = -vgst * cox / (2 * phi);
    *capgs = vgst * cox / (1.5 * phi) + cox / 3;
    *capgd = 0;
  } else {
    vds = vgs - vgd;
    if (vdsat <= vds) {
      *capgs = cox / 3;
      *capgd = 0;
      *capgb = 0;
    } else {
      vddif = 2.0 * vdsat - vds;
      vddif1 = vdsat - vds;
      vddif2 = vddif * vddif;
      *capgd = cox * (1.0 - vdsat * vdsat / vddif2) / 3;
      *capgs = cox * (1.0 - vddif1 * vddif1 / vddif2) / 3;
      *capgb = 0;
    }
  }
}

#endif
