// Device-side helpers shared by the MOSFET1 and BSIM3 kernels: the SPICE3
// junction/FET voltage limiters and thin wrappers over the FP64 libm.
//
// The limiters are pure functions of their arguments and must branch exactly
// as DeviceSupport::limvds / pnjlim / fetlim do (N_DEV_DeviceSupport.C:
// 186-203, 214-251, 337-421), since the branch taken decides the Newton path.
#pragma once

#include <hip/hip_runtime.h>

namespace xg {

__device__ __forceinline__ double dm_exp(double x) { return exp(x); }
__device__ __forceinline__ double dm_log(double x) { return log(x); }

__device__ __forceinline__ double dm_limvds(double vnew, double vold) {
  if (vold >= 3.5) {
    if (vnew > vold) vnew = fmin(vnew, (3.0 * vold) + 2.0);
    else if (vnew < 3.5) vnew = fmax(vnew, 2.0);
  } else {
    if (vnew > vold) vnew = fmin(vnew, 4.0);
    else vnew = fmax(vnew, -0.5);
  }
  return vnew;
}

__device__ __forceinline__ double dm_pnjlim(double vnew, double vold,
                                            double vt, double vcrit,
                                            int &icheck) {
  if ((vnew > vcrit) && (fabs(vnew - vold) > (vt + vt))) {
    if (vold > 0) {
      const double arg = 1 + (vnew - vold) / vt;
      vnew = arg > 0 ? vold + vt * dm_log(arg) : vcrit;
    } else {
      vnew = vt * dm_log(vnew / vt);
    }
    icheck = 1;
  } else {
    icheck = 0;
  }
  return vnew;
}

__device__ __forceinline__ double dm_fetlim(double vnew, double vold,
                                            double vto) {
  const double vtsthi = fabs(2 * (vold - vto)) + 2;
  const double vtstlo = vtsthi / 2 + 2;
  const double vtox = vto + 3.5;
  const double delv = vnew - vold;
  if (vold >= vto) {
    if (vold >= vtox) {
      if (delv <= 0) {
        if (vnew >= vtox) {
          if (-delv > vtstlo) vnew = vold - vtstlo;
        } else {
          vnew = fmax(vnew, vto + 2.0);
        }
      } else if (delv >= vtsthi) {
        vnew = vold + vtsthi;
      }
    } else {
      vnew = delv <= 0 ? fmax(vnew, vto - 0.5) : fmin(vnew, vto + 4.0);
    }
  } else {
    if (delv <= 0) {
      if (-delv > vtsthi) vnew = vold - vtsthi;
    } else {
      const double vtemp = vto + 0.5;
      if (vnew <= vtemp) {
        if (delv > vtstlo) vnew = vold + vtstlo;
      } else {
        vnew = vtemp;
      }
    }
  }
  return vnew;
}

}  // namespace xg
