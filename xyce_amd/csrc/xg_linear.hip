// Linear device load for gfx950: resistors, capacitors and independent
// voltage sources (DC or PULSE), one thread per device, devices grouped by
// type in whole waves.  Each thread computes the values of its device's
// Master-loop statements and writes them to the slot buffer, where the
// ordered reduction folds them after the MOSFETs' (device types load in the
// order MOSFET1, BSIM3, resistor, capacitor, source).
//
//   resistor   N_DEV_Resistor.C:1073-1174   i0 = G (v+ - v-)
//   capacitor  N_DEV_Capacitor.C:1821-2060  q0 = C (v+ - v-), state q0
//   source     N_DEV_Vsrc.C:941-1036        branch current, v+ - v- - V(t)
//   PULSE      N_DEV_SourceData.C:928-1000
#include "xg_internal.h"

namespace xg {

namespace {

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

__global__ void __launch_bounds__(256)
    linear_eval_kernel(LinArgs a, SolverStateDev ss, LoadPtrs p, double time,
                       double bp_tol, const double *dstep) {
  const int s = blockIdx.x * 256 + threadIdx.x;
  if (dstep) {  // time and breakpoint tolerance from device memory
    time = dstep[0];
    bp_tol = dstep[1];
  }
  const int S = a.n_slots;
  if (s >= S) return;
  const int32_t type = a.type[s];
  if (type < 0) return;
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

}  // namespace

hipError_t launch_linear_eval(const LinArgs &a, const SolverStateDev &ss,
                              const LoadPtrs &p, double time, double bp_tol,
                              const double *dstep, hipStream_t s) {
  if (a.n_slots == 0) return hipSuccess;
  hipLaunchKernelGGL(linear_eval_kernel, dim3((a.n_slots + 255) / 256),
                     dim3(256), 0, s, a, ss, p, time, bp_tol, dstep);
  return hipGetLastError();
}

}  // namespace xg
