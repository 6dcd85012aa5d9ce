// MOSFET level-1 fused load kernel for gfx950.
//
// One thread per instance.  Instances are grouped by card (model at a
// temperature) and each group is padded to whole 64-lane wavefronts, so a
// wave reads exactly one card; the block's cards are staged in LDS.  The
// thread evaluates the reference's updateState and computes every value its
// loadDAEVectors and loadDAEMatrices statements would add, writing them to
// the slot buffer (one array per statement, coalesced).  The ordered
// reduction kernel then folds the slots into F/Q/dFdxdVp/dQdxdVp and the CSR
// values in the reference's order.
//
// Physics follows, statement for statement (FP64, no FMA contraction):
//   Instance::updateIntermediateVars  N_DEV_MOSFET1.C:2618-3427
//   Master::updateState               N_DEV_MOSFET1.C:4103-4192
//   Master::loadDAEVectors            N_DEV_MOSFET1.C:4202-4430
//   Master::loadDAEMatrices           N_DEV_MOSFET1.C:4441-4722
//   DeviceSupport fetlim/limvds/pnjlim/qmeyer N_DEV_DeviceSupport.C:186-614
// with newMeyerFlag=false and no homotopy.
#include "xg_internal.h"
#include "xg_devmath.h"

namespace xg {

namespace {

constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / 64;

struct MeyerCaps {
  double gs, gd, gb;
};

// DeviceSupport::qmeyer with (vgs, vgd) as given; the caller swaps the
// roles of source and drain for mode < 0 (N_DEV_MOSFET1.C:3187-3228).
__device__ __forceinline__ MeyerCaps qmeyer(double vgs, double vgd, double von,
                                            double vdsat, double phi,
                                            double cox) {
  double capgs, capgd, capgb;
  const double vgst = vgs - von;
  if (vgst <= -phi) {
    capgb = cox / 2;
    capgs = 0;
    capgd = 0;
  } else if (vgst <= -phi / 2) {
    capgb = -vgst * cox / (2 * phi);
    capgs = 0;
    capgd = 0;
  } else if (vgst <= 0) {
    capgb = -vgst * cox / (2 * phi);
    capgs = vgst * cox / (1.5 * phi) + cox / 3;
    capgd = 0;
  } else {
    const double vds = vgs - vgd;
    if (vdsat <= vds) {
      capgs = cox / 3;
      capgd = 0;
      capgb = 0;
    } else {
      const double vddif = 2.0 * vdsat - vds;
      const double vddif1 = vdsat - vds;
      const double vddif2 = vddif * vddif;
      capgd = cox * (1.0 - vdsat * vdsat / vddif2) / 3;
      capgs = cox * (1.0 - vddif1 * vddif1 / vddif2) / 3;
      capgb = 0;
    }
  }
  return MeyerCaps{capgs, capgd, capgb};
}

// Depletion charge and capacitance of one bulk junction.
__device__ __forceinline__ void junction_charge(
    double v, double Cb, double Cbsw, double f2, double f3, double f4,
    double tDepCap, double tBulkPot, double mj, double mjsw, bool mj_half,
    bool mjsw_half, bool mj_eq_mjsw, bool source_side, double &q,
    double &cap) {
  if (Cb != 0 || Cbsw != 0) {
    if (v < tDepCap) {
      const double arg = 1 - v / tBulkPot;
      double sarg, sargsw;
      // The source side tests MJ == MJSW first; the drain side tests both
      // equal to 0.5 (N_DEV_MOSFET1.C:3062-3091 vs 3117-3142).
      if (source_side && mj_eq_mjsw) {
        if (mj_half) sarg = sargsw = 1 / sqrt(arg);
        else sarg = sargsw = dm_exp(-mj * dm_log(arg));
      } else if (!source_side && mj_half && mjsw_half) {
        sarg = sargsw = 1 / sqrt(arg);
      } else {
        sarg = mj_half ? 1 / sqrt(arg) : dm_exp(-mj * dm_log(arg));
        sargsw = mjsw_half ? 1 / sqrt(arg) : dm_exp(-mjsw * dm_log(arg));
      }
      q = tBulkPot * (Cb * (1 - arg * sarg) / (1 - mj) +
                      Cbsw * (1 - arg * sargsw) / (1 - mjsw));
      cap = Cb * sarg + Cbsw * sargsw;
    } else {
      if (source_side) q = f4 + v * (f2 + v * (f3 / 2));
      else q = f4 + v * (f2 + v * f3 / 2);
      cap = source_side ? f2 + f3 * v : f2 + v * f3;
    }
  } else {
    q = 0;
    cap = 0;
  }
}

__global__ void __launch_bounds__(kBlock)
    mos1_eval_kernel(Mos1Args a, SolverStateDev ss, LoadPtrs p) {
  __shared__ double card_lds[kWavesPerBlock][M1C_N];
  if (p.gate && *p.gate) return;
  const int tid = threadIdx.x;
  const int wave = tid >> 6;
  const int slot = blockIdx.x * kBlock + tid;
  const int S = a.n_slots;
  {
    const int nwaves = S >> 6;
    for (int k = tid; k < kWavesPerBlock * M1C_N; k += kBlock) {
      const int w = k / M1C_N, f = k % M1C_N;
      const int gw = blockIdx.x * kWavesPerBlock + w;
      card_lds[w][f] = gw < nwaves ? a.cards[a.wave_card[gw] * M1C_N + f] : 0.0;
    }
  }
  __syncthreads();
  if (slot >= S) return;
  if (a.inst_of[slot] < 0) return;
  if (skip_slot(p, slot)) return;
  const double *cd = card_lds[wave];
  const double dtype = cd[M1C_DTYPE];
  const double gamma = cd[M1C_GAMMA];
  const double lambda = cd[M1C_LAMBDA];
  const double tPhi = cd[M1C_TPHI];
  const double tVbi = cd[M1C_TVBI];
  const double vt = cd[M1C_VT];
  const double gmin = cd[M1C_GMIN];
  const bool limiter = cd[M1C_LIMITER] != 0.0;

  auto F = [&](int f) { return a.fields[(size_t)f * S + slot]; };
  const int32_t flags = a.flags[slot];
  const bool OFF = flags & XF_OFF;
  const int ld = a.lid4[slot], lg = a.lid4[S + slot], ls = a.lid4[2 * S + slot],
            lb = a.lid4[3 * S + slot];
  const int ldp = (flags & XF_HAS_DP) ? a.lid_int[slot] : ld;
  const int lsp = (flags & XF_HAS_SP) ? a.lid_int[S + slot] : ls;
  const int sta = a.sta_base[slot], sto = a.sto_base[slot];

  // ---- updateIntermediateVars
  const double DrainSatCur = F(M1F_DSATCUR), SourceSatCur = F(M1F_SSATCUR);
  const double Beta = F(M1F_BETA);
  const double Vd = p.x[ld], Vg = p.x[lg], Vs = p.x[ls], Vb = p.x[lb];
  const double Vsp = p.x[lsp], Vdp = p.x[ldp];
  const double Vddp = Vd - Vdp, Vssp = Vs - Vsp, Vbsp = Vb - Vsp;
  const double Vgsp = Vg - Vsp, Vdpsp = Vdp - Vsp;
  double vbs = dtype * Vbsp, vgs = dtype * Vgsp, vds = dtype * Vdpsp;
  double vbd = vbs - vds, vgd = vgs - vds;
  bool origFlag = true;
  bool limitedFlag = false;
  const double vgs_orig = vgs, vds_orig = vds, vbs_orig = vbs,
               vbd_orig = vbd, vgd_orig = vgd;
  if (ss.initJct && !OFF && limiter) {
    if (flags & XF_IC) {
      vds = dtype * F(M1F_ICVDS);
      vgs = dtype * F(M1F_ICVGS);
      vbs = dtype * F(M1F_ICVBS);
      vbd = vbs - vds;
      vgd = vgs - vds;
      origFlag = false;
    } else {
      vbs = -1;
      vgs = dtype * cd[M1C_TVTO];
      vds = 0;
      vbd = vbs - vds;
      vgd = vgs - vds;
    }
  } else if ((ss.initFix || ss.initJct) && OFF) {
    vbs = vgs = vds = 0;
    vbd = vgd = 0;
  }
  double vbs_old, vbd_old, vgs_old, vds_old, Von;
  if (ss.newtonIter == 0) {
    if (!ss.dcop) {
      vbs_old = p.sto_curr[sto + 1];
      vbd_old = p.sto_curr[sto + 0];
      vgs_old = p.sto_curr[sto + 2];
      vds_old = p.sto_curr[sto + 3];
      Von = dtype * p.sto_curr[sto + 4];
    } else {
      vbs_old = vbs;
      vbd_old = vbd;
      vgs_old = vgs;
      vds_old = vds;
      Von = 0.0;
    }
  } else {
    vbs_old = p.sto_next[sto + 1];
    vbd_old = p.sto_next[sto + 0];
    vgs_old = p.sto_next[sto + 2];
    vds_old = p.sto_next[sto + 3];
    Von = dtype * p.sto_next[sto + 4];
  }
  const double vgd_old = vgs_old - vds_old;
  if (limiter && !(ss.initFix && OFF)) {
    int Check = 1;
    if (vds_old >= 0) {
      vgs = dm_fetlim(vgs, vgs_old, Von);
      vds = vgs - vgd;
      vds = dm_limvds(vds, vds_old);
      vgd = vgs - vds;
    } else {
      vgd = dm_fetlim(vgd, vgd_old, Von);
      vds = vgs - vgd;
      vds = -dm_limvds(-vds, -vds_old);
      vgs = vgd + vds;
    }
    if (vds >= 0.0) {
      vbs = dm_pnjlim(vbs, vbs_old, vt, F(M1F_SVCRIT), Check);
      vbd = vbs - vds;
    } else {
      vbd = dm_pnjlim(vbd, vbd_old, vt, F(M1F_DVCRIT), Check);
      vbs = vbd + vds;
    }
    if (Check == 1) limitedFlag = true;
  }
  vbd = vbs - vds;
  vgd = vgs - vds;
  const double Vgb = vgs - vbs;
  if (vgs_orig != vgs || vds_orig != vds || vbs_orig != vbs ||
      vbd_orig != vbd || vgd_orig != vgd)
    origFlag = false;

  double gbs, cbs, gbd, cbd;
  if (vbs <= 0) {
    gbs = SourceSatCur / vt;
    gbs += gmin;
    cbs = gbs * vbs;
  } else {
    const double evbs = dm_exp(fmin(100.0, vbs / vt));
    gbs = (SourceSatCur * evbs / vt + gmin);
    cbs = (SourceSatCur * (evbs - 1) + gmin * vbs);
  }
  if (vbd <= 0) {
    gbd = DrainSatCur / vt;
    gbd += gmin;
    cbd = gbd * vbd;
  } else {
    const double evbd = dm_exp(fmin(100.0, vbd / vt));
    gbd = (DrainSatCur * evbd / vt + gmin);
    cbd = (DrainSatCur * (evbd - 1) + gmin * vbd);
  }
  const int mode = (vds >= 0) ? 1 : -1;
  double cdrain, gm, gds, gmbs, Vdsat;
  {
    const double vbx = mode == 1 ? vbs : vbd;
    double sarg;
    if (vbx <= 0) {
      sarg = sqrt(tPhi - vbx);
    } else {
      sarg = sqrt(tPhi);
      sarg = sarg - vbx / (sarg + sarg);
      sarg = fmax(0.0, sarg);
    }
    Von = (tVbi * dtype) + gamma * sarg;
    const double vgst = (mode == 1 ? vgs : vgd) - Von;
    Vdsat = fmax(vgst, 0.0);
    const double arg = sarg <= 0 ? 0 : gamma / (sarg + sarg);
    if (vgst <= 0) {
      cdrain = 0;
      gm = 0;
      gds = 0;
      gmbs = 0;
    } else {
      const double vdsm = vds * mode;
      const double betap = Beta * (1 + lambda * vdsm);
      if (vgst <= vdsm) {
        cdrain = betap * vgst * vgst * .5;
        gm = betap * vgst;
        gds = lambda * Beta * vgst * vgst * .5;
        gmbs = gm * arg;
      } else {
        cdrain = betap * vdsm * (vgst - .5 * vdsm);
        gm = betap * vdsm;
        gds = betap * (vgst - vdsm) + lambda * Beta * vdsm * (vgst - .5 * vdsm);
        gmbs = gm * arg;
      }
    }
  }
  const double von = dtype * Von;

  const bool mj_half = cd[M1C_MJ_EQ_HALF] != 0.0,
             mjsw_half = cd[M1C_MJSW_EQ_HALF] != 0.0,
             mj_eq = cd[M1C_MJ_EQ_MJSW] != 0.0;
  double qbs, capbs, qbd, capbd;
  junction_charge(vbs, F(M1F_CBS), F(M1F_CBSSW), F(M1F_F2S), F(M1F_F3S),
                  F(M1F_F4S), cd[M1C_TDEPCAP], cd[M1C_TBULKPOT], cd[M1C_MJ],
                  cd[M1C_MJSW], mj_half, mjsw_half, mj_eq, true, qbs, capbs);
  junction_charge(vbd, F(M1F_CBD), F(M1F_CBDSW), F(M1F_F2D), F(M1F_F3D),
                  F(M1F_F4D), cd[M1C_TDEPCAP], cd[M1C_TBULKPOT], cd[M1C_MJ],
                  cd[M1C_MJSW], mj_half, mjsw_half, mj_eq, false, qbd, capbd);

  const double OxideCap = F(M1F_OXIDECAP);
  const MeyerCaps mc = mode > 0 ? qmeyer(vgs, vgd, Von, Vdsat, tPhi, OxideCap)
                                : qmeyer(vgd, vgs, Von, Vdsat, tPhi, OxideCap);
  const double capgs = mode > 0 ? mc.gs : mc.gd;
  const double capgd = mode > 0 ? mc.gd : mc.gs;
  const double capgb = mc.gb;

  double Capgs, Capgd, Capgb;
  if (ss.dcop) {
    Capgs = 2.0 * capgs + F(M1F_GSO);
    Capgd = 2.0 * capgd + F(M1F_GDO);
    Capgb = 2.0 * capgb + F(M1F_GBO);
  } else {
    Capgs = (capgs + p.sta_curr[sta + 3] + F(M1F_GSO));
    Capgd = (capgd + p.sta_curr[sta + 4] + F(M1F_GDO));
    Capgb = (capgb + p.sta_curr[sta + 5] + F(M1F_GBO));
  }
  Capgs *= (Capgs < 0.0) ? -1.0 : 1.0;
  Capgd *= (Capgd < 0.0) ? -1.0 : 1.0;
  Capgb *= (Capgb < 0.0) ? -1.0 : 1.0;
  const double gdr = F(M1F_GD), gsr = F(M1F_GS);
  const double Idrain = gdr * Vddp;
  const double Isource = gsr * Vssp;
  double Gm, Gmbs, nrmsum, revsum, cdreq;
  if (mode >= 0) {
    Gm = gm;
    Gmbs = gmbs;
    nrmsum = Gm + Gmbs;
    revsum = 0;
    cdreq = dtype * cdrain;
  } else {
    Gm = -gm;
    Gmbs = -gmbs;
    nrmsum = 0;
    revsum = -(Gm + Gmbs);
    cdreq = -(dtype)*cdrain;
  }

  // ---- Master::updateState: store and state writes, Meyer charges
  p.sto_next[sto + 0] = vbd;
  p.sto_next[sto + 1] = vbs;
  p.sto_next[sto + 2] = vgs;
  p.sto_next[sto + 3] = vds;
  p.sto_next[sto + 4] = von;
  p.sta_next[sta + 3] = capgs;
  p.sta_next[sta + 4] = capgd;
  p.sta_next[sta + 5] = capgb;
  double qgs, qgd, qgb;
  if (ss.dcop) {
    qgs = Capgs * vgs;
    qgd = Capgd * vgd;
    qgb = Capgb * Vgb;
  } else {
    qgs = p.sta_curr[sta + 0];
    qgd = p.sta_curr[sta + 1];
    qgb = p.sta_curr[sta + 2];
    const double vgs1 = p.sto_curr[sto + 2];
    const double vbs1 = p.sto_curr[sto + 1];
    const double vds1 = p.sto_curr[sto + 3];
    const double vgb1 = vgs1 - vbs1;
    const double vgd1 = vgs1 - vds1;
    qgs += Capgs * (vgs - vgs1);
    qgd += Capgd * (vgd - vgd1);
    qgb += Capgb * ((vgs - vbs) - vgb1);
  }
  p.sta_next[sta + 0] = qgs;
  p.sta_next[sta + 1] = qgd;
  p.sta_next[sta + 2] = qgb;
  p.sta_next[sta + 6] = qbd;
  p.sta_next[sta + 7] = qbs;
  // ---- Master::loadDAEVectors values
  double *out = a.slots + slot;
  double chk = 0.0;  // running sum of everything written, for the finiteness flag
  auto W = [&](int k, double v) {
    out[(size_t)k * S] = v;
    chk = chk + v;
  };
  const double np = F(M1F_M);
  const double ceqbs = dtype * (cbs), ceqbd = dtype * (cbd);
  const double ceqgb = 0.0, ceqgs = 0.0, ceqgd = 0.0;
  if (flags & XF_HAS_DP) W(M1S_F_D, Idrain * np);
  W(M1S_F_G, (ceqgs + ceqgd + ceqgb) * np);
  if (flags & XF_HAS_SP) W(M1S_F_S, Isource * np);
  W(M1S_F_B, (ceqbs + ceqbd - ceqgb) * np);
  W(M1S_F_DP, (-Idrain - (ceqbd - cdreq + ceqgd)) * np);
  W(M1S_F_SP, (-Isource - (ceqbs + cdreq + ceqgs)) * np);
  const double Qeqbs = dtype * (qbs), Qeqbd = dtype * (qbd),
               Qeqgb = dtype * (qgb), Qeqgs = dtype * (qgs),
               Qeqgd = dtype * (qgd);
  W(M1S_Q_G, (Qeqgs + Qeqgd + Qeqgb) * np);
  W(M1S_Q_B, (Qeqbs + Qeqbd - Qeqgb) * np);
  W(M1S_Q_DP, (-(Qeqbd + Qeqgd)) * np);
  W(M1S_Q_SP, (-(Qeqbs + Qeqgs)) * np);
  const bool cap_on = ss.tranop || ss.acop || ss.transient;
  const double gcgd = cap_on ? Capgd : 0.0, gcgs = cap_on ? Capgs : 0.0,
               gcgb = cap_on ? Capgb : 0.0, gcbs = cap_on ? capbs : 0.0,
               gcbd = cap_on ? capbd : 0.0;
  if (!origFlag) {
    const double dvgs = vgs - vgs_orig, dvgd = vgd - vgd_orig,
                 dvbs = vbs - vbs_orig, dvbd = vbd - vbd_orig,
                 dvds = vds - vds_orig;
    const double jx4 = dtype * (+((gbd - gmin)) * dvbd + ((gbs - gmin)) * dvbs);
    const double jx5 =
        dtype * (-((gbd - gmin)) * dvbd + gds * dvds +
                 Gm * ((mode > 0) ? dvgs : dvgd) + Gmbs * ((mode > 0) ? dvbs : dvbd));
    const double jx6 =
        dtype * (-((gbs - gmin)) * dvbs - gds * dvds -
                 Gm * ((mode > 0) ? dvgs : dvgd) - Gmbs * ((mode > 0) ? dvbs : dvbd));
    W(M1S_FL_B, jx4 * np);
    W(M1S_FL_DP, jx5 * np);
    W(M1S_FL_SP, jx6 * np);
    const double q2 = dtype * (gcgd * dvgd + gcgs * dvgs +
                               gcgb * (vgs - vgs_orig - vbs + vbs_orig));
    const double q4 = dtype * (-(gcgb) * (vgs - vgs_orig - vbs + vbs_orig) +
                               (gcgb)*dvbd + (gcbs)*dvbs);
    const double q5 = dtype * (-(gcgd)*dvgd - (gcbd)*dvbd);
    const double q6 = dtype * (-gcgs * dvgs - (gcbs)*dvbs);
    W(M1S_QL_G, q2 * np);
    W(M1S_QL_B, q4 * np);
    W(M1S_QL_DP, q5 * np);
    W(M1S_QL_SP, q6 * np);
  } else {
    // A skipped "+=" is the identity: -0.0 is the additive identity for every
    // accumulator value, including -0.0 itself.
    W(M1S_FL_B, -0.0);
    W(M1S_FL_DP, -0.0);
    W(M1S_FL_SP, -0.0);
    W(M1S_QL_G, -0.0);
    W(M1S_QL_B, -0.0);
    W(M1S_QL_DP, -0.0);
    W(M1S_QL_SP, -0.0);
  }

  // ---- Master::loadDAEMatrices values ("x -= v" stored as "+ (-v)")
  const int JF = M1S_JF0, JQ = M1S_JQ0;
  if (flags & XF_HAS_DP) {
    W(JF + 0, gdr * np);
    W(JF + 1, -(gdr * np));
    W(JF + 7, -(gdr * np));
  }
  if (flags & XF_HAS_SP) {
    W(JF + 2, gsr * np);
    W(JF + 3, -(gsr * np));
    W(JF + 13, -(gsr * np));
  }
  W(JF + 4, (gbs + gbd) * np);
  W(JF + 5, -(gbd * np));
  W(JF + 6, -(gbs * np));
  W(JF + 8, (Gm)*np);
  W(JF + 9, (-gbd + Gmbs) * np);
  W(JF + 10, (gdr + gds + gbd + revsum) * np);
  W(JF + 11, (-gds - nrmsum) * np);
  W(JF + 12, -((Gm)*np));
  W(JF + 14, -((gbs + Gmbs) * np));
  W(JF + 15, -((gds + revsum) * np));
  W(JF + 16, (gsr + gds + gbs + nrmsum) * np);

  W(JQ + 0, (gcgd + gcgs + gcgb) * np);
  W(JQ + 1, -(gcgb * np));
  W(JQ + 2, -(gcgd * np));
  W(JQ + 3, -(gcgs * np));
  W(JQ + 4, -(gcgb * np));
  W(JQ + 5, (+gcbs + gcbd + gcgb) * np);
  W(JQ + 6, -(+gcbd * np));
  W(JQ + 7, -(+gcbs * np));
  W(JQ + 8, -gcgd * np);
  W(JQ + 9, -gcbd * np);
  W(JQ + 10, (+gcbd + gcgd) * np);
  W(JQ + 11, -(gcgs * np));
  W(JQ + 12, -(+gcbs * np));
  W(JQ + 13, (+gcbs + gcgs) * np);
  // Bit 0: junction voltages limited (the reference's limitedFlag).  Bit 1:
  // a non-finite statement value or next-state value (their sum is finite
  // exactly when every term is: the magnitudes are far from overflow).
  chk = chk + qgs + qgd + qgb + qbd + qbs;
  a.limited[slot] = (limitedFlag ? 1 : 0) | ((chk - chk) != 0.0 ? 2 : 0);
}

}  // namespace

hipError_t launch_mos1_eval(const Mos1Args &a, const SolverStateDev &ss,
                            const LoadPtrs &p, hipStream_t s) {
  if (a.n_slots == 0) return hipSuccess;
  const int grid = (a.n_slots + kBlock - 1) / kBlock;
  hipLaunchKernelGGL(mos1_eval_kernel, dim3(grid), dim3(kBlock), 0, s, a, ss,
                     p);
  return hipGetLastError();
}

}  // namespace xg
