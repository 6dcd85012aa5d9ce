// Data layout of the BSIM3 load: card fields (per model and temperature),
// size-dependent record fields (per model, L, W — the reference's
// SizeDependParam cache key) and per-instance geometry fields, plus the
// stamp-value slots in the reference's statement order.
#pragma once

namespace xg {

// Card: model-level and temperature-level quantities the load reads.  One
// card per distinct content, staged in LDS by the eval kernel.
enum B3Card : int {
  B3C_DTYPE = 0, B3C_MOBMOD, B3C_XPART, B3C_VCRIT, B3C_NVTM, B3C_VTM,
  B3C_IJTH, B3C_TEMPRATIO, B3C_PHIB, B3C_PHIBSW, B3C_PHIBSWG, B3C_MJ,
  B3C_MJSW, B3C_MJSWG, B3C_GMIN, B3C_LIMITER, B3C_CAPMOD,
  B3C_N = 17
};

// Size-dependent record (SizeDependParam plus the model fields a
// Monte-Carlo sample may perturb: tox, cox, factor1).  B3R_VFBQ is the
// flat-band voltage of the selected charge model: vfbcv for capMod 0,
// vfbzb for capMod 1-3 (each model reads only its own).
enum B3SdpField : int {
  B3R_LEFF = 0, B3R_WEFF, B3R_LEFFCV, B3R_WEFFCV, B3R_VTH0, B3R_VBSC,
  B3R_PHI, B3R_PHIS3, B3R_XDEP0, B3R_SQRTPHI, B3R_VBI, B3R_DVT2, B3R_DVT2W,
  B3R_DVT1, B3R_DVT0, B3R_DVT1W, B3R_DVT0W, B3R_NLX, B3R_K1OX, B3R_KT1,
  B3R_KT1L, B3R_KT2, B3R_W0, B3R_ETA0, B3R_ETAB, B3R_THETA0VB0, B3R_K1,
  B3R_K2OX, B3R_K3, B3R_K3B, B3R_NFACTOR, B3R_CDSC, B3R_CDSCB, B3R_CDSCD,
  B3R_CIT, B3R_VFB, B3R_NGATE, B3R_VOFF, B3R_CDEP0, B3R_DWG, B3R_DWB,
  B3R_PRWG, B3R_PRWB, B3R_RDS0, B3R_XJ, B3R_A0, B3R_B1, B3R_B0, B3R_AGS,
  B3R_KETA, B3R_UA, B3R_UC, B3R_UB, B3R_U0TEMP, B3R_VSATTEMP, B3R_A1,
  B3R_A2, B3R_DELTA, B3R_PCLM, B3R_LITL, B3R_THETAROUT, B3R_PDIBLB,
  B3R_PVAG, B3R_PSCBE1, B3R_PSCBE2, B3R_ALPHA0, B3R_ALPHA1, B3R_BETA0,
  B3R_VFBQ, B3R_ACDE, B3R_LDEB, B3R_MOIN, B3R_ABULKCVF, B3R_NOFF,
  B3R_VOFFCV, B3R_CGBO, B3R_CGDO, B3R_CGSO, B3R_CGDL, B3R_CGSL, B3R_CKAPPA,
  B3R_TOX, B3R_COX, B3R_FACTOR1,
  B3R_N
};

// Per-instance geometry / junction fields.
enum B3Inst : int {
  B3I_SSATCUR = 0, B3I_DSATCUR, B3I_VJSM, B3I_ISEVJSM, B3I_VJDM, B3I_ISEVJDM,
  B3I_CZBS, B3I_CZBSSW, B3I_CZBSSWG, B3I_CZBD, B3I_CZBDSW, B3I_CZBDSWG,
  B3I_GD, B3I_GS, B3I_M,
  B3I_N
};

// Stamp-value slots (N_DEV_MOSFET_B3.C:9306-9643 vectors, 9654-9928
// matrices).
enum B3Slot : int {
  B3S_F_D = 0, B3S_F_S, B3S_F_B, B3S_F_DP, B3S_F_SP,
  B3S_FL_B, B3S_FL_DP, B3S_FL_SP,
  B3S_Q_G, B3S_Q_B, B3S_Q_DP, B3S_Q_SP,
  B3S_QL_G, B3S_QL_B, B3S_QL_DP, B3S_QL_SP,
  B3S_JF0,                 // 18 dF/dx statements
  B3S_JQ0 = B3S_JF0 + 18,  // 16 dQ/dx statements
  B3S_N = B3S_JQ0 + 16
};

}  // namespace xg
