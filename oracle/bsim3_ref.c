/*
 * ORACLE — test infrastructure only.  Nothing in the product links this.
 *
 * CPU restatement of Xyce's BSIM3 v3.2.2 device load
 * (src/DeviceModelPKG/OpenModels/src/N_DEV_MOSFET_B3.C) in the reference's
 * own shape: a Model struct holding the whole parameter card and its
 * size-dependent-parameter cache, array-of-structs instances, a Master that
 * loops over them in insertion order, and read-modify-write stamping
 * through cached pointers into the CSR value arrays.
 *
 * Restated: Model::Model/processParams (N_DEV_MOSFET_B3.C:8481-9100),
 * Instance::Instance/processParams (1845-2785, stamp part handled by
 * stamp_ref.py), updateTemperature with the SizeDependParam cache
 * (3611-4518), updateIntermediateVars (4526-7273) for capMod 0/1/2/3
 * and every mobMod, setupCapacitors_oldDAE/newDAE (7558-7885),
 * auxChargeCalculations (7465-7555), Master::updateState /
 * loadDAEVectors / loadDAEMatrices (9222-9928).
 * Not restated: nqsMod, IC= (extra branch-current unknowns),
 * homotopy (artParameterFlag), lead currents.  Instances using them are
 * rejected at setup.
 */
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include "ref_common.h"

#define RC_Eg0 (1.16)
#define RC_alphaEg (7.02e-4)
#define RC_betaEg (1108.0)
#define RC_Ni0 (1.45e10)
#define RC_vt0 (RC_BOLTZ * (27.0 + 273.15) / RC_Q)
#define RC_PI 3.14159265358979323846

enum {
#define XG_B3P(N, f, d) B3P_##N,
#include "b3_params.def"
#undef XG_B3P
  B3P_COUNT
};

typedef struct ref_b3_sdp {
  double Width, Length;
  double cdsc, cdscb, cdscd, cit, nfactor, xj, vsat, at, a0, ags, a1, a2,
      keta, nsub, npeak, ngate, gamma1, gamma2, vbx, vbi, vbm, vbsc, xt, phi,
      litl, k1, kt1, kt1l, kt2, k2, k3, k3b, w0, nlx, dvt0, dvt1, dvt2, dvt0w,
      dvt1w, dvt2w, drout, dsub, vth0, ua, ua1, ub, ub1, uc, uc1, u0, ute,
      voff, vfb, delta, rdsw, rds0, prwg, prwb, prt, eta0, etab, pclm,
      pdibl1, pdibl2, pdiblb, pscbe1, pscbe2, pvag, wr, dwg, dwb, b0, b1,
      alpha0, alpha1, beta0;
  double elm, cgsl, cgdl, ckappa, cf, clc, cle, vfbcv, noff, voffcv, acde,
      moin;
  double dw, dl, leff, weff, dwc, dlc, leffCV, weffCV, abulkCVfactor, cgso,
      cgdo, cgbo, tconst, u0temp, vsattemp, sqrtPhi, phis3, Xdep0, sqrtXdep0,
      theta0vb0, thetaRout, cof1, cof2, cof3, cof4, cdep0, vfbzb, ldeb, k1ox,
      k2ox;
  double referenceTemperature;
  struct ref_b3_sdp *next;
} ref_b3_sdp;

typedef struct ref_b3_model {
  int dtype;
#define XG_B3P(N, f, d) double f;
#include "b3_params.def"
#undef XG_B3P
  unsigned char given[B3P_COUNT];
  double cox, vcrit, factor1, Vtm0, Eg0, ni;
  ref_b3_sdp *sdp_list;
} ref_b3_model;

#define GIVEN(m, N) ((m)->given[B3P_##N])

typedef struct ref_b3_inst {
  ref_b3_model *model;
  ref_b3_sdp *paramPtr;
  int OFF;
  double l, w, drainArea, sourceArea, drainSquares, sourceSquares,
      drainPerimeter, sourcePerimeter, numberParallel, temp;
  double drainConductance, sourceConductance;
  /* temperature */
  double vtm, jctTempSatCurDensity, jctSidewallTempSatCurDensity,
      unitAreaJctCapTemp, unitLengthSidewallJctCapTemp,
      unitLengthGateSidewallJctCapTemp, PhiBTemp, PhiBSWTemp, PhiBSWGTemp;
  double vjsm, IsEvjsm, vjdm, IsEvjdm, cgso, cgdo;
  /* eval results */
  double ueff, thetavth, von, vdsat;
  double Vd, Vs, Vg, Vb, Vsp, Vdp, Vddp, Vssp, Vbsp, Vbdp, Vgsp, Vgdp, Vgb,
      Vdpsp;
  double vbs, vgs, vds, vbd, vgd, vgb;
  double vbs_orig, vgs_orig, vds_orig, vbd_orig, vgd_orig;
  double vbs_old, vgs_old, vds_old, vbd_old;
  int origFlag, limitedFlag, mode, ChargeComputationNeeded;
  double gbs, cbs, gbd, cbd, cdrain, gds, gm, gmbs, gbbs, gbgs, gbds, csub,
      cd, qgate, qdrn, qsrc, qbulk, qinv, cggb, cgsb, cgdb, cdgb, cdsb, cddb,
      cbgb, cbsb, cbdb, cqdb, cqsb, cqgb, cqbb, gtau, qbs, qbd, capbs, capbd,
      qgdo, qgso, qgd, qgs, qgb, qg, qd, qb, qcheq, qcdump, qdef, cqdef,
      Idrain, Isource, T1global;
  double CAPcggb, CAPcgdb, CAPcgsb, CAPcdgb, CAPcddb, CAPcdsb, CAPcsgb,
      CAPcsdb, CAPcssb, CAPcbgb, CAPcbdb, CAPcbsb;
  double gcggb, gcgdb, gcgsb, gcdgb, gcddb, gcdsb, gcsgb, gcsdb, gcssb,
      gcbgb, gcbdb, gcbsb;
  double ggtg, ggtd, ggtb, ggts, sxpart, dxpart, ddxpart_dVd, ddxpart_dVg,
      ddxpart_dVb, ddxpart_dVs, dsxpart_dVd, dsxpart_dVg, dsxpart_dVb,
      dsxpart_dVs;
  double Qeqqg_Jdxp, Qeqqb_Jdxp, Qeqqd_Jdxp;
  /* loadDAEVectors quantities */
  double Gm, Gmbs, FwdSum, RevSum, cdreq, ceqbd, ceqbs, gbbdp, gbbsp, gbdpg,
      gbdpdp, gbdpb, gbdpsp, gbspg, gbspdp, gbspb, gbspsp, cdreq_Jdxp,
      ceqbd_Jdxp, ceqbs_Jdxp;
  int li_Drain, li_DrainPrime, li_Source, li_SourcePrime, li_Gate, li_Bulk;
  int li_store_vbd, li_store_vbs, li_store_vgs, li_store_vds, li_store_von;
  int li_state_qb, li_state_qg, li_state_qd, li_state_qbs, li_state_qbd,
      li_state_qcheq, li_state_qcdump, li_state_qdef;
  int off[22];
  double *fp[22], *qp[22];
} ref_b3_inst;

enum {
  O_D_D, O_D_DP, O_G_G, O_G_B, O_G_DP, O_G_SP, O_S_S, O_S_SP, O_B_G, O_B_B,
  O_B_DP, O_B_SP, O_DP_D, O_DP_G, O_DP_B, O_DP_DP, O_DP_SP, O_SP_G, O_SP_S,
  O_SP_B, O_SP_DP, O_SP_SP
};

typedef struct ref_b3_master {
  int n, cap;
  ref_b3_inst *inst;
  ref_dev_options opt;
  ref_solver_state ss;
  double *nextSol, *currSta, *currSto, *nextSta, *nextSto;
} ref_b3_master;

static const char *b3_names[B3P_COUNT] = {
#define XG_B3P(N, f, d) #N,
#include "b3_params.def"
#undef XG_B3P
};

/* Model::processParams, N_DEV_MOSFET_B3.C:8481-8571 */
static void b3_model_process(ref_b3_model *m) {
  m->cox = 3.453133e-11 / m->tox;
  if (!GIVEN(m, TOXM)) m->toxm = m->tox;
  if (!GIVEN(m, DSUB)) m->dsub = m->drout;
  if (!GIVEN(m, LLC)) m->Llc = m->Ll;
  if (!GIVEN(m, LWC)) m->Lwc = m->Lw;
  if (!GIVEN(m, LWLC)) m->Lwlc = m->Lwl;
  if (!GIVEN(m, WLC)) m->Wlc = m->Wl;
  if (!GIVEN(m, WWL)) m->Wwlc = m->Wwl;   /* as written in the reference */
  if (!GIVEN(m, WWLC)) m->Wwlc = m->Wwl;
  if (!GIVEN(m, DWC)) m->dwc = m->Wint;
  if (!GIVEN(m, DLC)) m->dlc = m->Lint;
  if (!GIVEN(m, CF)) {
    double C1 = 2.0 * RC_EPSOX;
    double C5 = RC_PI;
    double C2 = 1.0 + (0.4e-6 / m->tox);
    double C3 = log(C2);
    m->cf = C1 * C3 / C5;
  }
  if (!GIVEN(m, CGDO)) {
    if (GIVEN(m, DLC) && (m->dlc > 0.0)) m->cgdo = m->dlc * m->cox - m->cgdl;
    else m->cgdo = 0.6 * m->xj * m->cox;
  }
  if (!GIVEN(m, CGSO)) {
    if (GIVEN(m, DLC) && (m->dlc > 0.0)) m->cgso = m->dlc * m->cox - m->cgsl;
    else m->cgso = 0.6 * m->xj * m->cox;
  }
  if (!GIVEN(m, CGBO)) m->cgbo = 2.0 * m->dwc * m->cox;
  if (!GIVEN(m, CJSWG)) m->unitLengthGateSidewallJctCap = m->unitLengthSidewallJctCap;
  if (!GIVEN(m, PBSWG)) m->GatesidewallJctPotential = m->sidewallJctPotential;
  if (!GIVEN(m, MJSWG)) m->bulkJctGateSideGradingCoeff = m->bulkJctSideGradingCoeff;
  if (m->bulkJctPotential < 0.1) m->bulkJctPotential = 0.1;
  if (m->sidewallJctPotential < 0.1) m->sidewallJctPotential = 0.1;
  if (m->GatesidewallJctPotential < 0.1) m->GatesidewallJctPotential = 0.1;
  m->vcrit = RC_vt0 * log(RC_vt0 / (RC_ROOT2 * 1.0e-14));
  m->factor1 = sqrt(RC_EPSSI / RC_EPSOX * m->tox);
  m->Vtm0 = RC_KoverQ * m->tnom;
  m->Eg0 = RC_Eg0 - RC_alphaEg * m->tnom * m->tnom / (m->tnom + RC_betaEg);
  m->ni = RC_Ni0 * (m->tnom / RC_REFTEMP) * sqrt(m->tnom / RC_REFTEMP) *
          exp(21.5565981 - m->Eg0 / (2.0 * m->Vtm0));
}

/* Model::Model, N_DEV_MOSFET_B3.C:8596-9100.  TNOM is given in Celsius as
   on a .MODEL line (N_DEV_DeviceEntity.C:863-866).  Returns NULL on an
   unknown parameter name. */
ref_b3_model *ref_b3_model_new(int dtype, int n, const char **keys,
                               const double *vals, double tnom_default) {
  ref_b3_model *m = (ref_b3_model *)calloc(1, sizeof(*m));
  int i, k;
  m->dtype = dtype;
#define XG_B3P(N, f, d) m->f = (d);
#include "b3_params.def"
#undef XG_B3P
  for (i = 0; i < n; ++i) {
    for (k = 0; k < B3P_COUNT; ++k)
      if (!strcmp(keys[i], b3_names[k])) break;
    if (k == B3P_COUNT) {
      if (!strcmp(keys[i], "LEVEL")) continue;
      free(m);
      return NULL;
    }
    m->given[k] = 1;
    switch (k) {
#define XG_B3P(N, f, d) case B3P_##N: m->f = vals[i]; break;
#include "b3_params.def"
#undef XG_B3P
    }
  }
  if (GIVEN(m, TNOM)) m->tnom += 273.15;
  else m->tnom = tnom_default;
  if (!GIVEN(m, VTH0)) m->vth0 = (dtype == 1) ? 0.7 : -0.7;
  if (!GIVEN(m, UC)) m->uc = (m->mobMod == 3) ? -0.0465 : -0.0465e-9;
  if (!GIVEN(m, UC1)) m->uc1 = (m->mobMod == 3) ? -0.056 : -0.056e-9;
  if (!GIVEN(m, U0)) m->u0 = (dtype == 1) ? 0.067 : 0.025;
  b3_model_process(m);
  return m;
}

void ref_b3_model_free(ref_b3_model *m) {
  ref_b3_sdp *p = m->sdp_list, *q;
  while (p) {
    q = p->next;
    free(p);
    p = q;
  }
  free(m);
}

/* Instance::updateTemperature, N_DEV_MOSFET_B3.C:3611-4518 */
static void b3_update_temperature(ref_b3_inst *in) {
  ref_b3_model *m = in->model;
  double tmp, tmp1, tmp2, tmp3, Eg, T0, T1, T2, T3, T4, T5, Ldrn, Wdrn;
  double delTemp, TRatio, Inv_L, Inv_W, Inv_LW, Tnom, Nvtm, SourceSatCurrent,
      DrainSatCurrent;
  double Eg0 = m->Eg0, ni = m->ni, Vtm0 = m->Vtm0;
  ref_b3_sdp *p, *it;

  Tnom = m->tnom;
  TRatio = in->temp / Tnom;
  in->vtm = RC_KoverQ * in->temp;
  Eg = RC_Eg0 - RC_alphaEg * in->temp * in->temp / (in->temp + RC_betaEg);
  if (in->temp != Tnom) {
    T0 = Eg0 / Vtm0 - Eg / in->vtm + m->jctTempExponent * log(in->temp / Tnom);
    T1 = exp(T0 / m->jctEmissionCoeff);
    in->jctTempSatCurDensity = m->jctSatCurDensity * T1;
    in->jctSidewallTempSatCurDensity = m->jctSidewallSatCurDensity * T1;
  } else {
    in->jctTempSatCurDensity = m->jctSatCurDensity;
    in->jctSidewallTempSatCurDensity = m->jctSidewallSatCurDensity;
  }
  if (in->jctTempSatCurDensity < 0.0) in->jctTempSatCurDensity = 0.0;
  if (in->jctSidewallTempSatCurDensity < 0.0) in->jctSidewallTempSatCurDensity = 0.0;
  delTemp = in->temp - Tnom;
  T0 = m->tcj * delTemp;
  if (T0 >= -1.0) in->unitAreaJctCapTemp = m->unitAreaJctCap * (1.0 + T0);
  else if (in->unitAreaJctCapTemp > 0.0) in->unitAreaJctCapTemp = 0.0;
  T0 = m->tcjsw * delTemp;
  if (T0 >= -1.0) in->unitLengthSidewallJctCapTemp = m->unitLengthSidewallJctCap * (1.0 + T0);
  else if (in->unitLengthSidewallJctCapTemp > 0.0) in->unitLengthSidewallJctCapTemp = 0.0;
  T0 = m->tcjswg * delTemp;
  if (T0 >= -1.0) in->unitLengthGateSidewallJctCapTemp = m->unitLengthGateSidewallJctCap * (1.0 + T0);
  else if (in->unitLengthGateSidewallJctCapTemp > 0.0) in->unitLengthGateSidewallJctCapTemp = 0.0;
  in->PhiBTemp = m->bulkJctPotential - m->tpb * delTemp;
  if (in->PhiBTemp < 0.01) in->PhiBTemp = 0.01;
  in->PhiBSWTemp = m->sidewallJctPotential - m->tpbsw * delTemp;
  if (in->PhiBSWTemp <= 0.01) in->PhiBSWTemp = 0.01;
  in->PhiBSWGTemp = m->GatesidewallJctPotential - m->tpbswg * delTemp;
  if (in->PhiBSWGTemp <= 0.01) in->PhiBSWGTemp = 0.01;

  p = NULL;
  for (it = m->sdp_list; it; it = it->next)
    if (it->Length == in->l && it->Width == in->w) p = it;
  if (!p) {
    p = (ref_b3_sdp *)calloc(1, sizeof(*p));
    /* appended at the end of the model's list, as push_back does */
    if (!m->sdp_list) m->sdp_list = p;
    else {
      for (it = m->sdp_list; it->next; it = it->next) {}
      it->next = p;
    }
    p->referenceTemperature = in->temp;
    Ldrn = in->l;
    Wdrn = in->w;
    p->Length = Ldrn;
    p->Width = Wdrn;
    T0 = pow(Ldrn, m->Lln);
    T1 = pow(Wdrn, m->Lwn);
    tmp1 = m->Ll / T0 + m->Lw / T1 + m->Lwl / (T0 * T1);
    p->dl = m->Lint + tmp1;
    tmp2 = m->Llc / T0 + m->Lwc / T1 + m->Lwlc / (T0 * T1);
    p->dlc = m->dlc + tmp2;
    T2 = pow(Ldrn, m->Wln);
    T3 = pow(Wdrn, m->Wwn);
    tmp1 = m->Wl / T2 + m->Ww / T3 + m->Wwl / (T2 * T3);
    p->dw = m->Wint + tmp1;
    tmp2 = m->Wlc / T2 + m->Wwc / T3 + m->Wwlc / (T2 * T3);
    p->dwc = m->dwc + tmp2;
    p->leff = in->l - 2.0 * p->dl;
    p->weff = in->w - 2.0 * p->dw;
    p->leffCV = in->l - 2.0 * p->dlc;
    p->weffCV = in->w - 2.0 * p->dwc;
    if (m->binUnit == 1) {
      Inv_L = 1.0e-6 / p->leff;
      Inv_W = 1.0e-6 / p->weff;
      Inv_LW = 1.0e-12 / (p->leff * p->weff);
    } else {
      Inv_L = 1.0 / p->leff;
      Inv_W = 1.0 / p->weff;
      Inv_LW = 1.0 / (p->leff * p->weff);
    }
#define BIN(x) p->x = m->x + m->l##x * Inv_L + m->w##x * Inv_W + m->p##x * Inv_LW
    BIN(cdsc); BIN(cdscb); BIN(cdscd); BIN(cit); BIN(nfactor); BIN(xj);
    BIN(vsat); BIN(at); BIN(a0); BIN(ags); BIN(a1); BIN(a2); BIN(keta);
    BIN(nsub); BIN(npeak); BIN(ngate); BIN(gamma1); BIN(gamma2); BIN(vbx);
    BIN(vbm); BIN(xt); BIN(vfb); BIN(k1); BIN(kt1); BIN(kt1l); BIN(k2);
    BIN(kt2); BIN(k3); BIN(k3b); BIN(w0); BIN(nlx); BIN(dvt0); BIN(dvt1);
    BIN(dvt2); BIN(dvt0w); BIN(dvt1w); BIN(dvt2w); BIN(drout); BIN(dsub);
    BIN(vth0); BIN(ua); BIN(ua1); BIN(ub); BIN(ub1); BIN(uc); BIN(uc1);
    BIN(u0); BIN(ute); BIN(voff); BIN(delta); BIN(rdsw); BIN(prwg);
    BIN(prwb); BIN(prt); BIN(eta0); BIN(etab); BIN(pclm); BIN(pdibl1);
    BIN(pdibl2); BIN(pdiblb); BIN(pscbe1); BIN(pscbe2); BIN(pvag); BIN(wr);
    BIN(dwg); BIN(dwb); BIN(b0); BIN(b1); BIN(alpha0); BIN(alpha1);
    BIN(beta0); BIN(elm); BIN(cgsl); BIN(cgdl); BIN(ckappa); BIN(cf);
    BIN(clc); BIN(cle); BIN(vfbcv); BIN(acde); BIN(moin); BIN(noff);
    BIN(voffcv);
#undef BIN
    p->abulkCVfactor = 1.0 + pow((p->clc / p->leffCV), p->cle);
    T0 = (TRatio - 1.0);
    p->ua = p->ua + p->ua1 * T0;
    p->ub = p->ub + p->ub1 * T0;
    p->uc = p->uc + p->uc1 * T0;
    if (p->u0 > 1.0) p->u0 = p->u0 / 1.0e4;
    p->u0temp = p->u0 * pow(TRatio, p->ute);
    p->vsattemp = p->vsat - p->at * T0;
    p->rds0 = (p->rdsw + p->prt * T0) / pow(p->weff * 1E6, p->wr);
    p->cgdo = (m->cgdo + p->cf) * p->weffCV;
    p->cgso = (m->cgso + p->cf) * p->weffCV;
    p->cgbo = m->cgbo * p->leffCV;
    T0 = p->leffCV * p->leffCV;
    p->tconst = p->u0temp * p->elm / (m->cox * p->weffCV * p->leffCV * T0);
    if (!GIVEN(m, NCH) && GIVEN(m, GAMMA1)) {
      T0 = p->gamma1 * m->cox;
      p->npeak = 3.021E22 * T0 * T0;
    }
    p->phi = 2.0 * Vtm0 * log(p->npeak / ni);
    p->sqrtPhi = sqrt(p->phi);
    p->phis3 = p->sqrtPhi * p->phi;
    p->Xdep0 = sqrt(2.0 * RC_EPSSI / (RC_Q * p->npeak * 1.0e6)) * p->sqrtPhi;
    p->sqrtXdep0 = sqrt(p->Xdep0);
    p->litl = sqrt(3.0 * p->xj * m->tox);
    p->vbi = Vtm0 * log(1.0e20 * p->npeak / (ni * ni));
    p->cdep0 = sqrt(RC_Q * RC_EPSSI * p->npeak * 1.0e6 / 2.0 / p->phi);
    p->ldeb = sqrt(RC_EPSSI * Vtm0 / (RC_Q * p->npeak * 1.0e6)) / 3.0;
    p->acde *= pow((p->npeak / 2.0e16), -0.25);
    if (GIVEN(m, K1) || GIVEN(m, K2)) {
      if (!GIVEN(m, K1)) p->k1 = 0.53;
      if (!GIVEN(m, K2)) p->k2 = -0.0186;
    } else {
      if (!GIVEN(m, VBX)) p->vbx = p->phi - 7.7348e-4 * p->npeak * p->xt * p->xt;
      if (p->vbx > 0.0) p->vbx = -p->vbx;
      if (p->vbm > 0.0) p->vbm = -p->vbm;
      if (!GIVEN(m, GAMMA1)) p->gamma1 = 5.753e-12 * sqrt(p->npeak) / m->cox;
      if (!GIVEN(m, GAMMA2)) p->gamma2 = 5.753e-12 * sqrt(p->nsub) / m->cox;
      T0 = p->gamma1 - p->gamma2;
      T1 = sqrt(p->phi - p->vbx) - p->sqrtPhi;
      T2 = sqrt(p->phi * (p->phi - p->vbm)) - p->phi;
      p->k2 = T0 * T1 / (2.0 * T2 + p->vbm);
      p->k1 = p->gamma2 - 2.0 * p->k2 * sqrt(p->phi - p->vbm);
    }
    if (p->k2 < 0.0) {
      T0 = 0.5 * p->k1 / p->k2;
      p->vbsc = 0.9 * (p->phi - T0 * T0);
      if (p->vbsc > -3.0) p->vbsc = -3.0;
      else if (p->vbsc < -30.0) p->vbsc = -30.0;
    } else {
      p->vbsc = -30.0;
    }
    if (p->vbsc > p->vbm) p->vbsc = p->vbm;
    if (!GIVEN(m, VFB)) {
      if (GIVEN(m, VTH0)) p->vfb = m->dtype * p->vth0 - p->phi - p->k1 * p->sqrtPhi;
      else p->vfb = -1.0;
    }
    if (!GIVEN(m, VTH0)) p->vth0 = m->dtype * (p->vfb + p->phi + p->k1 * p->sqrtPhi);
    p->k1ox = p->k1 * m->tox / m->toxm;
    p->k2ox = p->k2 * m->tox / m->toxm;
    T1 = sqrt(RC_EPSSI / RC_EPSOX * m->tox * p->Xdep0);
    T0 = exp(-0.5 * p->dsub * p->leff / T1);
    p->theta0vb0 = (T0 + 2.0 * T0 * T0);
    T0 = exp(-0.5 * p->drout * p->leff / T1);
    T2 = (T0 + 2.0 * T0 * T0);
    p->thetaRout = p->pdibl1 * T2 + p->pdibl2;
    tmp = sqrt(p->Xdep0);
    tmp1 = p->vbi - p->phi;
    tmp2 = m->factor1 * tmp;
    T0 = -0.5 * p->dvt1w * p->weff * p->leff / tmp2;
    if (T0 > -RC_EXP_THRESHOLD) {
      T1 = exp(T0);
      T2 = T1 * (1.0 + 2.0 * T1);
    } else {
      T1 = RC_MIN_EXP;
      T2 = T1 * (1.0 + 2.0 * T1);
    }
    T0 = p->dvt0w * T2;
    T2 = T0 * tmp1;
    T0 = -0.5 * p->dvt1 * p->leff / tmp2;
    if (T0 > -RC_EXP_THRESHOLD) {
      T1 = exp(T0);
      T3 = T1 * (1.0 + 2.0 * T1);
    } else {
      T1 = RC_MIN_EXP;
      T3 = T1 * (1.0 + 2.0 * T1);
    }
    T3 = p->dvt0 * T3 * tmp1;
    T4 = m->tox * p->phi / (p->weff + p->w0);
    T0 = sqrt(1.0 + p->nlx / p->leff);
    T5 = p->k1ox * (T0 - 1.0) * p->sqrtPhi + (p->kt1 + p->kt1l / p->leff) * (TRatio - 1.0);
    tmp3 = m->dtype * p->vth0 - T2 - T3 + p->k3 * T4 + T5;
    p->vfbzb = tmp3 - p->phi - p->k1 * p->sqrtPhi;
  }
  in->paramPtr = p;
  in->cgso = p->cgso;
  in->cgdo = p->cgdo;
  Nvtm = in->vtm * m->jctEmissionCoeff;
  if ((in->sourceArea <= 0.0) && (in->sourcePerimeter <= 0.0)) SourceSatCurrent = 1.0e-14;
  else SourceSatCurrent = in->sourceArea * in->jctTempSatCurDensity +
                          in->sourcePerimeter * in->jctSidewallTempSatCurDensity;
  if ((SourceSatCurrent > 0.0) && (m->ijth > 0.0)) {
    in->vjsm = Nvtm * log(m->ijth / SourceSatCurrent + 1.0);
    in->IsEvjsm = SourceSatCurrent * exp(in->vjsm / Nvtm);
  }
  if ((in->drainArea <= 0.0) && (in->drainPerimeter <= 0.0)) DrainSatCurrent = 1.0e-14;
  else DrainSatCurrent = in->drainArea * in->jctTempSatCurDensity +
                         in->drainPerimeter * in->jctSidewallTempSatCurDensity;
  if ((DrainSatCurrent > 0.0) && (m->ijth > 0.0)) {
    in->vjdm = Nvtm * log(m->ijth / DrainSatCurrent + 1.0);
    in->IsEvjdm = DrainSatCurrent * exp(in->vjdm / Nvtm);
  }
}

ref_b3_master *ref_b3_master_new(double gmin, int limiter) {
  ref_b3_master *M = (ref_b3_master *)calloc(1, sizeof(*M));
  M->opt.gmin = gmin;
  M->opt.voltageLimiterFlag = limiter;
  return M;
}

void ref_b3_master_free(ref_b3_master *M) {
  free(M->inst);
  free(M);
}

/* Instance::Instance + processParams (N_DEV_MOSFET_B3.C:1845-1880,
   2440-2470), registerLIDs, registerStateLIDs (3011-3060) and
   registerStoreLIDs (3076-3100).  TEMP is in Celsius.  Returns the index,
   -1 on an unknown parameter, -2 for IC=/NQSMOD (not restated). */
int ref_b3_add_instance(ref_b3_master *M, ref_b3_model *m, int n,
                        const char **keys, const double *vals,
                        const int *lids, int sta_base, int sto_base,
                        double temp) {
  ref_b3_inst *in;
  int i, gL = 0, gW = 0;
  if (M->n == M->cap) {
    M->cap = M->cap ? 2 * M->cap : 64;
    M->inst = (ref_b3_inst *)realloc(M->inst, M->cap * sizeof(ref_b3_inst));
  }
  in = &M->inst[M->n];
  memset(in, 0, sizeof(*in));
  in->model = m;
  in->l = 1.0e-4; in->w = 1.0e-4;
  in->drainSquares = 1.0; in->sourceSquares = 1.0;
  in->numberParallel = 1.0;
  in->temp = temp;
  for (i = 0; i < n; ++i) {
    const char *k = keys[i];
    double v = vals[i];
    if (!strcmp(k, "L")) { in->l = v; gL = 1; }
    else if (!strcmp(k, "W")) { in->w = v; gW = 1; }
    else if (!strcmp(k, "AD")) in->drainArea = v;
    else if (!strcmp(k, "AS")) in->sourceArea = v;
    else if (!strcmp(k, "NRD")) in->drainSquares = v;
    else if (!strcmp(k, "NRS")) in->sourceSquares = v;
    else if (!strcmp(k, "PD")) in->drainPerimeter = v;
    else if (!strcmp(k, "PS")) in->sourcePerimeter = v;
    else if (!strcmp(k, "M")) in->numberParallel = v;
    else if (!strcmp(k, "OFF")) in->OFF = (v != 0.0);
    else if (!strcmp(k, "TEMP")) in->temp = v + 273.15;
    else if (!strcmp(k, "IC1") || !strcmp(k, "IC2") || !strcmp(k, "IC3") ||
             !strcmp(k, "NQSMOD")) return -2;
    else return -1;
  }
  if (!gL) in->l = m->model_l;
  if (!gW) in->w = m->model_w;
  in->drainConductance = m->sheetResistance * in->drainSquares;
  if (in->drainConductance > 0.0) in->drainConductance = 1.0 / in->drainConductance;
  else in->drainConductance = 0.0;
  in->sourceConductance = m->sheetResistance * in->sourceSquares;
  if (in->sourceConductance > 0.0) in->sourceConductance = 1.0 / in->sourceConductance;
  else in->sourceConductance = 0.0;
  b3_update_temperature(in);
  in->li_Drain = lids[0]; in->li_Gate = lids[1];
  in->li_Source = lids[2]; in->li_Bulk = lids[3];
  in->li_DrainPrime = in->drainConductance != 0.0 ? lids[4] : in->li_Drain;
  in->li_SourcePrime = in->sourceConductance != 0.0 ? lids[5] : in->li_Source;
  in->li_state_qb = sta_base + 0; in->li_state_qg = sta_base + 1;
  in->li_state_qd = sta_base + 2; in->li_state_qbs = sta_base + 3;
  in->li_state_qbd = sta_base + 4; in->li_state_qcheq = sta_base + 5;
  in->li_state_qcdump = sta_base + 6; in->li_state_qdef = sta_base + 7;
  in->li_store_vbd = sto_base + 0; in->li_store_vbs = sto_base + 1;
  in->li_store_vgs = sto_base + 2; in->li_store_vds = sto_base + 3;
  in->li_store_von = sto_base + 4;
  return M->n++;
}

void ref_b3_set_offsets(ref_b3_master *M, int i, const int *off22) {
  memcpy(M->inst[i].off, off22, sizeof(int) * 22);
}

void ref_b3_setup_pointers(ref_b3_master *M, const int *rowptr, double *dFdx,
                           double *dQdx) {
  int i, k;
  for (i = 0; i < M->n; ++i) {
    ref_b3_inst *in = &M->inst[i];
    int row[22] = {in->li_Drain, in->li_Drain, in->li_Gate, in->li_Gate,
                   in->li_Gate, in->li_Gate, in->li_Source, in->li_Source,
                   in->li_Bulk, in->li_Bulk, in->li_Bulk, in->li_Bulk,
                   in->li_DrainPrime, in->li_DrainPrime, in->li_DrainPrime,
                   in->li_DrainPrime, in->li_DrainPrime, in->li_SourcePrime,
                   in->li_SourcePrime, in->li_SourcePrime, in->li_SourcePrime,
                   in->li_SourcePrime};
    for (k = 0; k < 22; ++k) {
      in->fp[k] = dFdx + rowptr[row[k]] + in->off[k];
      in->qp[k] = dQdx + rowptr[row[k]] + in->off[k];
    }
  }
}

void ref_b3_set_solver_state(ref_b3_master *M, const int *flags) {
  M->ss.dcopFlag = flags[0];
  M->ss.initJctFlag = flags[1];
  M->ss.initFixFlag = flags[2];
  M->ss.initTranFlag = flags[3];
  M->ss.newtonIter = flags[4];
  M->ss.tranopFlag = flags[5];
  M->ss.acopFlag = flags[6];
  M->ss.transientFlag = flags[7];
}

/* Junction depletion charge/capacitance of one side (N_DEV_MOSFET_B3.C:
   7009-7134). */
static void b3_junction(ref_b3_inst *in, double v, double czb, double czbsw,
                        double czbswg, double *q, double *cap) {
  ref_b3_model *m = in->model;
  double MJ = m->bulkJctBotGradingCoeff, MJSW = m->bulkJctSideGradingCoeff,
         MJSWG = m->bulkJctGateSideGradingCoeff, arg, sarg, T0, T1;
  if (v == 0.0) {
    *q = 0.0;
    *cap = czb + czbsw + czbswg;
  } else if (v < 0.0) {
    if (czb > 0.0) {
      arg = 1.0 - v / in->PhiBTemp;
      if (MJ == 0.5) sarg = 1.0 / sqrt(arg);
      else sarg = exp(-MJ * log(arg));
      *q = in->PhiBTemp * czb * (1.0 - arg * sarg) / (1.0 - MJ);
      *cap = czb * sarg;
    } else {
      *q = 0.0;
      *cap = 0.0;
    }
    if (czbsw > 0.0) {
      arg = 1.0 - v / in->PhiBSWTemp;
      if (MJSW == 0.5) sarg = 1.0 / sqrt(arg);
      else sarg = exp(-MJSW * log(arg));
      *q += in->PhiBSWTemp * czbsw * (1.0 - arg * sarg) / (1.0 - MJSW);
      *cap += czbsw * sarg;
    }
    if (czbswg > 0.0) {
      arg = 1.0 - v / in->PhiBSWGTemp;
      if (MJSWG == 0.5) sarg = 1.0 / sqrt(arg);
      else sarg = exp(-MJSWG * log(arg));
      *q += in->PhiBSWGTemp * czbswg * (1.0 - arg * sarg) / (1.0 - MJSWG);
      *cap += czbswg * sarg;
    }
  } else {
    T0 = czb + czbsw + czbswg;
    T1 = v * (czb * MJ / in->PhiBTemp + czbsw * MJSW / in->PhiBSWTemp +
              czbswg * MJSWG / in->PhiBSWGTemp);
    *q = v * (T0 + 0.5 * T1);
    *cap = T0 + T1;
  }
}

/* Channel charge and its capacitances for capMod 0 (N_DEV_MOSFET_B3.C:
   6005-6308): the long-channel model on the flat-band voltage vfbcv.  The
   body bias is Vbs when Vbseff < 0, else phi - Phis.  Accumulation
   (Arg1 <= 0) gives qgate = CoxWL*Arg1; depletion (Vgst <= 0) the
   square-root body charge; inversion the saturation/linear forms of the
   0/100 (xpart > 0.5), 40/60 (xpart < 0.5) or 50/50 drain/source partition.
   qsrc is left alone here: the companion step rebuilds it from the others. */
static void b3_charge_cap0(ref_b3_model *m, ref_b3_sdp *p, ref_b3_inst *in,
                           double Vbs, double Vbseff, double Phis,
                           double dPhis_dVb, double sqrtPhis,
                           double dsqrtPhis_dVb, double Vgs_eff,
                           double dVgs_eff_dVg, double Vds, double Abulk0,
                           double dAbulk0_dVb) {
  double vb, dvb_dVb;
  if (Vbseff < 0.0) {
    vb = Vbs;
    dvb_dVb = 1.0;
  } else {
    vb = p->phi - Phis;
    dvb_dVb = -dPhis_dVb;
  }
  const double Vfb = p->vfbcv;
  const double Vth = Vfb + p->phi + p->k1ox * sqrtPhis;
  const double Vgst = Vgs_eff - Vth;
  const double dVth_dVb = p->k1ox * dsqrtPhis_dVb;
  const double CoxWL = m->cox * p->weffCV * p->leffCV;
  const double Arg1 = Vgs_eff - vb - Vfb;

  in->cgdb = in->cdgb = in->cddb = in->cdsb = in->cbdb = 0.0;
  if (Arg1 <= 0.0) {
    in->qgate = CoxWL * Arg1;
    in->qbulk = -in->qgate;
    in->qdrn = 0.0;
    in->cggb = CoxWL * dVgs_eff_dVg;
    in->cgsb = CoxWL * (dvb_dVb - dVgs_eff_dVg);
    in->cbgb = -CoxWL * dVgs_eff_dVg;
    in->cbsb = -in->cgsb;
    in->qinv = 0.0;
    return;
  }
  if (Vgst <= 0.0) {
    const double h = 0.5 * p->k1ox;
    const double r = sqrt(h * h + Arg1);
    in->qgate = CoxWL * p->k1ox * (r - h);
    in->qbulk = -in->qgate;
    in->qdrn = 0.0;
    const double c = CoxWL * h / r;
    in->cggb = c * dVgs_eff_dVg;
    in->cgsb = c * (dvb_dVb - dVgs_eff_dVg);
    in->cbgb = -in->cggb;
    in->cbsb = -in->cgsb;
    in->qinv = 0.0;
    return;
  }

  const double c3 = CoxWL / 3.0, c23 = 2.0 * c3;
  const double AbulkCV = Abulk0 * p->abulkCVfactor;
  const double dAbulkCV_dVb = p->abulkCVfactor * dAbulk0_dVb;
  const double Vdsat = Vgst / AbulkCV;
  const double dVdsat_dVg = dVgs_eff_dVg / AbulkCV;
  const double dVdsat_dVb = -(Vdsat * dAbulkCV_dVb + dVth_dVb) / AbulkCV;
  const int part = (m->xpart > 0.5) ? 0 : (m->xpart < 0.5) ? 1 : 2;

  if (Vdsat <= Vds) {
    /* saturation: the gate and bulk charges do not depend on the split */
    in->qgate = CoxWL * (Vgs_eff - Vfb - p->phi - Vdsat / 3.0);
    const double qi = -c23 * Vgst;
    in->qbulk = -(in->qgate + qi);
    in->cggb = c3 * (3.0 - dVdsat_dVg) * dVgs_eff_dVg;
    const double gb = -c3 * dVdsat_dVb;
    in->cgsb = -(in->cggb + gb);
    in->cbgb = -(in->cggb - c23 * dVgs_eff_dVg);
    const double bb = -(gb + c23 * dVth_dVb);
    in->cbsb = -(in->cbgb + bb);
    if (part == 0) {
      in->qdrn = 0.0;
    } else if (part == 1) {
      in->qdrn = 0.4 * qi;
      const double k = 0.4 * c23;
      in->cdgb = -k * dVgs_eff_dVg;
      in->cdsb = -(k * dVth_dVb + in->cdgb);
    } else {
      in->qdrn = 0.5 * qi;
      in->cdgb = -c3 * dVgs_eff_dVg;
      in->cdsb = -(c3 * dVth_dVb + in->cdgb);
    }
    in->qinv = -(in->qgate + in->qbulk);
    return;
  }

  /* linear region */
  const double Alphaz = Vgst / Vdsat;
  const double T1 = 2.0 * Vdsat - Vds;
  const double T2 = Vds / (3.0 * T1);
  const double T3 = T2 * Vds;
  const double q4 = 0.25 * CoxWL;
  const double T4 = q4 * Alphaz;
  in->qgate = CoxWL * (Vgs_eff - Vfb - p->phi - 0.5 * (Vds - T3));
  const double T5 = T3 / T1;
  in->cggb = CoxWL * (1.0 - T5 * dVdsat_dVg) * dVgs_eff_dVg;
  const double gb = -CoxWL * T5 * dVdsat_dVb;
  in->cgdb = CoxWL * (T2 - 0.5 + 0.5 * T5);
  if (part == 0) in->cgsb = -(in->cggb + gb + in->cgdb);
  else in->cgsb = -(in->cggb + in->cgdb + gb);
  const double rV = 1.0 / Vdsat;
  const double dAz_dVg = rV * (1.0 - Alphaz * dVdsat_dVg);
  const double dAz_dVb = -rV * (dVth_dVb + Alphaz * dVdsat_dVb);

  if (part == 0) {
    const double a = 2.0 * Vds - T1 - 3.0 * T3;
    const double b = T3 - T1 - 2.0 * Vds;
    in->qdrn = T4 * a;
    in->qbulk = -(in->qgate + in->qdrn + T4 * b);
    const double qa = q4 * a, qb = q4 * b;
    double k = 2.0 * T4 * (1.0 - 3.0 * T5);
    in->cdgb = (qa * dAz_dVg - k * dVdsat_dVg) * dVgs_eff_dVg;
    const double db = qa * dAz_dVb - k * dVdsat_dVb;
    in->cddb = T4 * (3.0 - 6.0 * T2 - 3.0 * T5);
    in->cdsb = -(in->cdgb + db + in->cddb);
    k = 2.0 * T4 * (1.0 + T5);
    const double bg = (qb * dAz_dVg - k * dVdsat_dVg) * dVgs_eff_dVg;
    const double bbv = qb * dAz_dVb - k * dVdsat_dVb;
    const double bd = T4 * (2.0 * T2 + T5 - 1.0);
    const double bs = -(bg + bbv + bd);
    in->cbgb = -(in->cggb + in->cdgb + bg);
    in->cbdb = -(in->cgdb + in->cddb + bd);
    in->cbsb = -(in->cgsb + in->cdsb + bs);
  } else if (part == 1) {
    const double w = 8.0 * Vdsat * Vdsat - 6.0 * Vdsat * Vds + 1.2 * Vds * Vds;
    const double u = T2 / T1;
    double a = Vds - T1 - u * w;
    in->qdrn = T4 * a;
    a *= q4;
    const double v = u / T1;
    const double k = T4 * (2.0 - 4.0 * v * w + u * (16.0 * Vdsat - 6.0 * Vds));
    in->cdgb = (a * dAz_dVg - k * dVdsat_dVg) * dVgs_eff_dVg;
    const double db = a * dAz_dVb - k * dVdsat_dVb;
    in->cddb = T4 * (2.0 - (1.0 / (3.0 * T1 * T1) + 2.0 * v) * w +
                     u * (6.0 * Vdsat - 2.4 * Vds));
    in->cdsb = -(in->cdgb + db + in->cddb);
    double e = 2.0 * (T1 + T3);
    in->qbulk = -(in->qgate - T4 * e);
    e *= q4;
    const double f = 4.0 * T4 * (1.0 - T5);
    const double bg = (-e * dAz_dVg - in->cdgb - f * dVdsat_dVg) * dVgs_eff_dVg;
    const double bbv = -e * dAz_dVb - db - f * dVdsat_dVb;
    const double bd = -4.0 * T4 * (T2 - 0.5 + 0.5 * T5) - in->cddb;
    const double bs = -(bd + bbv + bg);
    in->cbgb = -(in->cggb + in->cdgb + bg);
    /* the Vb term, as N_DEV_MOSFET_B3.C:6238 has it */
    in->cbdb = -(in->cgdb + in->cddb + bbv);
    in->cbsb = -(in->cgsb + in->cdsb + bs);
  } else {
    double a = T1 + T3;
    in->qdrn = -T4 * a;
    in->qbulk = -(in->qgate + in->qdrn + in->qdrn);
    a *= q4;
    const double k = T4 * (2.0 * T5 - 2.0);
    in->cdgb = (k * dVdsat_dVg - a * dAz_dVg) * dVgs_eff_dVg;
    const double db = k * dVdsat_dVb - a * dAz_dVb;
    in->cddb = T4 * (1.0 - 2.0 * T2 - T5);
    in->cdsb = -(in->cdgb + db + in->cddb);
    in->cbgb = -(in->cggb + 2.0 * in->cdgb);
    in->cbdb = -(in->cgdb + 2.0 * in->cddb);
    in->cbsb = -(in->cgsb + 2.0 * in->cdsb);
  }
  in->qinv = -(in->qgate + in->qbulk);
}

/* Channel charge for capMod 1 (N_DEV_MOSFET_B3.C:6360-6520) and capMod 2
   (6521-6685), on the smoothed Vgsteff of the capMod > 0 preamble and the
   flat-band voltage vfbzb.  capMod 1 is piecewise: a body charge from
   Arg1 = Vgs - Vbs - Vfb - Vgsteff (accumulation when <= 0), then the
   inversion charge in saturation (VdsatCV < Vds) or the linear form.
   capMod 2 is smooth: an accumulation charge Qac0 on the smoothed
   Vfbeff, a body charge Qsub0, and the inversion charge on the smoothed
   VdseffCV.  xpart selects the 0/100, 40/60 or 50/50 source share. */
static void b3_charge_cap12(ref_b3_model *m, ref_b3_sdp *p, ref_b3_inst *in,
                            int capmod, double VbseffCV, double dVbseffCV_dVb,
                            double dVbseff_dVb, double CoxWL, double Vgs_eff,
                            double dVgs_eff_dVg, double Vds, double Abulk0,
                            double dAbulk0_dVb, double Vgsteff,
                            double dVgsteff_dVg, double dVgsteff_dVd,
                            double dVgsteff_dVb) {
  const double k1ox = p->k1ox;
  const double AbulkCV = Abulk0 * p->abulkCVfactor;
  const double dAbulkCV_dVb = p->abulkCVfactor * dAbulk0_dVb;
  const double VdsatCV = Vgsteff / AbulkCV;
  const int part = (m->xpart > 0.5) ? 0 : (m->xpart < 0.5) ? 1 : 2;
  double qgate, qbulk, qsrc;
  double Cgg, Cgd, Cgb, Cbg, Cbd, Cbb, Csg, Csd, Csb;
  double Cgg1, Cgd1, Cgb1, Cbg1, Cbd1, Cbb1;
  double T0, T1, T2, T3, T4, T5, T6, T7;

  if (capmod == 1) {
    const double Arg1 = Vgs_eff - VbseffCV - p->vfbzb - Vgsteff;
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
    Cgb *= dVbseff_dVb;
    Cbb *= dVbseff_dVb;
    Csb *= dVbseff_dVb;
    in->qinv = -(qgate + qbulk);
  } else {
    const double vfbzb = p->vfbzb;
    const double V3 = vfbzb - Vgs_eff + VbseffCV - RC_DELTA_3;
    if (vfbzb <= 0.0) T0 = sqrt(V3 * V3 - 4.0 * RC_DELTA_3 * vfbzb);
    else T0 = sqrt(V3 * V3 + 4.0 * RC_DELTA_3 * vfbzb);
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

    const double V4 = VdsatCV - Vds - RC_DELTA_4;
    T0 = sqrt(V4 * V4 + 4.0 * RC_DELTA_4 * VdsatCV);
    double VdseffCV = VdsatCV - 0.5 * (V4 + T0);
    T1 = 0.5 * (1.0 + V4 / T0);
    T2 = RC_DELTA_4 / T0;
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
    in->qinv = -CoxWL * (Vgsteff - 0.5 * T0 + AbulkCV * T3);
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
    Cgb *= dVbseff_dVb;
    Cbb *= dVbseff_dVb;
    Csb *= dVbseff_dVb;
  }
  in->qgate = qgate;
  in->qbulk = qbulk;
  in->qsrc = qsrc;
  in->qdrn = -(qgate + qbulk + qsrc);
  in->cggb = Cgg;
  in->cgsb = -(Cgg + Cgd + Cgb);
  in->cgdb = Cgd;
  in->cdgb = -(Cgg + Cbg + Csg);
  in->cdsb = (Cgg + Cgd + Cgb + Cbg + Cbd + Cbb + Csg + Csd + Csb);
  in->cddb = -(Cgd + Cbd + Csd);
  in->cbgb = Cbg;
  in->cbsb = -(Cbg + Cbd + Cbb);
  in->cbdb = Cbd;
}

/* Gate overlap charge and capacitance of one side for capMod 0 (constant
   cgxo, N_DEV_MOSFET_B3.C:7164-7187) and capMod 1 (7188-7215: the bias-
   dependent overlap cgxl over weffCV, sqrt form for v < 0, linear above). */
static void b3_overlap_cap01(ref_b3_model *m, ref_b3_sdp *p, double v,
                             double cgxo, double cgxl, double *c, double *q) {
  if (m->capMod == 0) {
    *c = cgxo;
    *q = cgxo * v;
  } else if (v < 0.0) {
    const double r = sqrt(1.0 - 4.0 * v / p->ckappa);
    *c = cgxo + p->weffCV * cgxl / r;
    *q = cgxo * v - p->weffCV * 0.5 * cgxl * p->ckappa * (r - 1.0);
  } else {
    *c = cgxo + p->weffCV * cgxl;
    *q = (p->weffCV * cgxl + cgxo) * v;
  }
}

/* Instance::updateIntermediateVars, N_DEV_MOSFET_B3.C:4526-7273, capMod
   0-3, nqsMod 0, no homotopy.  Returns 0, or -1 for a model this restatement
   does not cover. */
static int b3_uiv(ref_b3_master *M, ref_b3_inst *in) {
  ref_b3_model *m = in->model;
  ref_b3_sdp *p = in->paramPtr;
  const ref_solver_state *ss = &M->ss;
  double gmin = M->opt.gmin;
  double SourceSatCurrent, DrainSatCurrent, vgdo, VgstNVt, ExpVgst;
  double czbd, czbdsw, czbdswg, czbs, czbssw, czbsswg, evbd, evbs;
  double Vfbeff, dVfbeff_dVg, dVfbeff_dVb, V3, V4;
  double Vds, Vgs, Vbs, Vgs_eff, dVgs_eff_dVg;
  double Phis, dPhis_dVb, sqrtPhis, dsqrtPhis_dVb, Vth, dVth_dVb, dVth_dVd;
  double Vgst, Nvtm, Vtm, n, dn_dVb, dn_dVd, voffcv, noff, dnoff_dVd,
      dnoff_dVb;
  double ExpArg, V0, CoxWLcen, QovCox, LINK, DeltaPhi, dDeltaPhi_dVg,
      dDeltaPhi_dVd, dDeltaPhi_dVb;
  double Cox, Tox, Tcen, dTcen_dVg, dTcen_dVd, dTcen_dVb, Ccen, Coxeff,
      dCoxeff_dVg, dCoxeff_dVd, dCoxeff_dVb;
  double Denomi, dDenomi_dVg, dDenomi_dVd, dDenomi_dVb, dueff_dVg, dueff_dVd,
      dueff_dVb, Esat, Vdsat, EsatL, dEsatL_dVg, dEsatL_dVd, dEsatL_dVb;
  double dVdsat_dVg, dVdsat_dVb, dVdsat_dVd, Vasat, dVasat_dVg, dVasat_dVb,
      dVasat_dVd, Va, dVa_dVd, dVa_dVg, dVa_dVb;
  double Vbseff, dVbseff_dVb, VbseffCV, dVbseffCV_dVb, CoxWL;
  double T0, dT0_dVg, dT0_dVd, dT0_dVb, T1, dT1_dVg, dT1_dVd, dT1_dVb, T2,
      dT2_dVg, dT2_dVd, dT2_dVb, T3, dT3_dVg, dT3_dVd, dT3_dVb, T4, T5, T6,
      T7, T8, T9, T10, T11, T12;
  double tmp, Abulk, dAbulk_dVb, Abulk0, dAbulk0_dVb, VACLM, dVACLM_dVg,
      dVACLM_dVd, dVACLM_dVb, VADIBL, dVADIBL_dVg, dVADIBL_dVd, dVADIBL_dVb;
  double Xdep, dXdep_dVb, lt1, dlt1_dVb, ltw, dltw_dVb, Delt_vth,
      dDelt_vth_dVb, Theta0, dTheta0_dVb, TempRatio, tmp1, tmp2, tmp3, tmp4;
  double DIBL_Sft, dDIBL_Sft_dVd, Lambda, dLambda_dVg, a1;
  double Vgsteff, dVgsteff_dVg, dVgsteff_dVd, dVgsteff_dVb, Vdseff,
      dVdseff_dVg, dVdseff_dVd, dVdseff_dVb, VdseffCV, dVdseffCV_dVg,
      dVdseffCV_dVd, dVdseffCV_dVb, diffVds, dAbulk_dVg;
  double beta, dbeta_dVg, dbeta_dVd, dbeta_dVb, gche, dgche_dVg, dgche_dVd,
      dgche_dVb, fgche1, dfgche1_dVg, dfgche1_dVd, dfgche1_dVb, fgche2,
      dfgche2_dVg, dfgche2_dVd, dfgche2_dVb;
  double Idl, dIdl_dVg, dIdl_dVd, dIdl_dVb, Idsa, dIdsa_dVg, dIdsa_dVd,
      dIdsa_dVb, Ids, Gm, Gds, Gmb, Isub, Gbd, Gbg, Gbb;
  double VASCBE, dVASCBE_dVg, dVASCBE_dVd, dVASCBE_dVb, CoxWovL, Rds,
      dRds_dVg, dRds_dVb, WVCox, WVCoxRds, Vgst2Vtm, VdsatCV, Leff, Weff,
      dWeff_dVg, dWeff_dVb, AbulkCV, dAbulkCV_dVb;
  double Cgg1, Cgb1, Cgd1, Cbg1, Cbb1, Cbd1, Qac0, Qsub0, dQac0_dVg,
      dQac0_dVb, dQsub0_dVg, dQsub0_dVd, dQsub0_dVb, von_local;
  double Cgg, Cgd, Cgb, Cbg, Cbd, Cbb, Csg, Csd, Csb, qinoi;
  double MJ, MJSW, MJSWG, arg, sarg;
  int Check = 1;
  (void)MJ; (void)MJSW; (void)MJSWG; (void)arg; (void)sarg;

  if (m->capMod != 0 && m->capMod != 1 && m->capMod != 2 && m->capMod != 3)
    return -1;

  in->ChargeComputationNeeded = (ss->tranopFlag || ss->acopFlag || ss->transientFlag) ? 1 : 0;

  in->Vd = M->nextSol[in->li_Drain];
  in->Vg = M->nextSol[in->li_Gate];
  in->Vs = M->nextSol[in->li_Source];
  in->Vb = M->nextSol[in->li_Bulk];
  in->Vsp = M->nextSol[in->li_SourcePrime];
  in->Vdp = M->nextSol[in->li_DrainPrime];
  in->Vddp = in->Vd - in->Vdp;
  in->Vssp = in->Vs - in->Vsp;
  in->Vbsp = in->Vb - in->Vsp;
  in->Vbdp = in->Vb - in->Vdp;
  in->Vgsp = in->Vg - in->Vsp;
  in->Vgdp = in->Vg - in->Vdp;
  in->Vgb = in->Vg - in->Vb;
  in->Vdpsp = in->Vdp - in->Vsp;

  in->vbs = m->dtype * in->Vbsp;
  in->vgs = m->dtype * in->Vgsp;
  in->vds = m->dtype * in->Vdpsp;
  in->qdef = m->dtype * 0.0;
  in->vbd = in->vbs - in->vds;
  in->vgd = in->vgs - in->vds;
  in->origFlag = 1;
  in->limitedFlag = 0;
  in->vgs_orig = in->vgs;
  in->vds_orig = in->vds;
  in->vbs_orig = in->vbs;
  in->vbd_orig = in->vbd;
  in->vgd_orig = in->vgd;

  if (ss->initJctFlag && !in->OFF && M->opt.voltageLimiterFlag) {
    in->vbs = 0.0;
    in->vgs = m->dtype * p->vth0 + 0.1;
    in->vds = 0.1;
    in->origFlag = 0;
    in->vbd = in->vbs - in->vds;
    in->vgd = in->vgs - in->vds;
  } else if ((ss->initFixFlag || ss->initJctFlag) && in->OFF) {
    in->qdef = in->vbs = in->vgs = in->vds = 0;
  }

  if (ss->newtonIter == 0) {
    if (!ss->dcopFlag || (ss->locaEnabledFlag && ss->dcopFlag)) {
      in->vbs_old = M->currSto[in->li_store_vbs];
      in->vbd_old = M->currSto[in->li_store_vbd];
      in->vgs_old = M->currSto[in->li_store_vgs];
      in->vds_old = M->currSto[in->li_store_vds];
      von_local = M->currSto[in->li_store_von];
    } else {
      in->vbs_old = in->vbs;
      in->vbd_old = in->vbd;
      in->vgs_old = in->vgs;
      in->vds_old = in->vds;
      von_local = 0.0;
    }
  } else {
    in->vbs_old = M->nextSto[in->li_store_vbs];
    in->vbd_old = M->nextSto[in->li_store_vbd];
    in->vgs_old = M->nextSto[in->li_store_vgs];
    in->vds_old = M->nextSto[in->li_store_vds];
    von_local = M->nextSto[in->li_store_von];
  }
  vgdo = in->vgs_old - in->vds_old;

  if (M->opt.voltageLimiterFlag && !(ss->initFixFlag && in->OFF)) {
    if (ss->newtonIter >= 0) {
      if (in->vds_old >= 0.0) {
        in->vgs = ref_fetlim(in->vgs, in->vgs_old, von_local);
        in->vds = in->vgs - in->vgd;
        in->vds = ref_limvds(in->vds, in->vds_old);
        in->vgd = in->vgs - in->vds;
      } else {
        in->vgd = ref_fetlim(in->vgd, vgdo, von_local);
        in->vds = in->vgs - in->vgd;
        in->vds = -ref_limvds(-in->vds, -in->vds_old);
        in->vgs = in->vgd + in->vds;
      }
      if (in->vds >= 0.0) {
        in->vbs = ref_pnjlim(in->vbs, in->vbs_old, RC_vt0, m->vcrit, &Check);
        in->vbd = in->vbs - in->vds;
      } else {
        in->vbd = ref_pnjlim(in->vbd, in->vbd_old, RC_vt0, m->vcrit, &Check);
        in->vbs = in->vbd + in->vds;
      }
    }
    if (in->vgs_orig != in->vgs || in->vds_orig != in->vds ||
        in->vbs_orig != in->vbs || in->vbd_orig != in->vbd ||
        in->vgd_orig != in->vgd)
      in->origFlag = 0;
    if (Check == 1) in->limitedFlag = 1;
  }

  in->vbd = in->vbs - in->vds;
  in->vgd = in->vgs - in->vds;
  in->vgb = in->vgs - in->vbs;

  Nvtm = in->vtm * m->jctEmissionCoeff;
  if ((in->sourceArea <= 0.0) && (in->sourcePerimeter <= 0.0)) SourceSatCurrent = 1.0e-14;
  else SourceSatCurrent = in->sourceArea * in->jctTempSatCurDensity +
                          in->sourcePerimeter * in->jctSidewallTempSatCurDensity;
  if (SourceSatCurrent <= 0.0) {
    in->gbs = gmin;
    in->cbs = in->gbs * in->vbs;
  } else if (m->ijth == 0.0 || in->vbs < in->vjsm) {
    evbs = exp(in->vbs / Nvtm);
    in->gbs = SourceSatCurrent * evbs / Nvtm + gmin;
    in->cbs = SourceSatCurrent * (evbs - 1.0) + gmin * in->vbs;
  } else {
    T0 = in->IsEvjsm / Nvtm;
    in->gbs = T0 + gmin;
    in->cbs = in->IsEvjsm - SourceSatCurrent + T0 * (in->vbs - in->vjsm) + gmin * in->vbs;
  }
  if ((in->drainArea <= 0.0) && (in->drainPerimeter <= 0.0)) DrainSatCurrent = 1.0e-14;
  else DrainSatCurrent = in->drainArea * in->jctTempSatCurDensity +
                         in->drainPerimeter * in->jctSidewallTempSatCurDensity;
  if (DrainSatCurrent <= 0.0) {
    in->gbd = gmin;
    in->cbd = in->gbd * in->vbd;
  } else if (m->ijth == 0.0 || in->vbd < in->vjdm) {
    evbd = exp(in->vbd / Nvtm);
    in->gbd = DrainSatCurrent * evbd / Nvtm + gmin;
    in->cbd = DrainSatCurrent * (evbd - 1.0) + gmin * in->vbd;
  } else {
    T0 = in->IsEvjdm / Nvtm;
    in->gbd = T0 + gmin;
    in->cbd = in->IsEvjdm - DrainSatCurrent + T0 * (in->vbd - in->vjdm) + gmin * in->vbd;
  }

  if (in->vds >= 0.0) {
    in->mode = 1;
    Vds = in->vds;
    Vgs = in->vgs;
    Vbs = in->vbs;
  } else {
    in->mode = -1;
    Vds = -in->vds;
    Vgs = in->vgd;
    Vbs = in->vbd;
  }

  T0 = Vbs - p->vbsc - 0.001;
  in->T1global = sqrt(T0 * T0 - 0.004 * p->vbsc);
  Vbseff = p->vbsc + 0.5 * (T0 + in->T1global);
  dVbseff_dVb = 0.5 * (1.0 + T0 / in->T1global);
  if (Vbseff < Vbs) Vbseff = Vbs;

  if (Vbseff > 0.0) {
    T0 = p->phi / (p->phi + Vbseff);
    Phis = p->phi * T0;
    dPhis_dVb = -T0 * T0;
    sqrtPhis = p->phis3 / (p->phi + 0.5 * Vbseff);
    dsqrtPhis_dVb = -0.5 * sqrtPhis * sqrtPhis / p->phis3;
  } else {
    Phis = p->phi - Vbseff;
    dPhis_dVb = -1.0;
    sqrtPhis = sqrt(Phis);
    dsqrtPhis_dVb = -0.5 / sqrtPhis;
  }
  Xdep = p->Xdep0 * sqrtPhis / p->sqrtPhi;
  dXdep_dVb = (p->Xdep0 / p->sqrtPhi) * dsqrtPhis_dVb;
  Leff = p->leff;
  Vtm = in->vtm;
  T3 = sqrt(Xdep);
  V0 = p->vbi - p->phi;
  T0 = p->dvt2 * Vbseff;
  if (T0 >= -0.5) {
    T1 = 1.0 + T0;
    T2 = p->dvt2;
  } else {
    T4 = 1.0 / (3.0 + 8.0 * T0);
    T1 = (1.0 + 3.0 * T0) * T4;
    T2 = p->dvt2 * T4 * T4;
  }
  lt1 = m->factor1 * T3 * T1;
  dlt1_dVb = m->factor1 * (0.5 / T3 * T1 * dXdep_dVb + T3 * T2);
  T0 = p->dvt2w * Vbseff;
  if (T0 >= -0.5) {
    T1 = 1.0 + T0;
    T2 = p->dvt2w;
  } else {
    T4 = 1.0 / (3.0 + 8.0 * T0);
    T1 = (1.0 + 3.0 * T0) * T4;
    T2 = p->dvt2w * T4 * T4;
  }
  ltw = m->factor1 * T3 * T1;
  dltw_dVb = m->factor1 * (0.5 / T3 * T1 * dXdep_dVb + T3 * T2);
  T0 = -0.5 * p->dvt1 * Leff / lt1;
  if (T0 > -RC_EXP_THRESHOLD) {
    T1 = exp(T0);
    Theta0 = T1 * (1.0 + 2.0 * T1);
    dT1_dVb = -T0 / lt1 * T1 * dlt1_dVb;
    dTheta0_dVb = (1.0 + 4.0 * T1) * dT1_dVb;
  } else {
    T1 = RC_MIN_EXP;
    Theta0 = T1 * (1.0 + 2.0 * T1);
    dTheta0_dVb = 0.0;
  }
  in->thetavth = p->dvt0 * Theta0;
  Delt_vth = in->thetavth * V0;
  dDelt_vth_dVb = p->dvt0 * dTheta0_dVb * V0;
  T0 = -0.5 * p->dvt1w * p->weff * Leff / ltw;
  if (T0 > -RC_EXP_THRESHOLD) {
    T1 = exp(T0);
    T2 = T1 * (1.0 + 2.0 * T1);
    dT1_dVb = -T0 / ltw * T1 * dltw_dVb;
    dT2_dVb = (1.0 + 4.0 * T1) * dT1_dVb;
  } else {
    T1 = RC_MIN_EXP;
    T2 = T1 * (1.0 + 2.0 * T1);
    dT2_dVb = 0.0;
  }
  T0 = p->dvt0w * T2;
  T2 = T0 * V0;
  dT2_dVb = p->dvt0w * dT2_dVb * V0;
  TempRatio = in->temp / m->tnom - 1.0;
  T0 = sqrt(1.0 + p->nlx / Leff);
  T1 = p->k1ox * (T0 - 1.0) * p->sqrtPhi +
       (p->kt1 + p->kt1l / Leff + p->kt2 * Vbseff) * TempRatio;
  tmp2 = m->tox * p->phi / (p->weff + p->w0);
  T3 = p->eta0 + p->etab * Vbseff;
  if (T3 < 1.0e-4) {
    T9 = 1.0 / (3.0 - 2.0e4 * T3);
    T3 = (2.0e-4 - T3) * T9;
    T4 = T9 * T9;
  } else {
    T4 = 1.0;
  }
  dDIBL_Sft_dVd = T3 * p->theta0vb0;
  DIBL_Sft = dDIBL_Sft_dVd * Vds;
  Vth = m->dtype * p->vth0 - p->k1 * p->sqrtPhi + p->k1ox * sqrtPhis -
        p->k2ox * Vbseff - Delt_vth - T2 + (p->k3 + p->k3b * Vbseff) * tmp2 +
        T1 - DIBL_Sft;
  in->von = Vth;
  dVth_dVb = p->k1ox * dsqrtPhis_dVb - p->k2ox - dDelt_vth_dVb - dT2_dVb +
             p->k3b * tmp2 - p->etab * Vds * p->theta0vb0 * T4 +
             p->kt2 * TempRatio;
  dVth_dVd = -dDIBL_Sft_dVd;
  tmp2 = p->nfactor * RC_EPSSI / Xdep;
  tmp3 = p->cdsc + p->cdscb * Vbseff + p->cdscd * Vds;
  tmp4 = (tmp2 + tmp3 * Theta0 + p->cit) / m->cox;
  if (tmp4 >= -0.5) {
    n = 1.0 + tmp4;
    dn_dVb = (-tmp2 / Xdep * dXdep_dVb + tmp3 * dTheta0_dVb + p->cdscb * Theta0) / m->cox;
    dn_dVd = p->cdscd * Theta0 / m->cox;
  } else {
    T0 = 1.0 / (3.0 + 8.0 * tmp4);
    n = (1.0 + 3.0 * tmp4) * T0;
    T0 *= T0;
    dn_dVb = (-tmp2 / Xdep * dXdep_dVb + tmp3 * dTheta0_dVb + p->cdscb * Theta0) / m->cox * T0;
    dn_dVd = p->cdscd * Theta0 / m->cox * T0;
  }
  T0 = p->vfb + p->phi;
  if ((p->ngate > 1.e18) && (p->ngate < 1.e25) && (Vgs > T0)) {
    T1 = 1.0e6 * RC_Q * RC_EPSSI * p->ngate / (m->cox * m->cox);
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
  Vgst = Vgs_eff - Vth;
  T10 = 2.0 * n * Vtm;
  VgstNVt = Vgst / T10;
  ExpArg = (2.0 * p->voff - Vgst) / T10;
  if (VgstNVt > RC_EXP_THRESHOLD) {
    Vgsteff = Vgst;
    dVgsteff_dVg = dVgs_eff_dVg;
    dVgsteff_dVd = -dVth_dVd;
    dVgsteff_dVb = -dVth_dVb;
  } else if (ExpArg > RC_EXP_THRESHOLD) {
    T0 = (Vgst - p->voff) / (n * Vtm);
    ExpVgst = exp(T0);
    Vgsteff = Vtm * p->cdep0 / m->cox * ExpVgst;
    dVgsteff_dVg = Vgsteff / (n * Vtm);
    dVgsteff_dVd = -dVgsteff_dVg * (dVth_dVd + T0 * Vtm * dn_dVd);
    dVgsteff_dVb = -dVgsteff_dVg * (dVth_dVb + T0 * Vtm * dn_dVb);
    dVgsteff_dVg *= dVgs_eff_dVg;
  } else {
    ExpVgst = exp(VgstNVt);
    T1 = T10 * log(1.0 + ExpVgst);
    dT1_dVg = ExpVgst / (1.0 + ExpVgst);
    dT1_dVb = -dT1_dVg * (dVth_dVb + Vgst / n * dn_dVb) + T1 / n * dn_dVb;
    dT1_dVd = -dT1_dVg * (dVth_dVd + Vgst / n * dn_dVd) + T1 / n * dn_dVd;
    dT2_dVg = -m->cox / (Vtm * p->cdep0) * exp(ExpArg);
    T2 = 1.0 - T10 * dT2_dVg;
    dT2_dVd = -dT2_dVg * (dVth_dVd - 2.0 * Vtm * ExpArg * dn_dVd) + (T2 - 1.0) / n * dn_dVd;
    dT2_dVb = -dT2_dVg * (dVth_dVb - 2.0 * Vtm * ExpArg * dn_dVb) + (T2 - 1.0) / n * dn_dVb;
    Vgsteff = T1 / T2;
    T3 = T2 * T2;
    dVgsteff_dVg = (T2 * dT1_dVg - T1 * dT2_dVg) / T3 * dVgs_eff_dVg;
    dVgsteff_dVd = (T2 * dT1_dVd - T1 * dT2_dVd) / T3;
    dVgsteff_dVb = (T2 * dT1_dVb - T1 * dT2_dVb) / T3;
  }
  T9 = sqrtPhis - p->sqrtPhi;
  Weff = p->weff - 2.0 * (p->dwg * Vgsteff + p->dwb * T9);
  dWeff_dVg = -2.0 * p->dwg;
  dWeff_dVb = -2.0 * p->dwb * dsqrtPhis_dVb;
  if (Weff < 2.0e-8) {
    T0 = 1.0 / (6.0e-8 - 2.0 * Weff);
    Weff = 2.0e-8 * (4.0e-8 - Weff) * T0;
    T0 *= T0 * 4.0e-16;
    dWeff_dVg *= T0;
    dWeff_dVb *= T0;
  }
  T0 = p->prwg * Vgsteff + p->prwb * T9;
  if (T0 >= -0.9) {
    Rds = p->rds0 * (1.0 + T0);
    dRds_dVg = p->rds0 * p->prwg;
    dRds_dVb = p->rds0 * p->prwb * dsqrtPhis_dVb;
  } else {
    T1 = 1.0 / (17.0 + 20.0 * T0);
    Rds = p->rds0 * (0.8 + T0) * T1;
    T1 *= T1;
    dRds_dVg = p->rds0 * p->prwg * T1;
    dRds_dVb = p->rds0 * p->prwb * dsqrtPhis_dVb * T1;
  }
  T1 = 0.5 * p->k1ox / sqrtPhis;
  dT1_dVb = -T1 / sqrtPhis * dsqrtPhis_dVb;
  T9 = sqrt(p->xj * Xdep);
  tmp1 = Leff + 2.0 * T9;
  T5 = Leff / tmp1;
  tmp2 = p->a0 * T5;
  tmp3 = p->weff + p->b1;
  tmp4 = p->b0 / tmp3;
  T2 = tmp2 + tmp4;
  dT2_dVb = -T9 / tmp1 / Xdep * dXdep_dVb;
  T6 = T5 * T5;
  T7 = T5 * T6;
  Abulk0 = 1.0 + T1 * T2;
  dAbulk0_dVb = T1 * tmp2 * dT2_dVb + T2 * dT1_dVb;
  T8 = p->ags * p->a0 * T7;
  dAbulk_dVg = -T1 * T8;
  Abulk = Abulk0 + dAbulk_dVg * Vgsteff;
  dAbulk_dVb = dAbulk0_dVb - T8 * Vgsteff * (dT1_dVb + 3.0 * T1 * dT2_dVb);
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
  T2 = p->keta * Vbseff;
  if (T2 >= -0.9) {
    T0 = 1.0 / (1.0 + T2);
    dT0_dVb = -p->keta * T0 * T0;
  } else {
    T1 = 1.0 / (0.8 + T2);
    T0 = (17.0 + 20.0 * T2) * T1;
    dT0_dVb = -p->keta * T1 * T1;
  }
  dAbulk_dVg *= T0;
  dAbulk_dVb = dAbulk_dVb * T0 + Abulk * dT0_dVb;
  dAbulk0_dVb = dAbulk0_dVb * T0 + Abulk0 * dT0_dVb;
  Abulk *= T0;
  Abulk0 *= T0;
  if (m->mobMod == 1) {
    T0 = Vgsteff + Vth + Vth;
    T2 = p->ua + p->uc * Vbseff;
    T3 = T0 / m->tox;
    T5 = T3 * (T2 + p->ub * T3);
    dDenomi_dVg = (T2 + 2.0 * p->ub * T3) / m->tox;
    dDenomi_dVd = dDenomi_dVg * 2.0 * dVth_dVd;
    dDenomi_dVb = dDenomi_dVg * 2.0 * dVth_dVb + p->uc * T3;
  } else if (m->mobMod == 2) {
    T5 = Vgsteff / m->tox * (p->ua + p->uc * Vbseff + p->ub * Vgsteff / m->tox);
    dDenomi_dVg = (p->ua + p->uc * Vbseff + 2.0 * p->ub * Vgsteff / m->tox) / m->tox;
    dDenomi_dVd = 0.0;
    dDenomi_dVb = Vgsteff * p->uc / m->tox;
  } else {
    T0 = Vgsteff + Vth + Vth;
    T2 = 1.0 + p->uc * Vbseff;
    T3 = T0 / m->tox;
    T4 = T3 * (p->ua + p->ub * T3);
    T5 = T4 * T2;
    dDenomi_dVg = (p->ua + 2.0 * p->ub * T3) * T2 / m->tox;
    dDenomi_dVd = dDenomi_dVg * 2.0 * dVth_dVd;
    dDenomi_dVb = dDenomi_dVg * 2.0 * dVth_dVb + p->uc * T4;
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
  in->ueff = p->u0temp / Denomi;
  T9 = -in->ueff / Denomi;
  dueff_dVg = T9 * dDenomi_dVg;
  dueff_dVd = T9 * dDenomi_dVd;
  dueff_dVb = T9 * dDenomi_dVb;
  WVCox = Weff * p->vsattemp * m->cox;
  WVCoxRds = WVCox * Rds;
  Esat = 2.0 * p->vsattemp / in->ueff;
  EsatL = Esat * Leff;
  T0 = -EsatL / in->ueff;
  dEsatL_dVg = T0 * dueff_dVg;
  dEsatL_dVd = T0 * dueff_dVd;
  dEsatL_dVb = T0 * dueff_dVb;
  a1 = p->a1;
  if (a1 == 0.0) {
    Lambda = p->a2;
    dLambda_dVg = 0.0;
  } else if (a1 > 0.0) {
    T0 = 1.0 - p->a2;
    T1 = T0 - p->a1 * Vgsteff - 0.0001;
    T2 = sqrt(T1 * T1 + 0.0004 * T0);
    Lambda = p->a2 + T0 - 0.5 * (T1 + T2);
    dLambda_dVg = 0.5 * p->a1 * (1.0 + T1 / T2);
  } else {
    T1 = p->a2 + p->a1 * Vgsteff - 0.0001;
    T2 = sqrt(T1 * T1 + 0.0004 * p->a2);
    Lambda = 0.5 * (T1 + T2);
    dLambda_dVg = 0.5 * p->a1 * (1.0 + T1 / T2);
  }
  Vgst2Vtm = Vgsteff + 2.0 * Vtm;
  if (Rds > 0) {
    tmp2 = dRds_dVg / Rds + dWeff_dVg / Weff;
    tmp3 = dRds_dVb / Rds + dWeff_dVb / Weff;
  } else {
    tmp2 = dWeff_dVg / Weff;
    tmp3 = dWeff_dVb / Weff;
  }
  if ((Rds == 0.0) && (Lambda == 1.0)) {
    T0 = 1.0 / (Abulk * EsatL + Vgst2Vtm);
    tmp1 = 0.0;
    T1 = T0 * T0;
    T2 = Vgst2Vtm * T0;
    T3 = EsatL * Vgst2Vtm;
    Vdsat = T3 * T0;
    dT0_dVg = -(Abulk * dEsatL_dVg + EsatL * dAbulk_dVg + 1.0) * T1;
    dT0_dVd = -(Abulk * dEsatL_dVd) * T1;
    dT0_dVb = -(Abulk * dEsatL_dVb + dAbulk_dVb * EsatL) * T1;
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
    dT0_dVg = 2.0 * (T8 * tmp2 - Abulk * tmp1 + (2.0 * T9 + 1.0 / Lambda - 1.0) * dAbulk_dVg);
    dT0_dVb = 2.0 * (T8 * (2.0 / Abulk * dAbulk_dVb + tmp3) + (1.0 / Lambda - 1.0) * dAbulk_dVb);
    dT0_dVd = 0.0;
    T1 = Vgst2Vtm * (2.0 / Lambda - 1.0) + Abulk * EsatL + 3.0 * T7;
    dT1_dVg = (2.0 / Lambda - 1.0) - 2.0 * Vgst2Vtm * tmp1 + Abulk * dEsatL_dVg +
              EsatL * dAbulk_dVg + 3.0 * (T9 + T7 * tmp2 + T6 * dAbulk_dVg);
    dT1_dVb = Abulk * dEsatL_dVb + EsatL * dAbulk_dVb + 3.0 * (T6 * dAbulk_dVb + T7 * tmp3);
    dT1_dVd = Abulk * dEsatL_dVd;
    T2 = Vgst2Vtm * (EsatL + 2.0 * T6);
    dT2_dVg = EsatL + Vgst2Vtm * dEsatL_dVg + T6 * (4.0 + 2.0 * Vgst2Vtm * tmp2);
    dT2_dVb = Vgst2Vtm * (dEsatL_dVb + 2.0 * T6 * tmp3);
    dT2_dVd = Vgst2Vtm * dEsatL_dVd;
    T3 = sqrt(T1 * T1 - 2.0 * T0 * T2);
    Vdsat = (T1 - T3) / T0;
    dT3_dVg = (T1 * dT1_dVg - 2.0 * (T0 * dT2_dVg + T2 * dT0_dVg)) / T3;
    dT3_dVd = (T1 * dT1_dVd - 2.0 * (T0 * dT2_dVd + T2 * dT0_dVd)) / T3;
    dT3_dVb = (T1 * dT1_dVb - 2.0 * (T0 * dT2_dVb + T2 * dT0_dVb)) / T3;
    (void)dT3_dVg; (void)dT3_dVd; (void)dT3_dVb;
    dVdsat_dVg = (dT1_dVg - (T1 * dT1_dVg - dT0_dVg * T2 - T0 * dT2_dVg) / T3 - Vdsat * dT0_dVg) / T0;
    dVdsat_dVb = (dT1_dVb - (T1 * dT1_dVb - dT0_dVb * T2 - T0 * dT2_dVb) / T3 - Vdsat * dT0_dVb) / T0;
    dVdsat_dVd = (dT1_dVd - (T1 * dT1_dVd - T0 * dT2_dVd) / T3) / T0;
  }
  in->vdsat = Vdsat;
  T1 = Vdsat - Vds - p->delta;
  dT1_dVg = dVdsat_dVg;
  dT1_dVd = dVdsat_dVd - 1.0;
  dT1_dVb = dVdsat_dVb;
  T2 = sqrt(T1 * T1 + 4.0 * p->delta * Vdsat);
  T0 = T1 / T2;
  T3 = 2.0 * p->delta / T2;
  dT2_dVg = T0 * dT1_dVg + T3 * dVdsat_dVg;
  dT2_dVd = T0 * dT1_dVd + T3 * dVdsat_dVd;
  dT2_dVb = T0 * dT1_dVb + T3 * dVdsat_dVb;
  Vdseff = Vdsat - 0.5 * (T1 + T2);
  dVdseff_dVg = dVdsat_dVg - 0.5 * (dT1_dVg + dT2_dVg);
  dVdseff_dVd = dVdsat_dVd - 0.5 * (dT1_dVd + dT2_dVd);
  dVdseff_dVb = dVdsat_dVb - 0.5 * (dT1_dVb + dT2_dVb);
  if (Vds == 0.0) {
    Vdseff = 0.0;
    dVdseff_dVg = 0.0;
    dVdseff_dVb = 0.0;
  }
  tmp4 = 1.0 - 0.5 * Abulk * Vdsat / Vgst2Vtm;
  T9 = WVCoxRds * Vgsteff;
  T8 = T9 / Vgst2Vtm;
  T0 = EsatL + Vdsat + 2.0 * T9 * tmp4;
  T7 = 2.0 * WVCoxRds * tmp4;
  dT0_dVg = dEsatL_dVg + dVdsat_dVg + T7 * (1.0 + tmp2 * Vgsteff) -
            T8 * (Abulk * dVdsat_dVg - Abulk * Vdsat / Vgst2Vtm + Vdsat * dAbulk_dVg);
  dT0_dVb = dEsatL_dVb + dVdsat_dVb + T7 * tmp3 * Vgsteff -
            T8 * (dAbulk_dVb * Vdsat + Abulk * dVdsat_dVb);
  dT0_dVd = dEsatL_dVd + dVdsat_dVd - T8 * Abulk * dVdsat_dVd;
  T9 = WVCoxRds * Abulk;
  T1 = 2.0 / Lambda - 1.0 + T9;
  dT1_dVg = -2.0 * tmp1 + WVCoxRds * (Abulk * tmp2 + dAbulk_dVg);
  dT1_dVb = dAbulk_dVb * WVCoxRds + T9 * tmp3;
  Vasat = T0 / T1;
  dVasat_dVg = (dT0_dVg - Vasat * dT1_dVg) / T1;
  dVasat_dVb = (dT0_dVb - Vasat * dT1_dVb) / T1;
  dVasat_dVd = dT0_dVd / T1;
  if (Vdseff > Vds) Vdseff = Vds;
  diffVds = Vds - Vdseff;
  if ((p->pclm > 0.0) && (diffVds > 1.0e-10)) {
    T0 = 1.0 / (p->pclm * Abulk * p->litl);
    dT0_dVb = -T0 / Abulk * dAbulk_dVb;
    dT0_dVg = -T0 / Abulk * dAbulk_dVg;
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
    VACLM = RC_MAX_EXP;
    dVACLM_dVd = dVACLM_dVg = dVACLM_dVb = 0.0;
  }
  if (p->thetaRout > 0.0) {
    T8 = Abulk * Vdsat;
    T0 = Vgst2Vtm * T8;
    dT0_dVg = Vgst2Vtm * Abulk * dVdsat_dVg + T8 + Vgst2Vtm * Vdsat * dAbulk_dVg;
    dT0_dVb = Vgst2Vtm * (dAbulk_dVb * Vdsat + Abulk * dVdsat_dVb);
    dT0_dVd = Vgst2Vtm * Abulk * dVdsat_dVd;
    T1 = Vgst2Vtm + T8;
    dT1_dVg = 1.0 + Abulk * dVdsat_dVg + Vdsat * dAbulk_dVg;
    dT1_dVb = Abulk * dVdsat_dVb + dAbulk_dVb * Vdsat;
    dT1_dVd = Abulk * dVdsat_dVd;
    T9 = T1 * T1;
    T2 = p->thetaRout;
    VADIBL = (Vgst2Vtm - T0 / T1) / T2;
    dVADIBL_dVg = (1.0 - dT0_dVg / T1 + T0 * dT1_dVg / T9) / T2;
    dVADIBL_dVb = (-dT0_dVb / T1 + T0 * dT1_dVb / T9) / T2;
    dVADIBL_dVd = (-dT0_dVd / T1 + T0 * dT1_dVd / T9) / T2;
    T7 = p->pdiblb * Vbseff;
    if (T7 >= -0.9) {
      T3 = 1.0 / (1.0 + T7);
      VADIBL *= T3;
      dVADIBL_dVg *= T3;
      dVADIBL_dVb = (dVADIBL_dVb - VADIBL * p->pdiblb) * T3;
      dVADIBL_dVd *= T3;
    } else {
      T4 = 1.0 / (0.8 + T7);
      T3 = (17.0 + 20.0 * T7) * T4;
      dVADIBL_dVg *= T3;
      dVADIBL_dVb = dVADIBL_dVb * T3 - VADIBL * p->pdiblb * T4 * T4;
      dVADIBL_dVd *= T3;
      VADIBL *= T3;
    }
  } else {
    VADIBL = RC_MAX_EXP;
    dVADIBL_dVd = dVADIBL_dVg = dVADIBL_dVb = 0.0;
  }
  T8 = p->pvag / EsatL;
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
  Va = Vasat + T0 * T1;
  dVa_dVg = dVasat_dVg + T1 * dT0_dVg + T0 * dT1_dVg;
  dVa_dVd = dVasat_dVd + T1 * dT0_dVd + T0 * dT1_dVd;
  dVa_dVb = dVasat_dVb + T1 * dT0_dVb + T0 * dT1_dVb;
  if (p->pscbe2 > 0.0) {
    if (diffVds > p->pscbe1 * p->litl / RC_EXP_THRESHOLD) {
      T0 = p->pscbe1 * p->litl / diffVds;
      VASCBE = Leff * exp(T0) / p->pscbe2;
      T1 = T0 * VASCBE / diffVds;
      dVASCBE_dVg = T1 * dVdseff_dVg;
      dVASCBE_dVd = -T1 * (1.0 - dVdseff_dVd);
      dVASCBE_dVb = T1 * dVdseff_dVb;
    } else {
      VASCBE = RC_MAX_EXP * Leff / p->pscbe2;
      dVASCBE_dVg = dVASCBE_dVd = dVASCBE_dVb = 0.0;
    }
  } else {
    VASCBE = RC_MAX_EXP;
    dVASCBE_dVg = dVASCBE_dVd = dVASCBE_dVb = 0.0;
  }
  CoxWovL = m->cox * Weff / Leff;
  beta = in->ueff * CoxWovL;
  dbeta_dVg = CoxWovL * dueff_dVg + beta * dWeff_dVg / Weff;
  dbeta_dVd = CoxWovL * dueff_dVd;
  dbeta_dVb = CoxWovL * dueff_dVb + beta * dWeff_dVb / Weff;
  T0 = 1.0 - 0.5 * Abulk * Vdseff / Vgst2Vtm;
  dT0_dVg = -0.5 * (Abulk * dVdseff_dVg - Abulk * Vdseff / Vgst2Vtm + Vdseff * dAbulk_dVg) / Vgst2Vtm;
  dT0_dVd = -0.5 * Abulk * dVdseff_dVd / Vgst2Vtm;
  dT0_dVb = -0.5 * (Abulk * dVdseff_dVb + dAbulk_dVb * Vdseff) / Vgst2Vtm;
  fgche1 = Vgsteff * T0;
  dfgche1_dVg = Vgsteff * dT0_dVg + T0;
  dfgche1_dVd = Vgsteff * dT0_dVd;
  dfgche1_dVb = Vgsteff * dT0_dVb;
  T9 = Vdseff / EsatL;
  fgche2 = 1.0 + T9;
  dfgche2_dVg = (dVdseff_dVg - T9 * dEsatL_dVg) / EsatL;
  dfgche2_dVd = (dVdseff_dVd - T9 * dEsatL_dVd) / EsatL;
  dfgche2_dVb = (dVdseff_dVb - T9 * dEsatL_dVb) / EsatL;
  gche = beta * fgche1 / fgche2;
  dgche_dVg = (beta * dfgche1_dVg + fgche1 * dbeta_dVg - gche * dfgche2_dVg) / fgche2;
  dgche_dVd = (beta * dfgche1_dVd + fgche1 * dbeta_dVd - gche * dfgche2_dVd) / fgche2;
  dgche_dVb = (beta * dfgche1_dVb + fgche1 * dbeta_dVb - gche * dfgche2_dVb) / fgche2;
  T0 = 1.0 + gche * Rds;
  T9 = Vdseff / T0;
  Idl = gche * T9;
  dIdl_dVg = (gche * dVdseff_dVg + T9 * dgche_dVg) / T0 - Idl * gche / T0 * dRds_dVg;
  dIdl_dVd = (gche * dVdseff_dVd + T9 * dgche_dVd) / T0;
  dIdl_dVb = (gche * dVdseff_dVb + T9 * dgche_dVb - Idl * dRds_dVb * gche) / T0;
  T9 = diffVds / Va;
  T0 = 1.0 + T9;
  Idsa = Idl * T0;
  dIdsa_dVg = T0 * dIdl_dVg - Idl * (dVdseff_dVg + T9 * dVa_dVg) / Va;
  dIdsa_dVd = T0 * dIdl_dVd + Idl * (1.0 - dVdseff_dVd - T9 * dVa_dVd) / Va;
  dIdsa_dVb = T0 * dIdl_dVb - Idl * (dVdseff_dVb + T9 * dVa_dVb) / Va;
  T9 = diffVds / VASCBE;
  T0 = 1.0 + T9;
  Ids = Idsa * T0;
  Gm = T0 * dIdsa_dVg - Idsa * (dVdseff_dVg + T9 * dVASCBE_dVg) / VASCBE;
  Gds = T0 * dIdsa_dVd + Idsa * (1.0 - dVdseff_dVd - T9 * dVASCBE_dVd) / VASCBE;
  Gmb = T0 * dIdsa_dVb - Idsa * (dVdseff_dVb + T9 * dVASCBE_dVb) / VASCBE;
  Gds += Gm * dVgsteff_dVd;
  Gmb += Gm * dVgsteff_dVb;
  Gm *= dVgsteff_dVg;
  Gmb *= dVbseff_dVb;
  tmp = p->alpha0 + p->alpha1 * Leff;
  if ((tmp <= 0.0) || (p->beta0 <= 0.0)) {
    Isub = Gbd = Gbb = Gbg = 0.0;
  } else {
    T2 = tmp / Leff;
    if (diffVds > p->beta0 / RC_EXP_THRESHOLD) {
      T0 = -p->beta0 / diffVds;
      T1 = T2 * diffVds * exp(T0);
      T3 = T1 / diffVds * (T0 - 1.0);
      dT1_dVg = T3 * dVdseff_dVg;
      dT1_dVd = T3 * (dVdseff_dVd - 1.0);
      dT1_dVb = T3 * dVdseff_dVb;
    } else {
      T3 = T2 * RC_MIN_EXP;
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
  in->cdrain = Ids;
  in->gds = Gds;
  in->gm = Gm;
  in->gmbs = Gmb;
  in->gbbs = Gbb;
  in->gbgs = Gbg;
  in->gbds = Gbd;
  in->csub = Isub;

  if ((m->xpart < 0) || (!in->ChargeComputationNeeded)) {
    in->qgate = in->qdrn = in->qsrc = in->qbulk = 0.0;
    in->cggb = in->cgsb = in->cgdb = 0.0;
    in->cdgb = in->cdsb = in->cddb = 0.0;
    in->cbgb = in->cbsb = in->cbdb = 0.0;
    in->cqdb = in->cqsb = in->cqgb = in->cqbb = 0.0;
    in->gtau = 0.0;
    goto finished;
  }
  if (m->capMod == 0) {
    b3_charge_cap0(m, p, in, Vbs, Vbseff, Phis, dPhis_dVb, sqrtPhis,
                   dsqrtPhis_dVb, Vgs_eff, dVgs_eff_dVg, Vds, Abulk0, dAbulk0_dVb);
    goto finished;
  }
  /* capMod > 0 preamble (N_DEV_MOSFET_B3.C:6310-6358) */
  if (Vbseff < 0.0) {
    VbseffCV = Vbseff;
    dVbseffCV_dVb = 1.0;
  } else {
    VbseffCV = p->phi - Phis;
    dVbseffCV_dVb = -dPhis_dVb;
  }
  CoxWL = m->cox * p->weffCV * p->leffCV;
  noff = n * p->noff;
  dnoff_dVd = p->noff * dn_dVd;
  dnoff_dVb = p->noff * dn_dVb;
  T0 = Vtm * noff;
  voffcv = p->voffcv;
  VgstNVt = (Vgst - voffcv) / T0;
  if (VgstNVt > RC_EXP_THRESHOLD) {
    Vgsteff = Vgst - voffcv;
    dVgsteff_dVg = dVgs_eff_dVg;
    dVgsteff_dVd = -dVth_dVd;
    dVgsteff_dVb = -dVth_dVb;
  } else if (VgstNVt < -RC_EXP_THRESHOLD) {
    Vgsteff = T0 * log(1.0 + RC_MIN_EXP);
    dVgsteff_dVg = 0.0;
    dVgsteff_dVd = Vgsteff / noff;
    dVgsteff_dVb = dVgsteff_dVd * dnoff_dVb;
    dVgsteff_dVd *= dnoff_dVd;
  } else {
    ExpVgst = exp(VgstNVt);
    Vgsteff = T0 * log(1.0 + ExpVgst);
    dVgsteff_dVg = ExpVgst / (1.0 + ExpVgst);
    dVgsteff_dVd = -dVgsteff_dVg * (dVth_dVd + (Vgst - voffcv) / noff * dnoff_dVd) + Vgsteff / noff * dnoff_dVd;
    dVgsteff_dVb = -dVgsteff_dVg * (dVth_dVb + (Vgst - voffcv) / noff * dnoff_dVb) + Vgsteff / noff * dnoff_dVb;
    dVgsteff_dVg *= dVgs_eff_dVg;
  }
  if (m->capMod == 1 || m->capMod == 2) {
    b3_charge_cap12(m, p, in, (int)m->capMod, VbseffCV, dVbseffCV_dVb,
                    dVbseff_dVb, CoxWL, Vgs_eff, dVgs_eff_dVg, Vds, Abulk0,
                    dAbulk0_dVb, Vgsteff, dVgsteff_dVg, dVgsteff_dVd,
                    dVgsteff_dVb);
    goto finished;
  }
  /* capMod 3 (charge-thickness model), N_DEV_MOSFET_B3.C:6688-6959 */
  V3 = p->vfbzb - Vgs_eff + VbseffCV - RC_DELTA_3;
  if (p->vfbzb <= 0.0) {
    T0 = sqrt(V3 * V3 - 4.0 * RC_DELTA_3 * p->vfbzb);
    T2 = -RC_DELTA_3 / T0;
  } else {
    T0 = sqrt(V3 * V3 + 4.0 * RC_DELTA_3 * p->vfbzb);
    T2 = RC_DELTA_3 / T0;
  }
  T1 = 0.5 * (1.0 + V3 / T0);
  Vfbeff = p->vfbzb - 0.5 * (V3 + T0);
  dVfbeff_dVg = T1 * dVgs_eff_dVg;
  dVfbeff_dVb = -T1 * dVbseffCV_dVb;
  Cox = m->cox;
  Tox = 1.0e8 * m->tox;
  T0 = (Vgs_eff - VbseffCV - p->vfbzb) / Tox;
  dT0_dVg = dVgs_eff_dVg / Tox;
  dT0_dVb = -dVbseffCV_dVb / Tox;
  tmp = T0 * p->acde;
  if ((-RC_EXP_THRESHOLD < tmp) && (tmp < RC_EXP_THRESHOLD)) {
    Tcen = p->ldeb * exp(tmp);
    dTcen_dVg = p->acde * Tcen;
    dTcen_dVb = dTcen_dVg * dT0_dVb;
    dTcen_dVg *= dT0_dVg;
  } else if (tmp <= -RC_EXP_THRESHOLD) {
    Tcen = p->ldeb * RC_MIN_EXP;
    dTcen_dVg = dTcen_dVb = 0.0;
  } else {
    Tcen = p->ldeb * RC_MAX_EXP;
    dTcen_dVg = dTcen_dVb = 0.0;
  }
  LINK = 1.0e-3 * m->tox;
  V3 = p->ldeb - Tcen - LINK;
  V4 = sqrt(V3 * V3 + 4.0 * LINK * p->ldeb);
  Tcen = p->ldeb - 0.5 * (V3 + V4);
  T1 = 0.5 * (1.0 + V3 / V4);
  dTcen_dVg *= T1;
  dTcen_dVb *= T1;
  Ccen = RC_EPSSI / Tcen;
  T2 = Cox / (Cox + Ccen);
  Coxeff = T2 * Ccen;
  T3 = -Ccen / Tcen;
  dCoxeff_dVg = T2 * T2 * T3;
  dCoxeff_dVb = dCoxeff_dVg * dTcen_dVb;
  dCoxeff_dVg *= dTcen_dVg;
  CoxWLcen = CoxWL * Coxeff / Cox;
  Qac0 = CoxWLcen * (Vfbeff - p->vfbzb);
  QovCox = Qac0 / Coxeff;
  dQac0_dVg = CoxWLcen * dVfbeff_dVg + QovCox * dCoxeff_dVg;
  dQac0_dVb = CoxWLcen * dVfbeff_dVb + QovCox * dCoxeff_dVb;
  T0 = 0.5 * p->k1ox;
  T3 = Vgs_eff - Vfbeff - VbseffCV - Vgsteff;
  if (p->k1ox == 0.0) {
    T1 = 0.0;
    T2 = 0.0;
  } else if (T3 < 0.0) {
    T1 = T0 + T3 / p->k1ox;
    T2 = CoxWLcen;
  } else {
    T1 = sqrt(T0 * T0 + T3);
    T2 = CoxWLcen * T0 / T1;
  }
  Qsub0 = CoxWLcen * p->k1ox * (T1 - T0);
  QovCox = Qsub0 / Coxeff;
  dQsub0_dVg = T2 * (dVgs_eff_dVg - dVfbeff_dVg - dVgsteff_dVg) + QovCox * dCoxeff_dVg;
  dQsub0_dVd = -T2 * dVgsteff_dVd;
  dQsub0_dVb = -T2 * (dVfbeff_dVb + dVbseffCV_dVb + dVgsteff_dVb) + QovCox * dCoxeff_dVb;
  if (p->k1ox <= 0.0) {
    Denomi = 0.25 * p->moin * Vtm;
    T0 = 0.5 * p->sqrtPhi;
  } else {
    Denomi = p->moin * Vtm * p->k1ox * p->k1ox;
    T0 = p->k1ox * p->sqrtPhi;
  }
  T1 = 2.0 * T0 + Vgsteff;
  DeltaPhi = Vtm * log(1.0 + T1 * Vgsteff / Denomi);
  dDeltaPhi_dVg = 2.0 * Vtm * (T1 - T0) / (Denomi + T1 * Vgsteff);
  dDeltaPhi_dVd = dDeltaPhi_dVg * dVgsteff_dVd;
  dDeltaPhi_dVb = dDeltaPhi_dVg * dVgsteff_dVb;
  (void)dDeltaPhi_dVd; (void)dDeltaPhi_dVb;
  T3 = 4.0 * (Vth - p->vfbzb - p->phi);
  Tox += Tox;
  if (T3 >= 0.0) {
    T0 = (Vgsteff + T3) / Tox;
    dT0_dVd = (dVgsteff_dVd + 4.0 * dVth_dVd) / Tox;
    dT0_dVb = (dVgsteff_dVb + 4.0 * dVth_dVb) / Tox;
  } else {
    T0 = (Vgsteff + 1.0e-20) / Tox;
    dT0_dVd = dVgsteff_dVd / Tox;
    dT0_dVb = dVgsteff_dVb / Tox;
  }
  tmp = exp(0.7 * log(T0));
  T1 = 1.0 + tmp;
  T2 = 0.7 * tmp / (T0 * Tox);
  Tcen = 1.9e-9 / T1;
  dTcen_dVg = -1.9e-9 * T2 / T1 / T1;
  dTcen_dVd = Tox * dTcen_dVg;
  dTcen_dVb = dTcen_dVd * dT0_dVb;
  dTcen_dVd *= dT0_dVd;
  dTcen_dVg *= dVgsteff_dVg;
  Ccen = RC_EPSSI / Tcen;
  T0 = Cox / (Cox + Ccen);
  Coxeff = T0 * Ccen;
  T1 = -Ccen / Tcen;
  dCoxeff_dVg = T0 * T0 * T1;
  dCoxeff_dVd = dCoxeff_dVg * dTcen_dVd;
  dCoxeff_dVb = dCoxeff_dVg * dTcen_dVb;
  dCoxeff_dVg *= dTcen_dVg;
  CoxWLcen = CoxWL * Coxeff / Cox;
  AbulkCV = Abulk0 * p->abulkCVfactor;
  dAbulkCV_dVb = p->abulkCVfactor * dAbulk0_dVb;
  VdsatCV = (Vgsteff - DeltaPhi) / AbulkCV;
  V4 = VdsatCV - Vds - RC_DELTA_4;
  T0 = sqrt(V4 * V4 + 4.0 * RC_DELTA_4 * VdsatCV);
  VdseffCV = VdsatCV - 0.5 * (V4 + T0);
  T1 = 0.5 * (1.0 + V4 / T0);
  T2 = RC_DELTA_4 / T0;
  T3 = (1.0 - T1 - T2) / AbulkCV;
  T4 = T3 * (1.0 - dDeltaPhi_dVg);
  dVdseffCV_dVg = T4;
  dVdseffCV_dVd = T1;
  dVdseffCV_dVb = -T3 * VdsatCV * dAbulkCV_dVb;
  if (Vds == 0.0) {
    VdseffCV = 0.0;
    dVdseffCV_dVg = 0.0;
    dVdseffCV_dVb = 0.0;
  }
  T0 = AbulkCV * VdseffCV;
  T1 = Vgsteff - DeltaPhi;
  T2 = 12.0 * (T1 - 0.5 * T0 + 1.0e-20);
  T3 = T0 / T2;
  T4 = 1.0 - 12.0 * T3 * T3;
  T5 = AbulkCV * (6.0 * T0 * (4.0 * T1 - T0) / (T2 * T2) - 0.5);
  T6 = T5 * VdseffCV / AbulkCV;
  in->qgate = qinoi = CoxWLcen * (T1 - T0 * (0.5 - T3));
  QovCox = in->qgate / Coxeff;
  Cgg1 = CoxWLcen * (T4 * (1.0 - dDeltaPhi_dVg) + T5 * dVdseffCV_dVg);
  Cgd1 = CoxWLcen * T5 * dVdseffCV_dVd + Cgg1 * dVgsteff_dVd + QovCox * dCoxeff_dVd;
  Cgb1 = CoxWLcen * (T5 * dVdseffCV_dVb + T6 * dAbulkCV_dVb) + Cgg1 * dVgsteff_dVb + QovCox * dCoxeff_dVb;
  Cgg1 = Cgg1 * dVgsteff_dVg + QovCox * dCoxeff_dVg;
  T7 = 1.0 - AbulkCV;
  T8 = T2 * T2;
  T9 = 12.0 * T7 * T0 * T0 / (T8 * AbulkCV);
  T10 = T9 * (1.0 - dDeltaPhi_dVg);
  T11 = -T7 * T5 / AbulkCV;
  T12 = -(T9 * T1 / AbulkCV + VdseffCV * (0.5 - T0 / T2));
  in->qbulk = CoxWLcen * T7 * (0.5 * VdseffCV - T0 * VdseffCV / T2);
  QovCox = in->qbulk / Coxeff;
  Cbg1 = CoxWLcen * (T10 + T11 * dVdseffCV_dVg);
  Cbd1 = CoxWLcen * T11 * dVdseffCV_dVd + Cbg1 * dVgsteff_dVd + QovCox * dCoxeff_dVd;
  Cbb1 = CoxWLcen * (T11 * dVdseffCV_dVb + T12 * dAbulkCV_dVb) + Cbg1 * dVgsteff_dVb + QovCox * dCoxeff_dVb;
  Cbg1 = Cbg1 * dVgsteff_dVg + QovCox * dCoxeff_dVg;
  if (m->xpart > 0.5) {
    in->qsrc = -CoxWLcen * (T1 / 2.0 + T0 / 4.0 - 0.5 * T0 * T0 / T2);
    QovCox = in->qsrc / Coxeff;
    T2 += T2;
    T3 = T2 * T2;
    T7 = -(0.25 - 12.0 * T0 * (4.0 * T1 - T0) / T3);
    T4 = -(0.5 + 24.0 * T0 * T0 / T3) * (1.0 - dDeltaPhi_dVg);
    T5 = T7 * AbulkCV;
    T6 = T7 * VdseffCV;
    Csg = CoxWLcen * (T4 + T5 * dVdseffCV_dVg);
    Csd = CoxWLcen * T5 * dVdseffCV_dVd + Csg * dVgsteff_dVd + QovCox * dCoxeff_dVd;
    Csb = CoxWLcen * (T5 * dVdseffCV_dVb + T6 * dAbulkCV_dVb) + Csg * dVgsteff_dVb + QovCox * dCoxeff_dVb;
    Csg = Csg * dVgsteff_dVg + QovCox * dCoxeff_dVg;
  } else if (m->xpart < 0.5) {
    T2 = T2 / 12.0;
    T3 = 0.5 * CoxWLcen / (T2 * T2);
    T4 = T1 * (2.0 * T0 * T0 / 3.0 + T1 * (T1 - 4.0 * T0 / 3.0)) - 2.0 * T0 * T0 * T0 / 15.0;
    in->qsrc = -T3 * T4;
    QovCox = in->qsrc / Coxeff;
    T8 = 4.0 / 3.0 * T1 * (T1 - T0) + 0.4 * T0 * T0;
    T5 = -2.0 * in->qsrc / T2 - T3 * (T1 * (3.0 * T1 - 8.0 * T0 / 3.0) + 2.0 * T0 * T0 / 3.0);
    T6 = AbulkCV * (in->qsrc / T2 + T3 * T8);
    T7 = T6 * VdseffCV / AbulkCV;
    Csg = T5 * (1.0 - dDeltaPhi_dVg) + T6 * dVdseffCV_dVg;
    Csd = Csg * dVgsteff_dVd + T6 * dVdseffCV_dVd + QovCox * dCoxeff_dVd;
    Csb = Csg * dVgsteff_dVb + T6 * dVdseffCV_dVb + T7 * dAbulkCV_dVb + QovCox * dCoxeff_dVb;
    Csg = Csg * dVgsteff_dVg + QovCox * dCoxeff_dVg;
  } else {
    in->qsrc = -0.5 * in->qgate;
    Csg = -0.5 * Cgg1;
    Csd = -0.5 * Cgd1;
    Csb = -0.5 * Cgb1;
  }
  in->qgate += Qac0 + Qsub0 - in->qbulk;
  in->qbulk -= (Qac0 + Qsub0);
  in->qdrn = -(in->qgate + in->qbulk + in->qsrc);
  Cbg = Cbg1 - dQac0_dVg - dQsub0_dVg;
  Cbd = Cbd1 - dQsub0_dVd;
  Cbb = Cbb1 - dQac0_dVb - dQsub0_dVb;
  Cgg = Cgg1 - Cbg;
  Cgd = Cgd1 - Cbd;
  Cgb = Cgb1 - Cbb;
  Cgb *= dVbseff_dVb;
  Cbb *= dVbseff_dVb;
  Csb *= dVbseff_dVb;
  in->cggb = Cgg;
  in->cgsb = -(Cgg + Cgd + Cgb);
  in->cgdb = Cgd;
  in->cdgb = -(Cgg + Cbg + Csg);
  in->cdsb = (Cgg + Cgd + Cgb + Cbg + Cbd + Cbb + Csg + Csd + Csb);
  in->cddb = -(Cgd + Cbd + Csd);
  in->cbgb = Cbg;
  in->cbsb = -(Cbg + Cbd + Cbb);
  in->cbdb = Cbd;
  in->qinv = -qinoi;

finished:
  in->cd = in->cdrain;
  if (in->ChargeComputationNeeded) {
    double ag0 = 0.0;
    czbd = in->unitAreaJctCapTemp * in->drainArea;
    czbs = in->unitAreaJctCapTemp * in->sourceArea;
    if (in->drainPerimeter < p->weff) {
      czbdswg = in->unitLengthGateSidewallJctCapTemp * in->drainPerimeter;
      czbdsw = 0.0;
    } else {
      czbdsw = in->unitLengthSidewallJctCapTemp * (in->drainPerimeter - p->weff);
      czbdswg = in->unitLengthGateSidewallJctCapTemp * p->weff;
    }
    if (in->sourcePerimeter < p->weff) {
      czbssw = 0.0;
      czbsswg = in->unitLengthGateSidewallJctCapTemp * in->sourcePerimeter;
    } else {
      czbssw = in->unitLengthSidewallJctCapTemp * (in->sourcePerimeter - p->weff);
      czbsswg = in->unitLengthGateSidewallJctCapTemp * p->weff;
    }
    b3_junction(in, in->vbs, czbs, czbssw, czbsswg, &in->qbs, &in->capbs);
    b3_junction(in, in->vbd, czbd, czbdsw, czbdswg, &in->qbd, &in->capbd);

    /* overlap capacitances (N_DEV_MOSFET_B3.C:7164-7238) */
    if (m->capMod == 0 || m->capMod == 1) {
      b3_overlap_cap01(m, p, in->vgd, p->cgdo, p->cgdl, &in->cgdo, &in->qgdo);
      b3_overlap_cap01(m, p, in->vgs, p->cgso, p->cgsl, &in->cgso, &in->qgso);
      goto overlaps_done;
    }
    T0 = in->vgd + RC_DELTA_1;
    T1 = sqrt(T0 * T0 + 4.0 * RC_DELTA_1);
    T2 = 0.5 * (T0 - T1);
    T3 = p->weffCV * p->cgdl;
    T4 = sqrt(1.0 - 4.0 * T2 / p->ckappa);
    in->cgdo = p->cgdo + T3 - T3 * (1.0 - 1.0 / T4) * (0.5 - 0.5 * T0 / T1);
    in->qgdo = (p->cgdo + T3) * in->vgd - T3 * (T2 + 0.5 * p->ckappa * (T4 - 1.0));
    T0 = in->vgs + RC_DELTA_1;
    T1 = sqrt(T0 * T0 + 4.0 * RC_DELTA_1);
    T2 = 0.5 * (T0 - T1);
    T3 = p->weffCV * p->cgsl;
    T4 = sqrt(1.0 - 4.0 * T2 / p->ckappa);
    in->cgso = p->cgso + T3 - T3 * (1.0 - 1.0 / T4) * (0.5 - 0.5 * T0 / T1);
    in->qgso = (p->cgso + T3) * in->vgs - T3 * (T2 + 0.5 * p->ckappa * (T4 - 1.0));
  overlaps_done:

    /* setupCapacitors_oldDAE, charge part (N_DEV_MOSFET_B3.C:7632-7800) */
    (void)ag0;
    if (in->mode > 0) {
      in->qgd = in->qgdo;
      in->qgs = in->qgso;
      in->qgb = p->cgbo * in->vgb;
      in->qgate += in->qgd + in->qgs + in->qgb;
      in->qbulk -= in->qgb;
      in->qdrn -= in->qgd;
      in->qsrc = -(in->qgate + in->qbulk + in->qdrn);
      in->sxpart = 0.6;
      in->dxpart = 0.4;
    } else {
      in->qgd = in->qgdo;
      in->qgs = in->qgso;
      in->qgb = p->cgbo * in->vgb;
      in->qgate += in->qgd + in->qgs + in->qgb;
      in->qbulk -= in->qgb;
      in->qsrc = in->qdrn - in->qgs;
      in->qdrn = -(in->qgate + in->qbulk + in->qsrc);
      in->sxpart = 0.4;
      in->dxpart = 0.6;
    }
    in->ggtg = in->ggtd = in->ggtb = in->ggts = 0.0;
    in->ddxpart_dVd = in->ddxpart_dVg = in->ddxpart_dVb = in->ddxpart_dVs = 0.0;
    in->dsxpart_dVd = in->dsxpart_dVg = in->dsxpart_dVb = in->dsxpart_dVs = 0.0;

    /* setupCapacitors_newDAE (N_DEV_MOSFET_B3.C:7558-7630) */
    if (in->mode > 0) {
      in->CAPcggb = (in->cggb + in->cgdo + in->cgso + p->cgbo);
      in->CAPcgdb = (in->cgdb - in->cgdo);
      in->CAPcgsb = (in->cgsb - in->cgso);
      in->CAPcdgb = (in->cdgb - in->cgdo);
      in->CAPcddb = (in->cddb + in->capbd + in->cgdo);
      in->CAPcdsb = in->cdsb;
      in->CAPcsgb = -(in->cggb + in->cbgb + in->cdgb + in->cgso);
      in->CAPcsdb = -(in->cgdb + in->cbdb + in->cddb);
      in->CAPcssb = (in->capbs + in->cgso - (in->cgsb + in->cbsb + in->cdsb));
      in->CAPcbgb = (in->cbgb - p->cgbo);
      in->CAPcbdb = (in->cbdb - in->capbd);
      in->CAPcbsb = (in->cbsb - in->capbs);
    } else {
      in->CAPcggb = (in->cggb + in->cgdo + in->cgso + p->cgbo);
      in->CAPcgdb = (in->cgsb - in->cgdo);
      in->CAPcgsb = (in->cgdb - in->cgso);
      in->CAPcdgb = -(in->cggb + in->cbgb + in->cdgb + in->cgdo);
      in->CAPcddb = (in->capbd + in->cgdo - (in->cgsb + in->cbsb + in->cdsb));
      in->CAPcdsb = -(in->cgdb + in->cbdb + in->cddb);
      in->CAPcsgb = (in->cdgb - in->cgso);
      in->CAPcsdb = in->cdsb;
      in->CAPcssb = (in->cddb + in->capbs + in->cgso);
      in->CAPcbgb = (in->cbgb - p->cgbo);
      in->CAPcbdb = (in->cbsb - in->capbd);
      in->CAPcbsb = (in->cbdb - in->capbs);
    }
    in->cqdef = 0.0;
    in->qg = in->qgate;
    in->qd = in->qdrn - in->qbd;
    in->qb = in->qbulk + in->qbd + in->qbs;
  }
  in->Idrain = in->drainConductance * in->Vddp;
  in->Isource = in->sourceConductance * in->Vssp;
  in->T1global = T1;
  return 0;
}

/* Master::updateState, N_DEV_MOSFET_B3.C:9222-9296 */
int ref_b3_update_state(ref_b3_master *M, double *nextSol, double *currSta,
                        double *currSto, double *nextSta, double *nextSto) {
  int i;
  M->nextSol = nextSol; M->currSta = currSta; M->currSto = currSto;
  M->nextSta = nextSta; M->nextSto = nextSto;
  for (i = 0; i < M->n; ++i) {
    ref_b3_inst *mi = &M->inst[i];
    if (b3_uiv(M, mi)) return -1;
    nextSto[mi->li_store_vbs] = mi->vbs;
    nextSto[mi->li_store_vgs] = mi->vgs;
    nextSto[mi->li_store_vds] = mi->vds;
    nextSto[mi->li_store_vbd] = mi->vbd;
    nextSto[mi->li_store_von] = mi->von;
    nextSta[mi->li_state_qb] = mi->qb;
    nextSta[mi->li_state_qg] = mi->qg;
    nextSta[mi->li_state_qd] = mi->qd;
    nextSta[mi->li_state_qbs] = mi->qbs;
    nextSta[mi->li_state_qbd] = mi->qbd;
    if (!(M->ss.dcopFlag) && M->ss.initTranFlag && M->ss.newtonIter == 0) {
      currSta[mi->li_state_qb] = mi->qb;
      currSta[mi->li_state_qg] = mi->qg;
      currSta[mi->li_state_qd] = mi->qd;
      currSta[mi->li_state_qbs] = mi->qbs;
      currSta[mi->li_state_qbd] = mi->qbd;
    }
  }
  return 0;
}

/* Instance::auxChargeCalculations, N_DEV_MOSFET_B3.C:7465-7555 (nqsMod 0) */
static void b3_aux_charge(ref_b3_inst *in) {
  if (!in->ChargeComputationNeeded) {
    in->sxpart = (1.0 - (in->dxpart = (in->mode > 0) ? 0.4 : 0.6));
    in->ddxpart_dVd = in->ddxpart_dVg = in->ddxpart_dVb = in->ddxpart_dVs = 0.0;
    in->dsxpart_dVd = in->dsxpart_dVg = in->dsxpart_dVb = in->dsxpart_dVs = 0.0;
    in->gtau = 0.0;
  } else {
    double vgb_orig = in->vgs_orig - in->vbs_orig;
    in->Qeqqg_Jdxp = 0.0;
    if (!in->origFlag)
      in->Qeqqg_Jdxp = -in->CAPcggb * (in->vgb - vgb_orig) +
                       in->CAPcgdb * (in->vbd - in->vbd_orig) +
                       in->CAPcgsb * (in->vbs - in->vbs_orig);
    in->Qeqqb_Jdxp = 0.0;
    if (!in->origFlag)
      in->Qeqqb_Jdxp = -in->CAPcbgb * (in->vgb - vgb_orig) +
                       in->CAPcbdb * (in->vbd - in->vbd_orig) +
                       in->CAPcbsb * (in->vbs - in->vbs_orig);
    in->Qeqqd_Jdxp = 0.0;
    if (!in->origFlag)
      in->Qeqqd_Jdxp = -in->CAPcdgb * (in->vgb - vgb_orig) +
                       in->CAPcddb * (in->vbd - in->vbd_orig) +
                       in->CAPcdsb * (in->vbs - in->vbs_orig);
  }
}

/* Master::loadDAEVectors, N_DEV_MOSFET_B3.C:9306-9643 (no IC=, no lead
   currents). */
void ref_b3_load_vectors(ref_b3_master *M, double *fVec, double *qVec,
                         double *dFdxdVp, double *dQdxdVp) {
  int i;
  for (i = 0; i < M->n; ++i) {
    ref_b3_inst *mi = &M->inst[i];
    int dtype = mi->model->dtype;
    double Qeqqg, Qeqqb, Qeqqd, np = mi->numberParallel;
    mi->cdreq_Jdxp = 0.0;
    mi->ceqbd_Jdxp = 0.0;
    mi->ceqbs_Jdxp = 0.0;
    if (mi->mode >= 0) {
      mi->Gm = mi->gm;
      mi->Gmbs = mi->gmbs;
      mi->FwdSum = mi->Gm + mi->Gmbs;
      mi->RevSum = 0.0;
      mi->cdreq = dtype * (mi->cd);
      mi->ceqbd = -dtype * (mi->csub);
      mi->ceqbs = 0.0;
      mi->gbbdp = -mi->gbds;
      mi->gbbsp = (mi->gbds + mi->gbgs + mi->gbbs);
      mi->gbdpg = mi->gbgs;
      mi->gbdpdp = mi->gbds;
      mi->gbdpb = mi->gbbs;
      mi->gbdpsp = -(mi->gbdpg + mi->gbdpdp + mi->gbdpb);
      mi->gbspg = 0.0;
      mi->gbspdp = 0.0;
      mi->gbspb = 0.0;
      mi->gbspsp = 0.0;
    } else {
      mi->Gm = -mi->gm;
      mi->Gmbs = -mi->gmbs;
      mi->FwdSum = 0.0;
      mi->RevSum = -(mi->Gm + mi->Gmbs);
      mi->cdreq = -dtype * (mi->cd);
      mi->ceqbs = -dtype * (mi->csub);
      mi->ceqbd = 0.0;
      mi->gbbsp = -mi->gbds;
      mi->gbbdp = (mi->gbds + mi->gbgs + mi->gbbs);
      mi->gbdpg = 0.0;
      mi->gbdpsp = 0.0;
      mi->gbdpb = 0.0;
      mi->gbdpdp = 0.0;
      mi->gbspg = mi->gbgs;
      mi->gbspsp = mi->gbds;
      mi->gbspb = mi->gbbs;
      mi->gbspdp = -(mi->gbspg + mi->gbspsp + mi->gbspb);
    }
    if (dtype > 0) {
      mi->ceqbs += (mi->cbs);
      mi->ceqbd += (mi->cbd);
    } else {
      mi->ceqbs -= (mi->cbs);
      mi->ceqbd -= (mi->cbd);
    }
    if (mi->drainConductance != 0.0) fVec[mi->li_Drain] += mi->Idrain * np;
    if (mi->sourceConductance != 0.0) fVec[mi->li_Source] += mi->Isource * np;
    fVec[mi->li_Bulk] += (mi->ceqbs + mi->ceqbd) * np;
    fVec[mi->li_DrainPrime] += (-(mi->ceqbd - mi->cdreq) - mi->Idrain) * np;
    fVec[mi->li_SourcePrime] += (-(mi->cdreq + mi->ceqbs) - mi->Isource) * np;
    if (M->opt.voltageLimiterFlag && !mi->origFlag) {
      double tmp;
      if (mi->mode >= 0) {
        tmp = dtype * (-mi->gds * (mi->vds - mi->vds_orig) -
                       mi->Gm * (mi->vgs - mi->vgs_orig) -
                       mi->Gmbs * (mi->vbs - mi->vbs_orig));
        mi->cdreq_Jdxp += tmp;
        mi->cdreq += tmp;
        tmp = -dtype * (-mi->gbds * (mi->vds - mi->vds_orig) -
                        mi->gbgs * (mi->vgs - mi->vgs_orig) -
                        mi->gbbs * (mi->vbs - mi->vbs_orig));
        mi->ceqbd_Jdxp += tmp;
        mi->ceqbd += tmp;
      } else {
        tmp = -dtype * (mi->gds * (mi->vds - mi->vds_orig) +
                        mi->Gm * (mi->vgd - mi->vgd_orig) +
                        mi->Gmbs * (mi->vbd - mi->vbd_orig));
        mi->cdreq_Jdxp += tmp;
        mi->cdreq += tmp;
        tmp = -dtype * (mi->gbds * (mi->vds - mi->vds_orig) -
                        mi->gbgs * (mi->vgd - mi->vgd_orig) -
                        mi->gbbs * (mi->vbd - mi->vbd_orig));
        mi->ceqbd_Jdxp += tmp;
        mi->ceqbd += tmp;
      }
      if (dtype > 0) {
        mi->ceqbs_Jdxp += (-mi->gbs * (mi->vbs - mi->vbs_orig));
        mi->ceqbs += (-mi->gbs * (mi->vbs - mi->vbs_orig));
        mi->ceqbd_Jdxp += (-mi->gbd * (mi->vbd - mi->vbd_orig));
        mi->ceqbd += (-mi->gbd * (mi->vbd - mi->vbd_orig));
      } else {
        mi->ceqbs_Jdxp -= (-mi->gbs * (mi->vbs - mi->vbs_orig));
        mi->ceqbs -= (-mi->gbs * (mi->vbs - mi->vbs_orig));
        mi->ceqbd_Jdxp -= (-mi->gbd * (mi->vbd - mi->vbd_orig));
        mi->ceqbd -= (-mi->gbd * (mi->vbd - mi->vbd_orig));
      }
      dFdxdVp[mi->li_Bulk] += -(mi->ceqbs_Jdxp + mi->ceqbd_Jdxp) * np;
      dFdxdVp[mi->li_DrainPrime] += (mi->ceqbd_Jdxp - mi->cdreq_Jdxp) * np;
      dFdxdVp[mi->li_SourcePrime] += (mi->cdreq_Jdxp + mi->ceqbs_Jdxp) * np;
    }
    b3_aux_charge(mi);
    if (dtype > 0) {
      Qeqqg = mi->qg;
      Qeqqb = mi->qb;
      Qeqqd = mi->qd;
    } else {
      Qeqqg = -mi->qg;
      Qeqqb = -mi->qb;
      Qeqqd = -mi->qd;
    }
    qVec[mi->li_Gate] += Qeqqg * np;
    qVec[mi->li_Bulk] += (Qeqqb) * np;
    qVec[mi->li_DrainPrime] += (-(-Qeqqd)) * np;
    qVec[mi->li_SourcePrime] += (-(+Qeqqg + Qeqqb + Qeqqd)) * np;
    if (M->opt.voltageLimiterFlag) {
      if (dtype < 0) {
        mi->Qeqqg_Jdxp = -mi->Qeqqg_Jdxp;
        mi->Qeqqb_Jdxp = -mi->Qeqqb_Jdxp;
        mi->Qeqqd_Jdxp = -mi->Qeqqd_Jdxp;
      }
      if (!mi->origFlag) {
        dQdxdVp[mi->li_Gate] += -mi->Qeqqg_Jdxp * np;
        dQdxdVp[mi->li_Bulk] += -(+mi->Qeqqb_Jdxp) * np;
        dQdxdVp[mi->li_DrainPrime] += (-mi->Qeqqd_Jdxp) * np;
        dQdxdVp[mi->li_SourcePrime] += (+mi->Qeqqg_Jdxp + mi->Qeqqb_Jdxp + mi->Qeqqd_Jdxp) * np;
      }
    }
  }
}

/* Master::loadDAEMatrices (pointer form), N_DEV_MOSFET_B3.C:9654-9928 */
void ref_b3_load_matrices(ref_b3_master *M) {
  int i;
  for (i = 0; i < M->n; ++i) {
    ref_b3_inst *mi = &M->inst[i];
    double np = mi->numberParallel, gd = mi->drainConductance,
           gs = mi->sourceConductance;
    *mi->fp[O_D_D] += gd * np;
    *mi->fp[O_D_DP] -= gd * np;
    *mi->fp[O_S_S] += gs * np;
    *mi->fp[O_S_SP] -= gs * np;
    *mi->fp[O_B_G] += (-mi->gbgs) * np;
    *mi->fp[O_B_B] += (mi->gbd + mi->gbs - mi->gbbs) * np;
    *mi->fp[O_B_DP] += (-mi->gbd + mi->gbbdp) * np;
    *mi->fp[O_B_SP] += (-mi->gbs + mi->gbbsp) * np;
    *mi->fp[O_DP_D] -= gd * np;
    *mi->fp[O_DP_B] -= (mi->gbd - mi->Gmbs - mi->dxpart * mi->ggtb -
                        mi->T1global * mi->ddxpart_dVb - mi->gbdpb) * np;
    *mi->fp[O_DP_G] += (mi->Gm + mi->dxpart * mi->ggtg +
                        mi->T1global * mi->ddxpart_dVg + mi->gbdpg) * np;
    *mi->fp[O_DP_DP] += (gd + mi->gds + mi->gbd + mi->RevSum + mi->dxpart * mi->ggtd +
                         mi->T1global * mi->ddxpart_dVd + mi->gbdpdp) * np;
    *mi->fp[O_DP_SP] -= (mi->gds + mi->FwdSum - mi->dxpart * mi->ggts -
                         mi->T1global * mi->ddxpart_dVs - mi->gbdpsp) * np;
    *mi->fp[O_SP_G] += (-mi->Gm + mi->sxpart * mi->ggtg +
                        mi->T1global * mi->dsxpart_dVg + mi->gbspg) * np;
    *mi->fp[O_SP_B] -= (mi->gbs + mi->Gmbs - mi->sxpart * mi->ggtb -
                        mi->T1global * mi->dsxpart_dVb - mi->gbspb) * np;
    *mi->fp[O_SP_S] -= gs * np;
    *mi->fp[O_SP_DP] -= (mi->gds + mi->RevSum - mi->sxpart * mi->ggtd -
                         mi->T1global * mi->dsxpart_dVd - mi->gbspdp) * np;
    *mi->fp[O_SP_SP] += (gs + mi->gds + mi->gbs + mi->FwdSum + mi->sxpart * mi->ggts +
                         mi->T1global * mi->dsxpart_dVs + mi->gbspsp) * np;
    if (!(!(M->ss.dcopFlag) && M->ss.initTranFlag && M->ss.newtonIter == 0)) {
      *mi->qp[O_G_G] += (mi->CAPcggb) * np;
      *mi->qp[O_G_B] -= (mi->CAPcggb + mi->CAPcgdb + mi->CAPcgsb) * np;
      *mi->qp[O_G_DP] += (mi->CAPcgdb) * np;
      *mi->qp[O_G_SP] += (mi->CAPcgsb) * np;
      *mi->qp[O_B_G] += (mi->CAPcbgb) * np;
      *mi->qp[O_B_B] += (-mi->CAPcbgb - mi->CAPcbdb - mi->CAPcbsb) * np;
      *mi->qp[O_B_DP] += (mi->CAPcbdb) * np;
      *mi->qp[O_B_SP] += (mi->CAPcbsb) * np;
      *mi->qp[O_DP_B] -= (+mi->CAPcdgb + mi->CAPcddb + mi->CAPcdsb) * np;
      *mi->qp[O_DP_G] += (mi->CAPcdgb) * np;
      *mi->qp[O_DP_DP] += (+mi->CAPcddb) * np;
      *mi->qp[O_DP_SP] -= (-mi->CAPcdsb) * np;
      *mi->qp[O_SP_G] += (mi->CAPcsgb) * np;
      *mi->qp[O_SP_B] -= (+mi->CAPcsgb + mi->CAPcsdb + mi->CAPcssb) * np;
      *mi->qp[O_SP_DP] -= (-mi->CAPcsdb) * np;
      *mi->qp[O_SP_SP] += (+mi->CAPcssb) * np;
    }
  }
}

void ref_b3_flags(const ref_b3_master *M, int *limited, int *orig) {
  int i;
  for (i = 0; i < M->n; ++i) {
    limited[i] = M->inst[i].limitedFlag;
    orig[i] = M->inst[i].origFlag;
  }
}

/* Size-dependent parameters of instance i, for checking other setups:
   out[0..]=leff, weff, leffCV, weffCV, vth0, vfb, vfbzb, k1, k2, k1ox,
   phi, u0temp, vsattemp, rds0, cgso, cgdo, cgbo, abulkCVfactor, litl,
   thetaRout, theta0vb0, cdep0, ldeb, acde, Xdep0, vbsc, vbi, and the
   temperature-adjusted mobility coefficients ua, ub, uc (27..29). */
void ref_b3_sdp_probe(const ref_b3_master *M, int i, double *out) {
  const ref_b3_sdp *p = M->inst[i].paramPtr;
  double v[30] = {p->leff, p->weff, p->leffCV, p->weffCV, p->vth0, p->vfb,
                  p->vfbzb, p->k1, p->k2, p->k1ox, p->phi, p->u0temp,
                  p->vsattemp, p->rds0, p->cgso, p->cgdo, p->cgbo,
                  p->abulkCVfactor, p->litl, p->thetaRout, p->theta0vb0,
                  p->cdep0, p->ldeb, p->acde, p->Xdep0, p->vbsc, p->vbi,
                  p->ua, p->ub, p->uc};
  memcpy(out, v, sizeof(v));
}
