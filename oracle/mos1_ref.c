/*
 * ORACLE — test infrastructure only.  Nothing in the product links this.
 *
 * CPU restatement of Xyce's MOSFET level-1 device load
 * (src/DeviceModelPKG/OpenModels/src/N_DEV_MOSFET1.C), kept in the
 * reference's own shape: array-of-structs instances, a Master that loops
 * over them in insertion order, and stamping by read-modify-write through
 * cached pointers into the CSR value arrays.  newMeyerFlag=false (the
 * default, N_DEV_DeviceOptions.C:147) is the only formulation restated;
 * homotopy (artParameterFlag) and lead currents are not.
 *
 * Every function names the reference lines it follows.  The order of
 * floating-point operations is the reference's, so differences against it
 * come from libm ulps only.
 */
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include "ref_common.h"

/* ---- Model (N_DEV_MOSFET1.h:704-772, defaults N_DEV_MOSFET1.C:117-263) ---- */
typedef struct ref_mos1_model {
  int dtype;
  double model_l, model_w, tnom, latDiff, jctSatCurDensity, jctSatCur;
  double drainResistance, sourceResistance, sheetResistance, transconductance;
  double gateSourceOverlapCapFactor, gateDrainOverlapCapFactor,
      gateBulkOverlapCapFactor, oxideCapFactor, vt0, capBD, capBS,
      bulkCapFactor, sideWallCapFactor, bulkJctPotential,
      bulkJctBotGradingCoeff, bulkJctSideGradingCoeff, fwdCapDepCoeff, phi,
      gamma, lambda, substrateDoping, surfaceStateDensity, oxideThickness,
      surfaceMobility, surfaceMobility0;
  int gateType;
  int capBDGiven, capBSGiven, bulkCapFactorGiven, sideWallCapFactorGiven;
  /* given flags used by processParams */
  int g_L, g_W, g_TNOM, g_TOX, g_KP, g_UO, g_U0, g_NSUB, g_PHI, g_TPG,
      g_GAMMA, g_VTO, g_NSS, g_RSH;
  double fact1, vtnom, egfet1, pbfact1;
} ref_mos1_model;

/* ---- Instance (N_DEV_MOSFET1.h:83-654) ---- */
typedef struct ref_mos1_inst {
  ref_mos1_model *model;
  int OFF, IC_GIVEN;
  double l, w, drainArea, sourceArea, drainSquares, sourceSquares,
      drainPerimeter, sourcePerimeter, sourceConductance, drainConductance,
      temp, numberParallel;
  double tTransconductance, tSurfMob, tPhi, tVto, tSatCur, tSatCurDens, tCbd,
      tCbs, tCj, tCjsw, tBulkPot, tDepCap, tVbi;
  double icVBS, icVDS, icVGS, von, vdsat, sourceVcrit, drainVcrit, cd, cbs,
      cbd, gmbs, gm, gds, gbd, gbs, capbd, capbs, Cbd, Cbdsw, Cbs, Cbssw, f2d,
      f3d, f4d, f2s, f3s, f4s;
  int mode;
  int limitedFlag, origFlag;
  double EffectiveLength, DrainSatCur, SourceSatCur, GateSourceOverlapCap,
      GateDrainOverlapCap, GateBulkOverlapCap, OxideCap;
  double Vd, Vs, Vg, Vb, Vdp, Vsp, Vddp, Vssp, Vbsp, Vbdp, Vgsp, Vgdp, Vgb,
      Vdpsp, vgd, vt;
  double Capgs, Capgd, Capgb, Isource, Idrain, cdrain, Gm, Gmbs, revsum,
      nrmsum, cdreq;
  int li_Drain, li_DrainPrime, li_Source, li_SourcePrime, li_Gate, li_Bulk;
  /* offsets, in the order registerJacLIDs assigns them
     (N_DEV_MOSFET1.C:1650-1673) */
  int off[22];
  double *fp[22], *qp[22];
  double vbd, vbs, vgs, vds;
  double vgs_orig, vds_orig, vbs_orig, vbd_orig, vgd_orig;
  double vgs_old, vds_old, vbs_old, vbd_old, vgd_old;
  double capgs, capgd, capgb, qgs, qgd, qgb, qbd, qbs;
  int li_store_vbd, li_store_vbs, li_store_vgs, li_store_vds, li_store_von;
  int li_state_capgs, li_state_capgd, li_state_capgb, li_state_qgs,
      li_state_qgd, li_state_qgb, li_state_qbd, li_state_qbs;
} ref_mos1_inst;

enum {
  O_D_D, O_D_DP, O_G_G, O_G_B, O_G_DP, O_G_SP, O_S_S, O_S_SP, O_B_G, O_B_B,
  O_B_DP, O_B_SP, O_DP_D, O_DP_G, O_DP_B, O_DP_DP, O_DP_SP, O_SP_G, O_SP_S,
  O_SP_B, O_SP_DP, O_SP_SP
};

typedef struct ref_mos1_master {
  int n, cap;
  ref_mos1_inst *inst;
  ref_dev_options opt;
  ref_solver_state ss;
  /* ExternData (N_DEV_ExternData.h:67-253): raw vector pointers */
  double *nextSol, *currSta, *currSto, *nextSta, *nextSto;
  double *dFdxdVp, *dQdxdVp;
} ref_mos1_master;

/* Model::processParams, N_DEV_MOSFET1.C:3786-3868 */
static void mos1_model_process(ref_mos1_model *m) {
  double wkfngs, wkfng, fermig, fermis, vfb, kt1, arg1;
  m->fact1 = m->tnom / RC_REFTEMP;
  m->vtnom = m->tnom * RC_KoverQ;
  kt1 = RC_BOLTZ * m->tnom;
  m->egfet1 = 1.16 - (7.02e-4 * m->tnom * m->tnom) / (m->tnom + 1108);
  arg1 = -m->egfet1 / (kt1 + kt1) + 1.1150877 / (RC_BOLTZ * (RC_REFTEMP + RC_REFTEMP));
  m->pbfact1 = -2 * m->vtnom * (1.5 * log(m->fact1) + RC_Q * arg1);
  if (!m->g_TOX || m->oxideThickness == 0) {
    m->oxideCapFactor = 0;
  } else {
    m->oxideCapFactor = 3.9 * 8.854214871e-12 / m->oxideThickness;
    if (!m->g_KP) {
      if (!m->g_UO && !m->g_U0) m->surfaceMobility = 600;
      m->transconductance = m->surfaceMobility * m->oxideCapFactor * 1e-4;
    }
    if (m->g_NSUB) {
      if (m->substrateDoping * 1e6 > 1.45e16) {
        if (!m->g_PHI) {
          m->phi = 2 * m->vtnom * log(m->substrateDoping * 1e6 / 1.45e16);
          m->phi = rc_max(0.1, m->phi);
        }
        fermis = m->dtype * .5 * m->phi;
        wkfng = 3.2;
        if (!m->g_TPG) m->gateType = 1;
        if (m->gateType != 0) {
          fermig = m->dtype * m->gateType * .5 * m->egfet1;
          wkfng = 3.25 + .5 * m->egfet1 - fermig;
        }
        wkfngs = wkfng - (3.25 + .5 * m->egfet1 + fermis);
        if (!m->g_GAMMA)
          m->gamma = sqrt(2 * 11.70 * 8.854214871e-12 * RC_Q *
                          m->substrateDoping * 1e6) / m->oxideCapFactor;
        if (!m->g_VTO) {
          if (!m->g_NSS) m->surfaceStateDensity = 0;
          vfb = wkfngs - m->surfaceStateDensity * 1e4 * RC_Q / m->oxideCapFactor;
          m->vt0 = vfb + m->dtype * (m->gamma * sqrt(m->phi) + m->phi);
        }
      } else {
        m->substrateDoping = 0;
      }
    }
  }
}

/* Model::Model + parameter defaults, N_DEV_MOSFET1.C:117-263, 3889-3980 */
ref_mos1_model *ref_mos1_model_new(int dtype, int n, const char **keys,
                                   const double *vals) {
  ref_mos1_model *m = (ref_mos1_model *)calloc(1, sizeof(*m));
  int i;
  m->dtype = dtype;
  m->model_l = 1e-4; m->model_w = 1e-4; m->vt0 = 0.0; m->transconductance = 2e-5;
  m->gamma = 0.0; m->phi = 0.6; m->lambda = 0.0; m->drainResistance = 0.0;
  m->sourceResistance = 0.0; m->capBD = 0.0; m->capBS = 0.0; m->jctSatCur = 1e-14;
  m->bulkJctPotential = 0.8; m->gateSourceOverlapCapFactor = 0.0;
  m->gateDrainOverlapCapFactor = 0.0; m->gateBulkOverlapCapFactor = 0.0;
  m->sheetResistance = 0.0; m->bulkCapFactor = 0.0; m->bulkJctBotGradingCoeff = 0.5;
  m->sideWallCapFactor = 0.0; m->bulkJctSideGradingCoeff = 0.5;
  m->jctSatCurDensity = 0.0; m->oxideThickness = 1e-7; m->latDiff = 0.0;
  m->surfaceMobility = 600.0; m->surfaceMobility0 = 600.0; m->fwdCapDepCoeff = 0.5;
  m->substrateDoping = 0.0; m->surfaceStateDensity = 0.0; m->tnom = RC_REFTEMP;
  m->gateType = 0;
  for (i = 0; i < n; ++i) {
    const char *k = keys[i];
    double v = vals[i];
    if (!strcmp(k, "L")) { m->model_l = v; m->g_L = 1; }
    else if (!strcmp(k, "W")) { m->model_w = v; m->g_W = 1; }
    else if (!strcmp(k, "VTO")) { m->vt0 = v; m->g_VTO = 1; }
    else if (!strcmp(k, "KP")) { m->transconductance = v; m->g_KP = 1; }
    else if (!strcmp(k, "GAMMA")) { m->gamma = v; m->g_GAMMA = 1; }
    else if (!strcmp(k, "PHI")) { m->phi = v; m->g_PHI = 1; }
    else if (!strcmp(k, "LAMBDA")) m->lambda = v;
    else if (!strcmp(k, "RD")) m->drainResistance = v;
    else if (!strcmp(k, "RS")) m->sourceResistance = v;
    else if (!strcmp(k, "CBD")) { m->capBD = v; m->capBDGiven = 1; }
    else if (!strcmp(k, "CBS")) { m->capBS = v; m->capBSGiven = 1; }
    else if (!strcmp(k, "IS")) m->jctSatCur = v;
    else if (!strcmp(k, "PB")) m->bulkJctPotential = v;
    else if (!strcmp(k, "CGSO")) m->gateSourceOverlapCapFactor = v;
    else if (!strcmp(k, "CGDO")) m->gateDrainOverlapCapFactor = v;
    else if (!strcmp(k, "CGBO")) m->gateBulkOverlapCapFactor = v;
    else if (!strcmp(k, "RSH")) { m->sheetResistance = v; m->g_RSH = 1; }
    else if (!strcmp(k, "CJ")) { m->bulkCapFactor = v; m->bulkCapFactorGiven = 1; }
    else if (!strcmp(k, "MJ")) m->bulkJctBotGradingCoeff = v;
    else if (!strcmp(k, "CJSW")) { m->sideWallCapFactor = v; m->sideWallCapFactorGiven = 1; }
    else if (!strcmp(k, "MJSW")) m->bulkJctSideGradingCoeff = v;
    else if (!strcmp(k, "JS")) m->jctSatCurDensity = v;
    else if (!strcmp(k, "TOX")) { m->oxideThickness = v; m->g_TOX = 1; }
    else if (!strcmp(k, "LD")) m->latDiff = v;
    else if (!strcmp(k, "UO")) { m->surfaceMobility = v; m->g_UO = 1; }
    else if (!strcmp(k, "U0")) { m->surfaceMobility0 = v; m->g_U0 = 1; }
    else if (!strcmp(k, "FC")) m->fwdCapDepCoeff = v;
    else if (!strcmp(k, "NSUB")) { m->substrateDoping = v; m->g_NSUB = 1; }
    else if (!strcmp(k, "NSS")) { m->surfaceStateDensity = v; m->g_NSS = 1; }
    /* TNOM and TEMP are given in Celsius: N_DEV_DeviceEntity.C:863-866 */
    else if (!strcmp(k, "TNOM")) { m->tnom = v + 273.15; m->g_TNOM = 1; }
    else if (!strcmp(k, "TPG")) { m->gateType = (int)v; m->g_TPG = 1; }
    else { free(m); return NULL; }
  }
  /* N_DEV_MOSFET1.C:3957-3966 */
  if (m->capBD != 0) m->capBDGiven = 1;
  if (m->capBS != 0) m->capBSGiven = 1;
  /* N_DEV_MOSFET1.C:3971-3978 */
  if (m->g_U0) m->surfaceMobility = m->surfaceMobility0;
  mos1_model_process(m);
  return m;
}

void ref_mos1_model_free(ref_mos1_model *m) { free(m); }

/* Instance::updateTemperature, N_DEV_MOSFET1.C:3437-3680 */
static void mos1_update_temperature(ref_mos1_inst *in) {
  ref_mos1_model *m = in->model;
  double czbd, czbdsw, czbs, czbssw, arg, sarg, sargsw, ratio, ratio4, fact2,
      kt, egfet, pbfact, capfact, phio, pbo, gmanew, gmaold, tnom;
  tnom = m->tnom;
  ratio = in->temp / tnom;
  in->vt = in->temp * RC_KoverQ;
  ratio = in->temp / tnom;
  fact2 = in->temp / RC_REFTEMP;
  kt = in->temp * RC_BOLTZ;
  egfet = 1.16 - (7.02e-4 * in->temp * in->temp) / (in->temp + 1108);
  arg = -egfet / (kt + kt) + 1.1150877 / (RC_BOLTZ * (RC_REFTEMP + RC_REFTEMP));
  pbfact = -2 * in->vt * (1.5 * log(fact2) + RC_Q * arg);

  ratio4 = ratio * sqrt(ratio);
  in->tTransconductance = m->transconductance / ratio4;
  in->tSurfMob = m->surfaceMobility / ratio4;
  phio = (m->phi - m->pbfact1) / m->fact1;
  in->tPhi = fact2 * phio + pbfact;
  in->tVbi = m->vt0 - m->dtype * (m->gamma * sqrt(m->phi)) +
             .5 * (m->egfet1 - egfet) + m->dtype * .5 * (in->tPhi - m->phi);
  in->tVto = in->tVbi + m->dtype * m->gamma * sqrt(in->tPhi);
  in->tSatCur = m->jctSatCur * exp(-egfet / in->vt + m->egfet1 / m->vtnom);
  in->tSatCurDens = m->jctSatCurDensity * exp(-egfet / in->vt + m->egfet1 / m->vtnom);
  pbo = (m->bulkJctPotential - m->pbfact1) / m->fact1;
  gmaold = (m->bulkJctPotential - pbo) / pbo;
  capfact = 1 / (1 + m->bulkJctBotGradingCoeff * (4e-4 * (m->tnom - RC_REFTEMP) - gmaold));
  in->tCbd = m->capBD * capfact;
  in->tCbs = m->capBS * capfact;
  in->tCj = m->bulkCapFactor * capfact;
  capfact = 1 / (1 + m->bulkJctSideGradingCoeff * (4e-4 * (m->tnom - RC_REFTEMP) - gmaold));
  in->tCjsw = m->sideWallCapFactor * capfact;
  in->tBulkPot = fact2 * pbo + pbfact;
  gmanew = (in->tBulkPot - pbo) / pbo;
  capfact = (1 + m->bulkJctBotGradingCoeff * (4e-4 * (in->temp - RC_REFTEMP) - gmanew));
  in->tCbd *= capfact;
  in->tCbs *= capfact;
  in->tCj *= capfact;
  capfact = (1 + m->bulkJctSideGradingCoeff * (4e-4 * (in->temp - RC_REFTEMP) - gmanew));
  in->tCjsw *= capfact;
  in->tDepCap = m->fwdCapDepCoeff * in->tBulkPot;

  if ((m->jctSatCurDensity == 0) || (in->drainArea == 0) || (in->sourceArea == 0)) {
    in->sourceVcrit = in->drainVcrit = in->vt * log(in->vt / (RC_ROOT2 * m->jctSatCur));
  } else {
    in->drainVcrit = in->vt * log(in->vt / (RC_ROOT2 * m->jctSatCurDensity * in->drainArea));
    in->sourceVcrit = in->vt * log(in->vt / (RC_ROOT2 * m->jctSatCurDensity * in->sourceArea));
  }
  if (m->capBDGiven) czbd = in->tCbd;
  else czbd = m->bulkCapFactorGiven ? in->tCj * in->drainArea : 0;
  czbdsw = m->sideWallCapFactorGiven ? in->tCjsw * in->drainPerimeter : 0;
  arg = 1 - m->fwdCapDepCoeff;
  sarg = exp((-m->bulkJctBotGradingCoeff) * log(arg));
  sargsw = exp((-m->bulkJctSideGradingCoeff) * log(arg));
  in->Cbd = czbd;
  in->Cbdsw = czbdsw;
  in->f2d = czbd * (1 - m->fwdCapDepCoeff * (1 + m->bulkJctBotGradingCoeff)) * sarg / arg +
            czbdsw * (1 - m->fwdCapDepCoeff * (1 + m->bulkJctSideGradingCoeff)) * sargsw / arg;
  in->f3d = czbd * m->bulkJctBotGradingCoeff * sarg / arg / in->tBulkPot +
            czbdsw * m->bulkJctSideGradingCoeff * sargsw / arg / in->tBulkPot;
  in->f4d = czbd * in->tBulkPot * (1 - arg * sarg) / (1 - m->bulkJctBotGradingCoeff) +
            czbdsw * in->tBulkPot * (1 - arg * sargsw) / (1 - m->bulkJctSideGradingCoeff) -
            in->f3d / 2 * (in->tDepCap * in->tDepCap) - in->tDepCap * in->f2d;
  if (m->capBSGiven) czbs = in->tCbs;
  else czbs = m->bulkCapFactorGiven ? in->tCj * in->sourceArea : 0;
  czbssw = m->sideWallCapFactorGiven ? in->tCjsw * in->sourcePerimeter : 0;
  arg = 1 - m->fwdCapDepCoeff;
  sarg = exp((-m->bulkJctBotGradingCoeff) * log(arg));
  sargsw = exp((-m->bulkJctSideGradingCoeff) * log(arg));
  in->Cbs = czbs;
  in->Cbssw = czbssw;
  in->f2s = czbs * (1 - m->fwdCapDepCoeff * (1 + m->bulkJctBotGradingCoeff)) * sarg / arg +
            czbssw * (1 - m->fwdCapDepCoeff * (1 + m->bulkJctSideGradingCoeff)) * sargsw / arg;
  in->f3s = czbs * m->bulkJctBotGradingCoeff * sarg / arg / in->tBulkPot +
            czbssw * m->bulkJctSideGradingCoeff * sargsw / arg / in->tBulkPot;
  in->f4s = czbs * in->tBulkPot * (1 - arg * sarg) / (1 - m->bulkJctBotGradingCoeff) +
            czbssw * in->tBulkPot * (1 - arg * sargsw) / (1 - m->bulkJctSideGradingCoeff) -
            in->f3s / 2 * (in->tDepCap * in->tDepCap) - in->tDepCap * in->f2s;
}

ref_mos1_master *ref_mos1_master_new(void) {
  ref_mos1_master *M = (ref_mos1_master *)calloc(1, sizeof(*M));
  M->opt.gmin = 1.0e-12;
  M->opt.voltageLimiterFlag = 1;
  return M;
}

void ref_mos1_master_free(ref_mos1_master *M) {
  free(M->inst);
  free(M);
}

/*
 * Instance::Instance, N_DEV_MOSFET1.C:313-1260 (old-Meyer stamp branch) and
 * registerLIDs / registerStateLIDs / registerStoreLIDs
 * (N_DEV_MOSFET1.C:1275-1350, 1486-1520, 1547-1566).
 * lids = {d, g, s, b, d', s'}: d'/s' are used only when the instance has a
 * drain / source conductance (internal node), as registerLIDs does.
 * Returns the instance index, or -1 on an unknown parameter.
 */
int ref_mos1_add_instance(ref_mos1_master *M, ref_mos1_model *m, int n,
                          const char **keys, const double *vals,
                          const int *lids, int sta_base, int sto_base,
                          double temp) {
  ref_mos1_inst *in;
  int i, gL = 0, gW = 0, gAD = 0, gAS = 0;
  if (M->n == M->cap) {
    M->cap = M->cap ? 2 * M->cap : 64;
    M->inst = (ref_mos1_inst *)realloc(M->inst, M->cap * sizeof(ref_mos1_inst));
  }
  in = &M->inst[M->n];
  memset(in, 0, sizeof(*in));
  in->model = m;
  in->l = 1.0e-4; in->w = 1.0e-4;               /* defl, defw */
  in->drainArea = 0.0; in->sourceArea = 0.0;    /* defad, defas */
  in->drainSquares = 1.0; in->sourceSquares = 1.0;
  in->numberParallel = 1;
  in->temp = temp;
  in->mode = 1;
  for (i = 0; i < n; ++i) {
    const char *k = keys[i];
    double v = vals[i];
    if (!strcmp(k, "L")) { in->l = v; gL = 1; }
    else if (!strcmp(k, "W")) { in->w = v; gW = 1; }
    else if (!strcmp(k, "AD")) { in->drainArea = v; gAD = 1; }
    else if (!strcmp(k, "AS")) { in->sourceArea = v; gAS = 1; }
    else if (!strcmp(k, "NRD")) in->drainSquares = v;
    else if (!strcmp(k, "NRS")) in->sourceSquares = v;
    else if (!strcmp(k, "PD")) in->drainPerimeter = v;
    else if (!strcmp(k, "PS")) in->sourcePerimeter = v;
    else if (!strcmp(k, "M")) in->numberParallel = v;
    else if (!strcmp(k, "OFF")) in->OFF = (v != 0.0);
    else if (!strcmp(k, "IC1")) { in->icVDS = v; in->IC_GIVEN = 1; }
    else if (!strcmp(k, "IC2")) { in->icVGS = v; in->IC_GIVEN = 1; }
    else if (!strcmp(k, "IC3")) { in->icVBS = v; in->IC_GIVEN = 1; }
    else if (!strcmp(k, "TEMP")) in->temp = v + 273.15;  /* Celsius, as TNOM */
    else return -1;
  }
  if (!gL) in->l = m->model_l;
  if (!gW) in->w = m->model_w;
  (void)gAD; (void)gAS;
  /* N_DEV_MOSFET1.C:1184-1225 */
  if (m->drainResistance != 0) in->drainConductance = 1 / m->drainResistance;
  else if (m->g_RSH) in->drainConductance = m->sheetResistance != 0 ? 1 / (m->sheetResistance * in->drainSquares) : 0;
  else in->drainConductance = 0;
  if (m->sourceResistance != 0) in->sourceConductance = 1 / m->sourceResistance;
  else if (m->g_RSH) in->sourceConductance = m->sheetResistance != 0 ? 1 / (m->sheetResistance * in->sourceSquares) : 0;
  else in->sourceConductance = 0;
  /* N_DEV_MOSFET1.C:1238-1248 */
  in->EffectiveLength = in->l - 2 * m->latDiff;
  in->GateSourceOverlapCap = m->gateSourceOverlapCapFactor * in->w;
  in->GateDrainOverlapCap = m->gateDrainOverlapCapFactor * in->w;
  in->GateBulkOverlapCap = m->gateBulkOverlapCapFactor * in->EffectiveLength;
  in->OxideCap = m->oxideCapFactor * in->EffectiveLength * in->w;
  mos1_update_temperature(in);
  /* registerLIDs */
  in->li_Drain = lids[0]; in->li_Gate = lids[1];
  in->li_Source = lids[2]; in->li_Bulk = lids[3];
  in->li_DrainPrime = in->drainConductance != 0.0 ? lids[4] : in->li_Drain;
  in->li_SourcePrime = in->sourceConductance != 0.0 ? lids[5] : in->li_Source;
  /* registerStateLIDs order */
  in->li_state_qgs = sta_base + 0; in->li_state_qgd = sta_base + 1;
  in->li_state_qgb = sta_base + 2; in->li_state_capgs = sta_base + 3;
  in->li_state_capgd = sta_base + 4; in->li_state_capgb = sta_base + 5;
  in->li_state_qbd = sta_base + 6; in->li_state_qbs = sta_base + 7;
  /* registerStoreLIDs order */
  in->li_store_vbd = sto_base + 0; in->li_store_vbs = sto_base + 1;
  in->li_store_vgs = sto_base + 2; in->li_store_vds = sto_base + 3;
  in->li_store_von = sto_base + 4;
  return M->n++;
}

/* Instance::registerJacLIDs result (N_DEV_MOSFET1.C:1650-1673): 22 offsets. */
void ref_mos1_set_offsets(ref_mos1_master *M, int i, const int *off22) {
  memcpy(M->inst[i].off, off22, sizeof(int) * 22);
}

/* Instance::setupPointers, N_DEV_MOSFET1.C:1710-1853 (dFdx and dQdx share
   the JACOBIAN_OVERLAP_GND graph; row pointer = &vals[rowptr[li]]). */
void ref_mos1_setup_pointers(ref_mos1_master *M, const int *rowptr,
                             double *dFdx, double *dQdx) {
  int i, k;
  for (i = 0; i < M->n; ++i) {
    ref_mos1_inst *in = &M->inst[i];
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

void ref_mos1_set_solver_state(ref_mos1_master *M, const int *flags) {
  M->ss.dcopFlag = flags[0];
  M->ss.initJctFlag = flags[1];
  M->ss.initFixFlag = flags[2];
  M->ss.initTranFlag = flags[3];
  M->ss.newtonIter = flags[4];
  M->ss.tranopFlag = flags[5];
  M->ss.acopFlag = flags[6];
  M->ss.transientFlag = flags[7];
}

void ref_mos1_set_options(ref_mos1_master *M, double gmin, int limiter) {
  M->opt.gmin = gmin;
  M->opt.voltageLimiterFlag = limiter;
}

/* Instance::updateIntermediateVars, N_DEV_MOSFET1.C:2618-3427
   (artParameterFlag=false, newMeyerFlag=false, inputOPFlag=false). */
static void mos1_uiv(ref_mos1_master *M, ref_mos1_inst *in) {
  ref_mos1_model *m = in->model;
  const ref_solver_state *ss = &M->ss;
  double Von, Vdsat, Beta, evbs, evbd, sarg, sargsw, arg;
  int Check = 1;
  double capgs_old, capgd_old, capgb_old;
  double vgs_save, vgd_save, vds_save;
  double *sol = M->nextSol;

  if ((in->tSatCurDens == 0) || (in->drainArea == 0) || (in->sourceArea == 0)) {
    in->DrainSatCur = in->tSatCur;
    in->SourceSatCur = in->tSatCur;
  } else {
    in->DrainSatCur = in->tSatCurDens * in->drainArea;
    in->SourceSatCur = in->tSatCurDens * in->sourceArea;
  }
  Beta = in->tTransconductance * in->w / in->EffectiveLength;

  in->Vd = sol[in->li_Drain];
  in->Vg = sol[in->li_Gate];
  in->Vs = sol[in->li_Source];
  in->Vb = sol[in->li_Bulk];
  in->Vsp = sol[in->li_SourcePrime];
  in->Vdp = sol[in->li_DrainPrime];

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
    if (in->IC_GIVEN) {
      in->vds = m->dtype * in->icVDS;
      in->vgs = m->dtype * in->icVGS;
      in->vbs = m->dtype * in->icVBS;
      in->vbd = in->vbs - in->vds;
      in->vgd = in->vgs - in->vds;
      in->origFlag = 0;
    } else {
      in->vbs = -1;
      in->vgs = m->dtype * in->tVto;
      in->vds = 0;
      in->vbd = in->vbs - in->vds;
      in->vgd = in->vgs - in->vds;
    }
  } else if ((ss->initFixFlag || ss->initJctFlag) && in->OFF) {
    in->vbs = in->vgs = in->vds = 0;
    in->vbd = in->vgd = 0;
  }

  if (ss->newtonIter == 0) {
    if (!(ss->dcopFlag) || (ss->locaEnabledFlag && ss->dcopFlag)) {
      in->vbs_old = M->currSto[in->li_store_vbs];
      in->vbd_old = M->currSto[in->li_store_vbd];
      in->vgs_old = M->currSto[in->li_store_vgs];
      in->vds_old = M->currSto[in->li_store_vds];
      Von = m->dtype * M->currSto[in->li_store_von];
    } else {
      in->vbs_old = in->vbs;
      in->vbd_old = in->vbd;
      in->vgs_old = in->vgs;
      in->vds_old = in->vds;
      Von = 0.0;
    }
    in->vgd_old = in->vgs_old - in->vds_old;
  } else {
    in->vbs_old = M->nextSto[in->li_store_vbs];
    in->vbd_old = M->nextSto[in->li_store_vbd];
    in->vgs_old = M->nextSto[in->li_store_vgs];
    in->vds_old = M->nextSto[in->li_store_vds];
    Von = m->dtype * M->nextSto[in->li_store_von];
    in->vgd_old = in->vgs_old - in->vds_old;
  }

  if (M->opt.voltageLimiterFlag) {
    if (!(ss->initFixFlag && in->OFF)) {
      if (in->vds_old >= 0) {
        in->vgs = ref_fetlim(in->vgs, in->vgs_old, Von);
        in->vds = in->vgs - in->vgd;
        in->vds = ref_limvds(in->vds, in->vds_old);
        in->vgd = in->vgs - in->vds;
      } else {
        in->vgd = ref_fetlim(in->vgd, in->vgd_old, Von);
        in->vds = in->vgs - in->vgd;
        in->vds = -ref_limvds(-in->vds, -in->vds_old);
        in->vgs = in->vgd + in->vds;
      }
      if (in->vds >= 0.0) {
        in->vbs = ref_pnjlim(in->vbs, in->vbs_old, in->vt, in->sourceVcrit, &Check);
        in->vbd = in->vbs - in->vds;
      } else {
        in->vbd = ref_pnjlim(in->vbd, in->vbd_old, in->vt, in->drainVcrit, &Check);
        in->vbs = in->vbd + in->vds;
      }
      if (Check == 1) in->limitedFlag = 1;
    }
  }

  in->vbd = in->vbs - in->vds;
  in->vgd = in->vgs - in->vds;
  in->Vgb = in->vgs - in->vbs;

  if (in->vgs_orig != in->vgs || in->vds_orig != in->vds ||
      in->vbs_orig != in->vbs || in->vbd_orig != in->vbd ||
      in->vgd_orig != in->vgd)
    in->origFlag = 0;

  if (in->vbs <= 0) {
    in->gbs = in->SourceSatCur / in->vt;
    in->gbs += M->opt.gmin;
    in->cbs = in->gbs * in->vbs;
  } else {
    evbs = exp(rc_min(RC_MAX_EXP_ARG, in->vbs / in->vt));
    in->gbs = (in->SourceSatCur * evbs / in->vt + M->opt.gmin);
    in->cbs = (in->SourceSatCur * (evbs - 1) + M->opt.gmin * in->vbs);
  }
  if (in->vbd <= 0) {
    in->gbd = in->DrainSatCur / in->vt;
    in->gbd += M->opt.gmin;
    in->cbd = in->gbd * in->vbd;
  } else {
    evbd = exp(rc_min(RC_MAX_EXP_ARG, in->vbd / in->vt));
    in->gbd = (in->DrainSatCur * evbd / in->vt + M->opt.gmin);
    in->cbd = (in->DrainSatCur * (evbd - 1) + M->opt.gmin * in->vbd);
  }

  in->mode = (in->vds >= 0) ? 1 : -1;

  {
    double arg2, betap, sarg2, vgst;
    vds_save = in->vds;
    vgs_save = in->vgs;
    vgd_save = in->vgd;
    if ((in->mode == 1 ? in->vbs : in->vbd) <= 0) {
      sarg2 = sqrt(in->tPhi - (in->mode == 1 ? in->vbs : in->vbd));
    } else {
      sarg2 = sqrt(in->tPhi);
      sarg2 = sarg2 - (in->mode == 1 ? in->vbs : in->vbd) / (sarg2 + sarg2);
      sarg2 = rc_max(0.0, sarg2);
    }
    Von = (in->tVbi * m->dtype) + m->gamma * sarg2;
    vgst = (in->mode == 1 ? in->vgs : in->vgd) - Von;
    Vdsat = rc_max(vgst, 0.0);
    if (sarg2 <= 0) arg2 = 0;
    else arg2 = m->gamma / (sarg2 + sarg2);
    if (vgst <= 0) {
      in->cdrain = 0;
      in->gm = 0;
      in->gds = 0;
      in->gmbs = 0;
    } else {
      betap = Beta * (1 + m->lambda * (in->vds * in->mode));
      if (vgst <= (in->vds * in->mode)) {
        in->cdrain = betap * vgst * vgst * .5;
        in->gm = betap * vgst;
        in->gds = m->lambda * Beta * vgst * vgst * .5;
        in->gmbs = in->gm * arg2;
      } else {
        in->cdrain = betap * (in->vds * in->mode) * (vgst - .5 * (in->vds * in->mode));
        in->gm = betap * (in->vds * in->mode);
        in->gds = betap * (vgst - (in->vds * in->mode)) +
                  m->lambda * Beta * (in->vds * in->mode) * (vgst - .5 * (in->vds * in->mode));
        in->gmbs = in->gm * arg2;
      }
    }
  }

  in->von = m->dtype * Von;
  in->vdsat = m->dtype * Vdsat;
  in->cd = in->mode * in->cdrain - in->cbd;

  if (in->Cbs != 0 || in->Cbssw != 0) {
    if (in->vbs < in->tDepCap) {
      arg = 1 - in->vbs / in->tBulkPot;
      if (m->bulkJctBotGradingCoeff == m->bulkJctSideGradingCoeff) {
        if (m->bulkJctBotGradingCoeff == .5) sarg = sargsw = 1 / sqrt(arg);
        else sarg = sargsw = exp(-m->bulkJctBotGradingCoeff * log(arg));
      } else {
        if (m->bulkJctBotGradingCoeff == .5) sarg = 1 / sqrt(arg);
        else sarg = exp(-m->bulkJctBotGradingCoeff * log(arg));
        if (m->bulkJctSideGradingCoeff == .5) sargsw = 1 / sqrt(arg);
        else sargsw = exp(-m->bulkJctSideGradingCoeff * log(arg));
      }
      in->qbs = in->tBulkPot * (in->Cbs * (1 - arg * sarg) / (1 - m->bulkJctBotGradingCoeff) +
                                in->Cbssw * (1 - arg * sargsw) / (1 - m->bulkJctSideGradingCoeff));
      in->capbs = in->Cbs * sarg + in->Cbssw * sargsw;
    } else {
      in->qbs = in->f4s + in->vbs * (in->f2s + in->vbs * (in->f3s / 2));
      in->capbs = in->f2s + in->f3s * in->vbs;
    }
  } else {
    in->qbs = 0;
    in->capbs = 0;
  }

  if (in->Cbd != 0 || in->Cbdsw != 0) {
    if (in->vbd < in->tDepCap) {
      arg = 1 - in->vbd / in->tBulkPot;
      if (m->bulkJctBotGradingCoeff == .5 && m->bulkJctSideGradingCoeff == .5) {
        sarg = sargsw = 1 / sqrt(arg);
      } else {
        if (m->bulkJctBotGradingCoeff == .5) sarg = 1 / sqrt(arg);
        else sarg = exp(-m->bulkJctBotGradingCoeff * log(arg));
        if (m->bulkJctSideGradingCoeff == .5) sargsw = 1 / sqrt(arg);
        else sargsw = exp(-m->bulkJctSideGradingCoeff * log(arg));
      }
      in->qbd = in->tBulkPot * (in->Cbd * (1 - arg * sarg) / (1 - m->bulkJctBotGradingCoeff) +
                                in->Cbdsw * (1 - arg * sargsw) / (1 - m->bulkJctSideGradingCoeff));
      in->capbd = in->Cbd * sarg + in->Cbdsw * sargsw;
    } else {
      in->qbd = in->f4d + in->vbd * (in->f2d + in->vbd * in->f3d / 2);
      in->capbd = in->f2d + in->vbd * in->f3d;
    }
  } else {
    in->qbd = 0;
    in->capbd = 0;
  }

  if (in->mode > 0)
    ref_qmeyer(in->vgs, in->vgd, in->Vgb, Von, Vdsat, &in->capgs, &in->capgd,
               &in->capgb, in->tPhi, in->OxideCap);
  else
    ref_qmeyer(in->vgd, in->vgs, in->Vgb, Von, Vdsat, &in->capgd, &in->capgs,
               &in->capgb, in->tPhi, in->OxideCap);

  if (ss->dcopFlag) {
    in->Capgs = 2.0 * in->capgs + in->GateSourceOverlapCap;
    in->Capgd = 2.0 * in->capgd + in->GateDrainOverlapCap;
    in->Capgb = 2.0 * in->capgb + in->GateBulkOverlapCap;
  } else {
    capgs_old = M->currSta[in->li_state_capgs];
    capgd_old = M->currSta[in->li_state_capgd];
    capgb_old = M->currSta[in->li_state_capgb];
    in->Capgs = (in->capgs + capgs_old + in->GateSourceOverlapCap);
    in->Capgd = (in->capgd + capgd_old + in->GateDrainOverlapCap);
    in->Capgb = (in->capgb + capgb_old + in->GateBulkOverlapCap);
  }
  in->Capgs *= (in->Capgs < 0.0) ? -1.0 : 1.0;
  in->Capgd *= (in->Capgd < 0.0) ? -1.0 : 1.0;
  in->Capgb *= (in->Capgb < 0.0) ? -1.0 : 1.0;

  in->Idrain = in->drainConductance * in->Vddp;
  in->Isource = in->sourceConductance * in->Vssp;

  if (in->mode >= 0) {
    in->Gm = in->gm;
    in->Gmbs = in->gmbs;
    in->nrmsum = in->Gm + in->Gmbs;
    in->revsum = 0;
    in->cdreq = m->dtype * in->cdrain;
  } else {
    in->Gm = -in->gm;
    in->Gmbs = -in->gmbs;
    in->nrmsum = 0;
    in->revsum = -(in->Gm + in->Gmbs);
    in->cdreq = -(m->dtype) * in->cdrain;
  }
  in->vds = vds_save;
  in->vgs = vgs_save;
  in->vgd = vgd_save;
}

/* Master::updateState, N_DEV_MOSFET1.C:4103-4192 */
void ref_mos1_update_state(ref_mos1_master *M, double *nextSol,
                           double *currSta, double *currSto, double *nextSta,
                           double *nextSto) {
  int i;
  M->nextSol = nextSol; M->currSta = currSta; M->currSto = currSto;
  M->nextSta = nextSta; M->nextSto = nextSto;
  for (i = 0; i < M->n; ++i) {
    ref_mos1_inst *mi = &M->inst[i];
    double *staVec = nextSta, *oldstaVec = currSta, *stoVec = nextSto,
           *oldstoVec = currSto;
    double vgs1 = 0.0, vgd1 = 0.0, vbs1 = 0.0, vgb1 = 0.0, vds1 = 0.0;
    mos1_uiv(M, mi);
    stoVec[mi->li_store_vbd] = mi->vbd;
    stoVec[mi->li_store_vbs] = mi->vbs;
    stoVec[mi->li_store_vgs] = mi->vgs;
    stoVec[mi->li_store_vds] = mi->vds;
    stoVec[mi->li_store_von] = mi->von;
    staVec[mi->li_state_capgs] = mi->capgs;
    staVec[mi->li_state_capgd] = mi->capgd;
    staVec[mi->li_state_capgb] = mi->capgb;
    if (M->ss.dcopFlag) {
      mi->qgs = mi->Capgs * mi->vgs;
      mi->qgd = mi->Capgd * mi->vgd;
      mi->qgb = mi->Capgb * mi->Vgb;
    } else {
      mi->qgs = oldstaVec[mi->li_state_qgs];
      mi->qgd = oldstaVec[mi->li_state_qgd];
      mi->qgb = oldstaVec[mi->li_state_qgb];
      vgs1 = oldstoVec[mi->li_store_vgs];
      vbs1 = oldstoVec[mi->li_store_vbs];
      vds1 = oldstoVec[mi->li_store_vds];
      vgb1 = vgs1 - vbs1;
      vgd1 = vgs1 - vds1;
      mi->qgs += mi->Capgs * (mi->vgs - vgs1);
      mi->qgd += mi->Capgd * (mi->vgd - vgd1);
      mi->qgb += mi->Capgb * ((mi->vgs - mi->vbs) - vgb1);
    }
    staVec[mi->li_state_qgs] = mi->qgs;
    staVec[mi->li_state_qgd] = mi->qgd;
    staVec[mi->li_state_qgb] = mi->qgb;
    staVec[mi->li_state_qbd] = mi->qbd;
    staVec[mi->li_state_qbs] = mi->qbs;
  }
}

/* Master::loadDAEVectors, N_DEV_MOSFET1.C:4202-4430 (no lead currents). */
void ref_mos1_load_vectors(ref_mos1_master *M, double *fVec, double *qVec,
                           double *dFdxdVp, double *dQdxdVp) {
  double gmin1 = M->opt.gmin;
  int i;
  for (i = 0; i < M->n; ++i) {
    ref_mos1_inst *mi = &M->inst[i];
    int Dtype = mi->model->dtype;
    double ceqbs, ceqbd, ceqgb, ceqgs, ceqgd, Qeqbs, Qeqbd, Qeqgb, Qeqgs, Qeqgd, coef;
    ceqbs = Dtype * (mi->cbs);
    ceqbd = Dtype * (mi->cbd);
    ceqgb = 0.0; ceqgs = 0.0; ceqgd = 0.0;
    if (mi->drainConductance != 0.0) fVec[mi->li_Drain] += mi->Idrain * mi->numberParallel;
    coef = (ceqgs + ceqgd + ceqgb);
    fVec[mi->li_Gate] += coef * mi->numberParallel;
    if (mi->sourceConductance != 0.0) fVec[mi->li_Source] += mi->Isource * mi->numberParallel;
    coef = ceqbs + ceqbd - ceqgb;
    fVec[mi->li_Bulk] += coef * mi->numberParallel;
    coef = -mi->Idrain - (ceqbd - mi->cdreq + ceqgd);
    fVec[mi->li_DrainPrime] += coef * mi->numberParallel;
    coef = -mi->Isource - (ceqbs + mi->cdreq + ceqgs);
    fVec[mi->li_SourcePrime] += coef * mi->numberParallel;

    Qeqbs = Dtype * (mi->qbs);
    Qeqbd = Dtype * (mi->qbd);
    Qeqgb = Dtype * (mi->qgb);
    Qeqgs = Dtype * (mi->qgs);
    Qeqgd = Dtype * (mi->qgd);
    coef = (Qeqgs + Qeqgd + Qeqgb);
    qVec[mi->li_Gate] += coef * mi->numberParallel;
    coef = Qeqbs + Qeqbd - Qeqgb;
    qVec[mi->li_Bulk] += coef * mi->numberParallel;
    coef = -(Qeqbd + Qeqgd);
    qVec[mi->li_DrainPrime] += coef * mi->numberParallel;
    coef = -(Qeqbs + Qeqgs);
    qVec[mi->li_SourcePrime] += coef * mi->numberParallel;

    if (!mi->origFlag) {
      double coef_Jdxp4 = Dtype * (+((mi->gbd - gmin1)) * (mi->vbd - mi->vbd_orig) +
                                   ((mi->gbs - gmin1)) * (mi->vbs - mi->vbs_orig));
      double coef_Jdxp5 = Dtype * (-((mi->gbd - gmin1)) * (mi->vbd - mi->vbd_orig) +
                                   mi->gds * (mi->vds - mi->vds_orig) +
                                   mi->Gm * ((mi->mode > 0) ? (mi->vgs - mi->vgs_orig) : (mi->vgd - mi->vgd_orig)) +
                                   mi->Gmbs * ((mi->mode > 0) ? (mi->vbs - mi->vbs_orig) : (mi->vbd - mi->vbd_orig)));
      double coef_Jdxp6 = Dtype * (-((mi->gbs - gmin1)) * (mi->vbs - mi->vbs_orig) -
                                   mi->gds * (mi->vds - mi->vds_orig) -
                                   mi->Gm * ((mi->mode > 0) ? (mi->vgs - mi->vgs_orig) : (mi->vgd - mi->vgd_orig)) -
                                   mi->Gmbs * ((mi->mode > 0) ? (mi->vbs - mi->vbs_orig) : (mi->vbd - mi->vbd_orig)));
      double gcgd = 0.0, gcgs = 0.0, gcgb = 0.0, gcbs = 0.0, gcbd = 0.0;
      double coef_Jdxp2, coef_Jdxp4q, coef_Jdxp5q, coef_Jdxp6q;
      dFdxdVp[mi->li_Bulk] += coef_Jdxp4 * mi->numberParallel;
      dFdxdVp[mi->li_DrainPrime] += coef_Jdxp5 * mi->numberParallel;
      dFdxdVp[mi->li_SourcePrime] += coef_Jdxp6 * mi->numberParallel;
      if (M->ss.tranopFlag || M->ss.acopFlag || M->ss.transientFlag) {
        gcgd = mi->Capgd; gcgs = mi->Capgs; gcgb = mi->Capgb;
        gcbs = mi->capbs; gcbd = mi->capbd;
      }
      coef_Jdxp2 = Dtype * (gcgd * (mi->vgd - mi->vgd_orig) + gcgs * (mi->vgs - mi->vgs_orig) +
                            gcgb * (mi->vgs - mi->vgs_orig - mi->vbs + mi->vbs_orig));
      coef_Jdxp4q = Dtype * (-(gcgb) * (mi->vgs - mi->vgs_orig - mi->vbs + mi->vbs_orig) +
                             (gcgb) * (mi->vbd - mi->vbd_orig) + (gcbs) * (mi->vbs - mi->vbs_orig));
      coef_Jdxp5q = Dtype * (-(gcgd) * (mi->vgd - mi->vgd_orig) - (gcbd) * (mi->vbd - mi->vbd_orig));
      coef_Jdxp6q = Dtype * (-gcgs * (mi->vgs - mi->vgs_orig) - (gcbs) * (mi->vbs - mi->vbs_orig));
      dQdxdVp[mi->li_Gate] += coef_Jdxp2 * mi->numberParallel;
      dQdxdVp[mi->li_Bulk] += coef_Jdxp4q * mi->numberParallel;
      dQdxdVp[mi->li_DrainPrime] += coef_Jdxp5q * mi->numberParallel;
      dQdxdVp[mi->li_SourcePrime] += coef_Jdxp6q * mi->numberParallel;
    }
  }
}

/* Master::loadDAEMatrices (pointer form), N_DEV_MOSFET1.C:4441-4722 */
void ref_mos1_load_matrices(ref_mos1_master *M) {
  int i;
  for (i = 0; i < M->n; ++i) {
    ref_mos1_inst *mi = &M->inst[i];
    double np = mi->numberParallel;
    double gcgd = 0.0, gcgs = 0.0, gcgb = 0.0, gcbs = 0.0, gcbd = 0.0;
    *mi->fp[O_D_D] += mi->drainConductance * np;
    *mi->fp[O_D_DP] -= mi->drainConductance * np;
    *mi->fp[O_S_S] += mi->sourceConductance * np;
    *mi->fp[O_S_SP] -= mi->sourceConductance * np;
    *mi->fp[O_B_B] += (mi->gbs + mi->gbd) * np;
    *mi->fp[O_B_DP] -= mi->gbd * np;
    *mi->fp[O_B_SP] -= mi->gbs * np;
    *mi->fp[O_DP_D] -= mi->drainConductance * np;
    *mi->fp[O_DP_G] += (mi->Gm) * np;
    *mi->fp[O_DP_B] += (-mi->gbd + mi->Gmbs) * np;
    *mi->fp[O_DP_DP] += (mi->drainConductance + mi->gds + mi->gbd + mi->revsum) * np;
    *mi->fp[O_DP_SP] += (-mi->gds - mi->nrmsum) * np;
    *mi->fp[O_SP_G] -= (mi->Gm) * np;
    *mi->fp[O_SP_S] -= mi->sourceConductance * np;
    *mi->fp[O_SP_B] -= (mi->gbs + mi->Gmbs) * np;
    *mi->fp[O_SP_DP] -= (mi->gds + mi->revsum) * np;
    *mi->fp[O_SP_SP] += (mi->sourceConductance + mi->gds + mi->gbs + mi->nrmsum) * np;

    if (M->ss.tranopFlag || M->ss.acopFlag || M->ss.transientFlag) {
      gcgd = mi->Capgd; gcgs = mi->Capgs; gcgb = mi->Capgb;
      gcbs = mi->capbs; gcbd = mi->capbd;
    }
    *mi->qp[O_G_G] += (gcgd + gcgs + gcgb) * np;
    *mi->qp[O_G_B] -= gcgb * np;
    *mi->qp[O_G_DP] -= gcgd * np;
    *mi->qp[O_G_SP] -= gcgs * np;
    *mi->qp[O_B_G] -= gcgb * np;
    *mi->qp[O_B_B] += (+gcbs + gcbd + gcgb) * np;
    *mi->qp[O_B_DP] -= +gcbd * np;
    *mi->qp[O_B_SP] -= +gcbs * np;
    *mi->qp[O_DP_G] += -gcgd * np;
    *mi->qp[O_DP_B] += -gcbd * np;
    *mi->qp[O_DP_DP] += (+gcbd + gcgd) * np;
    *mi->qp[O_SP_G] -= gcgs * np;
    *mi->qp[O_SP_B] -= +gcbs * np;
    *mi->qp[O_SP_SP] += (+gcbs + gcgs) * np;
  }
}

int ref_mos1_count(const ref_mos1_master *M) { return M->n; }

/* Per-instance flags after the last updateState: limitedFlag, origFlag. */
void ref_mos1_flags(const ref_mos1_master *M, int *limited, int *orig) {
  int i;
  for (i = 0; i < M->n; ++i) {
    limited[i] = M->inst[i].limitedFlag;
    orig[i] = M->inst[i].origFlag;
  }
}

/* ---- Probes for pinning tests (no counterpart in the reference) ---- */

/* The SPICE3 limiters and Meyer capacitances of ref_common.h as callable
   functions (DeviceSupport::fetlim / limvds / pnjlim / qmeyer,
   N_DEV_DeviceSupport.C:186-251, 337-421, 545-614). */
double ref_x_fetlim(double vnew, double vold, double vto) {
  return ref_fetlim(vnew, vold, vto);
}
double ref_x_limvds(double vnew, double vold) { return ref_limvds(vnew, vold); }
double ref_x_pnjlim(double vnew, double vold, double vt, double vcrit, int *icheck) {
  return ref_pnjlim(vnew, vold, vt, vcrit, icheck);
}
void ref_x_qmeyer(double vgs, double vgd, double vgb, double von, double vdsat,
                  double phi, double cox, double *out3) {
  ref_qmeyer(vgs, vgd, vgb, von, vdsat, &out3[0], &out3[1], &out3[2], phi, cox);
}

/* Temperature- and geometry-derived fields of instance i
   (Instance::updateTemperature, N_DEV_MOSFET1.C:3437-3680): out[0..13] =
   vt, tPhi, tVto, tVbi, tSatCur, tTransconductance, tSurfMob, tBulkPot,
   tDepCap, tCbd, tCbs, sourceVcrit, EffectiveLength, OxideCap. */
void ref_mos1_probe(const ref_mos1_master *M, int i, double *out) {
  const ref_mos1_inst *in = &M->inst[i];
  double v[14] = {in->vt, in->tPhi, in->tVto, in->tVbi, in->tSatCur,
                  in->tTransconductance, in->tSurfMob, in->tBulkPot, in->tDepCap,
                  in->tCbd, in->tCbs, in->sourceVcrit, in->EffectiveLength,
                  in->OxideCap};
  int k;
  for (k = 0; k < 14; ++k) out[k] = v[k];
}
