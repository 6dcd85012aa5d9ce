"""ORACLE — test infrastructure only; the product never imports this module.

Restatement of the level-1 junction diode with RS = 0 (N_DEV_Diode.C), in
Python floats (IEEE doubles), statement for statement:

* Model::processParams (:1772-1793): M <= 0.9, EG >= 0.1, FC <= 0.95,
  F2 = exp((1+M) log(1-FC)), F3 = 1 - FC (1+M);
* Instance::updateTemperature (:1332-1470): junction potential and
  capacitance, saturation current, tVcrit, tIRF, breakdown voltage (the
  default Picard iteration, :1456-1500);
* Master::updateState (:2007-2025): the store gets the instance's junction
  voltage from the previous evaluation, then updateIntermediateVars
  (:873-1316): initial junction voltage, pnjlim (N_DEV_DeviceSupport.C:
  214-263), the normal-exponential / linear-reverse / breakdown current
  (:1083-1158), depletion and diffusion charge (:1264-1308);
* Master::loadDAEVectors (:2034-2080) and loadDAEMatrices (:2088-2108) with
  the internal node merged into the positive one.
"""
import math

CONST_Q = 1.6021918e-19
CONST_BOLTZ = 1.3806226e-23
CONST_KOVERQ = CONST_BOLTZ / CONST_Q
CONST_REFTEMP = 300.15
MAX_EXP_ARG = 100.0


def pnjlim(vnew, vold, vt, vcrit):
    """DeviceSupport::pnjlim; returns (v, limited)."""
    if vnew > vcrit and abs(vnew - vold) > (vt + vt):
        if vold > 0:
            arg = 1 + (vnew - vold) / vt
            vnew = vold + vt * math.log(arg) if arg > 0 else vcrit
        else:
            vnew = vt * math.log(vnew / vt)
        return vnew, True
    return vnew, False


def diode_fields(mp, ip, temp, tnom):
    """Temperature- and area-scaled fields of one instance (TNOM / TEMP
    parameters in Celsius; temp, tnom the circuit's, in kelvin)."""
    IS = mp["IS"] if "IS" in mp else mp.get("JS", 1.0e-14)
    N = mp.get("N", 1.0)
    TT = mp.get("TT", 0.0)
    CJO = mp["CJO"] if "CJO" in mp else mp["CJ"] if "CJ" in mp else mp.get("CJ0", 0.0)
    VJ = mp.get("VJ", 1.0)
    M, EG, FC = mp.get("M", 0.5), mp.get("EG", 1.11), mp.get("FC", 0.5)
    XTI = mp.get("XTI", 3.0)
    bv_given = "BV" in mp or "VB" in mp
    BV = mp["BV"] if "BV" in mp else mp.get("VB", 1e99)
    IBV, IRF = mp.get("IBV", 1.0e-3), mp.get("IRF", 1.0)
    TNOM = mp["TNOM"] + 273.15 if "TNOM" in mp else tnom
    Temp = ip["TEMP"] + 273.15 if "TEMP" in ip else temp
    Area = ip.get("AREA", 1.0)
    M = min(M, 0.9)
    EG = max(EG, 0.1)
    FC = min(FC, 0.95)
    F2 = math.exp((1.0 + M) * math.log(1.0 - FC))
    F3 = 1.0 - FC * (1.0 + M)
    vtnom = CONST_KOVERQ * TNOM
    xfc = math.log(1.0 - FC)
    vt = CONST_KOVERQ * Temp
    fact2 = Temp / CONST_REFTEMP
    egfet = 1.16 - (7.02e-4 * Temp * Temp) / (Temp + 1108.0)
    arg = -egfet / (2.0 * CONST_BOLTZ * Temp) + \
        1.1150877 / (CONST_BOLTZ * (CONST_REFTEMP + CONST_REFTEMP))
    pbfact = -2.0 * vt * (1.5 * math.log(fact2) + CONST_Q * arg)
    egfet1 = 1.16 - (7.02e-4 * TNOM * TNOM) / (TNOM + 1108.0)
    arg1 = -egfet1 / (2.0 * CONST_BOLTZ * TNOM) + 1.1150877 / (2.0 * CONST_BOLTZ * CONST_REFTEMP)
    fact1 = TNOM / CONST_REFTEMP
    pbfact1 = -2.0 * vtnom * (1.5 * math.log(fact1) + CONST_Q * arg1)
    pbo = (VJ - pbfact1) / fact1
    gmaold = (VJ - pbo) / pbo
    tJctCap = CJO / (1.0 + M * (4.0e-4 * (TNOM - CONST_REFTEMP) - gmaold))
    tJctPot = pbfact + fact2 * pbo
    gmanew = (tJctPot - pbo) / pbo
    tJctCap *= 1.0 + M * (4.0e-4 * (Temp - CONST_REFTEMP) - gmanew)
    tSatCur = IS * math.exp(((Temp / TNOM) - 1.0) * EG / (N * vt) +
                            XTI / N * math.log(Temp / TNOM))
    tF1 = tJctPot * (1.0 - math.exp((1.0 - M) * xfc)) / (1.0 - M)
    tDepCap = FC * tJctPot
    vte = N * vt
    tVcrit = vte * math.log(vte / (math.sqrt(2.0) * tSatCur))
    tIRF = IRF * math.pow(fact2, 1.6)
    tBrkdwnV = 0.0
    if bv_given:
        e = math.exp(1.0)
        a1 = 3.0 * vte / (e * BV)
        a1 = a1 * a1 * a1
        lo = tSatCur * tIRF * (1 - a1)
        hi = tSatCur * tIRF * (1 - 1.0 / (e * e * e)) * math.exp(-1.0 * (3.0 * vte - BV) / vte)
        if IBV >= hi:
            tBrkdwnV = 3.0 * vte
        elif IBV <= lo:
            tBrkdwnV = BV
        else:
            xbv = BV
            for _ in range(8):
                a2 = 3.0 * vte / (e * xbv)
                a2 = a2 * a2 * a2
                xbv = BV - vte * math.log(IBV / (tSatCur * tIRF)) + vte * math.log(1 - a2)
            tBrkdwnV = xbv
    return dict(Isat=tSatCur * Area, Vte=vte, Vcrit=tVcrit, IRF=tIRF, Brk=tBrkdwnV,
                BVGiven=bv_given, BV=BV, tJctCap=tJctCap, Czero=tJctCap * Area,
                tJctPot=tJctPot, tDepCap=tDepCap, tF1=tF1, M=M, F2=F2, F3=F3, TT=TT,
                Gspr=0.0, off=bool(ip.get("OFF", 0.0)), ic="IC" in ip, IC=ip.get("IC", 0.0))


class RefDiode:
    def __init__(self, f, p, q, sto, pos_f, pos_q):
        self.f = f
        self.p, self.q, self.sto = p, q, sto
        self.pos_f, self.pos_q = pos_f, pos_q   # 7 dF/dx, 4 dQ/dx CSR positions
        self.Vd = 0.0                            # the instance's junction voltage

    def update_state(self, x, sto_c, sto_n, flags, gmin, limiter):
        f = self.f
        sto_n[self.sto] = self.Vd                # Master::updateState
        Vpp, Vn = float(x[self.p]), float(x[self.q])
        Vd = Vpp - Vn
        Vd_orig = Vd
        orig = True
        Vte, vcrit = f["Vte"], f["Vcrit"]
        if flags["newtonIter"] == 0:
            if flags["initJct"] and limiter:
                Vd = f["IC"] if f["ic"] else 0.0 if f["off"] else vcrit
                orig = False
            Vd_old = Vd
            if not flags["dcop"]:
                Vd_old = float(sto_c[self.sto])
        else:
            Vd_old = float(sto_n[self.sto])
        if limiter:
            BV = f["BV"]
            if f["BVGiven"] and Vd < min(0.0, -BV + 10.0 * Vte):
                Vdtmp = -(BV + Vd)
                Vdtmp, lim = pnjlim(Vdtmp, -(Vd_old + BV), Vte, vcrit)
                Vd = -(Vdtmp + BV)
            else:
                Vd, lim = pnjlim(Vd, Vd_old, Vte, vcrit)
            if lim:
                orig = False
        Isat = f["Isat"] * (1.0 if Vd >= -3.0 * Vte else f["IRF"])
        brk = f["Brk"]
        if Vd >= -3.0 * Vte:
            evd = math.exp(min(MAX_EXP_ARG, Vd / Vte))
            Id = Isat * (evd - 1.0) + gmin * Vd
            Gd = Isat * evd / Vte + gmin
        elif not brk or Vd >= -brk:
            arg = 3.0 * Vte / (Vd * math.exp(1.0))
            arg = arg * arg * arg
            Id = -Isat * (1.0 + arg) + gmin * Vd
            Gd = Isat * 3.0 * arg / Vd + gmin
        else:
            evrev = math.exp(min(MAX_EXP_ARG, -(brk + Vd) / Vte))
            arg2 = 3.0 * Vte / (math.exp(1.0) * brk)
            arg2 = arg2 * arg2 * arg2
            Isat_b = Isat * (1 - arg2)
            Id = -Isat_b * evrev + gmin * Vd
            Gd = Isat_b * evrev / Vte + gmin
        Vc = Vd
        Qd = Cd = 0.0
        if f["tJctCap"] != 0.0:
            Czero, pot, dep, M, TT = f["Czero"], f["tJctPot"], f["tDepCap"], f["M"], f["TT"]
            if Vc < dep:
                arg = 1.0 - Vc / pot
                sarg = math.exp(min(MAX_EXP_ARG, -M * math.log(arg)))
                Qd = TT * Id + pot * Czero * (1.0 - arg * sarg) / (1.0 - M)
                Cd = TT * Gd + Czero * sarg
            else:
                Czof2 = Czero / f["F2"]
                MotJctPot = M / pot
                Qd = TT * Id + Czero * f["tF1"] + Czof2 * (
                    f["F3"] * (Vc - dep) + (0.5 * MotJctPot) * (Vc * Vc - dep * dep))
                Cd = TT * Gd + Czof2 * (f["F3"] + MotJctPot * Vc)
        self.Vd, self.Vd_orig, self.orig = Vd, Vd_orig, orig
        self.Id, self.Gd, self.Qd, self.Cd = Id, Gd, Qd, Cd

    def load_vectors(self, F, Q, FL, QL, limiter):
        p, q, Gspr = self.p, self.q, self.f["Gspr"]
        Ir = Gspr * (float(0.0))                 # Vp - Vpp = 0 with RS = 0
        F[p] -= -Ir
        F[q] -= self.Id
        F[p] -= (-self.Id + Ir)
        Q[q] -= self.Qd
        Q[p] -= -self.Qd
        if limiter:
            Vd_diff = self.Vd - self.Vd_orig
            Cd_Jdxp = -(self.Cd) * Vd_diff
            Gd_Jdxp = -(self.Gd) * Vd_diff
            FL[q] += Gd_Jdxp
            FL[p] -= Gd_Jdxp
            QL[q] += Cd_Jdxp
            QL[p] -= Cd_Jdxp

    def load_matrices(self, dFdx, dQdx):
        a = self.pos_f
        Gspr, Gd, Cd = self.f["Gspr"], self.Gd, self.Cd
        dFdx[a[0]] += Gspr
        dFdx[a[1]] -= Gspr
        dFdx[a[2]] += Gd
        dFdx[a[3]] -= Gd
        dFdx[a[4]] -= Gspr
        dFdx[a[5]] -= Gd
        dFdx[a[6]] += Gspr + Gd
        b = self.pos_q
        dQdx[b[0]] += Cd
        dQdx[b[1]] -= Cd
        dQdx[b[2]] -= Cd
        dQdx[b[3]] += Cd
