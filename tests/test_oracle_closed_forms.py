"""Closed-form pins of the oracle's BSIM3, MOSFET1, limiter and diode
functions: each compares the oracle (the CPU restatement of the reference)
with a formula written here from the published model equations, in a limit
where that formula is exact or its error is bounded and stated, so that a
plausible mistake in the restatement (a dropped term, a wrong sign, a
transposed exponent) fails.

Sources of the formulas (not of the code under test):
* BSIM3v3.2.2 MOSFET Model User's Manual (UC Berkeley, 1999), Appendix A
  (parameter list) and Appendix B (equations): surface potential, effective
  geometry and binning, threshold voltage, subthreshold swing, Vgsteff,
  mobility, Vdsat, substrate current and the temperature laws;
* SPICE3 level-1 MOSFET and junction-diode temperature dependence
  (Vladimirescu, "The SPICE Book", ch. 3-4; Sze, "Physics of Semiconductor
  Devices": n_i ~ T^1.5 exp(-Eg/2kT), Eg(T) = Eg(0) - aT^2/(T+b));
* Meyer's gate-capacitance model (Meyer, RCA Review 1971): Cgs = 2/3 Cox W L
  in saturation, Cgs = Cgd = 1/2 Cox W L at Vds = 0, Cgb = Cox W L in
  accumulation;
* the SPICE3 fetlim / limvds definitions, evaluated by hand into tables.

Physical constants are the reference's (N_DEV_Const.h:50-106).
"""
import math
import os
import re

import numpy as np
import pytest

from xyce_amd import circuits
from xyce_amd.circuits import Netlist

Q = 1.6021918e-19
KB = 1.3806226e-23
EPSOX = 3.453133e-11
EPSSI = 1.03594e-10
# level 1 takes the oxide permittivity as 3.9 eps0 (CONSTpermOxide,
# N_DEV_Const.h:60-61); BSIM3 takes CONSTEPSOX (:78), 3.1e-6 smaller
EPSOX_M1 = 3.9 * 8.854214871e-12
T27 = 300.15


def vt_of(T):
    return KB * T / Q


def eg_of(T):
    """Silicon band gap (eV) of Varshni's form with the SPICE constants."""
    return 1.16 - 7.02e-4 * T * T / (T + 1108.0)


def phi_at(phi_nom, tnom, T):
    """A 2-phi_F-type potential (surface potential, junction potential) at T
    from its value at tnom: phi = 2 vt ln(N / n_i) with n_i ~ T^1.5 exp(-Eg /
    2kT) gives phi(T) = T/tnom (phi_nom - Eg(tnom)) - 3 vt(T) ln(T/tnom) +
    Eg(T)."""
    return T / tnom * (phi_nom - eg_of(tnom)) - 3.0 * vt_of(T) * math.log(T / tnom) + eg_of(T)


# ------------------------------------------------------------------ helpers
def single(kind, dtype, card, inst, nodes=("d", "g", "s", "b")):
    nl = Netlist("one %s" % kind)
    nl.add_model("M", kind, dtype, **card)
    nl.add_mos("X1", *nodes, "M", **inst)
    nl.layout()
    return nl


def oracle_load(nl, volts, ss, gmin=0.0, limiter=False, temp=T27, sta=None):
    """Full oracle load of a one-transistor netlist at the given terminal
    voltages; the old store equals the new voltages, so no limiter moves
    them.  Returns (F, Q, JF, JQ, sta_next, sto_next, ref)."""
    from oracle.oracle import RefLoad
    x = np.zeros(nl.n_nodes + 1)
    for name, v in volts.items():
        x[nl.node(name)] = v
    ref = RefLoad(nl, gmin=gmin, limiter=limiter, temp=temp)
    ref.set_solver_state(**ss)
    sta_c = np.zeros(nl.n_state) if sta is None else sta.copy()
    sto = np.zeros(nl.n_store)
    sta_n = np.zeros(nl.n_state)
    out = ref.full_load(x, sta_c, sto.copy(), sta_n, sto)
    n1 = nl.n_nodes + 1
    JF = np.zeros((n1, n1))
    JQ = np.zeros((n1, n1))
    for r in range(n1):
        for k in range(ref.rowptr[r], ref.rowptr[r + 1]):
            JF[r, ref.colidx[k]] += out[4][k]
            JQ[r, ref.colidx[k]] += out[5][k]
    return out[0], out[1], JF, JQ, sta_n, sto, ref


DC = dict(dcop=1, newtonIter=1)


# ------------------------------------------------------------------ BSIM3
# A card whose short-channel, narrow-width, DIBL and temperature terms are
# switched off (K3 = K3B = NLX = DVT0 = DVT0W = ETA0 = ETAB = 0, CDSC* = 0,
# long channel so the remaining exp(-L/lt) factors underflow), with no
# series resistance, no junction currents, no substrate current and no
# mobility degradation, unless a test turns one of them back on.
B3_BASE = dict(TOX=4e-9, NCH=1.7e17, VTH0=0.4, K1=0.5, K2=0.02, K3=0.0, K3B=0.0,
               NLX=0.0, DVT0=0.0, DVT0W=0.0, ETA0=0.0, ETAB=0.0, U0=400.0, UA=0.0,
               UB=0.0, UC=0.0, VSAT=1.0e5, A0=1.0, AGS=0.0, B0=0.0, B1=0.0, KETA=0.0,
               RDSW=0.0, PRWG=0.0, PRWB=0.0, WINT=0.0, LINT=0.0, DWG=0.0, DWB=0.0,
               VOFF=-0.08, NFACTOR=1.0, CIT=0.0, CDSC=0.0, CDSCB=0.0, CDSCD=0.0,
               PCLM=0.0, PDIBLC1=0.0, PDIBLC2=0.0, PDIBLCB=0.0, PVAG=0.0, PSCBE2=0.0,
               DELTA=1e-10, ALPHA0=0.0, ALPHA1=0.0, JS=0.0, JSW=0.0, XJ=1.5e-7)
# junction area and perimeter given with JS = JSW = 0: no junction current
# (with neither given the reference falls back to Is = 1e-14,
# N_DEV_MOSFET_B3.C:5039-5042)
JCT = dict(AD=1e-12, AS=1e-12, PD=4e-6, PS=4e-6)
LONG = dict(W=10e-6, L=10e-6, **JCT)


def manual_phi(nch, tnom=T27):
    """Surface potential 2 Vtm0 ln(NCH / n_i), n_i = 1.45e10 (Tnom/300.15)^1.5
    exp(21.5565981 - Eg0 / 2 Vtm0) (NCH and n_i in cm^-3)."""
    vtm0 = vt_of(tnom)
    ni = 1.45e10 * (tnom / T27) ** 1.5 * math.exp(21.5565981 - eg_of(tnom) / (2 * vtm0))
    return 2 * vtm0 * math.log(nch / ni)


def cox_of(tox):
    return EPSOX / tox


def _b3(card, inst, volts, ss=DC, temp=T27):
    nl = single("bsim3", 1, card, inst)
    F, Qv, JF, JQ, sta, sto, ref = oracle_load(nl, volts, ss, temp=temp)
    return nl, F, Qv, JF, JQ, sta, sto, ref


def _field_index():
    """Size-dependent record field names of the load library's BSIM3 layout
    (read from its header), for xg_get_bsim3_fields."""
    here = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    text = open(os.path.join(here, "xyce_amd", "csrc", "xg_bsim3_layout.h")).read()
    body = re.search(r"enum B3SdpField : int \{(.*?)\};", text, re.S).group(1)
    names = [t.strip() for t in body.replace("\n", " ").split(",") if t.strip()]
    idx = {}
    for k, nm in enumerate(names):
        nm = nm.split("=")[0].strip()
        if nm == "B3R_N":
            break
        idx[nm[4:]] = k
    return idx


def _library_fields(nl, temp=T27):
    from xyce_amd.loader import DeviceLoad
    dev = DeviceLoad(nl, plan_only=True, temp=temp, tnom=T27)
    f = dev.bsim3_fields()
    dev.close()
    return f[0], _field_index()


GEO_CARD = dict(B3_BASE, LINT=2e-8, LL=1e-14, LLN=1.5, LW=3e-14, LWN=0.8, LWL=2e-21,
                WINT=3e-8, WL=4e-14, WLN=1.2, WW=5e-14, WWN=0.9, WWL=1e-21,
                LVTH0=0.02, WVTH0=-0.01, PVTH0=0.003, LU0=30.0, WU0=-20.0, PU0=5.0,
                NCH=2.3e17)
GEO_INST = dict(W=1.3e-6, L=0.35e-6, **JCT)


def _manual_geometry(card, inst):
    """Effective L and W (dL, dW of the drawn sizes) and binned parameters
    P = P0 + PL/L' + PW/W' + PP/(L'W'), L' and W' in microns (BINUNIT 1)."""
    L, W = inst["L"], inst["W"]
    dl = (card["LINT"] + card["LL"] / L ** card["LLN"] + card["LW"] / W ** card["LWN"]
          + card["LWL"] / (L ** card["LLN"] * W ** card["LWN"]))
    dw = (card["WINT"] + card["WL"] / L ** card["WLN"] + card["WW"] / W ** card["WWN"]
          + card["WWL"] / (L ** card["WLN"] * W ** card["WWN"]))
    leff, weff = L - 2 * dl, W - 2 * dw
    il, iw = 1e-6 / leff, 1e-6 / weff

    def binned(p):
        return card[p] + card["L" + p] * il + card["W" + p] * iw + card["P" + p] * il * iw
    return leff, weff, binned


def test_bsim3_geometry_binning_and_surface_potential(oracle_lib):
    """Leff, Weff, binned VTH0 and U0 and the surface potential, in the
    oracle's size-dependent record and in the load library's fields."""
    nl, *_, ref = _b3(GEO_CARD, GEO_INST, dict(d=0.1))
    leff, weff, binned = _manual_geometry(GEO_CARD, GEO_INST)
    p = ref.b3.sdp(0)
    phi = manual_phi(GEO_CARD["NCH"])
    want = dict(LEFF=leff, WEFF=weff, VTH0=binned("VTH0"), PHI=phi,
                U0TEMP=binned("U0") / 1e4)
    got = dict(LEFF=p[0], WEFF=p[1], VTH0=p[4], PHI=p[10], U0TEMP=p[11])
    for k in want:
        assert abs(got[k] - want[k]) <= 1e-13 * abs(want[k]), (k, got[k], want[k])
    # Xdep0 = sqrt(2 eps_si phi / (q NCH)), NCH in m^-3
    xdep0 = math.sqrt(2 * EPSSI * phi / (Q * GEO_CARD["NCH"] * 1e6))
    assert abs(p[24] - xdep0) <= 1e-12 * xdep0
    ref.close()
    f, ix = _library_fields(nl)
    for k in want:
        assert abs(f[ix[k]] - want[k]) <= 1e-13 * abs(want[k]), ("library", k)


def test_bsim3_k1_k2_from_doping_profile(oracle_lib):
    """K1 and K2 derived from the two-layer doping when not given:
    gamma_i = sqrt(2 q eps_si N_i) / Cox, Vbx = phi - q NCH XT^2 / (2 eps_si),
    K2 = (g1 - g2)(sqrt(phi - Vbx) - sqrt(phi)) /
         (2 sqrt(phi)(sqrt(phi - Vbm) - sqrt(phi)) + Vbm),
    K1 = g2 - 2 K2 sqrt(phi - Vbm).  The reference's gamma and Vbx
    prefactors are 4-digit roundings of these constants (5.753e-12 vs
    5.7615e-12, 7.7348e-4 vs 7.7331e-4), hence the 3e-3 tolerance."""
    card = {k: v for k, v in B3_BASE.items() if k not in ("K1", "K2")}
    card.update(NSUB=6e16, XT=1.55e-7, VBM=-3.0)
    nl, *_, ref = _b3(card, LONG, dict(d=0.1))
    p = ref.b3.sdp(0)
    cox = cox_of(card["TOX"])
    phi = manual_phi(card["NCH"])
    g1 = math.sqrt(2 * Q * EPSSI * card["NCH"] * 1e6) / cox
    g2 = math.sqrt(2 * Q * EPSSI * card["NSUB"] * 1e6) / cox
    vbx = phi - Q * card["NCH"] * 1e6 * card["XT"] ** 2 / (2 * EPSSI)
    vbm = card["VBM"]
    sp = math.sqrt(phi)
    k2 = (g1 - g2) * (math.sqrt(phi - vbx) - sp) / (2 * sp * (math.sqrt(phi - vbm) - sp) + vbm)
    k1 = g2 - 2 * k2 * math.sqrt(phi - vbm)
    assert abs(p[7] - k1) <= 3e-3 * abs(k1), (p[7], k1)
    assert abs(p[8] - k2) <= 3e-3 * abs(k2), (p[8], k2)
    ref.close()


def test_bsim3_temperature_laws(oracle_lib):
    """At TEMP != TNOM: mu0(T) = U0 (T/Tnom)^UTE; vsat(T) = VSAT - AT (T/Tnom
    - 1); Ua,b,c(T) = Ua,b,c + Ua1,b1,c1 (T/Tnom - 1); Rdsw(T) = RDSW + PRT
    (T/Tnom - 1) and rds0 = Rdsw(T) / (Weff[um])^WR."""
    card = dict(B3_BASE, TNOM=27.0, UTE=-1.5, AT=3.3e4, UA=1e-9, UB=1e-19, UC=-2e-11,
                UA1=4.3e-9, UB1=-7.6e-18, UC1=-5.6e-11, RDSW=200.0, PRT=50.0, WR=0.9)
    T = 273.15 + 85.0
    nl, *_, ref = _b3(card, LONG, dict(d=0.1), temp=T)
    p = ref.b3.sdp(0)
    r = T / T27 - 1.0
    weff = LONG["W"]
    want = dict(U0TEMP=card["U0"] / 1e4 * (T / T27) ** card["UTE"],
                VSATTEMP=card["VSAT"] - card["AT"] * r,
                RDS0=(card["RDSW"] + card["PRT"] * r) / (weff * 1e6) ** card["WR"],
                UA=card["UA"] + card["UA1"] * r, UB=card["UB"] + card["UB1"] * r,
                UC=card["UC"] + card["UC1"] * r)
    got = dict(U0TEMP=p[11], VSATTEMP=p[12], RDS0=p[13], UA=p[27], UB=p[28], UC=p[29])
    for k in want:
        assert abs(got[k] - want[k]) <= 1e-13 * abs(want[k]), (k, got[k], want[k])
    ref.close()
    f, ix = _library_fields(nl, temp=T)
    for k in want:
        assert abs(f[ix[k]] - want[k]) <= 1e-13 * abs(want[k]), ("library", k)


@pytest.mark.parametrize("vbs", [0.0, -0.5, -1.0, -2.0])
def test_bsim3_threshold_body_effect(oracle_lib, vbs):
    """With the short-channel, narrow-width and DIBL terms off, the stored
    von is Vth = VTH0 + K1 (sqrt(phi - Vbs) - sqrt(phi)) - K2 Vbs.  The
    model uses a smoothed Vbseff = Vbs (1 - ~1e-3 / (Vbs - vbsc)), vbsc =
    -30 V here, so the limit holds to 1e-4 |Vbs| volts (a missing K2 term
    is 0.02 |Vbs|)."""
    card = B3_BASE
    _, F, Qv, JF, JQ, sta, sto, ref = _b3(card, LONG, dict(d=0.1, g=0.0, s=0.0, b=vbs))
    phi = manual_phi(card["NCH"])
    vth = card["VTH0"] + card["K1"] * (math.sqrt(phi - vbs) - math.sqrt(phi)) - card["K2"] * vbs
    von = sto[4]
    assert abs(von - vth) <= 1e-4 * abs(vbs) + 1e-12, (von, vth)
    ref.close()


def test_bsim3_threshold_temperature_shift(oracle_lib):
    """Vth(T) - Vth(Tnom) = (KT1 + KT1L / Leff + KT2 Vbs)(T/Tnom - 1) with
    the surface potential itself held at its Tnom value by the model."""
    card = dict(B3_BASE, TNOM=27.0, KT1=-0.11, KT1L=2e-9, KT2=0.022)
    T = 273.15 + 100.0
    volts = dict(d=0.1, g=0.0, s=0.0, b=0.0)
    *_, sto0, r0 = _b3(card, LONG, volts)
    *_, sto1, r1 = _b3(card, LONG, volts, temp=T)
    want = (card["KT1"] + card["KT1L"] / LONG["L"]) * (T / T27 - 1.0)
    assert abs((sto1[4] - sto0[4]) - want) <= 1e-12, (sto1[4] - sto0[4], want)
    r0.close()
    r1.close()


def test_bsim3_subthreshold_swing(oracle_lib):
    """Below threshold Ids ~ exp(Vgs / (n vt)) with n = 1 + (NFACTOR eps_si /
    Xdep + CIT) / Cox (CDSC terms off): the central-difference slope of
    ln Ids at Vgs = Vth - 0.8 V matches 1 / (n vt) to 1e-4 (the Vgsteff
    interpolation leaves a relative error below 1e-5 there)."""
    card = dict(B3_BASE, NFACTOR=1.2, CIT=2e-4)
    phi = manual_phi(card["NCH"])
    xdep = math.sqrt(2 * EPSSI * phi / (Q * card["NCH"] * 1e6))
    n = 1 + (card["NFACTOR"] * EPSSI / xdep + card["CIT"]) / cox_of(card["TOX"])
    vg0 = card["VTH0"] - 0.8
    h = 1e-3
    ids = []
    for vg in (vg0 - h, vg0 + h):
        _, F, *_, ref = _b3(card, LONG, dict(d=1.0, g=vg, s=0.0, b=0.0))
        ids.append(F[0])
        ref.close()
    assert min(ids) > 0
    slope = (math.log(ids[1]) - math.log(ids[0])) / (2 * h)
    want = 1.0 / (n * vt_of(T27))
    assert abs(slope - want) <= 1e-4 * want, (slope, want, n)


def test_bsim3_linear_region_conductance(oracle_lib):
    """As Vds -> 0 with RDSW = 0 and no mobility degradation, Ids / Vds ->
    mu0 Cox (Weff / Leff) Vgsteff, and Vgsteff -> Vgs - Vth in strong
    inversion (1.2 V above threshold the smoothing is below 1e-7)."""
    card = B3_BASE
    vgst = 1.2
    vds = 1e-6
    _, F, *_, ref = _b3(card, LONG, dict(d=vds, g=card["VTH0"] + vgst, s=0.0, b=0.0))
    d = 0
    g = card["U0"] / 1e4 * cox_of(card["TOX"]) * LONG["W"] / LONG["L"] * vgst
    assert abs(F[d] / vds - g) <= 1e-5 * g, (F[d] / vds, g)
    ref.close()


SAT_CARD = dict(B3_BASE, VSAT=1e10, XJ=1e-8)
SAT_INST = dict(W=1e-3, L=1e-3, **JCT)


def _abulk(card, phi):
    """Bulk-charge factor of a long channel: 1 + A0 K1 / (2 sqrt(phi))."""
    return 1 + card["A0"] * card["K1"] / (2 * math.sqrt(phi))


def test_bsim3_long_channel_saturation(oracle_lib):
    """Long channel, no velocity saturation (Esat L -> inf) and no output
    conductance terms: Vdsat = (Vgsteff + 2 vt) / Abulk and
    Ids = mu0 Cox W/(2L) Vgsteff (Vgsteff + 2 vt) / Abulk.  The channel of
    1 mm leaves Leff / (Leff + 2 sqrt(XJ Xdep)) within 6e-5 of 1, so Abulk
    and Ids hold to 3e-5."""
    card = SAT_CARD
    phi = manual_phi(card["NCH"])
    vgst = 1.2
    _, F, *_, ref = _b3(card, SAT_INST, dict(d=2.5, g=card["VTH0"] + vgst, s=0.0, b=0.0))
    vt = vt_of(T27)
    ab = _abulk(card, phi)
    ids = (card["U0"] / 1e4 * cox_of(card["TOX"]) * SAT_INST["W"] / (2 * SAT_INST["L"])
           * vgst * (vgst + 2 * vt) / ab)
    assert abs(F[0] - ids) <= 3e-5 * ids, (F[0], ids)
    ref.close()


def test_bsim3_substrate_current(oracle_lib):
    """Isub / Ids = (ALPHA0 / Leff)(Vds - Vdsat) exp(-BETA0 / (Vds - Vdsat))
    with Vdsat of the long-channel limit above; the substrate current leaves
    through the bulk terminal (F at bulk = -Isub, F at drain = Ids +
    Isub)."""
    card = dict(SAT_CARD, ALPHA0=1e-5, BETA0=20.0)
    phi = manual_phi(card["NCH"])
    vgst, vds = 1.2, 3.0
    _, F, *_, ref = _b3(card, SAT_INST, dict(d=vds, g=card["VTH0"] + vgst, s=0.0, b=0.0))
    b = 3
    isub = -F[b]
    ids = F[0] + F[b]
    vdsat = (vgst + 2 * vt_of(T27)) / _abulk(card, phi)
    dv = vds - vdsat
    want = card["ALPHA0"] / SAT_INST["L"] * dv * math.exp(-card["BETA0"] / dv)
    assert isub > 0
    assert abs(isub / ids - want) <= 1e-3 * want, (isub / ids, want)
    ref.close()


# ------------------------------------------------------------------ MOSFET1
M1_TEMP_CARD = dict(VTO=0.7, KP=5e-5, GAMMA=0.45, PHI=0.65, IS=2e-15, PB=0.75, TNOM=27.0,
                    TOX=2e-8)


@pytest.mark.parametrize("temp_c", [-20.0, 27.0, 85.0, 150.0])
@pytest.mark.parametrize("dtype", [1, -1])
def test_mos1_temperature_closed_forms(oracle_lib, temp_c, dtype):
    """updateTemperature against the physics it encodes: phi(T) and PB(T)
    from n_i(T); Vto(T) - Vto = type (gamma (sqrt(phi(T)) - sqrt(phi)) +
    (phi(T) - phi) / 2) - (Eg(T) - Eg(Tnom)) / 2 (flat band follows half the
    gap and half the Fermi-potential change); Is(T) = Is exp(Eg(Tnom) /
    vt(Tnom) - Eg(T) / vt(T)); KP(T) = KP (T / Tnom)^-1.5."""
    from oracle.oracle import lib
    card = dict(M1_TEMP_CARD)
    if dtype < 0:
        card["VTO"] = -card["VTO"]
    T = 273.15 + temp_c
    nl = single("mos1", dtype, card, dict(W=10e-6, L=2e-6))
    _, _, _, _, _, _, ref = oracle_load(nl, dict(d=0.0), DC, temp=T)
    out = np.zeros(14)
    lib().ref_mos1_probe.argtypes = [__import__("ctypes").c_void_p, __import__("ctypes").c_int,
                                     np.ctypeslib.ndpointer(np.float64)]
    lib().ref_mos1_probe(ref.m1, 0, out)
    vt, tphi, tvto, _, tsat, tkp, _, tpb = out[:8]
    tn = T27
    phi_t = phi_at(card["PHI"], tn, T)
    pb_t = phi_at(card["PB"], tn, T)
    dvto = (dtype * (card["GAMMA"] * (math.sqrt(phi_t) - math.sqrt(card["PHI"]))
                     + 0.5 * (phi_t - card["PHI"])) - 0.5 * (eg_of(T) - eg_of(tn)))
    isat = card["IS"] * math.exp(eg_of(tn) / vt_of(tn) - eg_of(T) / vt_of(T))
    assert abs(vt - vt_of(T)) <= 1e-15
    assert abs(tphi - phi_t) <= 1e-12, (tphi, phi_t)
    assert abs(tpb - pb_t) <= 1e-12, (tpb, pb_t)
    assert abs(tvto - (card["VTO"] + dvto)) <= 1e-12, (tvto, card["VTO"] + dvto)
    assert abs(tsat - isat) <= 1e-11 * isat, (tsat, isat)
    assert abs(tkp - card["KP"] * (T / tn) ** -1.5) <= 1e-13 * card["KP"]
    ref.close()


MEYER_CARD = dict(VTO=0.7, KP=5e-5, GAMMA=0.0, PHI=0.6, LAMBDA=0.0, TOX=2e-8, IS=0.0)
MEYER_INST = dict(W=10e-6, L=2e-6)


@pytest.mark.parametrize("region,volts,cgs,cgd,cgb", [
    # saturation: Cgs = 2/3 Cox W L, Cgd = Cgb = 0
    ("saturation", dict(d=3.0, g=1.7, s=0.0, b=0.0), 2 / 3, 0.0, 0.0),
    # linear at Vds = 0: Cgs = Cgd = 1/2 Cox W L
    ("linear Vds=0", dict(d=0.0, g=2.0, s=0.0, b=0.0), 1 / 2, 1 / 2, 0.0),
    # strong accumulation (Vgs - Vto <= -phi): Cgb = Cox W L
    ("accumulation", dict(d=0.5, g=-0.5, s=0.0, b=0.0), 0.0, 0.0, 1.0),
    # depletion, -phi < Vgs - Vto <= -phi/2: Cgb = -(Vgs - Vto)/phi Cox W L
    ("depletion", dict(d=0.5, g=0.3, s=0.0, b=0.0), 0.0, 0.0, 0.4 / 0.6),
])
def test_mos1_meyer_capacitances(oracle_lib, region, volts, cgs, cgd, cgb):
    """Meyer's gate capacitances of a level-1 device: a DC load stores half
    of each in the state (the caps of two time points are averaged); the DC
    gate charge is Q = C v per branch; a transient load at the same
    voltages with that state as the previous time point puts the whole
    capacitances in the gate row of dQ/dx: dQg/dVs = -Cgs, dQg/dVd = -Cgd,
    dQg/dVb = -Cgb (no overlap capacitances given)."""
    from oracle.oracle import RefLoad
    nl = single("mos1", 1, MEYER_CARD, MEYER_INST)
    _, Qv, _, _, sta, sto, ref = oracle_load(nl, volts, dict(dcop=1, newtonIter=1))
    ref.close()
    cox = EPSOX_M1 / MEYER_CARD["TOX"] * MEYER_INST["W"] * MEYER_INST["L"]
    want = np.array([cgs, cgd, cgb]) * cox
    got_state = sta[3:6] * 2.0
    assert np.allclose(got_state, want, rtol=1e-13, atol=1e-13 * cox), (region, got_state, want)
    d, g, s, b = (nl.node(k) for k in "dgsb")
    vg = volts["g"]
    qg = want[0] * (vg - volts["s"]) + want[1] * (vg - volts["d"]) + want[2] * (vg - volts["b"])
    assert abs(Qv[g] - qg) <= 1e-13 * cox * 3, (Qv[g], qg)
    # transient at the same point, previous time point = the DC load
    x = np.zeros(nl.n_nodes + 1)
    for k, v in volts.items():
        x[nl.node(k)] = v
    ref = RefLoad(nl, gmin=0.0, limiter=False)
    ref.set_solver_state(transient=1, newtonIter=1)
    out = ref.full_load(x, sta.copy(), sto.copy(), np.zeros(nl.n_state), sto.copy())
    JQ = np.zeros((nl.n_nodes + 1,) * 2)
    for r in range(nl.n_nodes + 1):
        for k in range(ref.rowptr[r], ref.rowptr[r + 1]):
            JQ[r, ref.colidx[k]] += out[5][k]
    ref.close()
    got_jq = -np.array([JQ[g, s], JQ[g, d], JQ[g, b]])
    assert np.allclose(got_jq, want, rtol=1e-13, atol=1e-13 * cox), (region, got_jq, want)


def _tran_fd_q(nl, x0, h=1e-6):
    """dQ/dx of a transient load against central differences of Q, at a
    point where the previous time step's voltages equal the current ones
    (so the back-averaged Meyer charge q = q_old + C_avg (v - v_old) has
    dq/dv = C_avg exactly there)."""
    from oracle.oracle import RefLoad
    ref = RefLoad(nl, gmin=1e-12, limiter=False)
    n1 = nl.n_nodes + 1

    def run(x, sto_c, sta_c):
        ref.set_solver_state(transient=1, newtonIter=1)
        sto_n = sto_c.copy()
        return ref.full_load(x, sta_c.copy(), sto_c.copy(), np.zeros(nl.n_state), sto_n)

    # previous time point: a DC load at x0 gives its store and state
    ref.set_solver_state(dcop=1, newtonIter=1)
    sto0 = np.zeros(nl.n_store)
    sta0 = np.zeros(nl.n_state)
    ref.full_load(x0, np.zeros(nl.n_state), np.zeros(nl.n_store), sta0, sto0)
    base = run(x0, sto0, sta0)
    JQ = np.zeros((n1, n1))
    for r in range(n1):
        for k in range(ref.rowptr[r], ref.rowptr[r + 1]):
            JQ[r, ref.colidx[k]] += base[5][k]
    n = nl.n_nodes
    scale = max(np.abs(JQ[:n, :n]).max(), 1e-300)
    bad = []
    for c in range(n):
        xp, xm = x0.copy(), x0.copy()
        xp[c] += h
        xm[c] -= h
        qp = run(xp, sto0, sta0)[1]
        qm = run(xm, sto0, sta0)[1]
        num = (qp[:n] - qm[:n]) / (2 * h)
        # |analytic - numerical| <= relTol |numerical| + absTol, absTol on
        # the block's scale (N_DEV_NumericalJacobian.C:407-477 form)
        err = np.abs(JQ[:n, c] - num) - (1e-6 * np.abs(num) + 1e-9 * scale)
        bad += [(r, c, JQ[r, c], num[r]) for r in np.flatnonzero(err > 0)]
    ref.close()
    return bad, scale


@pytest.mark.parametrize("volts", [dict(d=1.5, g=1.2, s=0.1, b=-0.5),
                                   dict(d=0.2, g=2.0, s=0.0, b=-1.0),
                                   dict(d=1.0, g=0.3, s=0.4, b=-0.3),
                                   dict(d=-0.4, g=1.1, s=0.6, b=-1.2)])
def test_mos1_transient_charge_jacobian(oracle_lib, volts):
    """MOSFET1 in transient: dQ/dx (Meyer gate charges, overlap charges,
    junction depletion charges) against central differences of Q, in the
    reference's numerical-Jacobian error form, at 1e-6 relative."""
    card = dict(VTO=0.6, KP=6e-5, GAMMA=0.4, PHI=0.7, LAMBDA=0.02, TOX=2e-8,
                CGSO=3e-10, CGDO=3e-10, CGBO=2e-10, CBD=2e-15, CBS=2.5e-15, PB=0.8,
                MJ=0.45, FC=0.5, CJSW=1e-10, MJSW=0.33)
    nl = single("mos1", 1, card, dict(W=4e-6, L=1e-6, PD=6e-6, PS=6e-6))
    x0 = np.zeros(nl.n_nodes + 1)
    for k, v in volts.items():
        x0[nl.node(k)] = v
    bad, scale = _tran_fd_q(nl, x0)
    assert not bad and scale > 1e-16, bad


# ------------------------------------------------------------------ limiters
# (vnew, vold, vto) -> limited vnew, evaluated by hand from the SPICE3
# definition of fetlim: vtsthi = |2 (vold - vto)| + 2, vtstlo = vtsthi / 2 +
# 2, vtox = vto + 3.5, delv = vnew - vold.  Two of its branches cannot fire
# (an "off" rise that stays below vto + 0.5 never exceeds vtstlo, and an
# "on" fall that stays above vtox never exceeds vtstlo); their rows check
# that the value passes unchanged.
FETLIM_CASES = [
    # off (vold < vto), rising past vto + 0.5: clipped to vto + 0.5
    (5.0, 0.0, 0.5, 1.0),
    # off, rising but staying below vto + 0.5: unchanged
    (0.9, -10.0, 0.5, 0.9),
    # off, falling by more than vtsthi = 3: vold - vtsthi
    (-10.0, 0.0, 0.5, -3.0),
    # off, falling by less: unchanged
    (-2.0, 0.0, 0.5, -2.0),
    # off, falling: vold = -1, vtsthi = 5 -> -6
    (-20.0, -1.0, 0.5, -6.0),
    # middle (vto <= vold < vto + 3.5), rising: at most vto + 4
    (9.0, 1.0, 0.5, 4.5),
    (2.0, 1.0, 0.5, 2.0),
    # middle, falling: at least vto - 0.5
    (-3.0, 1.0, 0.5, 0.0),
    (0.2, 1.0, 0.5, 0.2),
    # middle: vold = 8, vto = 5 (vtox = 8.5), falling to 0 -> vto - 0.5
    (0.0, 8.0, 5.0, 4.5),
    # on (vold >= vtox), rising by vtsthi or more: vold = 4.5, vto = 0.5,
    # vtsthi = 10 -> 14.5
    (20.0, 4.5, 0.5, 14.5),
    (10.0, 4.5, 0.5, 10.0),
    # on: vold = 5, vtsthi = 11 -> 16
    (30.0, 5.0, 0.5, 16.0),
    # on, falling below vtox: at least vto + 2
    (0.0, 4.5, 0.5, 2.5),
    (3.0, 4.5, 0.5, 3.0),
    (-5.0, 9.0, 1.0, 3.0),
    (0.0, 30.0, 10.0, 12.0),
    (35.0, 60.0, 50.0, 52.0),
    # on, falling but staying at or above vtox: unchanged
    (4.1, 12.0, 0.5, 4.1),
    (4.6, 40.0, 0.5, 4.6),
    (14.0, 30.0, 10.0, 14.0),
    (54.0, 80.0, 50.0, 54.0),
]


def test_fetlim_table(oracle_lib):
    from oracle.oracle import lib
    import ctypes
    L = lib()
    L.ref_x_fetlim.restype = ctypes.c_double
    L.ref_x_fetlim.argtypes = [ctypes.c_double] * 3
    for vnew, vold, vto, want in FETLIM_CASES:
        got = L.ref_x_fetlim(vnew, vold, vto)
        assert got == want, (vnew, vold, vto, got, want)


# (vnew, vold) -> limited vds, from the SPICE3 limvds definition: for vold
# >= 3.5 a rise is capped at 3 vold + 2 and a fall below 3.5 stops at 2;
# otherwise a rise is capped at 4 and a fall at -0.5.
LIMVDS_CASES = [
    (20.0, 4.0, 14.0), (10.0, 4.0, 10.0), (1.0, 4.0, 2.0), (3.0, 4.0, 3.0),
    (3.8, 5.0, 3.8), (9.0, 1.0, 4.0), (2.0, 1.0, 2.0), (-3.0, 1.0, -0.5),
    (-0.2, 1.0, -0.2), (5.0, 3.49, 4.0), (-1.0, -0.2, -0.5), (0.0, -3.0, 0.0),
]


def test_limvds_table(oracle_lib):
    from oracle.oracle import lib
    import ctypes
    L = lib()
    L.ref_x_limvds.restype = ctypes.c_double
    L.ref_x_limvds.argtypes = [ctypes.c_double] * 2
    for vnew, vold, want in LIMVDS_CASES:
        assert L.ref_x_limvds(vnew, vold) == want, (vnew, vold, want)


def test_pnjlim_closed_forms(oracle_lib):
    """pnjlim: a junction step larger than 2 vt above vcrit is replaced by
    vold + vt ln(1 + dv / vt) (vold > 0) or vt ln(vnew / vt) (vold <= 0);
    smaller steps or voltages below vcrit pass unchanged."""
    from oracle.oracle import lib
    import ctypes
    L = lib()
    L.ref_x_pnjlim.restype = ctypes.c_double
    L.ref_x_pnjlim.argtypes = [ctypes.c_double] * 4 + [ctypes.POINTER(ctypes.c_int)]
    vt, vcrit = 0.025865, 0.6
    chk = ctypes.c_int()
    for vnew, vold, want, lim in [
            (1.5, 0.7, 0.7 + vt * math.log(1 + 0.8 / vt), 1),
            (1.5, 0.0, vt * math.log(1.5 / vt), 1),
            (1.5, -0.3, vt * math.log(1.5 / vt), 1),
            (0.72, 0.7, 0.72, 0),
            (0.5, -2.0, 0.5, 0)]:
        got = L.ref_x_pnjlim(vnew, vold, vt, vcrit, ctypes.byref(chk))
        assert abs(got - want) <= 1e-15 and chk.value == lim, (vnew, vold, got, want)


# ------------------------------------------------------------------ diode
@pytest.mark.parametrize("temp_c", [-40.0, 27.0, 60.0, 125.0])
def test_diode_temperature_laws(temp_c):
    """SPICE junction-diode temperature dependence (TNOM = 27 C):
    Is(T) = Is (T/Tnom)^(XTI/N) exp((T/Tnom - 1) EG / (N vt(T)));
    VJ(T) from n_i(T) as phi_at; CJO(T) = CJO (1 + M (4e-4 (T - Tnom) -
    (VJ(T) - VJ) / VJ)); and the forward current at T through the load."""
    from oracle.diode_ref import diode_fields
    mp = dict(IS=2e-14, N=1.15, EG=1.11, XTI=3.0, CJO=3e-15, VJ=0.75, M=0.4, TNOM=27.0)
    T = 273.15 + temp_c
    f = diode_fields(mp, dict(TEMP=temp_c), T27, T27)
    vt = vt_of(T)
    isat = mp["IS"] * (T / T27) ** (mp["XTI"] / mp["N"]) * math.exp(
        (T / T27 - 1.0) * mp["EG"] / (mp["N"] * vt))
    vj = phi_at(mp["VJ"], T27, T)
    cj = mp["CJO"] * (1 + mp["M"] * (4e-4 * (T - T27) - (vj - mp["VJ"]) / mp["VJ"]))
    assert abs(f["Isat"] - isat) <= 1e-12 * isat, (f["Isat"], isat)
    assert abs(f["tJctPot"] - vj) <= 1e-12, (f["tJctPot"], vj)
    # the reference writes Eg(300.15 K) as the literal 1.1150877 where
    # Varshni's form gives 1.11508782: the difference cancels in VJ(T) but
    # leaves ~1e-7 V in the grading term's reference potential, i.e. a few
    # 1e-9 relative in CJO(T)
    assert abs(f["tJctCap"] - cj) <= 2e-8 * cj, (f["tJctCap"], cj)
    assert abs(f["Vte"] - mp["N"] * vt) <= 1e-16

    # the load at T: F = Is(T) (exp(V / (N vt)) - 1) + gmin V
    nl = Netlist("diode at T")
    nl.add_model("D", "diode", 1, **mp)
    nl.add_vsrc("V1", "a", "0", 0.0)
    nl.add_diode("D1", "a", "0", "D", TEMP=temp_c)
    nl.layout()
    x = np.zeros(nl.n_nodes + 1)
    v = 0.45
    x[nl.node("a")] = v
    F, *_ = oracle_load(nl, dict(a=v), DC, gmin=1e-12)[:1]
    want = isat * math.expm1(v / (mp["N"] * vt)) + 1e-12 * v
    # the source's branch current also enters node a
    got = F[nl.node("a")] - x[nl.vsrc_branch[0]]
    assert abs(got - want) <= 1e-12 * abs(want), (got, want)
