"""Circuit descriptions and the seeded benchmark netlists.

A ``Netlist`` holds what a SPICE deck would: named nodes (``"0"`` is ground),
MOSFET level-1 and BSIM3 models and instances, and the linear devices
(resistors, capacitors, independent voltage sources).  ``layout()`` assigns
solution ids (LIDs), internal drain'/source' nodes, voltage-source branch
currents, and the per-instance state/store slots, the way Xyce's topology
hands them to devices (node ids in order of first appearance here; ground is
GID -1, stored in the trailing ground slot).

Generators (sizes and seeds of the benchmark configurations):
  ring_oscillator_mos1   101-stage CMOS ring, MOSFET level 1
  ring_oscillator_bsim3  ring of BSIM3 inverters (Monte-Carlo samples)
  random_logic_bsim3     seeded NAND/NOR gate netlist with RC interconnect
"""
from dataclasses import dataclass, field
import numpy as np

GROUND_NAMES = ("0", "gnd", "GND")


@dataclass
class Model:
    name: str
    kind: str          # "mos1" | "bsim3"
    dtype: int         # +1 NMOS, -1 PMOS
    params: dict


@dataclass
class Mosfet:
    name: str
    d: str
    g: str
    s: str
    b: str
    model: str
    params: dict = field(default_factory=dict)


@dataclass
class Pulse:
    """PULSE(V1 V2 TD TR TF PW PER); None = the reference's default."""
    v1: float
    v2: float
    td: float = 0.0
    tr: float = None
    tf: float = None
    pw: float = None
    per: float = None


@dataclass
class Vsource:
    name: str
    p: str
    n: str
    dc: float
    pulse: Pulse = None


class Netlist:
    def __init__(self, title=""):
        self.title = title
        self.node_ids = {}
        self.models = {}
        self.mos = []          # Mosfet, both kinds, in netlist order
        self.caps = []         # (name, a, b, C)
        self.res = []          # (name, a, b, R)
        self.vsrc = []         # Vsource
        self.diodes = []       # (name, a, b, model, instance params)
        self.ic = {}           # node -> initial voltage (.IC)
        self._laid_out = False

    # ---- building
    def node(self, name):
        if name in GROUND_NAMES:
            return -1
        if name not in self.node_ids:
            self.node_ids[name] = len(self.node_ids)
        return self.node_ids[name]

    def probe_id(self, name):
        """Solution id of an existing node, for reading results: never
        creates a node.  KeyError for a name the netlist does not hold;
        ValueError for ground, which has no unknown of its own."""
        if name in GROUND_NAMES:
            raise ValueError("ground (%r) is not a probe: it has no unknown" % name)
        if name not in self.node_ids:
            raise KeyError("no node named %r in %r" % (name, self.title))
        i = self.node_ids[name]
        n = getattr(self, "n_nodes", len(self.node_ids))
        if not 0 <= i < n:
            raise KeyError("node %r has id %d outside 0..%d" % (name, i, n - 1))
        return i

    def add_model(self, name, kind, dtype, **params):
        self.models[name] = Model(name, kind, dtype, dict(params))

    def add_mos(self, name, d, g, s, b, model, **params):
        for n in (d, g, s, b):
            self.node(n)
        self.mos.append(Mosfet(name, d, g, s, b, model, dict(params)))

    def add_cap(self, name, a, b, C):
        self.node(a), self.node(b)
        self.caps.append((name, a, b, float(C)))

    def add_res(self, name, a, b, R):
        self.node(a), self.node(b)
        self.res.append((name, a, b, float(R)))

    def add_diode(self, name, a, b, model, **params):
        self.node(a), self.node(b)
        self.diodes.append((name, a, b, model, dict(params)))

    def add_vsrc(self, name, p, n, dc, pulse=None):
        self.node(p), self.node(n)
        self.vsrc.append(Vsource(name, p, n, float(dc), pulse))

    # ---- layout
    def _has_series(self, mdl, inst, which):
        """drainConductance / sourceConductance != 0 (N_DEV_MOSFET1.C:1184-1225)."""
        p = mdl.params
        r = p.get("RD" if which == "d" else "RS", 0.0)
        if r != 0:
            return True
        if "RSH" in p and p["RSH"] != 0:
            return True
        return False

    def layout(self):
        """Assign ids.  Returns self; fills n_nodes, n_state, n_store and the
        per-device tables mos1/bsim3 (lists of dicts)."""
        if self._laid_out:
            return self
        n = len(self.node_ids)
        self.mos1, self.bsim3 = [], []
        sta = sto = 0
        for m in self.mos:
            mdl = self.models[m.model]
            lids = [self.node(m.d), self.node(m.g), self.node(m.s), self.node(m.b), -1, -1]
            if mdl.kind == "mos1":
                if self._has_series(mdl, m, "d"):
                    lids[4] = n
                    n += 1
                if self._has_series(mdl, m, "s"):
                    lids[5] = n
                    n += 1
                rec = dict(name=m.name, model=m.model, params=m.params, lids=lids,
                           sta=sta, sto=sto)
                sta += 8
                sto += 5
                self.mos1.append(rec)
            elif mdl.kind == "bsim3":
                p = mdl.params
                rsh = p.get("RSH", 0.0)
                nrd = m.params.get("NRD", 1.0)
                nrs = m.params.get("NRS", 1.0)
                if rsh * nrd != 0:
                    lids[4] = n
                    n += 1
                if rsh * nrs != 0:
                    lids[5] = n
                    n += 1
                rec = dict(name=m.name, model=m.model, params=m.params, lids=lids,
                           sta=sta, sto=sto)
                sta += 12
                sto += 5
                self.bsim3.append(rec)
            else:
                raise ValueError("unknown model kind %r" % mdl.kind)
        # one store slot per diode (li_storevd)
        self.diode_sto = []
        for _ in self.diodes:
            self.diode_sto.append(sto)
            sto += 1
        self.vsrc_branch = []
        for v in self.vsrc:
            self.vsrc_branch.append(n)
            n += 1
        # one charge state slot per capacitor (N_DEV_Capacitor.C li_QState)
        self.cap_sta = []
        for _ in self.caps:
            self.cap_sta.append(sta)
            sta += 1
        self.n_nodes, self.n_state, self.n_store = n, sta, sto
        self._laid_out = True
        return self

    def pulse_params(self, v, tstep=1e-12, tstop=1e-9):
        """The seven PULSE values with the reference's defaults for omitted
        ones (N_DEV_SourceData.C:902-916: TR, TF = starting step; PW, PER =
        final time), or None for a DC source."""
        if v.pulse is None:
            return None
        p = v.pulse
        tr = p.tr if p.tr is not None else tstep
        tf = p.tf if p.tf is not None else tstep
        pw = p.pw if p.pw is not None else tstop
        per = p.per if p.per is not None else tstop
        return [p.v1, p.v2, p.td or 0.0, tr, tf, pw, per]

    def voltage_unknowns(self):
        """Solution ids of the voltage unknowns (nodes and internal drain' /
        source' nodes): every unknown but the voltage sources' branch
        currents."""
        self.layout()
        br = set(self.vsrc_branch)
        return [i for i in range(self.n_nodes) if i not in br]

    def breakpoints(self, tstop):
        """Sorted corner times in (0, tstop) of every PULSE source (TD, TD+TR,
        TD+TR+PW, TD+TR+PW+TF and TD+PER of each period, PulseData::
        getBreakPoints, N_DEV_SourceData.C:1114-1160), with the source values
        the load uses (pulse_params with its defaults)."""
        out = set()
        for v in self.vsrc:
            pp = self.pulse_params(v)
            if pp is None:
                continue
            _, _, td, tr, tf, pw, per = pp
            base = 0.0
            while True:
                for c in (td, td + tr, td + tr + pw, td + tr + pw + tf, td + per):
                    if 0.0 < base + c < tstop:
                        out.add(base + c)
                if per <= 0.0:
                    break
                base += per
                if base + td >= tstop:
                    break
        return sorted(out)

    def linear_devices(self):
        """(kind, pos, neg, branch, value, pulse, sta) in load order:
        resistors, capacitors, diodes, sources, each in netlist order.  For
        a diode, value = (model parameters, instance parameters) and sta is
        its store slot."""
        out = []
        for name, a, b, r in self.res:
            out.append(("R", self.node(a), self.node(b), -1, r, None, 0))
        for (name, a, b, c), st in zip(self.caps, self.cap_sta):
            out.append(("C", self.node(a), self.node(b), -1, c, None, st))
        for (name, a, b, model, params), st in zip(self.diodes, self.diode_sto):
            out.append(("D", self.node(a), self.node(b), -1,
                        (self.models[model].params, params), None, st))
        for v, br in zip(self.vsrc, self.vsrc_branch):
            out.append(("V", self.node(v.p), self.node(v.n), br, v.dc, self.pulse_params(v), 0))
        return out

    def foreign_stamps(self):
        """(gids, rows, cols) of every device the load library does not
        evaluate: capacitor and resistor 2x2 blocks
        (N_DEV_Capacitor.C / N_DEV_Resistor.C jacStamp), voltage source
        [[0 0 1],[0 0 1],[1 1 0]]-shaped stamp (N_DEV_Vsrc.C)."""
        out = []
        for _, a, b, _C in self.caps + self.res:
            out.append(([self.node(a), self.node(b)], [0, 0, 1, 1], [0, 1, 0, 1]))
        for _, a, b, _m, _p in self.diodes:
            out.append(([self.node(a), self.node(b)], [0, 0, 1, 1], [0, 1, 0, 1]))
        for v, br in zip(self.vsrc, self.vsrc_branch):
            out.append(([self.node(v.p), self.node(v.n), br], [0, 1, 2, 2], [2, 2, 0, 1]))
        return out


# ---------------------------------------------------------------------------
# Benchmark netlists

MOS1_NMOS = dict(VTO=0.45, KP=120e-6)
MOS1_PMOS = dict(VTO=-0.45, KP=40e-6)


def supply(vdd, ramp):
    """Vdd source: DC, or (ramp > 0) a PULSE from 0 to vdd in `ramp` seconds
    that stays on (the start kick of a transient: the DC operating point is
    then all-zero and the unequal stage loads break the ring's symmetry)."""
    if not ramp:
        return dict(dc=vdd)
    return dict(dc=0.0, pulse=Pulse(0.0, vdd, 0.0, ramp, ramp, 1.0, 2.0))


def ring_oscillator_mos1(stages=101, seed=1, vdd=1.8, ramp=0.0):
    """101-stage CMOS ring oscillator of MOSFET level-1 devices.

    NMOS W=1u L=0.18u VTO=0.45 KP=120u; PMOS W=2u L=0.18u VTO=-0.45 KP=40u;
    load capacitor per stage output drawn from U[5, 15] fF (seeded); Vdd 1.8 V
    (ramped from 0 over `ramp` seconds when given).
    """
    rng = np.random.default_rng(seed)
    nl = Netlist("ring oscillator, %d stages, MOSFET level 1" % stages)
    nl.add_model("NMOS1", "mos1", 1, **MOS1_NMOS)
    nl.add_model("PMOS1", "mos1", -1, **MOS1_PMOS)
    sv = supply(vdd, ramp)
    nl.add_vsrc("VDD", "vdd", "0", sv["dc"], sv.get("pulse"))
    for i in range(stages):
        inp = "out%d" % ((i - 1) % stages)
        out = "out%d" % i
        nl.add_mos("MN%d" % i, out, inp, "0", "0", "NMOS1", W=1e-6, L=0.18e-6)
        nl.add_mos("MP%d" % i, out, inp, "vdd", "vdd", "PMOS1", W=2e-6, L=0.18e-6)
    caps = rng.uniform(5e-15, 15e-15, stages)
    for i in range(stages):
        nl.add_cap("C%d" % i, "out%d" % i, "0", caps[i])
    return nl


def mos1_variety(seed=0, n_nodes=24, n_dev=96):
    """Parity workload covering the MOSFET1 options a netlist can set:
    series RD/RS and RSH (internal d'/s' nodes), TOX (Meyer capacitances),
    CBD/CBS and CJ/CJSW junctions with MJ != MJSW, JS with areas, GAMMA,
    LAMBDA, NSUB-derived parameters, LD, OFF, IC=, M=, model TNOM and
    instance TEMP (both in Celsius), both polarities and devices tied to
    ground on any terminal."""
    rng = np.random.default_rng(seed)
    nl = Netlist("MOSFET1 variety")
    base = dict(TOX=1.5e-8, LAMBDA=0.05, GAMMA=0.4, PHI=0.65, CGSO=2e-10,
                CGDO=2e-10, CGBO=1e-10, CJ=2e-4, CJSW=1e-9, MJ=0.45, MJSW=0.33,
                JS=1e-4, LD=1e-8)
    nl.add_model("N_A", "mos1", 1, VTO=0.5, KP=110e-6, **base)
    nl.add_model("P_A", "mos1", -1, VTO=-0.55, KP=45e-6, **base)
    nl.add_model("N_R", "mos1", 1, VTO=0.4, KP=100e-6, RD=15.0, RS=12.0, CBD=2e-15,
                 CBS=3e-15, LAMBDA=0.02)
    nl.add_model("P_SH", "mos1", -1, VTO=-0.4, KP=40e-6, RSH=20.0, TOX=2e-8,
                 NSUB=1e16, MJ=0.5, MJSW=0.5, CJ=1e-4, CJSW=5e-10)
    nl.add_model("N_PLAIN", "mos1", 1, VTO=0.45, KP=120e-6)
    nl.add_model("N_HOT", "mos1", 1, VTO=0.45, KP=100e-6, TNOM=40.0, NSUB=2e16,
                 CJ=1.5e-4, CJSW=8e-10, JS=2e-4, TOX=1.8e-8)
    names = list(nl.models)
    nodes = ["n%d" % k for k in range(n_nodes)] + ["0"]
    for i in range(n_dev):
        d, g, s, b = (nodes[k] for k in rng.integers(0, len(nodes), 4))
        mdl = names[i % len(names)]
        p = dict(W=float(rng.uniform(0.5e-6, 4e-6)), L=float(rng.uniform(0.15e-6, 1e-6)))
        if rng.random() < 0.5:
            p.update(AD=float(rng.uniform(1e-13, 5e-12)), AS=float(rng.uniform(1e-13, 5e-12)),
                     PD=float(rng.uniform(1e-6, 8e-6)), PS=float(rng.uniform(1e-6, 8e-6)))
        if rng.random() < 0.3:
            p.update(NRD=float(rng.uniform(0.5, 3)), NRS=float(rng.uniform(0.5, 3)))
        if rng.random() < 0.2:
            p["M"] = float(rng.integers(2, 4))
        if rng.random() < 0.08:
            p["OFF"] = 1.0
        if rng.random() < 0.15:
            p["TEMP"] = float(rng.choice([-20.0, 60.0, 110.0]))
        if rng.random() < 0.08:
            p.update(IC1=float(rng.uniform(0, 1.8)), IC2=float(rng.uniform(0, 1.8)),
                     IC3=float(rng.uniform(-0.5, 0)))
        nl.add_mos("M%d" % i, d, g, s, b, mdl, **p)
    return nl


# 130 nm-class BSIM3v3 cards (level 9): every parameter not listed keeps the
# BSIM3 default of the model's parameter table.
BSIM3_NMOS_130 = dict(TOX=3.2e-9, VTH0=0.36, U0=350.0, VSAT=1.1e5, K1=0.52, K2=-0.02,
                      NCH=2.6e17, XJ=1.0e-7, RDSW=180.0, PRWG=0.1, PRWB=-0.1,
                      DVT0=2.5, DVT1=0.6, DVT2=-0.03, ETA0=0.06, ETAB=-0.05,
                      PCLM=1.1, PDIBLC1=0.3, PDIBLC2=0.004, PSCBE1=4.0e8, PSCBE2=1.0e-5,
                      ALPHA0=1.0e-6, BETA0=25.0, DELTA=0.01, A0=1.2, AGS=0.2, KETA=-0.04,
                      CJ=1.0e-3, CJSW=2.2e-10, CJSWG=3.0e-10, PB=0.8, PBSW=0.8, MJ=0.42,
                      MJSW=0.31, CGSO=2.4e-10, CGDO=2.4e-10, CGBO=1.0e-12, CGSL=1.0e-10,
                      CGDL=1.0e-10, CKAPPA=0.6, LINT=1.0e-8, WINT=5.0e-9, JS=1.0e-6,
                      JSW=1.0e-12, NFACTOR=1.2, VOFF=-0.09, XPART=0.5)
BSIM3_PMOS_130 = dict(TOX=3.2e-9, VTH0=-0.34, U0=110.0, VSAT=8.0e4, K1=0.50, K2=-0.01,
                      NCH=2.2e17, XJ=1.0e-7, RDSW=400.0, PRWG=0.05, PRWB=-0.05,
                      DVT0=2.2, DVT1=0.55, DVT2=-0.03, ETA0=0.05, ETAB=-0.04,
                      PCLM=0.9, PDIBLC1=0.25, PDIBLC2=0.003, PSCBE1=4.5e8, PSCBE2=1.0e-5,
                      ALPHA0=5.0e-7, BETA0=28.0, DELTA=0.01, A0=1.1, AGS=0.25, KETA=-0.03,
                      CJ=1.1e-3, CJSW=2.0e-10, CJSWG=2.8e-10, PB=0.85, PBSW=0.85, MJ=0.45,
                      MJSW=0.33, CGSO=2.2e-10, CGDO=2.2e-10, CGBO=1.0e-12, CGSL=1.0e-10,
                      CGDL=1.0e-10, CKAPPA=0.6, LINT=1.0e-8, WINT=5.0e-9, JS=1.0e-6,
                      JSW=1.0e-12, NFACTOR=1.3, VOFF=-0.1, XPART=0.5)


def _diffusion(w, ext=0.3e-6):
    """Drain/source diffusion geometry of a transistor of width w."""
    return dict(AD=w * ext, AS=w * ext, PD=2 * (w + ext), PS=2 * (w + ext))


def bsim3_variety(seed=0, n_nodes=30, n_dev=128, capmod=None):
    """Parity workload covering the BSIM3 options a card can set: mobMod
    1/2/3, xpart 0 / 0.4 / 0.5 / 1 (the three charge partitions), RSH sheet
    resistance (internal d'/s'), IJTH junction linearisation, NGATE
    poly depletion, K1/K2 given vs. derived from NCH/GAMMA1/GAMMA2/VBX,
    VFB given, length/width binning (L*, W*, P* coefficients), DWG/DWB,
    TNOM and instance TEMP away from nominal, diffusion perimeters smaller
    than Weff, OFF, M=, both polarities and grounded terminals.  With
    capmod given, every card uses that charge model (VFBCV set per card
    for capMod 0)."""
    rng = np.random.default_rng(seed)
    nl = Netlist("BSIM3 variety")
    nl.add_model("N1", "bsim3", 1, **BSIM3_NMOS_130)
    nl.add_model("P1", "bsim3", -1, **BSIM3_PMOS_130)
    n2 = dict(BSIM3_NMOS_130, MOBMOD=2, XPART=0.0, IJTH=0.05, RSH=5.0, NGATE=5e19,
              LVTH0=2e-3, WU0=3.0, PVSAT=1e3, DWG=-2e-9, DWB=1e-9, TNOM=35.0)
    nl.add_model("N2", "bsim3", 1, **n2)
    p3 = dict(BSIM3_PMOS_130, MOBMOD=3, XPART=1.0, VFB=-0.6, LK1=0.01, WETA0=1e-3,
              TCJ=1e-3, TPB=1e-3, IJTH=0.2)
    nl.add_model("P3", "bsim3", -1, **p3)
    n4 = {k: v for k, v in BSIM3_NMOS_130.items() if k not in ("K1", "K2", "VTH0")}
    n4.update(GAMMA1=0.6, GAMMA2=0.45, VBX=-1.5, XPART=0.4, CGSO=0.0, CGDO=0.0)
    nl.add_model("N4", "bsim3", 1, **n4)
    if capmod is not None:
        for k, name in enumerate(nl.models):
            nl.models[name].params["CAPMOD"] = float(capmod)
            if capmod == 0:
                nl.models[name].params["VFBCV"] = (-1.0, -0.8, 0.3, -0.95)[k % 4]
    names = list(nl.models)
    nodes = ["n%d" % k for k in range(n_nodes)] + ["0"]
    for i in range(n_dev):
        d, g, s, b = (nodes[k] for k in rng.integers(0, len(nodes), 4))
        mdl = names[i % len(names)]
        w = float(rng.choice([0.4e-6, 0.5e-6, 1.0e-6, float(rng.uniform(0.3e-6, 3e-6))]))
        l = float(rng.choice([0.13e-6, float(rng.uniform(0.12e-6, 0.5e-6))]))
        p = dict(W=w, L=l)
        if rng.random() < 0.7:
            ext = float(rng.uniform(0.05e-6, 0.4e-6))
            p.update(_diffusion(w, ext))
            if rng.random() < 0.3:
                p.update(PD=float(rng.uniform(0.05e-6, w)), PS=float(rng.uniform(0.05e-6, w)))
        if rng.random() < 0.3:
            p.update(NRD=float(rng.uniform(0.5, 3)), NRS=float(rng.uniform(0.5, 3)))
        if rng.random() < 0.2:
            p["M"] = float(rng.integers(2, 4))
        if rng.random() < 0.06:
            p["OFF"] = 1.0
        if rng.random() < 0.1:
            p["TEMP"] = float(rng.choice([50.0, 85.0]))
        nl.add_mos("M%d" % i, d, g, s, b, mdl, **p)
    return nl


def ring_oscillator_bsim3(stages=11, seed=3, vdd=1.2, dvth0=0.0, dtox=0.0, ramp=0.0,
                          rng=None, prefix=""):
    """Ring oscillator of BSIM3 inverters.  With dvth0 / dtox nonzero every
    device gets its own model card with VTH0 += N(0, dvth0) and
    TOX *= U[1 - dtox, 1 + dtox] (a Monte-Carlo sample)."""
    rng = rng if rng is not None else np.random.default_rng(seed)
    nl = Netlist("BSIM3 ring oscillator, %d stages" % stages)
    per_device = dvth0 != 0.0 or dtox != 0.0
    if not per_device:
        nl.add_model("NB", "bsim3", 1, **BSIM3_NMOS_130)
        nl.add_model("PB", "bsim3", -1, **BSIM3_PMOS_130)
    sv = supply(vdd, ramp)
    nl.add_vsrc("VDD", "vdd", "0", sv["dc"], sv.get("pulse"))
    caps = rng.uniform(1e-15, 3e-15, stages)
    for i in range(stages):
        inp = "out%d" % ((i - 1) % stages)
        out = "out%d" % i
        wn, wp = 0.5e-6, 1.0e-6
        mn, mp = "NB", "PB"
        if per_device:
            mn, mp = "N%s%d" % (prefix, i), "P%s%d" % (prefix, i)
            for name, base, dt in ((mn, BSIM3_NMOS_130, 1), (mp, BSIM3_PMOS_130, -1)):
                card = dict(base)
                card["VTH0"] = base["VTH0"] + float(rng.normal(0.0, dvth0))
                card["TOX"] = base["TOX"] * float(rng.uniform(1 - dtox, 1 + dtox))
                nl.add_model(name, "bsim3", dt, **card)
        nl.add_mos("MN%d" % i, out, inp, "0", "0", mn, W=wn, L=0.13e-6, **_diffusion(wn))
        nl.add_mos("MP%d" % i, out, inp, "vdd", "vdd", mp, W=wp, L=0.13e-6, **_diffusion(wp))
        nl.add_cap("C%d" % i, out, "0", caps[i])
    return nl


def random_logic_bsim3(n_gates=5000, seed=2, vdd=1.2, n_inputs=64, ramp=0.0):
    """Seeded random CMOS logic: n_gates 2-input NAND / NOR gates (50/50),
    i.e. 4 * n_gates BSIM3 transistors.  Each gate output fans out to 1-4
    gate inputs; W and L are drawn independently from U[0.8, 1.2] x
    (0.5u, 0.13u); every net is an RC line of 1-3 segments with R from
    U[10, 200] ohm and C from U[1, 20] fF per segment.  Gate inputs that no
    output reaches are driven by n_inputs primary-input sources.  ramp > 0:
    the supply rises from 0 in `ramp` seconds (a trivial DC operating point
    for circuits whose feedback paths leave the DC problem without a
    Newton-reachable solution)."""
    rng = np.random.default_rng(seed)
    nl = Netlist("random BSIM3 logic, %d gates" % n_gates)
    nl.add_model("NB", "bsim3", 1, **BSIM3_NMOS_130)
    nl.add_model("PB", "bsim3", -1, **BSIM3_PMOS_130)
    sv = supply(vdd, ramp)
    nl.add_vsrc("VDD", "vdd", "0", sv["dc"], sv.get("pulse"))
    for k in range(n_inputs):
        nl.add_vsrc("VIN%d" % k, "in%d" % k, "0", 0.0,
                    Pulse(0.0, vdd, float(rng.uniform(0, 1e-9)), 50e-12, 50e-12,
                          float(rng.uniform(0.3e-9, 1e-9)), float(rng.uniform(1.5e-9, 3e-9))))
    fanout = rng.integers(1, 5, n_gates)
    slots = np.repeat(np.arange(n_gates), fanout)
    rng.shuffle(slots)
    n_in = 2 * n_gates
    # driver of every gate input: a gate output, else a primary input
    drivers = [("g", int(slots[k])) if k < len(slots) else ("i", int(rng.integers(0, n_inputs)))
               for k in range(n_in)]
    order = rng.permutation(n_in)
    drivers = [drivers[k] for k in order]
    # RC line per gate output net: out -> seg1 -> ... -> end (fan-out point)
    end_node = {}
    for gidx in range(n_gates):
        prev = "g%d" % gidx
        nseg = int(rng.integers(1, 4))
        for s in range(nseg):
            node = "g%d_w%d" % (gidx, s)
            nl.add_res("R%d_%d" % (gidx, s), prev, node, float(rng.uniform(10, 200)))
            nl.add_cap("C%d_%d" % (gidx, s), node, "0", float(rng.uniform(1e-15, 20e-15)))
            prev = node
        end_node[gidx] = prev
    for gidx in range(n_gates):
        a_kind, a_idx = drivers[2 * gidx]
        b_kind, b_idx = drivers[2 * gidx + 1]
        a = end_node[a_idx] if a_kind == "g" else "in%d" % a_idx
        b = end_node[b_idx] if b_kind == "g" else "in%d" % b_idx
        out, mid = "g%d" % gidx, "g%d_m" % gidx
        geo = []
        for _ in range(4):
            w = 0.5e-6 * float(rng.uniform(0.8, 1.2))
            l = 0.13e-6 * float(rng.uniform(0.8, 1.2))
            geo.append(dict(W=w, L=l, **_diffusion(w)))
        if rng.random() < 0.5:   # NAND2: parallel PMOS, series NMOS
            nl.add_mos("MPA%d" % gidx, out, a, "vdd", "vdd", "PB", **geo[0])
            nl.add_mos("MPB%d" % gidx, out, b, "vdd", "vdd", "PB", **geo[1])
            nl.add_mos("MNA%d" % gidx, out, a, mid, "0", "NB", **geo[2])
            nl.add_mos("MNB%d" % gidx, mid, b, "0", "0", "NB", **geo[3])
        else:                    # NOR2: series PMOS, parallel NMOS
            nl.add_mos("MPA%d" % gidx, mid, a, "vdd", "vdd", "PB", **geo[0])
            nl.add_mos("MPB%d" % gidx, out, b, mid, "vdd", "PB", **geo[1])
            nl.add_mos("MNA%d" % gidx, out, a, "0", "0", "NB", **geo[2])
            nl.add_mos("MNB%d" % gidx, out, b, "0", "0", "NB", **geo[3])
    return nl


def random_node_voltages(nl, seed=0, lo=-0.2, hi=2.0):
    """Solution vector (length n_nodes + 1, ground slot last = 0) with
    node voltages drawn from U[lo, hi] so that every region is exercised."""
    rng = np.random.default_rng(seed)
    x = rng.uniform(lo, hi, nl.n_nodes + 1)
    x[-1] = 0.0
    return x


def power_grid_bsim3(nx=500, ny=500, n_inv=100000, seed=5, vdd=1.2, period=1e-9):
    """Resistive power-grid mesh driving BSIM3 inverter loads: an nx x ny
    mesh (R ~ U[0.05, 0.2] ohm between neighbours, 10 fF from every node to
    ground) fed by Vdd sources at its four corners; n_inv inverters at
    evenly spread mesh taps (PMOS source and bulk on the tap, NMOS on
    ground), each input a clock PULSE(0 vdd) of the given period with a
    phase drawn from U[0, period], each output loaded with 2 fF.  Seeded."""
    rng = np.random.default_rng(seed)
    nl = Netlist("power grid %dx%d, %d BSIM3 inverters" % (nx, ny, n_inv))
    nl.add_model("NB", "bsim3", 1, **BSIM3_NMOS_130)
    nl.add_model("PB", "bsim3", -1, **BSIM3_PMOS_130)
    g = lambda i, j: "g%d_%d" % (i, j)
    rh = rng.uniform(0.05, 0.2, (nx, ny - 1))
    rv = rng.uniform(0.05, 0.2, (nx - 1, ny))
    for i in range(nx):
        for j in range(ny):
            if j + 1 < ny:
                nl.add_res("RH%d_%d" % (i, j), g(i, j), g(i, j + 1), rh[i, j])
            if i + 1 < nx:
                nl.add_res("RV%d_%d" % (i, j), g(i, j), g(i + 1, j), rv[i, j])
            nl.add_cap("CG%d_%d" % (i, j), g(i, j), "0", 10e-15)
    for k, (i, j) in enumerate(((0, 0), (0, ny - 1), (nx - 1, 0), (nx - 1, ny - 1))):
        nl.add_vsrc("VDD%d" % k, g(i, j), "0", vdd)
    phase = rng.uniform(0.0, period, n_inv)
    taps = np.linspace(0, nx * ny - 1, n_inv).round().astype(int)
    wn, wp = 0.5e-6, 1.0e-6
    for k in range(n_inv):
        tap = g(taps[k] // ny, taps[k] % ny)
        inp, out = "in%d" % k, "out%d" % k
        nl.add_vsrc("VCLK%d" % k, inp, "0", 0.0,
                    Pulse(0.0, vdd, float(phase[k]), 50e-12, 50e-12,
                          period / 2 - 50e-12, period))
        nl.add_mos("MN%d" % k, out, inp, "0", "0", "NB", W=wn, L=0.13e-6, **_diffusion(wn))
        nl.add_mos("MP%d" % k, out, inp, tap, tap, "PB", W=wp, L=0.13e-6, **_diffusion(wp))
        nl.add_cap("CL%d" % k, out, "0", 2e-15)
    return nl


def rc_ladder_diodes(n_nodes=10000, n_diodes=1000, seed=4, vdd=1.0, r=(10.0, 100.0),
                     c=1e-15, model=None):
    """RC ladder of n_nodes nodes (R ~ U[r], C from every node to ground)
    driven at its head by a PULSE(0 vdd) source, with n_diodes junction
    diodes (IS = 1e-14, N = 1.0 unless `model` says otherwise) from evenly
    spread ladder nodes to ground.  Seeded."""
    rng = np.random.default_rng(seed)
    nl = Netlist("RC ladder, %d nodes, %d diodes" % (n_nodes, n_diodes))
    nl.add_model("DL", "diode", 1, **(model or dict(IS=1e-14, N=1.0)))
    nl.add_vsrc("VIN", "n0", "0", 0.0, Pulse(0.0, vdd, 0.0, 20e-12, 20e-12, 1e-9, 2e-9))
    rs = rng.uniform(r[0], r[1], n_nodes - 1)
    for k in range(n_nodes - 1):
        nl.add_res("R%d" % k, "n%d" % k, "n%d" % (k + 1), rs[k])
        nl.add_cap("C%d" % k, "n%d" % (k + 1), "0", c)
    taps = np.linspace(1, n_nodes - 1, n_diodes).round().astype(int)
    for j, t in enumerate(taps):
        nl.add_diode("D%d" % j, "n%d" % t, "0", "DL")
    return nl


def diode_variety():
    """Diodes in every regime the load distinguishes: forward, linear
    reverse, breakdown (BV), depletion and diffusion charge (CJO, TT, FC),
    AREA, OFF, IC, TNOM / TEMP, grounded and floating terminals."""
    nl = Netlist("diode variety")
    nl.add_model("D0", "diode", 1, IS=1e-14, N=1.0)
    nl.add_model("D1", "diode", 1, IS=2e-15, N=1.3, CJO=2e-15, VJ=0.8, M=0.4, FC=0.6, TT=1e-11)
    nl.add_model("D2", "diode", 1, IS=1e-13, BV=3.0, IBV=1e-4, CJO=1e-15, TNOM=40.0)
    nl.add_model("D3", "diode", 1, IS=5e-14, N=1.1, BV=5.0, IBV=10.0, IRF=2.0, XTI=2.0, EG=1.0)
    nl.add_vsrc("V1", "a", "0", 0.7)
    nl.add_res("R1", "a", "b", 50.0)
    nl.add_diode("DA", "b", "0", "D0")
    nl.add_diode("DB", "b", "c", "D1", AREA=2.0)
    nl.add_diode("DC", "0", "c", "D2")
    nl.add_diode("DD", "c", "d", "D3", TEMP=60.0)
    nl.add_diode("DE", "d", "b", "D1", OFF=1)
    nl.add_diode("DF", "a", "d", "D0", IC=0.3)
    nl.add_res("R2", "c", "0", 1e3)
    nl.add_cap("C1", "d", "0", 1e-15)
    return nl


def merge(parts, prefixes=None):
    """One netlist holding several independent circuits (they share only
    ground): node, model and device names get a per-part prefix.  The
    device load and the Newton solve then run every part in the same
    launches, one LU per part."""
    prefixes = prefixes or ["s%d_" % k for k in range(len(parts))]
    nl = Netlist("batch of %d circuits" % len(parts))
    ids = nl.node_ids
    for src, pre in zip(parts, prefixes):
        # keep each part's node order, so every sub-circuit has the same
        # unknown ordering (and so the same pivots) as when it runs alone;
        # every device node of a part is one of its nodes or ground
        ren = {g: g for g in GROUND_NAMES}
        for name in sorted(src.node_ids, key=src.node_ids.get):
            ren[name] = pre + name
            ids.setdefault(pre + name, len(ids))
        for name, m in src.models.items():
            nl.add_model(pre + name, m.kind, m.dtype, **m.params)
        nl.mos.extend(Mosfet(pre + m.name, ren[m.d], ren[m.g], ren[m.s], ren[m.b],
                             pre + m.model, dict(m.params)) for m in src.mos)
        nl.res.extend((pre + name, ren[a], ren[b], float(r)) for name, a, b, r in src.res)
        nl.caps.extend((pre + name, ren[a], ren[b], float(c)) for name, a, b, c in src.caps)
        nl.diodes.extend((pre + name, ren[a], ren[b], pre + model, dict(params))
                         for name, a, b, model, params in src.diodes)
        nl.vsrc.extend(Vsource(pre + v.name, ren[v.p], ren[v.n], float(v.dc), v.pulse)
                       for v in src.vsrc)
    return nl


def monte_carlo_rings(n_samples=256, seed=3, stages=(11, 301), dvth0=0.030, dtox=0.02,
                      ramp=50e-12):
    """Monte-Carlo samples of the BSIM3 ring oscillator: stage counts drawn
    uniformly from `stages` (inclusive), every device with its own card
    (VTH0 += N(0, dvth0), TOX *= U[1 - dtox, 1 + dtox]), seeded.  Returns
    the list of sample netlists."""
    rng = np.random.default_rng(seed)
    counts = rng.integers(stages[0], stages[1] + 1, n_samples)
    return [ring_oscillator_bsim3(int(s), dvth0=dvth0, dtox=dtox, ramp=ramp, rng=rng,
                                  prefix="m")
            for s in counts]
