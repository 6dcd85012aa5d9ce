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
    v1: float
    v2: float
    td: float
    tr: float
    tf: float
    pw: float
    per: float


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
        self.ic = {}           # node -> initial voltage (.IC)
        self._laid_out = False

    # ---- building
    def node(self, name):
        if name in GROUND_NAMES:
            return -1
        if name not in self.node_ids:
            self.node_ids[name] = len(self.node_ids)
        return self.node_ids[name]

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
        self.vsrc_branch = []
        for v in self.vsrc:
            self.vsrc_branch.append(n)
            n += 1
        self.n_nodes, self.n_state, self.n_store = n, sta, sto
        self._laid_out = True
        return self

    def foreign_stamps(self):
        """(gids, rows, cols) of every device the load library does not
        evaluate: capacitor and resistor 2x2 blocks
        (N_DEV_Capacitor.C / N_DEV_Resistor.C jacStamp), voltage source
        [[0 0 1],[0 0 1],[1 1 0]]-shaped stamp (N_DEV_Vsrc.C)."""
        out = []
        for _, a, b, _C in self.caps + self.res:
            out.append(([self.node(a), self.node(b)], [0, 0, 1, 1], [0, 1, 0, 1]))
        for v, br in zip(self.vsrc, self.vsrc_branch):
            out.append(([self.node(v.p), self.node(v.n), br], [0, 1, 2, 2], [2, 2, 0, 1]))
        return out


# ---------------------------------------------------------------------------
# Benchmark netlists

MOS1_NMOS = dict(VTO=0.45, KP=120e-6)
MOS1_PMOS = dict(VTO=-0.45, KP=40e-6)


def ring_oscillator_mos1(stages=101, seed=1, vdd=1.8):
    """101-stage CMOS ring oscillator of MOSFET level-1 devices.

    NMOS W=1u L=0.18u VTO=0.45 KP=120u; PMOS W=2u L=0.18u VTO=-0.45 KP=40u;
    load capacitor per stage output drawn from U[5, 15] fF (seeded); Vdd 1.8 V.
    """
    rng = np.random.default_rng(seed)
    nl = Netlist("ring oscillator, %d stages, MOSFET level 1" % stages)
    nl.add_model("NMOS1", "mos1", 1, **MOS1_NMOS)
    nl.add_model("PMOS1", "mos1", -1, **MOS1_PMOS)
    nl.add_vsrc("VDD", "vdd", "0", vdd)
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
    LAMBDA, NSUB-derived parameters, LD, OFF, IC=, M=, both polarities and
    devices tied to ground on any terminal."""
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
        if rng.random() < 0.08:
            p.update(IC1=float(rng.uniform(0, 1.8)), IC2=float(rng.uniform(0, 1.8)),
                     IC3=float(rng.uniform(-0.5, 0)))
        nl.add_mos("M%d" % i, d, g, s, b, mdl, **p)
    return nl


def random_node_voltages(nl, seed=0, lo=-0.2, hi=2.0):
    """Solution vector (length n_nodes + 1, ground slot last = 0) with
    node voltages drawn from U[lo, hi] so that every region is exercised."""
    rng = np.random.default_rng(seed)
    x = rng.uniform(lo, hi, nl.n_nodes + 1)
    x[-1] = 0.0
    return x
