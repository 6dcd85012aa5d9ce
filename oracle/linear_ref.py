"""ORACLE — test infrastructure only; the product never imports this module.

Restatement of the linear devices' Master loops (statement for statement,
in Python floats, which are IEEE doubles):

* Resistor  N_DEV_Resistor.C:1073-1174: updateState i0 = (v+ - v-) * G;
  loadDAEVectors F[+] += i0, F[-] += -i0; loadDAEMatrices +G, -G, -G, +G
  into (+,+), (+,-), (-,+), (-,-).
* Capacitor N_DEV_Capacitor.C:1821-2060 (no IC=, constant C): updateState
  q0 = C (v+ - v-), state[li_QState] = q0; loadDAEVectors Q[+] += q0,
  Q[-] += -q0; dQ/dx +C, -C, -C, +C.
* Vsrc      N_DEV_Vsrc.C:941-1036: source = PULSE value at tranop /
  transient iterations, else the DC value; F[+] += i_br, F[-] += -i_br,
  F[br] += (v+ - v-), F[br] -= source; dF/dx +1 (+,br), -1 (-,br),
  +1 (br,+), -1 (br,-).
* PULSE     N_DEV_SourceData.C:928-1000 (PulseData::updateSource).
* Diode     oracle/diode_ref.py (level 1, RS = 0).

Device types load after the MOSFETs, in the order resistor, capacitor,
diode, source, each in netlist order.
"""
import math

from . import stamp_ref
from .diode_ref import RefDiode, diode_fields


def pulse_value(p, t, bp_tol):
    """PulseData::updateSource for p = [V1, V2, TD, TR, TF, PW, PER]."""
    V1, V2, TD, TR, TF, PW, PER = p
    time = t - TD
    if time > PER and PER != 0.0:
        time -= PER * math.floor(time / PER)
    if time <= 0 or (time > (TR + PW + TF) and abs(time - (TR + PW + TF)) > bp_tol):
        return V1
    if (time > TR and abs(time - TR) > bp_tol) and \
            (time < (TR + PW) or abs(time - (TR + PW)) < bp_tol):
        return V2
    if time > 0 and (time < TR or abs(time - TR) < bp_tol):
        return V1 + (V2 - V1) * time / TR if TR != 0.0 else V1
    return V2 + (V1 - V2) * (time - (TR + PW)) / TF if TF != 0.0 else V2


class RefLinear:
    def __init__(self, nl, rowptr, colidx, rowmaps=None, gmin=1e-12, limiter=True,
                 temp=300.15, tnom=300.15):
        self.n = nl.n_nodes
        gl = lambda g: self.n if g < 0 else g
        self.dev = []
        self.diodes = []
        self.gmin, self.limiter = gmin, limiter
        rowmaps = {} if rowmaps is None else rowmaps
        for kind, p, q, br, val, pulse, st in nl.linear_devices():
            if kind == "D":
                off = stamp_ref.stamp_offsets(rowptr, colidx, self.n, [p, q], [[0, 1], [0, 1]],
                                              rowmaps)
                P, N = gl(p), gl(q)
                pp, pn = rowptr[P] + off[0][0], rowptr[P] + off[0][1]
                np_, nn = rowptr[N] + off[1][0], rowptr[N] + off[1][1]
                # dF/dx: (P,P) (P,Pri) (N,N) (N,Pri) (Pri,P) (Pri,N) (Pri,Pri), Pri = P
                pos_f = [pp, pp, nn, np_, pp, pn, pp]
                pos_q = [nn, np_, pn, pp]
                self.diodes.append(RefDiode(diode_fields(val[0], val[1], temp, tnom), P, N, st,
                                            pos_f, pos_q))
                continue
            if kind == "V":
                gids = [p, q, br]
                stamp = [[2], [2], [0, 1]]
            else:
                gids = [p, q]
                stamp = [[0, 1], [0, 1]]
            off = stamp_ref.stamp_offsets(rowptr, colidx, self.n, gids, stamp, rowmaps)
            if kind == "V":
                pos = [rowptr[gl(p)] + off[0][0], rowptr[gl(q)] + off[1][0],
                       rowptr[gl(br)] + off[2][0], rowptr[gl(br)] + off[2][1]]
            else:
                pos = [rowptr[gl(p)] + off[0][0], rowptr[gl(p)] + off[0][1],
                       rowptr[gl(q)] + off[1][0], rowptr[gl(q)] + off[1][1]]
            g = 1.0 / val if kind == "R" else val
            self.dev.append(dict(kind=kind, p=gl(p), q=gl(q), br=gl(br) if br >= 0 else -1,
                                 val=g, pulse=pulse, sta=st, pos=pos))
        self.flags = dict(tranop=0, transient=0, dcop=0, initJct=0, newtonIter=0)
        self.time = 0.0
        self.bp_tol = 0.0

    def set_solver_state(self, tranop=0, transient=0, dcop=0, initJct=0, newtonIter=0, **_):
        self.flags = dict(tranop=tranop, transient=transient, dcop=dcop, initJct=initJct,
                          newtonIter=newtonIter)

    def set_time(self, t, bp_tol=0.0):
        self.time, self.bp_tol = float(t), float(bp_tol)

    def update_state(self, x, sta_n, sto_c=None, sto_n=None):
        for dio in self.diodes:
            dio.update_state(x, sto_c, sto_n, self.flags, self.gmin, self.limiter)
        for d in self.dev:
            vp, vn = float(x[d["p"]]), float(x[d["q"]])
            if d["kind"] == "R":
                d["i0"] = (vp - vn) * d["val"]
            elif d["kind"] == "C":
                d["q0"] = d["val"] * (vp - vn)
                sta_n[d["sta"]] = d["q0"]
            else:
                use_pulse = (self.flags["tranop"] or self.flags["transient"]) and d["pulse"]
                d["src"] = pulse_value(d["pulse"], self.time, self.bp_tol) if use_pulse \
                    else d["val"]
                d["ibr"] = float(x[d["br"]])
                d["drop"] = vp - vn

    def load_vectors(self, F, Q, FL=None, QL=None):
        for d in self.dev:
            if d["kind"] == "R":
                F[d["p"]] += d["i0"]
                F[d["q"]] += -d["i0"]
        for d in self.dev:
            if d["kind"] == "C":
                Q[d["p"]] += d["q0"]
                Q[d["q"]] += -d["q0"]
        for dio in self.diodes:
            dio.load_vectors(F, Q, FL, QL, self.limiter)
        for d in self.dev:
            if d["kind"] == "V":
                F[d["p"]] += d["ibr"]
                F[d["q"]] += -d["ibr"]
                F[d["br"]] += d["drop"]
                F[d["br"]] -= d["src"]

    def load_matrices(self, dFdx, dQdx):
        for kind in ("R", "C", "D", "V"):
            if kind == "D":
                for dio in self.diodes:
                    dio.load_matrices(dFdx, dQdx)
                continue
            for d in self.dev:
                if d["kind"] != kind:
                    continue
                a, b, c, e = d["pos"]
                if kind == "R":
                    G = d["val"]
                    dFdx[a] += G
                    dFdx[b] -= G
                    dFdx[c] -= G
                    dFdx[e] += G
                elif kind == "C":
                    C = d["val"]
                    dQdx[a] += C
                    dQdx[b] -= C
                    dQdx[c] -= C
                    dQdx[e] += C
                else:
                    dFdx[a] += 1.0
                    dFdx[b] -= 1.0
                    dFdx[c] += 1.0
                    dFdx[e] -= 1.0
