"""Pins of the oracle's junction-diode restatement (oracle/diode_ref.py)
against the mathematics the reference model fixes: the SPICE diode
equation and critical voltage, the derivatives of current and charge
(central differences), continuity of the piecewise current and charge at
their region boundaries, the depletion-capacitance law at the nominal
temperature, the pnjlim closed forms, KCL of the stamped F / Q / Jacobian,
and the store of the previous junction voltage."""
import math

import numpy as np
import pytest

from oracle.diode_ref import RefDiode, diode_fields, pnjlim
from xyce_amd.circuits import Netlist

T = 300.15
VT = 1.3806226e-23 / 1.6021918e-19 * T
GMIN = 1e-12


def _eval(mp, v, ip=None, limiter=False, flags=None):
    f = diode_fields(mp, ip or {}, T, T)
    d = RefDiode(f, 0, 1, 0, [0] * 7, [0] * 4)
    x = np.array([v, 0.0])
    sto = np.zeros(1)
    d.update_state(x, sto, sto.copy(), flags or dict(newtonIter=1, initJct=0, dcop=0),
                   GMIN, limiter)
    return d


def test_saturation_current_and_critical_voltage():
    f = diode_fields(dict(IS=3e-14, N=1.2), {}, T, T)
    vte = 1.2 * VT
    assert f["Isat"] == pytest.approx(3e-14, rel=1e-15)
    assert f["Vcrit"] == pytest.approx(vte * math.log(vte / (math.sqrt(2) * 3e-14)), rel=1e-14)
    f2 = diode_fields(dict(IS=3e-14), dict(AREA=2.5), T, T)
    assert f2["Isat"] == pytest.approx(2.5 * 3e-14, rel=1e-15)


@pytest.mark.parametrize("v", [-0.05, 0.0, 0.3, 0.65, 0.8])
def test_forward_current_is_the_diode_equation(v):
    d = _eval(dict(IS=1e-14), v)
    assert d.Id == pytest.approx(1e-14 * (math.exp(v / VT) - 1) + GMIN * v, rel=1e-13,
                                 abs=1e-27)


def _fd(fun, v, h):
    return (fun(v + h) - fun(v - h)) / (2 * h)


@pytest.mark.parametrize("v", [0.55, -0.02, -0.5, -2.0, -3.5, -4.5])
def test_conductance_is_the_current_derivative(v):
    mp = dict(IS=1e-13, N=1.1, BV=3.0, IBV=1e-4, IRF=2.0)
    g = _eval(mp, v).Gd
    num = _fd(lambda u: _eval(mp, u).Id, v, 1e-6)
    assert g == pytest.approx(num, rel=1e-5)


def test_current_is_continuous_at_region_boundaries():
    mp = dict(IS=1e-13, N=1.1, BV=3.0, IBV=1e-4)
    f = diode_fields(mp, {}, T, T)
    for edge in (-3.0 * f["Vte"], -f["Brk"]):
        a, b = _eval(mp, edge * (1 - 1e-12)).Id, _eval(mp, edge * (1 + 1e-12)).Id
        assert a == pytest.approx(b, rel=1e-9), edge


@pytest.mark.parametrize("v", [-1.0, -0.2, 0.2, 0.39, 0.41, 0.6])
def test_capacitance_is_the_charge_derivative(v):
    mp = dict(IS=1e-14, CJO=2e-15, VJ=0.8, M=0.4, FC=0.5, TT=1e-11)
    c = _eval(mp, v).Cd
    num = _fd(lambda u: _eval(mp, u).Qd, v, 1e-6)
    assert c == pytest.approx(num, rel=1e-5)


def test_depletion_capacitance_law_and_continuity():
    mp = dict(IS=1e-14, CJO=2e-15, VJ=0.8, M=0.4, FC=0.5)
    for v in (-1.0, -0.3, 0.0, 0.2, 0.35):
        assert _eval(mp, v).Cd == pytest.approx(2e-15 * (1 - v / 0.8) ** -0.4, rel=1e-12)
    edge = 0.5 * 0.8
    for name in ("Qd", "Cd"):
        a = getattr(_eval(mp, edge * (1 - 1e-12)), name)
        b = getattr(_eval(mp, edge * (1 + 1e-12)), name)
        assert a == pytest.approx(b, rel=1e-9), name


def test_breakdown_voltage_clips():
    vte = VT
    # IBV above the current at 3 N Vt of reverse bias: clipped to 3 N Vt
    assert diode_fields(dict(BV=0.2, IBV=1e-3), {}, T, T)["Brk"] == pytest.approx(3 * vte)
    assert diode_fields(dict(BV=5.0, IBV=1e-30), {}, T, T)["Brk"] == 5.0
    assert diode_fields(dict(IS=1e-14), {}, T, T)["Brk"] == 0.0


def test_pnjlim_closed_forms():
    v, lim = pnjlim(1.0, 0.6, 0.025, 0.7)
    assert lim and v == pytest.approx(0.6 + 0.025 * math.log(1 + 0.4 / 0.025), rel=1e-15)
    v, lim = pnjlim(1.0, -0.1, 0.025, 0.7)
    assert lim and v == pytest.approx(0.025 * math.log(1.0 / 0.025), rel=1e-15)
    v, lim = pnjlim(0.5, 0.0, 0.025, 0.7)
    assert not lim and v == 0.5


def test_kcl_store_and_limiter_terms():
    """Through the oracle's device load: F, Q and the limiter vectors
    conserve current and charge at the two terminals, every Jacobian row
    sums to zero, and the store holds the previous evaluation's junction
    voltage (Master::updateState writes it before the new evaluation)."""
    from oracle.oracle import RefLoad
    nl = Netlist("one diode")
    nl.add_model("D", "diode", 1, IS=1e-14, CJO=1e-15, TT=1e-12)
    nl.add_diode("D1", "a", "b", "D")
    nl.layout()
    ref = RefLoad(nl)
    n1 = nl.n_nodes + 1
    a, b = nl.node("a"), nl.node("b")
    x = np.zeros(n1)
    x[a], x[b] = 0.9, 0.1
    sto_c = np.array([0.3])
    sto_n = np.array([0.0])
    ref.set_solver_state(transient=1, newtonIter=0)
    F, Q, FL, QL, JF, JQ = ref.full_load(x, np.zeros(0), sto_c, np.zeros(0), sto_n)
    assert sto_n[0] == 0.0                      # instance Vd before any evaluation
    vd1 = ref.lin.diodes[0].Vd
    assert vd1 < 0.8                            # limited from the 0.3 V history
    assert F[a] + F[b] == 0.0 and Q[a] + Q[b] == 0.0
    assert FL[a] + FL[b] == 0.0 and QL[a] + QL[b] == 0.0
    J = np.zeros((n1, n1))
    rows = np.repeat(np.arange(n1), np.diff(ref.rowptr))
    np.add.at(J, (rows, ref.colidx), JF)
    assert np.allclose(J.sum(axis=1), 0.0, atol=1e-25)
    ref.set_solver_state(transient=1, newtonIter=1)
    ref.full_load(x, np.zeros(0), sto_c, np.zeros(0), sto_n)
    assert sto_n[0] == vd1
    ref.close()
