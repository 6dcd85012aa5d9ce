"""Pins of the oracle's variable-step ONESTEP (oracle/transient_ref.py
run_adaptive, the restatement of N_TIA_StepErrorControl / N_TIA_OneStep
step control) by what the mathematics fixes:

* an RC low-pass driven by a PULSE: the waveform follows the closed form
  1 - exp(-(t - t1) / RC) (t1 the end of the rise) to the order of the
  relative tolerance, and tightening the tolerance by 100x lowers the error
  and raises the step count between the second-order law (h ~ tol^(1/3),
  x 4.6) and the first-order one (h ~ tol^(1/2), x 10): the controller
  moves to order 2 only while the estimate stays below about 0.43 of the
  tolerance (completeStep's ratio test > 1.05), so both orders occur;
* the PULSE corners are time points (breakpoints reached exactly), the
  first step is min(TSTEP, 0.005 x time to the first breakpoint), no step
  exceeds 0.1 TSTOP, and the step doubles at most per accepted step away
  from breakpoints;
* a linear source ramp integrates exactly (the trapezoidal rule is exact
  for a piecewise-linear input into a capacitor charged through a
  resistor only up to its own error; a pure capacitor current source is
  exact), so the error estimate there is at rounding level.
"""
import math

import numpy as np

from xyce_amd.circuits import Netlist, Pulse

RC = 1e3 * 1e-12


def _rc(td=0.2e-9, tr=1e-12):
    nl = Netlist("rc")
    nl.add_vsrc("V1", "in", "0", 0.0, Pulse(0.0, 1.0, td, tr, tr, 10e-9, 20e-9))
    nl.add_res("R1", "in", "out", 1e3)
    nl.add_cap("C1", "out", "0", 1e-12)
    nl.layout()
    return nl


def _run(nl, tstop, tstep, reltol=1e-3):
    from oracle.transient_ref import RefTransient, run_adaptive
    tr = RefTransient(nl)
    return run_adaptive(tr, tstop, tstep, ["out", "in"], reltol=reltol)


def _exact(t, td=0.2e-9, tr=1e-12):
    """RC response to a ramp from td to td + tr then a step: closed form of
    v' = (u - v) / RC."""
    out = np.zeros_like(t)
    for k, tt in enumerate(t):
        if tt <= td:
            out[k] = 0.0
        elif tt <= td + tr:
            s = tt - td
            out[k] = (s - RC * (1 - math.exp(-s / RC))) / tr
        else:
            v1 = (tr - RC * (1 - math.exp(-tr / RC))) / tr
            out[k] = 1.0 + (v1 - 1.0) * math.exp(-(tt - td - tr) / RC)
    return out


def test_breakpoints_are_the_pulse_corners():
    nl = Netlist("p")
    nl.add_vsrc("V1", "a", "0", 0.0, Pulse(0.0, 1.0, 1e-9, 0.1e-9, 0.2e-9, 1e-9, 3e-9))
    nl.add_res("R1", "a", "0", 1e3)
    nl.layout()
    want = [1e-9, 1.1e-9, 2.1e-9, 2.3e-9, 4e-9, 4.1e-9, 5.1e-9, 5.3e-9, 7e-9, 7.1e-9]
    got = nl.breakpoints(7.5e-9)
    assert np.allclose(got, want, rtol=1e-12, atol=0), got


def test_rc_follows_closed_form_and_hits_breakpoints(oracle_lib):
    nl = _rc()
    t, w, st = _run(nl, 5e-9, 1e-11)
    err = np.abs(w[0] - _exact(t))
    assert err.max() <= 2e-3, err.max()
    for bp in (0.2e-9, 0.2e-9 + 1e-12):
        assert np.min(np.abs(t - bp)) <= 1e-22, bp
    h = np.diff(t)
    assert abs(h[0] - min(1e-11, 0.005 * 0.2e-9)) <= 1e-25
    assert h.max() <= 0.1 * 5e-9 * (1 + 1e-12)
    # away from breakpoints a step at most doubles the one before
    ratio = h[1:] / h[:-1]
    assert ratio.max() <= 2.0 + 1e-9
    assert st.accepted == len(t) - 1


def test_step_count_follows_second_order_law(oracle_lib):
    nl = _rc()
    t1, w1, s1 = _run(nl, 5e-9, 1e-11, reltol=1e-3)
    t2, w2, s2 = _run(_rc(), 5e-9, 1e-11, reltol=1e-5)
    e1 = np.abs(w1[0] - _exact(t1)).max()
    e2 = np.abs(w2[0] - _exact(t2)).max()
    assert e2 < e1 / 5, (e1, e2)
    # the smooth part of the run (after the rise) sets the count
    n1 = np.count_nonzero(t1 > 0.3e-9)
    n2 = np.count_nonzero(t2 > 0.3e-9)
    r = n2 / n1
    assert 100 ** (1 / 3) * 0.9 <= r <= 100 ** 0.5 * 1.1, (n1, n2, r)


def test_constant_current_into_capacitor_has_no_error_rejections(oracle_lib):
    """i = C dv/dt with a resistor-free ramp source: every step's estimate
    is at the level of the Q change only, so none is rejected and the
    waveform is the source itself."""
    nl = Netlist("ramp")
    nl.add_vsrc("V1", "in", "0", 0.0, Pulse(0.0, 1.0, 0.1e-9, 2e-9, 2e-9, 1e-9, 10e-9))
    nl.add_cap("C1", "in", "0", 1e-12)
    nl.add_res("R1", "in", "out", 1.0)
    nl.add_cap("C2", "out", "0", 1e-18)
    nl.layout()
    from oracle.transient_ref import RefTransient, run_adaptive
    tr = RefTransient(nl)
    t, w, st = run_adaptive(tr, 3e-9, 1e-11, ["in"])
    src = np.clip((t - 0.1e-9) / 2e-9, 0.0, 1.0)
    assert np.abs(w[0] - src).max() <= 1e-9
    assert st.rejected_newton == 0
