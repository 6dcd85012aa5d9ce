"""GPU parity of the native driver's variable-step ONESTEP
(xg_tran_adaptive_*) against the oracle's restatement of the reference's
step control (oracle/transient_ref.py run_adaptive).

The loads match the oracle to rounding and the linear solves differ from
LAPACK's by rounding, so the error estimates agree to ~1e-12 relative and
the step sequence is the same except where an estimate sits on a decision
threshold.  A linear RC circuit must give the same time points and
waveform; nonlinear circuits must give the same number of accepted steps
within a few and the same waveform within the time integrator's tolerance
(RELTOL 1e-3 of the swing, checked at 5e-3)."""
import numpy as np
import pytest

from xyce_amd.circuits import Netlist, Pulse

pytestmark = pytest.mark.gpu


def _pair(nl, tstop, tstep, probes, uic=False):
    from oracle.transient_ref import RefTransient, run_adaptive
    from xyce_amd.transient import Transient
    tr = Transient(nl)
    t_g, w_g, c_g = tr.run_adaptive(tstop, tstep, probes, uic=uic)
    tr.close()
    t_r, w_r, c_r = run_adaptive(RefTransient(nl), tstop, tstep, probes, uic=uic)
    print("steps: device %d (%d / %d rejected), oracle %d (%d / %d rejected)" % (
        c_g["accepted"], c_g["rejected_error"], c_g["rejected_newton"], c_r.accepted,
        c_r.rejected_error, c_r.rejected_newton))
    return (t_g, w_g, c_g), (t_r, w_r, c_r)


def test_rc_pulse_same_steps_and_waveform():
    nl = Netlist("rc")
    nl.add_vsrc("V1", "in", "0", 0.0, Pulse(0.0, 1.0, 0.2e-9, 20e-12, 20e-12, 1e-9, 3e-9))
    nl.add_res("R1", "in", "out", 1e3)
    nl.add_cap("C1", "out", "0", 1e-12)
    nl.add_res("R2", "out", "o2", 2e3)
    nl.add_cap("C2", "o2", "0", 0.5e-12)
    nl.layout()
    (t_g, w_g, c_g), (t_r, w_r, c_r) = _pair(nl, 5e-9, 1e-11, ["out", "o2"])
    assert c_g["accepted"] == c_r.accepted
    assert c_g["rejected_error"] == c_r.rejected_error
    assert np.abs(t_g - t_r).max() <= 1e-12 * 5e-9
    assert np.abs(w_g - w_r).max() <= 1e-9
    # the PULSE corners are time points of the device run
    for bp in nl.breakpoints(5e-9):
        assert np.min(np.abs(t_g - bp)) <= 1e-21


def _compare_nonlinear(nl, tstop, tstep, probes, uic, swing):
    (t_g, w_g, c_g), (t_r, w_r, c_r) = _pair(nl, tstop, tstep, probes, uic=uic)
    assert abs(c_g["accepted"] - c_r.accepted) <= max(3, c_r.accepted // 50)
    for k in range(len(probes)):
        on_ref = np.interp(t_r, t_g, w_g[k])
        err = np.abs(on_ref - w_r[k]).max()
        assert err <= 5e-3 * swing, (probes[k], err)
    assert abs(t_g[-1] - tstop) <= 1e-10 * tstop


def test_diode_ladder_adaptive_matches_oracle():
    from xyce_amd import circuits
    nl = circuits.rc_ladder_diodes(n_nodes=60, n_diodes=12, seed=3, vdd=1.0)
    nl.layout()
    _compare_nonlinear(nl, 3e-9, 1e-11, ["n1", "n30", "n59"], False, 1.0)


def test_mos1_ring_adaptive_matches_oracle():
    from xyce_amd import circuits
    nl = circuits.ring_oscillator_mos1(stages=11, seed=2, ramp=50e-12)
    nl.layout()
    _compare_nonlinear(nl, 1.5e-9, 1e-12, ["out0", "out5"], True, 1.8)


def test_bsim3_ring_adaptive_matches_oracle():
    from xyce_amd import circuits
    nl = circuits.ring_oscillator_bsim3(stages=11, seed=4, ramp=50e-12)
    nl.layout()
    _compare_nonlinear(nl, 0.6e-9, 1e-12, ["out0", "out5"], True, 1.2)
