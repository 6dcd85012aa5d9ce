"""GPU parity of the native driver's gmin stepping
(xg_tran_dcop_continuation) against the oracle's restatement: random BSIM3
logic where plain Newton fails at the operating point (20 gates, seed 2)
reaches it by continuation on the device, with the same number of
continuation solves and the same operating point to the DC Newton
tolerance; where Newton converges, mode "auto" gives Newton's point."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _dc(nl, mode):
    from xyce_amd.transient import NewtonFailure, Transient
    tr = Transient(nl)
    try:
        tr.dcop(mode)
        x = tr.x_curr.cpu().numpy().copy()
        ok = True
    except NewtonFailure:
        x, ok = None, False
    import ctypes
    from xyce_amd._lib import lib
    n = ctypes.c_int64()
    lib().xg_tran_gmin_steps(tr._tr, ctypes.byref(n))
    tr.close()
    return ok, x, n.value


def test_gmin_stepping_reaches_operating_point_newton_misses():
    from oracle.transient_ref import RefTransient
    from xyce_amd import circuits
    nl = circuits.random_logic_bsim3(n_gates=20, seed=2)
    nl.layout()
    ok, _, _ = _dc(nl, "none")
    assert not ok
    ok, x, steps = _dc(nl, "gmin")
    assert ok
    ref = RefTransient(nl)
    ref.dcop(continuation="gmin")
    print("continuation solves: device %d, oracle %d" % (steps, ref.gmin_steps))
    assert abs(steps - ref.gmin_steps) <= 2
    np.testing.assert_allclose(x[:nl.n_nodes], ref.x_curr[:nl.n_nodes], rtol=5e-3, atol=1e-6)
    ok, x2, _ = _dc(nl, "auto")
    assert ok
    np.testing.assert_allclose(x2, x, rtol=5e-3, atol=1e-6)


def test_auto_mode_is_newton_when_newton_converges():
    from xyce_amd import circuits
    nl = circuits.random_logic_bsim3(n_gates=20, seed=1)
    nl.layout()
    ok1, x1, _ = _dc(nl, "none")
    ok2, x2, steps = _dc(nl, "auto")
    assert ok1 and ok2 and steps == 0
    assert np.array_equal(x1, x2)
