"""GPU parity of the junction-diode load (with resistors, capacitors and
sources around it) against the oracle, through the C ABI: F, Q, the
limiter vectors, both Jacobians and the store, in every solver state a
transient visits, over two consecutive evaluations (the store carries the
previous evaluation's junction voltage); and transient waveforms of an RC
ladder with diode loads."""
import numpy as np
import pytest

from conftest import assert_close, compare_loads

pytestmark = pytest.mark.gpu

STATES = [
    dict(dcop=1, tranop=1, initJct=1, newtonIter=0),
    dict(dcop=1, tranop=1, newtonIter=1),
    dict(transient=1, initTran=1, newtonIter=0),
    dict(transient=1, newtonIter=0),
    dict(transient=1, newtonIter=2),
]


def _nets():
    from xyce_amd import circuits
    return [circuits.diode_variety,
            lambda: circuits.rc_ladder_diodes(300, 40, seed=6,
                                              model=dict(IS=1e-14, N=1.0, CJO=1e-15,
                                                         BV=2.0, IBV=1e-5))]


@pytest.mark.parametrize("si", range(len(STATES)))
@pytest.mark.parametrize("ni", range(2))
def test_diode_load_parity(ni, si):
    import torch
    from oracle.oracle import RefLoad
    from xyce_amd.loader import DeviceLoad
    nl = _nets()[ni]()
    nl.layout()
    ss = STATES[si]
    n1 = nl.n_nodes + 1
    rng = np.random.default_rng(10 + ni)
    sto_c = rng.uniform(-3.0, 0.9, nl.n_store)
    ref = RefLoad(nl)
    dev = DeviceLoad(nl)
    cu = lambda a: torch.tensor(a, dtype=torch.float64, device="cuda")
    sto_n_r = rng.uniform(-1, 1, nl.n_store)
    TN = cu(sto_n_r)
    for call in range(2):
        x = rng.uniform(-4.0, 1.2, n1)
        x[-1] = 0.0
        ref.set_solver_state(**ss)
        out_r = ref.full_load(x, np.zeros(nl.n_state), sto_c.copy(), np.zeros(nl.n_state),
                              sto_n_r)
        lim_r = sum(0 if d.orig else 1 for d in ref.lin.diodes)
        X, SC, TC, SN = cu(x), cu(np.zeros(nl.n_state)), cu(sto_c), cu(np.zeros(nl.n_state))
        F, Q, FL, QL = (cu(np.zeros(n1)) for _ in range(4))
        JF, JQ = cu(np.zeros(dev.nnz)), cu(np.zeros(dev.nnz))
        dev.set_solver_state(**ss)
        dev.load_all(X, SC, TC, SN, TN, F, Q, FL, QL, JF, JQ)
        dev.sync()
        out_g = [t.cpu().numpy() for t in (F, Q, FL, QL, JF, JQ)]
        compare_loads(out_g, out_r, ref.rowptr, ref.colidx, nl.n_nodes)
        assert_close(TN.cpu().numpy(), sto_n_r, what="store, call %d" % call)
        lim_g, bad = dev.stats()
        assert lim_g == lim_r and bad == 0
    ref.close()


def test_diode_ladder_transient_parity():
    from oracle.transient_ref import RefTransient
    from xyce_amd import circuits
    from xyce_amd.transient import Transient
    mk = lambda: circuits.rc_ladder_diodes(200, 20, seed=4, vdd=1.0, r=(50.0, 200.0),
                                           c=2e-15)
    g = Transient(mk())
    tg, wg = g.run(0.3e-9, 1e-12, ["n0", "n50", "n199"], every=5)
    r = RefTransient(mk())
    tr_, wr = r.run(0.3e-9, 1e-12, ["n0", "n50", "n199"], every=5)
    err = np.abs(wg - wr).max()
    print("diode ladder: max waveform difference %.3g V, Newton %d / %d"
          % (err, g.newton_total, r.newton_total), flush=True)
    assert np.allclose(tg, tr_)
    assert err <= 1e-3 * 1.0 + 1e-6
    assert wg[1].max() > 0.1     # the pulse travels down the ladder
