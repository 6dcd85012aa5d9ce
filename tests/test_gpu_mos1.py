"""GPU parity of the MOSFET level-1 load (C ABI on cuda:0) against the oracle.

For each solver state the reference visits during a transient (DCOP with
initJct, DCOP iterations, transient first iteration and later iterations)
both implementations get the same solution / state / store vectors, and every
output (F, Q, dFdxdVp, dQdxdVp, dF/dx, dQ/dx values, next state, next store)
is compared with the FP64 parity rule of conftest.assert_close.
"""
import numpy as np
import pytest

from conftest import assert_close, compare_loads

pytestmark = pytest.mark.gpu

SOLVER_STATES = [
    dict(dcop=1, initJct=1, newtonIter=0),
    dict(dcop=1, newtonIter=1),
    dict(dcop=1, tranop=1, initFix=1, newtonIter=2),
    dict(transient=1, newtonIter=0),
    dict(transient=1, newtonIter=3),
]


def _mos1_circuit(kind, seed):
    from xyce_amd import circuits
    if kind == "ring":
        return circuits.ring_oscillator_mos1(101, seed=seed)
    return circuits.mos1_variety(seed=seed)


def run_pair(nl, ss, seed, torch):
    from oracle.oracle import RefLoad
    from xyce_amd.loader import DeviceLoad
    rng = np.random.default_rng(seed)
    nl.layout()
    n1 = nl.n_nodes + 1
    x = rng.uniform(-0.2, 2.0, n1)
    x[-1] = 0.0
    sta_c = rng.uniform(-1e-14, 1e-14, nl.n_state)
    sto_c = rng.uniform(-0.5, 2.0, nl.n_store)
    sto_n0 = rng.uniform(-0.5, 2.0, nl.n_store)
    ref = RefLoad(nl)
    ref.set_solver_state(**ss)
    sta_n_r = np.zeros(nl.n_state)
    sto_n_r = sto_n0.copy()
    out_r = ref.full_load(x, sta_c.copy(), sto_c.copy(), sta_n_r, sto_n_r)

    dev = DeviceLoad(nl)
    cu = lambda a: torch.tensor(a, dtype=torch.float64, device="cuda")
    X, SC, TC = cu(x), cu(sta_c), cu(sto_c)
    SN, TN = cu(np.zeros(nl.n_state)), cu(sto_n0)
    F, Q, FL, QL = (cu(np.zeros(n1)) for _ in range(4))
    JF, JQ = cu(np.zeros(dev.nnz)), cu(np.zeros(dev.nnz))
    dev.set_solver_state(**ss)
    dev.update_state(X, SC, TC, SN, TN)
    dev.load_vectors(F, Q, FL, QL)
    dev.load_matrices(JF, JQ)
    dev.sync()
    out_g = [t.cpu().numpy() for t in (F, Q, FL, QL, JF, JQ)]
    lim_g, _ = dev.stats()
    lim_r, _ = ref.mos1_flags()
    return out_r, out_g, (sta_n_r, sto_n_r), (SN.cpu().numpy(), TN.cpu().numpy()), \
        int(lim_r.sum()), lim_g, (ref.rowptr, ref.colidx, nl.n_nodes)


@pytest.mark.parametrize("kind", ["ring", "variety"])
@pytest.mark.parametrize("si", range(len(SOLVER_STATES)))
def test_mos1_load_parity(kind, si):
    import torch
    nl = _mos1_circuit(kind, seed=11 + si)
    ss = SOLVER_STATES[si]
    print("case", kind, ss, flush=True)
    out_r, out_g, st_r, st_g, lim_r, lim_g, pat = run_pair(nl, ss, 100 + si, torch)
    compare_loads(out_g, out_r, *pat)
    assert_close(st_g[0], st_r[0], what="state")
    assert_close(st_g[1], st_r[1], what="store")
    assert lim_g == lim_r


def test_mos1_bitwise_deterministic():
    import torch
    from xyce_amd.loader import DeviceLoad
    nl = _mos1_circuit("variety", seed=3)
    nl.layout()
    rng = np.random.default_rng(5)
    n1 = nl.n_nodes + 1
    cu = lambda a: torch.tensor(a, dtype=torch.float64, device="cuda")
    x = rng.uniform(-0.2, 2.0, n1)
    x[-1] = 0
    dev = DeviceLoad(nl)
    dev.set_solver_state(transient=1, newtonIter=1)
    outs = []
    for _ in range(3):
        X, SC, TC = cu(x), cu(np.zeros(nl.n_state)), cu(np.zeros(nl.n_store))
        SN, TN = cu(np.zeros(nl.n_state)), cu(np.ones(nl.n_store))
        F, Q, FL, QL = (cu(np.zeros(n1)) for _ in range(4))
        JF, JQ = cu(np.zeros(dev.nnz)), cu(np.zeros(dev.nnz))
        dev.load_all(X, SC, TC, SN, TN, F, Q, FL, QL, JF, JQ)
        dev.sync()
        outs.append(np.concatenate([t.cpu().numpy() for t in (F, Q, FL, QL, JF, JQ, SN, TN)]))
    assert outs[0].tobytes() == outs[1].tobytes() == outs[2].tobytes()
