"""The host-pointer loads (xg_update_state_host, xg_load_vectors_host,
xg_load_matrices_host), as a caller whose vectors stay in host memory uses
them: the same bits as the device-pointer calls on the same inputs, in
both output modes, in solver states that write the current state
(initTran) and that read the previous iteration's store (the limiters),
over two consecutive iterations."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

STATES = [dict(transient=1, initTran=1, newtonIter=0), dict(transient=1, newtonIter=1),
          dict(dcop=1, tranop=1, initJct=1, newtonIter=0)]


def _cases():
    from xyce_amd import circuits
    return [circuits.random_logic_bsim3(n_gates=100, seed=4),
            circuits.mos1_variety(seed=2),
            circuits.diode_variety()]


@pytest.mark.parametrize("ci", range(3))
@pytest.mark.parametrize("overwrite", [False, True])
def test_host_loads_equal_device_loads(ci, overwrite):
    import ctypes
    import torch
    from xyce_amd._lib import check, lib
    from xyce_amd.loader import DeviceLoad
    L = lib()
    nl = _cases()[ci].layout()
    print("case", nl.title, "overwrite", overwrite, flush=True)
    n1 = nl.n_nodes + 1
    rng = np.random.default_rng(ci)
    dh, dd = DeviceLoad(nl), DeviceLoad(nl)
    for d in (dh, dd):
        d.set_output_mode(overwrite)
    sta_c = rng.uniform(-1e-14, 1e-14, nl.n_state)
    sto_c = rng.uniform(-0.2, 1.4, nl.n_store)
    sto_n = rng.uniform(-0.2, 1.4, nl.n_store)
    for ss in STATES:
        x = rng.uniform(-0.2, 1.4, n1)
        x[-1] = 0.0
        init = [rng.uniform(-1, 1, n1) for _ in range(4)] + \
               [rng.uniform(-1, 1, dh.nnz) for _ in range(2)]
        # host-pointer calls on numpy arrays
        hs = [a.copy() for a in (x, sta_c, sto_c, np.zeros(nl.n_state), sto_n)]
        ho = [a.copy() for a in init]
        dh.set_solver_state(**ss)
        p = lambda a: a.ctypes.data_as(ctypes.c_void_p)
        check(L.xg_update_state_host(dh.h, *[p(a) for a in hs]), "update host")
        check(L.xg_load_vectors_host(dh.h, *[p(a) for a in ho[:4]]), "vectors host")
        check(L.xg_load_matrices_host(dh.h, *[p(a) for a in ho[4:]]), "matrices host")
        # device-pointer calls
        cu = lambda a: torch.tensor(a, dtype=torch.float64, device="cuda")
        ds = [cu(a) for a in (x, sta_c, sto_c, np.zeros(nl.n_state), sto_n)]
        do = [cu(a) for a in init]
        dd.set_solver_state(**ss)
        dd.update_state(*ds)
        dd.load_vectors(*do[:4])
        dd.load_matrices(*do[4:])
        dd.sync()
        for a, b, nm in zip(hs[1:] + ho, ds[1:] + do,
                            ["sta_curr", "sto_curr", "sta_next", "sto_next", "F", "Q",
                             "dFdxdVp", "dQdxdVp", "dFdx", "dQdx"]):
            assert a.tobytes() == b.cpu().numpy().tobytes(), (nl.title, ss, nm)
        # the next iteration starts from this one's store
        sto_n = hs[4].copy()
        sta_c = hs[1].copy()
    dh.close()
    dd.close()


def test_host_loads_need_update_first():
    import ctypes
    from xyce_amd import circuits
    from xyce_amd._lib import XG_E_STATE, lib
    from xyce_amd.loader import DeviceLoad
    nl = circuits.ring_oscillator_bsim3(5, seed=1).layout()
    d = DeviceLoad(nl)
    v = [np.zeros(nl.n_nodes + 1) for _ in range(4)]
    assert lib().xg_load_vectors_host(d.h, *[a.ctypes.data_as(ctypes.c_void_p)
                                            for a in v]) == XG_E_STATE
    d.close()
