"""xg_load_all_noalloc (the load with its statement values and reduction
partials in the library's static device array) against xg_load_all on the
same handle and inputs: the same bits in every output and in the next
state / store, on two calls in a row of which the second has the smaller
circuit, in both output modes and in the solver states that write
different statement sets (the initTran call writes the current state and
zero dQ/dx statements).  Called through the C ABI on cuda:0."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

STATES = [dict(transient=1, initTran=1, newtonIter=0), dict(transient=1, newtonIter=2),
          dict(dcop=1, tranop=1, initJct=1, newtonIter=0)]


def _inputs(nl, seed, torch):
    rng = np.random.default_rng(seed)
    n1 = nl.n_nodes + 1
    x = rng.uniform(-0.2, 1.4, n1)
    x[-1] = 0.0
    sta_c = rng.uniform(-1e-14, 1e-14, nl.n_state)
    sto_c = rng.uniform(-0.5, 1.4, nl.n_store)
    sto_n = rng.uniform(-0.5, 1.4, nl.n_store)
    cu = lambda a: torch.tensor(a, dtype=torch.float64, device="cuda")
    return [cu(x), cu(sta_c), cu(sto_c), cu(np.zeros(nl.n_state)), cu(sto_n)]


def _run(dev, fn, ins, ss, torch, overwrite):
    from xyce_amd._lib import check
    from xyce_amd.loader import _ptr
    n1 = dev.nl.n_nodes + 1
    bufs = [t.clone() for t in ins]
    # outputs start as garbage in overwrite mode, as zeros when accumulating
    fill = float("nan") if overwrite else 0.0
    outs = [torch.full((n1,), fill, dtype=torch.float64, device="cuda") for _ in range(4)]
    outs += [torch.full((dev.nnz,), fill, dtype=torch.float64, device="cuda")
             for _ in range(2)]
    dev.set_output_mode(overwrite)
    dev.set_solver_state(**ss)
    check(fn(dev.h, *[_ptr(t) for t in bufs + outs]), fn.__name__)
    dev.sync()
    return [t.cpu().numpy() for t in outs + bufs[1:]]


def test_noalloc_load_equals_load_all_on_two_calls():
    import torch
    from xyce_amd import circuits
    from xyce_amd._lib import lib
    from xyce_amd.loader import DeviceLoad

    big = circuits.random_logic_bsim3(n_gates=300, seed=2).layout()
    small = circuits.merge([circuits.ring_oscillator_mos1(11, seed=1),
                            circuits.diode_variety()]).layout()
    assert len(small.mos) < len(big.mos) and small.n_nodes < big.n_nodes
    L = lib()
    for ss in STATES:
        for overwrite in (False, True):
            print("case: state %s, overwrite %s" % (ss, overwrite), flush=True)
            got, want = {}, {}
            devs = {name: DeviceLoad(nl) for name, nl in (("big", big), ("small", small))}
            ins = {name: _inputs(d.nl, 5, torch) for name, d in devs.items()}
            # the no-allocation calls back to back, larger circuit first
            for name in ("big", "small"):
                got[name] = _run(devs[name], L.xg_load_all_noalloc, ins[name], ss, torch,
                                 overwrite)
            # references on fresh handles: a diode keeps its last junction
            # voltage in the handle and stores it at the next call, so a
            # second call on the same handle is not the same load
            fresh = {name: DeviceLoad(d.nl) for name, d in devs.items()}
            for name in ("big", "small"):
                want[name] = _run(fresh[name], L.xg_load_all, ins[name], ss, torch, overwrite)
            for name in ("big", "small"):
                labels = ["F", "Q", "dFdxdVp", "dQdxdVp", "dFdx", "dQdx", "sta_curr",
                          "sto_curr", "sta_next", "sto_next"]
                for lab, g, w in zip(labels, got[name], want[name]):
                    assert np.array_equal(g.view(np.uint64), w.view(np.uint64)), (name, lab)
            for d in list(devs.values()) + list(fresh.values()):
                d.close()


def test_noalloc_refuses_an_active_mask():
    import torch
    from xyce_amd import circuits
    from xyce_amd._lib import XG_E_STATE, lib
    from xyce_amd.loader import DeviceLoad, _iptr, _ptr

    nl = circuits.ring_oscillator_bsim3(11, seed=3).layout()
    dev = DeviceLoad(nl)
    L = lib()
    n_comp = ctypes.c_int32()
    L.xg_components(dev.h, ctypes.byref(n_comp), None)
    mask = torch.ones(max(n_comp.value, 1), dtype=torch.int32, device="cuda")
    assert L.xg_set_active_mask(dev.h, _iptr(mask)) == 0
    ins = _inputs(nl, 1, torch)
    outs = [torch.zeros(nl.n_nodes + 1, dtype=torch.float64, device="cuda") for _ in range(4)]
    outs += [torch.zeros(dev.nnz, dtype=torch.float64, device="cuda") for _ in range(2)]
    rc = L.xg_load_all_noalloc(dev.h, *[_ptr(t) for t in ins + outs])
    assert rc == XG_E_STATE, rc
    assert b"active mask" in L.xg_last_error()
    assert L.xg_set_active_mask(dev.h, None) == 0
    assert L.xg_load_all_noalloc(dev.h, *[_ptr(t) for t in ins + outs]) == 0
    dev.sync()
    dev.close()
