"""GPU parity of the BSIM3 load (C ABI on cuda:0) against the oracle.

Every solver state the reference visits in a transient is exercised (DCOP
with initJct, DCOP iterations with and without charge computation, the
first transient iteration at initTran, later iterations), on a netlist that
covers mobMod 1/2/3, the three charge partitions, sheet resistance, IJTH,
NGATE, binning and off-nominal temperatures, and on the random-logic
workload.  All outputs and the next state/store vectors are compared with
the FP64 rule of conftest.assert_close.
"""
import numpy as np
import pytest

from conftest import assert_close, compare_loads

pytestmark = pytest.mark.gpu

SOLVER_STATES = [
    dict(dcop=1, tranop=1, initJct=1, newtonIter=0),
    dict(dcop=1, tranop=1, newtonIter=1),
    dict(dcop=1, newtonIter=2),                      # no charge computation
    dict(transient=1, initTran=1, newtonIter=0),
    dict(transient=1, newtonIter=0),
    dict(transient=1, newtonIter=3),
]


def run_pair(nl, ss, seed, torch, vhi=1.5, variant="auto"):
    from oracle.oracle import RefLoad
    from xyce_amd.loader import DeviceLoad
    rng = np.random.default_rng(seed)
    nl.layout()
    n1 = nl.n_nodes + 1
    x = rng.uniform(-0.2, vhi, n1)
    x[-1] = 0.0
    sta_c = rng.uniform(-1e-14, 1e-14, nl.n_state)
    sto_c = rng.uniform(-0.5, vhi, nl.n_store)
    sto_n0 = rng.uniform(-0.5, vhi, nl.n_store)
    ref = RefLoad(nl)
    ref.set_solver_state(**ss)
    sta_c_r = sta_c.copy()
    sta_n_r = np.zeros(nl.n_state)
    sto_n_r = sto_n0.copy()
    out_r = ref.full_load(x, sta_c_r, sto_c.copy(), sta_n_r, sto_n_r)
    lim_r, _ = ref.b3.flags()

    dev = DeviceLoad(nl)
    dev.set_eval_variant(variant)
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
    st_g = (SN.cpu().numpy(), TN.cpu().numpy(), SC.cpu().numpy())
    return (out_r, out_g, (sta_n_r, sto_n_r, sta_c_r), st_g, int(lim_r.sum()), lim_g,
            (ref.rowptr, ref.colidx, nl.n_nodes))


def check(out_r, out_g, st_r, st_g, lim_r, lim_g, pat):
    compare_loads(out_g, out_r, *pat)
    assert_close(st_g[0], st_r[0], what="next state")
    assert_close(st_g[1], st_r[1], what="next store")
    assert_close(st_g[2], st_r[2], what="current state (initTran copy)")
    assert lim_g == lim_r


@pytest.mark.parametrize("seed", [1, 2])
@pytest.mark.parametrize("si", range(len(SOLVER_STATES)))
def test_bsim3_variety_parity(seed, si):
    import torch
    from xyce_amd import circuits
    nl = circuits.bsim3_variety(seed=seed)
    ss = SOLVER_STATES[si]
    print("case variety", seed, ss, flush=True)
    check(*run_pair(nl, ss, 100 * seed + si, torch))


@pytest.mark.parametrize("si", [0, 5])
def test_bsim3_random_logic_parity(si):
    import torch
    from xyce_amd import circuits
    nl = circuits.random_logic_bsim3(n_gates=300, seed=2)
    ss = SOLVER_STATES[si]
    print("case random logic 300 gates", ss, flush=True)
    check(*run_pair(nl, ss, 7 + si, torch, vhi=1.3))


def test_bsim3_random_logic_full_size_parity():
    """The benchmark workload itself (5,000 gates, 20,000 BSIM3 instances):
    the supply rail collects contributions from 10,000 devices, which is the
    wavefront-folded long-segment path of the reduction."""
    import torch
    from xyce_amd import circuits
    nl = circuits.random_logic_bsim3(n_gates=5000, seed=2)
    print("case random logic 5000 gates", flush=True)
    check(*run_pair(nl, SOLVER_STATES[5], 11, torch, vhi=1.5))


def test_bsim3_two_eval_waves_per_workgroup_odd_count():
    """Above 256 eval waves the split kernel packs two eval waves into each
    workgroup; with an odd count the last workgroup's second group has no
    instances.  4,200 gates give 2 x 132 waves; one extra card of ten
    NMOS adds the 265th."""
    import torch
    from xyce_amd import circuits
    logic = circuits.random_logic_bsim3(n_gates=4200, seed=5)
    extra = circuits.Netlist("one card")
    extra.add_model("NX", "bsim3", 1, **dict(circuits.BSIM3_NMOS_130, MOBMOD=2, TNOM=40.0))
    for i in range(10):
        extra.add_mos("MX%d" % i, "x%d" % i, "xg", "0", "0", "NX", W=1e-6, L=0.13e-6)
    nl = circuits.merge([logic, extra])
    counts = {}
    for m in nl.mos:
        if nl.models[m.model].kind == "bsim3":
            counts[m.model] = counts.get(m.model, 0) + 1
    waves = sum((c + 63) // 64 for c in counts.values())
    assert waves > 256 and waves % 2 == 1, waves
    print("case %d eval waves" % waves, flush=True)
    check(*run_pair(nl, SOLVER_STATES[5], 13, torch, vhi=1.5))


def test_bsim3_monte_carlo_cards_parity():
    """Per-device model cards (VTH0 / TOX perturbed) share one LDS card per
    polarity; results still match the oracle's per-model evaluation."""
    import torch
    from xyce_amd import circuits
    nl = circuits.ring_oscillator_bsim3(stages=31, seed=3, dvth0=0.03, dtox=0.02)
    check(*run_pair(nl, SOLVER_STATES[5], 9, torch, vhi=1.3))


def test_bsim3_bitwise_deterministic():
    import torch
    from xyce_amd import circuits
    from xyce_amd.loader import DeviceLoad
    nl = circuits.bsim3_variety(seed=4)
    nl.layout()
    rng = np.random.default_rng(5)
    n1 = nl.n_nodes + 1
    cu = lambda a: torch.tensor(a, dtype=torch.float64, device="cuda")
    x = rng.uniform(-0.2, 1.5, n1)
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


def test_nonfinite_inputs_are_counted():
    """A NaN node voltage reaches every instance that reads the node; the
    stats report exactly those instances as non-finite (the load itself
    completes: no evaluation step loops on data)."""
    import torch
    from xyce_amd import circuits
    from xyce_amd.loader import DeviceLoad
    nl = circuits.random_logic_bsim3(n_gates=200, seed=6)
    nl.layout()
    n1 = nl.n_nodes + 1
    rng = np.random.default_rng(8)
    x = rng.uniform(-0.2, 1.5, n1)
    x[-1] = 0.0
    bad = nl.node(nl.mos[7].g)
    x[bad] = np.nan
    expect = sum(1 for r in nl.bsim3 if bad in r["lids"])
    assert expect > 0
    cu = lambda a: torch.tensor(a, dtype=torch.float64, device="cuda")
    dev = DeviceLoad(nl)
    dev.set_solver_state(transient=1, newtonIter=1)
    F, Q, FL, QL = (cu(np.zeros(n1)) for _ in range(4))
    JF, JQ = cu(np.zeros(dev.nnz)), cu(np.zeros(dev.nnz))
    dev.load_all(cu(x), cu(np.zeros(nl.n_state)), cu(np.zeros(nl.n_store)),
                 cu(np.zeros(nl.n_state)), cu(np.zeros(nl.n_store)), F, Q, FL, QL, JF, JQ)
    dev.sync()
    _, n_nonfinite = dev.stats()
    assert n_nonfinite == expect
    x[bad] = 0.5
    dev.load_all(cu(x), cu(np.zeros(nl.n_state)), cu(np.zeros(nl.n_store)),
                 cu(np.zeros(nl.n_state)), cu(np.zeros(nl.n_store)), F, Q, FL, QL, JF, JQ)
    dev.sync()
    assert dev.stats()[1] == 0


@pytest.mark.parametrize("variant", ["hbm", "staged", "staged_split"])
@pytest.mark.parametrize("si", [0, 2, 3, 5])
def test_bsim3_every_kernel_form(variant, si):
    """The three forms of the evaluation kernel (records from HBM; staged in
    LDS; staged and split into DC and charge waves) each match the oracle,
    whatever the circuit size would select."""
    import torch
    from xyce_amd import circuits
    nl = circuits.bsim3_variety(seed=7)
    print("case", variant, SOLVER_STATES[si], flush=True)
    check(*run_pair(nl, SOLVER_STATES[si], 50 + si, torch, variant=variant))


def test_overwrite_mode_equals_zero_then_accumulate():
    """XG_OUT_OVERWRITE writes every output entry: the bits equal zeroing
    the outputs and accumulating, except the ground slot/row/column, which
    it sets to +0.0 (the accumulate mode leaves them to the caller)."""
    import torch
    from xyce_amd import circuits
    from xyce_amd.loader import DeviceLoad
    nl = circuits.mos1_variety(seed=5, n_dev=40)
    v = circuits.bsim3_variety(seed=5, n_dev=60)
    for name, m in v.models.items():
        nl.add_model("B" + name, m.kind, m.dtype, **m.params)
    for m in v.mos:
        nl.add_mos("B" + m.name, m.d, m.g, m.s, m.b, "B" + m.model, **m.params)
    nl.add_res("R1", "n0", "n1", 100.0)
    nl.add_cap("C1", "x_only_cap", "0", 1e-15)   # an entry no MOSFET reaches
    nl.layout()
    n1 = nl.n_nodes + 1
    rng = np.random.default_rng(21)
    x = rng.uniform(-0.2, 1.5, n1)
    x[-1] = 0.0
    cu = lambda a: torch.tensor(a, dtype=torch.float64, device="cuda")
    dev = DeviceLoad(nl)
    dev.set_solver_state(transient=1, newtonIter=1)
    sto = rng.uniform(-0.2, 1.5, nl.n_store)

    def run(fill, overwrite):
        # fresh state/store each time: the load reads the next store it wrote
        state = [cu(np.zeros(nl.n_state)), cu(sto), cu(np.zeros(nl.n_state)), cu(sto)]
        dev.set_output_mode(overwrite)
        outs = [cu(np.full(n1, fill)) for _ in range(4)] + \
               [cu(np.full(dev.nnz, fill)) for _ in range(2)]
        dev.load_all(cu(x), *state, *outs)
        dev.sync()
        return [t.cpu().numpy() for t in outs]

    acc = run(0.0, False)
    ovw = run(7.5, True)
    ovw2 = run(-3.0, True)
    rowptr, colidx = dev.pattern()
    rows = np.repeat(np.arange(n1), np.diff(rowptr))
    gmask = (rows == nl.n_nodes) | (colidx == nl.n_nodes)
    for k in range(6):
        a, o, o2 = acc[k].copy(), ovw[k], ovw2[k]
        g = np.zeros(len(a), bool)
        if k < 4:
            g[nl.n_nodes] = True
        else:
            g = gmask
        assert np.all(o[g] == 0.0) and not np.signbit(o[g]).any()
        a[g] = 0.0
        assert a.tobytes() == o.tobytes() == o2.tobytes(), k


def test_graph_replay_matches_eager():
    """The whole load captured in a hipGraph (the capture records only kernel
    launches: no memset, no allocation) and replayed several times, with
    eager work and synchronisation between replays, gives the eager result
    bit for bit on every replay (both output modes)."""
    import torch
    from xyce_amd import circuits
    from xyce_amd.loader import DeviceLoad
    nl = circuits.random_logic_bsim3(n_gates=400, seed=12)
    nl.layout()
    n1 = nl.n_nodes + 1
    rng = np.random.default_rng(13)
    x = rng.uniform(-0.2, 1.5, n1)
    x[-1] = 0.0
    sto = rng.uniform(-0.2, 1.5, nl.n_store)
    cu = lambda a: torch.tensor(a, dtype=torch.float64, device="cuda")
    dev = DeviceLoad(nl)
    dev.set_solver_state(transient=1, newtonIter=1)
    X = cu(x)
    SC, TC, SN, TN = cu(np.zeros(nl.n_state)), cu(sto), cu(np.zeros(nl.n_state)), cu(sto)
    outs = [cu(np.zeros(n1)) for _ in range(4)] + [cu(np.zeros(dev.nnz)) for _ in range(2)]
    for overwrite in (False, True):
        dev.set_output_mode(overwrite)

        def step():
            if not overwrite:
                for t in outs:
                    t.zero_()
            TN.copy_(TC)
            dev.load_all(X, SC, TC, SN, TN, *outs)

        step()
        torch.cuda.synchronize()
        eager = [t.cpu().numpy().tobytes() for t in outs + [SN, TN]]
        stream = torch.cuda.Stream()
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.stream(stream):
            dev.set_stream(stream.cuda_stream)
            with torch.cuda.graph(graph, stream=stream):
                step()
        dev.set_stream(torch.cuda.current_stream().cuda_stream)
        for rep in range(3):
            for t in outs:
                t.fill_(float(rep + 1))       # eager work between replays
            torch.cuda.synchronize()
            graph.replay()
            torch.cuda.synchronize()
            got = [t.cpu().numpy().tobytes() for t in outs + [SN, TN]]
            assert got == eager, (overwrite, rep)


def test_update_state_by_type_matches_all():
    """xg_update_state_type for MOS1, BSIM3 and the linear devices in turn
    writes exactly what one xg_update_state writes (states, stores and
    every load that follows)."""
    import torch
    from xyce_amd import circuits
    from xyce_amd.loader import DeviceLoad
    nl = circuits.merge([circuits.mos1_variety(seed=3), circuits.bsim3_variety(seed=3),
                         circuits.diode_variety()])
    nl.layout()
    n1 = nl.n_nodes + 1
    rng = np.random.default_rng(7)
    x = rng.uniform(-0.2, 1.5, n1)
    x[-1] = 0.0
    sto = rng.uniform(-0.5, 1.5, nl.n_store)
    cu = lambda a: torch.tensor(a, dtype=torch.float64, device="cuda")
    res = []
    for by_type in (False, True):
        # a fresh handle each: the diode keeps its last junction voltage
        dev = DeviceLoad(nl)
        dev.set_solver_state(transient=1, newtonIter=1)
        X, SC, TC = cu(x), cu(np.zeros(nl.n_state)), cu(sto)
        SN, TN = cu(np.zeros(nl.n_state)), cu(sto)
        outs = [cu(np.zeros(n1)) for _ in range(4)] + [cu(np.zeros(dev.nnz)) for _ in range(2)]
        if by_type:
            for kind in ("mos1", "bsim3", "linear"):
                dev.update_state_type(kind, X, SC, TC, SN, TN)
        else:
            dev.update_state(X, SC, TC, SN, TN)
        dev.load_vectors(*outs[:4])
        dev.load_matrices(*outs[4:])
        dev.sync()
        res.append([t.cpu().numpy().tobytes() for t in outs + [SN, TN, SC]])
        dev.close()
    assert res[0] == res[1]
