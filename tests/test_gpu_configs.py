"""The benchmark configurations at their stated sizes on the GPU, through
the C ABI:

* the 101-stage MOSFET1 ring, 20 ns at 1 ps steps, against the oracle's
  waveform of the same run (tests/golden/c1_ring101_20ns.npz, made by
  tests/golden/make_c1_ring.py);
* the 256-sample Monte-Carlo BSIM3 ring batch (11-301 stages, seed 3): the
  merged load against the oracle entry by entry, three samples (the largest
  among them) bit-equal alone and in the batch, and the largest sample's
  waveform against the oracle's transient;
* the serving mix at its stated sizes (rings of 11-101 stages, random logic
  of 500-5,000 FETs, 10k-node ladders with 1k diodes), each kind's result
  equal alone and in the mix;
* the 500 x 500 power grid with 100,000 BSIM3 inverters: the full-size load
  finite and reproducible, 300 sampled inverters' rows against the oracle,
  the Jacobians against directional differences, and the same circuit
  family at 50 x 50 with 1,000 inverters against the oracle entry by entry.

Waveform tolerance: RELTOL 1e-3 of the supply plus 1e-6 V."""
import os

import numpy as np
import pytest

from conftest import compare_loads

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def _cu(torch, a):
    return torch.tensor(a, dtype=torch.float64, device="cuda")


# ---------------------------------------------------------------- C1
def test_c1_ring_101_stages_20ns_against_oracle():
    from xyce_amd import circuits
    from xyce_amd.transient import Transient
    g = np.load(os.path.join(HERE, "golden", "c1_ring101_20ns.npz"))
    probes = [str(p) for p in g["probes"]]
    tr = Transient(circuits.ring_oscillator_mos1(101, seed=1, ramp=50e-12))
    t, w = tr.run(float(g["tstop"]), float(g["dt"]), probes, every=int(g["every"]))
    tr.close()
    assert np.allclose(t, g["t"], rtol=0, atol=1e-18)
    err = np.abs(w - g["w"]).max()
    print("C1: max waveform difference %.3g V over %d samples; Newton iterations GPU %d, "
          "oracle %d" % (err, w.shape[1], tr.newton_total, int(g["newton"])), flush=True)
    assert err <= 1e-3 * 1.8 + 1e-6
    # it oscillates over the whole run (a stalled ring would also "match")
    late = w[0, w.shape[1] // 2:]
    assert late.max() > 1.6 and late.min() < 0.2


# ---------------------------------------------------------------- C3
def _c3_parts():
    from xyce_amd import circuits
    return circuits.monte_carlo_rings(256, seed=3)


def test_c3_batch_load_against_oracle():
    """One full load of the merged 256-sample netlist (80,686 BSIM3) at
    seeded voltages, every output entry against the oracle (1e-12)."""
    import torch
    from oracle.oracle import RefLoad
    from xyce_amd import circuits
    from xyce_amd.loader import DeviceLoad
    nl = circuits.merge(_c3_parts()).layout()
    assert len(nl.bsim3) == 80686
    rng = np.random.default_rng(31)
    n1 = nl.n_nodes + 1
    x = rng.uniform(-0.2, 1.4, n1)
    x[-1] = 0.0
    sta = rng.uniform(-1e-14, 1e-14, nl.n_state)
    sto = rng.uniform(-0.2, 1.4, nl.n_store)
    ss = dict(transient=1, newtonIter=1)
    ref = RefLoad(nl)
    ref.set_solver_state(**ss)
    want = ref.full_load(x, sta.copy(), sto.copy(), np.zeros(nl.n_state), sto.copy())
    dev = DeviceLoad(nl)
    dev.set_solver_state(**ss)
    bufs = [_cu(torch, a) for a in (x, sta, sto, np.zeros(nl.n_state), sto)]
    outs = [_cu(torch, np.zeros(n1)) for _ in range(4)] + \
           [_cu(torch, np.zeros(dev.nnz)) for _ in range(2)]
    dev.load_all(*bufs, *outs)
    dev.sync()
    compare_loads([o.cpu().numpy() for o in outs], want, ref.rowptr, ref.colidx, nl.n_nodes)
    dev.close()
    ref.close()


def test_c3_batch_samples_equal_alone_and_largest_matches_oracle():
    """The 256-sample batch, 0.1 ns at 1 ps after the DC point: three
    samples (the smallest, a middle one and the largest) give the bits they
    get alone; the largest sample's waveform matches the oracle's."""
    from oracle.transient_ref import RefTransient
    from xyce_amd.montecarlo import BatchedTransient
    from xyce_amd.transient import Transient
    parts = _c3_parts()
    sizes = [len(p.mos) // 2 for p in parts]
    big = int(np.argmax(sizes))
    picks = [int(np.argmin(sizes)), 128, big]
    tstop, dt = 0.1e-9, 1e-12
    b = BatchedTransient(parts)
    tb, wb = b.run(tstop, dt, ["out0", "out1"], every=5)
    print("C3 batch: %d Newton iterations, %d step failures; largest sample %d stages"
          % (b.tr.newton_total, b.tr.step_failures, sizes[big]), flush=True)
    b.tr.close()
    for k in picks:
        alone = Transient(_c3_parts()[k])
        ta, wa = alone.run(tstop, dt, ["out0", "out1"], every=5)
        alone.close()
        assert np.array_equal(ta, tb)
        if b.tr.step_failures == 0 and alone.step_failures == 0:
            assert wa.tobytes() == wb[k].tobytes(), k
        else:
            assert np.abs(wa - wb[k]).max() <= 1e-3 * 1.2 + 1e-6, k
    ref = RefTransient(_c3_parts()[big])
    tr_, wr = ref.run(tstop, dt, ["out0", "out1"], every=5)
    err = np.abs(wr - wb[big]).max()
    print("largest sample vs oracle: max difference %.3g V" % err, flush=True)
    assert np.allclose(tr_, tb)
    assert err <= 1e-3 * 1.2 + 1e-6


# ---------------------------------------------------------------- C4
def test_c4_serving_at_config_sizes():
    """Requests of the stated mix and sizes (seeded), served with
    continuous batching; one request of each kind is simulated again alone
    and must give its batched result (bit for bit unless a step of either
    run was halved, then within RELTOL)."""
    from xyce_amd.serving import UIC, poisson_requests, serve, summary
    from xyce_amd.transient import Transient
    reqs = poisson_requests(rate=200.0, n=24, seed=4, logic_gates=(125, 1250))
    kinds = {r.kind for r in reqs}
    assert kinds == {"ring", "logic", "ladder"}, kinds
    logic = [4 * r.size for r in reqs if r.kind == "logic"]
    assert min(logic) >= 500 and max(logic) <= 5000
    done = serve(reqs, max_batch=8, workers=3, window=0.002)
    s = summary(done)
    print("C4 serving:", s, flush=True)
    assert s["failed"] == 0, [r.error for r in done if r.error][:2]
    checked = set()
    for r in done:
        if r.kind in checked:
            continue
        checked.add(r.kind)
        alone = Transient(r._build())
        t, w = alone.run(r.tstop, r.dt, r.probes, every=10, uic=UIC[r.kind])
        alone.close()
        tb, wb = r.result
        assert np.array_equal(t, tb)
        if r.batch_info["step_failures"] == 0 and alone.step_failures == 0:
            assert np.array_equal(w, wb), r.kind
        else:
            vmax = 1.2 if r.kind != "ladder" else 1.0
            assert np.abs(w - wb).max() <= 1e-3 * vmax + 1e-6, r.kind
    assert checked == kinds


# ---------------------------------------------------------------- C5
def test_c5_power_grid_50x50_against_oracle():
    import torch
    from oracle.oracle import RefLoad
    from xyce_amd import circuits
    from xyce_amd.loader import DeviceLoad
    nl = circuits.power_grid_bsim3(nx=50, ny=50, n_inv=1000, seed=5).layout()
    rng = np.random.default_rng(51)
    n1 = nl.n_nodes + 1
    for ss, t in ((dict(dcop=1, tranop=1, newtonIter=1), 0.0),
                  (dict(transient=1, newtonIter=2), 0.37e-9)):
        x = rng.uniform(-0.2, 1.4, n1)
        x[-1] = 0.0
        sta = rng.uniform(-1e-14, 1e-14, nl.n_state)
        sto = rng.uniform(-0.2, 1.4, nl.n_store)
        ref = RefLoad(nl)
        ref.set_solver_state(**ss)
        ref.set_time(t)
        want = ref.full_load(x, sta.copy(), sto.copy(), np.zeros(nl.n_state), sto.copy())
        dev = DeviceLoad(nl)
        dev.set_solver_state(**ss)
        dev.set_time(t)
        bufs = [_cu(torch, a) for a in (x, sta, sto, np.zeros(nl.n_state), sto)]
        outs = [_cu(torch, np.zeros(n1)) for _ in range(4)] + \
               [_cu(torch, np.zeros(dev.nnz)) for _ in range(2)]
        dev.load_all(*bufs, *outs)
        dev.sync()
        compare_loads([o.cpu().numpy() for o in outs], want, ref.rowptr, ref.colidx,
                      nl.n_nodes)
        dev.close()
        ref.close()


def _inverter_sample(nl, ks):
    """The devices of inverters ks of the power grid (both transistors, the
    load capacitor, the clock source), each tap as a node of its own."""
    from xyce_amd.circuits import Netlist
    mini = Netlist("sampled inverters")
    for name, m in nl.models.items():
        mini.add_model(name, m.kind, m.dtype, **m.params)
    mos = {m.name: m for m in nl.mos}
    caps = {c[0]: c for c in nl.caps}
    srcs = {v.name: v for v in nl.vsrc}
    rename = {}
    for k in ks:
        tap = mos["MP%d" % k].s
        rename[tap] = "t%d" % k
        ren = lambda n, k=k: "t%d" % k if n == mos["MP%d" % k].s else n
        for nm in ("MN%d" % k, "MP%d" % k):
            m = mos[nm]
            mini.add_mos(nm, ren(m.d), ren(m.g), ren(m.s), ren(m.b), m.model, **m.params)
        _, a, b, c = caps["CL%d" % k]
        mini.add_cap("CL%d" % k, a, b, c)
        v = srcs["VCLK%d" % k]
        mini.add_vsrc(v.name, v.p, v.n, v.dc, v.pulse)
    return mini.layout()


def _compare_sampled_inverters(nl, ks, x, sta, sto, t, ss, a, rowptr, colidx):
    """Rows of inverters ks (inputs, outputs, clock branches) of a load `a`
    of the power grid, against the oracle evaluating those inverters'
    devices alone at the same voltages, state and store.  Returns the
    numbers of rows and matrix entries compared."""
    from conftest import assert_close
    from oracle.oracle import RefLoad
    F, Q, FL, QL, JF, JQ = a
    mini = _inverter_sample(nl, ks)
    big_id = dict(nl.node_ids)
    tap_of = {m.name[2:]: m.s for m in nl.mos if m.name.startswith("MP")}
    xm = np.zeros(mini.n_nodes + 1)
    for name, i in mini.node_ids.items():
        src = name if name in big_id else tap_of[name[1:]]
        xm[i] = x[big_id[src]]
    br_big = dict(zip([v.name for v in nl.vsrc], nl.vsrc_branch))
    for v, br in zip(mini.vsrc, mini.vsrc_branch):
        xm[br] = x[br_big[v.name]]
    # the sampled devices' state and store slots hold the full-size values
    sta_m, sto_m = np.zeros(mini.n_state), np.zeros(mini.n_store)
    recs_big = {r["name"]: r for r in nl.bsim3}
    for r in mini.bsim3:
        rb = recs_big[r["name"]]
        sta_m[r["sta"]:r["sta"] + 12] = sta[rb["sta"]:rb["sta"] + 12]
        sto_m[r["sto"]:r["sto"] + 5] = sto[rb["sto"]:rb["sto"] + 5]
    cap_big = {c[0]: st for c, st in zip(nl.caps, nl.cap_sta)}
    for c, st in zip(mini.caps, mini.cap_sta):
        sta_m[st] = sta[cap_big[c[0]]]
    ref = RefLoad(mini, limiter=False)
    ref.set_solver_state(**ss)
    ref.set_time(t)
    want = ref.full_load(xm, sta_m.copy(), sto_m.copy(), np.zeros(mini.n_state),
                         sto_m.copy())
    rows_big, rows_mini = [], []
    for k in ks:
        for nm in ("in%d" % k, "out%d" % k):
            rows_big.append(big_id[nm])
            rows_mini.append(mini.node_ids[nm])
    for v, br in zip(mini.vsrc, mini.vsrc_branch):
        rows_big.append(br_big[v.name])
        rows_mini.append(br)
    # column names: nodes by name (taps renamed back), branches by source
    mini_name = {i: (n if n in big_id else tap_of[n[1:]]) for n, i in mini.node_ids.items()}
    mini_name.update({br: ("br", v.name) for v, br in zip(mini.vsrc, mini.vsrc_branch)})
    inv_big = {i: n for n, i in big_id.items()}
    inv_big.update({br: ("br", nm) for nm, br in br_big.items()})
    for vec_g, vec_r, nm in zip(a[:4], want[:4], ["F", "Q", "dFdxdVp", "dQdxdVp"]):
        assert_close(vec_g[rows_big], vec_r[rows_mini], what="sampled " + nm)
    got_m, want_m = [], []
    for rb, rm in zip(rows_big, rows_mini):
        cols_g = {inv_big[colidx[kk]]: kk for kk in range(rowptr[rb], rowptr[rb + 1])
                  if colidx[kk] < nl.n_nodes}
        for kk in range(ref.rowptr[rm], ref.rowptr[rm + 1]):
            c = ref.colidx[kk]
            if c >= mini.n_nodes:
                continue
            kg = cols_g[mini_name[c]]
            got_m.append((JF[kg], JQ[kg]))
            want_m.append((want[4][kk], want[5][kk]))
    got_m, want_m = np.array(got_m), np.array(want_m)
    assert_close(got_m[:, 0], want_m[:, 0], what="sampled dF/dx")
    assert_close(got_m[:, 1], want_m[:, 1], what="sampled dQ/dx")
    ref.close()
    return len(rows_big), len(got_m)


def test_c5_full_size_load_properties():
    """The 500 x 500 mesh with 100,000 inverters (550,004 unknowns): every
    output finite and the same bits on a second load; the rows of 300
    sampled inverters' inputs, outputs and clock branches (which hold only
    that inverter's transistors, load capacitor and source) against the
    oracle evaluating those devices alone at the same voltages (1e-12);
    and dF/dx, dQ/dx applied to a random direction against central
    differences of F and Q (limiter off), to the 1e-3 the analytic BSIM3
    derivatives hold."""
    import scipy.sparse as sp
    import torch
    from xyce_amd import circuits
    from xyce_amd.loader import DeviceLoad
    nl = circuits.power_grid_bsim3().layout()
    print("C5: %d unknowns, %d BSIM3" % (nl.n_nodes, len(nl.bsim3)), flush=True)
    dev = DeviceLoad(nl, limiter=False)
    n1 = nl.n_nodes + 1
    rng = np.random.default_rng(55)
    x = rng.uniform(-0.2, 1.4, n1)
    x[-1] = 0.0
    sta = rng.uniform(-1e-14, 1e-14, nl.n_state)
    sto = rng.uniform(-0.2, 1.4, nl.n_store)
    t = 0.41e-9
    ss = dict(transient=1, newtonIter=1)
    dev.set_solver_state(**ss)
    dev.set_time(t)

    def load(xv):
        bufs = [_cu(torch, a) for a in (xv, sta, sto, np.zeros(nl.n_state), sto)]
        outs = [_cu(torch, np.zeros(n1)) for _ in range(4)] + \
               [_cu(torch, np.zeros(dev.nnz)) for _ in range(2)]
        dev.load_all(*bufs, *outs)
        dev.sync()
        return [o.cpu().numpy() for o in outs]

    a = load(x)
    b = load(x)
    for u, v in zip(a, b):
        assert np.all(np.isfinite(u))
        assert u.tobytes() == v.tobytes()
    assert dev.stats()[1] == 0
    F, Q, FL, QL, JF, JQ = a
    rowptr, colidx = dev.pattern()
    # sampled inverters against the oracle
    ks = np.sort(rng.choice(100000, 300, replace=False))
    n_rows, n_ent = _compare_sampled_inverters(nl, ks, x, sta, sto, t, ss, a, rowptr, colidx)
    print("C5: %d sampled rows, %d matrix entries match the oracle" % (n_rows, n_ent),
          flush=True)
    # directional derivatives (ground row and column dropped)
    rows = np.repeat(np.arange(n1), np.diff(rowptr))
    m = (rows < nl.n_nodes) & (colidx < nl.n_nodes)
    keep = slice(0, nl.n_nodes)
    d = np.zeros(n1)
    d[:nl.n_nodes] = rng.standard_normal(nl.n_nodes)
    h = 1e-6
    Fp, Qp = load(x + h * d)[:2]
    Fm, Qm = load(x - h * d)[:2]
    for J, vp, vm, nm in ((JF, Fp, Fm, "F"), (JQ, Qp, Qm, "Q")):
        A = sp.csr_matrix((J[m], (rows[m], colidx[m])), shape=(nl.n_nodes, nl.n_nodes))
        jd = A @ d[:nl.n_nodes]
        num = (vp[keep] - vm[keep]) / (2 * h)
        err = np.abs(jd - num)
        tol = 1e-3 * np.abs(num) + 1e-6 * np.abs(jd).max()
        assert np.all(err <= tol), (nm, int((err > tol).sum()), float(err.max()))
    dev.close()
