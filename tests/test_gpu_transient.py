"""GPU transient analysis (device load + assembly + dense LU per component
+ Newton update, all through the C ABI) against the oracle's CPU transient
and closed forms.

Waveform tolerance: 1e-3 relative to the supply plus 1e-6 V (the reference's
RELTOL and VNTOL); both sides run the same step sequence, so in practice the
waveforms agree far more closely."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rc():
    from xyce_amd.circuits import Netlist, Pulse
    nl = Netlist("rc")
    nl.add_vsrc("V1", "a", "0", 0.0, Pulse(0.0, 1.0, 0.0, 1e-12, 1e-12, 1.0, 2.0))
    nl.add_res("R1", "a", "b", 1e3)
    nl.add_cap("C1", "b", "0", 1e-12)
    return nl


def test_rc_step_matches_discrete_closed_form():
    """Backward Euler for two steps, then trapezoidal: the exact recurrence
    of the RC step response."""
    from xyce_amd.transient import Transient
    h, tau = 10e-12, 1e-9
    tr = Transient(_rc())
    t, w = tr.run(1e-9, h, ["a", "b"])
    vb = [0.0]
    for k in range(1, len(t)):
        va1, va0 = w[0, k], w[0, k - 1]
        if k <= 2:
            vb.append((vb[-1] + (h / tau) * va1) / (1 + h / tau))
        else:
            vb.append((vb[-1] * (1 - h / (2 * tau)) + (h / (2 * tau)) * (va1 + va0))
                      / (1 + h / (2 * tau)))
    assert np.max(np.abs(np.array(vb) - w[1])) < 1e-13


def _compare(nl_fn, tstop, dt, probes, vdd):
    from oracle.transient_ref import RefTransient
    from xyce_amd.transient import Transient
    g = Transient(nl_fn())
    tg, wg = g.run(tstop, dt, probes, every=5)
    r = RefTransient(nl_fn())
    tr_, wr = r.run(tstop, dt, probes, every=5)
    assert np.allclose(tg, tr_)
    err = np.max(np.abs(wg - wr))
    print("max waveform difference %.3g V, Newton iterations gpu %d oracle %d"
          % (err, g.newton_total, r.newton_total), flush=True)
    assert err <= 1e-3 * vdd + 1e-6
    return wg


def test_mos1_ring_waveform_parity():
    from xyce_amd import circuits
    w = _compare(lambda: circuits.ring_oscillator_mos1(11, seed=1, ramp=50e-12),
                 1.0e-9, 1e-12, ["out0", "out5", "vdd"], 1.8)
    assert w[0].max() > 1.5 and w[0].min() < 0.1       # it oscillates


def test_bsim3_ring_waveform_parity():
    from xyce_amd import circuits
    w = _compare(lambda: circuits.ring_oscillator_bsim3(11, seed=3, ramp=50e-12),
                 0.6e-9, 1e-12, ["out0", "out5"], 1.2)
    assert w[0].max() > 1.0


def _mos1_rings(stage_counts):
    """MOS1 rings (one connected component each) in one netlist."""
    from xyce_amd import circuits
    nl = circuits.Netlist("rings")
    for k, stages in enumerate(stage_counts):
        src, pre = circuits.ring_oscillator_mos1(stages, seed=k + 1), "r%d_" % k
        for name, m in src.models.items():
            nl.add_model(pre + name, m.kind, m.dtype, **m.params)
        for m in src.mos:
            ren = lambda n: n if n == "0" else pre + n
            nl.add_mos(pre + m.name, ren(m.d), ren(m.g), ren(m.s), ren(m.b), pre + m.model,
                       **m.params)
        for v in src.vsrc:
            nl.add_vsrc(pre + v.name, pre + v.p, "0", v.dc, v.pulse)
    nl.layout()
    return nl


def _dense_solve(nl, vals, rhs):
    """xg_dense_solve of one Newton update from x = 0 (weights 1):
    returns (x, per-component stats, component of each node)."""
    import torch
    from xyce_amd._lib import check, lib
    from xyce_amd.loader import DeviceLoad, _ptr
    dev = DeviceLoad(nl)
    n = nl.n_nodes
    cu = lambda x: torch.tensor(x, dtype=torch.float64, device="cuda")
    J, R, X = cu(vals), cu(rhs), cu(np.zeros(n + 1))
    W = cu(np.ones(n + 1))
    comp = _components(dev)
    stats = cu(np.zeros(4 * (comp.max() + 1)))
    check(lib().xg_dense_solve(dev.h, _ptr(J), _ptr(R), _ptr(X), _ptr(W), _ptr(W), 0.0, 1.0,
                               None, _ptr(stats)), "solve")
    out = X.cpu().numpy(), stats.cpu().numpy().reshape(-1, 4), comp
    dev.close()
    return out


def _components(dev):
    import ctypes
    from xyce_amd._lib import check, lib
    nc = ctypes.c_int32()
    comp = np.zeros(dev.n_rows - 1, np.int32)
    check(lib().xg_components(dev.h, ctypes.byref(nc), comp.ctypes.data), "components")
    return comp


def _component_system(rp, ci, vals, rhs, idx):
    n1 = len(rp) - 1
    rows = np.repeat(np.arange(n1), np.diff(rp))
    loc = -np.ones(n1, int)
    loc[idx] = np.arange(len(idx))
    A = np.zeros((len(idx), len(idx)))
    m = (loc[rows] >= 0) & (loc[ci] >= 0)
    A[loc[rows[m]], loc[ci[m]]] = vals[m]
    return A, rhs[idx]


def test_dense_solve_against_numpy():
    """The per-component LU (LDS and global-memory forms) solves random
    diagonally dominant systems on the circuit pattern like LAPACK."""
    from xyce_amd.loader import DeviceLoad
    nl = _mos1_rings([21, 301])        # one small component, one > 128 unknowns
    dev = DeviceLoad(nl, plan_only=True)
    rp, ci = dev.pattern()
    n = nl.n_nodes
    rng = np.random.default_rng(3)
    vals = rng.uniform(-1, 1, len(ci))
    rows = np.repeat(np.arange(n + 1), np.diff(rp))
    vals[rows == ci] += 20.0
    rhs = rng.uniform(-1, 1, n + 1)
    rhs[n] = 0.0
    x, st, comp = _dense_solve(nl, vals, rhs)
    assert comp.max() + 1 == 2
    for c in range(2):
        idx = np.flatnonzero(comp == c)
        A, b = _component_system(rp, ci, vals, rhs, idx)
        ref = np.linalg.solve(A, b)
        assert np.allclose(x[idx], ref, rtol=1e-12, atol=1e-14), c
    assert not st[:, 2].any()


def test_dense_solve_with_row_interchanges():
    """Random values with no dominant diagonal (partial pivoting swaps rows
    at most steps) on components in all four size classes of the dense
    LU: factored in LDS (<= 128 unknowns), in global memory with 16-column
    panels (<= 1024), 8-column panels (<= 1934) and 4-column panels
    (<= 2048, the largest with a dense form).  Each solution has
    a normwise backward error of a few ulps, and one column of zeros makes
    its component report a singular matrix."""
    from xyce_amd.loader import DeviceLoad
    nl = _mos1_rings([40, 500, 1200, 90, 1990])
    dev = DeviceLoad(nl, plan_only=True)
    rp, ci = dev.pattern()
    n = nl.n_nodes
    rng = np.random.default_rng(11)
    vals = rng.uniform(-1, 1, len(ci))
    rhs = rng.uniform(-1, 1, n + 1)
    rhs[n] = 0.0
    comp = _components(dev)
    sizes = np.bincount(comp)
    order = np.argsort(sizes)          # 40-, 90-, 500-, 1200- and 1990-stage rings
    assert sizes[order[1]] <= 128 < sizes[order[2]] <= 1024 < sizes[order[3]] <= 1934
    assert 1934 < sizes[order[4]] <= 2048
    # the 90-stage ring's component loses one column: singular
    sing = order[1]
    vals[np.isin(ci, np.flatnonzero(comp == sing)[5:6])] = 0.0
    x, st, comp2 = _dense_solve(nl, vals, rhs)
    assert np.array_equal(comp, comp2)
    for c in range(5):
        idx = np.flatnonzero(comp == c)
        if c == sing:
            assert st[c, 2] == 1.0
            continue
        assert st[c, 2] == 0.0
        A, b = _component_system(rp, ci, vals, rhs, idx)
        r = b - A @ x[idx]
        berr = np.abs(r).max() / (np.abs(A).sum(1).max() * np.abs(x[idx]).max() + np.abs(b).max())
        print("component %d: %d unknowns, backward error %.2e" % (c, len(idx), berr), flush=True)
        assert berr < 1e-13, c


def test_batched_monte_carlo_equals_each_alone():
    """Monte-Carlo BSIM3 rings of uneven sizes (per-device VTH0/TOX) run as
    one batch -- one device-load launch and one dense-solve launch per
    Newton iteration for all samples -- give each sample the waveform it
    gets when simulated alone, bit for bit: each sample's entries sum only
    its own devices' statements in the same order, and each converges on
    its own test."""
    from xyce_amd import circuits
    from xyce_amd.transient import Transient
    parts = circuits.monte_carlo_rings(4, seed=9, stages=(11, 41))
    batch = Transient(circuits.merge(parts))
    probes = ["s%d_out0" % k for k in range(len(parts))] + \
             ["s%d_out3" % k for k in range(len(parts))]
    tb, wb = batch.run(0.3e-9, 1e-12, probes, every=3)
    print("batch: %d Newton iterations, %d step failures"
          % (batch.newton_total, batch.step_failures), flush=True)
    for k, p in enumerate(parts):
        alone = Transient(circuits.monte_carlo_rings(4, seed=9, stages=(11, 41))[k])
        ta, wa = alone.run(0.3e-9, 1e-12, ["out0", "out3"], every=3)
        print("sample %d alone: %d Newton iterations, %d step failures, max diff %.3g"
              % (k, alone.newton_total, alone.step_failures,
                 np.abs(wa[0] - wb[k]).max()), flush=True)
        assert np.array_equal(tb, ta)
        assert wa[0].tobytes() == wb[k].tobytes(), k
        assert wa[1].tobytes() == wb[len(parts) + k].tobytes(), k


def _two_rings():
    from xyce_amd import circuits
    return circuits.merge([circuits.ring_oscillator_mos1(21, seed=1),
                           circuits.ring_oscillator_bsim3(301, seed=2)])


@pytest.mark.parametrize("mode", ["dominant", "fallback"])
def test_sparse_lu_against_numpy(mode):
    """xg_lu_solve solves per component like LAPACK.  "fallback": one
    component gets a tiny pivot under large off-diagonal entries, so its
    static pivots fail the backward-error test and the partial-pivoting LU
    takes it over in the same call; the other stays on the sparse path."""
    import ctypes
    import torch
    from xyce_amd._lib import check, lib
    from xyce_amd.loader import DeviceLoad, _ptr
    nl = _two_rings()
    nl.layout()
    dev = DeviceLoad(nl)
    rp, ci = dev.pattern()
    n = nl.n_nodes
    rng = np.random.default_rng(4)
    vals = rng.uniform(-1, 1, len(ci))
    rows = np.repeat(np.arange(n + 1), np.diff(rp))
    vals[rows == ci] += 20.0
    nc = ctypes.c_int32()
    comp = np.zeros(n, np.int32)
    check(lib().xg_components(dev.h, ctypes.byref(nc), comp.ctypes.data), "components")
    assert nc.value == 2
    if mode == "fallback":
        # component 0's first static pivot (nothing updates it): 1e-14, the
        # rest of its column 5.0: element growth ~5e14 ruins the static
        # factorisation's backward error
        P = dev.lu_plan()
        r0, c0 = int(P["prow"][0]), int(P["qcol"][0])  # position 0 of component 0
        assert comp[r0] == 0
        vals[(rows == r0) & (ci == c0)] = 1e-14
        col = (ci == c0) & (rows != r0) & (rows < n)
        assert col.any()
        vals[col] = 5.0
    rhs = rng.uniform(-1, 1, n + 1)
    rhs[n] = 0.0
    cu = lambda x: torch.tensor(x, dtype=torch.float64, device="cuda")
    J, R, X = cu(vals), cu(rhs), cu(np.zeros(n + 1))
    W = cu(np.ones(n + 1))
    stats = cu(np.zeros(4 * nc.value))
    check(lib().xg_lu_solve(dev.h, _ptr(J), _ptr(R), _ptr(X), _ptr(W), _ptr(W), 0.0, 1.0,
                            None, _ptr(stats)), "solve")
    info = dev.lu_info()
    print(mode, info, flush=True)
    assert info["fallbacks"] == (1 if mode == "fallback" else 0)
    x = X.cpu().numpy()
    for c in range(nc.value):
        idx = np.flatnonzero(comp == c)
        loc = -np.ones(n + 1, int)
        loc[idx] = np.arange(len(idx))
        A = np.zeros((len(idx), len(idx)))
        m = (loc[rows] >= 0) & (loc[ci] >= 0)
        A[loc[rows[m]], loc[ci[m]]] = vals[m]
        ref = np.linalg.solve(A, rhs[idx])
        assert np.allclose(x[idx], ref, rtol=1e-10, atol=1e-12 * np.abs(ref).max()), c
    st = stats.cpu().numpy().reshape(-1, 4)
    assert not st[:, 2].any()
    assert np.array_equal(st[:, 3], np.bincount(comp))


def test_sparse_lu_global_memory_form():
    """A component too large for LDS (an RC ladder of 6000 nodes) takes
    the global-memory form of the task-list LU when task lists are asked
    for (automatically it takes dense fronts); checked against a sparse
    direct solve."""
    import ctypes
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    import torch
    from xyce_amd import circuits
    from xyce_amd._lib import check, lib
    from xyce_amd.loader import DeviceLoad, _ptr
    nl = circuits.Netlist("long ladder")
    nl.add_vsrc("V1", "n0", "0", 1.0)
    for k in range(6000):
        nl.add_res("R%d" % k, "n%d" % k, "n%d" % (k + 1), 100.0)
        nl.add_cap("C%d" % k, "n%d" % (k + 1), "0", 1e-15)
    nl.layout()
    assert DeviceLoad(nl, plan_only=True).lu_info()["global_components"] == 0
    dev = DeviceLoad(nl, lu_mode="tasks")
    info = dev.lu_info()
    assert info["global_components"] == 1
    rp, ci = dev.pattern()
    n = nl.n_nodes
    rng = np.random.default_rng(7)
    vals = rng.uniform(-1, 1, len(ci))
    rows = np.repeat(np.arange(n + 1), np.diff(rp))
    vals[rows == ci] += 20.0
    rhs = rng.uniform(-1, 1, n + 1)
    rhs[n] = 0.0
    cu = lambda x: torch.tensor(x, dtype=torch.float64, device="cuda")
    J, R, X, W = cu(vals), cu(rhs), cu(np.zeros(n + 1)), cu(np.ones(n + 1))
    stats = cu(np.zeros(4))
    check(lib().xg_lu_solve(dev.h, _ptr(J), _ptr(R), _ptr(X), _ptr(W), _ptr(W), 0.0, 1.0,
                            None, _ptr(stats)), "solve")
    assert dev.lu_info()["fallbacks"] == 0
    m = (rows < n) & (ci < n)
    A = sp.csr_matrix((vals[m], (rows[m], ci[m])), shape=(n, n))
    ref = spla.spsolve(A.tocsc(), rhs[:n])
    x = X.cpu().numpy()[:n]
    assert np.allclose(x, ref, rtol=1e-10, atol=1e-12 * np.abs(ref).max())


def test_sparse_and_dense_solvers_give_the_same_waveforms():
    """A ring transient with the sparse LU and with the dense pivoting LU:
    the solvers round differently, the waveforms agree to far below the
    tolerance."""
    from xyce_amd import circuits
    from xyce_amd.transient import Transient
    mk = lambda: circuits.ring_oscillator_bsim3(31, seed=3, ramp=50e-12)
    a = Transient(mk(), solver="lu")
    ta, wa = a.run(0.3e-9, 1e-12, ["out0", "out7"], every=5)
    b = Transient(mk(), solver="dense")
    tb, wb = b.run(0.3e-9, 1e-12, ["out0", "out7"], every=5)
    err = np.abs(wa - wb).max()
    print("sparse vs dense: max difference %.3g V, Newton %d / %d, fallbacks now %d"
          % (err, a.newton_total, b.newton_total, a.dev.lu_info()["fallbacks"]), flush=True)
    assert np.array_equal(ta, tb)
    assert err <= 1e-6


@pytest.mark.parametrize("graphs", [True, False])
def test_native_driver_equals_python_driver(graphs):
    """The library's device-side Newton driver (with and without hipGraph
    replay) and the Python driver run the same kernels in the same order:
    the waveforms, the time grid and the Newton iteration counts agree bit
    for bit, on a ring and on a batch of Monte-Carlo samples."""
    from xyce_amd import circuits
    from xyce_amd.transient import Transient
    cases = [(lambda: circuits.ring_oscillator_mos1(11, seed=1, ramp=50e-12),
              ["out0", "out5", "vdd"]),
             (lambda: circuits.merge(circuits.monte_carlo_rings(3, seed=9, stages=(11, 31))),
              ["s0_out0", "s1_out2", "s2_out1"])]
    for mk, probes in cases:
        a = Transient(mk(), driver="native", graphs=graphs)
        ta, wa = a.run(0.2e-9, 1e-12, probes, every=4)
        b = Transient(mk(), driver="python")
        tb, wb = b.run(0.2e-9, 1e-12, probes, every=4)
        print("native (graphs %s) %d iterations, python %d"
              % (graphs, a.newton_total, b.newton_total), flush=True)
        assert np.array_equal(ta, tb)
        assert wa.tobytes() == wb.tobytes()
        assert a.newton_total == b.newton_total
        assert a.step_failures == b.step_failures
        a.close()


def test_checkpoint_restart_continues_bit_for_bit(tmp_path):
    """A run checkpointed half way and restarted in a fresh library
    instance ends in the state of the uninterrupted run, bit for bit
    (vectors, history, the diodes' held junction voltages, time, order)."""
    from xyce_amd import circuits
    from xyce_amd.transient import Transient
    mk = lambda: circuits.merge([circuits.ring_oscillator_bsim3(11, seed=3, ramp=50e-12),
                                 circuits.rc_ladder_diodes(60, 8, seed=2, c=2e-15)])
    full = Transient(mk())
    full.dcop()
    for _ in range(60):
        full.step(1e-12)
    a = Transient(mk())
    a.dcop()
    for _ in range(30):
        a.step(1e-12)
    a.checkpoint(tmp_path / "ck.npz")
    a.close()
    b = Transient(mk())
    b.restore(tmp_path / "ck.npz")
    for _ in range(30):
        b.step(1e-12)
    b._sync_counters()
    full._sync_counters()
    assert b.t == full.t and b.n_steps == full.n_steps
    assert b.x_curr.cpu().numpy().tobytes() == full.x_curr.cpu().numpy().tobytes()
    assert b.sto_c.cpu().numpy().tobytes() == full.sto_c.cpu().numpy().tobytes()
    assert b.newton_total == full.newton_total


def test_uic_start_logic_parity():
    """A random-logic transient started from zero (.TRAN UIC, supply
    ramping up) on the device against the oracle, and the two drivers bit
    for bit."""
    from oracle.transient_ref import RefTransient
    from xyce_amd import circuits
    from xyce_amd.transient import Transient
    mk = lambda: circuits.random_logic_bsim3(n_gates=40, seed=9, ramp=50e-12)
    g = Transient(mk())
    tg, wg = g.run(0.1e-9, 1e-12, ["g0", "g5"], every=5, uic=True)
    p = Transient(mk(), driver="python")
    tp, wp = p.run(0.1e-9, 1e-12, ["g0", "g5"], every=5, uic=True)
    r = RefTransient(mk())
    tr_, wr = r.run(0.1e-9, 1e-12, ["g0", "g5"], every=5, uic=True)
    err = np.abs(wg - wr).max()
    print("logic UIC: max waveform difference %.3g V, Newton %d / %d"
          % (err, g.newton_total, r.newton_total), flush=True)
    assert np.allclose(tg, tr_) and np.array_equal(tg, tp)
    assert wg.tobytes() == wp.tobytes()
    assert err <= 1e-3 * 1.2 + 1e-6


@pytest.mark.parametrize("dt,every", [(1e-12, 1), (1e-12, 3), (5e-12, 7), (2e-11, 2)])
def test_fused_steps_equal_one_step_per_launch(dt, every):
    """Fixed steps in groups of 8 per graph launch (xg_tran_run after
    xg_tran_prepare_run) against one step per launch: the same kernels in
    the same order, so the waveforms, the time grid, the step and Newton
    counts and the halvings agree bit for bit.  The larger steps need more
    iterations than a group's graph holds (the group halts and the host
    continues that step), and the largest fail and are halved; the run with
    progress reports goes in calls of 50 steps.  Fewer control-word reads
    with the groups."""
    import ctypes
    from xyce_amd import circuits
    from xyce_amd._lib import lib
    from xyce_amd.transient import Transient
    mk = lambda: circuits.merge(circuits.monte_carlo_rings(6, seed=11, stages=(11, 61)))
    probes = ["s0_out0", "s2_out3", "s5_out1", "s3_out0"]
    tstop = 60 * dt * every + 5 * dt
    res = []
    for fused in (True, False):
        tr = Transient(mk(), driver="native", fused_steps=fused)
        seen = []
        t, w = tr.run(tstop, dt, probes, every=every,
                      progress=(lambda s, n: seen.append(s)) if dt == 1e-12 else None)
        c = (ctypes.c_int64 * 4)()
        assert lib().xg_tran_counters(tr._tr, c) == 0
        res.append((t, w, tr.n_steps, tr.newton_total, tr.step_failures, c[0], seen))
        print("fused %s: %d steps, %d Newton iterations, %d halvings, %d control reads"
              % (fused, tr.n_steps, tr.newton_total, tr.step_failures, c[0]), flush=True)
        tr.close()
    (ta, wa, *ca, ra, sa), (tb, wb, *cb, rb, sb) = res
    assert len(ta) == int(round(tstop / dt)) // every + 1
    assert np.array_equal(ta, tb)
    assert wa.tobytes() == wb.tobytes()
    assert ca == cb
    assert sa == sb
    assert ra < rb
