"""The dense-front (multifrontal) LU on the GPU (xg_lu_solve through the
C ABI): solves like LAPACK on the circuit families of the CPU plan test
with every component forced through fronts, agrees with the task-list
solver, gives the same transient waveforms, and solves the 20,000-FET
random logic's Newton system, which is too large for task lists, like
SuperLU."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _cases():
    from xyce_amd import circuits
    return {
        "mos1 ring": circuits.ring_oscillator_mos1(11, seed=1),
        "bsim3 ring": circuits.ring_oscillator_bsim3(21, seed=3),
        "random logic": circuits.random_logic_bsim3(n_gates=60, seed=5),
        "diode ladder": circuits.rc_ladder_diodes(120, 12, seed=2),
        "power grid": circuits.power_grid_bsim3(nx=12, ny=12, n_inv=40, seed=5),
        "batch": circuits.merge([circuits.ring_oscillator_mos1(7, seed=1),
                                 circuits.ring_oscillator_bsim3(9, seed=2)]),
    }


def _solve(dev, vals, rhs, torch):
    from xyce_amd._lib import check, lib
    from xyce_amd.loader import _ptr
    n = dev.n_rows - 1
    nc = ctypes.c_int32()
    check(lib().xg_components(dev.h, ctypes.byref(nc), None), "components")
    cu = lambda a: torch.tensor(a, dtype=torch.float64, device="cuda")
    J, R, X, W = cu(vals), cu(rhs), cu(np.zeros(n + 1)), cu(np.ones(n + 1))
    st = cu(np.zeros(4 * nc.value))
    check(lib().xg_lu_solve(dev.h, _ptr(J), _ptr(R), _ptr(X), _ptr(W), _ptr(W), 0.0, 1.0,
                            None, _ptr(st)), "lu solve")
    dev.sync()
    return X.cpu().numpy(), st.cpu().numpy().reshape(-1, 4)


def _dense(rp, ci, vals, n):
    A = np.zeros((n, n))
    for r in range(n):
        for k in range(rp[r], rp[r + 1]):
            if ci[k] < n:
                A[r, ci[k]] += vals[k]
    return A


@pytest.mark.parametrize("name", list(_cases()))
def test_fronts_solve_like_lapack_and_the_task_lists(name):
    import torch
    from xyce_amd.loader import DeviceLoad
    nl = _cases()[name].layout()
    print("case:", name, flush=True)
    fr = DeviceLoad(nl, lu_fronts=True)
    tk = DeviceLoad(nl, lu_mode="tasks")
    info = np.zeros(6, np.int64)
    from xyce_amd._lib import lib
    lib().xg_lu_front_info(fr.h, info.ctypes.data)
    assert info[1] > 0
    rp, ci = fr.pattern()
    n = fr.n_rows - 1
    rng = np.random.default_rng(11)
    rows_of = np.repeat(np.arange(len(rp) - 1), np.diff(rp))
    for trial in range(3):
        vals = rng.standard_normal(len(ci))
        if trial == 0:
            vals[rows_of == ci] += 20.0
        rhs = rng.standard_normal(n + 1)
        rhs[n] = 0.0
        x, st = _solve(fr, vals, rhs, torch)
        A = _dense(rp, ci, vals, n)
        if np.linalg.matrix_rank(A) < n:
            continue
        assert np.all(st[:, 2] == 0), st[:, 2]
        if trial == 0:
            # a dominant diagonal: the fronts solve it, no pivoting fallback
            assert fr.lu_info()["fallbacks"] == 0
        want = np.linalg.solve(A, rhs[:n])
        berr = np.abs(A @ x[:n] - rhs[:n]).max() / (np.abs(A).sum(1).max() *
                                                    np.abs(x[:n]).max() + np.abs(rhs[:n]).max())
        assert berr < 1e-13, (name, trial, berr)
        assert np.abs(x[:n] - want).max() <= 1e-8 * np.abs(want).max(), (name, trial)
        xt, stt = _solve(tk, vals, rhs, torch)
        if np.all(stt[:, 2] == 0):
            assert np.abs(x[:n] - xt[:n]).max() <= 1e-9 * np.abs(want).max()
        # the same bits on a second solve
        x2, _ = _solve(fr, vals, rhs, torch)
        assert x2.tobytes() == x.tobytes()
    fr.close()
    tk.close()


def test_fronts_transient_matches_task_lists():
    """A BSIM3 ring's transient with every Newton solve on dense fronts
    matches the task-list run within the Newton tolerance."""
    from xyce_amd import circuits
    from xyce_amd.transient import Transient
    a = Transient(circuits.ring_oscillator_bsim3(21, seed=3, ramp=50e-12), lu_fronts=True)
    ta, wa = a.run(0.3e-9, 1e-12, ["out0", "out7"], every=5)
    a.close()
    b = Transient(circuits.ring_oscillator_bsim3(21, seed=3, ramp=50e-12))
    tb, wb = b.run(0.3e-9, 1e-12, ["out0", "out7"], every=5)
    b.close()
    err = np.abs(wa - wb).max()
    print("fronts vs task lists: %.3g V, Newton %d vs %d" % (err, a.newton_total,
                                                              b.newton_total), flush=True)
    assert np.array_equal(ta, tb)
    assert err <= 1e-3 * 1.2 + 1e-6
    assert np.ptp(wa) > 0.3   # the supply ramp moves the nodes


def test_random_logic_5000_gates_newton_system():
    """The 20,000-FET random logic (one 20k-unknown component, beyond the
    task lists): a transient Newton system assembled on the device, solved
    by the fronts, against SuperLU on the host."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    import torch
    from xyce_amd import circuits
    from xyce_amd._lib import check, lib
    from xyce_amd.loader import DeviceLoad, _ptr
    nl = circuits.random_logic_bsim3(n_gates=5000, seed=2).layout()
    dev = DeviceLoad(nl)
    info = np.zeros(6, np.int64)
    lib().xg_lu_front_info(dev.h, info.ctypes.data)
    print("fronts:", info, flush=True)
    assert info[0] == 1
    n1 = nl.n_nodes + 1
    rng = np.random.default_rng(21)
    x = rng.uniform(0.0, 1.2, n1)
    x[-1] = 0.0
    cu = lambda a: torch.tensor(a, dtype=torch.float64, device="cuda")
    bufs = [cu(a) for a in (x, np.zeros(nl.n_state), np.zeros(nl.n_store),
                            np.zeros(nl.n_state), np.zeros(nl.n_store))]
    outs = [cu(np.zeros(n1)) for _ in range(4)] + [cu(np.zeros(dev.nnz)) for _ in range(2)]
    dev.set_solver_state(transient=1, newtonIter=1)
    dev.load_all(*bufs, *outs)
    rhs, J = cu(np.zeros(n1)), cu(np.zeros(dev.nnz))
    qn0, fprev = cu(np.zeros(n1)), cu(np.zeros(n1))
    check(lib().xg_tia_assemble(dev.h, 1, 1, 1e-12, *[_ptr(t) for t in (
        qn0, fprev, *outs, rhs, J)]), "assemble")
    dev.sync()
    vals, r = J.cpu().numpy(), rhs.cpu().numpy()
    xs, st = _solve(dev, vals, r, torch)
    assert st[0, 2] == 0, st
    rp, ci = dev.pattern()
    n = nl.n_nodes
    rows = np.repeat(np.arange(n1), np.diff(rp))
    m = (rows < n) & (ci < n)
    A = sp.csc_matrix((vals[m], (rows[m], ci[m])), shape=(n, n))
    want = spla.spsolve(A, r[:n])
    berr = np.abs(A @ xs[:n] - r[:n]).max() / (abs(A).sum(1).max() * np.abs(xs[:n]).max() +
                                               np.abs(r[:n]).max())
    print("backward error %.3g, max |x - SuperLU| %.3g (scale %.3g)"
          % (berr, np.abs(xs[:n] - want).max(), np.abs(want).max()), flush=True)
    assert berr < 1e-13
    assert np.abs(xs[:n] - want).max() <= 1e-7 * np.abs(want).max()
    dev.close()


def _waveform_vs_oracle(nl_fn, tstop, dt, probes, vmax, uic):
    from oracle.transient_ref import RefTransient
    from xyce_amd.transient import Transient
    g = Transient(nl_fn(), lu_fronts=True)
    info = np.zeros(6, np.int64)
    from xyce_amd._lib import lib
    lib().xg_lu_front_info(g.dev.h, info.ctypes.data)
    assert info[0] >= 1
    tg, wg = g.run(tstop, dt, probes, every=5, uic=uic)
    g.close()
    r = RefTransient(nl_fn())
    tr_, wr = r.run(tstop, dt, probes, every=5, uic=uic)
    err = np.abs(wg - wr).max()
    print("fronts transient vs oracle: %.3g V over %d samples, Newton GPU %d oracle %d"
          % (err, wg.shape[1], g.newton_total, r.newton_total), flush=True)
    assert np.allclose(tg, tr_)
    assert err <= 1e-3 * vmax + 1e-6
    return wg


def test_random_logic_transient_on_fronts_matches_oracle():
    """A reduced random-logic circuit (150 gates, 600 BSIM3) from a zero
    start with the supply ramped, every Newton solve on dense fronts,
    against the oracle's transient (LAPACK solves)."""
    from xyce_amd import circuits
    w = _waveform_vs_oracle(lambda: circuits.random_logic_bsim3(n_gates=150, seed=8,
                                                                 ramp=50e-12),
                            0.25e-9, 2e-12, ["g0", "g1", "g2"], 1.2, True)
    assert np.ptp(w) > 0.2   # the gates move as the supply comes up


def test_power_grid_transient_on_fronts_matches_oracle():
    """A reduced power grid (20 x 20 mesh, 200 BSIM3 inverters with clock
    inputs) from its DC operating point, on dense fronts, against the
    oracle's transient."""
    from xyce_amd import circuits
    _waveform_vs_oracle(lambda: circuits.power_grid_bsim3(nx=20, ny=20, n_inv=200, seed=5),
                        0.3e-9, 2e-12, ["out0", "out77", "g10_10"], 1.2, False)


@pytest.mark.parametrize("gates,big_m", [(5000, 100000), (600, 48)])
def test_many_workgroup_fronts_give_the_same_bits(gates, big_m):
    """Fronts above the many-workgroup threshold (default 256 rows: the
    5,000-gate logic's top fronts) against the one-workgroup form of the
    same fronts (threshold raised above every front), and, on 600 gates,
    every front above 48 rows in the many-workgroup form against the
    default: identical solutions, bit for bit."""
    import torch
    from xyce_amd import circuits
    from xyce_amd._lib import check, lib
    from xyce_amd.loader import DeviceLoad, _ptr
    nl = circuits.random_logic_bsim3(n_gates=gates, seed=2).layout()
    print("case: %d gates, threshold %d" % (gates, big_m), flush=True)
    n1 = nl.n_nodes + 1
    rng = np.random.default_rng(5)
    x = rng.uniform(0.0, 1.2, n1)
    x[-1] = 0.0
    out = []
    for thr in (None, big_m):
        dev = DeviceLoad(nl, lu_fronts=True)
        info = np.zeros(6, np.int64)
        lib().xg_lu_front_info(dev.h, info.ctypes.data)
        if thr is not None:
            check(lib().xg_lu_front_set_big(dev.h, thr), "front threshold")
        cu = lambda a: torch.tensor(a, dtype=torch.float64, device="cuda")
        bufs = [cu(a) for a in (x, np.zeros(nl.n_state), np.zeros(nl.n_store),
                                np.zeros(nl.n_state), np.zeros(nl.n_store))]
        outs = [cu(np.zeros(n1)) for _ in range(4)] + [cu(np.zeros(dev.nnz)) for _ in range(2)]
        dev.set_solver_state(transient=1, newtonIter=1)
        dev.load_all(*bufs, *outs)
        rhs, J = cu(np.zeros(n1)), cu(np.zeros(dev.nnz))
        qn0, fprev = cu(np.zeros(n1)), cu(np.zeros(n1))
        check(lib().xg_tia_assemble(dev.h, 1, 1, 1e-12, *[_ptr(t) for t in (
            qn0, fprev, *outs, rhs, J)]), "assemble")
        dev.sync()
        xs, st = _solve(dev, J.cpu().numpy(), rhs.cpu().numpy(), torch)
        print("largest front %d rows, singular %s" % (info[4], st[:, 2].max()), flush=True)
        assert st[:, 2].max() == 0
        out.append(xs)
        dev.close()
    assert np.array_equal(out[0], out[1])


@pytest.mark.parametrize("big_m", [256, 48])
def test_fronts_pivot_on_random_values_without_fallback(big_m):
    """Random normal values on a 700-gate logic pattern (2,800 unknowns,
    beyond the dense fallback's 2048): the fronts must pivot at most
    columns, and the solve must stand on its own -- not singular, backward
    error at rounding level, equal to SuperLU -- in both the one-workgroup
    and the many-workgroup forms, with the same bits in both."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    import torch
    from xyce_amd import circuits
    from xyce_amd._lib import check, lib
    from xyce_amd.loader import DeviceLoad
    nl = circuits.random_logic_bsim3(n_gates=700, seed=9).layout()
    fr = DeviceLoad(nl, lu_fronts=True)
    check(lib().xg_lu_front_set_big(fr.h, big_m), "set big")
    rp, ci = fr.pattern()
    n = fr.n_rows - 1
    rng = np.random.default_rng(5)
    vals = rng.standard_normal(len(ci))
    rhs = rng.standard_normal(n + 1)
    rhs[n] = 0.0
    x, st = _solve(fr, vals, rhs, torch)
    comp = np.zeros(n, np.int32)
    nc = ctypes.c_int32()
    check(lib().xg_components(fr.h, ctypes.byref(nc), comp.ctypes.data), "components")
    big = int(np.argmax(np.bincount(comp)))
    assert np.bincount(comp)[big] > 2048
    assert st[big, 2] == 0, st[big]
    rows_of = np.repeat(np.arange(len(rp) - 1), np.diff(rp))
    m = (rows_of < n) & (ci < n)
    A = sp.csr_matrix((vals[m], (rows_of[m], ci[m])), shape=(n, n))
    idx = np.flatnonzero(comp == big)
    Ab = A[idx][:, idx]
    xb = x[idx]
    r = Ab @ xb - rhs[idx]
    berr = np.abs(r).max() / (abs(Ab).sum(1).max() * np.abs(xb).max() + np.abs(rhs[idx]).max())
    print("big_m %d: backward error %.3g" % (big_m, berr), flush=True)
    assert berr < 1e-12
    ref = spla.spsolve(Ab.tocsc(), rhs[idx])
    assert np.abs(xb - ref).max() <= 1e-6 * np.abs(ref).max()
    # the one-workgroup form gives the same bits
    check(lib().xg_lu_front_set_big(fr.h, 1 << 30), "set big")
    x1, _ = _solve(fr, vals, rhs, torch)
    assert x1.tobytes() == x.tobytes()
    fr.close()

