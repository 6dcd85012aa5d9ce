"""Static-pivot sparse LU plan (xg_lu_export), checked on the CPU.

The task lists are executed here level by level exactly as the device
kernel runs them (one value per task, terms in list order), with every
index bounds-checked and every read checked to come from an earlier level.
The solutions are compared with LAPACK (numpy.linalg.solve) on random
diagonally dominant values and on real Newton Jacobians of the oracle's
transients; the acceptance test (finite factors, nonzero pivots, normwise
backward error <= 1e-12) is evaluated as the device evaluates it."""
import numpy as np
import pytest

from xyce_amd import circuits
from xyce_amd.loader import DeviceLoad


FWD = -3   # t_div of a forward-substitution row scheduled in the factor phase


def _comp_starts(dev):
    import ctypes
    from xyce_amd._lib import check, lib
    n = dev.n_rows - 1
    nc = ctypes.c_int32()
    comp = np.zeros(max(n, 1), np.int32)
    check(lib().xg_components(dev.h, ctypes.byref(nc), comp.ctypes.data), "components")
    cstart = np.concatenate([[0], np.cumsum(np.bincount(comp[:n], minlength=nc.value))])
    return comp[:n], cstart


def run_plan(dev, vals, rhs):
    """Executes the plan; returns (x, components whose static pivots the
    device would reject)."""
    P = dev.lu_plan()
    rp, ci = dev.pattern()
    n = dev.n_rows - 1
    comp, cstart = _comp_starts(dev)
    x = np.zeros(n + 1)
    failed = []
    for c in range(len(cstart) - 1):
        beg, m = cstart[c], cstart[c + 1] - cstart[c]
        assert P["ok"][c] == 1
        nv = int(P["voff"][c + 1] - P["voff"][c])
        prow, qcol = P["prow"][beg:beg + m], P["qcol"][beg:beg + m]
        members = np.flatnonzero(comp == c)
        assert np.array_equal(np.sort(prow), members)
        assert np.array_equal(np.sort(qcol), members)
        val = np.zeros(nv)
        for g in prow:
            for k in range(rp[g], rp[g + 1]):
                p = P["lu_pos"][k]
                if p >= 0:
                    assert p < nv
                    val[p] = vals[k]
        vec = rhs[prow].astype(float)
        lv = P["lv_comp"][3 * c:3 * c + 4]
        bad = False
        for ph in range(3):
            for L in range(lv[ph], lv[ph + 1]):
                # values written in this level, factors and vector apart: no
                # task of a level may read what another writes in it
                new = {"val": {}, "vec": {}}
                for k in range(P["lv_ptr"][L], P["lv_ptr"][L + 1]):
                    d, dv = P["t_dst"][k], P["t_div"][k]
                    # a forward-substitution row (t_div -3) runs among the
                    # factor tasks and works on the vector
                    on_vec = ph > 0 or dv == FWD
                    src, lim, key = (vec, m, "vec") if on_vec else (val, nv, "val")
                    assert 0 <= d < lim and d not in new[key]
                    s = src[d]
                    for o in range(P["t_op"][k], P["t_op"][k + 1]):
                        a, b = P["ops"][o]
                        assert 0 <= a < nv and 0 <= b < lim
                        assert b not in new[key] and a not in new["val"]
                        s = s - val[a] * src[b]
                    if ph == 0 and dv >= 0:
                        assert dv < nv and dv not in new["val"]
                        s = s / val[dv]
                        bad |= not np.isfinite(s)
                    elif ph == 0 and dv == -2:
                        bad |= not (abs(s) > 0.0 and np.isfinite(s))
                    elif ph == 2:
                        s = s / val[dv]
                    new[key][d] = s
                for d, s in new["val"].items():
                    val[d] = s
                for d, s in new["vec"].items():
                    vec[d] = s
        # backward error with the unfactored rows
        qpos = {int(g): i for i, g in enumerate(qcol)}
        rm = am = 0.0
        for g in prow:
            r, asum = rhs[g], 0.0
            for k in range(rp[g], rp[g + 1]):
                if ci[k] in qpos:
                    r = r - vals[k] * vec[qpos[ci[k]]]
                    asum += abs(vals[k])
            rm, am = max(rm, abs(r)), max(am, asum)
        bmax = np.abs(rhs[prow]).max() if m else 0.0
        xmax = np.abs(vec).max() if m else 0.0
        bad |= not rm <= 1e-12 * (am * xmax + bmax)
        if bad:
            failed.append(c)
        x[qcol] = vec
    return x, failed


def _dense(dev, vals, idx):
    rp, ci = dev.pattern()
    n = dev.n_rows - 1
    rows = np.repeat(np.arange(n + 1), np.diff(rp))
    loc = -np.ones(n + 1, int)
    loc[idx] = np.arange(len(idx))
    A = np.zeros((len(idx), len(idx)))
    m = (loc[rows] >= 0) & (loc[ci] >= 0)
    A[loc[rows[m]], loc[ci[m]]] = vals[m]
    return A


def _rc_ladder():
    nl = circuits.Netlist("rc ladder")
    nl.add_vsrc("V1", "n0", "0", 1.0)
    for k in range(40):
        nl.add_res("R%d" % k, "n%d" % k, "n%d" % (k + 1), 100.0 + k)
        nl.add_cap("C%d" % k, "n%d" % (k + 1), "0", 1e-15)
    return nl


CASES = [
    lambda: circuits.ring_oscillator_mos1(11, seed=1),
    lambda: circuits.ring_oscillator_mos1(101, seed=1),
    lambda: circuits.ring_oscillator_bsim3(301, seed=3),
    lambda: circuits.merge(circuits.monte_carlo_rings(4, seed=9, stages=(11, 41))),
    _rc_ladder,
    # random logic (its components take the minimum-degree order)
    lambda: circuits.random_logic_bsim3(n_gates=150, seed=4),
]


@pytest.mark.parametrize("case", range(len(CASES)))
def test_plan_solves_like_lapack(case):
    nl = CASES[case]()
    nl.layout()
    dev = DeviceLoad(nl, plan_only=True, lu_mode="tasks")
    rp, ci = dev.pattern()
    n = dev.n_rows - 1
    rng = np.random.default_rng(case)
    vals = rng.uniform(-1, 1, len(ci))
    rows = np.repeat(np.arange(n + 1), np.diff(rp))
    vals[rows == ci] += 20.0
    # a source's branch row has no diagonal: keep its +-1 entries dominant
    rhs = rng.uniform(-1, 1, n + 1)
    rhs[n] = 0.0
    x, failed = run_plan(dev, vals, rhs)
    comp, cstart = _comp_starts(dev)
    info = dev.lu_info()
    assert info["static_fail"] == 0 and info["fallbacks"] == -1
    for c in range(len(cstart) - 1):
        if c in failed:
            continue
        idx = np.flatnonzero(comp == c)
        ref = np.linalg.solve(_dense(dev, vals, idx), rhs[idx])
        assert np.allclose(x[idx], ref, rtol=1e-11, atol=1e-13 * np.abs(ref).max())


def test_levels_stay_shallow():
    """Nested dissection keeps a 301-stage ring's factorisation to a few
    dozen dependent levels (the natural order would need ~300)."""
    nl = circuits.ring_oscillator_bsim3(301, seed=3)
    nl.layout()
    info = DeviceLoad(nl, plan_only=True, lu_mode="tasks").lu_info()
    assert info["max_factor_levels"] <= 48
    assert info["fnnz"] < 12 * 303


def test_same_plan_alone_and_in_batch():
    parts = circuits.monte_carlo_rings(3, seed=5, stages=(11, 31))
    batch = circuits.merge(parts)
    batch.layout()
    B = DeviceLoad(batch, plan_only=True, lu_mode="tasks")
    PB = B.lu_plan()
    _, cs = _comp_starts(B)
    for k in range(3):
        p = circuits.monte_carlo_rings(3, seed=5, stages=(11, 31))[k]
        p.layout()
        A = DeviceLoad(p, plan_only=True, lu_mode="tasks")
        PA = A.lu_plan()
        la = PA["lv_comp"]
        lb = PB["lv_comp"][3 * k:3 * k + 4]
        assert np.array_equal(np.diff(la), np.diff(lb))
        ta = slice(PA["lv_ptr"][la[0]], PA["lv_ptr"][la[3]])
        tb = slice(PB["lv_ptr"][lb[0]], PB["lv_ptr"][lb[3]])
        assert np.array_equal(PA["t_dst"][ta], PB["t_dst"][tb])
        assert np.array_equal(PA["t_div"][ta], PB["t_div"][tb])
        oa = PA["ops"][PA["t_op"][ta.start]:PA["t_op"][ta.stop]]
        ob = PB["ops"][PB["t_op"][tb.start]:PB["t_op"][tb.stop]]
        assert np.array_equal(oa, ob)


@pytest.mark.parametrize("which", ["mos1", "bsim3"])
def test_real_jacobians(which):
    """Every Newton Jacobian of a short ring transient (DC operating point
    included): the static pivots either pass the backward-error test and
    solve like LAPACK (to the conditioning of the matrix), or fail it (then
    the device refactors with pivoting)."""
    from oracle.transient_ref import RefTransient
    mk = {"mos1": lambda: circuits.ring_oscillator_mos1(11, seed=1, ramp=50e-12),
          "bsim3": lambda: circuits.ring_oscillator_bsim3(11, seed=3, ramp=50e-12)}[which]
    r = RefTransient(mk())
    mats = []
    orig = r._assemble

    def hook(transient, order, dt):
        orig(transient, order, dt)
        mats.append((r.J.copy(), r.rhs.copy()))
    r._assemble = hook
    r.run(0.1e-9, 1e-12, ["out0"], every=50)
    nl = mk()
    nl.layout()
    dev = DeviceLoad(nl, plan_only=True, lu_mode="tasks")
    n = dev.n_rows - 1
    idx = np.arange(n)
    passed = 0
    for J, rhs in mats[::5]:
        x, failed = run_plan(dev, J, rhs)
        if failed:
            continue
        passed += 1
        A = _dense(dev, J, idx)
        ref = np.linalg.solve(A, rhs[idx])
        tol = max(1e-12, 1e-15 * np.linalg.cond(A))
        assert np.max(np.abs(x[idx] - ref)) <= tol * np.abs(ref).max()
    assert passed >= 0.8 * len(mats[::5])


def test_automatic_form_keeps_task_lists_in_lds():
    """By default a component takes task lists only when they fit in one
    workgroup's LDS; a 400-gate random-logic request (its task lists would
    need ~1 MB) is factored by dense fronts, a 301-stage ring by task
    lists."""
    nl = circuits.random_logic_bsim3(n_gates=400, seed=7)
    nl.layout()
    dev = DeviceLoad(nl, plan_only=True)
    info = dev.lu_info()
    fp = dev.front_plan()
    # the request's one large component is on fronts; only its few small
    # components (sources with their own nodes) keep task lists
    assert info["fnnz"] < 1000
    assert int(fp["front_comp"].sum()) == 1 and int(fp["m"].max()) > 100
    ring = circuits.ring_oscillator_bsim3(301, seed=3)
    ring.layout()
    info = DeviceLoad(ring, plan_only=True).lu_info()
    assert info["fnnz"] > 0 and info["global_components"] == 0


def test_random_logic_takes_the_shorter_chain():
    """Random logic is expander-like: the minimum-degree order, chosen where
    its elimination chain is shorter than the dissection's (about a third
    of the depth and a quarter of the fill on such requests), keeps a
    400-gate request's factorisation under 400 levels and 40k entries."""
    nl = circuits.random_logic_bsim3(n_gates=400, seed=7)
    nl.layout()
    info = DeviceLoad(nl, plan_only=True, lu_mode="tasks").lu_info()
    print(info)
    assert info["static_fail"] == 0
    assert info["max_factor_levels"] < 400
    assert info["fnnz"] < 40000


def _cache_stats():
    from xyce_amd._lib import check, lib
    st = np.zeros(3, np.int64)
    check(lib().xg_lu_cache_stats(st.ctypes.data), "cache stats")
    return st


def test_plan_cache_reuses_analysed_components():
    """A component whose structure was analysed before (a repeated ring in
    the same batch, or the same circuit again) takes the stored analysis:
    counted as a hit, and the plan -- orderings, positions, task lists,
    fronts -- is the one a fresh analysis gives."""
    from xyce_amd._lib import check, lib
    L = lib()
    mk = lambda: circuits.merge([circuits.ring_oscillator_bsim3(s, seed=3)
                                 for s in (11, 31, 11, 31, 51)])
    logic = lambda: circuits.random_logic_bsim3(n_gates=400, seed=7)
    try:
        check(L.xg_lu_cache_set_limit(0), "cache off")
        ref = DeviceLoad(mk(), plan_only=True, lu_mode="tasks").lu_plan()
        ref_fr = DeviceLoad(logic(), plan_only=True).front_plan()
        check(L.xg_lu_cache_set_limit(64 << 20), "cache on")
        a = DeviceLoad(mk(), plan_only=True, lu_mode="tasks").lu_plan()
        h1, m1, _ = _cache_stats()
        assert h1 >= 2 and m1 >= 3, (h1, m1)   # the second 11- and 31-stage rings
        b = DeviceLoad(mk(), plan_only=True, lu_mode="tasks").lu_plan()
        h2, m2, held = _cache_stats()
        assert m2 == m1 and h2 == h1 + h1 + m1 and held > 0
        for plan in (a, b):
            for k in ref:
                assert np.array_equal(ref[k], plan[k]), k
        f1 = DeviceLoad(logic(), plan_only=True).front_plan()
        h3, m3, _ = _cache_stats()
        f2 = DeviceLoad(logic(), plan_only=True).front_plan()
        h4, m4, _ = _cache_stats()
        assert m4 == m3 and h4 > h3
        for plan in (f1, f2):
            for k in ref_fr:
                assert np.array_equal(ref_fr[k], plan[k]), k
    finally:
        check(L.xg_lu_cache_set_limit(256 << 20), "cache default")
