"""Dense-front (multifrontal) LU plan (xg_lu_front_export), checked on the
CPU: its structure (every position is a pivot of exactly one front, rows
ascending, contribution rows present in the parent, children below their
parent in the level order, matrix entries inside their front), then the
factorisation and solve carried out here front by front as the device does
them -- assemble the matrix entries, add the children's contribution blocks
in child order, partial LU of the fully summed columns with pivots chosen
among the fully summed rows, forward solve bottom-up passing each front's
right-hand-side remainder to its parent, backward solve top-down -- and
compared with LAPACK on random and diagonally weak systems."""
import numpy as np
import pytest

from xyce_amd import circuits
from xyce_amd.loader import DeviceLoad


def _cases():
    return {
        "mos1 ring": circuits.ring_oscillator_mos1(11, seed=1),
        "bsim3 ring": circuits.ring_oscillator_bsim3(21, seed=3),
        "random logic": circuits.random_logic_bsim3(n_gates=60, seed=5),
        "diode ladder": circuits.rc_ladder_diodes(120, 12, seed=2),
        "power grid": circuits.power_grid_bsim3(nx=12, ny=12, n_inv=40, seed=5),
        "batch": circuits.merge([circuits.ring_oscillator_mos1(7, seed=1),
                                 circuits.ring_oscillator_bsim3(9, seed=2)]),
    }


def _plan(nl):
    dev = DeviceLoad(nl, plan_only=True, lu_fronts=True)
    P, F = dev.lu_plan(), dev.front_plan()
    rp, ci = dev.pattern()
    n = dev.n_rows - 1
    dev.close()
    return P, F, rp, ci, n


def _check_structure(P, F, rp, ci, n):
    nf = len(F["m"])
    assert nf > 0 and np.all(F["front_comp"] == 1)
    owner = -np.ones(n, int)
    height_of = np.zeros(nf, int)
    for L in range(len(F["lv_ptr"]) - 1):
        for s in F["lv"][F["lv_ptr"][L]:F["lv_ptr"][L + 1]]:
            height_of[s] = L
    for s in range(nf):
        m, k = F["m"][s], F["k"][s]
        assert 1 <= k <= m and F["ro"][s + 1] - F["ro"][s] == m
        assert F["fo"][s + 1] - F["fo"][s] == m * m
        rows = F["rows"][F["ro"][s]:F["ro"][s + 1]]
        assert np.all(np.diff(rows) > 0) and rows[0] >= 0 and rows[-1] < n
        assert np.array_equal(rows[:k], rows[0] + np.arange(k))  # its columns
        assert np.all(owner[rows[:k]] < 0)
        owner[rows[:k]] = s
        for c in F["ch"][F["ch_ptr"][s]:F["ch_ptr"][s + 1]]:
            assert height_of[c] < height_of[s]
            crow = F["rows"][F["ro"][c]:F["ro"][c + 1]]
            cmap = F["cbmap"][F["ro"][c] + F["k"][c]:F["ro"][c + 1]]
            assert np.all((cmap >= 0) & (cmap < m))
            assert np.array_equal(rows[cmap], crow[F["k"][c]:])
        assert np.all(np.diff(F["ch"][F["ch_ptr"][s]:F["ch_ptr"][s + 1]]) > 0)
    assert np.all(owner >= 0)
    # every matrix entry outside ground maps into its front, once
    fpos = F["ent_fpos"]
    assert len(np.unique(fpos)) == len(fpos)
    assert np.all((fpos >= 0) & (fpos < F["fo"][-1]))
    rows_of = np.repeat(np.arange(len(rp) - 1), np.diff(rp))
    keep = (rows_of < n) & (ci < n)
    assert np.array_equal(np.sort(F["ent_k"]), np.flatnonzero(keep))
    return owner


def _solve_fronts(P, F, vals, rhs, n):
    """The device's algorithm in numpy; returns (dx, pivot failure)."""
    nf = len(F["m"])
    fronts = [np.zeros((F["m"][s], F["m"][s])) for s in range(nf)]
    starts = F["fo"][:-1]
    sof = np.searchsorted(starts, F["ent_fpos"], side="right") - 1
    for e, k in enumerate(F["ent_k"]):
        s = sof[e]
        off = F["ent_fpos"][e] - F["fo"][s]
        fronts[s][off // F["m"][s], off % F["m"][s]] = vals[k]
    perms = [None] * nf
    order = F["lv"]
    for s in order:  # bottom-up: children in earlier levels
        m, k = F["m"][s], F["k"][s]
        A = fronts[s]
        for c in F["ch"][F["ch_ptr"][s]:F["ch_ptr"][s + 1]]:
            kc = F["k"][c]
            cmap = F["cbmap"][F["ro"][c] + kc:F["ro"][c + 1]]
            A[np.ix_(cmap, cmap)] += fronts[c][kc:, kc:]
        perm = np.arange(k)
        for j in range(k):
            p = j + int(np.argmax(np.abs(A[j:k, j])))
            if not np.abs(A[p, j]) > 0:
                return None, True
            if p != j:
                A[[j, p], :] = A[[p, j], :]
            perm[j] = p
            A[j + 1:, j] /= A[j, j]
            A[j + 1:, j + 1:] -= np.outer(A[j + 1:, j], A[j, j + 1:])
        perms[s] = perm
    bp = rhs[P["prow"]]
    v = np.zeros(n)
    wres = [None] * nf
    for s in order:
        m, k = F["m"][s], F["k"][s]
        rows = F["rows"][F["ro"][s]:F["ro"][s + 1]]
        w = np.zeros(m)
        w[:k] = bp[rows[:k]]
        for c in F["ch"][F["ch_ptr"][s]:F["ch_ptr"][s + 1]]:
            kc = F["k"][c]
            cmap = F["cbmap"][F["ro"][c] + kc:F["ro"][c + 1]]
            w[cmap] += wres[c]
        for j in range(k):
            p = perms[s][j]
            w[j], w[p] = w[p], w[j]
        A = fronts[s]
        for j in range(k):
            w[j + 1:] -= A[j + 1:, j] * w[j]
        v[rows[:k]] = w[:k]
        wres[s] = w[k:].copy()
    for s in order[::-1]:
        m, k = F["m"][s], F["k"][s]
        rows = F["rows"][F["ro"][s]:F["ro"][s + 1]]
        A = fronts[s]
        t = v[rows[:k]] - A[:k, k:] @ v[rows[k:]]
        for j in range(k - 1, -1, -1):
            t[j] /= A[j, j]
            t[:j] -= A[:j, j] * t[j]
        v[rows[:k]] = t
    dx = np.zeros(n + 1)
    dx[P["qcol"]] = v
    return dx, False


def _dense(rp, ci, vals, n):
    A = np.zeros((n, n))
    for r in range(n):
        for k in range(rp[r], rp[r + 1]):
            if ci[k] < n:
                A[r, ci[k]] += vals[k]
    return A


@pytest.mark.parametrize("name", list(_cases()))
def test_front_plan_structure_and_solve(name):
    nl = _cases()[name]
    P, F, rp, ci, n = _plan(nl)
    _check_structure(P, F, rp, ci, n)
    rng = np.random.default_rng(7)
    for trial in range(3):
        vals = rng.standard_normal(len(ci))
        if trial == 0:   # diagonally dominant: static pivots would do too
            rows_of = np.repeat(np.arange(len(rp) - 1), np.diff(rp))
            vals[rows_of == ci] += 20.0
        rhs = rng.standard_normal(n + 1)
        dx, failed = _solve_fronts(P, F, vals, rhs, n)
        A = _dense(rp, ci, vals, n)
        if failed:
            assert np.linalg.matrix_rank(A) < n
            continue
        want = np.linalg.solve(A, rhs[:n])
        err = np.abs(dx[:n] - want).max() / max(np.abs(want).max(), 1e-300)
        assert err < 1e-8, (name, trial, err)
        res = np.abs(A @ dx[:n] - rhs[:n]).max() / (np.abs(A).sum(1).max() *
                                                     np.abs(dx[:n]).max() + np.abs(rhs[:n]).max())
        assert res < 1e-13, (name, trial, res)


def test_large_circuits_choose_fronts_automatically():
    """Beyond the task lists' budgets a component gets dense fronts without
    being asked (the 5,000-gate logic), and its largest front stays within
    what one workgroup's panel buffer holds."""
    import ctypes
    from xyce_amd._lib import lib
    nl = circuits.random_logic_bsim3(n_gates=5000, seed=2).layout()
    dev = DeviceLoad(nl, plan_only=True)
    info = np.zeros(6, np.int64)
    lib().xg_lu_front_info(dev.h, info.ctypes.data)
    dev.close()
    assert info[0] == 1 and info[1] > 1000 and 0 < info[4] <= 3481
    assert info[5] < 5e9  # minimum-degree order: a few GFLOP per factorisation
