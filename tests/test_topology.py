"""CSR pattern and stamp offsets, on the CPU.

1. The restated jacStampMap / build_pattern / matrixGlobalToLocal chain is
   pinned by what the operations mean, not by a second copy of them:
   - merging an internal node (d' onto d, s' onto s, as
     N_DEV_DeviceInstance.C:663-883 does when RD/RS are zero) must send every
     entry (r, c) of the full 6-node stamp, read through map/map2 the way
     registerJacLIDs reads it (N_DEV_MOSFET1.C:1650-1673), to the merged
     entry (merge(r), merge(node(r, c))), and the merged stamp must be
     exactly the set of merged entries;
   - the pattern equals the sorted union of all stamp entries, built from a
     dense 0/1 incidence by brute force (ground row/column last,
     N_LAS_Builder.C:336-338);
   - every offset addresses its own (row, column) in the CSR.
2. The load library's own pattern, offsets and per-instance setup (built by
   the host plan, no GPU involved) equal the restatement bit for bit.
"""
import numpy as np
import pytest

from oracle import stamp_ref
from xyce_amd import circuits


FULL = stamp_ref._MOS1_DC_SC


def _merge_fn(has_dc, has_sc):
    """Node renaming of the merged stamp: eliminated d'(4) -> d(0),
    s'(5) -> s(2); the surviving internal nodes are renumbered densely."""
    keep = [0, 1, 2, 3] + ([4] if has_dc else []) + ([5] if has_sc else [])
    ren = {n: i for i, n in enumerate(keep)}
    tgt = {4: 0, 5: 2}
    return lambda n: ren[n] if n in ren else ren[tgt[n]]


@pytest.mark.parametrize("has_dc,has_sc", [(True, True), (True, False), (False, True),
                                           (False, False)])
def test_jac_stamp_map_semantics(has_dc, has_sc):
    st, mp, m2 = stamp_ref.mos1_stamps()[(has_dc, has_sc)]
    merge = _merge_fn(has_dc, has_sc)
    expect = {}
    for r, row in enumerate(FULL):
        for c in row:
            expect.setdefault(merge(r), set()).add(merge(c))
    assert len(st) == len(expect)
    for r in range(len(st)):
        assert st[r] == sorted(expect[r]), (r, st[r], sorted(expect[r]))
    for (r, c) in stamp_ref.MOS1_OFFSET_ORDER:
        assert st[mp[r]][m2[r][c]] == merge(FULL[r][c])
        assert mp[r] == merge(r)


def test_bsim3_stamp_equals_mos1_stamp():
    """BSIM3 without NQS uses the MOSFET1 stamp shape and offset order
    (N_DEV_MOSFET_B3.C:2378-2418 builds the same 6-node stamp)."""
    from oracle import bsim3_bind
    nl = circuits.bsim3_variety(seed=4, n_dev=40)
    nl.layout()
    for (gids, st, mp, m2), rec in zip(bsim3_bind.stamp_rows(nl), nl.bsim3):
        has = (rec["lids"][4] >= 0, rec["lids"][5] >= 0)
        st1, mp1, m21 = stamp_ref.mos1_stamps()[has]
        assert st == st1 and mp == mp1 and m2 == m21


def _brute_pattern(n_rows, device_rows):
    A = np.zeros((n_rows + 1, n_rows + 1), bool)
    gl = lambda g: n_rows if g == -1 else g
    for gids, stamp in device_rows:
        for i, row in enumerate(stamp):
            for c in row:
                A[gl(gids[i]), gl(gids[c])] = True
    rowptr = np.concatenate([[0], np.cumsum(A.sum(axis=1))])
    colidx = np.concatenate([np.flatnonzero(A[r]) for r in range(n_rows + 1)])
    return rowptr, colidx


@pytest.mark.parametrize("seed", range(6))
def test_build_pattern_brute_force(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(1, 12))
    rows = []
    for _ in range(int(rng.integers(1, 15))):
        k = int(rng.integers(1, 5))
        gids = [int(g) for g in rng.integers(-1, n, k)]
        stamp = [sorted(set(int(c) for c in rng.integers(0, k, int(rng.integers(1, k + 1)))))
                 for _ in range(k)]
        rows.append((gids, stamp))
    rp, ci = stamp_ref.build_pattern(n, rows)
    brp, bci = _brute_pattern(n, rows)
    assert list(rp) == list(brp) and list(ci) == list(bci)
    rowmaps = {}
    for gids, stamp in rows:
        off = stamp_ref.stamp_offsets(rp, ci, n, gids, stamp, rowmaps)
        for i, row in enumerate(stamp):
            r = n if gids[i] == -1 else gids[i]
            for j, c in enumerate(row):
                col = n if gids[c] == -1 else gids[c]
                assert ci[rp[r] + off[i][j]] == col


def test_no_series_resistance_gives_dense_4x4():
    """RD = RS = 0: every FET contributes the dense d,g,s,b block (the
    closed form of the merged stamp)."""
    nl = circuits.Netlist("one FET")
    nl.add_model("N", "mos1", 1, VTO=0.5, KP=1e-4)
    nl.add_mos("M1", "a", "b", "c", "e", "N", W=1e-6, L=1e-6)
    from oracle.oracle import pattern_and_offsets
    rp, ci, off, _ = pattern_and_offsets(nl)
    assert list(rp) == [0, 4, 8, 12, 16, 16]
    assert list(ci) == [0, 1, 2, 3] * 4


# ------------------------------------------------------------ product plan
def _circuits():
    return {
        "mos1_ring": lambda: circuits.ring_oscillator_mos1(21, seed=1),
        "mos1_variety": lambda: circuits.mos1_variety(seed=2),
        "bsim3_variety": lambda: circuits.bsim3_variety(seed=1),
        "bsim3_logic": lambda: circuits.random_logic_bsim3(n_gates=300, seed=5),
        "bsim3_mc_ring": lambda: circuits.ring_oscillator_bsim3(15, seed=4, dvth0=0.03,
                                                                dtox=0.02),
        "mixed": lambda: _mixed(),
    }


def _mixed():
    nl = circuits.mos1_variety(seed=9, n_dev=30)
    v = circuits.bsim3_variety(seed=9, n_dev=30)
    for name, m in v.models.items():
        nl.add_model("B" + name, m.kind, m.dtype, **m.params)
    for m in v.mos:
        nl.add_mos("B" + m.name, m.d, m.g, m.s, m.b, "B" + m.model, **m.params)
    return nl


@pytest.mark.parametrize("name", list(_circuits()))
def test_plan_pattern_and_offsets_match_restatement(name):
    from oracle.oracle import pattern_and_offsets
    from xyce_amd.loader import DeviceLoad
    nl = _circuits()[name]()
    rp_r, ci_r, m1_r, b3_r = pattern_and_offsets(nl)
    dev = DeviceLoad(nl, plan_only=True)
    rp, ci = dev.pattern()
    assert np.array_equal(rp, rp_r) and np.array_equal(ci, ci_r)
    if len(nl.mos1):
        assert np.array_equal(dev.mos1_offsets(), m1_r)
    if len(nl.bsim3):
        assert np.array_equal(dev.bsim3_offsets(), b3_r)
    dev.close()


def test_plan_accepts_caller_pattern():
    """A pattern handed in by the caller (the reference's topology built it)
    is used as is; offsets are located inside it."""
    from oracle.oracle import pattern_and_offsets
    from xyce_amd.loader import DeviceLoad
    nl = circuits.bsim3_variety(seed=3, n_dev=40)
    rp_r, ci_r, _, b3_r = pattern_and_offsets(nl)
    dev = DeviceLoad(nl, plan_only=True, pattern=(rp_r, ci_r))
    assert np.array_equal(dev.bsim3_offsets(), b3_r)
    dev.close()
