import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the C ABI on cuda:0)")


# Parity rule for FP64 loads: entry-wise relative error 1e-12, measured
# against the larger magnitude of the pair, with a floor of 1e-3 x the
# largest magnitude of the reference array (so that entries produced by
# cancellation down to ~0 are judged on the array's scale).
RTOL = 1e-12


def assert_close(got, ref, rtol=RTOL, floor_frac=1e-3, what=""):
    got = np.asarray(got, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    assert got.shape == ref.shape, (what, got.shape, ref.shape)
    assert np.all(np.isfinite(got)) == np.all(np.isfinite(ref)), what
    scale = np.max(np.abs(ref)) if ref.size else 0.0
    denom = np.maximum(np.maximum(np.abs(got), np.abs(ref)), floor_frac * scale)
    err = np.abs(got - ref)
    bad = err > rtol * denom
    if np.any(bad):
        i = int(np.argmax(np.where(bad, err / np.where(denom > 0, denom, 1), 0)))
        raise AssertionError("%s: %d entries differ; worst at %d: got %r ref %r (rel %.3g)"
                             % (what, int(bad.sum()), i, got.flat[i], ref.flat[i],
                                err.flat[i] / max(denom.flat[i], 1e-300)))


def ground_masks(rowptr, colidx, n_nodes):
    """Boolean keep-masks (vector, CSR values) that drop the ground slot and
    the ground row and column: the reference writes them but N_LAS discards
    them before the linear solve (N_LAS_Builder.C:424-434), and the load
    library does not write them."""
    vkeep = np.ones(n_nodes + 1, bool)
    vkeep[n_nodes] = False
    rows = np.repeat(np.arange(n_nodes + 1), np.diff(rowptr))
    mkeep = (rows != n_nodes) & (np.asarray(colidx) != n_nodes)
    return vkeep, mkeep


def compare_loads(got, ref, rowptr, colidx, n_nodes):
    """got/ref = (F, Q, dFdxdVp, dQdxdVp, dFdx, dQdx)."""
    vkeep, mkeep = ground_masks(rowptr, colidx, n_nodes)
    names = ["F", "Q", "dFdxdVp", "dQdxdVp", "dFdx", "dQdx"]
    for k, (nm, a, b) in enumerate(zip(names, got, ref)):
        keep = vkeep if k < 4 else mkeep
        assert_close(np.asarray(a)[keep], np.asarray(b)[keep], what=nm)


@pytest.fixture(scope="session", autouse=True)
def built_library():
    """The load library is built in-tree; build it once if a fresh checkout
    runs the tests before any build step did."""
    from xyce_amd import build
    if not os.path.exists(build.LIB):
        build.build()
    return build.LIB


@pytest.fixture(scope="session")
def oracle_lib():
    from oracle import oracle
    oracle.build()
    return oracle.lib()
