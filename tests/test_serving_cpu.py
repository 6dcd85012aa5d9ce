"""Host-side pieces of the serving loop: the seeded request stream and
the footprint estimate (no GPU)."""
import numpy as np

from xyce_amd.serving import footprint_bytes, poisson_requests


def test_stream_is_seeded_poisson_with_the_mix():
    a = poisson_requests(rate=200.0, n=2000, seed=4)
    b = poisson_requests(rate=200.0, n=2000, seed=4)
    assert [(r.kind, r.size, r.arrival) for r in a] == [(r.kind, r.size, r.arrival) for r in b]
    gaps = np.diff([0.0] + [r.arrival for r in a])
    assert abs(gaps.mean() - 1 / 200.0) < 0.1 / 200.0
    kinds = [r.kind for r in a]
    for k, p in (("ring", 0.6), ("logic", 0.3), ("ladder", 0.1)):
        assert abs(kinds.count(k) / len(kinds) - p) < 0.04
    rings = [r.size for r in a if r.kind == "ring"]
    assert min(rings) >= 11 and max(rings) <= 101


def test_footprint_grows_with_the_circuit():
    small = poisson_requests(n=50, seed=1)
    sizes = {}
    for r in small:
        sizes.setdefault(r.kind, []).append((r.size, footprint_bytes(r.netlist())))
    for kind, v in sizes.items():
        v.sort()
        assert all(b2 >= b1 for (_, b1), (_, b2) in zip(v, v[1:])), kind
