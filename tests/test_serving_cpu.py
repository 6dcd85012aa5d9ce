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
    """Per kind, with the generator's seed fixed, a larger request has a
    larger footprint (random logic of other seeds can have fewer nodes at
    more gates, so sizes alone do not order footprints across seeds)."""
    from xyce_amd import circuits
    gens = {
        "ring": lambda k: circuits.ring_oscillator_bsim3(k, seed=3, ramp=50e-12),
        "logic": lambda k: circuits.random_logic_bsim3(n_gates=k, seed=3, ramp=50e-12),
        "ladder": lambda k: circuits.rc_ladder_diodes(k, max(1, k // 10), seed=3),
    }
    sizes = {"ring": (11, 41, 101), "logic": (125, 500, 1250), "ladder": (1000, 4000, 10000)}
    for kind, gen in gens.items():
        b = [footprint_bytes(gen(k)) for k in sizes[kind]]
        assert b[0] < b[1] < b[2], (kind, b)


def test_oversized_request_is_refused_and_serve_returns():
    """A request whose footprint exceeds the whole HBM budget is refused at
    arrival (error set, completion time recorded) and the loop ends; no
    batch is admitted, so no GPU is needed."""
    import threading
    from xyce_amd.serving import serve

    reqs = poisson_requests(n=3, seed=4)
    budget = min(footprint_bytes(r.netlist()) for r in reqs) - 1
    done = []
    th = threading.Thread(target=lambda: done.append(serve(reqs, hbm_budget=budget,
                                                           realtime=False)), daemon=True)
    th.start()
    th.join(timeout=30)
    assert not th.is_alive(), "serve() did not return"
    out = done[0]
    assert len(out) == 3
    for r in out:
        assert r.error and r.error.startswith("refused"), r.error
        assert r.t_done is not None and r.result is None


def test_probe_lookup_is_read_only_and_checked():
    """Probes resolve against the laid-out node table without adding
    nodes: unknown names raise KeyError and ground raises ValueError."""
    import pytest
    from xyce_amd import circuits
    nl = circuits.ring_oscillator_bsim3(5, seed=1).layout()
    n_before = len(nl.node_ids)
    assert nl.probe_id("out0") == nl.node_ids["out0"]
    with pytest.raises(KeyError):
        nl.probe_id("out_typo")
    with pytest.raises(ValueError):
        nl.probe_id("0")
    assert len(nl.node_ids) == n_before
