"""The serving loop on the GPU: a Poisson stream of mixed requests (BSIM3
rings, random BSIM3 logic, RC ladders with diodes) admitted into batches as
they arrive, several batches in flight; every request completes, requests
share batches, and each result equals the request simulated alone."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_mixed_stream_batches_and_matches_alone():
    from xyce_amd.serving import poisson_requests, serve, summary
    from xyce_amd.transient import Transient
    reqs = poisson_requests(rate=400.0, n=24, seed=11, ring_stages=(11, 21),
                            logic_gates=(20, 40), ladder_nodes=200)
    done = serve(reqs, max_batch=8, workers=3, window=0.002)
    s = summary(done)
    print("serving:", s, flush=True)
    assert s["failed"] == 0, [r.error for r in done if r.error][:2]
    admits = {}
    for r in done:
        admits.setdefault(r.t_admit, []).append(r)
    assert max(len(v) for v in admits.values()) > 1       # requests shared batches
    checked = set()
    for r in done:
        if r.kind in checked:
            continue
        checked.add(r.kind)
        from xyce_amd.serving import UIC
        alone = Transient(r._build())
        t, w = alone.run(r.tstop, r.dt, r.probes, every=10, uic=UIC[r.kind])
        alone.close()
        tb, wb = r.result
        print("%s: batch %s, alone %d step failures, max difference %.3g V"
              % (r.kind, r.batch_info, alone.step_failures, np.abs(w - wb).max()), flush=True)
        assert np.array_equal(t, tb)
        if r.batch_info["step_failures"] == 0 and alone.step_failures == 0:
            # no step of either run was halved: the same bits
            assert np.array_equal(w, wb), r.kind
        else:
            vmax = 1.2 if r.kind != "ladder" else 1.0
            assert np.abs(w - wb).max() <= 1e-3 * vmax + 1e-6, r.kind
    assert checked == {"ring", "logic", "ladder"}
