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


def test_failing_request_is_evicted_without_disturbing_its_batch():
    """A request whose Newton solve cannot converge (a 0 -> 5 V source
    ramped in 10 ps straight across a junction diode) is evicted from its
    batch at the smallest step; the ring batched with it completes with the
    waveform it gets alone (to the Newton tolerance: the batch's steps were
    halved while the failing request was still in it)."""
    from xyce_amd import circuits
    from xyce_amd.circuits import Netlist, Pulse
    from xyce_amd.serving import Request, _Batch, _run_batch

    bad = Netlist("diode across a fast ramp")
    bad.add_model("DX", "diode", 1, IS=1e-14, N=1.0)
    bad.add_vsrc("V1", "a", "0", 0.0, Pulse(0.0, 5.0, 5e-12, 10e-12, 10e-12, 1.0, 2.0))
    bad.add_diode("D1", "a", "0", "DX")

    def req(rid, nl, probes):
        r = Request(rid, "ring", 11, 3, 0.0, 0.05e-9, 1e-12, probes)
        r._nl = nl
        return r

    ring = lambda: circuits.ring_oscillator_bsim3(11, seed=3, ramp=50e-12)
    alone = req(0, ring(), ["out0"])
    _run_batch(_Batch([alone]), None)
    good, evil = req(1, ring(), ["out0"]), req(2, bad, ["a"])
    _run_batch(_Batch([good, evil]), None)
    print("errors:", good.error, "|", evil.error)
    assert evil.error is not None and evil.error.startswith("evicted")
    assert good.error is None
    t_a, w_a = alone.result
    t_g, w_g = good.result
    # the halved steps sum to the same output times up to rounding
    assert t_a.shape == t_g.shape and np.abs(t_a - t_g).max() <= 1e-9 * 1e-12
    print("max difference to the run alone: %.3g V" % np.abs(w_g - w_a).max())
    assert np.abs(w_g - w_a).max() <= 1e-3 * 1.2
