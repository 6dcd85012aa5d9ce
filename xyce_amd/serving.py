"""Serving transient-simulation requests with continuous batching.

Requests (a netlist generator with its size and seed, a stop time and a
step) arrive over time.  The server admits them into device-resident
batches as memory and worker slots free up:

  * a dispatcher queues each request at its arrival time;
  * an admission step, every `window` seconds, takes the queued requests
    of one kind (same stop time and step) into a micro-batch, up to
    `max_batch` requests and within the HBM budget still free (footprints
    estimated from the netlist sizes);
  * a worker builds the batch (the requests merged into one netlist whose
    independent sub-circuits are the requests), runs it with the native
    transient driver on its own HIP stream -- its state vectors and LU
    factors stay on the device for the whole run -- and completes each
    request with its waveforms;
  * up to `workers` batches are on the GPU at once, so batches that arrive
    while others run overlap with them on the device.

Inside a batch every kernel runs once per Newton iteration for all its
requests, and each request converges on its own test, so a request's
waveform is the one it gets alone (bit for bit unless a step of the batch
had to be halved for another request, then within the Newton tolerance).

`serve(requests, max_batch=1, workers=1)` is the plain one-circuit-per-
launch form the batched server is compared with.
"""
import queue
import threading
import time
from dataclasses import dataclass

import numpy as np

from . import circuits


@dataclass
class Request:
    rid: int
    kind: str            # "ring", "logic" or "ladder"
    size: int            # stages, gates or ladder nodes
    seed: int
    arrival: float       # seconds after the stream starts
    tstop: float
    dt: float
    probes: list
    result: object = None
    t_admit: float = None
    t_done: float = None
    error: str = None

    def netlist(self):
        """The request's circuit (built once)."""
        if getattr(self, "_nl", None) is None:
            self._nl = self._build()
        return self._nl

    def _build(self):
        if self.kind == "ring":
            return circuits.ring_oscillator_bsim3(self.size, seed=self.seed, ramp=50e-12)
        if self.kind == "logic":
            return circuits.random_logic_bsim3(n_gates=self.size, seed=self.seed, ramp=50e-12)
        if self.kind == "ladder":
            return circuits.rc_ladder_diodes(self.size, max(1, self.size // 10),
                                             seed=self.seed)
        raise ValueError(self.kind)

    @property
    def latency(self):
        return None if self.t_done is None else self.t_done - self.arrival


# Transient spans per request kind (seconds) and the probes each returns.
SPANS = {"ring": (0.1e-9, 1e-12, ["out0"]),
         "logic": (0.05e-9, 1e-12, ["g0"]),
         "ladder": (0.05e-9, 1e-12, ["n0", "n1"])}
# Random logic starts from zero (.TRAN UIC) with its supply ramping up: its
# feedback paths can leave the plain DC Newton solve without convergence,
# and the gmin-stepping operating point (Transient.dcop("gmin")) costs
# more solves than the ramp.
UIC = {"ring": False, "logic": True, "ladder": False}


def poisson_requests(rate=200.0, n=200, seed=4, mix=(0.6, 0.3, 0.1),
                     ring_stages=(11, 101), logic_gates=(125, 1250), ladder_nodes=10000):
    """A seeded Poisson stream: rings of U{ring_stages} BSIM3 stages,
    random BSIM3 logic of U{logic_gates} gates (4 FETs each), RC ladders
    of ladder_nodes nodes with one diode per ten nodes."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, n))
    kinds = rng.choice(["ring", "logic", "ladder"], n, p=list(mix))
    out = []
    for k in range(n):
        kind = str(kinds[k])
        if kind == "ring":
            size = int(rng.integers(ring_stages[0], ring_stages[1] + 1))
        elif kind == "logic":
            size = int(rng.integers(logic_gates[0], logic_gates[1] + 1))
        else:
            size = ladder_nodes
        tstop, dt, probes = SPANS[kind]
        out.append(Request(k, kind, size, int(rng.integers(1 << 30)), float(t[k]), tstop, dt,
                           list(probes)))
    return out


def footprint_bytes(nl):
    """Estimated device bytes of one request: statement slots, loads and
    Jacobians, LU factors and vectors (per device and per unknown)."""
    nl.layout()
    n_fet = len(nl.mos)
    n_lin = len(nl.res) + len(nl.caps) + len(nl.vsrc) + len(nl.diodes)
    return 1400 * n_fet + 320 * n_lin + 900 * nl.n_nodes + (1 << 20)


class _Batch:
    def __init__(self, reqs):
        self.reqs = reqs
        self.bytes = 0


def _run_batch(batch, sync):
    from .transient import DEVICE_LOCK, Transient
    reqs = batch.reqs
    merged = circuits.merge([r.netlist() for r in reqs])
    tstop, dt = reqs[0].tstop, reqs[0].dt
    names = []
    for k, r in enumerate(reqs):
        names += ["s%d_%s" % (k, p) for p in r.probes]
    t0 = time.perf_counter()
    tr = Transient(merged)
    t1 = time.perf_counter()
    try:
        # a request whose Newton solve fails at dt / 1024 leaves the batch
        # (its sub-circuits frozen); the others run on
        tr.set_eviction(True, dt / 1024)
        t, w = tr.run(tstop, dt, names, every=10, uic=UIC[reqs[0].kind])
        evicted = set(tr.evicted_components())
        gone = set()
        if evicted:
            import ctypes
            from ._lib import lib
            nc = ctypes.c_int32()
            comp = np.zeros(max(merged.n_nodes, 1), np.int32)
            lib().xg_components(tr.dev.h, ctypes.byref(nc), comp.ctypes.data)
            for name, nid in merged.node_ids.items():
                if 0 <= nid < merged.n_nodes and int(comp[nid]) in evicted:
                    gone.add(int(name.split("_", 1)[0][1:]))
        for r in reqs:
            r.batch_info = dict(size=len(reqs), step_failures=tr.step_failures,
                                newton=tr.newton_total, setup_s=t1 - t0,
                                run_s=time.perf_counter() - t1)
    finally:
        with DEVICE_LOCK:
            tr.close()
            tr.dev.close()
    for r in reqs:
        r._nl = None
    j = 0
    for k, r in enumerate(reqs):
        m = len(r.probes)
        if k in gone:
            r.error = "evicted: its Newton solve failed at the smallest step"
        else:
            r.result = (t, w[j:j + m])
        j += m


def serve(requests, max_batch=32, workers=3, window=0.005, hbm_budget=64 << 30,
          realtime=True, log=None):
    """Runs the request stream; returns the requests with results,
    admission and completion times (seconds on the stream's clock)."""
    runtime = []

    def start_runtime():
        # the HIP runtime starts on this thread, not in a worker, and only
        # once a batch is admitted
        if not runtime:
            import torch
            torch.cuda.init()
            from . import _lib
            _lib.lib()
            runtime.append(True)

    sync = None
    pending = {}             # kind -> list of arrived requests
    lock = threading.Lock()
    work = queue.Queue()
    state = dict(free=hbm_budget, busy=0, done=0)
    reqs = sorted(requests, key=lambda r: r.arrival)
    # the clients' netlists (what a request carries) are ready before the
    # stream starts; the library setup, upload and run count in latency
    for r in reqs:
        r._bytes = footprint_bytes(r.netlist())
    t0 = time.perf_counter()
    now = lambda: time.perf_counter() - t0

    def worker():
        while True:
            b = work.get()
            if b is None:
                return
            try:
                _run_batch(b, sync)
            except Exception as e:  # a failed batch fails its requests
                for r in b.reqs:
                    r.error = repr(e)
            tdone = now()
            with lock:
                for r in b.reqs:
                    r.t_done = tdone
                state["free"] += b.bytes
                state["busy"] -= 1
                state["done"] += len(b.reqs)

    threads = [threading.Thread(target=worker, daemon=True) for _ in range(workers)]
    for th in threads:
        th.start()
    i = 0
    n = len(reqs)
    while True:
        t = now()
        # arrivals; a request larger than the whole budget can never be
        # admitted: it is refused at once instead of waiting forever
        while i < n and (reqs[i].arrival <= t or not realtime):
            r = reqs[i]
            i += 1
            if r._bytes > hbm_budget:
                r.error = ("refused: footprint %d bytes exceeds the HBM budget of %d bytes"
                           % (r._bytes, hbm_budget))
                with lock:
                    r.t_done = now()
                    state["done"] += 1
                if log:
                    log("t=%.3f refused request %d (%.1f MB > budget)" % (r.t_done, r.rid,
                                                                       r._bytes / 1e6))
                continue
            pending.setdefault(r.kind, []).append(r)
        # admission: the oldest kind first, as slots and memory allow
        with lock:
            kinds = sorted((q[0].arrival, k) for k, q in pending.items() if q)
            for _, kind in kinds:
                if state["busy"] >= workers:
                    break
                q = pending[kind]
                take, need = [], 0
                while q and len(take) < max_batch and need + q[0]._bytes <= state["free"]:
                    r = q.pop(0)
                    take.append(r)
                    need += r._bytes
                if not take:
                    continue
                start_runtime()
                b = _Batch(take)
                b.bytes = need
                state["free"] -= need
                state["busy"] += 1
                ta = now()
                for r in take:
                    r.t_admit = ta
                work.put(b)
                if log:
                    log("t=%.3f admitted %d %s requests (%.1f MB)" % (ta, len(take), kind,
                                                                    need / 1e6))
            finished = state["done"] >= n
        if finished:
            break
        time.sleep(window)
    for _ in threads:
        work.put(None)
    for th in threads:
        th.join()
    return reqs


def summary(reqs):
    lat = np.array([r.latency for r in reqs if r.error is None])
    done = [r for r in reqs if r.error is None]
    span = max(r.t_done for r in done) - min(r.arrival for r in reqs)
    out = dict(requests=len(reqs), failed=sum(r.error is not None for r in reqs),
               solves_per_s=len(done) / span, latency_p50_s=float(np.percentile(lat, 50)),
               latency_p99_s=float(np.percentile(lat, 99)), span_s=span)
    # per kind: batches, mean batch size, mean setup and run seconds per batch
    for kind in sorted({r.kind for r in done}):
        batches = {}
        for r in done:
            if r.kind == kind and getattr(r, "batch_info", None):
                batches[r.t_admit] = r.batch_info
        if batches:
            b = list(batches.values())
            out[kind] = dict(batches=len(b), mean_size=float(np.mean([x["size"] for x in b])),
                             setup_s=float(np.mean([x["setup_s"] for x in b])),
                             run_s=float(np.mean([x["run_s"] for x in b])))
    return out
