"""Batched Monte-Carlo transients: many independent circuits (samples of
uneven sizes) simulated as one batch.

The samples are merged into one netlist whose connected components are the
samples (node names prefixed "s<k>_").  Every Newton iteration then runs one
device-load launch per device type for all samples, one ordered reduction,
one assembly and one sparse-LU launch (a workgroup per sample), and each
sample converges on its own test (Transient masks converged samples out of
the solve), so each sample's waveform is bit for bit the one it gets when
simulated alone.  The batch shares one time grid.

`run_plain` is the one-circuit-per-launch form the batch is compared with:
the same code path, one sample at a time.
"""
import time

import numpy as np

from . import circuits
from .transient import Transient


class BatchedTransient:
    def __init__(self, parts, **kw):
        self.parts = list(parts)
        self.prefixes = ["s%d_" % k for k in range(len(self.parts))]
        t0 = time.perf_counter()
        self.tr = Transient(circuits.merge(self.parts, self.prefixes), **kw)
        self.setup_s = time.perf_counter() - t0

    def run(self, tstop, dt, probes, every=1):
        """Returns (times, waveforms[sample, probe, time])."""
        names = [pre + p for pre in self.prefixes for p in probes]
        t, w = self.tr.run(tstop, dt, names, every=every)
        return t, w.reshape(len(self.parts), len(probes), -1)


def run_batched(parts, tstop, dt, probes, every=1, sync=None):
    """One batch of every sample.  Returns dict(times, waves, seconds,
    setup_s, newton, step_failures, host_reads)."""
    b = BatchedTransient(parts)
    if sync:
        sync()
    t0 = time.perf_counter()
    t, w = b.run(tstop, dt, probes, every)
    if sync:
        sync()
    el = time.perf_counter() - t0
    return dict(times=t, waves=w, seconds=el, setup_s=b.setup_s,
                newton=b.tr.newton_total, step_failures=b.tr.step_failures,
                host_reads=b.tr.host_reads())


def run_plain(parts, tstop, dt, probes, every=1, sync=None):
    """Each sample alone, one after the other.  Returns dict(times, waves,
    seconds[per sample], newton[per sample])."""
    waves, secs, newton = [], [], []
    t = None
    for p in parts:
        tr = Transient(p)
        if sync:
            sync()
        t0 = time.perf_counter()
        t, w = tr.run(tstop, dt, probes, every)
        if sync:
            sync()
        secs.append(time.perf_counter() - t0)
        waves.append(w)
        newton.append(tr.newton_total)
    return dict(times=t, waves=np.array(waves), seconds=np.array(secs),
                newton=np.array(newton))
