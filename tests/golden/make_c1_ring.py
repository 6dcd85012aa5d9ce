"""Writes c1_ring101_20ns.npz: the oracle's transient of the 101-stage
MOSFET level-1 ring oscillator (seed 1, loads U[5, 15] fF, Vdd 1.8 V with a
50 ps supply ramp as the start kick), .TRAN 20 ns at fixed 1 ps steps, three
probes sampled every 10 ps, and the oracle's Newton iteration count.

The oracle is the CPU restatement of the reference (oracle/); this fixture
lets the GPU test compare the full-length waveform without the minute of
CPU time the oracle needs.  Run from the repository root:

    python tests/golden/make_c1_ring.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

PROBES = ["out0", "out50", "out100"]
TSTOP, DT, EVERY = 20e-9, 1e-12, 10


def main():
    from oracle.transient_ref import RefTransient
    from xyce_amd import circuits
    r = RefTransient(circuits.ring_oscillator_mos1(101, seed=1, ramp=50e-12))
    t, w = r.run(TSTOP, DT, PROBES, every=EVERY)
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "c1_ring101_20ns.npz")
    np.savez_compressed(out, t=t, w=w, probes=np.array(PROBES), newton=r.newton_total,
                        tstop=TSTOP, dt=DT, every=EVERY)
    print("wrote", out, w.shape, "Newton iterations", r.newton_total)


if __name__ == "__main__":
    main()
