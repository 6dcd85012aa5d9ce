"""ORACLE — test infrastructure only.  Only bench.py's CPU-baseline leg and
tests/ use this module.

The CPU baseline of the BSIM3 device load: the oracle's restatement of the
reference Master loops (MOSFET_B3::Master::updateState / loadDAEVectors /
loadDAEMatrices, N_DEV_MOSFET_B3.C:9222-9928: arrays of structs, pointer
read-modify-write stamping) timed

* serially on one host core, and
* partitioned over P host threads the way Xyce distributes devices over
  MPI ranks (N_DEV_DeviceMgr.C:2230-2232, 2277-2282): each partition owns a
  contiguous block of instances and loads into its own copy of the vectors
  and matrices; the copies are added in fixed partition order (the
  Export(Add) of fillComplete, N_LAS_Matrix.C:197-204).

The timing loop is C (baseline_ref.c) with POSIX threads, so no Python runs
inside the timed region.  `native_library()` compiles the sources with -O3
-march=native on the host that runs it, into a temporary directory.
"""
import ctypes
import os
import subprocess
import tempfile

import numpy as np

from . import bsim3_bind, oracle

HERE = os.path.dirname(os.path.abspath(__file__))
_dp = ctypes.POINTER(ctypes.c_double)


def _declare(L):
    bsim3_bind.declare(L)
    L.ref_b3_partitioned_loads.restype = ctypes.c_double
    L.ref_b3_partitioned_loads.argtypes = [
        ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)] + [_dp] * 5 + [
        ctypes.POINTER(_dp), ctypes.POINTER(_dp), ctypes.c_int64, ctypes.c_int64, ctypes.c_int]
    return L


def native_library():
    """(ctypes library, path): the oracle built -O3 -march=native here."""
    out = os.path.join(tempfile.mkdtemp(prefix="xg_oracle_native_"), "liboracle_native.so")
    subprocess.run(["make", "-s", "-B", "-C", HERE, "native", "NATIVE_SO=" + out], check=True)
    return _declare(ctypes.CDLL(out)), out


def portable_library():
    return _declare(oracle.lib())


class _Part:
    """What RefBsim3 reads from its parent: the CSR pattern, the offsets of
    its instances, and the value arrays its stamps point into."""

    def __init__(self, rowptr, offsets, nnz):
        self.rowptr = rowptr
        self.b3_off = offsets
        self.dFdx = np.zeros(nnz)
        self.dQdx = np.zeros(nnz)


class _View:
    def __init__(self, nl, recs):
        self.models = nl.models
        self.bsim3 = recs
        self.n_nodes = nl.n_nodes


class PartitionedBsim3:
    """P partitions of a netlist's BSIM3 instances, each a Master over a
    contiguous block with its own output copies."""

    def __init__(self, nl, P, L, gmin=1e-12, limiter=True, temp=300.15):
        nl.layout()
        self.nl, self.P, self.L = nl, P, L
        rowptr, colidx, _, b3_off = oracle.pattern_and_offsets(nl)
        self.rowptr = rowptr
        self.n1, self.nnz = nl.n_nodes + 1, len(colidx)
        n = len(nl.bsim3)
        bounds = [n * t // P for t in range(P + 1)]
        self.parts, self.masters, self.vecs = [], [], []
        for t in range(P):
            lo, hi = bounds[t], bounds[t + 1]
            part = _Part(rowptr, np.ascontiguousarray(b3_off[lo:hi]), self.nnz)
            m = bsim3_bind.RefBsim3(part, _View(nl, nl.bsim3[lo:hi]), gmin, limiter, temp, L=L)
            self.parts.append(part)
            self.masters.append(m)
            self.vecs.append([np.zeros(self.n1) for _ in range(4)])
        self.out = [np.zeros(self.n1) for _ in range(4)] + [np.zeros(self.nnz) for _ in range(2)]

    def set_solver_state(self, **flags):
        f = np.array([flags.get(k, 0) for k in ("dcop", "initJct", "initFix", "initTran",
                                                 "newtonIter", "tranop", "acop", "transient")],
                     np.int32)
        for m in self.masters:
            m.set_solver_state(f)

    def run(self, x, sta_c, sto_c, sta_n, sto_n, reps):
        """reps full loads; returns the wall seconds and the outputs (F, Q,
        dFdxdVp, dQdxdVp, dF/dx, dQ/dx) of the last one."""
        P = self.P
        M = (ctypes.c_void_p * P)(*[m.M for m in self.masters])
        part = []
        for t in range(P):
            part += self.vecs[t] + [self.parts[t].dFdx, self.parts[t].dQdx]
        part_p = (_dp * (6 * P))(*[a.ctypes.data_as(_dp) for a in part])
        out_p = (_dp * 6)(*[a.ctypes.data_as(_dp) for a in self.out])
        args = [a.ctypes.data_as(_dp) for a in (x, sta_c, sto_c, sta_n, sto_n)]
        el = self.L.ref_b3_partitioned_loads(P, M, *args, part_p, out_p, self.n1, self.nnz,
                                             int(reps))
        if el < 0:
            raise RuntimeError("could not start %d threads" % P)
        outs = part[:6] if P == 1 else self.out
        return el, [o.copy() for o in outs]

    def close(self):
        for m in self.masters:
            m.close()
        self.masters = []


def host_description():
    """Host facts that explain a per-core rate: logical CPUs, the CPUs this
    process may run on, the cgroup CPU quota and the CPU model."""
    d = dict(nproc=os.cpu_count(), affinity=len(os.sched_getaffinity(0)))
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            q, per = f.read().split()
            d["cgroup_cpus"] = None if q == "max" else float(q) / float(per)
    except (OSError, ValueError):
        d["cgroup_cpus"] = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    d["cpu_model"] = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return d
