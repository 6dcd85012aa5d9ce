"""ORACLE — test infrastructure only.  Only tests/, __graft_entry__.smoke()
and bench.py's CPU-baseline leg may import this module, and only as the
checker.

Drives the CPU restatement of the reference (liboracle.so, built from
mos1_ref.c / bsim3_ref.c by oracle/Makefile) from a laid-out Netlist:

* the CSR pattern and per-instance offsets come from ``stamp_ref`` (the
  restated jacStampMap / generateRowColData / matrixGlobalToLocal chain);
* ``RefLoad`` mirrors the reference DeviceMgr's per-iteration sequence:
  updateState -> loadDAEVectors -> loadDAEMatrices, device types in the
  order MOSFET1 then BSIM3, instances in netlist order.
"""
import ctypes
import os
import subprocess

import numpy as np

from . import stamp_ref

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)
_L = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _L
    if _L is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        vp = ctypes.c_void_p
        L.ref_mos1_model_new.restype = vp
        L.ref_mos1_model_new.argtypes = [ctypes.c_int, ctypes.c_int,
                                         ctypes.POINTER(ctypes.c_char_p), _dp]
        L.ref_mos1_model_free.argtypes = [vp]
        L.ref_mos1_master_new.restype = vp
        L.ref_mos1_master_free.argtypes = [vp]
        L.ref_mos1_add_instance.restype = ctypes.c_int
        L.ref_mos1_add_instance.argtypes = [vp, vp, ctypes.c_int, ctypes.POINTER(ctypes.c_char_p),
                                            _dp, _ip, ctypes.c_int, ctypes.c_int, ctypes.c_double]
        L.ref_mos1_set_offsets.argtypes = [vp, ctypes.c_int, _ip]
        L.ref_mos1_setup_pointers.argtypes = [vp, _ip, _dp, _dp]
        L.ref_mos1_set_solver_state.argtypes = [vp, _ip]
        L.ref_mos1_set_options.argtypes = [vp, ctypes.c_double, ctypes.c_int]
        L.ref_mos1_update_state.argtypes = [vp, _dp, _dp, _dp, _dp, _dp]
        L.ref_mos1_load_vectors.argtypes = [vp, _dp, _dp, _dp, _dp]
        L.ref_mos1_load_matrices.argtypes = [vp]
        L.ref_mos1_flags.argtypes = [vp, _ip, _ip]
        if hasattr(L, "ref_b3_model_new"):
            from . import bsim3_bind
            bsim3_bind.declare(L)
        _L = L
    return _L


def _kv(params):
    items = list(params.items())
    n = len(items)
    keys = (ctypes.c_char_p * max(n, 1))(*[k.upper().encode() for k, _ in items])
    vals = (ctypes.c_double * max(n, 1))(*[float(v) for _, v in items])
    return n, keys, vals


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


def mos1_stamp_rows(nl):
    """(gids, merged stamp, map, map2) per MOSFET1 instance."""
    stamps = stamp_ref.mos1_stamps()
    out = []
    for rec in nl.mos1:
        has_dc, has_sc = rec["lids"][4] >= 0, rec["lids"][5] >= 0
        st, mp, m2 = stamps[(has_dc, has_sc)]
        out.append((stamp_ref.mos1_gids(rec["lids"], has_dc, has_sc), st, mp, m2))
    return out


def pattern_and_offsets(nl):
    """Restated topology: CSR pattern over all devices, MOSFET1 offsets
    (n x 22) and BSIM3 offsets."""
    nl.layout()
    rows = []
    m1 = mos1_stamp_rows(nl)
    for gids, st, _, _ in m1:
        rows.append((gids, st))
    b3 = []
    if nl.bsim3:
        from . import bsim3_bind
        b3 = bsim3_bind.stamp_rows(nl)
        for gids, st, _, _ in b3:
            rows.append((gids, st))
    for gids, r, c in nl.foreign_stamps():
        st = [[] for _ in gids]
        for a, b in zip(r, c):
            st[a].append(b)
        rows.append((gids, st))
    rowptr, colidx = stamp_ref.build_pattern(nl.n_nodes, rows)
    rowmaps = {}
    m1_off = []
    for gids, st, mp, m2 in m1:
        jl = stamp_ref.stamp_offsets(rowptr, colidx, nl.n_nodes, gids, st, rowmaps)
        m1_off.append(stamp_ref.mos1_register_offsets(jl, mp, m2))
    b3_off = []
    if b3:
        from . import bsim3_bind
        for gids, st, mp, m2 in b3:
            jl = stamp_ref.stamp_offsets(rowptr, colidx, nl.n_nodes, gids, st, rowmaps)
            b3_off.append(bsim3_bind.register_offsets(jl, mp, m2))
    return (np.array(rowptr, np.int32), np.array(colidx, np.int32),
            np.array(m1_off, np.int32).reshape(-1, 22),
            np.array(b3_off, np.int32).reshape(len(b3_off), -1) if b3_off else np.zeros((0, 0), np.int32))


class RefLoad:
    """The reference's device load for one netlist, on the CPU."""

    def __init__(self, nl, gmin=1e-12, limiter=True, temp=300.15):
        L = lib()
        nl.layout()
        self.nl = nl
        self.rowptr, self.colidx, self.m1_off, self.b3_off = pattern_and_offsets(nl)
        self.nnz = len(self.colidx)
        self.dFdx = np.zeros(self.nnz)
        self.dQdx = np.zeros(self.nnz)
        self._models = {}
        self.m1 = L.ref_mos1_master_new()
        L.ref_mos1_set_options(self.m1, gmin, int(limiter))
        for name, m in nl.models.items():
            if m.kind == "mos1":
                n, k, v = _kv(m.params)
                p = L.ref_mos1_model_new(m.dtype, n, k, v)
                if not p:
                    raise ValueError("oracle rejected MOSFET1 model %s" % name)
                self._models[name] = p
        for i, rec in enumerate(nl.mos1):
            n, k, v = _kv(rec["params"])
            lids = np.array([nl.n_nodes if x < 0 else x for x in rec["lids"]], np.int32)
            idx = L.ref_mos1_add_instance(self.m1, self._models[rec["model"]], n, k, v,
                                          _i(lids), rec["sta"], rec["sto"], temp)
            if idx < 0:
                raise ValueError("oracle rejected MOSFET1 instance %s" % rec["name"])
            off = np.ascontiguousarray(self.m1_off[i])
            L.ref_mos1_set_offsets(self.m1, idx, _i(off))
        L.ref_mos1_setup_pointers(self.m1, _i(self.rowptr), _d(self.dFdx), _d(self.dQdx))
        self.b3 = None
        if nl.bsim3:
            from . import bsim3_bind
            self.b3 = bsim3_bind.RefBsim3(self, nl, gmin, limiter, temp)
        from .linear_ref import RefLinear
        self.lin = RefLinear(nl, self.rowptr, self.colidx, gmin=gmin, limiter=limiter,
                             temp=temp)

    def set_solver_state(self, dcop=0, initJct=0, initFix=0, initTran=0, newtonIter=0,
                         tranop=0, acop=0, transient=0):
        f = np.array([dcop, initJct, initFix, initTran, newtonIter, tranop, acop, transient],
                     np.int32)
        self._flags = f
        lib().ref_mos1_set_solver_state(self.m1, _i(f))
        if self.b3:
            self.b3.set_solver_state(f)
        self.lin.set_solver_state(tranop=tranop, transient=transient, dcop=dcop,
                                  initJct=initJct, newtonIter=newtonIter)

    def set_time(self, t, bp_tol=0.0):
        self.lin.set_time(t, bp_tol)

    def update_state(self, x, sta_c, sto_c, sta_n, sto_n):
        lib().ref_mos1_update_state(self.m1, _d(x), _d(sta_c), _d(sto_c), _d(sta_n), _d(sto_n))
        if self.b3:
            self.b3.update_state(x, sta_c, sto_c, sta_n, sto_n)
        self.lin.update_state(x, sta_n, sto_c, sto_n)

    def load_vectors(self, F, Q, dFdxdVp, dQdxdVp):
        lib().ref_mos1_load_vectors(self.m1, _d(F), _d(Q), _d(dFdxdVp), _d(dQdxdVp))
        if self.b3:
            self.b3.load_vectors(F, Q, dFdxdVp, dQdxdVp)
        self.lin.load_vectors(F, Q, dFdxdVp, dQdxdVp)

    def load_matrices(self):
        """Adds into self.dFdx / self.dQdx (zero them first, as the
        assembler does)."""
        lib().ref_mos1_load_matrices(self.m1)
        if self.b3:
            self.b3.load_matrices()
        self.lin.load_matrices(self.dFdx, self.dQdx)

    def mos1_flags(self):
        n = len(self.nl.mos1)
        lim = np.zeros(max(n, 1), np.int32)
        org = np.zeros(max(n, 1), np.int32)
        lib().ref_mos1_flags(self.m1, _i(lim), _i(org))
        return lim[:n], org[:n]

    def full_load(self, x, sta_c, sto_c, sta_n, sto_n):
        """Zeroed outputs + the three calls; returns (F, Q, dFdxdVp,
        dQdxdVp, dFdx, dQdx) as new arrays."""
        n1 = self.nl.n_nodes + 1
        F, Q, fl, ql = (np.zeros(n1) for _ in range(4))
        self.dFdx[:] = 0.0
        self.dQdx[:] = 0.0
        self.update_state(x, sta_c, sto_c, sta_n, sto_n)
        self.load_vectors(F, Q, fl, ql)
        self.load_matrices()
        return F, Q, fl, ql, self.dFdx.copy(), self.dQdx.copy()

    def close(self):
        L = lib()
        if self.m1:
            L.ref_mos1_master_free(self.m1)
            self.m1 = None
        for p in self._models.values():
            L.ref_mos1_model_free(p)
        self._models = {}
        if self.b3:
            self.b3.close()
            self.b3 = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
