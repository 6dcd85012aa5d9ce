"""ORACLE — test infrastructure only.

ctypes side of the BSIM3 restatement (bsim3_ref.c).  The BSIM3 stamps are
built exactly like MOSFET1's: the same 6x6 (d, g, s, b, d', s') base stamp
and the same three jacStampMap merges (N_DEV_MOSFET_B3.C:2378-2418), and
registerJacLIDs assigns the same 22 offsets in the same order
(N_DEV_MOSFET_B3.C:3150-3330) when neither IC= nor NQSMOD is given.
"""
import ctypes

import numpy as np

from . import stamp_ref

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)


def declare(L):
    vp = ctypes.c_void_p
    L.ref_b3_model_new.restype = vp
    L.ref_b3_model_new.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_char_p),
                                   _dp, ctypes.c_double]
    L.ref_b3_model_free.argtypes = [vp]
    L.ref_b3_master_new.restype = vp
    L.ref_b3_master_new.argtypes = [ctypes.c_double, ctypes.c_int]
    L.ref_b3_master_free.argtypes = [vp]
    L.ref_b3_add_instance.restype = ctypes.c_int
    L.ref_b3_add_instance.argtypes = [vp, vp, ctypes.c_int, ctypes.POINTER(ctypes.c_char_p), _dp,
                                      _ip, ctypes.c_int, ctypes.c_int, ctypes.c_double]
    L.ref_b3_set_offsets.argtypes = [vp, ctypes.c_int, _ip]
    L.ref_b3_setup_pointers.argtypes = [vp, _ip, _dp, _dp]
    L.ref_b3_set_solver_state.argtypes = [vp, _ip]
    L.ref_b3_update_state.restype = ctypes.c_int
    L.ref_b3_update_state.argtypes = [vp, _dp, _dp, _dp, _dp, _dp]
    L.ref_b3_load_vectors.argtypes = [vp, _dp, _dp, _dp, _dp]
    L.ref_b3_load_matrices.argtypes = [vp]
    L.ref_b3_flags.argtypes = [vp, _ip, _ip]
    L.ref_b3_sdp_probe.argtypes = [vp, ctypes.c_int, _dp]


def stamp_rows(nl):
    stamps = stamp_ref.mos1_stamps()
    out = []
    for rec in nl.bsim3:
        has_dc, has_sc = rec["lids"][4] >= 0, rec["lids"][5] >= 0
        st, mp, m2 = stamps[(has_dc, has_sc)]
        out.append((stamp_ref.mos1_gids(rec["lids"], has_dc, has_sc), st, mp, m2))
    return out


def register_offsets(jac_lid_vec, mp, map2):
    return stamp_ref.mos1_register_offsets(jac_lid_vec, mp, map2)


def _kv(params):
    items = list(params.items())
    n = len(items)
    keys = (ctypes.c_char_p * max(n, 1))(*[k.upper().encode() for k, _ in items])
    vals = (ctypes.c_double * max(n, 1))(*[float(v) for _, v in items])
    return n, keys, vals


class RefBsim3:
    def __init__(self, parent, nl, gmin, limiter, temp, L=None):
        from .oracle import lib
        L = L if L is not None else lib()
        self.L = L
        self.nl = nl
        self.models = {}
        self.M = L.ref_b3_master_new(gmin, int(limiter))
        for name, m in nl.models.items():
            if m.kind == "bsim3":
                n, k, v = _kv(m.params)
                p = L.ref_b3_model_new(m.dtype, n, k, v, 300.15)
                if not p:
                    raise ValueError("oracle rejected BSIM3 model %s" % name)
                self.models[name] = p
        for i, rec in enumerate(nl.bsim3):
            n, k, v = _kv(rec["params"])
            lids = np.array([nl.n_nodes if x < 0 else x for x in rec["lids"]], np.int32)
            idx = L.ref_b3_add_instance(self.M, self.models[rec["model"]], n, k, v,
                                        lids.ctypes.data_as(_ip), rec["sta"], rec["sto"], temp)
            if idx < 0:
                raise ValueError("oracle rejected BSIM3 instance %s (%d)" % (rec["name"], idx))
            off = np.ascontiguousarray(parent.b3_off[i])
            L.ref_b3_set_offsets(self.M, idx, off.ctypes.data_as(_ip))
        L.ref_b3_setup_pointers(self.M, parent.rowptr.ctypes.data_as(_ip),
                                parent.dFdx.ctypes.data_as(_dp), parent.dQdx.ctypes.data_as(_dp))

    def set_solver_state(self, f):
        self.L.ref_b3_set_solver_state(self.M, f.ctypes.data_as(_ip))

    def update_state(self, x, sta_c, sto_c, sta_n, sto_n):
        rc = self.L.ref_b3_update_state(self.M, *[a.ctypes.data_as(_dp) for a in
                                                  (x, sta_c, sto_c, sta_n, sto_n)])
        if rc:
            raise ValueError("oracle BSIM3 load does not cover this model (capMod outside 0-3)")

    def load_vectors(self, F, Q, fl, ql):
        self.L.ref_b3_load_vectors(self.M, *[a.ctypes.data_as(_dp) for a in (F, Q, fl, ql)])

    def load_matrices(self):
        self.L.ref_b3_load_matrices(self.M)

    def flags(self):
        n = len(self.nl.bsim3)
        lim = np.zeros(max(n, 1), np.int32)
        org = np.zeros(max(n, 1), np.int32)
        self.L.ref_b3_flags(self.M, lim.ctypes.data_as(_ip), org.ctypes.data_as(_ip))
        return lim[:n], org[:n]

    def sdp(self, i):
        out = np.zeros(30)
        self.L.ref_b3_sdp_probe(self.M, i, out.ctypes.data_as(_dp))
        return out

    def close(self):
        if self.M:
            self.L.ref_b3_master_free(self.M)
            self.M = None
        for p in self.models.values():
            self.L.ref_b3_model_free(p)
        self.models = {}
