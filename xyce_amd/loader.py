"""Host-side mirror of the reference's device-load interface over the C ABI.

``DeviceLoad`` plays the role of the MOSFET1/BSIM3 ``Master`` objects behind
DeviceMgr (src/DeviceModelPKG/Core/src/N_DEV_DeviceMgr.C:2201-2661): it is
built once from a laid-out Netlist, then per Newton iteration

    set_solver_state(...)                 DeviceMgr::setupSolverInfo_
    update_state(x, sta_c, sto_c, sta_n, sto_n)   Device::updateState
    load_vectors(F, Q, dFdxdVp, dQdxdVp)          Device::loadDAEVectors
    load_matrices(dFdx_vals, dQdx_vals)           Device::loadDAEMatrices

Arguments are torch tensors resident on the GPU (float64, contiguous); torch
is used only to own device memory and streams.
"""
import ctypes
import threading
import numpy as np

from . import _lib
from ._lib import check, keys_vals, lib

# Device allocations, uploads, synchronous copies and graph captures of
# concurrent users (the serving loop runs several transients at once) are
# serialised by this lock; a graph capture must not overlap another thread's
# allocation or synchronous copy.  Kernel launches are not serialised.
DEVICE_LOCK = threading.RLock()


def _ptr(t, dtype="torch.float64"):
    if not t.is_cuda or str(t.dtype) != dtype or not t.is_contiguous():
        raise ValueError("expected a contiguous %s CUDA tensor" % dtype[6:])
    return ctypes.c_void_p(t.data_ptr())


def _iptr(t):
    return _ptr(t, "torch.int32")


LU_MODES = {"auto": 0, "fronts": 1, "tasks": 2}


class DeviceLoad:
    def __init__(self, nl, gmin=1e-12, limiter=True, temp=300.15, tnom=300.15,
                 pattern=None, plan_only=False, lu_fronts=False, lu_mode="auto"):
        nl.layout()
        L = lib()
        self.nl = nl
        c = ctypes.c_void_p()
        check(L.xg_circuit_new(ctypes.byref(c), nl.n_nodes, nl.n_state, nl.n_store),
              "xg_circuit_new")
        try:
            check(L.xg_circuit_set_options(c, gmin, int(limiter), temp, tnom),
                  "xg_circuit_set_options")
            # linear-solve form: "auto" (task lists where they fit in LDS,
            # dense fronts elsewhere), "fronts" (every component) or
            # "tasks" (task lists wherever their budgets allow)
            mode = LU_MODES["fronts" if lu_fronts else lu_mode]
            if mode:
                check(L.xg_circuit_set_lu_mode(c, mode), "xg_circuit_set_lu_mode")
            mid = {}
            for name, m in nl.models.items():
                if m.kind == "diode":
                    continue
                n, k, v = keys_vals(m.params)
                out = ctypes.c_int32()
                fn = L.xg_add_mos1_model if m.kind == "mos1" else L.xg_add_bsim3_model
                check(fn(c, m.dtype, n, k, v, ctypes.byref(out)), "add model " + name)
                mid[name] = out.value
            # instances in runs of one parameter-name list, one bulk call
            # per run (the same as one call per instance, in order)
            for recs, fn, what in ((nl.mos1, L.xg_add_mos1_instances, "MOSFET1"),
                                   (nl.bsim3, L.xg_add_bsim3_instances, "BSIM3")):
                i = 0
                while i < len(recs):
                    r0 = recs[i]
                    names = tuple(r0["params"])
                    j = i + 1
                    while j < len(recs) and tuple(recs[j]["params"]) == names:
                        j += 1
                    run = recs[i:j]
                    n, k, _ = keys_vals(r0["params"])
                    vals = np.array([list(r["params"].values()) for r in run],
                                    dtype=np.float64).reshape(len(run), n)
                    models = np.array([mid[r["model"]] for r in run], dtype=np.int32)
                    lids = np.array([r["lids"] for r in run], dtype=np.int32)
                    sta = np.array([r["sta"] for r in run], dtype=np.int32)
                    sto = np.array([r["sto"] for r in run], dtype=np.int32)
                    ids = np.zeros(len(run), np.int32)
                    check(fn(c, models.ctypes.data, len(run), n, k, vals.ctypes.data,
                             lids.ctypes.data, sta.ctypes.data, sto.ctypes.data,
                             ids.ctypes.data),
                          "add %s %s..%s" % (what, r0["name"], run[-1]["name"]))
                    i = j
            # resistors and capacitors in two calls (each type keeps its order)
            rs, cs = [], []
            for kind, p, q, br, val, pulse, st in nl.linear_devices():
                if kind == "R":
                    rs.append((p, q, val))
                elif kind == "C":
                    cs.append((p, q, val, st))
                elif kind == "D":
                    nm, km, vm = keys_vals(val[0])
                    ni, ki, vi = keys_vals(val[1])
                    check(L.xg_add_diode(c, p, q, nm, km, vm, ni, ki, vi, st), "add diode")
                else:
                    pp = (ctypes.c_double * 7)(*pulse) if pulse else None
                    check(L.xg_add_vsource(c, p, q, br, val, pp), "add source")
            if rs:
                a = np.array(rs, dtype=np.float64)
                ip, iq = (np.ascontiguousarray(a[:, k], dtype=np.int32) for k in (0, 1))
                vr = np.ascontiguousarray(a[:, 2])
                check(L.xg_add_resistors(c, len(rs), ip.ctypes.data, iq.ctypes.data,
                                         vr.ctypes.data), "add resistors")
            if cs:
                a = np.array(cs, dtype=np.float64)
                ip, iq, ist = (np.ascontiguousarray(a[:, k], dtype=np.int32) for k in (0, 1, 3))
                vc = np.ascontiguousarray(a[:, 2])
                check(L.xg_add_capacitors(c, len(cs), ip.ctypes.data, iq.ctypes.data,
                                          vc.ctypes.data, ist.ctypes.data), "add capacitors")
            if pattern is not None:
                rp, ci = (np.ascontiguousarray(a, dtype=np.int32) for a in pattern)
                check(L.xg_circuit_set_pattern(
                    c, rp.ctypes.data_as(_lib.c_int32_p), ci.ctypes.data_as(_lib.c_int32_p)),
                    "set pattern")
            h = ctypes.c_void_p()
            # host analysis first, then the device half under the lock that
            # keeps allocations and uploads apart from graph captures
            check(L.xg_finalize_plan(c, ctypes.byref(h)), "xg_finalize")
            if not plan_only:
                with DEVICE_LOCK:
                    rc = L.xg_upload(h)
                if rc:
                    L.xg_destroy(h)
                    check(rc, "xg_upload")
        finally:
            L.xg_circuit_free(c)
        self.h = h
        nr = ctypes.c_int32()
        nnz = ctypes.c_int64()
        check(L.xg_pattern_size(h, ctypes.byref(nr), ctypes.byref(nnz)), "pattern size")
        self.n_rows = nr.value
        self.nnz = nnz.value
        self.n_mos1 = len(nl.mos1)
        self.n_bsim3 = len(nl.bsim3)

    def close(self):
        if getattr(self, "h", None):
            with DEVICE_LOCK:
                lib().xg_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- setup queries
    def pattern(self):
        rp = np.zeros(self.n_rows + 1, np.int32)
        ci = np.zeros(max(self.nnz, 1), np.int32)
        check(lib().xg_get_pattern(self.h, rp.ctypes.data_as(_lib.c_int32_p),
                                   ci.ctypes.data_as(_lib.c_int32_p)), "get pattern")
        return rp, ci[: self.nnz]

    def front_plan(self):
        """Host copy of the dense-front plan (xg_lu_front_export) as a dict
        of numpy arrays."""
        sz = np.zeros(6, np.int64)
        check(lib().xg_lu_front_export(self.h, sz.ctypes.data, None, None), "front export")
        nf, nr, nch, nlev, ne, nc = (int(v) for v in sz)
        i32 = dict(comp=nf, m=nf, k=nf, rows=nr, cbmap=nr, ch_ptr=nf + 1, ch=nch,
                   lv_ptr=nlev + 1, lv=nf, ent_k=ne, front_comp=nc)
        i64 = dict(fo=nf + 1, ro=nf + 1, ent_fpos=ne)
        out = {k: np.zeros(max(v, 1), np.int32) for k, v in i32.items()}
        out.update({k: np.zeros(max(v, 1), np.int64) for k, v in i64.items()})
        p32 = (ctypes.c_void_p * 11)(*[out[k].ctypes.data for k in i32])
        p64 = (ctypes.c_void_p * 3)(*[out[k].ctypes.data for k in i64])
        check(lib().xg_lu_front_export(self.h, sz.ctypes.data, p32, p64), "front export")
        for k, v in list(i32.items()) + list(i64.items()):
            out[k] = out[k][:v]
        return out

    def lu_plan(self):
        """Host copy of the static-pivot sparse LU plan (xg_lu_export) as a
        dict of numpy arrays."""
        sz = np.zeros(6, np.int64)
        check(lib().xg_lu_export(self.h, sz.ctypes.data, None, None), "lu export")
        n, nnz, nc, nlev, nt, nops = (int(v) for v in sz)
        names = ["prow", "qcol", "lu_pos", "ok", "lv_comp", "lv_ptr", "t_dst", "t_div",
                 "t_op", "ops"]
        lens = [n, n, nnz, nc, 3 * nc + 1, nlev, nt, nt, nt + 1, 2 * nops]
        arrs = [np.zeros(max(k, 1), np.int32) for k in lens]
        ptrs = (ctypes.c_void_p * 10)(*[a.ctypes.data for a in arrs])
        voff = np.zeros(nc + 1, np.int64)
        check(lib().xg_lu_export(self.h, sz.ctypes.data, ptrs, voff.ctypes.data), "lu export")
        out = {k: a[:m] for k, a, m in zip(names, arrs, lens)}
        out["ops"] = out["ops"].reshape(-1, 2)
        out["voff"] = voff
        return out

    def lu_info(self):
        """xg_lu_info as a dict."""
        info = np.zeros(8, np.int64)
        check(lib().xg_lu_info(self.h, info.ctypes.data), "lu info")
        keys = ["fnnz", "factor_ops", "max_factor_levels", "max_solve_levels",
                "static_fail", "global_components", "fallbacks", "tasks"]
        return dict(zip(keys, (int(v) for v in info)))

    def mos1_offsets(self):
        off = np.zeros((max(self.n_mos1, 1), 22), np.int32)
        check(lib().xg_get_mos1_offsets(self.h, off.ctypes.data_as(_lib.c_int32_p)), "offsets")
        return off[: self.n_mos1]

    def bsim3_offsets(self):
        off = np.zeros((max(self.n_bsim3, 1), 22), np.int32)
        check(lib().xg_get_bsim3_offsets(self.h, off.ctypes.data_as(_lib.c_int32_p)), "offsets")
        return off[: self.n_bsim3]

    def mos1_fields(self):
        nf = ctypes.c_int32()
        check(lib().xg_get_mos1_fields(self.h, None, ctypes.byref(nf)), "fields")
        out = np.zeros((max(self.n_mos1, 1), nf.value))
        check(lib().xg_get_mos1_fields(self.h, out.ctypes.data_as(_lib.c_double_p),
                                       ctypes.byref(nf)), "fields")
        return out[: self.n_mos1]

    def bsim3_fields(self):
        nf = ctypes.c_int32()
        check(lib().xg_get_bsim3_fields(self.h, None, ctypes.byref(nf)), "fields")
        out = np.zeros((max(self.n_bsim3, 1), max(nf.value, 1)))
        check(lib().xg_get_bsim3_fields(self.h, out.ctypes.data_as(_lib.c_double_p),
                                        ctypes.byref(nf)), "fields")
        return out[: self.n_bsim3, : nf.value]

    # ---- per Newton iteration
    def set_solver_state(self, dcop=0, initJct=0, initFix=0, initTran=0, newtonIter=0,
                         tranop=0, acop=0, transient=0):
        ss = _lib.SolverState(dcop, initJct, initFix, initTran, newtonIter, tranop,
                              acop, transient)
        check(lib().xg_set_solver_state(self.h, ctypes.byref(ss)), "set solver state")

    EVAL_VARIANTS = {"auto": 0, "hbm": 1, "staged": 2, "staged_split": 3}

    def set_output_mode(self, overwrite):
        """overwrite=False: loads add into the outputs (the reference's
        "+=" into zeroed vectors); True: loads write every entry, so the
        outputs need no zeroing first."""
        check(lib().xg_set_output_mode(self.h, 1 if overwrite else 0), "output mode")

    def set_eval_variant(self, name):
        """Force one BSIM3 evaluation kernel form (all give identical
        results); "auto" chooses by circuit size."""
        check(lib().xg_set_eval_variant(self.h, self.EVAL_VARIANTS[name]), "eval variant")

    def set_time(self, t, bp_tol=0.0):
        """Time at which the next update_state evaluates PULSE sources."""
        check(lib().xg_set_time(self.h, float(t), float(bp_tol)), "set time")

    def set_stream(self, stream_handle):
        check(lib().xg_set_stream(self.h, ctypes.c_void_p(stream_handle)), "set stream")

    def update_state(self, x, sta_curr, sto_curr, sta_next, sto_next):
        check(lib().xg_update_state(self.h, _ptr(x), _ptr(sta_curr), _ptr(sto_curr),
                                    _ptr(sta_next), _ptr(sto_next)), "xg_update_state")

    DEVICE_TYPES = {"mos1": 1, "bsim3": 2, "linear": 3}

    def update_state_type(self, kind, x, sta_curr, sto_curr, sta_next, sto_next):
        """One device type's updateState alone ("mos1", "bsim3" or "linear")."""
        check(lib().xg_update_state_type(self.h, self.DEVICE_TYPES[kind], _ptr(x),
                                         _ptr(sta_curr), _ptr(sto_curr), _ptr(sta_next),
                                         _ptr(sto_next)), "xg_update_state_type")

    def load_vectors(self, F, Q, dFdxdVp, dQdxdVp):
        check(lib().xg_load_vectors(self.h, _ptr(F), _ptr(Q), _ptr(dFdxdVp),
                                    _ptr(dQdxdVp)), "xg_load_vectors")

    def load_matrices(self, dFdx, dQdx):
        check(lib().xg_load_matrices(self.h, _ptr(dFdx), _ptr(dQdx)), "xg_load_matrices")

    def load_all(self, x, sta_curr, sto_curr, sta_next, sto_next, F, Q, dFdxdVp, dQdxdVp,
                 dFdx, dQdx):
        check(lib().xg_load_all(self.h, *[_ptr(t) for t in (
            x, sta_curr, sto_curr, sta_next, sto_next, F, Q, dFdxdVp, dQdxdVp, dFdx, dQdx)]),
            "xg_load_all")

    def stats(self):
        st = _lib.Stats()
        check(lib().xg_get_stats(self.h, ctypes.byref(st)), "stats")
        return st.n_limited, st.n_nonfinite

    def sync(self):
        check(lib().xg_sync(self.h), "xg_sync")
