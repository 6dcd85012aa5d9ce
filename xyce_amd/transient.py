"""Transient analysis on the GPU: DC operating point, then fixed-step
ONESTEP integration (backward Euler for the first two steps, trapezoidal
after, as N_TIA_OneStep::completeStep raises the order,
N_TIA_OneStep.C:1945-2010), with a full Newton solve per step.

Two drivers run the same kernels in the same order:
  "native"  the library's xg_tran_* driver: each Newton iteration is one
            hipGraph launch ending in the convergence test on the device;
            the host reads four control words per iteration;
  "python"  the loop below, one C-ABI call per kernel group, the
            convergence test on the host.
Both give bit-identical results.

Every Newton iteration runs on the device through the C ABI:
  xg_update_state + ordered reduction (xg_load_all, overwrite mode),
  xg_tia_assemble (N_TIA_OneStep::obtainResidual / obtainJacobian),
  xg_lu_solve (static-pivot sparse LU per independent sub-circuit, with
  the partial-pivoting dense LU for components that fail the threshold
  test; + update + norms),
  xg_rhs_norms (residual norm of the convergence test).
The host only keeps the step sequence and reads one small stats array per
iteration.

Newton control follows N_NLS_DampedNewton / N_NLS_NLParams:
  weights   reltol * max(|x_old|, |x_new|) + abstol (DampedNewton.C:311-335),
            fixed per step in a transient, refreshed every DC iteration;
  converged every device converged (no junction limiting in the last load)
            and  max|rhs| <= RHSTol  and  wRMS(dx) <= deltaXTol,
            or wRMS(dx) <= 1e-6 (DampedNewton.C:1787-1830);
  DC:       reltol 1e-3, abstol 1e-12, RHSTol 1e-6, deltaXTol 1, 200 steps;
  TRAN:     reltol 1e-2, abstol 1e-6, RHSTol 1e-2, deltaXTol 0.33, 20 steps
            (N_NLS_NLParams.C:108-125, NLParams.h defaults).
The predictor is ONESTEP's (N_TIA_OneStep.C:115-133 with the constant-step
coefficients of updateCoeffs, N_TIA_OneStep.C:1355-1410).  Time steps are
fixed (the reference's LTE step control is not reproduced); a step whose
Newton solve fails is retried with half the step.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import check, lib
from .loader import DEVICE_LOCK, DeviceLoad, _iptr, _ptr

DC_PARAMS = dict(reltol=1e-3, abstol=1e-12, rhstol=1e-6, dxtol=1.0, maxit=200)
TRAN_PARAMS = dict(reltol=1e-2, abstol=1e-6, rhstol=1e-2, dxtol=0.33, maxit=20)
SMALL_UPDATE = 1e-6


class NewtonFailure(RuntimeError):
    pass




class _TranBuffers(ctypes.Structure):
    _fields_ = [(name, ctypes.c_void_p) for name in (
        "x", "x_curr", "xh1", "xh2", "xpred", "sta_c", "sta_n", "sto_c", "sto_n",
        "F", "Q", "FL", "QL", "JF", "JQ", "J", "rhs", "qn0", "fprev", "stats", "nstats",
        "lim", "active", "conv", "counts", "ctl", "step")]


_FAIL_WORDS = {1: "singular Jacobian", 2: "non-finite device values", 3: "no convergence"}


class Transient:
    def __init__(self, nl, gmin=1e-12, limiter=True, bp_tol=None, device="cuda",
                 solver="lu", driver=None, graphs=True, lu_fronts=False, lu_mode="auto",
                 fused_steps=True):
        # the device load's host analysis runs unlocked; its upload, the
        # buffers and the driver's graph captures under DEVICE_LOCK.
        # lu_fronts: every component through dense fronts; lu_mode: the
        # linear-solve form otherwise (loader.LU_MODES); fused_steps: the
        # native driver's fixed steps in groups per graph launch
        # (xg_tran_run), with graphs
        dev = DeviceLoad(nl, gmin=gmin, limiter=limiter, lu_fronts=lu_fronts, lu_mode=lu_mode)
        with DEVICE_LOCK:
            self._init(nl, dev, bp_tol, device, solver, driver, graphs, limiter)
            if self._tr is not None and graphs and fused_steps:
                check(lib().xg_tran_prepare_run(self._tr), "fused-step graphs")

    def _init(self, nl, dev, bp_tol, device, solver, driver, graphs, limiter):
        """solver: "lu" (static-pivot sparse LU with pivoting fallback) or
        "dense" (dense partial-pivoting LU for every component; python
        driver only).  driver: "native" (default with "lu") or "python"."""
        import torch
        if solver not in ("lu", "dense"):
            raise ValueError("solver must be 'lu' or 'dense'")
        if driver is None:
            driver = "native" if solver == "lu" else "python"
        if driver not in ("native", "python"):
            raise ValueError("driver must be 'native' or 'python'")
        if driver == "native" and solver != "lu":
            raise ValueError("the native driver solves with the sparse LU")
        self.solver = solver
        self.driver = driver
        self.torch = torch
        self.nl = nl.layout()
        self.dev = dev
        self.dev.set_output_mode(True)
        n1 = nl.n_nodes + 1
        nc = ctypes.c_int32()
        check(lib().xg_components(self.dev.h, ctypes.byref(nc), None), "components")
        self.n_comp = nc.value
        z = lambda n: torch.zeros(n, dtype=torch.float64, device=device)
        self.x = z(n1)            # Newton iterate (next solution)
        self.x_curr = z(n1)       # accepted solution at t_n
        self.xh1 = z(n1)          # x_n - x_{n-1}
        self.xh2 = z(n1)          # x_{n-1} - x_{n-2}
        self.xpred = z(n1)
        self.sta_c, self.sta_n = z(nl.n_state), z(nl.n_state)
        self.sto_c, self.sto_n = z(nl.n_store), z(nl.n_store)
        self.vec = z(4 * n1)
        self.F, self.Q, self.FL, self.QL = (self.vec[k * n1:(k + 1) * n1] for k in range(4))
        nnz = self.dev.nnz
        self.JF, self.JQ, self.J = z(nnz), z(nnz), z(nnz)
        self.rhs = z(n1)
        self.qn0, self.fprev = z(n1), z(n1)
        self.stats = z(4 * self.n_comp)
        self.nstats = z(4 * self.n_comp)
        self.lim = torch.zeros(self.n_comp, dtype=torch.int32, device=device)
        self.active = torch.ones(self.n_comp, dtype=torch.int32, device=device)
        # loads skip the instances of converged sub-circuits
        check(lib().xg_set_active_mask(self.dev.h, self.active.data_ptr()), "active mask")
        self.bp_tol = bp_tol
        self.t = 0.0
        self.n_steps = 0
        self.newton_total = 0
        self.step_failures = 0
        self.limiter = limiter
        self._tr = None
        if driver == "native":
            self.conv = torch.zeros(self.n_comp, dtype=torch.int32, device=device)
            self.counts = torch.zeros(128, dtype=torch.int64, device=device)
            self.ctl = torch.zeros(8, dtype=torch.int32, device=device)
            self.stepv = torch.zeros(4, dtype=torch.float64, device=device)
            b = _TranBuffers(*[t.data_ptr() if t.numel() else None for t in (
                self.x, self.x_curr, self.xh1, self.xh2, self.xpred, self.sta_c, self.sta_n,
                self.sto_c, self.sto_n, self.F, self.Q, self.FL, self.QL, self.JF, self.JQ,
                self.J, self.rhs, self.qn0, self.fprev, self.stats, self.nstats, self.lim,
                self.active, self.conv, self.counts, self.ctl, self.stepv)])
            torch.cuda.current_stream().synchronize()
            tr = ctypes.c_void_p()
            check(lib().xg_tran_create(self.dev.h, ctypes.byref(b), bp_tol or 0.0,
                                       ctypes.byref(tr)), "transient driver")
            self._tr = tr
            check(lib().xg_tran_set_graphs(tr, 1 if graphs else 0), "graphs")

    def close(self):
        with DEVICE_LOCK:
            if self._tr is not None:
                lib().xg_tran_destroy(self._tr)
                self._tr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _native(self, rc, what):
        if rc > 0:
            self._sync_counters()
            raise NewtonFailure("%s: %s at t=%g" % (what, _FAIL_WORDS.get(rc, rc), self.t))
        check(rc, what)
        self._sync_counters()

    def host_reads(self):
        """The native driver's control-word reads so far (host round trips
        of the Newton loop); None without the native driver."""
        if self._tr is None:
            return None
        c = (ctypes.c_int64 * 4)()
        check(lib().xg_tran_counters(self._tr, c), "driver counters")
        return int(c[0])

    def _sync_counters(self):
        t = ctypes.c_double()
        c = (ctypes.c_int64 * 4)()
        check(lib().xg_tran_info(self._tr, ctypes.byref(t), c), "driver info")
        self.t, self.n_steps, self.newton_total, self.step_failures = \
            t.value, c[0], c[1], c[2]

    # ---- one device load + assembly
    def _load(self, it, transient, first_tran=False, dcop=False, init_jct=False, t=0.0):
        dev = self.dev
        dev.set_solver_state(dcop=int(dcop), tranop=int(dcop), initJct=int(init_jct),
                             initTran=int(first_tran), newtonIter=it,
                             transient=int(transient))
        dev.set_time(t, self.bp_tol if self.bp_tol is not None else 0.0)
        dev.load_all(self.x, self.sta_c, self.sto_c, self.sta_n, self.sto_n,
                     self.F, self.Q, self.FL, self.QL, self.JF, self.JQ)

    def _assemble(self, transient, order=1, dt=0.0):
        check(lib().xg_tia_assemble(
            self.dev.h, int(transient), order, dt,
            *[_ptr(t) for t in (self.qn0, self.fprev, self.F, self.Q, self.FL, self.QL,
                               self.JF, self.JQ, self.rhs, self.J)]), "assemble")

    def _solve(self, prm, wa, wb):
        fn = lib().xg_lu_solve if self.solver == "lu" else lib().xg_dense_solve
        check(fn(self.dev.h, _ptr(self.J), _ptr(self.rhs), _ptr(self.x),
                 _ptr(wa), _ptr(wb), prm["reltol"], prm["abstol"],
                 _iptr(self.active), _ptr(self.stats)), "linear solve")

    def _newton(self, prm, transient, order, dt, t, first_tran=False, dcop=False,
                wa=None, wb=None):
        """Full Newton from self.x, one convergence test per independent
        sub-circuit: a converged component is left out of later solves
        (its iterate, and so its load, no longer changes), so each circuit
        of a batch follows exactly the iterations it would take alone.
        Returns the iteration count."""
        self.active.fill_(1)
        conv = np.zeros(self.n_comp, bool)
        self._load(0, transient, first_tran=first_tran, dcop=dcop, init_jct=dcop, t=t)
        self._assemble(transient, order, dt)
        it = 0
        while True:
            if dcop:
                # DC weights follow the iterate (DampedNewton.C:520-528); the
                # solve reads each weight before it updates that entry
                wb = self.x
            self._solve(prm, wa, wb)
            it += 1
            self._load(it, transient, first_tran=first_tran, dcop=dcop, t=t)
            self._assemble(transient, order, dt)
            L = lib()
            check(L.xg_rhs_norms(self.dev.h, _ptr(self.rhs), _ptr(self.nstats)), "norms")
            check(L.xg_component_limited(self.dev.h, _iptr(self.lim)), "limited")
            st = self.stats.view(-1, 4).cpu().numpy()
            nrm = self.nstats.view(-1, 4).cpu().numpy()
            lim = self.lim.cpu().numpy()
            if st[:, 2].any():
                raise NewtonFailure("singular Jacobian at t=%g" % t)
            _, n_bad = self.dev.stats()
            if n_bad:
                raise NewtonFailure("non-finite device values at t=%g" % t)
            max_rhs, upd = nrm[:, 0], st[:, 1]
            devices_ok = (lim == 0) | (it >= prm["maxit"])
            now = devices_ok & (((max_rhs <= prm["rhstol"]) & (upd <= prm["dxtol"])) |
                                (upd <= SMALL_UPDATE))
            conv |= now
            if conv.all():
                break
            if it >= prm["maxit"]:
                bad = np.flatnonzero(~conv)
                raise NewtonFailure("Newton did not converge at t=%g in %d of %d sub-circuits "
                                    "(|rhs| %g, dx %g)" % (t, len(bad), self.n_comp,
                                                           max_rhs[bad].max(), upd[bad].max()))
            self.active.copy_(self.torch.from_numpy((~conv).astype(np.int32)))
        self.newton_total += it
        return it

    # ---- analyses
    def uic(self):
        """Start from zero without an operating point (.TRAN UIC): one load
        at x = 0 with operating-point flags starts the history."""
        if self._tr is not None:
            self._native(lib().xg_tran_uic(self._tr), "UIC start")
            return
        self.x.zero_()
        self.sta_c.zero_()
        self.sto_c.zero_()
        self.active.fill_(1)
        self._load(1, False, dcop=True, t=0.0)
        self.x_curr.copy_(self.x)
        self.sta_c.copy_(self.sta_n)
        self.sto_c.copy_(self.sto_n)
        self.qn0.copy_(self.Q)
        self.fprev.copy_(self.F)
        self.xh1.zero_()
        self.xh2.zero_()
        self.t = 0.0
        self.n_steps = 0

    def dcop(self, continuation="none"):
        """DC operating point at t = 0 (the tranop before a transient).
        continuation="gmin": by gmin stepping (.options nonlin
        continuation=gmin: a conductance 10^p - 10^-4 from every voltage
        unknown to ground, p stepped from 4 to -4); "auto": Newton first,
        gmin stepping if it fails (no counterpart in the reference, which
        stops at a failed operating point).  Continuation needs the native
        driver."""
        modes = {"none": 0, "gmin": 1, "auto": 2}
        if continuation not in modes:
            raise ValueError("continuation must be one of %s" % sorted(modes))
        if self._tr is not None:
            self._native(lib().xg_tran_dcop_continuation(self._tr, modes[continuation]),
                         "DC operating point")
            return
        if continuation != "none":
            raise RuntimeError("DC continuation needs the native transient driver")
        self.x.zero_()
        self.sta_c.zero_()
        self.sto_c.zero_()
        self._newton(DC_PARAMS, transient=False, order=1, dt=0.0, t=0.0, dcop=True,
                     wa=self.x_curr)
        # the converged point starts the history (N_TIA_OneStep::initialize)
        self.x_curr.copy_(self.x)
        self.sta_c.copy_(self.sta_n)
        self.sto_c.copy_(self.sto_n)
        self.qn0.copy_(self.Q)
        self.fprev.copy_(self.F)
        self.xh1.zero_()
        self.xh2.zero_()
        self.t = 0.0
        self.n_steps = 0

    def step(self, dt):
        """One ONESTEP time step of size dt (retried with dt/2 on failure)."""
        if self._tr is not None:
            self._native(lib().xg_tran_step(self._tr, dt), "time step")
            return
        order = 2 if self.n_steps >= 2 else 1
        t_next = self.t + dt
        # predictor xn0 = x_n + beta1 (x_n - x_{n-1}) [+ beta2 (x_{n-1} - x_{n-2})],
        # constant-step coefficients: order 1 beta1 = 1; order 2 beta1 = 1.5,
        # beta2 = -0.5
        b1, b2 = (1.5, -0.5) if order == 2 else (1.0, 0.0)
        check(lib().xg_tia_predict(self.dev.h, order, b1, b2, _ptr(self.x_curr),
                                   _ptr(self.xh1), _ptr(self.xh2), _ptr(self.xpred)),
              "predict")
        self.x.copy_(self.xpred)
        try:
            self._newton(TRAN_PARAMS, transient=True, order=order, dt=dt, t=t_next,
                         first_tran=self.n_steps == 0, wa=self.x_curr, wb=self.xpred)
        except NewtonFailure:
            if dt < 1e-18:
                raise
            self.step_failures += 1
            self.x.copy_(self.x_curr)
            self.step(dt / 2)
            self.step(dt / 2)
            return
        # completeStep: history, states
        self.xh2.copy_(self.xh1)
        self.xh1.copy_(self.x).sub_(self.x_curr)
        self.x_curr.copy_(self.x)
        self.qn0.copy_(self.Q)
        self.fprev.copy_(self.F)
        self.sta_c.copy_(self.sta_n)
        self.sto_c.copy_(self.sto_n)
        self.t = t_next
        self.n_steps += 1

    def _probe_ids(self, probes):
        """Solution ids of the probed nodes, checked on the host before any
        device read: unknown names raise KeyError, ground ValueError."""
        ids = [self.nl.probe_id(p) for p in probes]
        n = self.nl.n_nodes
        for p, i in zip(probes, ids):
            if not 0 <= i < n:
                raise KeyError("probe %r maps to id %d outside 0..%d" % (p, i, n - 1))
        return ids

    def run(self, tstop, dt, probes, every=1, uic=False, dc="none", progress=None):
        """DC operating point (or, uic=True, the zero start; dc selects the
        operating point's continuation, see dcop), then steps of dt to
        tstop.  Returns (times, waveforms[probe, time]) for the given node
        names.  progress(step, n_steps), if given, is called every 50
        steps."""
        torch = self.torch
        if self._tr is not None:
            return self._run_native(tstop, dt, probes, every, uic, dc, progress)
        idx = torch.tensor(self._probe_ids(probes), device=self.x.device)
        self.uic() if uic else self.dcop()
        n = int(round(tstop / dt))
        out = torch.zeros(len(probes), n // every + 1, dtype=torch.float64,
                          device=self.x.device)
        times = [0.0]
        out[:, 0] = self.x_curr[idx]
        k = 1
        for s in range(1, n + 1):
            self.step(dt)
            if s % every == 0:
                out[:, k] = self.x_curr[idx]
                times.append(self.t)
                k += 1
        return np.array(times), out[:, :k].cpu().numpy()

    def _run_native(self, tstop, dt, probes, every, uic=False, dc="none", progress=None):
        torch = self.torch
        L = lib()
        n = int(round(tstop / dt))
        with DEVICE_LOCK:
            ids = torch.tensor(self._probe_ids(probes), dtype=torch.int32,
                               device=self.x.device)
            out = torch.zeros(n // every + 1, len(probes), dtype=torch.float64,
                              device=self.x.device)
            torch.cuda.current_stream().synchronize()
        self.uic() if uic else self.dcop(dc)
        check(L.xg_tran_probe(self._tr, ids.data_ptr(), len(probes), out[0].data_ptr()),
              "probe")
        # the steps in calls of 50 when progress is reported, else one call
        rec = n // every
        times = np.zeros(rec + 1)
        row1 = out.data_ptr() + out.element_size() * len(probes)
        chunk = 50 if progress is not None else max(n, 1)
        s = 0
        while s < n:
            m = min(chunk, n - s)
            rc = L.xg_tran_run(self._tr, dt, m, ids.data_ptr(), len(probes), every, s, row1,
                               times[1:].ctypes.data)
            if rc:
                self._native(rc, "time step")
            s += m
            if progress is not None and s % 50 == 0:
                progress(s, n)
        check(L.xg_tran_sync(self._tr), "sync")
        self._sync_counters()
        with DEVICE_LOCK:
            w = out.T.cpu().numpy()
        return times, w

    def run_adaptive(self, tstop, tstep, probes, uic=False, reltol=1e-3, abstol=1e-6,
                     max_points=1_000_000, dc="none", progress=None):
        """DC operating point (or the zero start), then the reference's
        variable-step ONESTEP to tstop: step size and order from the local
        error estimate, steps rejected above it, the PULSE corners as
        breakpoints at which the integration restarts (N_TIA_StepErrorControl,
        N_TIA_OneStep completeStep / rejectStep / initialize; .TRAN TSTEP
        TSTOP with RELTOL / ABSTOL of the time integrator).  Returns (times,
        waveforms[probe, accepted step], counts) where counts holds the
        accepted steps, the steps rejected by the error test and by Newton
        failures.  progress(steps, t), if given, is called every 50
        accepted steps.  Needs the native driver."""
        if self._tr is None:
            raise RuntimeError("run_adaptive needs the native transient driver")
        torch = self.torch
        L = lib()
        with DEVICE_LOCK:
            ids = torch.tensor(self._probe_ids(probes), dtype=torch.int32,
                               device=self.x.device)
            buf = torch.zeros(4096, len(probes), dtype=torch.float64, device=self.x.device)
            torch.cuda.current_stream().synchronize()
        self.uic() if uic else self.dcop(dc)
        bps = np.ascontiguousarray(self.nl.breakpoints(tstop), dtype=np.float64)
        self._native(L.xg_tran_adaptive_begin(self._tr, tstop, tstep, bps.ctypes.data,
                                              len(bps), reltol, abstol), "adaptive begin")
        times, chunks = [0.0], []
        check(L.xg_tran_probe(self._tr, ids.data_ptr(), len(probes), buf[0].data_ptr()),
              "probe")
        k = 1
        tnow = ctypes.c_double()
        while abs(times[-1] - tstop) >= 1e-10 * tstop:
            if len(times) >= max_points:
                raise RuntimeError("more than %d accepted steps" % max_points)
            rc = L.xg_tran_adaptive_step(self._tr, ctypes.byref(tnow))
            if rc:
                self._native(rc, "adaptive step")
            if k == buf.shape[0]:
                with DEVICE_LOCK:
                    chunks.append(buf.cpu().numpy())
                k = 0
            check(L.xg_tran_probe(self._tr, ids.data_ptr(), len(probes), buf[k].data_ptr()),
                  "probe")
            times.append(tnow.value)
            k += 1
            if progress is not None and len(times) % 50 == 0:
                progress(len(times) - 1, tnow.value)
        check(L.xg_tran_sync(self._tr), "sync")
        self._sync_counters()
        with DEVICE_LOCK:
            chunks.append(buf[:k].cpu().numpy())
        counts = np.zeros(4, np.int64)
        check(L.xg_tran_adaptive_info(self._tr, None, counts.ctypes.data), "adaptive info")
        return np.array(times), np.concatenate(chunks).T, dict(
            accepted=int(counts[0]), rejected_error=int(counts[1]),
            rejected_newton=int(counts[2]))

    def set_eviction(self, on=True, min_dt=1e-18):
        """Batches of independent circuits: a step whose Newton solve still
        fails at min_dt takes the sub-circuits that did not converge out of
        the solve (frozen at their last accepted values) instead of failing
        the run.  Needs the native driver."""
        if self._tr is None:
            raise RuntimeError("eviction needs the native transient driver")
        check(lib().xg_tran_set_eviction(self._tr, 1 if on else 0, min_dt), "eviction")

    def evicted_components(self):
        """Ids of the evicted sub-circuits (components of xg_components)."""
        import ctypes as C
        nc = C.c_int32()
        check(lib().xg_components(self.dev.h, C.byref(nc), None), "components")
        flags = np.zeros(max(nc.value, 1), np.int32)
        n = C.c_int64()
        check(lib().xg_tran_evicted(self._tr, flags.ctypes.data, C.byref(n)), "evicted")
        return [int(c) for c in np.flatnonzero(flags[:nc.value])]

    # ---- checkpoint / restart (N_IO_RestartMgr's role for a transient)
    _CKPT = ("x_curr", "xh1", "xh2", "sta_c", "sto_c", "qn0", "fprev")

    def checkpoint(self, path):
        """Writes the accepted state (solution and its history, state and
        store vectors, charge and F history, device-held instance state,
        time and counters) to an .npz file."""
        if self._tr is not None:
            check(lib().xg_tran_sync(self._tr), "sync")
            self._sync_counters()
        self.torch.cuda.synchronize()
        n = ctypes.c_int64()
        check(lib().xg_instance_state(self.dev.h, 0, None, ctypes.byref(n)), "instance state")
        inst = np.zeros(max(n.value, 1))
        check(lib().xg_instance_state(self.dev.h, 1, inst.ctypes.data, ctypes.byref(n)),
              "instance state")
        arrays = {k: getattr(self, k).cpu().numpy() for k in self._CKPT}
        np.savez(path, instance=inst[:n.value], t=self.t, n_steps=self.n_steps,
                 newton_total=self.newton_total, step_failures=self.step_failures,
                 sizes=np.array([self.dev.nnz, len(self.x), len(self.sta_c), len(self.sto_c)]),
                 **arrays)

    def restore(self, path):
        """Loads a checkpoint of the same circuit; the next step continues
        exactly where the checkpointed run was."""
        d = np.load(path)
        sizes = [self.dev.nnz, len(self.x), len(self.sta_c), len(self.sto_c)]
        if list(d["sizes"]) != sizes:
            raise ValueError("checkpoint is of another circuit")
        for k in self._CKPT:
            getattr(self, k).copy_(self.torch.from_numpy(d[k]))
        inst = np.ascontiguousarray(d["instance"], dtype=np.float64)
        n = ctypes.c_int64(len(inst))
        if len(inst):
            check(lib().xg_instance_state(self.dev.h, 2, inst.ctypes.data, ctypes.byref(n)),
                  "instance state")
        self.t, self.n_steps = float(d["t"]), int(d["n_steps"])
        self.newton_total, self.step_failures = int(d["newton_total"]), int(d["step_failures"])
        self.torch.cuda.synchronize()
        if self._tr is not None:
            check(lib().xg_tran_restore(self._tr, self.t, self.n_steps, self.newton_total,
                                        self.step_failures), "restore")
