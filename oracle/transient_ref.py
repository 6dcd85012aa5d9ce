"""ORACLE — test infrastructure only; the product never imports this module.

CPU restatement of a transient analysis as the reference runs it, over the
oracle's device load (RefLoad):

* DC operating point (tranop): Newton on F with the DC parameters;
* ONESTEP steps (N_TIA_OneStep.C): order 1 for the first two steps, order 2
  (trapezoidal) after (completeStep, :1945-2010); predictor (:115-133) with
  the constant-step coefficients of updateCoeffs (:1355-1410); residual and
  Jacobian of obtainResidual / obtainJacobian (:215-330) with the same
  operation order; history update (updateHistory, :1209-1250);
* Newton (N_NLS_DampedNewton.C): full steps, weights reltol * max(|x_old|,
  |x_new|) + abstol (:311-335), convergence test converged_ (:1787-1830) with
  the DC and transient parameters of N_NLS_NLParams.C:108-125;
* linear solve: LAPACK LU with partial pivoting (numpy.linalg.solve) per
  connected component of the circuit graph without ground.

`run` takes fixed steps (a failed Newton solve halves the step, as the
product does); `run_adaptive` restates the reference's variable-step
ONESTEP control:

* step size and order from the local error estimate: the error-weight
  vectors (N_TIA_DataStore.C:1174-1330 with newLte: one weight
  reltol * max|x_n| + abstol for every unknown, reltol * |Q_i| + abstol per
  Q entry), the Newton corrections x - x_pred and Q - Q_n
  (stepLinearCombo, :1782-1800), the estimate
  ck * sqrt((|dx|_w^2 + |dQ|_w^2) / 2) (WRMS_errorNorm, :1377-1420;
  N_TIA_OneStep.h:97) and the acceptance test estimate <= 1
  (N_TIA_StepErrorControl.C:668-760; not on a step that begins the
  integration);
* updateCoeffs (N_TIA_OneStep.C:1355-1410: psi, the variable-step
  predictor coefficients beta and ck), completeStep (:1945-2190: ratio
  (0.5 / (est + 1e-4))^(1/(order+1)), doubling at >= 2, [0.25, 0.9] below
  1, order 2 from the second step when the ratio exceeds 1.05),
  rejectStep (:1695-1930: Newton failure dt/8, first error failure the
  clamped ratio, later ones 0.25, order back to 1, psi restored) and
  initialize (:1482-1640: first step min(TSTEP, 0.005 (stop - t0)), after
  a breakpoint 0.1 min(saved step, time to stop), order 1, constant
  history);
* breakpoints (the PULSE corners) as stop times
  (N_TIA_StepErrorControl.C:520-620, 916-1160), the integration restarted
  on each one reached (N_ANP_Transient.C:1039-1130), maximum step
  0.1 (tstop - t0) and minimum step t * 10 * eps.
"""
import math

import numpy as np

from .oracle import RefLoad

DC_PARAMS = dict(reltol=1e-3, abstol=1e-12, rhstol=1e-6, dxtol=1.0, maxit=200)
TRAN_PARAMS = dict(reltol=1e-2, abstol=1e-6, rhstol=1e-2, dxtol=0.33, maxit=20)
SMALL_UPDATE = 1e-6


class NewtonFailure(RuntimeError):
    pass


def components(rowptr, colidx, n):
    par = list(range(n))

    def find(a):
        while par[a] != a:
            par[a] = par[par[a]]
            a = par[a]
        return a
    for r in range(n):
        for k in range(rowptr[r], rowptr[r + 1]):
            c = colidx[k]
            if c < n:
                a, b = find(r), find(c)
                if a != b:
                    par[max(a, b)] = min(a, b)
    groups = {}
    for v in range(n):
        groups.setdefault(find(v), []).append(v)
    return [np.array(g) for g in sorted(groups.values(), key=lambda g: g[0])]


class RefTransient:
    def __init__(self, nl, gmin=1e-12, limiter=True):
        self.nl = nl.layout()
        self.ref = RefLoad(nl, gmin=gmin, limiter=limiter)
        self.limiter = limiter
        self.n = nl.n_nodes
        self.rowptr, self.colidx = self.ref.rowptr, self.ref.colidx
        self.comps = components(self.rowptr, self.colidx, self.n)
        self.rows = np.repeat(np.arange(self.n + 1), np.diff(self.rowptr))
        n1 = self.n + 1
        self.x = np.zeros(n1)
        self.x_curr = np.zeros(n1)
        self.xh1 = np.zeros(n1)
        self.xh2 = np.zeros(n1)
        self.sta_c, self.sta_n = np.zeros(nl.n_state), np.zeros(nl.n_state)
        self.sto_c, self.sto_n = np.zeros(nl.n_store), np.zeros(nl.n_store)
        self.qn0, self.fprev = np.zeros(n1), np.zeros(n1)
        self.t = 0.0
        self.n_steps = 0
        self.newton_total = 0
        self.bp_tol = 0.0   # breakpoint tolerance of the source evaluation
        self.gmin_g = 0.0   # gmin-stepping conductance (0: none)
        self.vnodes = np.array(nl.voltage_unknowns(), dtype=int)
        self.vdiag = np.array([self.rowptr[i] + int(np.searchsorted(
            self.colidx[self.rowptr[i]:self.rowptr[i + 1]], i)) for i in self.vnodes],
            dtype=int)
        self.gmin_steps = 0

    def _load(self, it, transient, first_tran=False, dcop=False, init_jct=False, t=0.0):
        self.ref.set_solver_state(dcop=int(dcop), tranop=int(dcop), initJct=int(init_jct),
                                  initTran=int(first_tran), newtonIter=it,
                                  transient=int(transient))
        self.ref.set_time(t, self.bp_tol)
        (self.F, self.Q, self.FL, self.QL, self.JF, self.JQ) = self.ref.full_load(
            self.x, self.sta_c, self.sto_c, self.sta_n, self.sto_n)
        lim, _ = self.ref.mos1_flags()
        n_lim = int(lim.sum())
        if self.ref.b3:
            l3, _ = self.ref.b3.flags()
            n_lim += int(l3.sum())
        n_lim += sum(0 if d.orig else 1 for d in self.ref.lin.diodes)
        self.n_lim = n_lim

    def _assemble(self, transient, order, dt):
        F, Q, FL, QL = self.F, self.Q, self.FL, self.QL
        if transient:
            inv_dt = 1.0 / dt
            r = 1.0 * Q + -1.0 * self.qn0
            if order == 2:
                r = inv_dt * r + 0.5 * F
                r = 1.0 * r + 0.5 * self.fprev
            else:
                r = inv_dt * r + 1.0 * F
            r = -r
            if self.limiter:
                r = r + 1.0 * (QL * inv_dt)
                r = r + (0.5 if order == 2 else 1.0) * FL
            J = inv_dt * self.JQ + (0.5 if order == 2 else 1.0) * self.JF
        else:
            r = -F
            if self.limiter:
                r = r + 1.0 * FL
            J = self.JF.copy()
        if self.gmin_g:
            # gmin stepping: a conductance from every voltage unknown to
            # ground, in the residual and on the Jacobian diagonal
            # (N_NLS_NOX_AugmentLinSys_GStepping.C:augmentResidual /
            # augmentJacobian)
            v = self.vnodes
            r = r.copy()
            r[v] = r[v] - self.gmin_g * self.x[v]
            J = J.copy()
            J[self.vdiag] = J[self.vdiag] + self.gmin_g
        r[self.n] = 0.0
        self.rhs, self.J = r, J

    def _solve(self, prm, wa, wb):
        n = self.n
        dense_rows, cols, vals = self.rows, self.colidx, self.J
        upd = 0.0
        x_new = self.x.copy()
        for comp in self.comps:
            loc = -np.ones(n + 1, int)
            loc[comp] = np.arange(len(comp))
            A = np.zeros((len(comp), len(comp)))
            mask = (loc[dense_rows] >= 0) & (loc[cols] >= 0)
            A[loc[dense_rows[mask]], loc[cols[mask]]] = vals[mask]
            dx = np.linalg.solve(A, self.rhs[comp])
            w = prm["reltol"] * np.maximum(np.abs(wa[comp]), np.abs(wb[comp])) + prm["abstol"]
            upd = max(upd, float(np.sqrt(np.mean((dx / w) ** 2))))
            x_new[comp] = self.x[comp] + dx
        self.x = x_new
        return upd

    def _newton(self, prm, transient, order, dt, t, first_tran=False, dcop=False,
                wa=None, wb=None, init_jct=None):
        self._load(0, transient, first_tran, dcop,
                   init_jct=dcop if init_jct is None else init_jct, t=t)
        self._assemble(transient, order, dt)
        it = 0
        while True:
            if dcop:
                wb = self.x.copy()
            upd = self._solve(prm, wa, wb)
            it += 1
            self._load(it, transient, first_tran, dcop, t=t)
            self._assemble(transient, order, dt)
            max_rhs = max(float(np.abs(self.rhs[c]).max()) for c in self.comps)
            devices_ok = self.n_lim == 0 or it >= prm["maxit"]
            if devices_ok and max_rhs <= prm["rhstol"] and upd <= prm["dxtol"]:
                break
            if devices_ok and upd <= SMALL_UPDATE:
                break
            if it >= prm["maxit"]:
                raise NewtonFailure("no convergence at t=%g" % t)
        self.newton_total += it
        return it

    def _gmin_stepping(self):
        """DC operating point by gmin stepping (.options nonlin
        continuation=gmin, N_NLS_NOX_Interface.C:894-960): natural
        continuation in p from 4 to -4 with the conductance 10^p - 10^-4 from
        every voltage unknown to ground; first step -2, grown by
        1 + 0.01 ((20 - its) / 20)^2 after a converged solve and halved after
        a failed one (N_NLS_LOCA_StepSizeControl.C), the last one clipped to
        land on -4 (conductance 0); constant predictor; at most 400 steps."""
        p, step, p_end = 4.0, -2.0, -4.0
        x_prev = self.x.copy()
        first, n = True, 0
        while True:
            p_try = p if first else max(p + step, p_end)
            self.gmin_g = 10.0 ** p_try - 10.0 ** p_end + 0.0
            self.x = x_prev.copy()
            try:
                its = self._newton(DC_PARAMS, False, 1, 0.0, 0.0, dcop=True,
                                   init_jct=first, wa=self.x_curr)
                ok = True
            except NewtonFailure:
                ok = False
            n += 1
            if first and not ok:
                self.gmin_g = 0.0
                raise NewtonFailure("gmin stepping: no solution at the first conductance")
            if ok:
                p, x_prev = p_try, self.x.copy()
                if not first:
                    f = (20.0 - its) / 20.0
                    step *= 1.0 + 0.01 * f * f
                    step = max(step, -1e12)
                first = False
                if p == p_end:
                    break
            else:
                self.x = x_prev.copy()
                step *= 0.5
                if abs(step) < 1e-6:
                    self.gmin_g = 0.0
                    raise NewtonFailure("gmin stepping: step below the minimum at p=%g" % p)
            if n >= 400:
                self.gmin_g = 0.0
                raise NewtonFailure("gmin stepping: 400 continuation steps")
        self.gmin_g = 0.0
        self.gmin_steps = n

    def dcop(self, continuation="none"):
        """DC operating point: Newton from zero; continuation="gmin" solves
        by gmin stepping instead, "auto" falls back to it when Newton
        fails."""
        self.x[:] = 0.0
        self.sta_c[:] = 0.0
        self.sto_c[:] = 0.0
        if continuation == "gmin":
            self._gmin_stepping()
        else:
            try:
                self._newton(DC_PARAMS, False, 1, 0.0, 0.0, dcop=True, wa=self.x_curr)
            except NewtonFailure:
                if continuation != "auto":
                    raise
                self.x[:] = 0.0
                self._gmin_stepping()
        self.x_curr = self.x.copy()
        self.sta_c[:] = self.sta_n
        self.sto_c[:] = self.sto_n
        self.qn0 = self.Q.copy()
        self.fprev = self.F.copy()
        self.xh1[:] = 0.0
        self.xh2[:] = 0.0
        self.t = 0.0
        self.n_steps = 0

    def uic(self):
        """Zero start without an operating point: one load at x = 0 with
        operating-point flags (no junction initialisation) starts the
        history."""
        self.x[:] = 0.0
        self.sta_c[:] = 0.0
        self.sto_c[:] = 0.0
        self._load(1, False, dcop=True, t=0.0)
        self.x_curr = self.x.copy()
        self.sta_c[:] = self.sta_n
        self.sto_c[:] = self.sto_n
        self.qn0 = self.Q.copy()
        self.fprev = self.F.copy()
        self.xh1[:] = 0.0
        self.xh2[:] = 0.0
        self.t = 0.0
        self.n_steps = 0

    def step(self, dt):
        order = 2 if self.n_steps >= 2 else 1
        t_next = self.t + dt
        if order == 2:
            xpred = self.x_curr + 1.5 * self.xh1 + -0.5 * self.xh2
        else:
            xpred = self.x_curr + 1.0 * self.xh1
        self.x = xpred.copy()
        try:
            self._newton(TRAN_PARAMS, True, order, dt, t_next,
                         first_tran=self.n_steps == 0, wa=self.x_curr, wb=xpred)
        except NewtonFailure:
            if dt < 1e-18:
                raise
            self.x = self.x_curr.copy()
            self.step(dt / 2)
            self.step(dt / 2)
            return
        self.xh2 = self.xh1.copy()
        self.xh1 = self.x - self.x_curr
        self.x_curr = self.x.copy()
        self.qn0 = self.Q.copy()
        self.fprev = self.F.copy()
        self.sta_c[:] = self.sta_n
        self.sto_c[:] = self.sto_n
        self.t = t_next
        self.n_steps += 1

    def run(self, tstop, dt, probes, every=1, uic=False):
        idx = [self.nl.node(p) for p in probes]
        self.uic() if uic else self.dcop()
        n = int(round(tstop / dt))
        times, out = [0.0], [self.x_curr[idx].copy()]
        for s in range(1, n + 1):
            self.step(dt)
            if s % every == 0:
                times.append(self.t)
                out.append(self.x_curr[idx].copy())
        return np.array(times), np.array(out).T


class StepStats:
    def __init__(self):
        self.accepted = 0
        self.rejected_error = 0
        self.rejected_newton = 0


def run_adaptive(tr, tstop, tstep, probes, uic=False, reltol=1e-3, abstol=1e-6):
    """Variable-step ONESTEP transient of a RefTransient (see the module
    docstring).  Returns (times, waveforms, stats); a waveform column per
    accepted step, the DC or zero start first."""
    eps = np.finfo(float).eps
    n = tr.n
    idx = [tr.nl.node(p) for p in probes]
    tr.uic() if uic else tr.dcop()
    bps = [0.0] + tr.nl.breakpoints(tstop) + [tstop]
    st = StepStats()
    t0 = 0.0
    cur = 0.0                    # currentTime
    stop = 0.0                   # stopTime
    h = tstep                    # currentTimeStep
    last_h = h                   # lastTimeStep
    h_sum = 2.0 * h              # currentTimeStepSum
    saved_h = 1e-10              # savedTimeStep
    min_h, max_h = 0.0, 0.1 * (tstop - t0)
    bp_tol = 0.0
    nxt = 0.0                    # nextTime
    psi = [0.0, 0.0, 0.0]
    order, n_steps, nef = 1, 0, 0
    est = 0.0
    step_number = 0              # accepted transient steps
    beginning = True
    attempt_ok = True
    newton_ok = True
    qw = tr.Q.copy()             # Q of the last load, for the Q weights
    times, out = [0.0], [tr.x_curr[idx].copy()]
    xh2 = np.zeros_like(tr.x_curr)
    while abs(cur - tstop) >= 1e-10 * (tstop - t0):
        # updateStopTime
        old_stop = stop
        stop = min(tstop, next(b for b in bps if b > cur) if any(b > cur for b in bps)
                   else tstop)
        if abs(stop - old_stop) < bp_tol and beginning and stop != tstop:
            later = [b for b in bps if b > stop]
            stop = later[0] if later else tstop
        if beginning and attempt_ok:
            # initialize
            to_stop = stop - cur
            if cur == t0:
                hh = min(0.005 * abs(to_stop), tstep)
                h = hh
            else:
                hh = 0.1 * min(saved_h, abs(to_stop))
                h = min(h, hh)
            h_sum = 2.0 * h
            last_h = h
            nxt = cur + h
            tr.xh1[:] = 0.0
            xh2[:] = 0.0
            n_steps = 0
            order = 1
            psi[0] = h
        # updateCoeffs
        if order == 2:
            psi[2] = psi[1]
        psi[1] = psi[0]
        psi[0] = h
        if order == 2:
            t2 = psi[1]
            beta1 = h / t2 + (h / t2) * (h / t2) / 2
            beta2 = -1.0 * h * h / t2 / psi[2] / 2
            ck = (h / h_sum) / 3.0
        else:
            beta1 = h / psi[1]
            beta2 = 0.0
            ck = h / h_sum
        # predictor, error weights, Newton
        wx = reltol * float(np.abs(tr.x_curr[:n]).max(initial=0.0)) + abstol
        wq = reltol * np.abs(qw[:n]) + abstol
        xpred = tr.x_curr + beta1 * tr.xh1
        if order == 2:
            xpred = xpred + beta2 * xh2
        tr.x = xpred.copy()
        tr.bp_tol = bp_tol   # PulseData::updateSource's corner tolerance
        try:
            tr._newton(TRAN_PARAMS, True, order, h, nxt, first_tran=step_number <= 0,
                       wa=tr.x_curr, wb=xpred)
            newton_ok = True
        except NewtonFailure:
            newton_ok = False
        qw = tr.Q.copy()
        attempt_ok = newton_ok
        if newton_ok and step_number >= 1 and not beginning:
            dx = (tr.x[:n] - xpred[:n]) / wx
            dq = (tr.Q[:n] - tr.qn0[:n]) / wq
            ex = math.sqrt(float(np.mean(dx * dx))) if n else 0.0
            eq = math.sqrt(float(np.mean(dq * dq))) if n else 0.0
            est = ck * math.sqrt(0.5 * ex * ex + 0.5 * eq * eq)
            attempt_ok = est <= 1.0
        if attempt_ok:
            # processSuccessfulStep: min step, breakpoint tolerance, completeStep
            min_h = cur * 10.0 * eps
            bp_tol = 2.0 * min_h
            n_steps += 1
            nef = 0
            cur = nxt
            new_h = h
            last_h = h
            rr = (0.5 / (est + 0.0001)) ** (1.0 / (order + 1.0))
            if n_steps >= 2 and order == 1:
                order = 2
                rr = (0.5 / (est + 0.0001)) ** (1.0 / (order + 1.0))
                if rr <= 1.05:
                    order = 1
            if rr >= 2.0:
                new_h = 2.0 * h
            elif rr <= 1.0:
                new_h = max(0.25, min(0.9, rr)) * h
            # updateHistory (xHistory[2] moves only at order 2, where it is read)
            if order == 2:
                xh2[:] = tr.xh1
            tr.xh1 = tr.x - tr.x_curr
            tr.x_curr = tr.x.copy()
            tr.qn0 = tr.Q.copy()
            tr.fprev = tr.F.copy()
            tr.sta_c[:] = tr.sta_n
            tr.sto_c[:] = tr.sto_n
            new_h = min(max(new_h, min_h), max_h)
            if stop - cur >= min_h:
                npt = cur + new_h
                if npt > stop:
                    saved_h = new_h
                    npt = stop
                    new_h = stop - cur
                nxt = npt
                h_sum = new_h + last_h
            h = new_h
            step_number += 1
            st.accepted += 1
            beginning = abs(cur - stop) <= bp_tol and abs(cur - tstop) > bp_tol
            times.append(cur)
            out.append(tr.x_curr[idx].copy())
        else:
            # processFailedStep / rejectStep
            nef += 1
            for i in range(1, order + 1):
                psi[i - 1] = psi[i]
            if nef >= 15:
                raise NewtonFailure("too many local error test failures at t=%g" % cur)
            if not newton_ok:
                new_h = h / 8
                st.rejected_newton += 1
            elif nef == 1:
                rr = (0.5 / (est + 0.0001)) ** (1.0 / (order + 1.0))
                new_h = max(0.25, min(0.9, rr)) * h
                st.rejected_error += 1
            else:
                new_h = 0.25 * h
                st.rejected_error += 1
            order = 1
            new_h = min(max(new_h, min_h), max_h)
            npt = cur + new_h
            if npt > stop:
                npt = stop
                new_h = stop - cur
            nxt = npt
            h_sum = new_h + last_h
            h = new_h
            if h <= min_h:
                raise NewtonFailure("time step too small at t=%g" % cur)
    tr.t = cur
    return np.array(times), np.array(out).T, st
