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

Time steps are fixed (the reference's LTE step control is not restated);
a failed Newton solve halves the step, as the product does.
"""
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

    def _load(self, it, transient, first_tran=False, dcop=False, init_jct=False, t=0.0):
        self.ref.set_solver_state(dcop=int(dcop), tranop=int(dcop), initJct=int(init_jct),
                                  initTran=int(first_tran), newtonIter=it,
                                  transient=int(transient))
        self.ref.set_time(t, 0.0)
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
                wa=None, wb=None):
        self._load(0, transient, first_tran, dcop, init_jct=dcop, t=t)
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

    def dcop(self):
        self.x[:] = 0.0
        self.sta_c[:] = 0.0
        self.sto_c[:] = 0.0
        self._newton(DC_PARAMS, False, 1, 0.0, 0.0, dcop=True, wa=self.x_curr)
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
