"""In-process outer loops over prescribed-cilia runs: flow maps, concurrent ensembles,
periodic-orbit search by Newton-Krylov with hooksteps, Floquet exponents of an orbit
(Arnoldi on the linearised one-period map, or the whole monodromy matrix from one
concurrent ensemble), and bisection between states.

The reference runs these loops in Python around its executable: every evaluation of
the flow map writes psi.dat and globals.ini, launches the simulation
(pyfile/driver/func.py:544) and reads ``_true_states.dat`` back
(find_periodic/newton_solver.py:49-74, bisection/bisection.py:150-162).  Here each
evaluation is a call into the library on the GPU; several evaluations run at once on
separate contexts (one HIP stream each, the Python threads release the interpreter
while the library works).

State vectors follow the reference's Newton layout ``x = [T, phases (NFIL), angles
(NFIL)]`` (angles only with shape-angle dynamics); the flow map integrates ``ndts``
steps of length ``T / steps_per_period``.
"""
from __future__ import annotations

import math
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from .cilia import CiliaEnsemble, PrescribedCilia


def box(x, period):
    """x reduced to [0, period) (find_periodic/util.py)."""
    return np.asarray(x, float) - np.floor(np.asarray(x, float) / period) * period


def order_parameter(phases):
    """Kuramoto order parameter |sum exp(i phase)| / NFIL (bisection.py:135-141)."""
    z = np.exp(1j * np.asarray(phases, float))
    return float(np.abs(z.sum()) / z.size)


def interpolate_state(left, right, alpha, nfil):
    """Initial condition between two states (bisection.py:150-155): phases on the circle,
    the remaining entries linearly, weight alpha on the left state."""
    left, right = np.asarray(left, float), np.asarray(right, float)
    out = alpha * left + (1.0 - alpha) * right
    out[:nfil] = np.arctan2(alpha * np.sin(left[:nfil]) + (1 - alpha) * np.sin(right[:nfil]),
                            alpha * np.cos(left[:nfil]) + (1 - alpha) * np.cos(right[:nfil]))
    out[:nfil] = box(out[:nfil], 2.0 * math.pi)
    return out


class CiliaFlow:
    """Flow map of a lattice of prescribed cilia with dynamic phases (and shape angles)
    driven by recorded generalised forces (the reference's DYNAMIC_PHASE_EVOLUTION runs
    its drivers launch).  base: 3 n cilia bases; q_phase / q_angle: driving tables;
    spring: the torsional spring of the shape angle (the drivers' k); s_tol: tolerance of
    the segment-parameter bisection (0: the reference's 0.1 / nseg, which makes the flow
    map piecewise smooth on that scale -- a small value keeps finite-difference
    Jacobians meaningful)."""

    def __init__(self, base, q_phase, q_angle=None, *, spring=0.0, steps_per_period=300,
                 omega0=2.0 * math.pi, nseg=20, dl=2.6, precision="fp64", solve_mode="fp64",
                 rtol=1e-10, precond="diag", s_tol=0.0):
        self.base = np.asarray(base, float).reshape(-1)
        self.nfil = self.base.size // 3
        self.q_phase = np.asarray(q_phase, float)
        self.q_angle = None if q_angle is None else np.asarray(q_angle, float)
        self.angles = self.q_angle is not None
        self.spring = float(spring)
        self.spp = int(steps_per_period)
        self.omega0 = float(omega0)
        self.kw = dict(nseg=nseg, dl=dl, precision=precision, solve_mode=solve_mode, rtol=rtol,
                       precond=precond, s_tol=s_tol)

    def _split(self, x):
        x = np.asarray(x, float)
        n = self.nfil
        return float(x[0]), x[1:1 + n], (x[1 + n:1 + 2 * n] if self.angles else None)

    def run(self, x, ndts, record=False):
        """Integrate ndts steps from state x; returns the final state (same layout, T kept)
        and, with record, the true-state rows (phases, angles) of every step."""
        T, ph, an = self._split(x)
        dt = T / self.spp
        n = self.nfil
        rows = []
        with PrescribedCilia(self.base, ph, np.full(n, self.omega0), dt=dt, **self.kw) as C:
            if an is not None:
                C.set_phases(angle=an)
            C.set_driving(self.q_phase, self.q_angle, self.spring)
            C.step(0)          # solve at the initial state (the first step does not advance)
            for nt in range(1, ndts + 1):
                C.step(nt)
                if record:
                    sp = C.speeds()
                    rows.append(np.concatenate([sp["phase"], sp["angle"]] if self.angles else [sp["phase"]]))
            sp = C.speeds()
        y = np.concatenate([[T], sp["phase"]] + ([sp["angle"]] if self.angles else []))
        return (y, np.array(rows)) if record else y

    def ensemble(self, xs, ndts, workers=4, batched=True):
        """Flow maps of several states at once.  batched: every member in one
        CiliaEnsemble (one product launch and one lockstep GMRES per step for all of them;
        FP64 pair sums and solve); otherwise one context and stream per member on
        `workers` host threads."""
        xs = [np.asarray(x, float) for x in xs]
        kw = getattr(self, "kw", None)   # flows that override run() step on their own
        fp64 = kw is not None and kw["precision"] == "fp64" and kw["solve_mode"] == "fp64"
        if not batched or not fp64 or not xs:
            with ThreadPoolExecutor(max_workers=max(1, int(workers))) as ex:
                return list(ex.map(lambda x: self.run(x, ndts), xs))
        E, n = len(xs), self.nfil
        parts = [self._split(x) for x in xs]
        T = np.array([p[0] for p in parts])
        ph = np.stack([p[1] for p in parts])
        an = np.stack([p[2] for p in parts]) if self.angles else None
        with CiliaEnsemble(self.base, E, self.q_phase, self.q_angle, spring=self.spring,
                           nseg=kw["nseg"], dl=kw["dl"], precond=kw["precond"], rtol=kw["rtol"],
                           s_tol=kw["s_tol"]) as C:
            C.set_state(ph, np.full((E, n), self.omega0), T / self.spp, angle=an)
            C.step(0)          # solve at the initial states (the first step does not advance)
            for nt in range(1, ndts + 1):
                C.step(nt)
            sp = C.speeds()
        return [np.concatenate([[T[e]], sp["phase"][e]] + ([sp["angle"][e]] if self.angles else []))
                for e in range(E)]

    # -- the periodic-orbit residual of find_periodic/newton_solver.py:80-126 --------
    def steporbit(self, ndts, x):
        y = np.zeros_like(np.asarray(x, float))
        y[1:] = self.run(x, ndts)[1:]
        return y

    def getrhs(self, x):
        """Phi_T(x) - x - 2 pi on the phases, zero in the period slot."""
        y = self.steporbit(self.spp, x) - np.asarray(x, float)
        y[1:self.nfil + 1] -= 2.0 * math.pi
        y[0] = 0.0
        return y


def _dot(n_, a, b):
    """Dot product; n_ == -1 leaves out the period slot (newton_solver.py:35-41)."""
    n1 = 1 if n_ == -1 else 0
    return float(np.sum(np.asarray(a)[n1:] * np.asarray(b)[n1:]))


class HookGMRES:
    """Restarted GMRES with a hookstep (trust-region) solution, after Viswanath (2008,
    arXiv:0809.1498) as find_periodic/gmresc.py uses it: the step y minimises the
    Arnoldi residual subject to |y| ~ del through a Levenberg parameter on the SVD of
    the Hessenberg matrix."""

    def __init__(self):
        self.h = self.v = None
        self.beta = 0.0
        self.j = 0

    def hook(self, delta):
        a = self.h[:self.j + 1, :self.j]
        u, s, vt = np.linalg.svd(a)
        p = self.beta * u[0, :self.j]
        mu = max(s[self.j - 1] ** 2 * 1e-6, 1e-99)
        qn = 1e99
        while qn > delta:
            mu *= 1.1
            q = p * s / (mu + s * s)
            qn = math.sqrt(float(np.sum(q * q)))
        y = vt.T @ q
        r = -(self.h[:self.j + 1, :self.j] @ y)
        r[0] += self.beta
        return y, math.sqrt(float(np.sum(r * r)))

    def solve(self, m, n, x, b, matvec, psolve, dot, res, delta, its, info):
        if info == 2:     # another hookstep on the last Krylov space
            y, delta = self.hook(delta)
            x = psolve(self.v[:, :self.j] @ y[:self.j])
            return x, res, delta, its, 0
        tol, imx, its = res, its, 0
        self.v = np.zeros((n, m + 1))
        while True:
            res_prev = 1e99
            stag = 1.0 - 1e-14
            self.beta = math.sqrt(dot(n, x, x))
            w = np.zeros(n) if self.beta == 0.0 else matvec(x)
            w = b - w
            self.beta = math.sqrt(dot(n, w, w))
            self.v[:, 0] = w / self.beta
            self.h = np.zeros((m + 1, m))
            for j in range(m):
                self.j = j + 1
                its += 1
                w = matvec(psolve(self.v[:, j]))
                for i in range(j + 1):
                    self.h[i, j] = dot(n, w, self.v[:, i])
                    w = w - self.h[i, j] * self.v[:, i]
                self.h[j + 1, j] = math.sqrt(dot(n, w, w))
                self.v[:, j + 1] = w / self.h[j + 1, j]
                p = np.zeros(self.j + 1)
                p[0] = self.beta
                hj = self.h[:self.j + 1, :self.j]
                y = np.linalg.pinv(hj) @ p
                r = -(hj @ y)
                r[0] += self.beta
                res = math.sqrt(float(np.sum(r * r)))
                done = res <= tol or its == imx or res > res_prev
                if done or self.j == m:
                    if delta > 0.0:
                        y, delta = self.hook(delta)
                    # the update uses all self.j Arnoldi vectors
                    x = x + psolve(self.v[:, :self.j] @ y[:self.j])
                    if its == imx:
                        info = 2
                    if res > res_prev:
                        info = 1
                    if res <= tol:
                        info = 0
                    if done:
                        return x, res, delta, its, info
                res_prev = res * stag


class NewtonHook:
    """Newton-Krylov with hooksteps for periodic orbits (find_periodic/newton_solver.py:
    128-334): unknowns x = [T, state]; F(x) = flow.getrhs(x); J dx by a finite
    difference of F; with the period free, the first row is the phase condition
    dx . x_dot = 0."""

    def __init__(self, flow: CiliaFlow, x0, eps_j=1e-5, fix_t=False):
        self.flow = flow
        self.x = np.array(x0, float)
        self.eps_j = eps_j
        self.fix_t = fix_t
        self.fx = None
        self.tol = None
        self.nits = 0
        self.gits = 0
        self.history = []

    def multJ(self, dx):
        e = math.sqrt(_dot(-1, dx, dx))
        e = self.eps_j * math.sqrt(_dot(-1, self.x, self.x)) / e
        y = (self.flow.getrhs(self.x + e * dx) - self.fx) / e
        if self.fix_t:
            y[0] = 0.0
        else:
            s = self.flow.steporbit(1, self.x)
            dt = self.x[0] / self.flow.spp
            y[0] = _dot(-1, (s - self.x) / dt, dx)
        return y

    def solve(self, m=20, gtol=1e-3, tol=1e-8, delta=-1.0, mndl=1e-20, mxdl=1e20, nits=20, info=0):
        """Returns (x, residual, status): 0 converged, 2 max iterations, 3 trust region
        collapsed or no valid update."""
        n = self.x.size
        nf = self.flow.nfil
        g = HookGMRES()
        self.x[1:nf + 1] = box(self.x[1:nf + 1], 2.0 * math.pi)
        self.fx = self.flow.getrhs(self.x)
        self.tol = math.sqrt(_dot(-1, self.fx, self.fx))
        dl = delta
        mxdl_ = mxdl
        if delta < 0.0:
            dl = self.tol / 10.0
            mxdl_ = 1e99
        ginfo = info
        x_, fx_, tol_ = self.x.copy(), self.fx.copy(), self.tol
        tol__ = 1e99
        x__ = fx__ = None
        del__ = dl
        self.history.append(self.tol)
        if self.tol < tol:
            return self.x, self.tol, 0
        while True:
            if dl < mndl:
                return self.x, self.tol, 3
            s = np.zeros(n)
            if ginfo != 2:
                self.gits = m
            if delta == 0.0:
                self.gits = 9999
            s, gres, gdel, self.gits, ginfo = g.solve(m, n, s, fx_, self.multJ, lambda z: z, _dot,
                                                      gtol * self.tol, dl, self.gits, ginfo)
            ginfo = info
            self.x = x_ - s
            self.x[1:nf + 1] = box(self.x[1:nf + 1], 2.0 * math.pi)
            self.fx = self.flow.getrhs(self.x)
            self.tol = math.sqrt(_dot(-1, self.fx, self.fx))
            snrm = math.sqrt(_dot(-1, s, s))
            ared = tol_ - self.tol
            pred = tol_ - gdel
            if delta == 0.0:
                pass
            elif self.tol > tol__:       # back to the better earlier step
                self.x, self.fx, self.tol, dl = x__, fx__, tol__, del__
            elif ared < 0.0:             # norm increased: smaller step
                dl = snrm * 0.5
                ginfo = 2
                if pred != 0.0 and ared / pred < -50:
                    return self.x, self.tol, 3
            elif ared / pred < 0.75:     # acceptable; try a smaller one too
                x__, fx__, tol__ = self.x.copy(), self.fx.copy(), self.tol
                del__ = snrm if ared / pred > 0.1 else snrm * 0.5
                dl = snrm * 0.7
                ginfo = 2
            elif snrm < dl * 0.9:        # good full Newton step
                dl = min(mxdl_, snrm * 2.0)
            elif dl < mxdl_ * 0.9:       # good: try a larger one
                x__, fx__, tol__, del__ = self.x.copy(), self.fx.copy(), self.tol, dl
                dl = min(mxdl_, snrm * 2.0)
                ginfo = 2
            if ginfo == 2:
                continue
            self.nits += 1
            self.history.append(self.tol)
            x_, fx_, tol_ = self.x.copy(), self.fx.copy(), self.tol
            tol__ = 1e99
            if self.tol < tol:
                return self.x, self.tol, 0
            if self.nits == nits:
                return self.x, self.tol, 2


def _exponents(mu, T):
    """log(mu) / T sorted by descending real part (arnoldi.py:85-90)."""
    ev = np.log(np.asarray(mu, dtype=complex)) / T
    idx = ev.argsort()[::-1]
    return ev[idx], idx


def floquet_arnoldi(flow, x_star, n_evals=7, eps=1e-2, tol=1e-3, seed=0, start=None):
    """Leading Floquet exponents of the periodic orbit x_star = [T, state] by Arnoldi
    on the linearised one-period map (arnoldi.py:53-141): A q = (Phi_T(x* + eps q) - x*
    - 2 pi on the phases) / eps, one flow map per Krylov vector, until the leading
    min(k+1, n_evals) exponents change by less than tol (and at least n_evals vectors).
    Returns (exponents, eigenvectors (state x n_evals), Krylov dimension).  The start
    vector is a seeded unit random vector (or start); the iteration stops at the
    state dimension."""
    x_star = np.asarray(x_star, float)
    T = float(x_star[0])
    u_star = x_star[1:]
    n = u_star.size
    nf = flow.nfil
    Q = np.zeros((n, n + 1))
    H = np.zeros((n + 1, n))
    if start is None:
        b = np.random.default_rng(seed).random(n) - 0.5
    else:
        b = np.asarray(start, float).copy()
    Q[:, 0] = b / np.linalg.norm(b)
    old = np.full(n_evals, 10.0, dtype=complex)
    diff = 2.0 * tol
    k = 0
    while (abs(diff) > tol or k < n_evals) and k < n:
        y = flow.run(np.concatenate([[T], u_star + eps * Q[:, k]]), flow.spp)
        U = y[1:].copy()
        U[:nf] -= 2.0 * math.pi
        U = (U - u_star) / eps
        for i in range(k + 1):                 # modified Gram-Schmidt (arnoldi.py:70-80)
            H[i, k] = Q[:, i] @ U
            U = U - H[i, k] * Q[:, i]
        H[k + 1, k] = np.linalg.norm(U)
        if H[k + 1, k] > 0.0:
            Q[:, k + 1] = U / H[k + 1, k]
        ev, _ = _exponents(np.linalg.eigvals(H[:k + 1, :k + 1]), T)
        m = min(k + 1, n_evals)
        diff = float(np.linalg.norm(ev[:m] - old[:m]))
        old[:m] = ev[:m]
        k += 1
        if H[k, k - 1] == 0.0:                 # invariant subspace: exact on it
            break
    mu, vec = np.linalg.eig(H[:k, :k])
    ev, idx = _exponents(mu, T)
    return ev[:n_evals], Q[:, :k] @ vec[:, idx[:n_evals]], k


def monodromy(flow, x_star, eps=1e-4, workers=8):
    """The whole linearised one-period map at x_star by central differences, its 2 n
    flow maps run as one ensemble (batched when the flow solves in FP64, else concurrent
    contexts on `workers` threads; the batched counterpart of the Arnoldi iteration for
    lattices small enough to form it).  Returns (matrix, exponents)."""
    x_star = np.asarray(x_star, float)
    T = float(x_star[0])
    n = x_star.size - 1
    xs = []
    for j in range(n):
        for sgn in (1.0, -1.0):
            x = x_star.copy()
            x[1 + j] += sgn * eps
            xs.append(x)
    ys = flow.ensemble(xs, flow.spp, workers=workers)
    M = np.zeros((n, n))
    for j in range(n):
        M[:, j] = (ys[2 * j][1:] - ys[2 * j + 1][1:]) / (2.0 * eps)
    ev, _ = _exponents(np.linalg.eigvals(M), T)
    return M, ev


def bisection(flow: CiliaFlow, left, right, ndts, classify, iters=8, workers=4, sections=2):
    """Edge tracking between two states (bisection/bisection.py): each round runs the
    interior points of the current interval of alpha (weight of the left state)
    concurrently and keeps the section where classify(final order-parameter history)
    changes from True to False.  classify(r_history) -> True for the left side."""
    nf = flow.nfil
    lo, hi = 0.0, 1.0
    T = float(left[0])
    for _ in range(iters):
        alphas = [lo + (hi - lo) * k / sections for k in range(sections + 1)]
        xs = [np.concatenate([[T], interpolate_state(left[1:], right[1:], a, nf)]) for a in alphas]

        def side(x):
            _, rows = flow.run(x, ndts, record=True)
            return bool(classify(np.array([order_parameter(r[:nf]) for r in rows])))

        with ThreadPoolExecutor(max_workers=max(1, int(workers))) as ex:
            sides = list(ex.map(side, xs))
        # alpha = 1 is the left state: the side flips somewhere in between
        for k in range(sections):
            if sides[k] != sides[k + 1]:
                lo, hi = alphas[k], alphas[k + 1]
                break
        else:
            break
    return lo, hi
