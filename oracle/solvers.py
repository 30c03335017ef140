"""FP64 numpy restatements of the reference Krylov solvers.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Each solver takes ``mv(cwhat, v)`` returning A v ('N') or conj(A conj(v))
('C') on compact vectors, the right-hand side b and an initial guess, and
returns (x, iterations, relative residual, matvec count).  They follow the
reference drivers line by line in their recurrences and reliable-residual
policy (ddscat_cuda_unified/cuda/ddscat_matvec_cuda.cu):
  pbcgs2  cu:459-480, cu:1168-1278
  qmrccg  cu:434-443, cu:1071-1103
  petrkp  cu:482-493, cu:1280-1357
Pinning: every one is checked against numpy.linalg.solve on dense operators
(tests/test_oracle_solvers.py); iteration counts are algorithm facts used
only as a loose band for the device solvers.
"""
from __future__ import annotations

import numpy as np

RELIABLE_PERIOD = 20
RELIABLE_GAP = 1e-3


def _dotc(a, b):
    return np.vdot(a, b)


def _reliable(mv, rhs, x, rrec, rhsnorm, rec_res, tol, nc):
    ax = mv("N", x)
    nc += 1
    tr = np.linalg.norm(rhs - ax)
    gapn = np.linalg.norm(rhs - ax - rrec)
    true_res = tr / rhsnorm
    gap = gapn / tr if tr > 0 else (0.0 if gapn == 0 else np.inf)
    conv = true_res <= tol
    restart = gap >= RELIABLE_GAP or (rec_res <= tol and not conv)
    return ax, true_res, conv, restart, nc


def pbcgs2(mv, b, x0=None, maxit=300, tol=1e-5):
    """ZBCG2 / BiCGStab(2) with Fokkema's convex update and reliable updates."""
    b = np.asarray(b, np.complex128)
    n = b.size
    xp = np.zeros(n, complex) if x0 is None else np.asarray(x0, np.complex128).copy()
    nc = 0
    if np.linalg.norm(xp) > 0:
        r0 = b - mv("N", xp)
        nc += 1
    else:
        r0 = b.copy()
    rr = r0.copy()
    bp = r0.copy()
    x = np.zeros(n, complex)
    r = [r0, np.zeros(n, complex), np.zeros(n, complex)]
    u = [np.zeros(n, complex) for _ in range(3)]
    rnrm0 = np.linalg.norm(r0)
    rnrm = mxnrmx = mxnrmr = rnrm0
    alpha, omega, sigma, rho0 = 0.0, 1.0, 1.0, 1.0
    resid = 1.0 if rnrm0 > 0 else 0.0
    it = 0
    while resid > tol and it < maxit:
        it += 1
        rho0 = -omega * rho0
        for k in (1, 2):
            rho1 = _dotc(rr, r[k - 1])
            if rho0 == 0:
                raise ZeroDivisionError("rho0 == 0")
            beta = alpha * (rho1 / rho0)
            rho0 = rho1
            for j in range(k):
                u[j] = r[j] - beta * u[j]
            u[k] = mv("N", u[k - 1]); nc += 1
            sigma = _dotc(rr, u[k])
            alpha = rho1 / sigma
            x = x + alpha * u[0]
            for j in range(k):
                r[j] = r[j] - alpha * u[j + 1]
            r[k] = mv("N", r[k - 1]); nc += 1
            rnrm = np.linalg.norm(r[0])
            mxnrmx = max(mxnrmx, rnrm); mxnrmr = max(mxnrmr, rnrm)
        Z = np.array([[_dotc(r[i], r[j]) for j in range(3)] for i in range(3)])
        y0 = np.array([-1.0, Z[1, 0] / Z[1, 1], 0.0], complex)
        yl = np.array([0.0, Z[1, 2] / Z[1, 1], -1.0], complex)
        q0 = np.vdot(y0, Z @ y0); ql = np.vdot(yl, Z @ yl)
        k0 = np.sqrt(abs(q0)); kl = np.sqrt(abs(ql))
        vr = np.vdot(yl, Z @ y0) / (k0 * kl)
        avr = abs(vr)
        hg = vr * max(avr, 0.7) / avr
        y0 = y0 - hg * (k0 / kl) * yl
        omega = y0[2]
        u[0] = u[0] - y0[1] * u[1] - y0[2] * u[2]
        x = x + y0[1] * r[0] + y0[2] * r[1]
        r[0] = r[0] - y0[1] * r[1] - y0[2] * r[2]
        rnrm = np.sqrt(abs(np.vdot(y0, Z @ y0)))
        mxnrmx = max(mxnrmx, rnrm); mxnrmr = max(mxnrmr, rnrm)
        delta = 1e-2
        xpdt = rnrm < delta * rnrm0 and rnrm0 < mxnrmx
        rcmp = (rnrm < delta * mxnrmr and rnrm0 < mxnrmr) or xpdt
        resid = rnrm / rnrm0 if rnrm0 > 0 else 0.0
        if it % RELIABLE_PERIOD == 0 or resid <= tol:
            ax, tres, conv, restart, nc = _reliable(mv, bp, x, r[0], rnrm0, resid, tol, nc)
            if conv:
                resid = tres
                break
            if restart:
                r[0] = bp - ax
                xp = xp + x; x = np.zeros(n, complex); bp = r[0].copy()
                rr = r[0].copy()
                r[1] = np.zeros(n, complex); r[2] = np.zeros(n, complex)
                u = [np.zeros(n, complex) for _ in range(3)]
                rnrm = mxnrmx = mxnrmr = tres * rnrm0
                alpha, omega, sigma, rho0 = 0.0, 1.0, 1.0, 1.0
                resid = tres
                continue
        if rcmp:
            ax = mv("N", x); nc += 1
            r[0] = bp - ax
            if xpdt:
                xp = xp + x; x = np.zeros(n, complex); bp = r[0].copy()
            mxnrmr = rnrm
            if xpdt:
                mxnrmx = rnrm
    x = x + xp
    ax = mv("N", x); nc += 1
    resid = np.linalg.norm(b - ax) / rnrm0 if rnrm0 > 0 else 0.0
    return x, it, resid, nc


def qmrccg(mv, b, x0=None, maxit=300, tol=1e-5, policy="replace", log=None):
    """Complex-symmetric QMR (Freund) with unconjugated bilinear forms.
    policy: what a failed reliable check does, "replace" (the true residual
    replaces the recursive one and the recurrence goes on; the library's
    default) or "restart" (the reference's restart, cu:543-556).  log: an
    optional list that receives (iteration, true residual, gap, action 0 /
    1 restart / 2 replace)."""
    b = np.asarray(b, np.complex128)
    n = b.size
    x = np.zeros(n, complex) if x0 is None else np.asarray(x0, np.complex128).copy()
    bnorm = np.linalg.norm(b)
    nc = 0
    r = b - mv("N", x); nc += 1
    dotu = lambda a, c: np.sum(a * c)

    def init(rv):
        nonlocal nc
        vt = rv.copy(); wt = rv.copy()
        atw = mv("N", wt); nc += 1
        st = dict(lam=1.0, kap=-1.0, th=-1.0, g=np.linalg.norm(vt), k=np.linalg.norm(wt), rho=dotu(vt, wt),
                  eps=dotu(vt, atw), mu=0.0)
        st["tau"] = st["eps"] / st["rho"]
        z = np.zeros(n, complex)
        return vt, wt, atw, st, z.copy(), z.copy(), z.copy(), z.copy(), z.copy()

    vt, wt, atw, S, p, ap, q, d, sv = init(r)
    resid = np.linalg.norm(r) / bnorm
    it = 0
    while it < maxit and resid > tol:
        it += 1
        p = vt / S["g"] - S["mu"] * p
        q = (atw - S["g"] * S["mu"] * q) / S["k"]
        ap = mv("N", p); nc += 1
        vt = ap - (S["tau"] / S["g"]) * vt
        wt = q - (S["tau"] / S["k"]) * wt
        g = np.linalg.norm(vt); k = np.linalg.norm(wt); rho = dotu(vt, wt)
        atw = mv("N", wt); nc += 1
        eps = dotu(vt, atw)
        g0, k0, rho0, tau0, kap0 = S["g"], S["k"], S["rho"], S["tau"], S["kap"]
        S.update(g=g, k=k, rho=rho, eps=eps)
        mu = g0 * k0 * rho / (g * tau0 * rho0)
        tau = eps / rho - g * mu
        at0 = abs(tau0) ** 2
        den2 = S["lam"] * at0 + abs(g) ** 2
        th = (1 - S["lam"]) * at0 / den2
        kap = -(g0 * np.conj(tau0) * kap0) / den2
        S.update(mu=mu, tau=tau, th=th, kap=kap, lam=S["lam"] * at0 / den2)
        d = th * d + kap * p
        sv = th * sv + kap * ap
        x = x + d
        r = r - sv
        resid = np.linalg.norm(r) / bnorm
        if it % RELIABLE_PERIOD == 0 or resid <= tol:
            rprev = r
            ax, tres, conv, restart, nc = _reliable(mv, b, x, r, bnorm, resid, tol, nc)
            replace = restart and policy == "replace"
            if log is not None:
                gap = np.linalg.norm(b - ax - rprev) / max(np.linalg.norm(b - ax), 1e-300)
                log.append((it, tres, gap, 0 if (conv or not restart) else (2 if replace else 1)))
            if conv:
                resid = tres
                break
            if restart:
                r = b - ax
                if not replace:
                    vt, wt, atw, S, p, ap, q, d, sv = init(r)
                resid = tres
    if resid > tol:
        ax, tres, conv, restart, nc = _reliable(mv, b, x, r, bnorm, resid, tol, nc)
        resid = tres
    return x, it, resid, nc


def petrkp(mv, b, x0=None, maxit=300, tol=1e-5):
    """CG on the normal equations A^H A x = A^H b (PETR90VER2)."""
    b = np.asarray(b, np.complex128)
    n = b.size
    x = np.zeros(n, complex) if x0 is None else np.asarray(x0, np.complex128).copy()
    bb = np.vdot(b, b).real
    nc = 0
    ace = mv("C", b); nc += 1
    gi = ace.copy(); pi = ace.copy()
    qi = mv("N", pi); nc += 1
    gg = np.vdot(gi, gi).real; qq = np.vdot(qi, qi).real
    alpha = gg / qq
    x = x + alpha * pi
    axi = mv("N", x); nc += 1
    resid = 1.0
    it = 0
    for it in range(1, maxit + 1):
        oldgg = gg
        gi = ace - mv("C", axi); nc += 1
        gg = np.vdot(gi, gi).real
        beta = gg / oldgg
        pi = gi + beta * pi
        qi = mv("N", pi); nc += 1
        qq = np.vdot(qi, qi).real
        alpha = gg / qq
        x = x + alpha * pi
        if it % 10 != 0:
            axi = axi + alpha * qi
        else:
            axi = mv("N", x); nc += 1
        rv = axi - b
        resid = np.sqrt(max(np.vdot(rv, rv).real, 0.0) / bb)
        if resid < tol:
            break
    return x, min(it, maxit), resid, nc
