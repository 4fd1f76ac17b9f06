"""Exact solution of the 1-D special-relativistic Riemann problem for an ideal gas with
zero tangential velocity (Marti & Mueller, J. Fluid Mech. 258 (1994) 317; Living Rev.
Relativity 6 (2003) 7, sections 6.1-6.2).  Test infrastructure: the shock-tube runs are
checked against it (L1 norms, plateau states, wave positions).

Rarefactions keep the Riemann invariant J (+ for left-facing, - for right-facing fans)
constant along an isentrope; shocks follow the Taub adiabat with the invariant mass flux
j.  The star pressure p* solves v*_L(p*) = v*_R(p*) by bisection on log p.
"""
import math

import numpy as np


class State:
    def __init__(self, gamma, rho, p, v):
        self.g, self.rho, self.p, self.v = gamma, rho, p, v
        self.h = 1.0 + gamma / (gamma - 1.0) * p / rho
        self.cs = math.sqrt(gamma * p / (rho * self.h))
        self.W = 1.0 / math.sqrt(1.0 - v * v)


def _rare_v(a, p, sign):
    """velocity behind a rarefaction from ahead state a to pressure p (isentropic); sign
    -1 for a left-facing fan (J+ constant), +1 for a right-facing one (J- constant)."""
    g = a.g
    rho = a.rho * (p / a.p) ** (1.0 / g)
    h = 1.0 + g / (g - 1.0) * p / rho
    cs = math.sqrt(g * p / (rho * h))
    s = math.sqrt(g - 1.0)
    A = ((1 + a.v) / (1 - a.v)) * (((s + a.cs) / (s - a.cs)) * ((s - cs) / (s + cs))) ** (
        -2.0 * sign / s)
    return (A - 1.0) / (A + 1.0), rho


def _shock(a, p, sign):
    """post-shock (velocity, density, shock speed) for ahead state a, pressure p > p_a;
    sign -1 for a left-facing shock, +1 for a right-facing one (Taub adiabat)."""
    g = a.g
    dp = a.p - p
    # h_b^2 (1 + (g-1) dp/(g p)) - h_b (g-1) dp/(g p) + h_a dp/rho_a - h_a^2 = 0
    qa = 1.0 + (g - 1.0) * dp / (g * p)
    qb = -(g - 1.0) * dp / (g * p)
    qc = a.h * dp / a.rho - a.h * a.h
    hb = (-qb + math.sqrt(qb * qb - 4 * qa * qc)) / (2 * qa)
    rhob = g * p / ((g - 1.0) * (hb - 1.0))
    j2 = -(p - a.p) / (hb / rhob - a.h / a.rho)
    j = sign * math.sqrt(j2)
    r2w2 = a.rho ** 2 * a.W ** 2
    Vs = (r2w2 * a.v + j * math.sqrt(j2 + r2w2 * (1 - a.v ** 2))) / (r2w2 + j2)
    Ws = 1.0 / math.sqrt(1.0 - Vs * Vs)
    vb = (a.h * a.W * a.v + Ws * (p - a.p) / j) / (
        a.h * a.W + (p - a.p) * (Ws * a.v / j + 1.0 / (a.rho * a.W)))
    return vb, rhob, Vs


def _side(a, p, sign):
    if p > a.p:
        v, rho, _ = _shock(a, p, sign)
    else:
        v, rho = _rare_v(a, p, sign)
    return v, rho


def star_state(L, R):
    """(p*, v*, rho*_L, rho*_R)"""
    lo, hi = math.log(1e-12 * min(L.p, R.p) + 1e-300), math.log(1e3 * max(L.p, R.p))
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        p = math.exp(mid)
        if _side(L, p, -1)[0] > _side(R, p, +1)[0]:
            lo = mid
        else:
            hi = mid
    p = math.exp(0.5 * (lo + hi))
    vl, rl = _side(L, p, -1)
    vr, rr = _side(R, p, +1)
    return p, 0.5 * (vl + vr), rl, rr


def _fan_state(a, xi, sign):
    """state inside a rarefaction fan at similarity coordinate xi: the pressure where the
    characteristic speed (v - sign*... ) equals xi, found by bisection."""
    def speed(p):
        v, rho = _rare_v(a, p, sign)
        h = 1.0 + a.g / (a.g - 1.0) * p / rho
        cs = math.sqrt(a.g * p / (rho * h))
        c = -cs if sign < 0 else cs
        return (v + c) / (1.0 + v * c), v, rho

    lo, hi = 1e-14 * a.p, a.p
    for _ in range(200):
        mid = math.sqrt(lo * hi)
        s, _, _ = speed(mid)
        # left fan: speed decreases with rising pressure toward the head
        if (s > xi) == (sign < 0):
            lo = mid
        else:
            hi = mid
    p = math.sqrt(lo * hi)
    _, v, rho = speed(p)
    return rho, p, v


def solution(gamma, left, right, x0, t, x):
    """density, pressure, velocity at positions x and time t for left/right = (rho, p, v)."""
    L, R = State(gamma, *left), State(gamma, *right)
    ps, vs, rls, rrs = star_state(L, R)
    rho = np.empty_like(x)
    pr = np.empty_like(x)
    vel = np.empty_like(x)
    # wave speeds
    if ps <= L.p:
        head = (L.v - L.cs) / (1 - L.v * L.cs)
        Ls = State(gamma, rls, ps, vs)
        tail = (vs - Ls.cs) / (1 - vs * Ls.cs)
        lshock = None
    else:
        lshock = _shock(L, ps, -1)[2]
    if ps <= R.p:
        rhead = (R.v + R.cs) / (1 + R.v * R.cs)
        Rs = State(gamma, rrs, ps, vs)
        rtail = (vs + Rs.cs) / (1 + vs * Rs.cs)
        rshock = None
    else:
        rshock = _shock(R, ps, +1)[2]
    for n, xx in enumerate(x):
        xi = (xx - x0) / t
        if xi < vs:
            if lshock is not None:
                st = (L.rho, L.p, L.v) if xi < lshock else (rls, ps, vs)
            elif xi < head:
                st = (L.rho, L.p, L.v)
            elif xi < tail:
                st = _fan_state(L, xi, -1)
            else:
                st = (rls, ps, vs)
        else:
            if rshock is not None:
                st = (R.rho, R.p, R.v) if xi > rshock else (rrs, ps, vs)
            elif xi > rhead:
                st = (R.rho, R.p, R.v)
            elif xi > rtail:
                st = _fan_state(R, xi, +1)
            else:
                st = (rrs, ps, vs)
        rho[n], pr[n], vel[n] = st
    return rho, pr, vel, dict(p_star=ps, v_star=vs, rho_star_l=rls, rho_star_r=rrs,
                              shock=rshock if rshock is not None else lshock)
