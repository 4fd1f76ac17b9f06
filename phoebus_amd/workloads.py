"""Seeded synthetic workloads shared by the benchmark and the tests.

batched_con2prim_states: independent primitive states for the batched con2prim workload
(rho log-uniform [1e-8, 1e2], P/rho log-uniform [1e-6, 1e2], Lorentz factor ~ U(1, 10),
magnetization log-uniform [1e-4, 1e2], ideal gas Gamma = 4/3, spherical Kerr-Schild
a = 0.9375 at r ~ U(1.5, 50), theta = acos U(-1, 1) kept 1e-3 off the axis).  Vector
components are drawn isotropic in an orthonormal frame and scaled by 1/sqrt(gamma_ii)
into coordinate components, so W and B^2 = sigma rho hold up to the off-diagonal metric
terms.  Primitive layout: rho, v1..3 (W v), u, B1..3, P, gamma1.

shock_tube_prims: the 1-D special-relativistic shock tube (shock_tube.cpp:79-119) on a
Minkowski block, left state for cell centres x < 0.5, u = P/(Gamma-1) in closed form.
"""
import numpy as np

GAMMA_C2P = 4.0 / 3.0
SPIN = 0.9375


def ks_diag_gamma(r, th, a=SPIN):
    """Diagonal of the spherical Kerr-Schild spatial metric (spherical_kerr_schild.hpp)."""
    cth, sth = np.cos(th), np.sin(th)
    rho2 = r * r + a * a * cth * cth
    br = 2 * r / rho2
    g11 = 1 + br
    g22 = rho2
    g33 = (rho2 + a * a * (1 + br) * sth * sth) * sth * sth
    return g11, g22, g33


def batched_con2prim_states(rng, n, a=SPIN, gamma=GAMMA_C2P):
    r = rng.uniform(1.5, 50.0, n)
    th = np.arccos(rng.uniform(-1.0, 1.0, n))
    th = np.clip(th, 1e-3, np.pi - 1e-3)
    rho = 10.0 ** rng.uniform(-8.0, 2.0, n)
    P = rho * 10.0 ** rng.uniform(-6.0, 2.0, n)
    W = rng.uniform(1.0, 10.0, n)
    sig = 10.0 ** rng.uniform(-4.0, 2.0, n)
    g = ks_diag_gamma(r, th, a)
    nv = rng.normal(size=(3, n))
    nv /= np.linalg.norm(nv, axis=0)
    nb = rng.normal(size=(3, n))
    nb /= np.linalg.norm(nb, axis=0)
    prim = np.empty((10, n))
    prim[0] = rho
    for d in range(3):
        prim[1 + d] = np.sqrt(W * W - 1.0) * nv[d] / np.sqrt(g[d])
        prim[5 + d] = np.sqrt(sig * rho) * nb[d] / np.sqrt(g[d])
    prim[4] = P / (gamma - 1.0)
    prim[8] = P
    prim[9] = gamma
    return np.ascontiguousarray(np.stack([r, th])), np.ascontiguousarray(prim)


def shock_tube_prims(nx, ng=4, gamma=5.0 / 3.0, left=(1.0, 1.0, 0.0), right=(0.125, 0.1, 0.0),
                     x0=0.5, mhd=False):
    """Primitives [nprim][1][1][nx + 2 ng] over the entire block on [0, 1]; left/right =
    (rho, P, v).  The velocity primitive is W v (shock_tube.cpp:108-110), gamma1 =
    bulk modulus / P = Gamma for the ideal gas."""
    dx = 1.0 / nx
    x = (np.arange(nx + 2 * ng) - ng + 0.5) * dx
    L = x < x0
    rho = np.where(L, left[0], right[0])
    P = np.where(L, left[1], right[1])
    v = np.where(L, left[2], right[2])
    W = 1.0 / np.sqrt(1.0 - v * v)
    npr = 10 if mhd else 7
    prim = np.zeros((npr, 1, 1, x.size))
    prim[0, 0, 0] = rho
    prim[1, 0, 0] = W * v
    prim[4, 0, 0] = P / (gamma - 1.0)
    prim[npr - 2, 0, 0] = P
    prim[npr - 1, 0, 0] = gamma
    return prim
