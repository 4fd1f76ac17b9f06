"""Independent closed forms used to pin the oracle (test infrastructure).

These restate GRMHD definitions directly from u^mu, b^mu and T^{mu nu}, a formulation
different from the reference's p2c / prim_to_flux algebra, so that a dropped term, sign
or index in the oracle shows up as a mismatch.
"""
import numpy as np


def four_vectors(g, prim, mhd):
    """u^mu, b^mu, b^2 from phoebus primitives (v = W v_normal, B = cell-centred field)."""
    gcov = g.np("gcov")
    gam = gcov[1:, 1:]
    alpha = g.alpha
    beta = np.array(g.beta[:])
    vp = np.array(prim[1:4])
    W = np.sqrt(1.0 + vp @ gam @ vp)
    v = vp / W
    u = np.zeros(4)
    u[0] = W / alpha
    u[1:] = W * (v - beta / alpha)
    B = np.array(prim[5:8]) if mhd else np.zeros(3)
    Bv = B @ gam @ v
    b = np.zeros(4)
    b[0] = W * Bv / alpha
    b[1:] = (B + alpha * b[0] * u[1:]) / W
    bsq = b @ gcov @ b
    return u, b, bsq, W


def tmunu(g, gcon4, prim, mhd):
    prs = 8 if mhd else 5
    u, b, bsq, _ = four_vectors(g, prim, mhd)
    h = prim[0] + prim[4] + prim[prs]
    return (h + bsq) * np.outer(u, u) + (prim[prs] + 0.5 * bsq) * gcon4 - np.outer(b, b)


def valencia_cons(g, gcon4, prim, mhd):
    """D, S_j, tau, B^j densitized by sqrt(gamma) from T^{mu nu}."""
    gcov = g.np("gcov")
    T = tmunu(g, gcon4, prim, mhd)
    u, _, _, _ = four_vectors(g, prim, mhd)
    sg = g.detgamma
    D = sg * prim[0] * g.alpha * u[0]
    Tmix = T @ gcov  # T^mu_nu
    S = sg * g.alpha * Tmix[0, 1:]
    tau = sg * g.alpha ** 2 * T[0, 0] - D
    out = [D, *S, tau]
    if mhd:
        out += list(sg * np.array(prim[5:8]))
    return np.array(out)


def valencia_flux(g, gcon4, prim, mhd, d):
    """Physical flux along direction d (0-based) times sqrt(-g)."""
    gcov = g.np("gcov")
    T = tmunu(g, gcon4, prim, mhd)
    u, b, _, W = four_vectors(g, prim, mhd)
    sgg = g.detg
    dd = d + 1
    FD = sgg * prim[0] * u[dd]
    Tmix = T @ gcov
    FS = sgg * Tmix[dd, 1:]
    Ftau = sgg * (g.alpha * T[0, dd] - prim[0] * u[dd])
    out = [FD, *FS, Ftau]
    if mhd:
        out += list(sgg * (b[1:] * u[dd] - b[dd] * u[1:]))
    return np.array(out)


def random_prims(rng, n, mhd, gamma, wmax=5.0, sigmax=1.0):
    """Random physical primitive states (rho, W v, u, [B], P, gamma1) in Minkowski."""
    rho = np.exp(rng.normal(0, 1, n))
    P = rho * rng.uniform(0.01, 1.0, n)
    W = rng.uniform(1.0, wmax, n)
    nv = rng.normal(size=(3, n))
    nv /= np.linalg.norm(nv, axis=0)
    vp = np.sqrt(W * W - 1.0) * nv
    u = P / (gamma - 1.0)
    out = [rho, vp[0], vp[1], vp[2], u]
    if mhd:
        B = rng.normal(size=(3, n)) * np.sqrt(sigmax * rho)
        out += [B[0], B[1], B[2]]
    out += [P, np.full(n, gamma)]
    return np.array(out)


def to_coordinate_frame(prims, gam, mhd):
    """Scale velocity/field components by 1/sqrt(gamma_ii) so that a state drawn for an
    orthonormal frame keeps W and b^2/rho of the same order in curvilinear coordinates."""
    out = prims.copy()
    s = 1.0 / np.sqrt(np.diag(gam))
    out[1:4] *= s[:, None]
    if mhd:
        out[5:8] *= s[:, None]
    return out
