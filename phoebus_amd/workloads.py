"""Seeded synthetic workloads shared by the benchmark and the tests.

batched_con2prim_states: independent primitive states for the batched con2prim workload
(rho log-uniform [1e-8, 1e2], P/rho log-uniform [1e-6, 1e2], Lorentz factor ~ U(1, 10),
magnetization log-uniform [1e-4, 1e2], ideal gas Gamma = 4/3, spherical Kerr-Schild
a = 0.9375 at r ~ U(1.5, 50), theta = acos U(-1, 1) kept 1e-3 off the axis).  Vector
components are drawn isotropic in an orthonormal frame and scaled by 1/sqrt(gamma_ii)
into coordinate components, so W and B^2 = sigma rho hold up to the off-diagonal metric
terms.  Primitive layout: rho, v1..3 (W v), u, B1..3, P, gamma1.

grmhd_block_prims: the 128^3-class Minkowski GRMHD block of the stage benchmark: rho =
exp N(0, 1), P = rho U(0.01, 1), Lorentz factor ~ U(1, 5) in isotropic directions, B the
centred curl of a random vector potential modulated by a smooth amplitude so that
B^2/rho is log-uniform over [1e-3, 10] up to per-cell scatter, over the entire block
(ghosts included).

shock_tube_prims: the 1-D special-relativistic shock tube (shock_tube.cpp:79-119) on a
Minkowski block, left state for cell centres x < 0.5, u = P/(Gamma-1) in closed form.
"""
import numpy as np

GAMMA_C2P = 4.0 / 3.0
SPIN = 0.9375


def smooth_uniform_field(rng, sh, nmodes=6):
    """A field over the block whose values are uniform on (0, 1) cell by cell and which
    varies smoothly in space: a sum of nmodes random plane waves of wave numbers |k_d| <= 2
    per block, mapped through its own rank (a monotone map, so the field stays smooth)."""
    z, y, x = np.meshgrid(*(np.arange(m) / m for m in sh), indexing="ij")
    f = np.zeros(sh)
    for _ in range(nmodes):
        k = rng.integers(-2, 3, 3)
        if not k.any():
            k[0] = 1
        f += rng.normal() * np.sin(2 * np.pi * (k[0] * x + k[1] * y + k[2] * z)
                                   + rng.uniform(0, 2 * np.pi))
    r = np.empty(f.size)
    r[np.argsort(f.ravel(), kind="stable")] = (np.arange(f.size) + 0.5) / f.size
    return r.reshape(sh)


def grmhd_block_prims(rng, n_int, ng, gamma, log_sigma=(-3.0, 1.0)):
    """Primitives [10][n][n][n] (n = n_int + 2 ng) of the stage benchmark's block:
    rho = exp N(0, 1), P = rho U(0.01, 1), Lorentz factor U(1, 5) in isotropic directions,
    and B = the centred-difference curl of s A, with A i.i.d. U(-1, 1) per cell and
    component and s a smooth amplitude 10^(c / 2), c uniform on log_sigma cell by cell
    (smooth_uniform_field); B is then scaled so that the median of sigma / 10^c is 1,
    sigma = B^2 / rho.  The curl keeps B divergence-free on the centred stencil; sigma is
    log-uniform over log_sigma up to the per-cell scatter of rho and of the curl of A
    (log10 sigma percentiles 1/5/50/95/99 %: -3.7 / -3.1 / -1.0 / 1.0 / 1.6 at 136^3)."""
    n = n_int + 2 * ng
    sh = (n, n, n)
    rho = np.exp(rng.normal(0.0, 1.0, sh))
    P = rho * rng.uniform(0.01, 1.0, sh)
    W = rng.uniform(1.0, 5.0, sh)
    nv = rng.normal(size=(3,) + sh)
    nv /= np.linalg.norm(nv, axis=0)
    vp = np.sqrt(W * W - 1.0) * nv
    A = rng.uniform(-1.0, 1.0, (3,) + sh)
    c = log_sigma[0] + (log_sigma[1] - log_sigma[0]) * smooth_uniform_field(rng, sh)
    A *= 10.0 ** (c / 2)
    d = lambda f, ax: 0.5 * (np.roll(f, -1, axis=ax) - np.roll(f, 1, axis=ax))
    # axes of (k, j, i): x -> 2, y -> 1, z -> 0
    Bx = d(A[2], 1) - d(A[1], 0)
    By = d(A[0], 0) - d(A[2], 2)
    Bz = d(A[1], 2) - d(A[0], 1)
    B = np.stack([Bx, By, Bz])
    sig = np.sum(B * B, axis=0) / rho
    B *= np.sqrt(1.0 / np.median(sig / 10.0 ** c))
    prim = np.empty((10,) + sh)
    prim[0] = rho
    prim[1:4] = vp
    prim[4] = P / (gamma - 1.0)
    prim[5:8] = B
    prim[8] = P
    prim[9] = gamma
    return np.ascontiguousarray(prim)


def magnetization(prim):
    """sigma = b^2 / rho of Minkowski primitives (b^2 = (B^2 + (B.Wv)^2) / W^2)."""
    vp, B = prim[1:4], prim[5:8]
    W2 = 1.0 + np.sum(vp * vp, axis=0)
    bsq = (np.sum(B * B, axis=0) + np.sum(B * vp, axis=0) ** 2) / W2
    return bsq / prim[0]


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


# ---- Fishbone-Moncrief torus on an FMKS grid (the torus workloads) ----------------------
# Fishbone & Moncrief, ApJ 207 (1976) 962, in the form HARM-family codes use
# (torus.cpp:744-790 restates the same: lfish_calc, log_enthalpy; u^phi in Boyer-Lindquist).
TORUS = dict(a=0.9375, gamma=5.0 / 3.0, rin=6.0, rmax=12.0, kappa=0.01, beta=100.0,
             jitter=0.04, hslope=0.3, r_outer=250.0, ninside=5,
             rho0_floor=1e-8, sie0_floor=1e-3, rho_exp_floor=-2.0, sie_exp_floor=-1.0,
             rho_min_eos=1e-8, sie_min_eos=1e-8)


def fm_angular_momentum(a, r):
    sr = np.sqrt(r)
    return (((a * a - 2.0 * a * sr + r * r) *
             ((-2.0 * a * r * (a * a - 2.0 * a * sr + r * r)) /
              np.sqrt(2.0 * a * sr + (-3.0 + r) * r) +
              ((a + (-2.0 + r) * sr) * (r ** 3 + a * a * (2.0 + r))) /
              np.sqrt(1 + (2.0 * a) / r ** 1.5 - 3.0 / r))) /
            (r ** 3 * np.sqrt(2.0 * a * sr + (-3.0 + r) * r) * (a * a + (-2.0 + r) * r)))


def _fm_terms(r, th, a, l):
    sth, cth = np.sin(th), np.cos(th)
    DD = r * r - 2.0 * r + a * a
    AA = (r * r + a * a) ** 2 - DD * a * a * sth * sth
    SS = r * r + a * a * cth * cth
    q = np.sqrt(1.0 + 4.0 * (l * l * SS * SS) * DD / (AA * AA * sth * sth))
    return 0.5 * np.log((1.0 + q) / (SS * DD / AA)) - 0.5 * q - 2.0 * a * r * l / AA, (DD, AA, SS, sth)


def fm_log_enthalpy(r, th, a, rin, l):
    lnh, _ = _fm_terms(r, th, a, l)
    lin, _ = _fm_terms(np.asarray(rin, float), np.asarray(0.5 * np.pi), a, l)
    return lnh - lin


def fm_uphi(r, th, a, l):
    _, (DD, AA, SS, sth) = _fm_terms(r, th, a, l)
    expm2chi = SS * SS * DD / (AA * AA * sth * sth)
    up1 = np.sqrt((-1.0 + np.sqrt(1.0 + 4.0 * l * l * expm2chi)) / 2.0)
    return 2.0 * a * r * np.sqrt(1.0 + up1 * up1) / np.sqrt(AA * SS * DD) + np.sqrt(SS / AA) * up1 / sth


def torus_grid_geometry(nx=(256, 128, 128), a=TORUS["a"], r_outer=TORUS["r_outer"],
                        ninside=TORUS["ninside"]):
    """x1min, dx of the torus grid: x1max = log r_outer and ninside zones inside the
    horizon (torus.cpp:477-497)."""
    x1max = np.log(r_outer)
    xh = np.log(1.0 + np.sqrt(1.0 - a * a))
    dx1 = (x1max - xh) / (nx[0] - ninside)
    x1min = xh - (ninside + 0.5) * dx1
    return (x1min, 0.0, 0.0), ((x1max - x1min) / nx[0], 1.0 / nx[1], 2.0 * np.pi / nx[2])


def torus_floors(x1, P=None):
    """The torus deck's density and specific-energy floors at X1 = log r (inputs/torus.pin:
    ExpX1RhoSie, rho0 1e-8, sie0 1e-3, exponents -2 and -1, clamped below by the EOS
    minima rho_min = sie_min = 1e-8 as fixup.hpp:105-110)."""
    P = dict(TORUS, **(P or {}))
    rflr = np.maximum(P["rho0_floor"] * np.exp(P["rho_exp_floor"] * x1), P["rho_min_eos"])
    sflr = np.maximum(P["sie0_floor"] * np.exp(P["sie_exp_floor"] * x1), P["sie_min_eos"])
    return rflr, sflr


def torus_grid(rng, nx=(256, 128, 128), ng=4, **kw):
    """Primitives [10][nz+2ng][ny+2ng][nx+2ng] of the magnetized Fishbone-Moncrief torus
    of the torus problem generator (torus.cpp:130-474, 525-560, 843-921) on an FMKS grid
    (derefine_poles off, hslope 0.3), ideal gas:
      - inside the torus (r > r_in and ln h > 0): rho, u from the enthalpy on the
        polytrope K = kappa, normalized by the density at r_max on the equator
        (GetStateFromEnthalpy, torus.cpp:800-810), W v^i = u^i + u^t beta^i from the
        Boyer-Lindquist u^phi (u^r = u^theta = 0, torus.cpp:303-333);
      - u *= 1 + jitter (U - 1/2) in every cell (torus.cpp:339);
      - the floors of the deck everywhere: rho <- max(rho, rho_flr), u <- rho sie_flr
        where u / rho < sie_flr (torus.cpp:342-352), so the atmosphere is the floor
        state at rest;
      - A_phi = max(rho_av - 0.2, 0) on cell corners, rho_av the mean of the four cells
        around the corner (torus.cpp:432-449); B^1 = d_2 A / sqrt(gamma),
        B^2 = -d_1 A / sqrt(gamma) by the corner differences of torus.cpp:452-466 (the
        last ghost row and column keep B = 0);
      - B scaled by sqrt(beta_pmax / beta) with beta_pmax = max P / (max b^2 / 2) over
        the interior cells (HARM-style normalization, torus.cpp:525-547, 843-921; b^2
        the comoving field strength, relativity_utils.hpp:131-141).
    Returns (prim, xmin, dx) with xmin of the first interior cell face."""
    P = dict(TORUS, **kw)
    a, gam, rin, rmax, kappa = P["a"], P["gamma"], P["rin"], P["rmax"], P["kappa"]
    xmin, dx = torus_grid_geometry(nx, a, P["r_outer"], P["ninside"])
    # cell centres including ghosts; the flow is axisymmetric, so compute one phi slice
    X1 = xmin[0] + (np.arange(nx[0] + 2 * ng) - ng + 0.5) * dx[0]
    X2 = xmin[1] + (np.arange(nx[1] + 2 * ng) - ng + 0.5) * dx[1]
    x1, x2 = np.meshgrid(X1, X2, indexing="xy")  # [j][i]
    r = np.exp(x1)
    th = np.pi * x2 + ((1.0 - P["hslope"]) / 2.0) * np.sin(2.0 * np.pi * x2)
    dthdx2 = np.pi + np.pi * (1.0 - P["hslope"]) * np.cos(2.0 * np.pi * x2)
    l = fm_angular_momentum(a, rmax)
    with np.errstate(all="ignore"):
        lnh = np.where(r > rin, fm_log_enthalpy(r, th, a, rin, l), -1.0)
    inside = lnh > 0.0
    hm1 = np.where(inside, np.exp(np.where(inside, lnh, 0.0)) - 1.0, 0.0)
    rho = (hm1 * (gam - 1.0) / (kappa * gam)) ** (1.0 / (gam - 1.0))
    hm1_max = np.exp(fm_log_enthalpy(np.asarray(rmax), np.asarray(0.5 * np.pi), a, rin, l)) - 1.0
    rho_max = (hm1_max * (gam - 1.0) / (kappa * gam)) ** (1.0 / (gam - 1.0))
    u = kappa * rho ** gam / (gam - 1.0) / rho_max
    rho = rho / rho_max
    # Kerr-Schild metric pieces at (r, theta) (the FMKS map leaves t and phi alone)
    cth, sth = np.cos(th), np.sin(th)
    rho2 = r * r + a * a * cth * cth
    g00 = -1.0 + 2.0 * r / rho2
    g03 = -2.0 * a * r * sth * sth / rho2
    g33 = (rho2 + a * a * (1.0 + 2.0 * r / rho2) * sth * sth) * sth * sth
    g11 = 1.0 + 2.0 * r / rho2
    alpha = np.sqrt(rho2 / (rho2 + 2.0 * r))
    betar = 2.0 * r / (rho2 + 2.0 * r)
    # torus 4-velocity: u^r = u^th = 0 (Boyer-Lindquist and Kerr-Schild agree), u^t from
    # the normalization, W v^i = u^i + u^t beta^i in FMKS (beta^1 = beta^r / r)
    with np.errstate(all="ignore"):
        up = np.where(inside, fm_uphi(r, th, a, l), 0.0)
        A_ = g00
        B_ = 2.0 * g03 * up
        C_ = g33 * up * up + 1.0
        ut = np.where(inside, (-B_ - np.sqrt(B_ * B_ - 4.0 * A_ * C_)) / (2.0 * A_), 1.0 / alpha)
    v1 = np.where(inside, ut * betar / r, 0.0)
    v3 = np.where(inside, up, 0.0)
    nz = nx[2] + 2 * ng
    # jitter (drawn per cell of the whole grid), then the deck's floors
    u = u[None] * (1.0 + P["jitter"] * (rng.uniform(size=(nz,) + rho.shape) - 0.5))
    rflr, sflr = torus_floors(x1, P)
    rho = np.maximum(rho, rflr)
    u = np.where(u / rho[None] < sflr[None], rho[None] * sflr[None], u)
    # single loop from the corner vector potential (axisymmetric: one slice)
    Ac = np.zeros_like(rho)
    Ac[1:, 1:] = np.maximum(0.25 * (rho[1:, 1:] + rho[1:, :-1] + rho[:-1, 1:] + rho[:-1, :-1])
                            - 0.2, 0.0)
    gamdet = np.sqrt(rho2 * (rho2 + 2.0 * r)) * sth * (r * dthdx2)  # sqrt(gamma) of FMKS
    B1 = np.zeros_like(rho)
    B2 = np.zeros_like(rho)
    B1[:-1, :-1] = -(Ac[:-1, :-1] - Ac[1:, :-1] + Ac[:-1, 1:] - Ac[1:, 1:]) / (2.0 * dx[1] * gamdet[:-1, :-1])
    B2[:-1, :-1] = (Ac[:-1, :-1] + Ac[1:, :-1] - Ac[:-1, 1:] - Ac[1:, 1:]) / (2.0 * dx[0] * gamdet[:-1, :-1])
    gam11 = g11 * r * r  # gamma_{X1 X1} = g_rr (dr/dX1)^2
    gam22 = rho2 * dthdx2 * dthdx2
    gam13 = -a * (1.0 + 2.0 * r / rho2) * sth * sth * r  # gamma_{X1 X3} = g_r phi dr/dX1
    prim = np.empty((10, nz) + rho.shape)
    prim[0] = rho
    prim[1] = v1
    prim[2] = 0.0
    prim[3] = v3
    prim[4] = u
    prim[5] = B1
    prim[6] = B2
    prim[7] = 0.0
    prim[8] = (gam - 1.0) * u
    prim[9] = gam
    # HARM-style beta normalization over the interior cells
    I = (slice(ng, -ng), slice(ng, -ng))
    W = np.sqrt(1.0 + gam11 * v1 * v1 + 2.0 * gam13 * v1 * v3 + g33 * v3 * v3)
    Bsq = gam11 * B1 * B1 + gam22 * B2 * B2
    Bv = (gam11 * v1 * B1 + gam13 * v3 * B1) / W
    bsq = Bsq / (W * W) + Bv * Bv
    bsq_max = bsq[I].max()
    if bsq_max > 0:
        beta_pmax = prim[8][(slice(ng, -ng),) + I].max() / (0.5 * bsq_max)
        s = np.sqrt(beta_pmax / P["beta"])
        prim[5] *= s
        prim[6] *= s
    return prim, xmin, dx


def uneven_blocks(rng, nblocks=512, sizes=(16, 32, 64), probs=(0.5, 0.35, 0.15),
                  grid=(256, 128, 128), centre=None):
    """Block sizes drawn from sizes with probs, and origins (interior index of the first
    cell) in the grid, the small blocks drawn near the torus centre (an AMR-like mix:
    finest blocks where rho is largest).  Blocks may overlap: each has its own fields."""
    s = rng.choice(np.asarray(sizes), size=nblocks, p=np.asarray(probs))
    ci = centre if centre is not None else (grid[0] // 2, grid[1] // 2, grid[2] // 2)
    out = []
    for n in s:
        spread = {16: 0.15, 32: 0.3}.get(int(n), 0.5)
        o = []
        for d in range(3):
            lo, hi = 0, grid[d] - n
            mid = ci[d] - n // 2
            v = int(np.clip(round(mid + rng.normal() * spread * grid[d]), lo, hi))
            o.append(v)
        out.append((int(n), tuple(o)))
    return out
