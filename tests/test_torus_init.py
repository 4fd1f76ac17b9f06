"""Pins the Fishbone-Moncrief torus initial data (phoebus_amd.workloads.torus_grid) used by
the uneven-pack and torus-mesh workloads against the semantics of the torus problem
generator (torus.cpp:130-474, 525-560, 744-921) and against closed forms that do not share
its algebra.  CPU only."""
import math

import numpy as np
import pytest

from phoebus_amd import workloads as WL

A, RIN, RMAX = WL.TORUS["a"], WL.TORUS["rin"], WL.TORUS["rmax"]


def _bl(r, th, a):
    """Boyer-Lindquist g_tt, g_tphi, g_phiphi and the inverse g^tt, g^tphi."""
    s2, c2 = math.sin(th) ** 2, math.cos(th) ** 2
    S = r * r + a * a * c2
    gtt = -(1 - 2 * r / S)
    gtp = -2 * a * r * s2 / S
    gpp = (r * r + a * a + 2 * a * a * r * s2 / S) * s2
    det = gtt * gpp - gtp * gtp
    return gtt, gtp, gpp, gpp / det, -gtp / det


def test_fm_angular_momentum_is_keplerian_ut_uphi():
    """l = u^t u_phi of the prograde circular equatorial orbit at r_max (Bardeen, Press &
    Teukolsky 1972, ApJ 178, 347, eqs. 2.12-2.13 with M = 1), which is the constant of
    the Fishbone-Moncrief solution (torus.cpp:188, lfish_calc)."""
    for a, r in ((A, RMAX), (0.0, 12.0), (0.5, 8.0), (0.9, 20.0)):
        den = r ** 0.75 * math.sqrt(r ** 1.5 - 3 * math.sqrt(r) + 2 * a)
        ut_low = -(r ** 1.5 - 2 * math.sqrt(r) + a) / den
        up_low = (r * r - 2 * a * math.sqrt(r) + a * a) / den
        _, _, _, gitt, gitp = _bl(r, math.pi / 2, a)
        ut = gitt * ut_low + gitp * up_low
        assert abs(WL.fm_angular_momentum(a, r) - ut * up_low) < 1e-12 * abs(ut * up_low)


def test_fm_enthalpy_zero_at_rin_and_stationary_at_rmax():
    """ln h = 0 at the inner edge on the equator, and the pressure maximum of the torus is
    at r_max on the equator (d ln h / dr = 0 there, d ln h / d theta = 0 by symmetry)."""
    l = WL.fm_angular_momentum(A, RMAX)
    assert abs(WL.fm_log_enthalpy(np.array(RIN), np.array(math.pi / 2), A, RIN, l)) < 1e-14
    h = 1e-4
    f = lambda r, th: float(WL.fm_log_enthalpy(np.array(r), np.array(th), A, RIN, l))
    dr = (f(RMAX + h, math.pi / 2) - f(RMAX - h, math.pi / 2)) / (2 * h)
    assert abs(dr) < 1e-8
    assert f(RMAX, math.pi / 2) > f(RMAX - 1, math.pi / 2) and f(RMAX, math.pi / 2) > f(RMAX + 1, math.pi / 2)
    assert f(RMAX, math.pi / 2) > f(RMAX, math.pi / 2 + 0.1)
    # no bound matter inside r_in on the equator
    assert f(RIN - 0.1, math.pi / 2) < 0


@pytest.mark.parametrize("r,th", [(9.0, 1.5), (12.0, math.pi / 2), (15.0, 1.3), (20.0, 1.7)])
def test_fm_uphi_keeps_ut_uphi_constant(r, th):
    """u^phi of the torus (Boyer-Lindquist, u^r = u^theta = 0) normalized with u^t from
    u.u = -1 satisfies u^t u_phi = l everywhere in the torus."""
    l = WL.fm_angular_momentum(A, RMAX)
    assert WL.fm_log_enthalpy(np.array(r), np.array(th), A, RIN, l) > 0
    up = float(WL.fm_uphi(np.array(r), np.array(th), A, l))
    gtt, gtp, gpp, _, _ = _bl(r, th, A)
    # gtt ut^2 + 2 gtp ut up + gpp up^2 = -1, future-directed root
    ut = (-2 * gtp * up - math.sqrt(4 * gtp * gtp * up * up - 4 * gtt * (gpp * up * up + 1))) / (2 * gtt)
    u_phi = gtp * ut + gpp * up
    assert abs(ut * u_phi - l) < 1e-11 * l


@pytest.fixture(scope="module")
def grid():
    nx, ng = (96, 64, 4), 4
    prim, xmin, dx = WL.torus_grid(np.random.default_rng(5), nx=nx, ng=ng)
    X1 = xmin[0] + (np.arange(nx[0] + 2 * ng) - ng + 0.5) * dx[0]
    X2 = xmin[1] + (np.arange(nx[1] + 2 * ng) - ng + 0.5) * dx[1]
    x1, x2 = np.meshgrid(X1, X2, indexing="xy")
    h = WL.TORUS["hslope"]
    th = np.pi * x2 + (1 - h) / 2 * np.sin(2 * np.pi * x2)
    return prim, np.exp(x1), th, x1, x2, nx, ng, dx


def test_density_normalized_at_rmax(grid):
    """rho = 1 at r_max on the equator (torus.cpp:283-289): the densest cell is the one
    nearest (r_max, pi/2) and holds 1 to the grid's resolution, no cell exceeds it by more
    than that resolution."""
    prim, r, th, *_ = grid
    rho = prim[0, 0]
    j, i = np.unravel_index(np.argmax(rho), rho.shape)
    assert abs(r[j, i] - RMAX) < 0.5 and abs(th[j, i] - np.pi / 2) < 0.06
    assert 0.98 < rho.max() <= 1.0 + 1e-3


def test_no_matter_inside_rin_and_floors(grid):
    """Cells with r <= r_in (and all cells with ln h <= 0) hold the floor state at rest with
    no field; every cell is at or above the deck's floors (torus.cpp:342-352)."""
    prim, r, th, x1, _, _, ng, _ = grid
    rflr, sflr = WL.torus_floors(x1)
    l = WL.fm_angular_momentum(A, RMAX)
    with np.errstate(all="ignore"):
        lnh = np.where(r > RIN, WL.fm_log_enthalpy(r, th, A, RIN, l), -1.0)
    out = lnh <= 0
    assert np.any(out) and np.any(~out)
    for k in range(prim.shape[1]):
        assert np.all(prim[0, k][out] == rflr[out])
        assert np.all(prim[4, k][out] == rflr[out] * sflr[out])
        assert np.all(prim[1:4, k][:, out] == 0.0)
        assert np.all(prim[0, k] >= rflr) and np.all(prim[4, k] >= prim[0, k] * sflr * (1 - 1e-15))
    inner = r <= RIN
    assert np.all(prim[5:8, :, inner] == 0.0)


def test_field_divergence_free_and_beta(grid):
    """B = curl of the corner A_phi: the corner-stencil divergence
    d_1 (sqrt(gamma) B^1) + d_2 (sqrt(gamma) B^2) vanishes to round-off; after the
    normalization max P / max(b^2 / 2) over the interior = target beta (torus.cpp:525-547)
    with b^2 the comoving field strength, here recomputed from u^mu and b^mu
    (tests/_physics.py) on the FMKS metric built from the Kerr-Schild one."""
    prim, r, th, x1, x2, nx, ng, dx = grid
    a = A
    h = WL.TORUS["hslope"]
    dthdx2 = np.pi + np.pi * (1 - h) * np.cos(2 * np.pi * x2)
    rho2 = r * r + a * a * np.cos(th) ** 2
    sg = np.sqrt(rho2 * (rho2 + 2 * r)) * np.sin(th) * r * dthdx2
    F1, F2 = sg * prim[5, 0], sg * prim[6, 0]
    # divergence at the corner between cells (j-1, j) x (i-1, i): the discretization in
    # which the corner-A curl is exactly divergence-free (flux-CT's, Toth 2000)
    c = ((F1[1:, 1:] + F1[:-1, 1:]) - (F1[1:, :-1] + F1[:-1, :-1])) / (2 * dx[0]) \
        + ((F2[1:, 1:] + F2[1:, :-1]) - (F2[:-1, 1:] + F2[:-1, :-1])) / (2 * dx[1])
    scale = np.max(np.abs(F1)) / dx[0] + np.max(np.abs(F2)) / dx[1]
    assert scale > 0
    assert np.max(np.abs(c[:-1, :-1])) < 1e-12 * scale
    # comoving b^2 from the four-vectors
    import _physics as PH

    class G:
        pass

    I = (slice(ng, -ng), slice(ng, -ng))
    bsq_max = 0.0
    pmax = prim[8][(slice(ng, -ng),) + I].max()
    jj, ii = np.nonzero(prim[5, 0] ** 2 + prim[6, 0] ** 2 > 0)
    for j, i in zip(jj, ii):
        if not (ng <= j < ng + nx[1] and ng <= i < ng + nx[0]):
            continue
        rr, t = r[j, i], th[j, i]
        s2, c2 = np.sin(t) ** 2, np.cos(t) ** 2
        S = rr * rr + a * a * c2
        # Kerr-Schild (t, r, th, phi) then the FMKS Jacobian diag(1, r, dth/dx2, 1)
        g = np.zeros((4, 4))
        g[0, 0] = -(1 - 2 * rr / S)
        g[0, 1] = g[1, 0] = 2 * rr / S
        g[0, 3] = g[3, 0] = -2 * a * rr * s2 / S
        g[1, 1] = 1 + 2 * rr / S
        g[1, 3] = g[3, 1] = -a * (1 + 2 * rr / S) * s2
        g[2, 2] = S
        g[3, 3] = (rr * rr + a * a + 2 * a * a * rr * s2 / S) * s2
        J = np.diag([1.0, rr, dthdx2[j, i], 1.0])
        gc = J.T @ g @ J
        gi = np.linalg.inv(gc)
        gg = G()
        gg.alpha = 1 / math.sqrt(-gi[0, 0])
        gg.beta = list(gi[0, 1:] * gg.alpha ** 2)
        gg.np = lambda name, gc=gc: gc
        _, _, bsq, _ = PH.four_vectors(gg, prim[:, 0, j, i], True)
        bsq_max = max(bsq_max, bsq)
    assert abs(pmax / (0.5 * bsq_max) / WL.TORUS["beta"] - 1) < 1e-10
