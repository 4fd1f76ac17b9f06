"""Pins for the oracle functions that equal-state and Minkowski checks leave open.

- WENO5-Z's blend with the MC-limited linear reconstruction (reconstruction.hpp:159-182):
  exact rational evaluation of the published weights on integer stencils where the blend
  weight lies strictly inside (0, 1).
- HLL and LLF dissipation in curved metrics (riemann.hpp:257-301): the oracle's face flux
  must be a combination of the two physical fluxes sqrt(-g) F and the jump of the
  densitized state sqrt(gamma) U (both from tests/_physics.py, not from the oracle) whose
  coefficients are the HLL (or Rusanov) weights of ONE pair of wave speeds, the largest of
  which is the speed the oracle returns.
- HLL = LLF on mirror-symmetric faces (c_l = -c_r), which pins LLF's dissipation against
  HLL's.
- The Valencia energy source (fluid.cpp:612-640) against the 3+1 extrinsic-curvature form
  of Banyuls et al. (1997), ApJ 476, 221 (also Font, Living Rev. Relativ. 11 (2008) 7,
  eq. 36), with K_ij from finite differences of the pinned metric.

Each check fails under the mutations listed with it; runs on the CPU.
"""
import itertools
import math
from fractions import Fraction as Fr

import numpy as np
import pytest

import _oracle as O
import _physics as PH


# ---------------------------------------------------------------- WENO5-Z blend ---------
def _mc(dm, dp, alpha):
    dc = (dm + dp) / 2 if dm * dp > 0 else Fr(0)
    m = min(abs(dc), alpha * min(abs(dm), abs(dp)))
    return m if dc >= 0 else -m


def _weno5z_exact(q, mc_alpha=Fr(2), blend=True):
    """WENO-Z (Borges et al. 2008, tau5 = |beta2 - beta0|) in exact rationals, blended with
    MC-limited PLM by the harmonic mean of each side's 'linearity'
    A = 3 / ((sum_k w_k)(sum_k gamma_k / w_k)) (A = 1 exactly when the nonlinear weights
    equal the linear ones).  eps = 1e-100 of the reference is dropped (it moves O(1)
    results by ~1e-100 relative)."""
    q = [Fr(x) for x in q]
    t = Fr(13, 3)
    b = [t * (q[0] - 2 * q[1] + q[2]) ** 2 + (q[0] - 4 * q[1] + 3 * q[2]) ** 2,
         t * (q[1] - 2 * q[2] + q[3]) ** 2 + (q[3] - q[1]) ** 2,
         t * (q[2] - 2 * q[3] + q[4]) ** 2 + (q[4] - 4 * q[3] + 3 * q[2]) ** 2]
    tau = abs(b[2] - b[0])
    g = [Fr(1, 10), Fr(6, 10), Fr(3, 10)]
    s = [1 + tau / bk for bk in b]
    # third-order candidate edge values of the three sub-stencils
    up = [Fr(1, 3) * q[0] - Fr(7, 6) * q[1] + Fr(11, 6) * q[2],
          -Fr(1, 6) * q[1] + Fr(5, 6) * q[2] + Fr(1, 3) * q[3],
          Fr(1, 3) * q[2] + Fr(5, 6) * q[3] - Fr(1, 6) * q[4]]
    lo = [Fr(1, 3) * q[4] - Fr(7, 6) * q[3] + Fr(11, 6) * q[2],
          -Fr(1, 6) * q[3] + Fr(5, 6) * q[2] + Fr(1, 3) * q[1],
          Fr(1, 3) * q[2] + Fr(5, 6) * q[1] - Fr(1, 6) * q[0]]

    def side(w, cand):
        W = sum(w)
        return sum(wk * ck for wk, ck in zip(w, cand)) / W, 3 / (W * sum(gk / wk for gk, wk in zip(g, w)))

    ql, al = side([g[0] * s[0], g[1] * s[1], g[2] * s[2]], up)
    qr, ar = side([g[0] * s[2], g[1] * s[1], g[2] * s[0]], lo)
    A = 2 * al * ar / (al + ar)
    if not blend:
        return ql, qr, A
    dq = _mc(q[2] - q[1], q[3] - q[2], mc_alpha)
    return A * ql + (1 - A) * (q[2] + dq / 2), A * qr + (1 - A) * (q[2] - dq / 2), A


def _blend_cases():
    out = []
    for q in itertools.product(range(5), repeat=5):
        q2 = [Fr(x) for x in q]
        b = [(q2[0] - 2 * q2[1] + q2[2]) ** 2 + (q2[0] - 4 * q2[1] + 3 * q2[2]) ** 2,
             (q2[1] - 2 * q2[2] + q2[3]) ** 2 + (q2[3] - q2[1]) ** 2,
             (q2[2] - 2 * q2[3] + q2[4]) ** 2 + (q2[4] - 4 * q2[3] + 3 * q2[2]) ** 2]
        if min(b) == 0:
            continue  # eps = 1e-100 decides these; exact arithmetic has no such term
        ql, qr, A = _weno5z_exact(q)
        if Fr(1, 5) < A < Fr(4, 5):
            out.append((q, ql, qr, A))
    return out


def test_weno5z_blend_exact_rationals():
    cases = _blend_cases()
    assert len(cases) > 100
    discriminating = 0
    for q, ql, qr, A in cases:
        l, r = O.recon(O.RECON_WENO5, q)
        assert abs(l - float(ql)) <= 4e-15 * max(1.0, abs(float(ql))), (q, l, float(ql))
        assert abs(r - float(qr)) <= 4e-15 * max(1.0, abs(float(qr))), (q, r, float(qr))
        # the same stencil without the blend, or with the MC slope factor 1.5, gives an
        # answer the oracle does not
        l0, r0, _ = _weno5z_exact(q, blend=False)
        l15, r15, _ = _weno5z_exact(q, mc_alpha=Fr(3, 2))
        if (abs(float(l0 - ql)) + abs(float(r0 - qr)) > 1e-3
                and abs(float(l15 - ql)) + abs(float(r15 - qr)) > 1e-3):
            discriminating += 1
    assert discriminating >= 20, discriminating


def test_weno5z_blend_hand_case():
    # q = (0, 0, 1, 4, 0): beta = (16, 76/3, 592/3) (with 13/3 a^2 + b^2), tau5 = 544/3;
    # the blend weight A = 0.3588 lies well inside (0, 1), the MC slope is
    # mc(1, 3, 2) = min(2, 2 * 1) = 2, so the linear part is 1 +- 1.
    ql, qr, A = _weno5z_exact((0, 0, 1, 4, 0))
    assert 0.35 < A < 0.37
    l, r = O.recon(O.RECON_WENO5, (0, 0, 1, 4, 0))
    assert abs(l - float(ql)) < 1e-15 and abs(r - float(qr)) < 1e-15
    lin_l, lin_r = 2.0, 0.0
    assert abs(l - (float(A) * float(_weno5z_exact((0, 0, 1, 4, 0), blend=False)[0])
                    + (1 - float(A)) * lin_l)) < 1e-15
    assert abs(r - (float(A) * float(_weno5z_exact((0, 0, 1, 4, 0), blend=False)[1])
                    + (1 - float(A)) * lin_r)) < 1e-15


# ---------------------------------------------------------------- Riemann dissipation ---
def _states(rng, n, mhd, p, X, wmax=3.0, sigmax=1.0):
    g = O.point_geom(p, X)
    return PH.to_coordinate_frame(PH.random_prims(rng, n, mhd, 13 / 9, wmax=wmax, sigmax=sigmax),
                                  g.np("gam"), mhd)


def _fit(f, GL, GR, dQ):
    """Least-squares coefficients (a, b, c) of f = a GL + b GR + c dQ and the relative
    residual."""
    M = np.stack([GL, GR, dQ], axis=1)
    sc = np.max(np.abs(M), axis=0)
    coef, *_ = np.linalg.lstsq(M / sc, f, rcond=None)
    coef = coef / sc
    res = np.max(np.abs(M @ coef - f)) / np.max(np.abs(np.concatenate([GL, GR, dQ])))
    return coef, res


CURVED = [(O.GEOM_SPHKS, (5.0, 0.9, 0.0)), (O.GEOM_SPHKS, (2.2, 2.3, 0.0)),
          (O.GEOM_FMKS, (math.log(4.0), 0.3, 0.0))]


@pytest.mark.parametrize("geom,X", CURVED)
@pytest.mark.parametrize("d", [0, 1, 2])
def test_hll_curved_textbook_form(geom, X, d):
    """HLL (Harten, Lax & van Leer 1983) on the densitized system
    d_t(sqrt(gamma) U) + d_i(sqrt(-g) F^i) = S:
    F* = (c_r G_L - c_l G_R + c_r c_l (Q_R - Q_L)) / (c_r - c_l), c_l <= 0 <= c_r,
    returned speed max(-c_l, c_r).  Fails if sqrt(gamma) and sqrt(-g) are swapped in the
    dissipation term (alpha != 1 here), if its sign flips, or if the speeds of the flux
    and dissipation terms disagree."""
    p = O.default_params(mhd=True, gamma=13 / 9, geom=geom, a=0.9375, riemann=O.RIEMANN_HLL)
    g = O.point_geom(p, X)
    assert abs(g.alpha - 1.0) > 0.05
    gi = O.geom_gcon4(p, X)
    rng = np.random.default_rng(100 + d)
    L, R = _states(rng, 12, True, p, X), _states(rng, 12, True, p, X)
    for s in range(L.shape[1]):
        f, sig = O.face_flux(p, d, X, L[:, s], R[:, s])
        GL, GR = PH.valencia_flux(g, gi, L[:, s], True, d), PH.valencia_flux(g, gi, R[:, s], True, d)
        dQ = PH.valencia_cons(g, gi, R[:, s], True) - PH.valencia_cons(g, gi, L[:, s], True)
        (a, b, c), res = _fit(f, GL, GR, dQ)
        assert res < 1e-11, res
        assert abs(a + b - 1) < 1e-10
        assert a >= -1e-12 and b >= -1e-12
        if min(a, b) < 1e-9:
            # every wave moves one way (c_l = 0 or c_r = 0): pure upwinding, no
            # dissipation term, and the speed is not recoverable from the flux
            assert abs(c) * np.max(np.abs(dQ)) < 1e-11 * np.max(np.abs(np.concatenate([GL, GR])))
            continue
        delta = -c / (a * b)  # c_r - c_l
        cr, cl = a * delta, -b * delta
        assert delta > 0 and cl <= 1e-12 and cr >= -1e-12
        assert abs(max(-cl, cr) - sig) < 1e-9 * max(sig, 1e-3), (cl, cr, sig)


@pytest.mark.parametrize("geom,X", CURVED)
@pytest.mark.parametrize("d", [0, 1, 2])
def test_llf_curved_textbook_form(geom, X, d):
    """Rusanov / LLF: F* = (G_L + G_R)/2 - c_max (Q_R - Q_L)/2 with the returned speed
    c_max.  Fails on a flipped dissipation sign, a missing 1/2 or a swapped
    sqrt(gamma) / sqrt(-g)."""
    p = O.default_params(mhd=True, gamma=13 / 9, geom=geom, a=0.9375, riemann=O.RIEMANN_LLF)
    g = O.point_geom(p, X)
    gi = O.geom_gcon4(p, X)
    rng = np.random.default_rng(200 + d)
    L, R = _states(rng, 12, True, p, X), _states(rng, 12, True, p, X)
    for s in range(L.shape[1]):
        f, sig = O.face_flux(p, d, X, L[:, s], R[:, s])
        GL, GR = PH.valencia_flux(g, gi, L[:, s], True, d), PH.valencia_flux(g, gi, R[:, s], True, d)
        dQ = PH.valencia_cons(g, gi, R[:, s], True) - PH.valencia_cons(g, gi, L[:, s], True)
        ref = 0.5 * (GL + GR) - 0.5 * sig * dQ
        scale = np.max(np.abs(np.concatenate([GL, GR, sig * dQ])))
        assert np.max(np.abs(f - ref)) < 1e-11 * scale
        assert sig > 0


def _mirror(prim, d):
    """The state reflected through the face normal to direction d (Minkowski): v^d and
    B^d change sign."""
    m = prim.copy()
    m[1 + d] = -m[1 + d]
    m[5 + d] = -m[5 + d]
    return m


@pytest.mark.parametrize("mhd", [False, True])
@pytest.mark.parametrize("d", [0, 1, 2])
def test_hll_equals_llf_on_mirror_symmetric_faces(mhd, d):
    """With R the mirror image of L, v-_R = -v+_L and v+_R = -v-_L, so c_l = -c_r = -c_max
    and the HLL formula reduces to LLF's exactly (up to rounding).  Dissipation matters
    here: the jump U_R - U_L is non-zero in S_d and B^d."""
    ph = O.default_params(mhd=mhd, gamma=5 / 3, riemann=O.RIEMANN_HLL)
    pl = O.default_params(mhd=mhd, gamma=5 / 3, riemann=O.RIEMANN_LLF)
    rng = np.random.default_rng(300 + d + 3 * mhd)
    L = PH.random_prims(rng, 20, mhd, 5 / 3, wmax=3.0)
    for s in range(L.shape[1]):
        q = L[:, s]
        if not mhd:
            r = q.copy()
            r[1 + d] = -r[1 + d]
        else:
            r = _mirror(q, d)
        fh, sh = O.face_flux(ph, d, (0, 0, 0), q, r)
        fl, sl = O.face_flux(pl, d, (0, 0, 0), q, r)
        assert abs(sh - sl) <= 1e-15 * sh
        diss = np.max(np.abs(fl - 0.5 * (PH.valencia_flux(O.point_geom(pl, (0, 0, 0)),
                                                           O.geom_gcon4(pl, (0, 0, 0)), q, mhd, d)
                                          + PH.valencia_flux(O.point_geom(pl, (0, 0, 0)),
                                                             O.geom_gcon4(pl, (0, 0, 0)), r, mhd, d))))
        assert diss > 1e-3  # the dissipation term is not negligible on these faces
        assert np.max(np.abs(fh - fl)) < 1e-13 * np.max(np.abs(fl)), (fh, fl)


# ---------------------------------------------------------------- energy source ---------
def _d4(fn, X, c, h):
    """Fourth-order central difference of fn along coordinate c."""
    def at(off):
        Y = list(X)
        Y[c] += off
        return np.asarray(fn(Y), float)
    return (8 * (at(h) - at(-h)) - (at(2 * h) - at(-2 * h))) / (12 * h)


def _adm(p, X):
    g = O.point_geom(p, X)
    gam = g.np("gam")
    return g.alpha, np.array(g.beta[:]), gam


def _extrinsic_curvature(p, X, h):
    """K_ij = (D_i beta_j + D_j beta_i) / (2 alpha) for a stationary metric (sign
    convention d_t gamma_ij = -2 alpha K_ij + D_i beta_j + D_j beta_i)."""
    alpha, beta, gam = _adm(p, X)
    gami = np.linalg.inv(gam)
    dgam = np.stack([_d4(lambda Y: _adm(p, Y)[2], X, c, h) for c in range(3)])  # [c][i][j]
    blo = lambda Y: _adm(p, Y)[2] @ _adm(p, Y)[1]
    dblo = np.stack([_d4(blo, X, c, h) for c in range(3)])  # [c][j] = d_c beta_j
    # 3-Christoffel Gamma^k_ij
    G = 0.5 * np.einsum("kl,ilj->kij", gami, dgam) \
        + 0.5 * np.einsum("kl,jli->kij", gami, dgam) \
        - 0.5 * np.einsum("kl,lij->kij", gami, dgam)
    bl = gam @ beta
    Db = dblo - np.einsum("kij,k->ij", G, bl)  # D_i beta_j
    K = (Db + Db.T) / (2 * alpha)
    dalpha = np.array([_d4(lambda Y: _adm(p, Y)[0], X, c, h) for c in range(3)])
    return alpha, beta, K, dalpha


@pytest.mark.parametrize("geom,X,h", [(O.GEOM_SPHKS, (6.0, 1.0, 0.0), 1e-3),
                                      (O.GEOM_SPHKS, (2.5, 0.4, 0.0), 1e-3),
                                      (O.GEOM_FMKS, (math.log(8.0), 0.35, 0.0), 1e-3)])
def test_energy_source_matches_extrinsic_curvature_form(geom, X, h):
    """S_tau = sqrt(-g) [T^00 (beta^i beta^j K_ij - beta^i d_i alpha)
                          + T^0i (2 beta^j K_ij - d_i alpha) + T^ij K_ij].
    Fails if the T^{mu 0} d_mu ln(alpha) term is dropped, if Gamma^0_{mn} is contracted
    with the wrong index, or if the overall alpha is missing."""
    p = O.default_params(mhd=True, gamma=4 / 3, geom=geom, a=0.9375)
    g = O.point_geom(p, X)
    gi = O.geom_gcon4(p, X)
    alpha, beta, K, da = _extrinsic_curvature(p, X, h)
    assert np.max(np.abs(K)) > 1e-2  # a metric with a shift: K does not vanish
    rng = np.random.default_rng(17)
    prims = PH.to_coordinate_frame(PH.random_prims(rng, 8, True, 4 / 3, wmax=3.0),
                                   g.np("gam"), True)
    for s in range(prims.shape[1]):
        T = PH.tmunu(g, gi, prims[:, s], True)
        ref = g.detg * (T[0, 0] * (beta @ K @ beta - beta @ da)
                        + T[0, 1:] @ (2 * K @ beta - da)
                        + np.sum(T[1:, 1:] * K))
        src = O.sources(p, X, prims[:, s])
        scale = g.detg * np.max(np.abs(T)) * max(np.max(np.abs(K)), np.max(np.abs(da)))
        tol = 1e-7 if geom == O.GEOM_FMKS else 1e-9
        assert abs(src[4] - ref) < tol * scale, (src[4], ref, scale)
