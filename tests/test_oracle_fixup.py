"""Pins the oracle's fixups (orc_c2p_fixup_block, fixup_c2p.cpp:44-356; orc_apply_floors_block,
fixup.cpp:285-539) with closed forms: the average of a linear field over the face
neighbours is its centre value, the near-vacuum state, the ceiling rescale
sqrt(gmax^2-1)/(W^2-1), a floor injected at rest, the MHD b^2/rho ceiling, the
floor-state fallback when the inversion fails, and conserved variables checked against
an independent T^{mu nu} contraction (tests/_physics.py).  Runs on the CPU."""
import numpy as np
import pytest

import _oracle as O
import _physics as PH

SMALL = 10.0 * np.finfo(np.float64).tiny


def _block(ndim, n=4):
    nx = [n] * ndim
    return O.make_block(ndim, nx, 4, xmin=(0.1, 0.2, 0.3), dx=(0.1, 0.1, 0.1))


def _fields(p, b, prim):
    """cons, temp, sig, mu from prim by the oracle's own p2c (a consistent state)."""
    cons, sig = O.prim_to_con_block(p, b, prim)
    temp = prim[4] / prim[0]  # cv = 1
    sig3 = np.zeros((3,) + O.block_shape(b))
    sig3[:b.ndim] = sig
    return cons, np.ascontiguousarray(temp), np.ascontiguousarray(sig3[:b.ndim]), np.zeros(O.block_shape(b))


def _linear_prims(b, mhd, gamma):
    sh = O.block_shape(b)
    k, j, i = np.meshgrid(*[np.arange(s, dtype=float) for s in sh], indexing="ij")
    npr = O.nprim(mhd)
    prim = np.zeros((npr,) + sh)
    prim[0] = 1.0 + 0.01 * i + 0.02 * j + 0.03 * k
    prim[1] = 0.1 + 0.001 * i - 0.002 * j + 0.0005 * k
    prim[2] = -0.05 + 0.003 * i + 0.001 * j
    prim[3] = 0.02 - 0.001 * k + 0.002 * i
    prim[4] = 2.0 + 0.05 * i - 0.01 * j + 0.02 * k
    if mhd:
        prim[5], prim[6], prim[7] = 0.3, -0.2, 0.1
    prs = 8 if mhd else 5
    prim[prs] = (gamma - 1.0) * prim[4]
    prim[prs + 1] = gamma
    return np.ascontiguousarray(prim)


def _centre(b):
    return tuple(b.ng + 1 if d < b.ndim else 0 for d in (2, 1, 0))


@pytest.mark.parametrize("ndim", [1, 2, 3])
@pytest.mark.parametrize("mhd", [False, True])
def test_fixup_linear_field_average_is_centre_value(ndim, mhd):
    gamma = 4.0 / 3.0
    p = O.default_params(mhd=mhd, gamma=gamma)
    b = _block(ndim)
    prim = _linear_prims(b, mhd, gamma)
    cons, temp, sig, _ = _fields(p, b, prim)
    expect = prim.copy()
    fail = np.ones(O.block_shape(b))
    c = _centre(b)
    fail[c] = 0.0
    prim[(slice(None),) + c] = 7.5  # garbage the inversion left behind
    before = [a.copy() for a in (prim, cons, temp, sig)]
    O.c2p_fixup_block(p, O.fixup(), b, prim, cons, temp, sig, fail)
    for m in range(5):  # rho, v, u: the linear field's centre value
        assert abs(prim[(m,) + c] - expect[(m,) + c]) < 1e-14 * (1 + abs(expect[(m,) + c]))
    prs = 8 if mhd else 5
    assert abs(prim[(prs,) + c] - (gamma - 1) * prim[(4,) + c]) < 1e-14 * prim[(prs,) + c]
    assert abs(prim[(prs + 1,) + c] - gamma) < 1e-14
    assert abs(temp[c] - prim[(4,) + c] / prim[(0,) + c]) < 1e-14 * temp[c]
    # conserved variables of the repaired state, independent T^{mu nu} form (B not rewritten)
    g = O.point_geom(p, (0.25, 0.3, 0.4))  # Minkowski: the same metric everywhere
    ref = PH.valencia_cons(g, np.diag([-1.0, 1, 1, 1]), prim[(slice(None),) + c], mhd)
    got = cons[(slice(None),) + c]
    assert np.max(np.abs(got[:5] - ref[:5])) < 1e-12 * np.max(np.abs(ref))
    if mhd:
        assert np.array_equal(got[5:], before[1][(slice(5, 8),) + c])
    # every other cell untouched
    mask = np.ones(O.block_shape(b), bool)
    mask[c] = False
    for a0, a1 in zip(before, (prim, cons, temp, sig)):
        assert np.array_equal(a0[..., mask], a1[..., mask])


def test_fixup_skips_failed_neighbours():
    p = O.default_params(mhd=True, gamma=5.0 / 3.0)
    b = _block(3)
    prim = _linear_prims(b, True, 5.0 / 3.0)
    cons, temp, sig, _ = _fields(p, b, prim)
    ref = prim.copy()
    fail = np.ones(O.block_shape(b))
    c = _centre(b)
    fail[c] = 0.0
    fail[(c[0], c[1], c[2] + 1)] = 0.0  # a failed interior neighbour (i + 1)
    O.c2p_fixup_block(p, O.fixup(), b, prim, cons, temp, sig, fail)
    nbrs = [(c[0], c[1], c[2] - 1), (c[0], c[1] - 1, c[2]), (c[0], c[1] + 1, c[2]),
            (c[0] - 1, c[1], c[2]), (c[0] + 1, c[1], c[2])]
    for m in range(5):
        want = np.mean([ref[(m,) + q] for q in nbrs])
        assert abs(prim[(m,) + c] - want) < 1e-14 * (1 + abs(want))


@pytest.mark.parametrize("strategy", [O.FAILURE_INTERPOLATE, O.FAILURE_FLOORS])
def test_fixup_near_vacuum_state(strategy):
    p = O.default_params(mhd=False, gamma=5.0 / 3.0)
    b = _block(2)
    prim = _linear_prims(b, False, 5.0 / 3.0)
    cons, temp, sig, _ = _fields(p, b, prim)
    fail = np.ones(O.block_shape(b))
    c = _centre(b)
    fail[c] = 0.0
    if strategy == O.FAILURE_INTERPOLATE:  # no valid neighbour
        for q in [(c[0], c[1] - 1, c[2]), (c[0], c[1] + 1, c[2]), (c[0], c[1], c[2] - 1),
                  (c[0], c[1], c[2] + 1)]:
            fail[q] = 0.0
    O.c2p_fixup_block(p, O.fixup(strategy=strategy), b, prim, cons, temp, sig, fail)
    assert prim[(0,) + c] == 100 * SMALL and prim[(4,) + c] == 100 * SMALL
    assert prim[(1,) + c] == prim[(2,) + c] == prim[(3,) + c] == 0.0
    assert cons[(0,) + c] == 100 * SMALL  # D = rho at rest in Minkowski


def test_fixup_gamma_ceiling_rescale_quirk():
    gmax = 3.0
    p = O.default_params(mhd=False, gamma=5.0 / 3.0)
    p.bounds.gam_max = gmax
    b = _block(1)
    prim = _linear_prims(b, False, 5.0 / 3.0)
    prim[1] = 4.0  # W v^1 = 4 everywhere -> W = sqrt(17) > gmax
    prim[2] = prim[3] = 0.0
    cons, temp, sig, _ = _fields(p, b, prim)
    fail = np.ones(O.block_shape(b))
    c = _centre(b)
    fail[c] = 0.0
    O.c2p_fixup_block(p, O.fixup(), b, prim, cons, temp, sig, fail)
    W = np.sqrt(1.0 + 16.0)
    rescale = np.sqrt(gmax * gmax - 1.0) / (W * W - 1.0)  # not sqrt of the quotient
    assert abs(prim[(1,) + c] - 4.0 * rescale) < 1e-15 * 4.0 * rescale


# ------------------------------------------------------------------------------ floors ----
def _rest_state(b, mhd, rho, u, gamma, v=(0.0, 0.0, 0.0), B=(0.0, 0.0, 0.0)):
    sh = O.block_shape(b)
    npr = O.nprim(mhd)
    prim = np.zeros((npr,) + sh)
    prim[0], prim[4] = rho, u
    for d in range(3):
        prim[1 + d] = v[d]
    if mhd:
        for d in range(3):
            prim[5 + d] = B[d]
    prs = 8 if mhd else 5
    prim[prs] = (gamma - 1) * u
    prim[prs + 1] = gamma
    return np.ascontiguousarray(prim)


def _floor_params(mhd, gamma, rho0, sie0, gmax=1e300, emax=1e300):
    p = O.default_params(mhd=mhd, gamma=gamma)
    p.bounds.floor_type = O.FLOOR_CONSTANT
    p.bounds.r0, p.bounds.s0 = rho0, sie0
    p.bounds.gam_max, p.bounds.sie_max = gmax, emax
    p.c2p_floor_scale_fac = 0.5  # the inversion's own floors stay below the injected state
    return p


def test_floors_leave_valid_state_untouched():
    gamma = 4.0 / 3.0
    p = _floor_params(True, gamma, 1e-6, 1e-4, gmax=50.0, emax=100.0)
    b = _block(3)
    prim = _linear_prims(b, True, gamma)
    cons, temp, sig, mu = _fields(p, b, prim)
    before = [a.copy() for a in (prim, cons, temp, mu, sig)]
    O.apply_floors_block(p, O.fixup(), b, prim, cons, temp, mu, sig)
    for a0, a1 in zip(before, (prim, cons, temp, mu, sig)):
        assert np.array_equal(a0, a1)


def test_floor_injected_at_rest_closed_form():
    gamma, rho0, sie0 = 5.0 / 3.0, 1e-3, 0.5
    rho, u = 2e-4, 0.1  # below the density floor, above rho0 * sie0 in energy
    p = _floor_params(False, gamma, rho0, sie0)
    b = _block(1)
    prim = _rest_state(b, False, rho, u, gamma)
    cons, temp, sig, mu = _fields(p, b, prim)
    O.apply_floors_block(p, O.fixup(), b, prim, cons, temp, mu, sig)
    drho = rho0 - rho
    du = sie0 * drho  # max(rho0 sie0 - u, sie0 drho)
    assert np.allclose(prim[0], rho0, rtol=1e-9)
    assert np.allclose(prim[4], u + du, rtol=1e-9)
    assert np.max(np.abs(prim[1:4])) < 1e-12
    assert np.allclose(cons[0], rho + drho, rtol=1e-15)  # D += drho (W = 1, sqrt(gamma) = 1)
    assert np.allclose(cons[4], u + du, rtol=1e-12)
    assert np.allclose(prim[5], (gamma - 1) * prim[4], rtol=1e-14)


def test_mhd_bsq_over_rho_ceiling():
    gamma, bsqorho = 4.0 / 3.0, 50.0
    p = _floor_params(True, gamma, 1e-12, 1e-12)
    b = _block(1)
    B = (0.6, 0.0, 0.8)  # b^2 = 1 at rest
    prim = _rest_state(b, True, 1e-3, 1e-2, gamma, B=B)
    cons, temp, sig, mu = _fields(p, b, prim)
    O.apply_floors_block(p, O.fixup(mhd_floors=True, bsqorho_max=bsqorho), b, prim, cons, temp,
                         mu, sig)
    assert np.allclose(prim[0], 1.0 / bsqorho, rtol=1e-9)  # rho raised to b^2 / bsqorho_max
    assert np.array_equal(prim[5:8], _rest_state(b, True, 1, 1, gamma, B=B)[5:8])


def test_energy_ceiling_rescale_quirk():
    gamma, gmax, emax = 5.0 / 3.0, 10.0, 1.0
    p = _floor_params(False, gamma, 1e-12, 1e-12, gmax=gmax, emax=emax)
    b = _block(1)
    vp = (0.1, 0.0, 0.0)
    prim = _rest_state(b, False, 1.0, 3.0, gamma, v=vp)  # u / rho = 3 > emax, W < gmax
    cons, temp, sig, mu = _fields(p, b, prim)
    O.apply_floors_block(p, O.fixup(), b, prim, cons, temp, mu, sig)
    W = np.sqrt(1.0 + 0.1 * 0.1)
    rescale = np.sqrt(gmax * gmax - 1.0) / (W * W - 1.0)  # ~995: the quirk speeds the flow up
    assert np.allclose(prim[1], 0.1 * rescale, rtol=1e-14)
    assert np.allclose(prim[4], 3.0, rtol=0)  # the energy is not capped
    g = O.point_geom(p, (0.15, 0.0, 0.0))
    ref = PH.valencia_cons(g, np.diag([-1.0, 1, 1, 1]), prim[:, 0, 0, 4], False)
    assert np.max(np.abs(cons[:, 0, 0, 4] - ref)) < 1e-11 * np.max(np.abs(ref))


def test_floor_fallback_when_inversion_fails():
    gamma, rho0, sie0 = 5.0 / 3.0, 1e-3, 0.5
    p = _floor_params(False, gamma, rho0, sie0)
    b = _block(1)
    prim = _rest_state(b, False, 2e-4, 0.1, gamma)
    cons, temp, sig, mu = _fields(p, b, prim)
    cons[4] = np.nan  # the inversion rejects the state up front
    O.apply_floors_block(p, O.fixup(), b, prim, cons, temp, mu, sig)
    drho = rho0 - 2e-4
    du = sie0 * drho
    assert np.all(prim[0] == drho)  # the reference sets the increments, not the floors
    assert np.all(prim[4] == du)
    assert np.all(prim[1:4] == 0.0)
    assert np.all(cons[0] == drho)
    assert np.allclose(cons[4], du, rtol=1e-15)
    assert np.allclose(temp, du / drho, rtol=1e-15)


def test_floors_disabled_is_noop():
    p = _floor_params(False, 5.0 / 3.0, 1e-3, 0.5)
    b = _block(1)
    prim = _rest_state(b, False, 2e-4, 0.1, 5.0 / 3.0)
    cons, temp, sig, mu = _fields(p, b, prim)
    before = prim.copy()
    O.apply_floors_block(p, O.fixup(c2p_fixup=False, floors=False), b, prim, cons, temp, mu, sig)
    assert np.array_equal(prim, before)
    # enable_c2p_fixup forces the floors on (fixup.cpp:120-124)
    O.apply_floors_block(p, O.fixup(c2p_fixup=True, floors=False), b, prim, cons, temp, mu, sig)
    assert np.allclose(prim[0], 1e-3, rtol=1e-9)
