"""Fixups on the GPU (include/phx_fixup.h) vs the CPU oracle on identical seeded blocks:
the con2prim failure repair (fixup_c2p.cpp:44-356) and the floors / ceilings
(fixup.cpp:285-539).  Minkowski results bit-exact; FMKS within 1e-12 relative (device
sin/cos/exp/pow differ from glibc by an ulp), con2prim inside the floors included."""
import sys

import numpy as np
import pytest

import _oracle as O
import _physics as PH
from phoebus_amd import _abi as A
from phoebus_amd.api import BlockState, Phx, fixup_params
from phoebus_amd.device import synchronize
from test_gpu_parity import assert_match, geom_block, geom_params, oblock, oparams

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def phx():
    ctx = Phx(0)
    yield ctx
    ctx.close()


def ofixup(fx):
    return O.Fixup(fx.enable_c2p_fixup, fx.failure_strategy, fx.enable_floors,
                   fx.enable_mhd_floors, fx.bsqorho_max, fx.bsqou_max)


def random_block_state(rng, p, b, gamma):
    """Random primitives over the entire block (coordinate frame for curved metrics),
    their conserved variables and signal speeds by the oracle, temperatures, a warm mu."""
    op, ob = oparams(p), oblock(b)
    sh = O.block_shape(ob)
    n = int(np.prod(sh))
    mhd = bool(p.mhd)
    prim = PH.random_prims(rng, n, mhd, gamma, wmax=5.0, sigmax=10.0)
    if p.geom.kind != A.GEOM_MINKOWSKI:
        prim[1:4] *= 0.3
        if mhd:
            prim[5:8] *= 0.3
    prim = np.ascontiguousarray(prim.reshape((-1,) + sh))
    cons, sig = O.prim_to_con_block(op, ob, prim)
    temp = np.ascontiguousarray(prim[4] / prim[0] / p.cv)
    mu = np.ascontiguousarray(rng.uniform(0.0, 1.0, sh))
    return prim, cons, np.ascontiguousarray(sig), temp, mu


def upload(st, prim, cons, sig, temp, mu, fail=None):
    st.prim.upload(prim)
    st.cons.upload(cons)
    st.cell_sig.upload(sig)
    st.temp.upload(temp)
    st.c2p_mu.upload(mu)
    if fail is not None:
        st.fail.upload(fail)
    synchronize()


CASES = [(3, A.GEOM_MINKOWSKI, True), (3, A.GEOM_MINKOWSKI, False), (2, A.GEOM_MINKOWSKI, True),
         (1, A.GEOM_MINKOWSKI, False), (3, A.GEOM_FMKS, True)]


@pytest.mark.parametrize("ndim,geom,mhd", CASES)
@pytest.mark.parametrize("strategy", [A.FAILURE_INTERPOLATE, A.FAILURE_FLOORS])
def test_c2p_fixup_parity(phx, ndim, geom, mhd, strategy):
    say(f"[gpu] c2p fixup ndim={ndim} geom={geom} mhd={mhd} strategy={strategy}")
    rng = np.random.default_rng(7000 + 10 * ndim + 3 * geom + mhd + 100 * strategy)
    gamma = 13.0 / 9.0
    p = geom_params(geom, gamma=gamma, mhd=mhd, ceilings=(3.0, 1e300))
    nx = {1: (64,), 2: (16, 12), 3: (12, 10, 8)}[ndim]
    b = geom_block(ndim, nx, geom)
    prim, cons, sig, temp, mu = random_block_state(rng, p, b, gamma)
    fail = np.ascontiguousarray((rng.uniform(size=temp.shape) > 0.2).astype(float))
    fail.reshape(-1)[: fail.size // 8] = 0.0  # a slab of failures: cells with no valid neighbour
    st = BlockState(p, b)
    upload(st, prim, cons, sig, temp, mu, fail)
    fx = fixup_params(c2p_fixup=True, strategy=strategy)
    phx.c2p_fixup(p, fx, b, st.fields())
    synchronize()
    op, ob = oparams(p), oblock(b)
    O.c2p_fixup_block(op, ofixup(fx), ob, prim, cons, temp, sig, fail)
    exact = geom == A.GEOM_MINKOWSKI
    assert_match(st.prim.get(), prim, exact, "prim", atol_scale=1e-300)
    assert_match(st.cons.get(), cons, exact, "cons", atol_scale=np.max(np.abs(cons), axis=0))
    assert_match(st.temp.get(), temp, exact, "temp")
    assert_match(st.cell_sig.get(), sig, exact, "cell_sig")
    say(f"    repaired {int(np.sum(fail[O.interior(ob)] == 0))} cells")


def floor_params(geom, mhd, prim, gamma, floor_type=A.FLOOR_CONSTANT_RHO_SIE):
    """Floors and ceilings that a good fraction of the random cells cross."""
    rho0 = float(np.quantile(prim[0], 0.25))
    sie0 = float(np.quantile(prim[4] / prim[0], 0.2))
    emax = float(np.quantile(prim[4] / prim[0], 0.9))
    floors = dict(floor_type=floor_type, rho0=rho0, sie0=sie0, rho_exp=-2.0, sie_exp=-1.0)
    return geom_params(geom, gamma=gamma, mhd=mhd, floors=floors, ceilings=(3.0, emax),
                       c2p_tol=1e-10)


FLOOR_CASES = [(3, A.GEOM_MINKOWSKI, True, A.FLOOR_CONSTANT_RHO_SIE),
               (3, A.GEOM_MINKOWSKI, False, A.FLOOR_CONSTANT_RHO_SIE),
               (2, A.GEOM_MINKOWSKI, True, A.FLOOR_CONSTANT_RHO_SIE),
               (1, A.GEOM_MINKOWSKI, False, A.FLOOR_CONSTANT_RHO_SIE),
               (3, A.GEOM_FMKS, True, A.FLOOR_EXPX1_RHO_SIE)]


@pytest.mark.parametrize("ndim,geom,mhd,ftype", FLOOR_CASES)
def test_apply_floors_parity(phx, ndim, geom, mhd, ftype):
    say(f"[gpu] floors ndim={ndim} geom={geom} mhd={mhd} floor_type={ftype}")
    rng = np.random.default_rng(8000 + 10 * ndim + 3 * geom + mhd)
    gamma = 13.0 / 9.0
    p0 = geom_params(geom, gamma=gamma, mhd=mhd)
    nx = {1: (64,), 2: (16, 12), 3: (12, 10, 8)}[ndim]
    b = geom_block(ndim, nx, geom)
    prim, cons, sig, temp, mu = random_block_state(rng, p0, b, gamma)
    p = floor_params(geom, mhd, prim, gamma, ftype)
    cons[4].reshape(-1)[::97] = np.nan  # inversions that fail: the floor-state fallback
    st = BlockState(p, b)
    upload(st, prim, cons, sig, temp, mu)
    fx = fixup_params(c2p_fixup=True, mhd_floors=mhd, bsqorho_max=5.0, bsqou_max=100.0)
    phx.apply_floors(p, fx, b, st.fields())
    synchronize()
    before = prim.copy()
    op, ob = oparams(p), oblock(b)
    O.apply_floors_block(op, ofixup(fx), ob, prim, cons, temp, mu, sig)
    exact = geom == A.GEOM_MINKOWSKI
    got = st.prim.get()
    if exact:
        assert_match(got, prim, True, "prim")
        assert_match(st.cons.get(), cons, True, "cons")
        assert_match(st.temp.get(), temp, True, "temp")
        assert_match(st.c2p_mu.get(), mu, True, "mu")
        assert_match(st.cell_sig.get(), sig, True, "cell_sig")
    else:
        # the inversion inside the floors may stop one iteration apart: compare at its
        # tolerance where both converged, and require near-total agreement
        close = np.isclose(got, prim, rtol=1e-8, atol=0, equal_nan=True)
        assert close.mean() > 0.995, f"prim agreement {close.mean()}"
    changed = np.any(before != prim, axis=0).mean()
    say(f"    floors/ceilings changed {changed:.1%} of the cells")
    assert changed > 0.2


def test_fixup_rejects_bad_input(phx):
    p = geom_params(A.GEOM_MINKOWSKI)
    b = geom_block(3, (8, 8, 8), A.GEOM_MINKOWSKI)
    st = BlockState(p, b)
    bad = fixup_params(strategy=7)
    with pytest.raises(A.PhxError, match="strategy"):
        phx.c2p_fixup(p, bad, b, st.fields())
    with pytest.raises(A.PhxError, match="range"):
        phx.apply_floors(p, fixup_params(), b, st.fields(), lo=(0, 0, 0), hi=(99, 0, 0))
