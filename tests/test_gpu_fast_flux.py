"""The fast flux arithmetic (flux_arith = FLUX_ARITH_FAST: contraction and Newton-refined
reciprocals in the sweeps) against the oracle and against the exact arithmetic.

Bar: fluxes and face signal speeds within 1e-12 of the oracle relative to the largest
magnitude of each flux component over the block (the flux tolerance of the reference
comparison; the differences are below 5e-14: one Newton step leaves reciprocals and square
roots within 2.2e-15 and 4.3e-15 of IEEE); every y/z sweep shape gives the
same bits as every other (the same fast arithmetic per face); the fused stage and a block
pack in the fast arithmetic keep the fail flags of the exact arithmetic and agree with it
to the con2prim tolerance; the pack equals the single-block fast stage bit for bit.
"""
import sys

import numpy as np
import pytest

import _oracle as O
from phoebus_amd import _abi as A
from phoebus_amd.api import BlockState, Pack, Phx, block, fluid_params
from phoebus_amd.device import DeviceArray, synchronize
from test_gpu_parity import (CASES, case_id, consumed_face_slices, geom_block, geom_params,
                             oblock, oparams, random_block)

pytestmark = pytest.mark.gpu

RTOL_FLUX = 1e-12


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def phx():
    ctx = Phx(0)
    yield ctx
    ctx.close()


def _flux_close(g, r, what):
    """|g - r| <= 1e-12 max |r| per component (the last axes are the faces)."""
    scale = np.max(np.abs(r)) if r.size else 0.0
    err = np.max(np.abs(g - r)) if r.size else 0.0
    assert err <= RTOL_FLUX * scale, f"{what}: {err} vs scale {scale}"
    return err / scale if scale > 0 else 0.0


@pytest.mark.parametrize("case", CASES, ids=[case_id(c) for c in CASES])
def test_fast_fluxes_match_oracle(phx, case):
    ndim, nx, mhd, recon, riemann, geom = case
    say(f"[gpu] fast flux arithmetic {case_id(case)}")
    rng = np.random.default_rng(hash(case_id(case)) % 2 ** 32)
    b = geom_block(ndim, nx, geom)
    p = geom_params(geom, gamma=13.0 / 9.0, mhd=mhd, recon=recon, riemann=riemann)
    p.flux_arith = A.FLUX_ARITH_FAST
    op, ob = oparams(p), oblock(b)
    prim = random_block(rng, b, mhd, 13.0 / 9.0, geom)
    st = BlockState(p, b)
    st.prim.upload(prim)
    ref_flux, ref_fsig = O.calculate_fluxes(op, ob, prim)
    phx.calculate_fluxes(p, b, st.fields())
    synchronize()
    gflux, gfsig = st.flux.get(), st.face_sig.get()
    worst = 0.0
    for d in range(ndim):
        sl = consumed_face_slices(b, d)
        for m in range(ref_flux.shape[1]):
            worst = max(worst, _flux_close(gflux[d, m][sl], ref_flux[d, m][sl], f"d={d} m={m}"))
        worst = max(worst, _flux_close(gfsig[d][sl], ref_fsig[d][sl], f"face_sig d={d}"))
    say(f"    largest difference {worst:.2e} of the component scale")


def test_fast_sweep_shapes_agree(phx):
    say("[gpu] fast flux arithmetic: every y/z sweep shape, bit for bit")
    rng = np.random.default_rng(91)
    b = block(3, (24, 20, 18), 4, dx=(0.1, 0.2, 0.3))
    p = fluid_params(gamma=13.0 / 9.0, mhd=True, recon=A.RECON_WENO5Z, riemann=A.RIEMANN_HLL,
                     flux_arith=A.FLUX_ARITH_FAST)
    st = BlockState(p, b)
    st.prim.upload(random_block(rng, b, True, 13.0 / 9.0, A.GEOM_MINKOWSKI))
    f = st.fields()
    first = None
    for ti, ts in [(8, 64), (2, 16), (5, 29), (8, -1), (4, -3), (16, -2), (1, -5), (0, 0)]:
        for d in (1, 2):
            phx.set_tile_shape(p, b, d, ti, ts)
        st.flux.zero()
        st.face_sig.zero()
        phx.calculate_fluxes(p, b, f)
        synchronize()
        got = (st.flux.get(), st.face_sig.get())
        if first is None:
            first = got
        np.testing.assert_array_equal(got[0], first[0], err_msg=f"shape {(ti, ts)}")
        np.testing.assert_array_equal(got[1], first[1], err_msg=f"shape {(ti, ts)}")
    for d in (1, 2):
        phx.set_tile_shape(p, b, d, 0, 0)


def _stage(phx, p, b, prim, u_base, beta, dt):
    st = BlockState(p, b)
    st.prim.upload(prim)
    phx.prim_to_con(p, b, st.fields())
    synchronize()
    d_base = DeviceArray.from_host(u_base)
    d_out = DeviceArray(st.cons.shape)
    d_out.zero()
    d_it = DeviceArray(st.shape, np.int32)
    d_dt = DeviceArray((1,))
    curved = p.geom.kind != A.GEOM_MINKOWSKI
    phx.fluid_stage(p, b, st.fields(with_src=curved), beta, dt, d_base, d_out, st.prim,
                    iters=d_it, cfl=0.4, dt_out=d_dt)
    synchronize()
    return (d_out.get(), st.prim.get(), st.fail.get(), st.c2p_mu.get(), d_dt.get()[0],
            st.flux.get())


@pytest.mark.parametrize("geom", [A.GEOM_MINKOWSKI, A.GEOM_FMKS])
def test_fast_stage_matches_exact_stage(phx, geom):
    say(f"[gpu] fused stage, fast vs exact flux arithmetic, geom {geom}")
    rng = np.random.default_rng(40 + geom)
    b = geom_block(3, (20, 16, 12), geom)
    p = geom_params(geom, gamma=13.0 / 9.0, mhd=True, recon=A.RECON_WENO5Z,
                    riemann=A.RIEMANN_HLL)
    prim = random_block(rng, b, True, 13.0 / 9.0, geom)
    op, ob = oparams(p), oblock(b)
    u0, _ = O.prim_to_con_block(op, ob, prim)
    u_base = u0 * (1.0 + 0.01 * rng.uniform(-1, 1, u0.shape))
    ex = _stage(phx, p, b, prim, u_base, 0.5, 2e-3)
    p.flux_arith = A.FLUX_ARITH_FAST
    fa = _stage(phx, p, b, prim, u_base, 0.5, 2e-3)
    I = O.interior(ob)
    sl = (slice(None),) + I
    assert np.array_equal(ex[2][I], fa[2][I]), "fail flags"
    for m in range(8):
        _flux_close(fa[0][m][I], ex[0][m][I], f"U_out m={m}")
    # con2prim runs to its 1e-10 bracket on slightly different conserved states; zones it
    # fails in (the same in both) keep what the failed inversion left
    ok = ex[2][I] == 1.0
    assert np.max(np.abs(fa[3][I] - ex[3][I])[ok]) <= 4 * p.c2p_tol
    # (a zone on the internal-energy floor has P = 0 and Gamma_1 = 0/0 in both, as in the
    # reference)
    gp, ep = fa[1][sl][:, ok], ex[1][sl][:, ok]
    both_nan = np.isnan(gp) & np.isnan(ep)
    assert not np.any(np.isnan(gp) ^ np.isnan(ep)), "NaN primitives in one arithmetic only"
    scale = np.nanmax(np.abs(ep), axis=1)[:, None]
    assert np.max(np.where(both_nan, 0.0, np.abs(gp - ep)) / scale) <= 1e-8, "primitives"
    assert abs(fa[4] - ex[4]) <= 1e-10 * ex[4], "dt"


def test_fast_pack_matches_fast_single_block_stages(phx):
    say("[gpu] pack in the fast flux arithmetic vs single-block fast stages")
    from test_gpu_pack import MARCH_SIZES, Setup, _blocks, _equal, _params

    geom = A.GEOM_MINKOWSKI
    rng = np.random.default_rng(61)
    p = _params(geom)
    p.flux_arith = A.FLUX_ARITH_FAST
    blocks = _blocks(geom, MARCH_SIZES)
    S = Setup(phx, p, blocks, rng, geom)
    pk = Pack(phx, p, [(b, st, c1, it) for b, st, c1, it in
                       zip(blocks, S.states, S.cons1, S.iters)])
    dt = 1e-3
    d_dt = DeviceArray((1,))
    S.reset()
    pk.stage(1, dt=dt, cfl=0.4, dt_out=d_dt)
    synchronize()
    after = S.snapshot()
    dt_pack = d_dt.get()[0]
    S.reset()
    dts = []
    d_one = DeviceArray((1,))
    for b, st, c1, it in zip(blocks, S.states, S.cons1, S.iters):
        phx.fluid_stage(p, b, st.fields(), 1.0, dt, st.cons, c1, st.prim, iters=it, cfl=0.4,
                        dt_out=d_one)
        synchronize()
        dts.append(d_one.get()[0])
    _equal(S.snapshot(), after, "fast stage 1")
    assert dt_pack == min(dts)
    pk.close()
