"""The uneven torus pack of the stage benchmark (512 FMKS blocks of 16^3 / 32^3 / 64^3 cut
from the magnetized Fishbone-Moncrief torus, a = 0.9375, torus-deck floors and ceilings):

- a few mixed 16^3 / 32^3 blocks of it run one pack RK2 step (two stages) against the
  oracle's task sequence per block, at the FMKS bars (fail flags exact; conserved state and
  primitives to 1e-8 of each component's scale; con2prim roots within 4 tol of the
  oracle's after stage 1);
- the full 512-block pack runs one stage through property checks that need no oracle:
  the conserved update of every block equals the flux divergence of the pack's own fluxes
  and sources (which telescopes to the net boundary flux), except in zones con2prim
  re-derived at a floor or ceiling, where p2c of the primitives gives the conserved state
  back; every interior zone inverts; dt = cfl x the minimum CFL time of every block.
"""
import sys

import numpy as np
import pytest

import _oracle as O
from phoebus_amd import _abi as A
from phoebus_amd.api import BlockState, Pack, Phx, block, fluid_params, mgr
from phoebus_amd.device import DeviceArray, synchronize
from phoebus_amd.workloads import TORUS, torus_grid, uneven_blocks

pytestmark = pytest.mark.gpu

NG = 4
SEED = 20261202 + 4  # the benchmark's torus pack seed


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def phx():
    ctx = Phx(0)
    yield ctx
    ctx.close()


@pytest.fixture(scope="module")
def torus():
    rng = np.random.default_rng(SEED)
    prim, xmin, dx = torus_grid(rng)
    return rng, prim, xmin, dx


def _params(xmin, dx):
    x0 = min(xmin[0], max(0.0, xmin[0] - NG * dx[0]))
    floors = dict(floor_type=A.FLOOR_EXPX1_RHO_SIE, rho0=1e-8, sie0=1e-3, rho_exp=-2.0,
                  sie_exp=-1.0, rho_min_eos=1e-8, sie_min_eos=1e-8)
    p = fluid_params(gamma=5.0 / 3.0, mhd=True, recon=A.RECON_WENO5Z, riemann=A.RIEMANN_HLL,
                     geom=A.GEOM_FMKS, a=TORUS["a"], c2p_tol=1e-10, c2p_max_iter=100,
                     floors=floors, ceilings=(10.0, 1.0),
                     mgr_params=mgr(derefine_poles=False, h=TORUS["hslope"], x0=x0))
    op = O.default_params(gamma=5.0 / 3.0, mhd=True, recon=O.RECON_WENO5,
                          riemann=O.RIEMANN_HLL, geom=O.GEOM_FMKS, a=TORUS["a"], tol=1e-10,
                          max_iter=100,
                          mgr_params=O.mgr(derefine_poles=False, h=TORUS["hslope"], x0=x0))
    op.bounds.floor_type = O.FLOOR_EXPX1_RHO_SIE
    op.bounds.r0, op.bounds.s0, op.bounds.ralpha, op.bounds.salpha = 1e-8, 1e-3, -2.0, -1.0
    op.bounds.rho_min_eos, op.bounds.sie_min_eos = 1e-8, 1e-8
    op.bounds.gam_max, op.bounds.sie_max = 10.0, 1.0
    return p, op


def _window(prim, n, o):
    return np.ascontiguousarray(prim[:, o[2]:o[2] + n + 2 * NG, o[1]:o[1] + n + 2 * NG,
                                     o[0]:o[0] + n + 2 * NG])


def _block(n, o, xmin, dx):
    return block(3, (n, n, n), NG, xmin=tuple(xmin[d] + o[d] * dx[d] for d in range(3)), dx=dx)


def _oblock(n, o, xmin, dx):
    return O.make_block(3, [n, n, n], NG, tuple(xmin[d] + o[d] * dx[d] for d in range(3)), dx)


def _close(a, b, tol, what):
    """max |a - b| per component relative to the component's largest |b|."""
    scale = np.max(np.abs(b).reshape(b.shape[0], -1), axis=1) + 1e-300
    err = np.max(np.abs(a - b).reshape(b.shape[0], -1), axis=1) / scale
    print(f"    {what}: largest difference {np.max(err):.2e} of the component scale", flush=True)
    assert np.all(err <= tol), f"{what}: {err}"


def test_torus_blocks_rk2_step_match_oracle(phx, torus):
    """Six blocks of the torus pack's mix (16^3 and 32^3 where the disk and its atmosphere
    meet, and inside the disk) in one pack: one RK2 step against the oracle per block."""
    _, prim_g, xmin, dx = torus
    p, op = _params(xmin, dx)
    # interior index origins: equatorial disk body, the inner edge near r_in, the disk's
    # surface against the floor atmosphere, one well above it
    picks = [(16, (150, 56, 10)), (16, (128, 60, 40)), (32, (140, 40, 64)),
             (16, (170, 30, 100)), (32, (110, 48, 0)), (16, (150, 10, 70))]
    say(f"[gpu] torus pack of {len(picks)} blocks vs oracle, one RK2 step")
    states, cons1, its, wins, blocks, oblocks = [], [], [], [], [], []
    for n, o in picks:
        b = _block(n, o, xmin, dx)
        st = BlockState(p, b)
        w = _window(prim_g, n, o)
        st.prim.upload(w)
        phx.prim_to_con(p, b, st.fields())
        synchronize()
        c0 = st.cons.get()
        states.append(st)
        cons1.append(DeviceArray.from_host(c0))
        its.append(DeviceArray(st.shape, np.int32))
        wins.append((w, c0))
        blocks.append(b)
        oblocks.append(_oblock(n, o, xmin, dx))
    # dt: cfl 0.4 over the initial cell signal speeds (EstimateTimestep's first call)
    dt = min(O.estimate_dt(op, ob, O.prim_to_con_block(op, ob, w)[1], None, 0.4)
             for ob, (w, _) in zip(oblocks, wins))
    pk = Pack(phx, p, list(zip(blocks, states, cons1, its)))
    pk.stage(1, dt=dt)
    synchronize()
    after1 = [(st.prim.get(), c1.get(), st.c2p_mu.get(), st.fail.get())
              for st, c1 in zip(states, cons1)]
    pk.stage(2, dt=dt)
    synchronize()
    after2 = [(st.prim.get(), st.cons.get(), st.fail.get()) for st in states]
    pk.close()
    for n_b, (ob, (w, c0)) in enumerate(zip(oblocks, wins)):
        I = O.interior(ob)
        lo = [s.start for s in I[::-1]]
        hi = [s.stop - 1 for s in I[::-1]]
        # stage 1: U1 = U0 + dt L(P0), P1 = c2p(U1) on the interior
        flux, _ = O.calculate_fluxes(op, ob, w)
        O.flux_ct(op, ob, flux)
        src = O.source_terms_block(op, ob, w)
        u1 = O.flux_divergence_update(op, ob, flux, src, c0.copy(), c0, 1.0, dt)
        p1, mu1 = w.copy(), np.zeros(O.block_shape(ob))
        r1 = O.con2prim_block(op, ob, u1, p1, mu1, lo=lo, hi=hi)
        gp1, gc1, gmu1, gf1 = after1[n_b]
        assert np.array_equal(gf1[I], r1["fail"][I]), f"block {n_b} stage 1 fail flags"
        sl = (slice(None),) + I
        # FMKS sources differ at the 1e-12 level of their scale (finite-difference metric
        # derivatives amplify ulp-level metric differences); U1 inherits more where the
        # update is all source (the atmosphere's theta momentum), and P1 carries the root's
        # tolerance (measured on these six blocks: U 2.0e-9, P 2.9e-9 of the scale)
        _close(gc1[sl], u1[sl], 1e-8, f"block {n_b} U1")
        _close(gp1[sl], p1[sl], 1e-8, f"block {n_b} P1")
        assert np.max(np.abs(gmu1[I] - mu1[I])) <= 4 * op.c2p_tol
        # stage 2 from the oracle's own stage-1 state: U_s = (U1 + U0) / 2,
        # U2 = U_s + dt / 2 L(P1)
        flux, _ = O.calculate_fluxes(op, ob, p1)
        O.flux_ct(op, ob, flux)
        src = O.source_terms_block(op, ob, p1)
        u2 = O.flux_divergence_update(op, ob, flux, src, u1.copy(), c0, 0.5, dt)
        p2, mu2 = p1.copy(), mu1.copy()
        r2 = O.con2prim_block(op, ob, u2, p2, mu2, lo=lo, hi=hi)
        gp2, gc2, gf2 = after2[n_b]
        assert np.array_equal(gf2[I], r2["fail"][I]), f"block {n_b} stage 2 fail flags"
        _close(gc2[sl], u2[sl], 1e-8, f"block {n_b} U2")
        _close(gp2[sl], p2[sl], 1e-8, f"block {n_b} P2")
        assert np.all(r2["fail"][I] == 1.0)


def test_full_torus_pack_stage_properties(phx, torus):
    """The 512-block pack of the benchmark, one stage at its CFL step, checked block by
    block without the oracle."""
    rng, prim_g, xmin, dx = torus
    p, _ = _params(xmin, dx)
    blocks = uneven_blocks(rng, 512)  # the benchmark's draw (same generator state)
    say(f"[gpu] full torus pack: {len(blocks)} blocks, "
        f"{sum(n ** 3 for n, _ in blocks)} interior zones, one stage")
    ents, wins = [], []
    for n, o in blocks:
        b = _block(n, o, xmin, dx)
        st = BlockState(p, b)
        w = _window(prim_g, n, o)
        st.prim.upload(w)
        phx.prim_to_con(p, b, st.fields())
        c1 = DeviceArray(st.cons.shape)
        ents.append((b, st, c1, DeviceArray(st.shape, np.int32)))
    synchronize()
    # dt: a first pass's CFL estimate, then the stage at that dt from the same state
    pk = Pack(phx, p, ents)
    d_dt = DeviceArray((1,))
    pk.stage(1, dt=1e-5, cfl=0.4, dt_out=d_dt)
    synchronize()
    dt = float(d_dt.get()[0])
    for b, st, c1, it in ents:
        st.c2p_mu.zero()
    for (n, o), (b, st, c1, it) in zip(blocks, ents):
        st.prim.upload(_window(prim_g, n, o))
    synchronize()
    pk.stage(1, dt=dt, cfl=0.4, dt_out=d_dt)
    synchronize()
    dt_out = float(d_dt.get()[0])
    say(f"    dt {dt:.6e}, dt_out {dt_out:.6e}")
    scratch = {}
    lmin = np.inf
    rederived = 0
    worst_rt = 0.0
    for (n, o), (b, st, c1, it) in zip(blocks, ents):
        sh = st.shape
        I = (slice(NG, NG + n),) * 3
        sl = (slice(None),) + I
        fail = st.fail.get()
        assert np.all(fail[I] == 1.0), f"block {o}: interior con2prim failures"
        # the conserved update against the divergence of the pack's own fluxes + sources
        if sh not in scratch:
            scratch[sh] = (DeviceArray(st.cons.shape), DeviceArray(st.cons.shape),
                           DeviceArray(st.cell_sig.shape))
        u_div, u_tmp, scratch_sig = scratch[sh]
        u_tmp.fill_from(st.cons)
        # the pack forms FluxCT's B fluxes from the corner EMFs and leaves the Riemann B
        # fluxes in flux: apply FluxCT to them first
        phx.flux_ct(p, b, st.fields())
        phx.flux_divergence_update(p, b, st.flux, st.src, u_tmp, st.cons, 1.0, dt, u_div)
        synchronize()
        U0, U1, Ud = st.cons.get(), c1.get(), u_div.get()
        F, S = st.flux.get(), st.src.get()
        # telescoping: sum over the interior of (U_div - U0) / dt - S equals minus the net
        # flux through the block's boundary faces (uniform logical coordinates)
        vol = dx[0] * dx[1] * dx[2]
        area = (dx[1] * dx[2], dx[0] * dx[2], dx[0] * dx[1])
        eps = np.finfo(np.float64).eps
        for m in range(8):
            Sm = S[m][I] if 1 <= m <= 4 else 0.0  # the update adds the S_i and tau sources
            lhs = np.sum((Ud[m][I] - U0[m][I]) / dt - Sm) * vol
            # rounding: the terms' magnitudes, and U_div's own rounding divided by dt
            net = 0.0
            mag = (np.sum(np.abs(Sm)) * vol + np.sum(np.abs(Ud[m][I] - U0[m][I])) * vol / dt
                   + 1e3 * eps * np.sum(np.abs(U0[m][I])) * vol / dt)
            for d in range(3):
                hi_face = [slice(NG, NG + n)] * 3
                lo_face = [slice(NG, NG + n)] * 3
                ax = 2 - d  # array axis of direction d (k, j, i)
                hi_face[ax] = slice(NG + n, NG + n + 1)
                lo_face[ax] = slice(NG, NG + 1)
                fh, fl = F[d, m][tuple(hi_face)], F[d, m][tuple(lo_face)]
                net += area[d] * (np.sum(fh) - np.sum(fl))
                mag += area[d] * (np.sum(np.abs(fh)) + np.sum(np.abs(fl)))
            assert abs(lhs + net) <= 1e-11 * mag + 1e-300, (o, m, lhs, net, mag)
        # U_{s+1} is the divergence update, except where con2prim re-derived it at a bound
        diff = np.any(U1[sl] != Ud[sl], axis=0)
        rederived += int(np.sum(diff))
        # p2c of the inverted primitives gives U_{s+1} back (B exactly)
        f = st.fields()
        f.cons = u_tmp.ptr
        f.cell_sig = scratch_sig.ptr
        lo, hi = (NG, NG, NG), (NG + n - 1,) * 3
        phx.prim_to_con(p, b, f, lo=lo, hi=hi)
        synchronize()
        Ur = u_tmp.get()
        scale = np.max(np.abs(U1[sl]).reshape(8, -1), axis=1)[:, None, None, None] + 1e-300
        rt = np.max(np.abs(Ur[sl] - U1[sl]) / scale)
        worst_rt = max(worst_rt, rt)
        assert rt <= 1e-9, (o, rt)
        # the block's CFL times (fluid.cpp:913-928): max of the cell and both face speeds
        cs, fs = st.cell_sig.get(), st.face_sig.get()
        ldt = np.zeros((n, n, n))
        for d in range(3):
            ax = 2 - d
            up = [slice(NG, NG + n)] * 3
            up[ax] = slice(NG + 1, NG + n + 1)
            s_ = np.maximum(cs[d][I], np.maximum(fs[d][I], fs[d][tuple(up)]))
            ldt = ldt + s_ / dx[d]
        t = 1.0 / (ldt + 1e-50)
        t = t[~np.isnan(t)]
        lmin = min(lmin, float(np.min(t)))
    say(f"    zones re-derived at a bound {rederived}, worst p2c round trip {worst_rt:.2e}")
    assert dt_out == 0.4 * lmin
    pk.close()
