"""The 128^3 GRMHD stage of the benchmark at its full size, checked through properties that
do not depend on the size (the oracle comparisons run on small blocks in test_gpu_parity):
the flux divergence telescopes (the sum of U_{s+1} - U_s over the block is the net flux
through its boundary; checked on the update before the inversion, phx_flux_divergence_update
on the stage's own fluxes after phx_flux_ct, which the fused stage's U_{s+1} equals bit for
bit wherever the inversion did not re-derive it), the inverted primitives give U_{s+1} back through p2c to
the solver's tolerance, and the CFL time step is the minimum over cells recomputed from the
stage's own signal speeds, bit for bit."""
import sys

import numpy as np
import pytest

from phoebus_amd import _abi as A
from phoebus_amd.api import BlockState, Phx, block, fluid_params
from phoebus_amd.device import DeviceArray, synchronize
from phoebus_amd.workloads import grmhd_block_prims

pytestmark = pytest.mark.gpu

N, NG, GAMMA = 128, 4, 13.0 / 9.0


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def phx():
    ctx = Phx(0)
    yield ctx
    ctx.close()


def test_fullsize_stage_properties(phx):
    say("[gpu] 128^3 WENO5-Z + HLL stage: telescoping, p2c round trip, CFL minimum")
    prim = grmhd_block_prims(np.random.default_rng(20261202), N, NG, GAMMA)
    p = fluid_params(gamma=GAMMA, mhd=True, recon=A.RECON_WENO5Z, riemann=A.RIEMANN_HLL,
                     c2p_tol=1e-10, c2p_max_iter=100)
    dx = 1.0 / N
    b = block(3, (N, N, N), NG, dx=(dx,) * 3)
    st = BlockState(p, b)
    st.prim.upload(prim)
    phx.prim_to_con(p, b, st.fields())
    synchronize()
    u0 = st.cons.get()
    u_out = DeviceArray(st.cons.shape)
    prim_out = DeviceArray.from_host(prim)
    iters = DeviceArray(st.shape, np.int32)
    d_dt = DeviceArray((1,))
    dt, cfl = 1e-4, 0.4
    phx.fluid_stage(p, b, st.fields(), 1.0, dt, st.cons, u_out, prim_out, iters=iters, cfl=cfl,
                    dt_out=d_dt)
    synchronize()
    u1, flux = u_out.get(), st.flux.get()
    I = (slice(NG, NG + N),) * 3
    lo, hi = NG, NG + N  # first interior index, first index past the interior
    # the update alone on the stage's fluxes after FluxCT (the fused stage forms CT's B
    # fluxes from the corner EMFs and leaves the Riemann B fluxes in flux): U_{s+1} before
    # the inversion
    phx.flux_ct(p, b, st.fields())
    synchronize()
    flux = st.flux.get()
    d_us, d_ub = DeviceArray.from_host(u0), DeviceArray.from_host(u0)
    u_sep = DeviceArray(st.cons.shape)
    phx.flux_divergence_update(p, b, st.flux, None, d_us, d_ub, 1.0, dt, u_sep)
    synchronize()
    us = u_sep.get()
    # the inversion re-derives U where it used a bound; with the default bounds only the
    # internal-energy floor at 0 can act (u = 0 then)
    floored = prim_out.get()[4][I] == 0.0
    say(f"    zones on the energy floor: {int(floored.sum())}")
    for l in range(u0.shape[0]):
        same = ~floored if l < 5 else np.ones_like(floored)
        assert np.array_equal(u1[l][I][same], us[l][I][same]), l

    # telescoping: dU = -dt / vol * sum over faces of area * flux; the interior faces cancel
    a, vol = dx * dx, dx * dx * dx
    for l in range(u0.shape[0]):
        lhs = np.sum(us[l][I] - u0[l][I])
        F1, F2, F3 = flux[0, l], flux[1, l], flux[2, l]
        net = (a * (np.sum(F1[I[0], I[1], hi]) - np.sum(F1[I[0], I[1], lo]))
               + a * (np.sum(F2[I[0], hi, I[2]]) - np.sum(F2[I[0], lo, I[2]]))
               + a * (np.sum(F3[hi, I[1], I[2]]) - np.sum(F3[lo, I[1], I[2]])))
        rhs = -dt / vol * net
        scale = dt / vol * a * sum(np.sum(np.abs(F[k])) for F, k in
                                   ((F1, I), (F2, I), (F3, I)))
        assert abs(lhs - rhs) <= 1e-12 * scale + 1e-300, (l, lhs, rhs, scale)

    # p2c of the inverted primitives against U_{s+1} (floors and ceilings are off, so every
    # solved zone's U_{s+1} is exactly what it inverted)
    fail = st.fail.get()[I]
    ok = fail == 1.0
    assert ok.mean() > 0.999, ok.mean()
    st2 = BlockState(p, b)
    st2.prim.upload(prim_out.get())
    phx.prim_to_con(p, b, st2.fields())
    synchronize()
    u2 = st2.cons.get()
    for l, name in enumerate(("D", "S1", "S2", "S3", "tau")):
        ref, got = u1[l][I][ok], u2[l][I][ok]
        sc = np.abs(u1[0][I][ok]) if l else np.abs(ref)  # momenta and energy relative to D
        rel = np.abs(got - ref) / np.maximum(sc, np.abs(ref))
        say(f"    {name}: median {np.median(rel):.2e} max {rel.max():.2e}")
        # the solve stops once the root's bracket is 2e-10 wide, so U_{s+1} comes back to
        # ~1e-10 relative, more where W and the magnetization amplify it
        assert np.median(rel) < 1e-9 and rel.max() < 1e-5, name
    for l in range(5, 8):  # B: p2c returns the primitive field, which con2prim copied
        assert np.array_equal(u2[l][I][ok], u1[l][I][ok]), l

    # the CFL time step: cfl * min over interior cells of 1 / (sum_d smax(cell speed,
    # smax(the two face speeds)) / dx + 1e-50), the stage's own formula and operation order;
    # smax(a, b) = b if a < b else a (a NaN face speed of a zone with a failed state gives way
    # to the other operand as it does on the device)
    smax = lambda x, y: np.where(x < y, y, x)
    csig, fsig = st.cell_sig.get(), st.face_sig.get()
    ldt = np.zeros((N, N, N))
    for d in range(3):
        up = [I[0], I[1], I[2]]
        up[2 - d] = slice(NG + 1, NG + N + 1)  # the upper face of direction d
        s = smax(csig[d][I], smax(fsig[d][I], fsig[d][tuple(up)]))
        ldt = ldt + s / dx
    cell_dt = 1.0 / (ldt + 1.e-50)
    # zones whose speeds are NaN (P = 0 on the energy floor makes Gamma_1 = 0/0, as in the
    # reference) constrain nothing, as in the reference's serial minimum
    say(f"    zones with a NaN CFL time: {int(np.isnan(cell_dt).sum())}")
    dt_ref = cfl * np.nanmin(cell_dt)
    assert d_dt.get()[0] == dt_ref, (d_dt.get()[0], dt_ref)
    it = iters.get()[I]
    say(f"    mean con2prim iterations (cold) {it[ok].mean():.2f}, dt {dt_ref:.6e}")
