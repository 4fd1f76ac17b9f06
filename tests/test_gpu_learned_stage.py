"""The learned con2prim bracket inside the stage (phx_pack_set_learned, phx_mesh_set_learned):
the update inverts the cells that have a previous root (as the plain stage does) and a
network + bracketed-solve pass inverts the rest.  The solver still runs to the same
tolerance, so primitives agree with the plain path to what that tolerance allows and the
fail flags are identical, whether the bracket holds the root (a bracket scale of 1e6 on
Minkowski states) or misses it (a scale of 1e-9: the plain [0, 1] solve runs); cells with a
previous root get the plain path's results bit for bit."""
import sys

import numpy as np
import pytest

from phoebus_amd import _abi as A
from phoebus_amd.api import Mesh, Mlp, Pack, Phx, fixup_params, fluid_params, mgr
from phoebus_amd.device import DeviceArray, synchronize
from phoebus_amd.workloads import torus_grid
from test_gpu_pack import SIZES, Setup, _blocks, _params

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def phx():
    ctx = Phx(0)
    yield ctx
    ctx.close()


def _net(phx, width_scale, width_min):
    net = Mlp(phx, seed=11)
    net.set_normalization([0, 0, 0, 0], [1, 1, 1, 1], width_scale, width_min)
    return net


def _close(a, b, rtol):
    den = np.maximum(np.abs(b), np.max(np.abs(b), axis=tuple(range(1, b.ndim)), keepdims=True) * 1e-6)
    ok = (np.abs(a - b) <= rtol * den) | (np.isnan(a) & np.isnan(b))
    return ok.all(), float(np.nanmax(np.abs(a - b) / np.maximum(den, 1e-300)))


def test_pack_learned_nan_cell_keeps_its_cfl_time(phx):
    """A cell with no previous root whose conserved state is NaN is left by both inversions
    (fail flag 0) with its signal speeds, and its CFL time still counts: given the largest
    speeds of the pack it sets dt_out, the same in the plain and the learned stage."""
    say("[gpu] pack stage with a NaN cell: learned vs plain time step")
    geom = A.GEOM_MINKOWSKI
    rng = np.random.default_rng(43)
    p = _params(geom)
    blocks = _blocks(geom)
    S = Setup(phx, p, blocks, rng, geom)
    pk = Pack(phx, p, [(b, st, c1, it) for b, st, c1, it in zip(blocks, S.states, S.cons1, S.iters)])
    net = _net(phx, 1e6, 1e-9)
    st0 = S.states[0]  # an 8^3 block with 4 ghosts: (6, 6, 6) is interior

    def stage(learned):
        S.reset()
        cons = st0.cons.get()
        cons[0, 6, 6, 6] = np.nan  # U^n, hence U_1, of the cell
        st0.cons.upload(cons)
        sig = st0.cell_sig.get()
        sig[:, 6, 6, 6] = 50.0  # its speeds before the stage
        st0.cell_sig.upload(sig)
        pk.set_learned(net if learned else None)
        d_dt = DeviceArray((1,))
        pk.stage(1, dt=1e-3, cfl=0.4, dt_out=d_dt)
        synchronize()
        return float(d_dt.get()[0]), float(st0.fail.get()[6, 6, 6])

    dt_plain, fail_plain = stage(False)
    dt_learned, fail_learned = stage(True)
    assert fail_plain == 0.0 and fail_learned == 0.0
    assert dt_learned == dt_plain, (dt_learned, dt_plain)
    want = 0.4 / (50.0 / 0.1 + 50.0 / 0.2 + 50.0 / 0.3)  # dx = (0.1, 0.2, 0.3)
    assert abs(dt_plain - want) <= 1e-14 * want, (dt_plain, want)
    pk.close()
    net.close()


@pytest.mark.parametrize("width", ["wide", "narrow"])
def test_pack_learned_stage_matches_plain(phx, width):
    say(f"[gpu] pack stage, learned bracket ({width}) vs plain")
    geom = A.GEOM_MINKOWSKI
    rng = np.random.default_rng(41)
    p = _params(geom)
    blocks = _blocks(geom)
    S = Setup(phx, p, blocks, rng, geom)
    pk = Pack(phx, p, [(b, st, c1, it) for b, st, c1, it in zip(blocks, S.states, S.cons1, S.iters)])
    S.reset()
    d_dtp = DeviceArray((1,))
    pk.stage(1, dt=1e-3, cfl=0.4, dt_out=d_dtp)
    synchronize()
    plain = S.snapshot()
    net = _net(phx, 1e6, 1e-9) if width == "wide" else _net(phx, 1e-9, 1e-12)
    pk.set_learned(net)
    pk.learned_stats(reset=True)
    S.reset()
    d_dtl = DeviceArray((1,))
    pk.stage(1, dt=1e-3, cfl=0.4, dt_out=d_dtl)
    synchronize()
    learned = S.snapshot()
    misses, solves = pk.learned_stats()
    interior = sum(int(np.prod(s)) for s in SIZES)
    say(f"    bracket misses {misses} of {solves}")
    assert solves == interior
    if width == "wide":
        assert misses == 0
    else:
        assert misses > 0.9 * solves
    for n, (a, b) in enumerate(zip(learned, plain)):
        prim_l, prim_p = a[0][:, 4:-4, 4:-4, 4:-4], b[0][:, 4:-4, 4:-4, 4:-4]
        ok, err = _close(prim_l, prim_p, 1e-6)
        assert ok, f"block {n}: primitives differ by {err}"
        assert np.array_equal(a[4], b[4]), f"block {n}: fail flags"
    # the time step from the signal speeds of the learned pass's cells: those of roots
    # within the tolerance of the plain ones
    dtl, dtp = float(d_dtl.get()[0]), float(d_dtp.get()[0])
    assert abs(dtl - dtp) <= 1e-6 * dtp, (dtl, dtp)
    # stage 2 starts from stage 1's roots: every interior cell has one, the update inverts
    # it as the plain stage does and the network brackets nothing
    d_dt = DeviceArray((1,))
    S.reset()
    pk.set_learned(None)
    pk.stage(1, dt=1e-3)
    pk.stage(2, dt=1e-3, cfl=0.4, dt_out=d_dt)
    synchronize()
    plain2, dt_plain = S.snapshot(), float(d_dt.get()[0])
    S.reset()
    pk.set_learned(None)
    pk.stage(1, dt=1e-3)
    pk.set_learned(net)
    pk.learned_stats(reset=True)
    d_dt.zero()
    pk.stage(2, dt=1e-3, cfl=0.4, dt_out=d_dt)
    synchronize()
    misses, solves = pk.learned_stats()
    assert solves == 0 and misses == 0
    assert float(d_dt.get()[0]) == dt_plain
    for n, (a, b) in enumerate(zip(S.snapshot(), plain2)):
        for q, name in enumerate(("prim", "cons", "cons1", "mu", "fail", "iters")):
            assert np.array_equal(a[q], b[q], equal_nan=True), f"block {n}: {name}"
    pk.close()
    net.close()


def test_mesh_learned_matches_plain(phx):
    say("[gpu] FMKS torus mesh 32x16x8: two steps plain vs learned (untrained network)")
    N, bnx = (32, 16, 8), (16, 8, 8)
    prim_g, xmin, dx = torus_grid(np.random.default_rng(3), nx=N)
    xmax = tuple(xmin[d] + N[d] * dx[d] for d in range(3))
    x0 = min(xmin[0], max(0.0, xmin[0] - 4 * dx[0]))
    floors = dict(floor_type=A.FLOOR_EXPX1_RHO_SIE, rho0=1e-8, sie0=1e-3, rho_exp=-2.0,
                  sie_exp=-1.0, rho_min_eos=1e-8, sie_min_eos=1e-8)
    p = fluid_params(gamma=5 / 3, mhd=True, recon=A.RECON_WENO5Z, riemann=A.RIEMANN_HLL,
                     geom=A.GEOM_FMKS, a=0.9375, mgr_params=mgr(derefine_poles=False, x0=x0),
                     c2p_tol=1e-10, floors=floors, ceilings=(10.0, 1.0))
    bc = (A.BC_OUTFLOW,) * 4 + (A.BC_PERIODIC,) * 2
    out = []
    net = _net(phx, 1e6, 1e-9)
    for learned in (False, True):
        m = Mesh(phx, p, N, bnx, xmin, xmax, bc, fx=fixup_params(c2p_fixup=True, mhd_floors=True),
                 inflow_check=True, cfl=0.4)
        for g, b in enumerate(m.blocks):
            lx = [int(round((b.xmin[d] - xmin[d]) / dx[d])) for d in range(3)]
            m.set_prim(g, np.ascontiguousarray(prim_g[:, lx[2]:lx[2] + 16, lx[1]:lx[1] + 16,
                                                      lx[0]:lx[0] + 24]))
        if learned:  # the initial inversion (no previous roots) uses the network
            m.set_learned(net)
        m.initialize()
        for _ in range(2):
            m.step()
        out.append([m.get(g, A.DRIVER_PRIM) for g in range(m.nblocks)])
        if learned:
            misses, solves = m.learned_stats()
            say(f"    learned solves {solves}, misses {misses}")
            # an untrained network: brackets that hold the root and ones that miss it
            assert solves > 0 and 0 <= misses <= solves
        m.close()
    for g, (a, b) in enumerate(zip(out[1], out[0])):
        ok, err = _close(a[:, 4:-4, 4:-4, 4:-4], b[:, 4:-4, 4:-4, 4:-4], 1e-6)
        assert ok, f"block {g}: {err}"
    net.close()
