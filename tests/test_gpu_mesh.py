"""The device mesh driver (include/phx_mesh.h) against the oracle's mesh driver
(tests/_oracle.py MeshDriver) on identical seeded blocks: ghost exchange, periodic wrap,
gr_outflow boundaries, FixFluxes, boundary conversion, fixups and the time step.
Minkowski meshes bit for bit (block edges exact: dx a power of two); an FMKS torus
mesh to the solver tolerance (device sin/cos/exp/pow differ from glibc by an ulp)."""
import sys

import numpy as np
import pytest

import _oracle as O
from phoebus_amd import _abi as A
from phoebus_amd.api import Mesh, Phx, fixup_params, fluid_params, mgr
from phoebus_amd.workloads import TORUS, torus_grid
from test_oracle_mesh import _wave_prims

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def phx():
    ctx = Phx(0)
    yield ctx
    ctx.close()


def _ofx(fx):
    return O.Fixup(fx.enable_c2p_fixup, fx.failure_strategy, fx.enable_floors,
                   fx.enable_mhd_floors, fx.bsqorho_max, fx.bsqou_max)


def _blocks_of(G, loc, bnx, ng=4):
    for g, lx in enumerate(loc):
        sl = tuple(slice(lx[d] * bnx[d], lx[d] * bnx[d] + bnx[d] + 2 * ng) for d in (2, 1, 0))
        yield g, np.ascontiguousarray(G[(slice(None),) + sl])


def _compare(m, om, exact, what=""):
    for g in range(m.nblocks):
        gp, op = m.get(g, A.DRIVER_PRIM), om.prim[g]
        gc, oc = m.get(g, A.DRIVER_CONS), om.cons[g][0]
        if exact:
            assert np.array_equal(gp, op, equal_nan=True), f"{what} block {g} prim"
            assert np.array_equal(gc, oc, equal_nan=True), f"{what} block {g} cons"
            assert np.array_equal(m.get(g, A.DRIVER_FAIL), om.fail[g]), f"{what} block {g} fail"
        else:
            for a, b, name in ((gp, op, "prim"), (gc, oc, "cons")):
                scale = np.max(np.abs(b), axis=(1, 2, 3), keepdims=True) + 1e-300
                err = np.max(np.abs(a - b) / scale)
                say(f"    block {g} {name}: largest difference {err:.2e} of the component scale")
                # measured: up to 1.3e-10 on the FMKS torus mesh after two steps
                assert err < 1e-9, f"{what} block {g} {name}: {err}"


CASES = {
    # all periodic, fixups on with floors that act
    "periodic": dict(bc=(A.BC_PERIODIC,) * 6, fx=True),
    # gr_outflow in x1 and x2, periodic x3 (the torus' boundary kinds), FixFluxes on
    "outflow": dict(bc=(A.BC_OUTFLOW,) * 4 + (A.BC_PERIODIC,) * 2, fx=True),
}


@pytest.mark.parametrize("case", list(CASES))
def test_minkowski_mesh_bit_exact(phx, case):
    cfg = CASES[case]
    say(f"[gpu] mesh 16^3 in 8 blocks, {case} boundaries, vs oracle mesh")
    N, bnx = (16, 16, 16), (8, 8, 8)
    floors = dict(floor_type=A.FLOOR_CONSTANT_RHO_SIE, rho0=0.85, sie0=0.5)
    p = fluid_params(gamma=4 / 3, mhd=True, recon=A.RECON_WENO5Z, riemann=A.RIEMANN_HLL,
                     c2p_tol=1e-10, floors=floors, ceilings=(1.5, 1e300))
    op = O.default_params(gamma=4 / 3, mhd=True, recon=O.RECON_WENO5, riemann=O.RIEMANN_HLL,
                          tol=1e-10)
    op.bounds.floor_type = O.FLOOR_CONSTANT
    op.bounds.r0, op.bounds.s0 = 0.85, 0.5
    op.bounds.gam_max = 1.5
    fx = fixup_params(c2p_fixup=True, mhd_floors=True, bsqorho_max=0.6, bsqou_max=2.0)
    m = Mesh(phx, p, N, bnx, (0, 0, 0), (1, 1, 1), cfg["bc"], fx=fx, cfl=0.3)
    om = O.MeshDriver(op, N, bnx, (0, 0, 0), (1, 1, 1), cfg["bc"], fx=_ofx(fx), cfl=0.3)
    G = _wave_prims(N, 4, (0, 0, 0), (1, 1, 1))
    for g, pr in _blocks_of(G, om.loc, bnx):
        m.set_prim(g, pr)
        om.set_prim(g, pr)
    m.initialize()
    om.initialize()
    assert m.time()[1] == om.dt
    _compare(m, om, True, "initial")
    for s in range(3):
        m.step()
        om.step()
        _compare(m, om, True, f"step {s + 1}")
        assert m.time() == (om.time, om.dt, om.ncycle)
    floored = sum(int(np.sum(np.isclose(om.prim[g][0], 0.85))) for g in range(m.nblocks))
    say(f"    cycles {om.ncycle}, cells at the density floor {floored}")
    m.close()


def test_fmks_torus_mesh_to_solver_tolerance(phx):
    say("[gpu] FMKS torus mesh 32x16x8 in 4 blocks vs oracle mesh")
    N, bnx = (32, 16, 8), (16, 8, 8)
    prim_g, xmin, dx = torus_grid(np.random.default_rng(3), nx=N)
    xmax = tuple(xmin[d] + N[d] * dx[d] for d in range(3))
    T = TORUS
    x0 = min(xmin[0], max(0.0, xmin[0] - 4 * dx[0]))
    floors = dict(floor_type=A.FLOOR_EXPX1_RHO_SIE, rho0=1e-8, sie0=1e-3, rho_exp=-2.0,
                  sie_exp=-1.0, rho_min_eos=1e-8, sie_min_eos=1e-8)
    mg = mgr(derefine_poles=False, h=T["hslope"], x0=x0)
    p = fluid_params(gamma=5 / 3, mhd=True, recon=A.RECON_WENO5Z, riemann=A.RIEMANN_HLL,
                     geom=A.GEOM_FMKS, a=T["a"], mgr_params=mg, c2p_tol=1e-10,
                     floors=floors, ceilings=(10.0, 1.0))
    op = O.default_params(gamma=5 / 3, mhd=True, recon=O.RECON_WENO5, riemann=O.RIEMANN_HLL,
                          geom=O.GEOM_FMKS, a=T["a"], tol=1e-10,
                          mgr_params=O.mgr(derefine_poles=False, h=T["hslope"], x0=x0))
    op.bounds.floor_type = O.FLOOR_EXPX1_RHO_SIE
    op.bounds.r0, op.bounds.s0, op.bounds.ralpha, op.bounds.salpha = 1e-8, 1e-3, -2.0, -1.0
    op.bounds.rho_min_eos, op.bounds.sie_min_eos = 1e-8, 1e-8
    op.bounds.gam_max, op.bounds.sie_max = 10.0, 1.0
    fx = fixup_params(c2p_fixup=True, mhd_floors=True)
    bc = (A.BC_OUTFLOW,) * 4 + (A.BC_PERIODIC,) * 2
    m = Mesh(phx, p, N, bnx, xmin, xmax, bc, fx=fx, inflow_check=True, cfl=0.4)
    om = O.MeshDriver(op, N, bnx, xmin, xmax, bc, fx=_ofx(fx), inflow_check=True, cfl=0.4)
    for g, pr in _blocks_of(prim_g, om.loc, bnx):
        m.set_prim(g, pr)
        om.set_prim(g, pr)
    m.initialize()
    om.initialize()
    assert abs(m.time()[1] - om.dt) <= 1e-10 * om.dt
    for s in range(2):
        m.step()
        om.step()
    assert m.time()[2] == om.ncycle
    _compare(m, om, False, "after 2 steps")
    fails = sum(int(np.sum(om.fail[g][4:-4, 4:-4, 4:-4] == 0)) for g in range(m.nblocks))
    say(f"    interior con2prim failures (repaired) in the last pass: {fails}")
    m.close()


def test_mesh_rejects_bad_descriptions(phx):
    p = fluid_params(mhd=True)
    with pytest.raises(A.PhxError, match="divide"):
        Mesh(phx, p, (16, 16, 16), (6, 8, 8), (0, 0, 0), (1, 1, 1), (A.BC_PERIODIC,) * 6)
    with pytest.raises(A.PhxError, match="pairs"):
        Mesh(phx, p, (16, 16, 16), (8, 8, 8), (0, 0, 0), (1, 1, 1),
             (A.BC_PERIODIC, A.BC_OUTFLOW) + (A.BC_PERIODIC,) * 4)
    with pytest.raises(A.PhxError, match="3-D MHD"):
        Mesh(phx, fluid_params(mhd=False), (16, 16, 16), (8, 8, 8), (0, 0, 0), (1, 1, 1),
             (A.BC_PERIODIC,) * 6)
