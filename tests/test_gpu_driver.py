"""The one-block RK2 driver (include/phx_driver.h) against the oracle's task sequence
(tests/_oracle.py Driver): the 1-D shock tube to t = 0.4 bit for bit, a 2-D periodic MHD
block with constrained transport, and a spherical Kerr-Schild block to the solver
tolerance."""
import sys

import numpy as np
import pytest

import _oracle as O
from phoebus_amd import _abi as A
from phoebus_amd.api import Driver, Phx, block, fixup_params, fluid_params
from phoebus_amd.workloads import shock_tube_prims

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def phx():
    ctx = Phx(0)
    yield ctx
    ctx.close()


def _pair(gamma, mhd, recon, riemann, geom=A.GEOM_MINKOWSKI, a=0.0, tol=1e-8):
    p = fluid_params(gamma=gamma, mhd=mhd, recon=recon, riemann=riemann, geom=geom, a=a,
                     c2p_tol=tol, c2p_max_iter=100)
    orc_recon = {A.RECON_LINEAR: O.RECON_LINEAR, A.RECON_WENO5Z: O.RECON_WENO5,
                 A.RECON_MP5: O.RECON_MP5}[recon]
    op = O.default_params(gamma=gamma, mhd=mhd, recon=orc_recon, riemann=riemann,
                          geom=O.GEOM_SPHKS if geom == A.GEOM_SPHERICAL_KS else O.GEOM_MINKOWSKI,
                          a=a, tol=tol, max_iter=100)
    return p, op


def test_shock_tube_bit_exact(phx):
    say("[gpu] driver: shock tube 512 zones to t=0.4 vs oracle")
    nx = 512
    p, op = _pair(5 / 3, False, A.RECON_WENO5Z, A.RIEMANN_HLL)
    b = block(1, (nx,), ng=4, dx=(1.0 / nx, 1.0, 1.0))
    ob = O.make_block(1, (nx,), ng=4, dx=(1.0 / nx, 1.0, 1.0))
    prim0 = shock_tube_prims(nx)
    d = Driver(phx, p, b, cfl=0.4)
    d.set_prim(prim0)
    d.initialize()
    od = O.Driver(op, ob, (O.BC_OUTFLOW,) * 6, cfl=0.4)
    od.initialize(prim0)
    assert d.time()[1] == od.dt
    # a few steps first, so a divergence is caught where it starts
    for _ in range(5):
        d.step()
        od.step()
    np.testing.assert_array_equal(d.get(A.DRIVER_PRIM), od.prim)
    d.execute(0.4)
    od.execute(0.4)
    t, dt, n = d.time()
    say(f"    cycles {n} t {t}")
    assert (t, dt, n) == (od.time, od.dt, od.ncycle)
    np.testing.assert_array_equal(d.get(A.DRIVER_PRIM), od.prim)
    np.testing.assert_array_equal(d.get(A.DRIVER_CONS), od.cons[0])
    np.testing.assert_array_equal(d.get(A.DRIVER_FAIL), od.fail)
    d.close()


def _mhd_state(sh, rng):
    nk, nj, ni = sh
    j, i = np.meshgrid(np.arange(nj), np.arange(ni), indexing="ij")
    prim = np.zeros((10,) + sh)
    ph = 2 * np.pi * i / (ni - 8), 2 * np.pi * j / (nj - 8)
    prim[0] = 1.0 + 0.3 * np.sin(ph[0]) * np.cos(ph[1])
    prim[1] = 0.4 * np.sin(ph[1])
    prim[2] = 0.3 * np.cos(ph[0])
    prim[3] = 0.1
    P = 0.5 + 0.2 * np.cos(ph[0] + ph[1])
    prim[4] = P / (4 / 3 - 1)
    prim[5] = 0.3 + 0.1 * np.cos(ph[1])
    prim[6] = 0.2 * np.sin(ph[0])
    prim[7] = 0.05
    prim[8] = P
    prim[9] = 4 / 3
    return prim


@pytest.mark.parametrize("recon", [A.RECON_LINEAR, A.RECON_MP5])
def test_periodic_mhd_2d_bit_exact(phx, recon):
    say(f"[gpu] driver: 2-D periodic MHD recon {recon} vs oracle")
    p, op = _pair(4 / 3, True, recon, A.RIEMANN_HLL, tol=1e-10)
    nx = (24, 20)
    dx = (1.0 / 24, 1.0 / 20, 1.0)
    b = block(2, nx, ng=4, dx=dx)
    ob = O.make_block(2, nx, ng=4, dx=dx)
    prim0 = _mhd_state(O.block_shape(ob), None)
    kinds = (A.BC_PERIODIC,) * 4 + (A.BC_OUTFLOW,) * 2
    d = Driver(phx, p, b, bc=kinds, cfl=0.3)
    d.set_prim(prim0)
    d.initialize()
    od = O.Driver(op, ob, kinds, cfl=0.3)
    od.initialize(prim0)
    for _ in range(8):
        d.step()
        od.step()
    assert d.time()[2] == od.ncycle
    assert d.time()[0] == od.time
    np.testing.assert_array_equal(d.get(A.DRIVER_PRIM), od.prim)
    np.testing.assert_array_equal(d.get(A.DRIVER_CONS), od.cons[0])
    d.close()


def test_kerr_schild_block_to_solver_tolerance(phx):
    # sin/cos of the metric round differently on the device: compare to what the
    # c2p tolerance and a few steps of round-off growth allow
    say("[gpu] driver: spherical KS 1-D block vs oracle")
    p, op = _pair(4 / 3, False, A.RECON_WENO5Z, A.RIEMANN_LLF, geom=A.GEOM_SPHERICAL_KS,
                  a=0.5, tol=1e-10)
    nx = 64
    dx = (0.1, 1.0, 1.0)
    xmin = (4.0, np.pi / 2 - 0.5, 0.0)
    b = block(1, (nx,), ng=4, xmin=xmin, dx=dx)
    ob = O.make_block(1, (nx,), ng=4, xmin=xmin, dx=dx)
    sh = O.block_shape(ob)
    r = xmin[0] + (np.arange(sh[2]) - 4 + 0.5) * dx[0]
    prim0 = np.zeros((7,) + sh)
    prim0[0] = 1.0 / r ** 1.5
    prim0[1] = -0.2 * np.sqrt(2 / r)
    P = 0.1 * prim0[0] / r
    prim0[4] = P / (1 / 3)
    prim0[5] = P
    prim0[6] = 4 / 3
    d = Driver(phx, p, b, cfl=0.3)
    d.set_prim(prim0)
    d.initialize()
    od = O.Driver(op, ob, (O.BC_OUTFLOW,) * 6, cfl=0.3)
    od.initialize(prim0)
    for _ in range(10):
        d.step()
        od.step()
    assert d.time()[2] == od.ncycle
    assert abs(d.time()[0] - od.time) <= 1e-12 * od.time
    pg, po = d.get(A.DRIVER_PRIM), od.prim
    scale = np.max(np.abs(po), axis=(1, 2, 3), keepdims=True)
    assert np.max(np.abs(pg - po) / scale) < 1e-8
    np.testing.assert_array_equal(d.get(A.DRIVER_FAIL), od.fail)
    d.close()


def test_driver_rejects_bad_arguments(phx):
    p, _ = _pair(5 / 3, False, A.RECON_WENO5Z, A.RIEMANN_HLL)
    b = block(1, (32,), ng=4, dx=(1 / 32, 1.0, 1.0))
    with pytest.raises(A.PhxError, match="boundary"):
        Driver(phx, p, b, bc=(7,) * 6)
    with pytest.raises(A.PhxError, match="positive"):
        Driver(phx, p, b, cfl=0.0)
    with pytest.raises(A.PhxError):
        Driver(phx, p, block(1, (32,), ng=2, dx=(1 / 32, 1.0, 1.0)))  # WENO5 needs 4 ghosts


def test_driver_with_fixups_bit_exact(phx):
    """A relativistic blast into a near vacuum with the c2p fixup and floors on
    (fixup.cpp:285-539, fixup_c2p.cpp:44-356): the density floor holds the vacuum and the
    gamma ceiling (2) caps the blast's Lorentz factor; the device and oracle sequences
    agree bit for bit."""
    say("[gpu] driver: 1-D blast into vacuum with floors, ceilings and c2p fixup vs oracle")
    nx = 256
    gamma = 5 / 3
    p, op = _pair(gamma, False, A.RECON_WENO5Z, A.RIEMANN_HLL, tol=1e-10)
    floors = dict(floor_type=A.FLOOR_CONSTANT_RHO_SIE, rho0=1e-5, sie0=1e-2)
    p = fluid_params(gamma=gamma, mhd=False, recon=A.RECON_WENO5Z, riemann=A.RIEMANN_HLL,
                     c2p_tol=1e-10, floors=floors, ceilings=(2.0, 1e4))
    op.bounds.floor_type = O.FLOOR_CONSTANT
    op.bounds.r0, op.bounds.s0 = 1e-5, 1e-2
    op.bounds.gam_max, op.bounds.sie_max = 2.0, 1e4
    b = block(1, (nx,), ng=4, dx=(1.0 / nx, 1.0, 1.0))
    ob = O.make_block(1, (nx,), ng=4, dx=(1.0 / nx, 1.0, 1.0))
    sh = O.block_shape(ob)
    x = (np.arange(sh[2]) - 4 + 0.5) / nx
    prim0 = np.zeros((7,) + sh)
    prim0[0] = np.where(x < 0.4, 1.0, 1e-6)
    P = np.where(x < 0.4, 1000.0, 1e-8)
    prim0[1] = np.where(x < 0.4, 0.0, -0.5)
    prim0[4] = P / (gamma - 1)
    prim0[5] = P
    prim0[6] = gamma
    fx = fixup_params(c2p_fixup=True, floors=True)
    d = Driver(phx, p, b, cfl=0.4)
    d.set_fixup(fx)
    d.set_prim(prim0)
    d.initialize()
    od = O.Driver(op, ob, (O.BC_OUTFLOW,) * 6, cfl=0.4, fx=O.fixup(c2p_fixup=True, floors=True))
    od.initialize(prim0)
    np.testing.assert_array_equal(d.get(A.DRIVER_PRIM), od.prim)
    nfail = 0
    for _ in range(60):
        d.step()
        od.step()
        nfail += int(np.sum(od.fail == 0.0))
    say(f"    cycles {od.ncycle}, con2prim failures repaired along the way {nfail}")
    assert d.time()[:3] == (od.time, od.dt, od.ncycle)
    np.testing.assert_array_equal(d.get(A.DRIVER_PRIM), od.prim)
    np.testing.assert_array_equal(d.get(A.DRIVER_CONS), od.cons[0])
    np.testing.assert_array_equal(d.get(A.DRIVER_FAIL), od.fail)
    assert np.all(od.prim[0] >= 1e-5 * (1 - 1e-12))  # the floor holds
    d.close()
