"""The oracle's mesh driver (tests/_oracle.py MeshDriver) on the CPU: a periodic Minkowski
mesh conserves its totals to round-off (the exchange closes the flux telescoping across
block faces), and the same problem cut into 1 or 8 blocks agrees to round-off (ghost
exchange, periodic wrap, block coordinates) -- bit for bit here, with dx a power of two
so that block edges and cell centres are exact."""
import numpy as np

import _oracle as O


def _wave_prims(N, ng, xmin, xmax):
    """A smooth MHD wave on the mesh with ghosts: [10][Nz+2ng][Ny+2ng][Nx+2ng]."""
    ax = [xmin[d] + (np.arange(N[d] + 2 * ng) - ng + 0.5) * (xmax[d] - xmin[d]) / N[d]
          for d in range(3)]
    z, y, x = np.meshgrid(ax[2], ax[1], ax[0], indexing="ij")
    ph = 2 * np.pi * (x + y + z)
    prim = np.zeros((10,) + z.shape)
    prim[0] = 1.0 + 0.2 * np.sin(ph)
    prim[1] = 0.3 * np.cos(ph)
    prim[2] = 0.2 * np.sin(2 * np.pi * x)
    prim[3] = 0.1
    P = 1.0 + 0.1 * np.cos(ph)
    prim[4] = P / (4 / 3 - 1)
    prim[5], prim[6], prim[7] = 0.5, 0.3 * np.cos(2 * np.pi * z), 0.2
    prim[8] = P
    prim[9] = 4 / 3
    return prim


def _mesh(nb_cells, N=(16, 16, 16), bc=None, fx=None):
    p = O.default_params(gamma=4 / 3, mhd=True, recon=O.RECON_WENO5, riemann=O.RIEMANN_HLL)
    bc = bc if bc is not None else (O.BC_PERIODIC,) * 6
    m = O.MeshDriver(p, N, nb_cells, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), bc, fx=fx, cfl=0.3)
    G = _wave_prims(N, 4, (0, 0, 0), (1, 1, 1))
    for g, lx in enumerate(m.loc):
        sl = tuple(slice(lx[d] * nb_cells[d], lx[d] * nb_cells[d] + nb_cells[d] + 8)
                   for d in (2, 1, 0))
        m.set_prim(g, G[(slice(None),) + sl])
    m.initialize()
    return m


def _interior(m, g, arr):
    return arr[..., 4:-4, 4:-4, 4:-4]


def _assemble(m, what):
    """The interior of a field over the whole mesh."""
    nv = what[0].shape[0]
    out = np.zeros((nv, m.N[2], m.N[1], m.N[0]))
    for g, lx in enumerate(m.loc):
        sl = tuple(slice(lx[d] * m.bnx[d], (lx[d] + 1) * m.bnx[d]) for d in (2, 1, 0))
        out[(slice(None),) + sl] = _interior(m, g, what[g])
    return out


def test_periodic_mesh_conserves_totals():
    m = _mesh((8, 8, 8))
    tot0 = _assemble(m, [c[0] for c in m.cons]).sum(axis=(1, 2, 3))
    for _ in range(2):
        m.step()
    tot = _assemble(m, [c[0] for c in m.cons]).sum(axis=(1, 2, 3))
    assert np.all(np.abs(tot - tot0) <= 1e-12 * np.abs(tot0).max()), (tot - tot0)


def test_decomposition_invariance():
    one = _mesh((16, 16, 16))
    eight = _mesh((8, 8, 8))
    for _ in range(2):
        one.step()
        eight.step()
    assert one.ncycle == eight.ncycle
    assert abs(one.time - eight.time) <= 1e-14 * one.time
    a, b = _assemble(one, one.prim), _assemble(eight, eight.prim)
    assert np.array_equal(a, b)
    assert np.array_equal(_assemble(one, [c[0] for c in one.cons]),
                          _assemble(eight, [c[0] for c in eight.cons]))
