"""Per-component field pointers (phx_block_vars): every component of prim, cons, flux[d] and
src in an allocation of its own, the way parthenon allocates its variables, passed as
kernel-argument tables with no gather or copy.  Each per-block entry point and the fused
stage must give the contiguous-field results bit for bit."""
import ctypes as C
import sys

import numpy as np
import pytest

import _oracle as O
from phoebus_amd import _abi as A
from phoebus_amd.api import BlockState, Phx, fixup_params
from phoebus_amd.device import DeviceArray, synchronize
from test_gpu_parity import geom_block, geom_params, oblock, random_block

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def phx():
    ctx = Phx(0)
    yield ctx
    ctx.close()


class Split:
    """Separately allocated components of prim, cons, flux[d] and src, plus the single
    fields (temp, mu, fail, speeds, emf) of a block."""

    def __init__(self, st, ndim, np_, nc):
        sh = st.shape
        self.np, self.nc, self.ndim = np_, nc, ndim
        self.prim = [DeviceArray(sh) for _ in range(np_)]
        self.cons = [DeviceArray(sh) for _ in range(nc)]
        self.flux = [[DeviceArray(sh) for _ in range(nc)] for _ in range(ndim)]
        self.src = [DeviceArray(sh) for _ in range(nc)]
        for a in self.prim + self.cons + self.src + sum(self.flux, []):
            a.zero()
        self.single = BlockState(st.p, st.b)  # temp, mu, fail, speeds, emf
        self.vars = self.table(self.prim, self.cons, self.flux, self.src)

    @staticmethod
    def table(prim=(), cons=(), flux=(), src=()):
        t = A.BlockVars()
        for v, a in enumerate(prim):
            t.prim[v] = a.ptr
        for m, a in enumerate(cons):
            t.cons[m] = a.ptr
        for d, row in enumerate(flux):
            for m, a in enumerate(row):
                t.flux[d][m] = a.ptr
        for m, a in enumerate(src):
            t.src[m] = a.ptr
        return t

    def fields(self):
        s = self.single
        return Phx.fields(temp=s.temp, c2p_mu=s.c2p_mu, fail=s.fail, cell_sig=s.cell_sig,
                          face_sig=s.face_sig, emf=s.emf, vars=self.vars)

    def upload(self, name, arr):
        for v, a in enumerate(getattr(self, name)):
            a.upload(np.ascontiguousarray(arr[v]))

    def get(self, name):
        if name == "flux":
            return np.stack([np.stack([a.get() for a in row]) for row in self.flux])
        return np.stack([a.get() for a in getattr(self, name)])


def _same(a, b, what):
    assert np.array_equal(a, b, equal_nan=True), what


@pytest.mark.parametrize("geom", [A.GEOM_MINKOWSKI, A.GEOM_FMKS])
def test_block_vars_match_contiguous_fields(phx, geom):
    say(f"[gpu] per-component field tables vs contiguous fields, geom {geom}")
    rng = np.random.default_rng(70 + geom)
    b = geom_block(3, (20, 16, 12), geom)
    floors = dict(floor_type=A.FLOOR_CONSTANT_RHO_SIE, rho0=0.2, sie0=0.05)
    p = geom_params(geom, gamma=13.0 / 9.0, mhd=True, recon=A.RECON_WENO5Z,
                    riemann=A.RIEMANN_HLL, floors=floors, ceilings=(4.0, 1e300))
    prim = random_block(rng, b, True, 13.0 / 9.0, geom)
    cg = BlockState(p, b)                 # contiguous
    sp = Split(cg, 3, 10, 8)              # separately allocated
    cg.prim.upload(prim)
    sp.upload("prim", prim)
    fc, fs = cg.fields(with_src=True), sp.fields()
    S = sp.single
    # PrimitiveToConserved
    phx.prim_to_con(p, b, fc)
    phx.prim_to_con(p, b, fs)
    synchronize()
    _same(sp.get("cons"), cg.cons.get(), "p2c cons")
    _same(S.cell_sig.get(), cg.cell_sig.get(), "p2c cell speeds")
    # CalculateFluxes, FluxCT, sources
    phx.calculate_fluxes(p, b, fc)
    phx.calculate_fluxes(p, b, fs)
    phx.flux_ct(p, b, fc)
    phx.flux_ct(p, b, fs)
    phx.fluid_source_terms(p, b, fc)
    phx.fluid_source_terms(p, b, fs)
    synchronize()
    _same(sp.get("flux"), cg.flux.get(), "fluxes after CT")
    _same(S.face_sig.get(), cg.face_sig.get(), "face speeds")
    _same(S.emf.get(), cg.emf.get(), "EMFs")
    I = O.interior(oblock(b))
    _same(sp.get("src")[1:5][(slice(None),) + I], cg.src.get()[1:5][(slice(None),) + I],
          "sources")
    # con2prim of a perturbed state over the entire block, then the fixups and floors
    u = cg.cons.get() * (1.0 + 0.02 * rng.uniform(-1, 1, cg.cons.shape))
    cg.cons.upload(u)
    sp.upload("cons", u)
    it_c, it_s = DeviceArray(cg.shape, np.int32), DeviceArray(cg.shape, np.int32)
    phx.con2prim(p, b, fc, iters=it_c)
    phx.con2prim(p, b, fs, iters=it_s)
    fx = fixup_params(c2p_fixup=True, mhd_floors=True, bsqorho_max=5.0, bsqou_max=50.0)
    phx.c2p_fixup(p, fx, b, fc)
    phx.c2p_fixup(p, fx, b, fs)
    phx.apply_floors(p, fx, b, fc)
    phx.apply_floors(p, fx, b, fs)
    synchronize()
    _same(it_s.get(), it_c.get(), "iterations")
    _same(S.fail.get(), cg.fail.get(), "fail flags")
    _same(sp.get("prim"), cg.prim.get(), "primitives after c2p + fixups")
    _same(sp.get("cons"), cg.cons.get(), "conserved after c2p + fixups")
    _same(S.c2p_mu.get(), cg.c2p_mu.get(), "mu")
    _same(S.temp.get(), cg.temp.get(), "temperature")
    d_c, d_s = DeviceArray((1,)), DeviceArray((1,))
    phx.estimate_dt(p, b, fc, True, 0.4, d_c)
    phx.estimate_dt(p, b, fs, True, 0.4, d_s)
    synchronize()
    assert d_c.get()[0] == d_s.get()[0]


@pytest.mark.parametrize("geom", [A.GEOM_MINKOWSKI, A.GEOM_FMKS])
def test_fused_stage_with_block_vars(phx, geom):
    """phx_fluid_stage reading U_s, prims and fluxes through a table and writing U^n,
    U_{s+1} and the new primitives through two more (the stage-2 container pattern)."""
    say(f"[gpu] fused stage with per-component tables, geom {geom}")
    rng = np.random.default_rng(80 + geom)
    b = geom_block(3, (24, 16, 12), geom)
    p = geom_params(geom, gamma=13.0 / 9.0, mhd=True, recon=A.RECON_WENO5Z,
                    riemann=A.RIEMANN_HLL)
    prim = random_block(rng, b, True, 13.0 / 9.0, geom)
    cg = BlockState(p, b)
    sp = Split(cg, 3, 10, 8)
    cg.prim.upload(prim)
    sp.upload("prim", prim)
    phx.prim_to_con(p, b, cg.fields())
    synchronize()
    u0 = cg.cons.get()
    u_base = u0 * (1.0 + 0.01 * rng.uniform(-1, 1, u0.shape))
    sp.upload("cons", u0)
    # contiguous: U^n, U_{s+1}, prim_out arrays
    d_base, d_out, d_pout = (DeviceArray.from_host(u_base), DeviceArray(cg.cons.shape),
                             DeviceArray(cg.prim.shape))
    d_out.zero()
    d_pout.zero()
    # split: the same three as per-component allocations
    base_c = [DeviceArray.from_host(np.ascontiguousarray(u_base[m])) for m in range(8)]
    out_c = [DeviceArray(cg.shape) for _ in range(8)]
    pout_c = [DeviceArray(cg.shape) for _ in range(10)]
    for a in out_c + pout_c:
        a.zero()
    base_t = Split.table(cons=base_c)
    out_t = Split.table(prim=pout_c, cons=out_c)
    it_c, it_s = DeviceArray(cg.shape, np.int32), DeviceArray(cg.shape, np.int32)
    it_c.zero()  # the stage writes the interior's counts
    it_s.zero()
    dt_c, dt_s = DeviceArray((1,)), DeviceArray((1,))
    phx.fluid_stage(p, b, cg.fields(with_src=True), 0.5, 1e-3, d_base, d_out, d_pout,
                    iters=it_c, cfl=0.4, dt_out=dt_c)
    phx.fluid_stage(p, b, sp.fields(), 0.5, 1e-3, None, None, None, iters=it_s, cfl=0.4,
                    dt_out=dt_s, base_vars=base_t, out_vars=out_t)
    synchronize()
    _same(np.stack([a.get() for a in out_c]), d_out.get(), "U_{s+1}")
    _same(np.stack([a.get() for a in pout_c]), d_pout.get(), "primitives out")
    _same(sp.get("cons"), cg.cons.get(), "U_s averaged in place")
    _same(sp.get("flux"), cg.flux.get(), "fluxes")
    _same(it_s.get(), it_c.get(), "iterations")
    _same(sp.single.fail.get(), cg.fail.get(), "fail flags")
    assert dt_c.get()[0] == dt_s.get()[0]


def test_block_vars_missing_component_refused(phx):
    """A field given only as a table must name every component the call reads."""
    p = geom_params(A.GEOM_MINKOWSKI, gamma=13.0 / 9.0, mhd=True)
    b = geom_block(3, (8, 8, 8), A.GEOM_MINKOWSKI)
    cg = BlockState(p, b)
    sp = Split(cg, 3, 10, 8)
    sp.vars.prim[7] = None  # B3 missing, and no contiguous prim
    with pytest.raises(A.PhxError, match="null field"):
        phx.calculate_fluxes(p, b, sp.fields())
