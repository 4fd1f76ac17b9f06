"""Block packs (include/phx_pack.h): one launch per kernel over blocks of different sizes,
bit-identical per block to the single-block stage, and a captured graph replayed against
the eager result."""
import ctypes as C
import sys

import numpy as np
import pytest

import _physics as PH
from phoebus_amd import _abi as A
from phoebus_amd.api import BlockState, Pack, Phx, block, fixup_params, fluid_params, mgr
from phoebus_amd.device import DeviceArray, synchronize

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def phx():
    ctx = Phx(0)
    yield ctx
    ctx.close()


SIZES = [(8, 8, 8), (12, 10, 8), (16, 8, 12), (8, 12, 10), (10, 10, 10)]


def _blocks(geom, sizes=SIZES):
    out = []
    for n, nx in enumerate(sizes):
        if geom == A.GEOM_FMKS:  # X = (log r, x2, phi), blocks side by side in x2
            b = block(3, nx, 4, xmin=(1.5, 0.2 + 0.12 * n, 0.0), dx=(0.04, 0.01, 0.05))
        else:
            b = block(3, nx, 4, xmin=(0.3 * n, 0.0, 0.0), dx=(0.1, 0.2, 0.3))
        out.append(b)
    return out


def _state(rng, b, geom):
    sh = (b.nx[2] + 8, b.nx[1] + 8, b.nx[0] + 8)
    pr = PH.random_prims(rng, int(np.prod(sh)), True, 13.0 / 9.0, wmax=3.0)
    if geom != A.GEOM_MINKOWSKI:
        pr[1:4] *= 0.2
        pr[5:8] *= 0.2
    return np.ascontiguousarray(pr.reshape((-1,) + sh))


def _params(geom):
    return fluid_params(gamma=13.0 / 9.0, mhd=True, recon=A.RECON_WENO5Z,
                        riemann=A.RIEMANN_HLL, geom=geom, a=0.9375,
                        mgr_params=mgr(x0=0.2), c2p_tol=1e-10)


class Setup:
    """Pack fields, their host initial values, and a reset to them."""

    def __init__(self, phx, p, blocks, rng, geom):
        self.p, self.blocks = p, blocks
        self.states, self.cons1, self.iters, self.init = [], [], [], []
        for b in blocks:
            st = BlockState(p, b)
            prim = _state(rng, b, geom)
            st.prim.upload(prim)
            phx.prim_to_con(p, b, st.fields())
            synchronize()
            cons = st.cons.get()
            self.states.append(st)
            self.cons1.append(DeviceArray.from_host(cons))
            self.iters.append(DeviceArray(st.shape, np.int32))
            self.init.append((prim, cons))

    def reset(self):
        for st, c1, (prim, cons) in zip(self.states, self.cons1, self.init):
            st.prim.upload(prim)
            st.cons.upload(cons)
            c1.upload(cons)
            st.c2p_mu.zero()
            st.temp.zero()
        synchronize()

    def snapshot(self):
        return [(st.prim.get(), st.cons.get(), c1.get(), st.c2p_mu.get(), st.fail.get(),
                 it.get()) for st, c1, it in zip(self.states, self.cons1, self.iters)]


def _equal(a, b, what):
    for n, (x, y) in enumerate(zip(a, b)):
        for k, (u, v) in enumerate(zip(x, y)):
            # (zones whose inversion fails keep NaN in both paths)
            assert np.array_equal(u, v, equal_nan=True), f"{what}: block {n} field {k}"


# a 64^3 block makes the default y/z sweeps the marching waves (>= 1024 workgroups), for
# the pack (every block) and for that block's single-block stage
MARCH_SIZES = [(64, 64, 64), (8, 8, 8), (12, 10, 8)]


@pytest.mark.parametrize("geom,sizes", [(A.GEOM_MINKOWSKI, SIZES), (A.GEOM_FMKS, SIZES),
                                        (A.GEOM_MINKOWSKI, MARCH_SIZES),
                                        (A.GEOM_FMKS, MARCH_SIZES)],
                         ids=["mink", "fmks", "mink-march", "fmks-march"])
def test_pack_matches_single_block_stages(phx, geom, sizes):
    say(f"[gpu] pack of {len(sizes)} blocks {sizes[0]}... vs single-block stages, geom {geom}")
    rng = np.random.default_rng(31 + geom)
    p = _params(geom)
    blocks = _blocks(geom, sizes)
    S = Setup(phx, p, blocks, rng, geom)
    pk = Pack(phx, p, [(b, st, c1, it) for b, st, c1, it in
                       zip(blocks, S.states, S.cons1, S.iters)])
    assert pk.cells() == (len(sizes), sum(int(np.prod(s)) for s in sizes))
    if sizes is MARCH_SIZES:
        assert all(pk.tile_shape(d)[1] < 0 for d in (1, 2)), "expected marching sweeps"
        assert all(phx.tile_shape(p, blocks[0], d)[1] < 0 for d in (1, 2))
    dt = 1e-3
    d_dt = DeviceArray((1,))
    # pack: stage 1 then stage 2
    S.reset()
    pk.stage(1, dt=dt, cfl=0.4, dt_out=d_dt)
    synchronize()
    after1 = S.snapshot()
    dt1 = d_dt.get()[0]
    pk.stage(2, dt=dt, cfl=0.4, dt_out=d_dt)
    synchronize()
    after2 = S.snapshot()
    dt2 = d_dt.get()[0]
    # the same stages block by block through phx_fluid_stage
    S.reset()
    dts1, dts2 = [], []
    d_one = DeviceArray((1,))
    curved = geom != A.GEOM_MINKOWSKI
    for b, st, c1, it in zip(blocks, S.states, S.cons1, S.iters):
        phx.fluid_stage(p, b, st.fields(with_src=curved), 1.0, dt, st.cons, c1, st.prim,
                        iters=it, cfl=0.4, dt_out=d_one)
        synchronize()
        dts1.append(d_one.get()[0])
    _equal(S.snapshot(), after1, "stage 1")
    assert dt1 == min(dts1)
    for b, st, c1, it in zip(blocks, S.states, S.cons1, S.iters):
        f = st.fields(with_src=curved)
        f.cons = c1.ptr  # U_s = U_1
        phx.fluid_stage(p, b, f, 0.5, dt, st.cons, st.cons, st.prim, iters=it, cfl=0.4,
                        dt_out=d_one)
        synchronize()
        dts2.append(d_one.get()[0])
    _equal(S.snapshot(), after2, "stage 2")
    assert dt2 == min(dts2)
    fails = sum(int(np.sum(s[4][4:-4, 4:-4, 4:-4] == 0.0)) for s in after2)
    say(f"    con2prim failures after stage 2: {fails}")
    pk.close()


def test_pack_graph_replays_match_eager(phx):
    say("[gpu] pack graph (two stages, device dt) replayed 3x vs eager")
    geom = A.GEOM_MINKOWSKI
    rng = np.random.default_rng(5)
    p = _params(geom)
    blocks = _blocks(geom)
    S = Setup(phx, p, blocks, rng, geom)
    pk = Pack(phx, p, [(b, st, c1, it) for b, st, c1, it in
                       zip(blocks, S.states, S.cons1, S.iters)])
    d_dtin = DeviceArray.from_host(np.array([7e-4]))
    d_dt = DeviceArray((1,))
    s = C.c_void_p()
    assert phx.L.phx_rt_stream_create(C.byref(s)) == 0
    stream = s.value
    stages = [Pack.args(1, dt_dev=d_dtin, cfl=0.4, dt_out=d_dt),
              Pack.args(2, dt_dev=d_dtin, cfl=0.4, dt_out=d_dt)]
    S.reset()
    pk.stage(1, dt_dev=d_dtin, cfl=0.4, dt_out=d_dt, stream=stream)
    pk.stage(2, dt_dev=d_dtin, cfl=0.4, dt_out=d_dt, stream=stream)
    phx.L.phx_rt_stream_synchronize(stream)
    eager, eager_dt = S.snapshot(), d_dt.get()[0]
    g = pk.graph(stages, stream)
    for rep in range(3):
        S.reset()  # eager uploads and a device synchronize between replays
        d_dt.zero()
        g.launch(stream)
        phx.L.phx_rt_stream_synchronize(stream)
        _equal(S.snapshot(), eager, f"replay {rep}")
        assert d_dt.get()[0] == eager_dt
    g.close()
    phx.L.phx_rt_stream_destroy(stream)
    pk.close()


def test_pack_autotune_keeps_results(phx):
    say("[gpu] pack tile autotune")
    geom = A.GEOM_MINKOWSKI
    rng = np.random.default_rng(8)
    p = _params(geom)
    blocks = _blocks(geom)
    S = Setup(phx, p, blocks, rng, geom)
    pk = Pack(phx, p, [(b, st, c1, it) for b, st, c1, it in
                       zip(blocks, S.states, S.cons1, S.iters)])
    S.reset()
    pk.stage(1, dt=1e-3)
    synchronize()
    ref = S.snapshot()
    pk.autotune(reps=1)
    say(f"    tuned {[pk.tile_shape(d) for d in (1, 2)]}")
    S.reset()
    pk.stage(1, dt=1e-3)
    synchronize()
    _equal(S.snapshot(), ref, "after autotune")
    pk.close()


def test_pack_autotune_refused_while_graph_lives(phx):
    """Autotuning rewrites the pack's work lists in place, which a captured graph decodes
    with the grid sizes and shapes of its capture: refused (and destroy refused) while a
    graph of the pack exists; after the graph is closed, tuning and a new capture work and
    the replay matches the eager stage."""
    say("[gpu] pack autotune / destroy refused while a graph of the pack exists")
    geom = A.GEOM_MINKOWSKI
    rng = np.random.default_rng(9)
    p = _params(geom)
    blocks = _blocks(geom)
    S = Setup(phx, p, blocks, rng, geom)
    pk = Pack(phx, p, [(b, st, c1, it) for b, st, c1, it in
                       zip(blocks, S.states, S.cons1, S.iters)])
    s = C.c_void_p()
    assert phx.L.phx_rt_stream_create(C.byref(s)) == 0
    stream = s.value
    g = pk.graph([Pack.args(1, dt=1e-3)], stream)
    with pytest.raises(A.PhxError, match="graphs of the pack"):
        pk.autotune(reps=1)
    assert phx.L.phx_pack_destroy(pk.h) == A.PHX_ERR_INVALID
    g.close()
    pk.autotune(reps=1, stream=stream)
    S.reset()
    pk.stage(1, dt=1e-3, stream=stream)
    phx.L.phx_rt_stream_synchronize(stream)
    eager = S.snapshot()
    g2 = pk.graph([Pack.args(1, dt=1e-3)], stream)
    S.reset()
    g2.launch(stream)
    phx.L.phx_rt_stream_synchronize(stream)
    _equal(S.snapshot(), eager, "replay after re-tune")
    g2.close()
    phx.L.phx_rt_stream_destroy(stream)
    pk.close()


def test_pack_rejects_bad_input(phx):
    p = _params(A.GEOM_MINKOWSKI)
    b2 = block(2, (8, 8), 4)
    st = BlockState(p, b2)
    with pytest.raises(A.PhxError, match="3-D"):
        Pack(phx, p, [(b2, st, DeviceArray(st.cons.shape), None)])
    ph = fluid_params(mhd=False)
    b3 = block(3, (8, 8, 8), 4)
    st3 = BlockState(ph, b3)
    with pytest.raises(A.PhxError, match="MHD"):
        Pack(phx, ph, [(b3, st3, DeviceArray(st3.cons.shape), None)])


def test_pack_fixups_match_block_sequence(phx):
    """Pack stage with the c2p fixup and floors on == per block: the single-block stage,
    phx_c2p_fixup, phx_apply_floors on the interior; dt after the fixups."""
    say("[gpu] pack stage with fixups vs per-block stage + fixup + floors")
    geom = A.GEOM_MINKOWSKI
    rng = np.random.default_rng(77)
    floors = dict(floor_type=A.FLOOR_CONSTANT_RHO_SIE, rho0=0.3, sie0=0.05)
    p = fluid_params(gamma=13.0 / 9.0, mhd=True, recon=A.RECON_WENO5Z, riemann=A.RIEMANN_HLL,
                     geom=geom, c2p_tol=1e-10, floors=floors, ceilings=(2.5, 1e300))
    blocks = _blocks(geom)
    S = Setup(phx, p, blocks, rng, geom)
    pk = Pack(phx, p, [(b, st, c1, it) for b, st, c1, it in zip(blocks, S.states, S.cons1, S.iters)])
    fx = fixup_params(c2p_fixup=True, mhd_floors=True, bsqorho_max=20.0, bsqou_max=400.0)
    pk.set_fixup(fx)
    dt = 1e-3
    d_dt = DeviceArray((1,))
    S.reset()
    pk.stage(1, dt=dt, cfl=0.4, dt_out=d_dt)
    synchronize()
    after1, dt1 = S.snapshot(), d_dt.get()[0]
    S.reset()
    dts = []
    d_one = DeviceArray((1,))
    for b, st, c1, it in zip(blocks, S.states, S.cons1, S.iters):
        f = st.fields()
        phx.fluid_stage(p, b, f, 1.0, dt, st.cons, c1, st.prim, iters=it)
        f1 = st.fields()
        f1.cons = c1.ptr  # the fixups act on U_1
        phx.c2p_fixup(p, fx, b, f1)
        lo = tuple(b.ng for _ in range(3))
        hi = tuple(b.ng + b.nx[d] - 1 for d in range(3))
        phx.apply_floors(p, fx, b, f1, lo=lo, hi=hi)
        phx.estimate_dt(p, b, f1, True, 0.4, d_one)
        synchronize()
        dts.append(d_one.get()[0])
    _equal(S.snapshot(), after1, "stage 1 with fixups")
    assert dt1 == min(dts)
    nfail = sum(int(np.sum(s[4][4:-4, 4:-4, 4:-4] == 0.0)) for s in after1)
    say(f"    con2prim failures repaired {nfail}")
    pk.close()
