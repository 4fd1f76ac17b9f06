"""Python view of the C ABI for tests and the benchmark.

Each method forwards to one entry point of include/phx.h and raises PhxError with the
library's message on a nonzero status.  Arrays are DeviceArray objects (device memory).
"""
import ctypes as C

import numpy as np

from . import _abi as A
from .device import DeviceArray

NPRIM = {True: 10, False: 7}
NCONS = {True: 8, False: 5}


def mgr(derefine_poles=True, h=0.3, xt=0.82, alpha=14.0, x0=0.0, smooth=0.5, hexp_br=1e30,
        hexp_nsq=4.0, hexp_csq=4.0):
    """McKinney-Gammie-Ryan (FMKS) parameters; defaults of fmks.cpp:37-52."""
    return A.Mgr(int(derefine_poles), h, xt, alpha, x0, smooth, hexp_br, hexp_nsq, hexp_csq)


def fluid_params(gamma=5.0 / 3.0, cv=1.0, mhd=True, recon=A.RECON_WENO5Z,
                 riemann=A.RIEMANN_HLL, c2p_tol=1e-10, c2p_max_iter=100,
                 geom=A.GEOM_MINKOWSKI, a=0.0, floor_scale_fac=1.0, fail_on_floors=False,
                 fail_on_ceilings=False, floors=None, ceilings=None, mgr_params=None,
                 flux_arith=A.FLUX_ARITH_EXACT):
    """FluidParams with floors and ceilings disabled unless given (fixup.hpp defaults);
    flux_arith selects the arithmetic of the flux sweeps (exact: the reference's
    rounding; fast: contraction and Newton-refined reciprocals, about 1e-15 relative)."""
    b = A.Bounds(A.FLOOR_CONSTANT_RHO_SIE, -A.DBL_MAX, -A.DBL_MAX, 0.0, 0.0, 0.0, 0.0,
                 A.DBL_MAX, A.DBL_MAX)
    if floors:
        for k, v in floors.items():
            setattr(b, k, v)
    if ceilings:
        b.gam_max, b.sie_max = ceilings
    return A.FluidParams(gamma, cv, int(mhd), recon, riemann, c2p_tol, c2p_max_iter,
                         floor_scale_fac, int(fail_on_floors), int(fail_on_ceilings), b,
                         A.Geometry(geom, a, mgr_params if mgr_params is not None else mgr()),
                         flux_arith)


def fixup_params(c2p_fixup=True, strategy=A.FAILURE_INTERPOLATE, floors=True, mhd_floors=False,
                 bsqorho_max=50.0, bsqou_max=2500.0):
    """fixup package switches (fixup.cpp:63-135, 217-243; bsq ceilings default 50, 2500)."""
    return A.FixupParams(int(c2p_fixup), strategy, int(floors), int(mhd_floors), bsqorho_max,
                         bsqou_max)


def block(ndim, nx, ng=4, xmin=(0.0, 0.0, 0.0), dx=(1.0, 1.0, 1.0)):
    nxx = list(nx) + [1] * (3 - len(nx))
    return A.Block(ndim, (C.c_int * 3)(*nxx), ng, (C.c_double * 3)(*xmin),
                   (C.c_double * 3)(*dx))


def block_shape(b):
    """(nk, nj, ni) including ghost zones."""
    n = [b.nx[d] + (2 * b.ng if d < b.ndim else 0) for d in range(3)]
    return (n[2], n[1], n[0])


def _p(d):
    return None if d is None else d.ptr


class Phx:
    """One library context (phx_ctx) on one device."""

    def __init__(self, device=0):
        self.L = A.lib()
        h = C.c_void_p()
        rc = self.L.phx_create(C.byref(h), device)
        if rc != A.PHX_OK:
            raise A.PhxError(f"phx_create failed ({rc}): no usable HIP device")
        self.h = h.value

    def close(self):
        if self.h:
            self.L.phx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc, what):
        if rc != A.PHX_OK:
            msg = self.L.phx_last_error(self.h)
            raise A.PhxError(f"{what}: status {rc}: {msg.decode() if msg else ''}")
        return rc

    @staticmethod
    def fields(vars=None, **kw):
        """BlockFields of device arrays; vars: a BlockVars table of per-component pointers
        (kept alive by the caller)."""
        f = A.BlockFields()
        for k, v in kw.items():
            setattr(f, k, _p(v))
        if vars is not None:
            f.vars = C.pointer(vars)
        return f

    # ---- per-block tasks -------------------------------------------------------------------
    def prim_to_con(self, p, b, f, lo=None, hi=None, stream=None):
        lo_ = (C.c_int * 3)(*lo) if lo is not None else None
        hi_ = (C.c_int * 3)(*hi) if hi is not None else None
        return self._chk(self.L.phx_prim_to_con(self.h, C.byref(p), C.byref(b), C.byref(f),
                                                lo_, hi_, stream), "phx_prim_to_con")

    def calculate_fluxes(self, p, b, f, stream=None):
        return self._chk(self.L.phx_calculate_fluxes(self.h, C.byref(p), C.byref(b), C.byref(f),
                                                     stream), "phx_calculate_fluxes")

    def flux_ct(self, p, b, f, stream=None):
        return self._chk(self.L.phx_flux_ct(self.h, C.byref(p), C.byref(b), C.byref(f), stream),
                         "phx_flux_ct")

    def fluid_source_terms(self, p, b, f, stream=None):
        return self._chk(self.L.phx_fluid_source_terms(self.h, C.byref(p), C.byref(b),
                                                       C.byref(f), stream),
                         "phx_fluid_source_terms")

    def flux_divergence_update(self, p, b, flux, src, u_stage, u_base, beta, dt, u_out,
                               stream=None):
        return self._chk(self.L.phx_flux_divergence_update(
            self.h, C.byref(p), C.byref(b), _p(flux), _p(src), _p(u_stage), _p(u_base),
            beta, dt, _p(u_out), stream), "phx_flux_divergence_update")

    def con2prim(self, p, b, f, lo=None, hi=None, iters=None, stream=None):
        lo_ = (C.c_int * 3)(*lo) if lo is not None else None
        hi_ = (C.c_int * 3)(*hi) if hi is not None else None
        return self._chk(self.L.phx_con2prim(self.h, C.byref(p), C.byref(b), C.byref(f), lo_,
                                             hi_, _p(iters), stream), "phx_con2prim")

    def c2p_fixup(self, p, fx, b, f, stream=None):
        return self._chk(self.L.phx_c2p_fixup(self.h, C.byref(p), C.byref(fx), C.byref(b),
                                              C.byref(f), stream), "phx_c2p_fixup")

    def apply_floors(self, p, fx, b, f, lo=None, hi=None, stream=None):
        lo_ = (C.c_int * 3)(*lo) if lo is not None else None
        hi_ = (C.c_int * 3)(*hi) if hi is not None else None
        return self._chk(self.L.phx_apply_floors(self.h, C.byref(p), C.byref(fx), C.byref(b),
                                                 C.byref(f), lo_, hi_, stream),
                         "phx_apply_floors")

    def estimate_dt(self, p, b, f, use_face_speeds, cfl, dt_out, stream=None):
        return self._chk(self.L.phx_estimate_dt(self.h, C.byref(p), C.byref(b), C.byref(f),
                                                int(use_face_speeds), cfl, _p(dt_out), stream),
                         "phx_estimate_dt")

    def fluid_stage(self, p, b, f, beta, dt, u_base, u_out, prim_out, iters=None, cfl=0.4,
                    dt_out=None, stream=None, base_vars=None, out_vars=None):
        """base_vars / out_vars: BlockVars tables of U^n and of (U_{s+1}, prim_out) for
        separately allocated components (u_base / u_out / prim_out may then be None)."""
        a = A.StageArgs(beta, dt, _p(u_base), _p(u_out), _p(prim_out), _p(iters), cfl, _p(dt_out))
        if base_vars is not None:
            a.base_vars = C.pointer(base_vars)
        if out_vars is not None:
            a.out_vars = C.pointer(out_vars)
        return self._chk(self.L.phx_fluid_stage(self.h, C.byref(p), C.byref(b), C.byref(f),
                                                C.byref(a), stream), "phx_fluid_stage")

    # ---- batched con2prim ----------------------------------------------------------------
    def autotune(self, p, b, f, reps=3, stream=None):
        return self._chk(self.L.phx_autotune(self.h, C.byref(p), C.byref(b), C.byref(f), reps,
                                             stream), "phx_autotune")

    def tile_shape(self, p, b, d):
        ti, ts = C.c_int(), C.c_int()
        self._chk(self.L.phx_tile_shape(self.h, C.byref(p), C.byref(b), d, C.byref(ti),
                                        C.byref(ts)), "phx_tile_shape")
        return ti.value, ts.value

    def set_tile_shape(self, p, b, d, ti, ts):
        return self._chk(self.L.phx_set_tile_shape(self.h, C.byref(p), C.byref(b), d, ti, ts),
                         "phx_set_tile_shape")

    def con2prim_batch(self, p, n, coords, cons, prim, mu, fail, iters=None, stream=None):
        return self._chk(self.L.phx_con2prim_batch(self.h, C.byref(p), n, _p(coords), _p(cons),
                                                   _p(prim), _p(mu), _p(fail), _p(iters),
                                                   stream), "phx_con2prim_batch")

    def prim_to_con_batch(self, p, n, coords, prim, cons, stream=None):
        return self._chk(self.L.phx_prim_to_con_batch(self.h, C.byref(p), n, _p(coords),
                                                      _p(prim), _p(cons), stream),
                         "phx_prim_to_con_batch")


class BlockState:
    """Device fields of one meshblock (the MeshBlockData analogue for tests and bench)."""

    def __init__(self, p, b):
        self.p, self.b = p, b
        mhd = bool(p.mhd)
        sh = block_shape(b)
        nd = b.ndim
        self.shape = sh
        self.prim = DeviceArray((NPRIM[mhd],) + sh)
        self.cons = DeviceArray((NCONS[mhd],) + sh)
        self.temp = DeviceArray(sh)
        self.c2p_mu = DeviceArray(sh)
        self.fail = DeviceArray(sh)
        self.cell_sig = DeviceArray((nd,) + sh)
        self.face_sig = DeviceArray((nd,) + sh)
        self.flux = DeviceArray((nd, NCONS[mhd]) + sh)
        self.emf = DeviceArray((3,) + sh)
        self.src = DeviceArray((NCONS[mhd],) + sh)
        for a in (self.prim, self.cons, self.temp, self.c2p_mu, self.fail, self.cell_sig,
                  self.face_sig, self.flux, self.emf, self.src):
            a.zero()

    def fields(self, with_src=False):
        return Phx.fields(prim=self.prim, cons=self.cons, temp=self.temp, c2p_mu=self.c2p_mu,
                          fail=self.fail, cell_sig=self.cell_sig, face_sig=self.face_sig,
                          flux=self.flux, emf=self.emf, src=self.src if with_src else None)


class Mlp:
    """The learned con2prim bracket network (include/phx_mlp.h) on a context's device."""

    def __init__(self, phx, seed=1):
        self.phx, self.L = phx, phx.L
        h = C.c_void_p()
        phx._chk(self.L.phx_mlp_create(phx.h, C.byref(h), seed), "phx_mlp_create")
        self.h = h.value

    def close(self):
        if self.h:
            self.L.phx_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, w):
        w = np.ascontiguousarray(w, np.float32)
        assert w.size == A.MLP_NPARAM
        self.phx._chk(self.L.phx_mlp_set_params(self.phx.h, self.h,
                                                 w.ctypes.data_as(C.POINTER(C.c_float))),
                      "phx_mlp_set_params")

    def get_params(self):
        w = np.zeros(A.MLP_NPARAM, np.float32)
        self.phx._chk(self.L.phx_mlp_get_params(self.phx.h, self.h,
                                                 w.ctypes.data_as(C.POINTER(C.c_float))),
                      "phx_mlp_get_params")
        return w

    def set_normalization(self, mean, std, width_scale=4.0, width_min=1e-9):
        self.width = (float(width_scale), float(width_min))
        m = (C.c_float * 4)(*[float(x) for x in mean])
        s = (C.c_float * 4)(*[float(x) for x in std])
        self.phx._chk(self.L.phx_mlp_set_normalization(self.phx.h, self.h, m, s, width_scale,
                                                        width_min), "phx_mlp_set_normalization")

    def features(self, p, n, coords, cons, out, stream=None):
        self.phx._chk(self.L.phx_con2prim_features(self.phx.h, C.byref(p), self.h, n, _p(coords),
                                                    _p(cons), _p(out), stream),
                      "phx_con2prim_features")

    def forward(self, n, feat, out, stream=None):
        self.phx._chk(self.L.phx_mlp_forward(self.phx.h, self.h, n, _p(feat), _p(out), stream),
                      "phx_mlp_forward")

    def gradient(self, n, feat, target, lam, loss=None, stream=None):
        self.phx._chk(self.L.phx_mlp_gradient(self.phx.h, self.h, n, _p(feat), _p(target), lam,
                                               _p(loss), stream), "phx_mlp_gradient")

    def read_gradient(self):
        g = np.zeros(A.MLP_NPARAM, np.float32)
        self.phx._chk(self.L.phx_mlp_read_gradient(self.phx.h, self.h,
                                                    g.ctypes.data_as(C.POINTER(C.c_float))),
                      "phx_mlp_read_gradient")
        return g

    def zero_gradient(self):
        self.phx._chk(self.L.phx_mlp_zero_gradient(self.phx.h, self.h), "phx_mlp_zero_gradient")

    def adam_step(self, lr, stream=None):
        self.phx._chk(self.L.phx_mlp_adam_step(self.phx.h, self.h, lr, stream), "phx_mlp_adam_step")

    def train_step(self, n, feat, target, lr, lam, loss=None, stream=None):
        self.phx._chk(self.L.phx_mlp_train_step(self.phx.h, self.h, n, _p(feat), _p(target), lr,
                                                 lam, _p(loss), stream), "phx_mlp_train_step")

    def con2prim_batch(self, p, n, coords, cons, prim, mu, fail, iters=None, evals=None,
                       fallback=None, stream=None):
        self.phx._chk(self.L.phx_con2prim_batch_learned(
            self.phx.h, C.byref(p), self.h, n, _p(coords), _p(cons), _p(prim), _p(mu), _p(fail),
            _p(iters), _p(evals), _p(fallback), stream), "phx_con2prim_batch_learned")


class Driver:
    """One-block RK2 driver (include/phx_driver.h): the fluid task sequence of a phoebus
    cycle on device fields the driver owns."""

    def __init__(self, phx, p, b, bc=(A.BC_OUTFLOW,) * 6, cfl=0.4, dt_init_fact=1.0):
        self.phx, self.L, self.p, self.b = phx, phx.L, p, b
        h = C.c_void_p()
        phx._chk(self.L.phx_driver_create(phx.h, C.byref(p), C.byref(b), (C.c_int * 6)(*bc),
                                          cfl, dt_init_fact, C.byref(h)), "phx_driver_create")
        self.h = h.value
        self.shape = block_shape(b)

    def close(self):
        if self.h:
            self.L.phx_driver_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_fixup(self, fx):
        self.phx._chk(self.L.phx_driver_set_fixup(self.h, C.byref(fx)), "phx_driver_set_fixup")

    def set_prim(self, prim):
        prim = np.ascontiguousarray(prim, np.float64)
        assert prim.shape == (NPRIM[bool(self.p.mhd)],) + self.shape
        self.phx._chk(self.L.phx_driver_set_prim(self.h, prim.ctypes.data), "phx_driver_set_prim")

    def initialize(self):
        self.phx._chk(self.L.phx_driver_initialize(self.h), "phx_driver_initialize")

    def step(self):
        self.phx._chk(self.L.phx_driver_step(self.h), "phx_driver_step")

    def execute(self, tlim, max_steps=1 << 30):
        self.phx._chk(self.L.phx_driver_execute(self.h, tlim, max_steps), "phx_driver_execute")

    def time(self):
        t, dt, n = C.c_double(), C.c_double(), C.c_int64()
        self.phx._chk(self.L.phx_driver_time(self.h, C.byref(t), C.byref(dt), C.byref(n)),
                      "phx_driver_time")
        return t.value, dt.value, n.value

    def get(self, what):
        nv = {A.DRIVER_PRIM: NPRIM[bool(self.p.mhd)], A.DRIVER_CONS: NCONS[bool(self.p.mhd)]}
        out = np.zeros((nv.get(what, 1),) + self.shape)
        self.phx._chk(self.L.phx_driver_get(self.h, what, out.ctypes.data), "phx_driver_get")
        return out if what in nv else out[0]


class Mesh:
    """A uniform 3-D mesh of equal MHD meshblocks on one device, advanced by the RK2 cycle
    with ghost exchange, boundaries and fixups (include/phx_mesh.h).  bc: six of
    BC_OUTFLOW / BC_PERIODIC (ix1, ox1, ix2, ox2, ix3, ox3)."""

    def __init__(self, phx, p, nx, block_nx, xmin, xmax, bc, fx=None, ng=4, flux_fixup=True,
                 inflow_check=False, cfl=0.4, dt_init_fact=1.0):
        self.phx, self.L, self.p = phx, phx.L, p
        d = A.MeshDesc((C.c_int * 3)(*nx), (C.c_int * 3)(*block_nx), ng,
                       (C.c_double * 3)(*xmin), (C.c_double * 3)(*xmax), (C.c_int * 6)(*bc),
                       int(flux_fixup), int(inflow_check), cfl, dt_init_fact)
        h = C.c_void_p()
        phx._chk(self.L.phx_mesh_create(phx.h, C.byref(p), C.byref(fx) if fx is not None else None,
                                        C.byref(d), C.byref(h)), "phx_mesh_create")
        self.h = h.value
        n, cells = C.c_int(), C.c_int64()
        self.L.phx_mesh_nblocks(self.h, C.byref(n), C.byref(cells))
        self.nblocks, self.interior_cells = n.value, cells.value
        self.blocks = []
        for g in range(self.nblocks):
            b = A.Block()
            self.L.phx_mesh_block(self.h, g, C.byref(b))
            self.blocks.append(b)
        self.shape = block_shape(self.blocks[0])

    def close(self):
        if self.h:
            self.L.phx_mesh_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_prim(self, gid, prim):
        prim = np.ascontiguousarray(prim, np.float64)
        assert prim.shape == (NPRIM[True],) + self.shape
        self.phx._chk(self.L.phx_mesh_set_prim(self.h, gid, prim.ctypes.data), "phx_mesh_set_prim")

    def initialize(self):
        self.phx._chk(self.L.phx_mesh_initialize(self.h), "phx_mesh_initialize")

    def step(self):
        self.phx._chk(self.L.phx_mesh_step(self.h), "phx_mesh_step")

    def execute(self, tlim, max_steps=1 << 30):
        self.phx._chk(self.L.phx_mesh_execute(self.h, tlim, max_steps), "phx_mesh_execute")

    def autotune(self, reps=2):
        self.phx._chk(self.L.phx_mesh_autotune(self.h, reps), "phx_mesh_autotune")

    def time(self):
        t, dt, n = C.c_double(), C.c_double(), C.c_int64()
        self.phx._chk(self.L.phx_mesh_time(self.h, C.byref(t), C.byref(dt), C.byref(n)),
                      "phx_mesh_time")
        return t.value, dt.value, n.value

    def get(self, gid, what):
        nv = {A.DRIVER_PRIM: NPRIM[True], A.DRIVER_CONS: NCONS[True]}
        out = np.zeros((nv.get(what, 1),) + self.shape)
        self.phx._chk(self.L.phx_mesh_get(self.h, gid, what, out.ctypes.data), "phx_mesh_get")
        return out if what in nv else out[0]

    def set_learned(self, net=None):
        """every con2prim with the learned bracket of net (an Mlp), or plain (None)."""
        self._net = net  # kept alive while the mesh uses it
        self.phx._chk(self.L.phx_mesh_set_learned(self.h, net.h if net is not None else None),
                      "phx_mesh_set_learned")

    def learned_stats(self, reset=False):
        """(bracket misses, learned solves) since the last reset."""
        a, b = C.c_int64(), C.c_int64()
        self.phx._chk(self.L.phx_mesh_learned_stats(self.h, int(reset), C.byref(a), C.byref(b)),
                      "phx_mesh_learned_stats")
        return a.value, b.value

    def get_iters(self, gid):
        out = np.zeros(self.shape, np.int32)
        self.phx._chk(self.L.phx_mesh_get_iters(self.h, gid, out.ctypes.data),
                      "phx_mesh_get_iters")
        return out


class Graph:
    """A captured, replayable sequence of pack stages (include/phx_pack.h).  Holds its pack
    alive: the library refuses to destroy or autotune a pack while its graphs exist."""

    def __init__(self, L, h, pack=None):
        self.L, self.h, self.pack = L, h, pack

    def launch(self, stream):
        rc = self.L.phx_graph_launch(self.h, stream)
        if rc != A.PHX_OK:
            raise A.PhxError(f"phx_graph_launch: status {rc}")

    def close(self):
        if self.h:
            self.L.phx_graph_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Pack:
    """The fluid stage over a pack of 3-D MHD blocks (include/phx_pack.h).  entries:
    (block, BlockState, cons1 DeviceArray, iters DeviceArray or None) per block."""

    def __init__(self, phx, p, entries):
        self.phx, self.L, self.p = phx, phx.L, p
        self.entries = entries  # keeps the device arrays alive
        arr = (A.PackBlock * len(entries))()
        curved = p.geom.kind != A.GEOM_MINKOWSKI
        for i, (b, st, cons1, iters) in enumerate(entries):
            arr[i].block = b
            arr[i].f = st.fields(with_src=curved)
            arr[i].cons1 = cons1.ptr
            arr[i].iters = _p(iters)
        h = C.c_void_p()
        phx._chk(self.L.phx_pack_create(phx.h, C.byref(p), len(entries), arr, C.byref(h)),
                 "phx_pack_create")
        self.h = h.value

    def close(self):
        if self.h:
            rc = self.L.phx_pack_destroy(self.h)
            if rc != A.PHX_OK:
                raise A.PhxError(f"phx_pack_destroy: status {rc} (close the pack's graphs first)")
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def args(stage, dt=0.0, dt_dev=None, cfl=0.4, dt_out=None):
        return A.PackStageArgs(stage, dt, _p(dt_dev), cfl, _p(dt_out))

    def stage(self, stage, dt=0.0, dt_dev=None, cfl=0.4, dt_out=None, stream=None):
        a = self.args(stage, dt, dt_dev, cfl, dt_out)
        self.phx._chk(self.L.phx_pack_stage(self.phx.h, self.h, C.byref(a), stream),
                      "phx_pack_stage")

    def set_learned(self, net=None):
        """the stage's con2prim with the learned bracket of net (an Mlp), or plain (None)."""
        self._net = net
        self.phx._chk(self.L.phx_pack_set_learned(self.phx.h, self.h,
                                                  net.h if net is not None else None),
                      "phx_pack_set_learned")

    def learned_stats(self, reset=False):
        a, b = C.c_int64(), C.c_int64()
        self.phx._chk(self.L.phx_pack_learned_stats(self.h, int(reset), C.byref(a), C.byref(b)),
                      "phx_pack_learned_stats")
        return a.value, b.value

    def set_fixup(self, fx=None):
        """fixup switches after each stage's con2prim (None disables)."""
        self.phx._chk(self.L.phx_pack_set_fixup(self.phx.h, self.h,
                                                C.byref(fx) if fx is not None else None),
                      "phx_pack_set_fixup")

    def autotune(self, reps=3, stream=None):
        self.phx._chk(self.L.phx_pack_autotune(self.phx.h, self.h, reps, stream),
                      "phx_pack_autotune")

    def tile_shape(self, d):
        ti, ts = C.c_int(), C.c_int()
        self.phx._chk(self.L.phx_pack_tile_shape(self.h, d, C.byref(ti), C.byref(ts)),
                      "phx_pack_tile_shape")
        return ti.value, ts.value

    def cells(self):
        n, c = C.c_int(), C.c_int64()
        self.L.phx_pack_nblocks(self.h, C.byref(n), C.byref(c))
        return n.value, c.value

    def graph(self, stages, stream):
        """stages: list of PackStageArgs (Pack.args(...)); stream: a non-default stream."""
        arr = (A.PackStageArgs * len(stages))(*stages)
        h = C.c_void_p()
        self.phx._chk(self.L.phx_pack_graph_create(self.phx.h, self.h, len(stages), arr, stream,
                                                   C.byref(h)), "phx_pack_graph_create")
        return Graph(self.L, h.value, self)
