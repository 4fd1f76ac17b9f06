"""ctypes view of the CPU oracle (oracle/phx_oracle.h).

Test infrastructure only: the oracle is the parity checker of the HIP path, never part of
the product.  The shared library is built by oracle/Makefile (``__graft_entry__.build()``
runs it) into oracle/_build/.
"""
import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
# PHX_ORACLE_LIB: another build of the oracle's source (a planted-mutation test builds one)
LIBPATH = os.environ.get("PHX_ORACLE_LIB") or os.path.join(ORACLE_DIR, "_build",
                                                          "libphx_oracle.so")

DBL_MAX = np.finfo(np.float64).max

GEOM_MINKOWSKI, GEOM_SPHKS, GEOM_FMKS = 0, 1, 2
RECON_CONSTANT, RECON_LINEAR, RECON_WENO5, RECON_MP5 = 0, 1, 2, 3
RIEMANN_LLF, RIEMANN_HLL = 0, 1
FLOOR_CONSTANT, FLOOR_EXPX1_RHO_SIE, FLOOR_EXPX1_RHO_U, FLOOR_X1_RHO_SIE, FLOOR_R_RHO_SIE = range(5)


class Mgr(C.Structure):
    _fields_ = [("derefine_poles", C.c_int), ("h", C.c_double), ("xt", C.c_double),
                ("alpha", C.c_double), ("x0", C.c_double), ("smooth", C.c_double),
                ("hexp_br", C.c_double), ("hexp_nsq", C.c_double), ("hexp_csq", C.c_double)]


def mgr(derefine_poles=True, h=0.3, xt=0.82, alpha=14.0, x0=0.0, smooth=0.5, hexp_br=1e30,
        hexp_nsq=4.0, hexp_csq=4.0):
    """McKinney-Gammie-Ryan parameters (fmks.cpp:37-52 defaults)."""
    return Mgr(int(derefine_poles), h, xt, alpha, x0, smooth, hexp_br, hexp_nsq, hexp_csq)


class Geom(C.Structure):
    _fields_ = [("kind", C.c_int), ("a", C.c_double), ("mgr", Mgr)]


class Bounds(C.Structure):
    _fields_ = [("floor_type", C.c_int), ("r0", C.c_double), ("s0", C.c_double),
                ("ralpha", C.c_double), ("salpha", C.c_double),
                ("rho_min_eos", C.c_double), ("sie_min_eos", C.c_double),
                ("gam_max", C.c_double), ("sie_max", C.c_double)]


class Params(C.Structure):
    _fields_ = [("gm1", C.c_double), ("cv", C.c_double), ("mhd", C.c_int),
                ("recon", C.c_int), ("riemann", C.c_int), ("c2p_tol", C.c_double),
                ("c2p_max_iter", C.c_int), ("c2p_floor_scale_fac", C.c_double),
                ("fail_on_floors", C.c_int), ("fail_on_ceilings", C.c_int),
                ("bounds", Bounds), ("geom", Geom)]


FAILURE_INTERPOLATE, FAILURE_FLOORS = 0, 1


class Fixup(C.Structure):
    _fields_ = [("enable_c2p_fixup", C.c_int), ("failure_strategy", C.c_int),
                ("enable_floors", C.c_int), ("enable_mhd_floors", C.c_int),
                ("bsqorho_max", C.c_double), ("bsqou_max", C.c_double)]


def fixup(c2p_fixup=True, strategy=FAILURE_INTERPOLATE, floors=True, mhd_floors=False,
          bsqorho_max=50.0, bsqou_max=2500.0):
    return Fixup(int(c2p_fixup), strategy, int(floors), int(mhd_floors), bsqorho_max, bsqou_max)


class PointGeom(C.Structure):
    _fields_ = [("alpha", C.c_double), ("beta", C.c_double * 3),
                ("gcov", (C.c_double * 4) * 4), ("gcon", (C.c_double * 3) * 3),
                ("gam", (C.c_double * 3) * 3), ("detgamma", C.c_double), ("detg", C.c_double)]

    def np(self, name):
        return np.array(getattr(self, name)[:] if name == "beta" else
                        [list(r) for r in getattr(self, name)])


class Block(C.Structure):
    _fields_ = [("ndim", C.c_int), ("nx", C.c_int * 3), ("ng", C.c_int),
                ("xmin", C.c_double * 3), ("dx", C.c_double * 3)]


def build():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            build()
        L = C.CDLL(LIBPATH)
        P = C.POINTER
        dp = P(C.c_double)
        L.orc_ratio.restype = C.c_double
        L.orc_face_flux.restype = C.c_double
        L.orc_c2p.restype = C.c_double
        L.orc_c2p_residual.restype = C.c_double
        L.orc_estimate_dt.restype = C.c_double
        L.orc_regula_falsi.restype = C.c_double
        L.orc_regula_falsi.argtypes = [C.CFUNCTYPE(C.c_double, C.c_double, C.c_void_p),
                                       C.c_void_p, C.c_double, C.c_double, C.c_double,
                                       C.c_double, C.c_int, P(C.c_int), P(C.c_int)]
        for name in ("orc_plm",):
            getattr(L, name).argtypes = [C.c_double] * 3 + [dp, dp]
        for name in ("orc_weno5z", "orc_mp5"):
            getattr(L, name).argtypes = [C.c_double] * 5 + [dp, dp]
        L.orc_geom_point.argtypes = [P(Geom), C.c_double, C.c_double, C.c_double, P(PointGeom)]
        _lib = L
    return _lib


def ptr(a):
    """Pointer to the data of a contiguous numpy array (float64 or int32)."""
    assert a.flags.c_contiguous
    if a.dtype == np.float64:
        return a.ctypes.data_as(C.POINTER(C.c_double))
    if a.dtype == np.int32:
        return a.ctypes.data_as(C.POINTER(C.c_int))
    raise TypeError(a.dtype)


def default_params(gamma=5.0 / 3.0, mhd=True, recon=RECON_WENO5, riemann=RIEMANN_HLL,
                   geom=GEOM_MINKOWSKI, a=0.0, tol=1e-10, max_iter=100, mgr_params=None):
    """Floors and ceilings disabled (fixup.hpp default Floors()/Ceilings(); EOS minima 0)."""
    b = Bounds(FLOOR_CONSTANT, -DBL_MAX, -DBL_MAX, 0.0, 0.0, 0.0, 0.0, DBL_MAX, DBL_MAX)
    return Params(gamma - 1.0, 1.0, int(mhd), recon, riemann, tol, max_iter, 1.0, 0, 0, b,
                  Geom(geom, a, mgr_params if mgr_params is not None else mgr()))


def nprim(mhd):
    return 10 if mhd else 7


def ncons(mhd):
    return 8 if mhd else 5


def point_geom(params, X):
    g = PointGeom()
    lib().orc_geom_point(C.byref(params.geom), X[0], X[1], X[2], C.byref(g))
    return g


def recon(kind, q):
    ql, qr = C.c_double(), C.c_double()
    L = lib()
    if kind == RECON_LINEAR:
        L.orc_plm(q[1], q[2], q[3], C.byref(ql), C.byref(qr))
    elif kind == RECON_WENO5:
        L.orc_weno5z(*[float(x) for x in q], C.byref(ql), C.byref(qr))
    elif kind == RECON_MP5:
        L.orc_mp5(*[float(x) for x in q], C.byref(ql), C.byref(qr))
    else:
        return float(q[2]), float(q[2])
    return ql.value, qr.value


def face_flux(params, d, X, ql, qr):
    f = np.zeros(8)
    Xa = (C.c_double * 3)(*X)
    s = lib().orc_face_flux(C.byref(params), d, Xa, ptr(np.ascontiguousarray(ql, np.float64)),
                            ptr(np.ascontiguousarray(qr, np.float64)), ptr(f))
    return f[:ncons(params.mhd)], s


def p2c(params, X, prim):
    g = point_geom(params, X)
    cons = np.zeros(8)
    sig = np.zeros(3)
    lib().orc_p2c(C.byref(params), C.byref(g), ptr(np.ascontiguousarray(prim, np.float64)),
                  ptr(cons), ptr(sig))
    return cons[:ncons(params.mhd)], sig


def c2p(params, X, cons, mu_guess=0.0):
    g = point_geom(params, X)
    cs = np.zeros(8)
    cs[:len(cons)] = cons
    prim = np.zeros(10)
    t = C.c_double(0.0)
    mu = C.c_double(mu_guess)
    sig = np.zeros(3)
    it = C.c_int(0)
    Xa = (C.c_double * 3)(*X)
    st = lib().orc_c2p(C.byref(params), C.byref(g), Xa, ptr(cs), ptr(prim), C.byref(t),
                       C.byref(mu), ptr(sig), C.byref(it))
    return dict(fail=st, prim=prim[:nprim(params.mhd)], cons=cs[:ncons(params.mhd)],
                temp=t.value, mu=mu.value, sig=sig, iters=it.value)


def make_block(ndim, nx, ng=4, xmin=(0.0, 0.0, 0.0), dx=(1.0, 1.0, 1.0)):
    nxx = list(nx) + [1] * (3 - len(nx))
    return Block(ndim, (C.c_int * 3)(*nxx), ng, (C.c_double * 3)(*xmin), (C.c_double * 3)(*dx))


def block_shape(b):
    """(nk, nj, ni) including ghosts."""
    n = [b.nx[d] + (2 * b.ng if d < b.ndim else 0) for d in range(3)]
    return (n[2], n[1], n[0])


def interior(b):
    """slices (k, j, i) of the interior cells."""
    sl = []
    for d in (2, 1, 0):
        s = b.ng if d < b.ndim else 0
        sl.append(slice(s, s + b.nx[d]))
    return tuple(sl)


def calculate_fluxes(params, b, prim):
    sh = block_shape(b)
    nc = ncons(params.mhd)
    flux = np.zeros((b.ndim, nc) + sh)
    fsig = np.zeros((b.ndim,) + sh)
    lib().orc_calculate_fluxes(C.byref(params), C.byref(b), ptr(np.ascontiguousarray(prim)),
                               ptr(flux), ptr(fsig))
    return flux, fsig


def flux_ct(params, b, flux):
    emf = np.zeros((3,) + block_shape(b))
    lib().orc_flux_ct(C.byref(params), C.byref(b), ptr(flux), ptr(emf))
    return emf


def flux_divergence_update(params, b, flux, src, u_stage, u_base, beta, dt):
    u_out = u_stage.copy()
    lib().orc_flux_divergence_update(
        C.byref(params), C.byref(b), ptr(flux), ptr(src) if src is not None else None,
        ptr(u_stage), ptr(u_base), C.c_double(beta), C.c_double(dt), ptr(u_out))
    return u_out


def con2prim_block(params, b, cons, prim, mu, lo=None, hi=None):
    sh = block_shape(b)
    if lo is None:
        lo = (0, 0, 0)
        hi = (sh[2] - 1, sh[1] - 1, sh[0] - 1)
    temp = np.zeros(sh)
    sig = np.zeros((b.ndim,) + sh)
    fail = np.zeros(sh)
    iters = np.zeros(sh, np.int32)
    lib().orc_con2prim_block(C.byref(params), C.byref(b), (C.c_int * 3)(*lo),
                             (C.c_int * 3)(*hi), ptr(cons), ptr(prim), ptr(temp), ptr(mu),
                             ptr(sig), ptr(fail), ptr(iters))
    return dict(temp=temp, sig=sig, fail=fail, iters=iters)


def prim_to_con_block(params, b, prim, lo=None, hi=None):
    sh = block_shape(b)
    if lo is None:
        lo = (0, 0, 0)
        hi = (sh[2] - 1, sh[1] - 1, sh[0] - 1)
    cons = np.zeros((ncons(params.mhd),) + sh)
    sig = np.zeros((b.ndim,) + sh)
    lib().orc_prim_to_con_block(C.byref(params), C.byref(b), (C.c_int * 3)(*lo),
                                (C.c_int * 3)(*hi), ptr(np.ascontiguousarray(prim)),
                                ptr(cons), ptr(sig))
    return cons, sig


def c2p_fixup_block(params, fx, b, prim, cons, temp, cell_sig, fail):
    """orc_c2p_fixup_block in place on prim, cons, temp, cell_sig."""
    lib().orc_c2p_fixup_block(C.byref(params), C.byref(fx), C.byref(b), ptr(prim), ptr(cons),
                              ptr(temp), ptr(cell_sig), ptr(fail))


def apply_floors_block(params, fx, b, prim, cons, temp, mu, cell_sig, lo=None, hi=None):
    """orc_apply_floors_block in place over lo..hi (default: the entire block)."""
    sh = block_shape(b)
    if lo is None:
        lo = (0, 0, 0)
        hi = (sh[2] - 1, sh[1] - 1, sh[0] - 1)
    lib().orc_apply_floors_block(C.byref(params), C.byref(fx), C.byref(b), (C.c_int * 3)(*lo),
                                 (C.c_int * 3)(*hi), ptr(prim), ptr(cons), ptr(temp), ptr(mu),
                                 ptr(cell_sig))


def source_terms_block(params, b, prim):
    src = np.zeros((ncons(params.mhd),) + block_shape(b))
    lib().orc_source_terms_block(C.byref(params), C.byref(b), ptr(np.ascontiguousarray(prim)),
                                 ptr(src))
    return src


def estimate_dt(params, b, cell_sig, face_sig, cfl):
    return lib().orc_estimate_dt(C.byref(params), C.byref(b), ptr(cell_sig),
                                 ptr(face_sig) if face_sig is not None else None,
                                 C.c_double(cfl))


def c2p_batch(params, coords, cons, mu=None):
    n = coords.shape[1]
    cons = np.ascontiguousarray(cons).copy()
    prim = np.zeros((nprim(params.mhd), n))
    mu = np.zeros(n) if mu is None else np.ascontiguousarray(mu, np.float64).copy()
    fail = np.zeros(n)
    iters = np.zeros(n, np.int32)
    lib().orc_c2p_batch(C.byref(params), C.c_long(n), ptr(np.ascontiguousarray(coords)),
                        ptr(cons), ptr(prim), ptr(mu), ptr(fail), ptr(iters))
    return dict(cons=cons, prim=prim, mu=mu, fail=fail, iters=iters)


def p2c_batch(params, coords, prim):
    n = coords.shape[1]
    cons = np.zeros((ncons(params.mhd), n))
    lib().orc_p2c_batch(C.byref(params), C.c_long(n), ptr(np.ascontiguousarray(coords)),
                        ptr(np.ascontiguousarray(prim)), ptr(cons))
    return cons


def tmunu(params, X, prim):
    T = np.zeros((4, 4))
    lib().orc_tmunu(C.byref(params), C.c_double(X[0]), C.c_double(X[1]), C.c_double(X[2]),
                    ptr(np.ascontiguousarray(prim, np.float64)), ptr(T))
    return T


def sources(params, X, prim):
    s = np.zeros(5)
    lib().orc_sources(C.byref(params), C.c_double(X[0]), C.c_double(X[1]), C.c_double(X[2]),
                      ptr(np.ascontiguousarray(prim, np.float64)), ptr(s))
    return s


def geom_dg(params, X):
    dg = np.zeros((4, 4, 4))
    lib().orc_geom_dg(C.byref(params.geom), C.c_double(X[0]), C.c_double(X[1]),
                      C.c_double(X[2]), ptr(dg))
    return dg


def geom_gcon4(params, X):
    gi = np.zeros((4, 4))
    lib().orc_geom_gcon4(C.byref(params.geom), C.c_double(X[0]), C.c_double(X[1]),
                         C.c_double(X[2]), ptr(gi))
    return gi


def geom_gradlnalpha(params, X):
    da = np.zeros(4)
    lib().orc_geom_gradlnalpha(C.byref(params.geom), C.c_double(X[0]), C.c_double(X[1]),
                               C.c_double(X[2]), ptr(da))
    return da


def threads():
    """OpenMP threads the oracle uses."""
    return lib().orc_num_threads()


def set_threads(n):
    """Sets the OpenMP threads of the oracle's later parallel loops."""
    lib().orc_set_num_threads(int(n))


BC_OUTFLOW, BC_PERIODIC = 0, 1


def apply_bcs(params, b, q, kinds):
    """orc_apply_bcs in place on q[nvar][k][j][i]."""
    assert q.flags.c_contiguous and q.dtype == np.float64
    lib().orc_apply_bcs(C.byref(params), C.byref(b), q.shape[0], ptr(q), (C.c_int * 6)(*kinds))


class Driver:
    """The one-block RK2 task sequence of a phoebus cycle on the oracle's block functions
    (phoebus_driver.cpp:147-202 for the start, :204-738 per stage; rk2 containers "base"
    and "1" with beta = 1, 1/2): the CPU counterpart of include/phx_driver.h, step for
    step and container for container, so that results compare bit for bit in Minkowski.
    Each con2prim is followed by the c2p fixup and the floors (phoebus_driver.cpp:586-597,
    664-676) when fixup switches are given."""

    BETA = (1.0, 0.5)

    def __init__(self, params, b, kinds, cfl, dt_init_fact=1.0, fx=None):
        self.p, self.b, self.kinds, self.cfl = params, b, tuple(kinds), cfl
        self.fx = fx if fx is not None else fixup(c2p_fixup=False, floors=False)
        self.dt_init_fact = dt_init_fact
        sh = block_shape(b)
        nc, npr = ncons(params.mhd), nprim(params.mhd)
        self.cons = [np.zeros((nc,) + sh), np.zeros((nc,) + sh)]
        self.prim = np.zeros((npr,) + sh)
        self.temp = np.zeros(sh)
        self.mu = np.zeros(sh)
        self.fail = np.zeros(sh)
        self.cell_sig = np.zeros((3,) + sh)
        self.face_sig = None
        self.flux = np.zeros((3, nc) + sh)
        self.iters = np.zeros(sh, np.int32)
        self.time, self.dt, self.ncycle = 0.0, 0.0, 0
        self.has_face_speeds = False
        self.tlim = 1e300
        self.n = [sh[2], sh[1], sh[0]]

    def _c2p(self, container):
        lo = (0, 0, 0)
        hi = (self.n[0] - 1, self.n[1] - 1, self.n[2] - 1)
        lib().orc_con2prim_block(C.byref(self.p), C.byref(self.b), (C.c_int * 3)(*lo),
                                 (C.c_int * 3)(*hi), ptr(self.cons[container]), ptr(self.prim),
                                 ptr(self.temp), ptr(self.mu), ptr(self.cell_sig),
                                 ptr(self.fail), ptr(self.iters))

    def _fill_derived(self, container):
        self._c2p(container)
        lib().orc_c2p_fixup_block(C.byref(self.p), C.byref(self.fx), C.byref(self.b),
                                  ptr(self.prim), ptr(self.cons[container]), ptr(self.temp),
                                  ptr(self.cell_sig), ptr(self.fail))
        self._floors(container)

    def _floors(self, container):
        lo = (0, 0, 0)
        hi = (self.n[0] - 1, self.n[1] - 1, self.n[2] - 1)
        lib().orc_apply_floors_block(C.byref(self.p), C.byref(self.fx), C.byref(self.b),
                                     (C.c_int * 3)(*lo), (C.c_int * 3)(*hi), ptr(self.prim),
                                     ptr(self.cons[container]), ptr(self.temp), ptr(self.mu),
                                     ptr(self.cell_sig))

    def _dt(self):
        fs = self.face_sig if self.has_face_speeds else None
        d = lib().orc_estimate_dt(C.byref(self.p), C.byref(self.b), ptr(self.cell_sig),
                                  ptr(fs) if fs is not None else None, C.c_double(self.cfl))
        self.has_face_speeds = True
        return d

    def initialize(self, prim0):
        self.prim[...] = prim0
        lo = (0, 0, 0)
        hi = (self.n[0] - 1, self.n[1] - 1, self.n[2] - 1)
        sig = np.zeros_like(self.cell_sig)
        lib().orc_prim_to_con_block(C.byref(self.p), C.byref(self.b), (C.c_int * 3)(*lo),
                                    (C.c_int * 3)(*hi), ptr(self.prim), ptr(self.cons[0]),
                                    ptr(sig))
        self._floors(0)
        apply_bcs(self.p, self.b, self.cons[0], self.kinds)
        self._c2p(0)
        self.cons[1][...] = self.cons[0]
        self.dt = self.dt_init_fact * self._dt()

    def stage(self, stage):
        sc0, sc1 = (0, 1) if stage == 1 else (1, 0)
        nc = ncons(self.p.mhd)
        sh = block_shape(self.b)
        src = None
        if self.p.geom.kind != GEOM_MINKOWSKI:
            src = source_terms_block(self.p, self.b, self.prim)
        flux = np.zeros((self.b.ndim, nc) + sh)
        fsig = np.zeros((self.b.ndim,) + sh)
        lib().orc_calculate_fluxes(C.byref(self.p), C.byref(self.b), ptr(self.prim), ptr(flux),
                                   ptr(fsig))
        if self.p.mhd and self.b.ndim > 1:
            emf = np.zeros((3,) + sh)
            lib().orc_flux_ct(C.byref(self.p), C.byref(self.b), ptr(flux), ptr(emf))
        self.face_sig = np.zeros((3,) + sh)
        self.face_sig[:self.b.ndim] = fsig
        lib().orc_flux_divergence_update(
            C.byref(self.p), C.byref(self.b), ptr(flux), ptr(src) if src is not None else None,
            ptr(self.cons[sc0]), ptr(self.cons[0]), C.c_double(self.BETA[stage - 1]),
            C.c_double(self.dt), ptr(self.cons[sc1]))
        self._fill_derived(sc1)
        apply_bcs(self.p, self.b, self.cons[sc1], self.kinds)
        self._fill_derived(sc1)

    def step(self):
        if self.time + self.dt > self.tlim:
            self.dt = self.tlim - self.time
        for s in (1, 2):
            self.stage(s)
        self.time += self.dt
        self.ncycle += 1
        self.dt = min(self._dt(), 2.0 * self.dt)

    def execute(self, tlim, max_steps=1 << 30):
        self.tlim = tlim
        n = 0
        while n < max_steps and self.time < self.tlim:
            self.step()
            n += 1


def mgr_transform(m, X):
    Cc = (C.c_double * 3)()
    Jcov = ((C.c_double * 3) * 3)()
    Jcon = ((C.c_double * 3) * 3)()
    lib().orc_mgr_transform(C.byref(m), C.c_double(X[0]), C.c_double(X[1]), C.c_double(X[2]),
                            Cc, Jcov, Jcon)
    return np.array(Cc[:]), np.array([r[:] for r in Jcov]), np.array([r[:] for r in Jcon])


class MeshDriver:
    """The RK2 cycle of phoebus on a uniform mesh of equal 3-D blocks, on the oracle's block
    functions, in the reference's task order (phoebus_driver.cpp:147-202 for the start,
    :204-738 per stage): sources, fluxes, CT, FixFluxes (gr_outflow faces), update,
    FillDerived (con2prim over the entire blocks) + c2p fixup + floors (entire), the
    same-level ghost exchange of the FillGhost fields (bc_vars = primitive: primitives,
    temperature, con2prim root, fail flag, conserved), gr_outflow boundaries (copy of the
    first interior layer, faces in the order x1, x2, x3), ConvertBoundaryConditions (outer-x1
    inflow check for FMKS, p2c on the ghost slabs of every face that is not a block
    boundary), FillDerived + fixups again; EstimateTimestep after the last stage.  Block
    gid = lx1 + nb1 (lx2 + nb2 lx3); block edges xmin + (xmax - xmin) lx / nb."""

    BETA = (1.0, 0.5)
    NEX = 21  # exchanged fields: 10 prims, 8 cons, temp, mu, fail

    def __init__(self, params, nx, block_nx, xmin, xmax, bc, fx=None, ng=4, flux_fixup=True,
                 inflow_check=False, cfl=0.4, dt_init_fact=1.0):
        self.p, self.ng, self.cfl, self.dt_init_fact = params, ng, cfl, dt_init_fact
        self.fx = fx if fx is not None else fixup(c2p_fixup=False, floors=False)
        self.N, self.bnx = list(nx), list(block_nx)
        self.nb = [nx[d] // block_nx[d] for d in range(3)]
        self.bc = list(bc)
        self.periodic = [bc[2 * d] == BC_PERIODIC for d in range(3)]
        self.flux_fixup, self.inflow = flux_fixup, inflow_check and params.geom.kind == GEOM_FMKS
        self.blocks, self.loc = [], []
        for g in range(self.nb[0] * self.nb[1] * self.nb[2]):
            lx = (g % self.nb[0], (g // self.nb[0]) % self.nb[1], g // (self.nb[0] * self.nb[1]))
            bmin = [xmin[d] + (xmax[d] - xmin[d]) * (lx[d] / self.nb[d]) for d in range(3)]
            dx = [(xmax[d] - xmin[d]) / nx[d] for d in range(3)]
            self.blocks.append(make_block(3, block_nx, ng, tuple(bmin), tuple(dx)))
            self.loc.append(lx)
        sh = block_shape(self.blocks[0])
        self.sh = sh
        nb = len(self.blocks)
        self.prim = [np.zeros((10,) + sh) for _ in range(nb)]
        self.cons = [[np.zeros((8,) + sh), np.zeros((8,) + sh)] for _ in range(nb)]
        self.temp = [np.zeros(sh) for _ in range(nb)]
        self.mu = [np.zeros(sh) for _ in range(nb)]
        self.fail = [np.zeros(sh) for _ in range(nb)]
        self.sig = [np.zeros((3,) + sh) for _ in range(nb)]
        self.fsig = [np.zeros((3,) + sh) for _ in range(nb)]
        self.time, self.dt, self.ncycle = 0.0, 0.0, 0
        self.has_face_speeds = False
        self.tlim = 1e300

    # ---- block tasks -------------------------------------------------------------------
    def _entire(self):
        return (0, 0, 0), (self.sh[2] - 1, self.sh[1] - 1, self.sh[0] - 1)

    def _c2p(self, g, c):
        lo, hi = self._entire()
        it = np.zeros(self.sh, np.int32)
        lib().orc_con2prim_block(C.byref(self.p), C.byref(self.blocks[g]), (C.c_int * 3)(*lo),
                                 (C.c_int * 3)(*hi), ptr(self.cons[g][c]), ptr(self.prim[g]),
                                 ptr(self.temp[g]), ptr(self.mu[g]), ptr(self.sig[g]),
                                 ptr(self.fail[g]), ptr(it))

    def _fixups(self, g, c):
        b = self.blocks[g]
        lib().orc_c2p_fixup_block(C.byref(self.p), C.byref(self.fx), C.byref(b), ptr(self.prim[g]),
                                  ptr(self.cons[g][c]), ptr(self.temp[g]), ptr(self.sig[g]),
                                  ptr(self.fail[g]))
        lo, hi = self._entire()
        lib().orc_apply_floors_block(C.byref(self.p), C.byref(self.fx), C.byref(b),
                                     (C.c_int * 3)(*lo), (C.c_int * 3)(*hi), ptr(self.prim[g]),
                                     ptr(self.cons[g][c]), ptr(self.temp[g]), ptr(self.mu[g]),
                                     ptr(self.sig[g]))

    def _fill_derived(self, c):
        for g in range(len(self.blocks)):
            self._c2p(g, c)
            self._fixups(g, c)

    def _fields(self, g, c):
        return [self.prim[g], self.cons[g][c], self.temp[g][None], self.mu[g][None],
                self.fail[g][None]]

    def _exchange(self, c):
        ng = self.ng
        G = np.full((self.NEX, self.N[2] + 2 * ng, self.N[1] + 2 * ng, self.N[0] + 2 * ng), np.nan)
        for g, lx in enumerate(self.loc):
            F = np.concatenate(self._fields(g, c))
            sl = tuple(slice(ng + lx[d] * self.bnx[d], ng + (lx[d] + 1) * self.bnx[d])
                       for d in (2, 1, 0))
            G[(slice(None),) + sl] = F[:, ng:-ng, ng:-ng, ng:-ng]
        for d in range(3):  # periodic wrap of the padding (axis 3 - d)
            if not self.periodic[d]:
                continue
            ax, n = 3 - d, self.N[d]
            lo_src = [slice(None)] * 4
            hi_src = [slice(None)] * 4
            lo_dst = [slice(None)] * 4
            hi_dst = [slice(None)] * 4
            lo_dst[ax], lo_src[ax] = slice(0, ng), slice(n, n + ng)
            hi_dst[ax], hi_src[ax] = slice(n + ng, n + 2 * ng), slice(ng, 2 * ng)
            G[tuple(lo_dst)] = G[tuple(lo_src)]
            G[tuple(hi_dst)] = G[tuple(hi_src)]
        for g, lx in enumerate(self.loc):
            sl = tuple(slice(lx[d] * self.bnx[d], lx[d] * self.bnx[d] + self.bnx[d] + 2 * ng)
                       for d in (2, 1, 0))
            P = G[(slice(None),) + sl]
            valid = np.ones(self.sh, bool)
            ghost = np.zeros(self.sh, bool)
            for d in range(3):
                ax = 2 - d
                idx = np.arange(self.sh[ax])
                gi = lx[d] * self.bnx[d] + idx - ng
                ok = (gi >= 0) & (gi < self.N[d]) | self.periodic[d]
                shp = [1, 1, 1]
                shp[ax] = -1
                valid &= ok.reshape(shp)
                ghost |= ((idx < ng) | (idx >= ng + self.bnx[d])).reshape(shp)
            m = valid & ghost
            for F, rows in zip(self._fields(g, c), (slice(0, 10), slice(10, 18), slice(18, 19),
                                                   slice(19, 20), slice(20, 21))):
                F[:, m] = P[rows][:, m]

    def _outflow(self, c):
        ng = self.ng
        for d in range(3):
            ax = 3 - d  # axis of dimension d in a [var][k][j][i] array
            for side in (0, 1):
                if self.bc[2 * d + side] != BC_OUTFLOW:
                    continue
                for g, lx in enumerate(self.loc):
                    if lx[d] != (0 if side == 0 else self.nb[d] - 1):
                        continue
                    ref = ng if side == 0 else ng + self.bnx[d] - 1
                    gh = slice(0, ng) if side == 0 else slice(ng + self.bnx[d], self.sh[2 - d])
                    for F in self._fields(g, c):
                        dst = [slice(None)] * 4
                        src = [slice(None)] * 4
                        dst[ax], src[ax] = gh, slice(ref, ref + 1)
                        F[tuple(dst)] = F[tuple(src)]

    def _convert(self, c):
        ng = self.ng
        for g, lx in enumerate(self.loc):
            b = self.blocks[g]
            slabs = []
            for d in range(3):
                for side in (0, 1):
                    if lx[d] != (0 if side == 0 else self.nb[d] - 1):
                        continue
                    lo = [0, 0, 0]
                    hi = [self.sh[2] - 1, self.sh[1] - 1, self.sh[0] - 1]
                    if side == 0:
                        hi[d] = ng - 1
                    else:
                        lo[d] = ng + self.bnx[d]
                    slabs.append((d, side, lo, hi))
            for d, side, lo, hi in slabs:
                if d == 0 and side == 1 and self.inflow:
                    lib().orc_inflow_check_block(C.byref(self.p), C.byref(b), (C.c_int * 3)(*lo),
                                                 (C.c_int * 3)(*hi), ptr(self.prim[g]))
            for d, side, lo, hi in slabs:
                lib().orc_prim_to_con_block(C.byref(self.p), C.byref(b), (C.c_int * 3)(*lo),
                                            (C.c_int * 3)(*hi), ptr(self.prim[g]),
                                            ptr(self.cons[g][c]), ptr(self.sig[g]))

    def _fix_fluxes(self, g, flux):
        if not self.flux_fixup:
            return
        ng, lx = self.ng, self.loc[g]
        inn = [slice(ng, ng + self.bnx[d]) for d in range(3)]
        for d in range(3):
            for side in (0, 1):
                if self.bc[2 * d + side] != BC_OUTFLOW or lx[d] != (0 if side == 0 else self.nb[d] - 1):
                    continue
                face = ng if side == 0 else ng + self.bnx[d]
                sl = [inn[2], inn[1], inn[0]]
                sl[2 - d] = face
                F = flux[d, 0]
                F[tuple(sl)] = np.minimum(F[tuple(sl)], 0.0) if side == 0 else np.maximum(F[tuple(sl)], 0.0)

    def _dt(self):
        lo = np.inf
        for g, b in enumerate(self.blocks):
            fs = self.fsig[g] if self.has_face_speeds else None
            lo = min(lo, lib().orc_estimate_dt(C.byref(self.p), C.byref(b), ptr(self.sig[g]),
                                               ptr(fs) if fs is not None else None,
                                               C.c_double(self.cfl)))
        self.has_face_speeds = True
        return lo

    # ---- driver ------------------------------------------------------------------------
    def set_prim(self, g, prim):
        self.prim[g][...] = prim

    def initialize(self):
        lo, hi = self._entire()
        for g, b in enumerate(self.blocks):
            lib().orc_prim_to_con_block(C.byref(self.p), C.byref(b), (C.c_int * 3)(*lo),
                                        (C.c_int * 3)(*hi), ptr(self.prim[g]),
                                        ptr(self.cons[g][0]), ptr(self.sig[g]))
            lib().orc_apply_floors_block(C.byref(self.p), C.byref(self.fx), C.byref(b),
                                         (C.c_int * 3)(*lo), (C.c_int * 3)(*hi),
                                         ptr(self.prim[g]), ptr(self.cons[g][0]),
                                         ptr(self.temp[g]), ptr(self.mu[g]), ptr(self.sig[g]))
        self._exchange(0)
        self._outflow(0)
        self._convert(0)
        for g in range(len(self.blocks)):
            self._c2p(g, 0)
            self.cons[g][1][...] = self.cons[g][0]
        self.time, self.ncycle = 0.0, 0
        self.has_face_speeds = False
        self.dt = self.dt_init_fact * self._dt()

    def stage(self, st):
        sc0, sc1 = (0, 1) if st == 1 else (1, 0)
        for g, b in enumerate(self.blocks):
            src = None
            if self.p.geom.kind != GEOM_MINKOWSKI:
                src = source_terms_block(self.p, b, self.prim[g])
            flux = np.zeros((3, 8) + self.sh)
            fsig = np.zeros((3,) + self.sh)
            lib().orc_calculate_fluxes(C.byref(self.p), C.byref(b), ptr(self.prim[g]),
                                       ptr(flux), ptr(fsig))
            emf = np.zeros((3,) + self.sh)
            lib().orc_flux_ct(C.byref(self.p), C.byref(b), ptr(flux), ptr(emf))
            self._fix_fluxes(g, flux)
            self.fsig[g] = fsig
            lib().orc_flux_divergence_update(
                C.byref(self.p), C.byref(b), ptr(flux), ptr(src) if src is not None else None,
                ptr(self.cons[g][sc0]), ptr(self.cons[g][0]), C.c_double(self.BETA[st - 1]),
                C.c_double(self.dt), ptr(self.cons[g][sc1]))
        self._fill_derived(sc1)
        self._exchange(sc1)
        self._outflow(sc1)
        self._convert(sc1)
        self._fill_derived(sc1)

    def step(self):
        if self.time + self.dt > self.tlim:
            self.dt = self.tlim - self.time
        for st in (1, 2):
            self.stage(st)
        self.time += self.dt
        self.ncycle += 1
        self.dt = min(self._dt(), 2.0 * self.dt)
