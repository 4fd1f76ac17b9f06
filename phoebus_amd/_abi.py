"""ctypes binding of libphx.so (include/phx.h, phx_runtime.h, phx_mlp.h, phx_driver.h,
phx_pack.h, phx_fixup.h, phx_mesh.h).

The library is the product: HIP kernels for gfx950 behind a C ABI.  Loading fails loudly
when it has not been built; there is no CPU fallback.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
MLP_NPARAM = 8770
LIBPATH = os.path.join(HERE, "lib", "libphx.so")

PHX_OK, PHX_ERR_INVALID, PHX_ERR_UNSUPPORTED, PHX_ERR_DEVICE, PHX_ERR_NOMEM = 0, -1, -2, -3, -4
GEOM_MINKOWSKI, GEOM_SPHERICAL_KS, GEOM_FMKS = 0, 1, 2
RECON_CONSTANT, RECON_LINEAR, RECON_WENO5Z, RECON_MP5 = 0, 1, 2, 3
RIEMANN_LLF, RIEMANN_HLL = 0, 1
FLUX_ARITH_EXACT, FLUX_ARITH_FAST = 0, 1
BC_OUTFLOW, BC_PERIODIC = 0, 1
DRIVER_PRIM, DRIVER_CONS, DRIVER_FAIL, DRIVER_C2P_MU = 0, 1, 2, 3
FAILURE_INTERPOLATE, FAILURE_FLOORS = 0, 1
(FLOOR_CONSTANT_RHO_SIE, FLOOR_EXPX1_RHO_SIE, FLOOR_EXPX1_RHO_U, FLOOR_X1_RHO_SIE,
 FLOOR_R_RHO_SIE) = range(5)

DBL_MAX = 1.7976931348623157e308


class Mgr(C.Structure):
    _fields_ = [("derefine_poles", C.c_int), ("h", C.c_double), ("xt", C.c_double),
                ("alpha", C.c_double), ("x0", C.c_double), ("smooth", C.c_double),
                ("hexp_br", C.c_double), ("hexp_nsq", C.c_double), ("hexp_csq", C.c_double)]


class Geometry(C.Structure):
    _fields_ = [("kind", C.c_int), ("a", C.c_double), ("mgr", Mgr)]


class Bounds(C.Structure):
    _fields_ = [("floor_type", C.c_int), ("rho0", C.c_double), ("sie0", C.c_double),
                ("rho_exp", C.c_double), ("sie_exp", C.c_double),
                ("rho_min_eos", C.c_double), ("sie_min_eos", C.c_double),
                ("gam_max", C.c_double), ("sie_max", C.c_double)]


class FluidParams(C.Structure):
    _fields_ = [("gamma", C.c_double), ("cv", C.c_double), ("mhd", C.c_int),
                ("recon", C.c_int), ("riemann", C.c_int), ("c2p_tol", C.c_double),
                ("c2p_max_iter", C.c_int), ("c2p_floor_scale_fac", C.c_double),
                ("c2p_fail_on_floors", C.c_int), ("c2p_fail_on_ceilings", C.c_int),
                ("bounds", Bounds), ("geom", Geometry), ("flux_arith", C.c_int)]


class Block(C.Structure):
    _fields_ = [("ndim", C.c_int), ("nx", C.c_int * 3), ("ng", C.c_int),
                ("xmin", C.c_double * 3), ("dx", C.c_double * 3)]


class BlockVars(C.Structure):
    """phx_block_vars: per-component device pointers (separately allocated variables)."""
    _fields_ = [("prim", C.c_void_p * 10), ("cons", C.c_void_p * 8),
                ("flux", (C.c_void_p * 8) * 3), ("src", C.c_void_p * 8)]


class BlockFields(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in ("prim", "cons", "temp", "c2p_mu", "fail", "cell_sig",
                                           "face_sig", "flux", "emf", "src")] + [
        ("vars", C.POINTER(BlockVars))]


class StageArgs(C.Structure):
    _fields_ = [("beta", C.c_double), ("dt", C.c_double), ("u_base", C.c_void_p),
                ("u_out", C.c_void_p), ("prim_out", C.c_void_p), ("iters", C.c_void_p),
                ("cfl", C.c_double), ("dt_out", C.c_void_p),
                ("base_vars", C.POINTER(BlockVars)), ("out_vars", C.POINTER(BlockVars))]


class FixupParams(C.Structure):
    _fields_ = [("enable_c2p_fixup", C.c_int), ("failure_strategy", C.c_int),
                ("enable_floors", C.c_int), ("enable_mhd_floors", C.c_int),
                ("bsqorho_max", C.c_double), ("bsqou_max", C.c_double)]


class MeshDesc(C.Structure):
    _fields_ = [("nx", C.c_int * 3), ("block_nx", C.c_int * 3), ("ng", C.c_int),
                ("xmin", C.c_double * 3), ("xmax", C.c_double * 3), ("bc", C.c_int * 6),
                ("flux_fixup", C.c_int), ("ox1_inflow_check", C.c_int), ("cfl", C.c_double),
                ("dt_init_fact", C.c_double)]


class PackBlock(C.Structure):
    _fields_ = [("block", Block), ("f", BlockFields), ("cons1", C.c_void_p),
                ("iters", C.c_void_p)]


class PackStageArgs(C.Structure):
    _fields_ = [("stage", C.c_int), ("dt", C.c_double), ("dt_dev", C.c_void_p),
                ("cfl", C.c_double), ("dt_out", C.c_void_p)]


_lib = None


class PhxError(RuntimeError):
    pass


def lib():
    """Load libphx.so once.  Raises if it is missing: the HIP path has no fallback."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIBPATH):
        raise PhxError(f"{LIBPATH} is not built; run __graft_entry__.build()")
    L = C.CDLL(LIBPATH)
    vp, i32, dbl = C.c_void_p, C.c_int, C.c_double
    P = C.POINTER
    sig = {
        "phx_create": (i32, [P(vp), i32]),
        "phx_destroy": (i32, [vp]),
        "phx_last_error": (C.c_char_p, [vp]),
        "phx_prim_to_con": (i32, [vp, P(FluidParams), P(Block), P(BlockFields), P(i32), P(i32), vp]),
        "phx_calculate_fluxes": (i32, [vp, P(FluidParams), P(Block), P(BlockFields), vp]),
        "phx_flux_ct": (i32, [vp, P(FluidParams), P(Block), P(BlockFields), vp]),
        "phx_fluid_source_terms": (i32, [vp, P(FluidParams), P(Block), P(BlockFields), vp]),
        "phx_flux_divergence_update": (i32, [vp, P(FluidParams), P(Block), vp, vp, vp, vp, dbl,
                                             dbl, vp, vp]),
        "phx_con2prim": (i32, [vp, P(FluidParams), P(Block), P(BlockFields), P(i32), P(i32), vp, vp]),
        "phx_estimate_dt": (i32, [vp, P(FluidParams), P(Block), P(BlockFields), i32, dbl, vp, vp]),
        "phx_fluid_stage": (i32, [vp, P(FluidParams), P(Block), P(BlockFields), P(StageArgs), vp]),
        "phx_fluid_stage_noalloc": (i32, [P(FluidParams), P(Block), P(BlockFields), P(StageArgs), vp]),
        "phx_con2prim_batch": (i32, [vp, P(FluidParams), C.c_int64, vp, vp, vp, vp, vp, vp, vp]),
        "phx_prim_to_con_batch": (i32, [vp, P(FluidParams), C.c_int64, vp, vp, vp, vp]),
        "phx_mlp_create": (i32, [vp, P(vp), C.c_uint64]),
        "phx_mlp_destroy": (i32, [vp]),
        "phx_mlp_set_params": (i32, [vp, vp, P(C.c_float)]),
        "phx_mlp_get_params": (i32, [vp, vp, P(C.c_float)]),
        "phx_mlp_set_normalization": (i32, [vp, vp, P(C.c_float), P(C.c_float), C.c_float, C.c_float]),
        "phx_con2prim_features": (i32, [vp, P(FluidParams), vp, C.c_int64, vp, vp, vp, vp]),
        "phx_mlp_forward": (i32, [vp, vp, C.c_int64, vp, vp, vp]),
        "phx_mlp_gradient": (i32, [vp, vp, C.c_int64, vp, vp, C.c_float, vp, vp]),
        "phx_mlp_read_gradient": (i32, [vp, vp, P(C.c_float)]),
        "phx_mlp_zero_gradient": (i32, [vp, vp]),
        "phx_mlp_adam_step": (i32, [vp, vp, C.c_float, vp]),
        "phx_mlp_train_step": (i32, [vp, vp, C.c_int64, vp, vp, C.c_float, C.c_float, vp, vp]),
        "phx_con2prim_batch_learned": (i32, [vp, P(FluidParams), vp, C.c_int64, vp, vp, vp, vp, vp,
                                             vp, vp, vp, vp]),
        "phx_c2p_fixup": (i32, [vp, P(FluidParams), P(FixupParams), P(Block), P(BlockFields), vp]),
        "phx_apply_floors": (i32, [vp, P(FluidParams), P(FixupParams), P(Block), P(BlockFields),
                                   P(i32), P(i32), vp]),
        "phx_autotune": (i32, [vp, P(FluidParams), P(Block), P(BlockFields), i32, vp]),
        "phx_tile_shape": (i32, [vp, P(FluidParams), P(Block), i32, P(i32), P(i32)]),
        "phx_set_tile_shape": (i32, [vp, P(FluidParams), P(Block), i32, i32, i32]),
        "phx_pack_create": (i32, [vp, P(FluidParams), i32, vp, P(vp)]),
        "phx_pack_destroy": (i32, [vp]),
        "phx_pack_nblocks": (i32, [vp, P(i32), P(C.c_int64)]),
        "phx_pack_stage": (i32, [vp, vp, P(PackStageArgs), vp]),
        "phx_pack_set_fixup": (i32, [vp, vp, P(FixupParams)]),
        "phx_pack_autotune": (i32, [vp, vp, i32, vp]),
        "phx_pack_tile_shape": (i32, [vp, i32, P(i32), P(i32)]),
        "phx_pack_graph_create": (i32, [vp, vp, i32, P(PackStageArgs), vp, P(vp)]),
        "phx_graph_launch": (i32, [vp, vp]),
        "phx_graph_destroy": (i32, [vp]),
        "phx_driver_create": (i32, [vp, P(FluidParams), P(Block), P(i32), dbl, dbl, P(vp)]),
        "phx_driver_destroy": (i32, [vp]),
        "phx_driver_set_prim": (i32, [vp, vp]),
        "phx_driver_set_fixup": (i32, [vp, P(FixupParams)]),
        "phx_driver_initialize": (i32, [vp]),
        "phx_driver_step": (i32, [vp]),
        "phx_driver_execute": (i32, [vp, dbl, i32]),
        "phx_driver_time": (i32, [vp, P(dbl), P(dbl), P(C.c_int64)]),
        "phx_driver_get": (i32, [vp, i32, vp]),
        "phx_mesh_create": (i32, [vp, P(FluidParams), P(FixupParams), P(MeshDesc), P(vp)]),
        "phx_mesh_destroy": (i32, [vp]),
        "phx_mesh_nblocks": (i32, [vp, P(i32), P(C.c_int64)]),
        "phx_mesh_block": (i32, [vp, i32, P(Block)]),
        "phx_mesh_set_prim": (i32, [vp, i32, vp]),
        "phx_mesh_initialize": (i32, [vp]),
        "phx_mesh_step": (i32, [vp]),
        "phx_mesh_execute": (i32, [vp, dbl, i32]),
        "phx_mesh_time": (i32, [vp, P(dbl), P(dbl), P(C.c_int64)]),
        "phx_mesh_get": (i32, [vp, i32, i32, vp]),
        "phx_mesh_get_iters": (i32, [vp, i32, vp]),
        "phx_mesh_autotune": (i32, [vp, i32]),
        "phx_mesh_set_learned": (i32, [vp, vp]),
        "phx_mesh_learned_stats": (i32, [vp, i32, P(C.c_int64), P(C.c_int64)]),
        "phx_pack_set_learned": (i32, [vp, vp, vp]),
        "phx_pack_learned_stats": (i32, [vp, i32, P(C.c_int64), P(C.c_int64)]),
        "phx_rt_device_count": (i32, [P(i32)]),
        "phx_rt_set_device": (i32, [i32]),
        "phx_rt_malloc": (i32, [P(vp), C.c_size_t]),
        "phx_rt_free": (i32, [vp]),
        "phx_rt_memcpy_htod": (i32, [vp, vp, C.c_size_t]),
        "phx_rt_memcpy_dtoh": (i32, [vp, vp, C.c_size_t]),
        "phx_rt_memcpy_dtod": (i32, [vp, vp, C.c_size_t]),
        "phx_rt_memcpy_dtod_async": (i32, [vp, vp, C.c_size_t, vp]),
        "phx_rt_memset": (i32, [vp, i32, C.c_size_t]),
        "phx_rt_device_synchronize": (i32, []),
        "phx_rt_stream_create": (i32, [P(vp)]),
        "phx_rt_stream_destroy": (i32, [vp]),
        "phx_rt_stream_synchronize": (i32, [vp]),
        "phx_rt_event_create": (i32, [P(vp)]),
        "phx_rt_event_destroy": (i32, [vp]),
        "phx_rt_event_record": (i32, [vp, vp]),
        "phx_rt_event_elapsed_ms": (i32, [vp, vp, P(C.c_float)]),
        "phx_rt_profile_start": (i32, [vp]),
        "phx_rt_profile_stop": (i32, [vp]),
        "phx_rt_profile_reset": (i32, [vp]),
        "phx_rt_profile_count": (i32, [vp, P(i32)]),
        "phx_rt_profile_get": (i32, [vp, i32, C.c_char_p, i32, P(C.c_float)]),
        "phx_rt_device_info": (i32, [C.c_char_p, i32, P(i32), P(dbl)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def exported_symbols():
    """Names every header in include/ declares (checked by the CPU tests)."""
    return sorted(n for n in dir(lib()) if n.startswith("phx_"))
