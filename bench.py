#!/usr/bin/env python3
"""Benchmark: zone-updates/s of one GRMHD RK stage on a 128^3 meshblock (MI355X, FP64).

Workload (single 3-D GRMHD meshblock, Minkowski, ideal gas Gamma = 13/9, 128^3 interior
+ 4 ghost zones, SoA FP64): one RK stage = CalculateFluxes (WENO5-Z + HLL, x/y/z) ->
FluxCT -> FluxDivergence + RK update -> ConservedToPrimitive (robust, tol 1e-10) on the
interior -> EstimateTimestep.  Synthetic seeded input: rho = exp N(0,1), P = rho U(0.01,1),
W ~ U(1,5) isotropic, B = centred-difference curl of a random vector potential under a
smooth log-uniform amplitude, so the magnetization b^2/rho is log-uniform over about
1e-3..10 (percentiles reported).  The stage runs at the input state's own CFL step
(cfl 0.4 over its cell signal speeds).  Every step repeats the stage on the same input,
with the con2prim guess taken from the input state's own root (the warm start a
time-stepping run sees), so the K steps do identical work.  The same block is also timed
with PLM and MP5 reconstruction (recon_legs).

Prints ONE JSON line.  --gpus N with N > 1 exits: this build measures one GPU.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

NX = 128
NG = 4
GAMMA = 13.0 / 9.0
SEED = 20261202
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6      # MI355X dense FP64 vector
VALU_SIMDS = 256 * 4         # CUs x SIMDs; one wave64 VALU instruction per 4 cycles each
CLOCK_HZ = 2.4e9             # peak engine clock
# compulsory FP64 state traffic of a fused stage per zone-update (reads U_s, U^n, 10 prims,
# mu, T guess = 28 doubles; writes U_{s+1}, 11 prims incl. T, mu, fail, 3 cell speeds = 24)
B_ALG_STAGE = 52 * 8

# the bracket network on torus cells (the uneven pack and the torus mesh): their centre
# errors differ by decades between the floored atmosphere and the disc, so the half-width
# needs the shared layers shaped by its own loss (weight 0.05; the batched states' default
# 1e-5 leaves them to the centre, and 61-67 % of the pack's cold cells then fall outside
# the bracket, against 4-25 %)
TORUS_BRACKET_TRAINING = {"lam": 0.05, "width_scale": 4.0}


# the sources the single-block stage's kernels are compiled from (flux sweeps, FluxCT,
# fused update + con2prim, dt, p2c) with every header they include
STAGE_SOURCES = ["phoebus_amd/csrc/" + f for f in (
    "phx_kernels.hip", "phx_flux.hip", "phx_flux_fast.hip", "phx_flux_weno.hip",
    "phx_flux_fast_weno.hip", "phx_stage.hip", "phx_device.hpp",
    "phx_internal.hpp", "phx_cells.hpp", "phx_flux_shapes.hpp", "phx_flux_sweeps.hpp")] + [
    "include/phx.h", "include/phx_fixup.h"]


def sources_digest():
    """sha256 over the stage kernels' sources and headers (the build a stage profile
    measured)."""
    import hashlib

    h = hashlib.sha256()
    files = sorted(os.path.join(ROOT, f) for f in STAGE_SOURCES)
    for f in files:
        h.update(os.path.basename(f).encode())
        h.update(open(f, "rb").read())
    return h.hexdigest()


def profile_stale(prof):
    """True when a committed profile was taken on other kernel sources than these (its
    "sources_sha256" differs), None when it does not record them."""
    d = prof.get("sources_sha256")
    return None if d is None else d != sources_digest()


def kernel_bytes(nx):
    """Algorithmic bytes per launch (each input read once, each output written once)."""
    cells = nx ** 3
    faces = (nx + 2) ** 3  # per direction (the reference's face ranges)
    flux_dir = faces * (10 + 8 + 1) * 8  # 10 primitives in, 8 fluxes + 1 speed out
    return {
        "flux_x": flux_dir, "flux_y": flux_dir, "flux_z": flux_dir,
        # FluxCT's EMF pass: 6 transverse B fluxes in, 3 corner EMFs out
        "flux_ct": (nx + 1) ** 3 * (6 + 3) * 8,
        # 18 face fluxes (the B fluxes come from the EMFs), 3 EMFs, U_s, U^n, mu, 3 face
        # speeds in; U_s, U_{s+1}, 10 prims, T, mu, fail, 3 cell speeds, int32 iteration
        # count out
        "update_con2prim": cells * ((18 + 3 + 8 + 8 + 1 + 3) + (8 + 8 + 10 + 1 + 1 + 1 + 3)) * 8
                           + cells * 4,
        "dt_finish": 8 * ((cells + 255) // 256),
    }


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def native_oracle(log):
    """The oracle rebuilt for this host: -O3 -march=native -ffp-contract=off -fopenmp (the
    CPU baseline's build); None when no C compiler is usable here (the caller then times
    the -O2 build the tests use)."""
    import shutil
    import tempfile

    cc = shutil.which("gcc") or shutil.which("cc")
    if not cc:
        return None
    d = tempfile.mkdtemp(prefix="phx_oracle_native_")
    so = os.path.join(d, "libphx_oracle_native.so")
    cmd = [cc, "-O3", "-march=native", "-ffp-contract=off", "-fno-fast-math", "-fopenmp",
           "-fPIC", "-std=c99", "-shared", "-o", so,
           os.path.join(ROOT, "oracle", "phx_oracle.c"), "-lm"]
    try:
        import subprocess
        subprocess.run(cmd, check=True, capture_output=True, timeout=120)
    except Exception as e:  # noqa: BLE001 -- reported, the portable build is timed instead
        log(f"cpu: native oracle build failed ({e}); timing the portable build")
        return None
    return so


def cpu_baseline(prim, cons, mu0, dt, args, log=lambda m: None):
    """The CPU oracle's stage on the same input and dt, built -O3 -march=native
    -ffp-contract=off on this host: first touch, 3 untimed stages, then the median of
    cpu_stages (>= 10) timed stages, on all host threads (OMP_PROC_BIND=close) over an
    --cpu-nx^3 corner of the block and on one thread over a --cpu-nx1^3 corner (bounded
    samples of the same workload)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import _oracle as O

    so = native_oracle(log)
    if so:
        O.LIBPATH = so
        O._lib = None  # an earlier leg may have loaded the portable build
    O.lib()
    orecon = {"weno5z": O.RECON_WENO5, "plm": O.RECON_LINEAR, "mp5": O.RECON_MP5,
              "constant": O.RECON_CONSTANT}[args.recon]
    oriem = {"hll": O.RIEMANN_HLL, "llf": O.RIEMANN_LLF}[args.riemann]
    op = O.default_params(gamma=GAMMA, mhd=True, recon=orecon, riemann=oriem,
                          tol=1e-10, max_iter=100)
    all_threads = O.threads()

    def run(n, threads):
        O.set_threads(threads)
        sl = slice(0, n + 2 * NG)
        ob = O.make_block(3, [n, n, n], NG, dx=(1.0 / NX,) * 3)
        p = np.ascontiguousarray(prim[:, sl, sl, sl])
        u = np.ascontiguousarray(cons[:, sl, sl, sl])
        m0 = np.ascontiguousarray(mu0[sl, sl, sl])
        lo = (NG, NG, NG)
        hi = (NG + n - 1,) * 3
        times = []
        for it in range(3 + args.cpu_stages):
            t0 = time.perf_counter()
            flux, fsig = O.calculate_fluxes(op, ob, p)
            O.flux_ct(op, ob, flux)
            out = O.flux_divergence_update(op, ob, flux, None, u.copy(), u, 1.0, dt)
            pr = p.copy()
            mu = m0.copy()
            r = O.con2prim_block(op, ob, out, pr, mu, lo=lo, hi=hi)
            O.estimate_dt(op, ob, r["sig"], fsig, 0.4)
            if it >= 3:
                times.append(time.perf_counter() - t0)
        return float(np.median(times))

    t_all = run(args.cpu_nx, all_threads)
    t_one = run(args.cpu_nx1, 1)
    O.set_threads(all_threads)
    return {"value": args.cpu_nx ** 3 / t_all, "unit": "zone-updates/s", "cores": all_threads,
            "kind": "port",
            "cpu_model": cpu_model(),
            "build": ("gcc -O3 -march=native -ffp-contract=off -fopenmp" if so else
                      "gcc -O2 -ffp-contract=off -fopenmp (portable build; native build failed)"),
            "one_thread": {"value": args.cpu_nx1 ** 3 / t_one, "unit": "zone-updates/s",
                           "cores": 1, "median_stage_s": t_one, "zones": args.cpu_nx1 ** 3},
            "sample": f"{args.cpu_nx}^3 corner of the same block and dt, C99 restatement of the "
                      f"stage (fluxes, CT, update, con2prim, dt) on {all_threads} threads "
                      f"(OMP_PROC_BIND=close): median of {args.cpu_stages} stages after 3 "
                      f"warm-ups, {t_all:.3f} s; one thread on a {args.cpu_nx1}^3 corner, "
                      f"median {t_one:.3f} s"}


def bracket_centre_stats(net, feat, out, mu, ok, iters, fallback):
    """The network's bracket against the plain root on one chunk of states: percentiles of
    the relative centre error |c / mu - 1| and of the relative half-width, the share of
    roots outside the bracket, and the learned solve's mean iterations per decade of
    centre error (fallback solves included)."""
    mean, std = net.norm
    ws, wmin = net.width
    f0 = feat[0].astype(np.float64) * std[0] + mean[0]
    f2 = feat[2].astype(np.float64) * std[2] + mean[2]
    z0 = np.maximum(1.0 + (np.exp(f0) - 1e-12) - 0.5 * (np.exp(f2) - 1e-12), 1.0)
    c = np.exp(out[0].astype(np.float64)) / z0
    half = np.maximum(ws * np.exp(out[1].astype(np.float64)), wmin)
    sel = ok & (mu > 0.0)
    err = np.abs(c[sel] / mu[sel] - 1.0)
    outside = np.abs(c[sel] - mu[sel]) > half[sel] * c[sel]
    half, it, fb = half[sel], iters[sel], fallback[sel]
    q = [50, 90, 99]
    decades = {}
    for lo in range(-8, 1):
        m = (err >= 10.0 ** (lo - 1)) & (err < 10.0 ** lo) if lo > -8 else err < 10.0 ** lo
        if np.any(m):
            decades[f"1e{lo - 1}..1e{lo}" if lo > -8 else f"<1e{lo}"] = {
                "share": float(np.mean(m)), "mean_iters": float(np.mean(it[m])),
                "fallback": float(np.mean(fb[m]))}
    return {"states": int(sel.sum()),
            "centre_rel_err_p50_p90_p99": [float(x) for x in np.percentile(err, q)],
            "half_width_rel_p50_p90_p99": [float(x) for x in np.percentile(half, q)],
            "root_outside_bracket": float(np.mean(outside)),
            "iters_by_centre_err": decades}


def c3_learned_vs_plain(phx, args, log):
    """Batched con2prim on spherical Kerr-Schild (a = 0.9375, Gamma = 4/3, tol 1e-10): the
    plain solver and the learned-bracket solver on the same 2^26 states (seed SEED, in
    chunks), after training the network on the device from 2^22 plain-solver pairs of a
    disjoint seed.  Reports mean iterations, histograms, fallback rate, the largest
    difference of the two roots, states/s of each, and the CPU oracle's root on a sample."""
    from phoebus_amd import _abi as A
    from phoebus_amd.api import fluid_params
    from phoebus_amd.device import DeviceArray, Event, synchronize
    from phoebus_amd.learned import train_bracket
    from phoebus_amd.workloads import batched_con2prim_states

    p = fluid_params(gamma=4.0 / 3.0, mhd=True, geom=A.GEOM_SPHERICAL_KS, a=0.9375,
                     c2p_tol=1e-10, c2p_max_iter=100)
    loaded = bool(args.c3_net) and os.path.exists(args.c3_net)
    if loaded:
        # a network an earlier run of the same configuration trained and saved (profiling
        # passes reuse it instead of repeating the training launches)
        from phoebus_amd.api import Mlp

        z = np.load(args.c3_net)
        net = Mlp(phx)
        net.set_params(z["params"])
        net.set_normalization(z["mean"], z["std"], float(z["width"][0]), float(z["width"][1]))
        net.norm = (z["mean"], z["std"])
        losses, t_train = [float(x) for x in z["losses"]], float(z["seconds"])
        train_steps, train_pairs = int(z["steps"]), int(z["pairs"])
        log(f"c3: network of {args.c3_net} (trained {train_steps} steps in {t_train:.1f} s)")
    else:
        t0 = time.perf_counter()
        train_steps, train_pairs = args.c3_train_steps, args.c3_pairs
        net, losses = train_bracket(phx, p, np.random.default_rng(SEED + 1), train_pairs,
                                    train_steps)
        t_train = time.perf_counter() - t0
        log(f"c3: trained {args.c3_train_steps} steps on {args.c3_pairs} pairs in "
            f"{t_train:.1f} s")
        if args.c3_net:
            np.savez(args.c3_net, params=net.get_params(), mean=net.norm[0], std=net.norm[1],
                     width=np.array(net.width), losses=np.array(losses), seconds=t_train,
                     steps=train_steps, pairs=train_pairs)
    rng = np.random.default_rng(SEED)
    chunk = min(args.c3_states, 1 << 22)
    n_total = (args.c3_states // chunk) * chunk
    hp = np.zeros(128, np.int64)
    hl = np.zeros(128, np.int64)
    fb_count = ok_count = 0
    ms_plain = ms_learned = 0.0
    max_dmu = 0.0
    centre = None
    fail_mismatch = 0
    sample = None
    d_coords = DeviceArray((2, chunk))
    d_prim = DeviceArray((10, chunk))
    d_cons = DeviceArray((8, chunk))
    d_cons2 = DeviceArray((8, chunk))
    d_out = DeviceArray((10, chunk))
    d_mu = DeviceArray((chunk,))
    d_mu2 = DeviceArray((chunk,))
    d_fail = DeviceArray((chunk,))
    d_fail2 = DeviceArray((chunk,))
    d_it = DeviceArray((chunk,), np.int32)
    d_it2 = DeviceArray((chunk,), np.int32)
    d_fb = DeviceArray((chunk,), np.int32)
    e0, e1, e2 = Event(), Event(), Event()
    for c0 in range(0, n_total, chunk):
        coords, prim = batched_con2prim_states(rng, chunk)
        d_coords.upload(coords)
        d_prim.upload(prim)
        phx.prim_to_con_batch(p, chunk, d_coords, d_prim, d_cons)
        d_cons2.fill_from(d_cons)
        if sample is None:
            sample = (coords[:, :2048].copy(), d_cons.get()[:, :2048].copy())
        d_mu.zero()
        d_mu2.zero()
        if c0 == 0:  # the network's bracket for the first chunk, before any solve
            d_feat = DeviceArray((4, chunk), np.float32)
            d_o = DeviceArray((2, chunk), np.float32)
            net.features(p, chunk, d_coords, d_cons, d_feat)
            net.forward(chunk, d_feat, d_o)
        synchronize()
        e0.record()
        phx.con2prim_batch(p, chunk, d_coords, d_cons, d_out, d_mu, d_fail, d_it)
        e1.record()
        net.con2prim_batch(p, chunk, d_coords, d_cons2, d_out, d_mu2, d_fail2, d_it2, None, d_fb)
        e2.record()
        synchronize()
        ms_plain += e0.elapsed_ms(e1)
        ms_learned += e1.elapsed_ms(e2)
        fail, fail2 = d_fail.get(), d_fail2.get()
        ok = fail == 1.0
        fail_mismatch += int(np.sum(fail != fail2))
        ip, il = d_it.get()[ok], d_it2.get()[ok]
        hp += np.bincount(np.clip(ip, 0, 127), minlength=128)
        hl += np.bincount(np.clip(il, 0, 127), minlength=128)
        fb_count += int(np.sum(d_fb.get()[ok]))
        ok_count += int(np.sum(ok))
        max_dmu = max(max_dmu, float(np.max(np.abs(d_mu.get()[ok] - d_mu2.get()[ok]))))
        if c0 == 0:
            centre = bracket_centre_stats(net, d_feat.get(), d_o.get(), d_mu.get(), ok,
                                          d_it2.get(), d_fb.get())
            del d_feat, d_o
    if sample is not None:
        mean_p = float(np.dot(np.arange(128), hp) / max(ok_count, 1))
    mean_l = float(np.dot(np.arange(128), hl) / max(ok_count, 1))
    out = {"workload": f"batched con2prim, {n_total} states, spherical KS a=0.9375, "
                       "Gamma=4/3, tol 1e-10, cold start",
           "mean_iters_plain": mean_p, "mean_iters_learned": mean_l,
           "iters_ratio": mean_p / mean_l if mean_l > 0 else None,
           "fallback_rate": fb_count / max(ok_count, 1),
           "max_abs_dmu_learned_vs_plain": max_dmu,
           "fail_flag_mismatches": fail_mismatch,
           "states_per_s_plain": n_total / (ms_plain * 1e-3),
           "states_per_s_learned": n_total / (ms_learned * 1e-3),
           "hist_plain": np.trim_zeros(hp, "b").tolist(),
           "hist_learned": np.trim_zeros(hl, "b").tolist(),
           "train": {"pairs": train_pairs, "steps": train_steps,
                     "seconds": t_train, "loss_first_last": [losses[0], losses[-1]],
                     "loaded_from": args.c3_net if loaded else None},
           "bracket_first_chunk": centre}
    if not args.no_cpu and sample is not None:
        # FP64 oracle on a sample of the first chunk: its root against the learned one
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import _oracle as O
        op = O.default_params(gamma=4.0 / 3.0, mhd=True, geom=O.GEOM_SPHKS, a=0.9375,
                              tol=1e-10, max_iter=100)
        coords_s, cons_s = sample
        ref = O.c2p_batch(op, np.ascontiguousarray(coords_s), np.ascontiguousarray(cons_s))
        d_c = DeviceArray.from_host(coords_s)
        d_u = DeviceArray.from_host(cons_s)
        m = coords_s.shape[1]
        d_o, d_m, d_f = DeviceArray((10, m)), DeviceArray.from_host(np.zeros(m)), DeviceArray((m,))
        net.con2prim_batch(p, m, d_c, d_u, d_o, d_m, d_f)
        synchronize()
        okm = ref["fail"] == 1.0
        out["oracle_sample"] = {"states": m,
                                "max_abs_dmu_vs_oracle": float(np.max(np.abs(
                                    d_m.get()[okm] - ref["mu"][okm]))),
                                "fail_flags_equal": bool(np.array_equal(d_f.get(),
                                                                        ref["fail"]))}
    net.close()
    return out


class _ArenaBlock:
    """Device fields of one pack block as views into shared arenas."""

    def __init__(self, arenas, off, sh):
        n = int(np.prod(sh))
        self.shape = sh
        self.views = {}
        for name, (arr, nv) in arenas.items():
            self.views[name] = arr.view(off * nv, (nv,) + sh if nv > 1 else sh)

    def __getattr__(self, k):
        v = self.__dict__.get("views", {})
        if k in v:
            return v[k]
        raise AttributeError(k)

    def fields(self, with_src=False):
        from phoebus_amd.api import Phx
        v = self.views
        return Phx.fields(prim=v["prim"], cons=v["cons"], temp=v["temp"], c2p_mu=v["mu"],
                          fail=v["fail"], cell_sig=v["csig"], face_sig=v["fsig"],
                          flux=v["flux"], emf=v["emf"], src=v["src"] if with_src else None)


def c4_uneven_pack(phx, args, log):
    """Uneven batch: 512 blocks of {16^3: 50 %, 32^3: 35 %, 64^3: 15 %} cut from a
    magnetized Fishbone-Moncrief torus (r_in 6, r_max 12, beta 100, 4 % u jitter) on an
    FMKS grid (a = 0.9375, 256 x 128 x 128 spacing); WENO5-Z + HLL + CT + sources +
    update + con2prim.  Times full RK2 steps (two stages, device-resident CFL dt) as one
    HIP graph, as eager pack launches, and as one phx_fluid_stage per block."""
    from phoebus_amd import _abi as A
    from phoebus_amd.api import Pack, block, fluid_params, mgr
    from phoebus_amd.device import DeviceArray, Event, synchronize
    from phoebus_amd.workloads import torus_grid, uneven_blocks

    rng = np.random.default_rng(SEED + 4)
    grid_prim, xmin, dx = torus_grid(rng)
    blocks = uneven_blocks(rng, args.c4_blocks)
    x1min = xmin[0]
    # floors and ceilings of the torus deck (inputs/torus.pin: ExpX1RhoSie rho 1e-8 r^-2,
    # sie 1e-3 r^-1, EOS minima 1e-8, gamma <= 10, sie <= 1)
    floors = dict(floor_type=A.FLOOR_EXPX1_RHO_SIE, rho0=1e-8, sie0=1e-3, rho_exp=-2.0,
                  sie_exp=-1.0, rho_min_eos=1e-8, sie_min_eos=1e-8)
    p = fluid_params(gamma=5.0 / 3.0, mhd=True, recon=A.RECON_WENO5Z, riemann=A.RIEMANN_HLL,
                     geom=A.GEOM_FMKS, a=0.9375, c2p_tol=1e-10, c2p_max_iter=100,
                     floors=floors, ceilings=(10.0, 1.0),
                     flux_arith={"exact": A.FLUX_ARITH_EXACT,
                                 "fast": A.FLUX_ARITH_FAST}[args.flux_arith],
                     mgr_params=mgr(derefine_poles=False, x0=min(x1min, max(0.0, x1min - NG * dx[0]))))
    shapes = [(n + 2 * NG,) * 3 for n, _ in blocks]
    cells = [int(np.prod(s)) for s in shapes]
    offs = np.concatenate([[0], np.cumsum(cells)])
    total = int(offs[-1])
    per = {"prim": 10, "cons": 8, "cons1": 8, "temp": 1, "mu": 1, "fail": 1, "csig": 3,
           "fsig": 3, "flux": 24, "emf": 3, "src": 8}
    arenas = {k: (DeviceArray((nv * total,)), nv) for k, nv in per.items()}
    it_arena = DeviceArray((total,), np.int32)
    for k in ("temp", "mu", "fail", "csig", "fsig", "flux", "emf", "src"):
        arenas[k][0].zero()
    mu0 = DeviceArray((total,))
    entries, interior = [], 0
    for bidx, (n, o) in enumerate(blocks):
        sh = shapes[bidx]
        ab = _ArenaBlock(arenas, int(offs[bidx]), sh)
        b = block(3, (n, n, n), NG, xmin=tuple(xmin[d] + o[d] * dx[d] for d in range(3)), dx=dx)
        win = np.ascontiguousarray(grid_prim[:, o[2]:o[2] + n + 2 * NG, o[1]:o[1] + n + 2 * NG,
                                             o[0]:o[0] + n + 2 * NG])
        ab.prim.upload(win)
        phx.prim_to_con(p, b, ab.fields())
        it = it_arena.view(int(offs[bidx]), sh)
        phx.con2prim(p, b, ab.fields(), iters=it)  # the warm start: the input's own roots
        entries.append((b, ab, ab.cons1, it))
        interior += n ** 3
    synchronize()
    del grid_prim

    def interior_fails():
        fh = arenas["fail"][0].get()
        out = 0
        for i in range(len(blocks)):
            f = fh[int(offs[i]):int(offs[i + 1])].reshape(shapes[i])[NG:-NG, NG:-NG, NG:-NG]
            out += int(np.sum(f == 0.0))
        return out

    fails_setup = interior_fails()
    # snapshots to restart every timed variant from the same state
    snap = {k: DeviceArray(arenas[k][0].shape) for k in ("prim", "cons", "mu")}
    for k in snap:
        snap[k].fill_from(arenas[k][0])

    def restore():
        for k in snap:
            arenas[k][0].fill_from(snap[k])
        synchronize()

    pk = Pack(phx, p, entries)
    sp = C.c_void_p()
    phx.L.phx_rt_stream_create(C.byref(sp))
    s = sp.value
    pk.autotune(reps=2, stream=s)
    phx.L.phx_rt_stream_synchronize(s)
    tiles = [pk.tile_shape(d) for d in (1, 2)]
    dt_dev = DeviceArray((1,))
    # the first dt: the CFL estimate of a stage-1 pass
    restore()
    dt_dev.upload(np.array([1e-4]))
    pk.stage(1, dt_dev=dt_dev, cfl=0.4, dt_out=dt_dev, stream=s)
    phx.L.phx_rt_stream_synchronize(s)
    dt0 = dt_dev.get().copy()
    steps = args.c4_steps
    e0, e1 = Event(), Event()

    def timed(run):
        restore()
        dt_dev.upload(dt0)
        run()  # warm-up step
        phx.L.phx_rt_stream_synchronize(s)
        restore()
        dt_dev.upload(dt0)
        e0.record(s)
        for _ in range(steps):
            run()
        e1.record(s)
        phx.L.phx_rt_stream_synchronize(s)
        return e0.elapsed_ms(e1) / steps

    stages = [Pack.args(1, dt_dev=dt_dev, cfl=0.4, dt_out=None),
              Pack.args(2, dt_dev=dt_dev, cfl=0.4, dt_out=dt_dev)]
    g = pk.graph(stages, s)
    ms_graph = timed(lambda: g.launch(s))
    fail_graph = interior_fails()

    def interior_iters():
        it_h = it_arena.get()
        return np.concatenate([it_h[int(offs[i]):int(offs[i + 1])].reshape(shapes[i])[NG:-NG, NG:-NG, NG:-NG].ravel()
                               for i in range(len(blocks))])

    its = interior_iters()  # those of the graph steps (the learned run repeats them)

    def eager():
        pk.stage(1, dt_dev=dt_dev, cfl=0.4, stream=s)
        pk.stage(2, dt_dev=dt_dev, cfl=0.4, dt_out=dt_dev, stream=s)

    ms_eager = timed(eager)
    d_one = DeviceArray((1,))

    def per_block():
        # one phx_fluid_stage per block and stage (7 launches each), host dt
        for stage in (1, 2):
            for b, ab, c1, it in entries:
                f = ab.fields(with_src=True)
                if stage == 1:
                    phx.fluid_stage(p, b, f, 1.0, float(dt0[0]), ab.cons, c1, ab.prim,
                                    iters=it, cfl=0.4, dt_out=None, stream=s)
                else:
                    f.cons = c1.ptr
                    phx.fluid_stage(p, b, f, 0.5, float(dt0[0]), ab.cons, ab.cons, ab.prim,
                                    iters=it, cfl=0.4, dt_out=None, stream=s)

    ms_blocks = timed(per_block) if args.c4_per_block else None
    g.close()
    learned = None
    if not args.no_c4_learned:
        # the learned bracket: pairs (conserved state, plain root, (X1, X2)) from the
        # interiors of every 4th block after the plain steps, the network trained on them
        # on the device, then the same graph steps with every inversion through it
        from phoebus_amd.learned import train_bracket_on
        coords, cons, mu, fl = [], [], [], []
        for bidx in range(0, len(blocks), 4):
            b, ab, _, _ = entries[bidx]
            n = b.nx[0]
            I = (slice(NG, NG + n),) * 3
            cons.append(ab.cons.get()[(slice(None),) + I].reshape(8, -1))
            mu.append(ab.mu.get()[I].reshape(-1))
            fl.append(ab.fail.get()[I].reshape(-1))
            x1 = b.xmin[0] + (np.arange(n) + 0.5) * b.dx[0]
            x2 = b.xmin[1] + (np.arange(n) + 0.5) * b.dx[1]
            coords.append(np.stack([np.broadcast_to(x1[None, None, :], (n, n, n)).reshape(-1),
                                    np.broadcast_to(x2[None, :, None], (n, n, n)).reshape(-1)]))
        data = {"coords": np.ascontiguousarray(np.concatenate(coords, axis=1)),
                "cons": np.ascontiguousarray(np.concatenate(cons, axis=1)),
                "mu": np.concatenate(mu), "fail": np.concatenate(fl)}
        if data["mu"].size > args.c3_pairs:
            sel = np.random.default_rng(SEED + 6).choice(data["mu"].size, args.c3_pairs, replace=False)
            data = {k: np.ascontiguousarray(v[..., sel]) for k, v in data.items()}
        data["d_coords"] = DeviceArray.from_host(data["coords"])
        t0 = time.perf_counter()
        net, losses = train_bracket_on(phx, p, np.random.default_rng(SEED + 7), data, steps=6000,
                                       **TORUS_BRACKET_TRAINING)
        train_s = time.perf_counter() - t0
        pk.set_learned(net)
        gl = pk.graph(stages, s)
        pk.learned_stats(reset=True)
        ms_learned = timed(lambda: gl.launch(s))
        misses, solves = pk.learned_stats()
        its_l = interior_iters()
        gl.close()

        def cold_stage(with_net):
            # stage 1 from the same state with no previous roots (mu = 0, as after the
            # problem generator): the plain cold solve or the network's bracket; the second
            # of two runs is timed
            pk.set_learned(net if with_net else None)
            for _ in range(2):
                restore()
                arenas["mu"][0].zero()
                dt_dev.upload(dt0)
                pk.learned_stats(reset=True)
                e0.record(s)
                pk.stage(1, dt_dev=dt_dev, cfl=0.4, stream=s)
                e1.record(s)
                phx.L.phx_rt_stream_synchronize(s)
            cm, cs_ = pk.learned_stats()
            ci = interior_iters()
            return {"ms": e0.elapsed_ms(e1), "con2prim_mean_iters": float(np.mean(np.maximum(ci, 0))),
                    "network_solves": cs_, "bracket_misses": cm}

        cold = {"plain": cold_stage(False), "learned": cold_stage(True)}
        pk.set_learned(None)
        learned = {"pairs": int(data["mu"].size), "train_s": train_s,
                   "loss_first_last": [losses[0], losses[-1]],
                   "ms_per_rk2_step_graph": ms_learned,
                   "zone_updates_per_s_graph": 2 * interior / (ms_learned * 1e-3),
                   "con2prim_mean_iters": float(np.mean(np.maximum(its_l, 0))),
                   "con2prim_hist": np.bincount(np.maximum(its_l, 0)).tolist(),
                   # every cell of the timed steps has a previous root: the update inverts
                   # it from there as the plain stage does and the network brackets nothing
                   "network_solves": solves, "bracket_misses": misses,
                   "cold_first_stage": cold}
        log(f"c4: learned graph {ms_learned:.3f} ms/step, iters {learned['con2prim_mean_iters']:.2f}"
            f" vs plain {float(np.mean(np.maximum(its, 0))):.2f}, misses {misses}/{solves}; "
            f"cold stage 1: plain {cold['plain']['ms']:.3f} ms {cold['plain']['con2prim_mean_iters']:.2f} "
            f"iters, learned {cold['learned']['ms']:.3f} ms {cold['learned']['con2prim_mean_iters']:.2f} iters")
    phx.L.phx_rt_stream_destroy(s)
    pk.close()
    zu = 2 * interior
    counts = {str(n): int(sum(1 for m, _ in blocks if m == n)) for n in (16, 32, 64)}
    out = {"workload": f"{len(blocks)} FMKS blocks {counts} (interior {interior} zones) of a "
                       "magnetized FM torus, a=0.9375, Gamma=5/3, torus-deck floors/ceilings, "
                       "WENO5-Z + HLL + CT + sources + update + con2prim (tol 1e-10), full RK2 "
                       "steps, device-resident dt",
           "zone_updates_per_s_graph": zu / (ms_graph * 1e-3),
           "zone_updates_per_s_eager_pack": zu / (ms_eager * 1e-3),
           "zone_updates_per_s_per_block_launches": zu / (ms_blocks * 1e-3) if ms_blocks else None,
           "ms_per_rk2_step_graph": ms_graph, "ms_per_rk2_step_eager_pack": ms_eager,
           "ms_per_rk2_step_per_block_launches": ms_blocks,
           "hbm_frac_stage_bytes": zu / (ms_graph * 1e-3) * B_ALG_STAGE / 1e9 / HBM_PEAK_GBS,
           "flux_tiles_yz": tiles, "con2prim_mean_iters": float(np.mean(np.maximum(its, 0))),
           "con2prim_hist": np.bincount(np.maximum(its, 0)).tolist(),
           "learned": learned,
           "con2prim_fail_zones_interior": {"setup": fails_setup, f"after_{steps + 1}_steps": fail_graph},
           "steps": steps}
    log(f"c4: graph {ms_graph:.3f} ms/step, eager {ms_eager:.3f}, per-block {ms_blocks}")
    return out


def c5_cpu_totals(p, fx, N, bnx, xmin, xmax, bc, grid_prim, tot0, tot_gpu, log):
    """The oracle's mesh driver (the CPU restatement of the same RK2 cycle, all host
    threads) from the same initial data for len(tot_gpu) steps: the conserved totals
    (history reducers' sums of D, S_i, tau over the interior, fluid.cpp:319-349) after each
    step against the GPU run's, as relative differences."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import _oracle as O

    m = p.geom.mgr
    op = O.default_params(gamma=p.gamma, mhd=True, recon=O.RECON_WENO5, riemann=O.RIEMANN_HLL,
                          geom=O.GEOM_FMKS, a=p.geom.a, tol=p.c2p_tol, max_iter=p.c2p_max_iter,
                          mgr_params=O.mgr(derefine_poles=bool(m.derefine_poles), h=m.h,
                                           xt=m.xt, alpha=m.alpha, x0=m.x0, smooth=m.smooth))
    b = p.bounds
    op.bounds.floor_type = b.floor_type
    op.bounds.r0, op.bounds.s0, op.bounds.ralpha, op.bounds.salpha = (b.rho0, b.sie0,
                                                                      b.rho_exp, b.sie_exp)
    op.bounds.rho_min_eos, op.bounds.sie_min_eos = b.rho_min_eos, b.sie_min_eos
    op.bounds.gam_max, op.bounds.sie_max = b.gam_max, b.sie_max
    ofx = O.Fixup(fx.enable_c2p_fixup, fx.failure_strategy, fx.enable_floors,
                  fx.enable_mhd_floors, fx.bsqorho_max, fx.bsqou_max)
    om = O.MeshDriver(op, N, bnx, xmin, xmax, bc, fx=ofx, inflow_check=True, cfl=0.8,
                      dt_init_fact=1e-6)
    for g, lx in enumerate(om.loc):
        sl = tuple(slice(lx[d] * bnx[d], lx[d] * bnx[d] + bnx[d] + 2 * NG) for d in (2, 1, 0))
        om.set_prim(g, np.ascontiguousarray(grid_prim[(slice(None),) + sl]))
    om.initialize()

    def totals():
        t = np.zeros(5)
        for g in range(len(om.blocks)):
            t += om.cons[g][0][:5, NG:-NG, NG:-NG, NG:-NG].reshape(5, -1).sum(axis=1)
        return t

    c0 = totals()
    rel = lambda a, b_: ((a - b_) / np.maximum(np.abs(b_), 1e-300)).tolist()
    steps = []
    t0 = time.perf_counter()
    for s in range(len(tot_gpu)):
        om.step()
        c = totals()
        steps.append({"step": s + 1, "cpu_totals": c.tolist(),
                      "gpu_minus_cpu_rel": rel(tot_gpu[s], c),
                      "drift_cpu": (c - c0).tolist(), "drift_gpu": (tot_gpu[s] - tot0).tolist()})
        log(f"[c5] cpu oracle step {s + 1}: max rel |gpu - cpu| of the totals "
            f"{np.max(np.abs(rel(tot_gpu[s], c))):.2e}")
    return {"initial_rel": rel(tot0, c0), "steps": steps,
            "cpu_s": time.perf_counter() - t0, "threads": O.threads(),
            "what": "oracle mesh driver (C99 restatement, OpenMP) from the same initial data; "
                    "totals of D, S_1..3, tau over the interior after each step"}


def c5_torus_mesh(phx, args, log):
    """Full 3-D GRMHD torus: the magnetized Fishbone-Moncrief torus of the uneven-batch leg
    on one 256 x 128 x 128 FMKS grid cut into 32 meshblocks of 64 x 64 x 32, advanced
    args.c5_steps RK2 steps by the mesh driver: WENO5-Z + HLL + CT + sources + FixFluxes +
    update + con2prim + c2p fixup + floors (torus deck floors and ceilings), ghost exchange,
    gr_outflow boundaries in x1 and x2 with the FMKS inflow check, periodic x3; cfl 0.8 and
    dt_init_fact 1e-6 as inputs/torus.pin.  Two runs from the same initial data: plain
    (con2prim warm-started from the previous root, as the reference), and learned: plain
    for the first args.c5_log_steps steps while logging solver pairs from the blocks, then
    the bracket network trained on them on the device and every later con2prim with the
    learned bracket.  Reports zone-updates/s (all steps, and the steps after the switch for
    both runs), per-launch times of one step of each, iteration histograms of the last
    step, the learned bracket's miss rate, and the drift of the conserved totals."""
    from phoebus_amd import _abi as A
    from phoebus_amd.api import Mesh, fixup_params, fluid_params, mgr
    from phoebus_amd.learned import mesh_pairs, train_bracket_on
    from phoebus_amd.device import DeviceArray
    from phoebus_amd.workloads import torus_grid

    N, bnx = (256, 128, 128), (64, 64, 32)
    rng = np.random.default_rng(SEED + 4)
    grid_prim, xmin, dx = torus_grid(rng, nx=N)
    xmax = tuple(xmin[d] + N[d] * dx[d] for d in range(3))
    x1min = xmin[0]
    floors = dict(floor_type=A.FLOOR_EXPX1_RHO_SIE, rho0=1e-8, sie0=1e-3, rho_exp=-2.0,
                  sie_exp=-1.0, rho_min_eos=1e-8, sie_min_eos=1e-8)
    p = fluid_params(gamma=5.0 / 3.0, mhd=True, recon=A.RECON_WENO5Z, riemann=A.RIEMANN_HLL,
                     geom=A.GEOM_FMKS, a=0.9375, c2p_tol=1e-10, c2p_max_iter=100,
                     floors=floors, ceilings=(10.0, 1.0),
                     flux_arith={"exact": A.FLUX_ARITH_EXACT,
                                 "fast": A.FLUX_ARITH_FAST}[args.flux_arith],
                     mgr_params=mgr(derefine_poles=False, x0=min(x1min, max(0.0, x1min - NG * dx[0]))))
    fx = fixup_params(c2p_fixup=True, mhd_floors=True)
    bc = (A.BC_OUTFLOW,) * 4 + (A.BC_PERIODIC,) * 2

    def make():
        m = Mesh(phx, p, N, bnx, xmin, xmax, bc, fx=fx, inflow_check=True, cfl=0.8,
                 dt_init_fact=1e-6)
        for g, b in enumerate(m.blocks):
            lx = [int(round((b.xmin[d] - xmin[d]) / dx[d])) for d in range(3)]
            win = grid_prim[:, lx[2]:lx[2] + bnx[2] + 2 * NG, lx[1]:lx[1] + bnx[1] + 2 * NG,
                            lx[0]:lx[0] + bnx[0] + 2 * NG]
            m.set_prim(g, np.ascontiguousarray(win))
        m.initialize()
        if not args.no_tune:
            m.autotune(2)
        return m

    def totals(m):
        t = np.zeros(5)
        for g in range(m.nblocks):
            c = m.get(g, A.DRIVER_CONS)[:5, NG:-NG, NG:-NG, NG:-NG]
            t += c.reshape(5, -1).sum(axis=1)
        return t

    def profiled_step(m):
        phx.L.phx_rt_profile_start(phx.h)
        phx.L.phx_rt_profile_reset(phx.h)
        m.step()
        n = C.c_int()
        phx.L.phx_rt_profile_count(phx.h, C.byref(n))
        launch = {}
        for r in range(n.value):
            name = C.create_string_buffer(64)
            ms = C.c_float()
            phx.L.phx_rt_profile_get(phx.h, r, name, 64, C.byref(ms))
            k = name.value.decode()
            launch[k] = launch.get(k, 0.0) + ms.value / 2.0  # per stage
        phx.L.phx_rt_profile_stop(phx.h)
        return launch

    def timed(m, nsteps):
        t0 = time.perf_counter()
        m.execute(1e300, nsteps)
        return time.perf_counter() - t0

    def last_iters(m):
        it = np.concatenate([m.get_iters(g)[NG:-NG, NG:-NG, NG:-NG].ravel()
                             for g in range(m.nblocks)])
        fails = sum(int(np.sum(m.get(g, A.DRIVER_FAIL)[NG:-NG, NG:-NG, NG:-NG] == 0.0))
                    for g in range(m.nblocks))
        return {"mean_iters": float(np.mean(np.maximum(it, 0))),
                "hist": np.bincount(np.maximum(it, 0)).tolist(),
                "unbracketed": int(np.sum(it < 0)), "fail_count_after_fixup_pass": fails}

    K, L = args.c5_steps, args.c5_log_steps
    cells = None
    # ---- plain ------------------------------------------------------------------------
    log(f"[c5] mesh {N} in 32 blocks of {bnx}, {K} RK2 steps, plain")
    m = make()
    cells = m.interior_cells
    tot0 = totals(m)
    launch_plain = profiled_step(m)
    tot_steps = [totals(m)]
    wall_a = 0.0
    for _ in range(max(0, args.c5_cpu_steps - 1)):  # totals after each of the first steps
        wall_a += timed(m, 1)
        tot_steps.append(totals(m))
    wall_a += timed(m, L - len(tot_steps))
    wall_b = timed(m, K - L)
    t_plain, dt_plain, n_plain = m.time()
    tot_plain = totals(m)
    it_plain = last_iters(m)
    m.close()
    wall_all = wall_a + wall_b
    log(f"[c5] plain: {1e3 * wall_all / (K - 1):.2f} ms/step, steps {L + 1}-{K} "
        f"{1e3 * wall_b / (K - L):.2f} ms/step, t = {t_plain:.4g}")
    out = {"workload": "FMKS torus 256x128x128, 32 blocks of 64x64x32, WENO5Z + HLL, "
                       "fixups + floors (torus deck), gr_outflow x1/x2 + periodic x3",
           "steps": K, "interior_cells": cells,
           "plain": {"ms_per_step": 1e3 * wall_all / (K - 1),
                     "zone_updates_per_s": cells * 2 * (K - 1) / wall_all,
                     "hbm_frac_stage_bytes": cells * 2 * (K - 1) / wall_all * B_ALG_STAGE
                                             / 1e9 / HBM_PEAK_GBS,
                     "ms_per_step_after_switch": 1e3 * wall_b / (K - L),
                     "time": t_plain, "dt": dt_plain, "ncycle": n_plain,
                     "launch_ms_per_stage": launch_plain, "con2prim_last_stage": it_plain},
           "conserved_totals_initial": tot0.tolist(),
           "conserved_drift_plain": (tot_plain - tot0).tolist()}
    if not args.no_cpu and args.c5_cpu_steps > 0:
        out["conserved_vs_cpu"] = c5_cpu_totals(p, fx, N, bnx, xmin, xmax, bc, grid_prim,
                                                tot0, tot_steps, log)
    if args.no_c5_learned:
        return out
    # ---- learned: log pairs over the first L steps, train, switch ------------------------
    m = make()
    logged = []
    per_block = max(1, (args.c5_pairs // L) // 2)
    nb_int = bnx[0] * bnx[1] * bnx[2]
    stride = max(1, nb_int // per_block)
    for s_ in range(L):
        m.step()
        logged.append(mesh_pairs(m, [(2 * s_) % m.nblocks, (2 * s_ + 1) % m.nblocks], stride))
    data = {k: np.ascontiguousarray(np.concatenate([d[k] for d in logged], axis=-1))
            for k in logged[0]}
    data["d_coords"] = DeviceArray.from_host(data["coords"])
    t0 = time.perf_counter()
    net, losses = train_bracket_on(phx, p, np.random.default_rng(SEED + 5), data,
                                   steps=args.c5_train_steps, **TORUS_BRACKET_TRAINING)
    train_s = time.perf_counter() - t0
    log(f"[c5] trained on {data['cons'].shape[1]} pairs from steps 1-{L} in {train_s:.1f} s, "
        f"loss {losses[0]:.4f} -> {losses[-1]:.4f}")
    m.set_learned(net)
    m.learned_stats(reset=True)
    launch_learned = profiled_step(m)
    wall_l = timed(m, K - L - 1)
    misses, solves = m.learned_stats()
    t_l, dt_l, n_l = m.time()
    tot_l = totals(m)
    it_l = last_iters(m)
    m.close()
    log(f"[c5] learned: steps {L + 2}-{K} {1e3 * wall_l / (K - L - 1):.2f} ms/step, "
        f"bracket misses {misses}/{solves}")
    rel = lambda a, b: ((a - b) / np.maximum(np.abs(b), 1e-300)).tolist()
    out.update({
            "learned": {"pairs": int(data["cons"].shape[1]), "log_steps": L,
                        "train_steps": args.c5_train_steps, "train_s": train_s,
                        "loss_first_last": [losses[0], losses[-1]],
                        "ms_per_step_after_switch": 1e3 * wall_l / (K - L - 1),
                        "zone_updates_per_s_after_switch": cells * 2 * (K - L - 1) / wall_l,
                        "bracket_misses": misses, "learned_solves": solves,
                        "miss_rate": misses / max(solves, 1),
                        "time": t_l, "ncycle": n_l, "launch_ms_per_stage": launch_learned,
                        "con2prim_last_stage": it_l},
            "conserved_drift_learned": (tot_l - tot0).tolist(),
            "learned_vs_plain_totals_rel": rel(tot_l, tot_plain)})
    return out


def c2_stage(phx, args, recon_name, prim_h, step, full=True, arith=None):
    """One RK stage of the 128^3 GRMHD block with reconstruction recon_name: the stage is
    timed at the input state's own CFL step (EstimateTimestep's first-call form over the
    input's cell signal speeds, cfl = 0.4), with the con2prim guess of every repeat reset to
    the input state's own root (the warm start a time-stepping run sees), so the K steps do
    identical work.  Returns the leg's numbers; full = True adds the roofline ratings."""
    from phoebus_amd import _abi as A
    from phoebus_amd.api import BlockState, Phx, block, fluid_params
    from phoebus_amd.device import DeviceArray, synchronize
    from phoebus_amd.workloads import magnetization

    nx = args.nx
    step(f"{nx}^3 stage set-up ({recon_name})")
    recon = {"weno5z": A.RECON_WENO5Z, "plm": A.RECON_LINEAR, "mp5": A.RECON_MP5,
             "constant": A.RECON_CONSTANT}[recon_name]
    riemann = {"hll": A.RIEMANN_HLL, "llf": A.RIEMANN_LLF}[args.riemann]
    arith = arith or args.flux_arith
    p = fluid_params(gamma=GAMMA, mhd=True, recon=recon, riemann=riemann,
                     c2p_tol=1e-10, c2p_max_iter=100,
                     flux_arith={"exact": A.FLUX_ARITH_EXACT, "fast": A.FLUX_ARITH_FAST}[arith])
    b = block(3, (nx, nx, nx), NG, dx=(1.0 / nx,) * 3)
    st = BlockState(p, b)
    st.prim.upload(prim_h)
    f_all = st.fields()
    phx.prim_to_con(p, b, f_all)            # product p2c builds the conserved input
    # the stage's dt: cfl x the minimum over the input state's cell signal speeds
    dt_d = DeviceArray((1,))
    phx.estimate_dt(p, b, f_all, 0, 0.4, dt_d)
    # the input state's own con2prim root: the warm start of the timed stages
    prim_work = DeviceArray(st.prim.shape)
    mu0 = DeviceArray(st.c2p_mu.shape)
    mu0.zero()
    it0 = DeviceArray(st.shape, np.int32)
    cons_copy = DeviceArray.from_host(st.cons.get())
    f0 = Phx.fields(prim=prim_work, cons=cons_copy, temp=st.temp, c2p_mu=mu0, fail=st.fail,
                    cell_sig=st.cell_sig)
    I = (slice(NG, NG + nx),) * 3
    lo, hi = (NG, NG, NG), (NG + nx - 1,) * 3
    phx.con2prim(p, b, f0, lo, hi, iters=it0)
    synchronize()
    dt = float(dt_d.get()[0])
    cold_iters = it0.get()[I]
    del cons_copy, prim_work

    # per-step buffers: the stage reads st.prim / st.cons and writes u_out / prim_out
    u_out = DeviceArray(st.cons.shape)
    prim_out = DeviceArray(st.prim.shape)
    iters = DeviceArray(st.shape, np.int32)
    f_stage = st.fields()
    sp = C.c_void_p()
    phx.L.phx_rt_stream_create(C.byref(sp))
    s = sp.value

    # pick the y/z flux sweep shapes for this block and reconstruction (times every
    # admissible shape on these fields; results do not depend on it), or fix them
    if args.yz_shape:
        ti, ts = (int(x) for x in args.yz_shape.split(","))
        for d in (1, 2):
            phx.set_tile_shape(p, b, d, ti, ts)
    elif not args.no_tune:
        step(f"autotune y/z sweeps ({recon_name})")
        phx.autotune(p, b, f_stage, reps=3, stream=s)
        phx.L.phx_rt_stream_synchronize(s)
    tiles = {d: list(phx.tile_shape(p, b, d)) for d in (1, 2)}
    step(f"timed stages ({recon_name}, dt {dt:.4e}), shapes {tiles}")

    def stage():
        phx.L.phx_rt_memcpy_dtod_async(st.c2p_mu.ptr, mu0.ptr, mu0.nbytes, s)
        phx.fluid_stage(p, b, f_stage, 1.0, dt, st.cons, u_out, prim_out, iters=iters,
                        cfl=0.4, dt_out=dt_d, stream=s)

    for _ in range(args.warmup):
        stage()
    phx.L.phx_rt_stream_synchronize(s)
    synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        stage()
    phx.L.phx_rt_stream_synchronize(s)
    synchronize()
    wall = time.perf_counter() - t0
    ms_per_step = 1e3 * wall / args.steps

    # per-launch device time: HIP events the library records around each launch of the
    # stage, on the stream the launch uses
    phx.L.phx_rt_profile_start(phx.h)
    acc, cnt = {}, {}
    for _ in range(args.steps):
        phx.L.phx_rt_profile_reset(phx.h)
        stage()
        phx.L.phx_rt_stream_synchronize(s)
        n = C.c_int()
        phx.L.phx_rt_profile_count(phx.h, C.byref(n))
        for r in range(n.value):
            name = C.create_string_buffer(64)
            ms = C.c_float()
            phx.L.phx_rt_profile_get(phx.h, r, name, 64, C.byref(ms))
            k = name.value.decode()
            acc[k] = acc.get(k, 0.0) + ms.value
            cnt[k] = cnt.get(k, 0) + 1
    phx.L.phx_rt_profile_stop(phx.h)
    phx.L.phx_rt_stream_destroy(s)
    launch_ms = {k: acc[k] / cnt[k] for k in acc}
    warm_iters = iters.get()[I]
    value = nx ** 3 / (ms_per_step * 1e-3)
    leg = {"flux_arith": arith, "ms_per_step": ms_per_step, "zone_updates_per_s": value,
           "stage_hbm_frac": value * B_ALG_STAGE / 1e9 / HBM_PEAK_GBS,
           "dt": dt, "flux_tiles_yz": tiles, "launch_ms": launch_ms,
           "con2prim_mean_iters_warm": float(np.mean(warm_iters)),
           "fail_count": int(np.sum(st.fail.get()[I] == 0.0))}
    if not full:
        return leg

    kb = kernel_bytes(nx)
    dom = max(launch_ms, key=lambda k: launch_ms[k])
    traffic, stale = None, None
    tfile = os.path.join(ROOT, "profiles", "stage_128cubed_weno5_hll_traffic.json")
    if (nx == 128 and recon_name == "weno5z" and args.riemann == "hll"
            and os.path.exists(tfile)
            and json.load(open(tfile)).get("flux_arith", "exact") == arith):
        tj = json.load(open(tfile))
        stale = profile_stale(tj)
        tkern = tj["kernels"]
        key = dom
        if dom in ("flux_y", "flux_z"):
            t = tiles[1 if dom == "flux_y" else 2]
            key = f"{dom}@{t[0]}x{t[1]}"
            other = f"{'flux_z' if dom == 'flux_y' else 'flux_y'}@{t[0]}x{t[1]}"
            if key not in tkern and other in tkern:  # the same sweep code and face count
                key = other
        tk = tkern.get(key)
        traffic = tk["traffic_bytes"] if tk else None
    # the flux sweeps are bound by VALU issue: FP64 FLOP/s and issue occupancy from the
    # instruction counts of rocprofv3 SQ_INSTS_VALU* runs of this workload (per launch,
    # at the tile shape in use) over the live launch times
    valu = None
    vfile = os.path.join(ROOT, "profiles", "stage_128cubed_weno5_hll_valu.json")
    if (nx == 128 and recon_name == "weno5z" and args.riemann == "hll"
            and os.path.exists(vfile)
            and json.load(open(vfile)).get("flux_arith", "exact") == arith):
        vj = json.load(open(vfile))
        vk = vj["kernels"]
        valu = {"profile_stale": profile_stale(vj)}
        for name in ("flux_x", "flux_y", "flux_z", "update_con2prim"):
            key = name
            if name in ("flux_y", "flux_z"):
                t = tiles[1 if name == "flux_y" else 2]
                key = f"{name}@{t[0]}x{t[1]}"
                other = f"{'flux_z' if name == 'flux_y' else 'flux_y'}@{t[0]}x{t[1]}"
                # y and z sweeps of a shape run the same code on as many faces (their
                # counts differ by < 1.5 %): one direction's counts stand in for the other's
                if key not in vk and other in vk:
                    key = other
            if key not in vk or name not in launch_ms:
                continue
            sec = launch_ms[name] * 1e-3
            valu[name] = {"fp64_tflops": vk[key]["fp64_flop"] / sec / 1e12,
                          "fp64_frac": vk[key]["fp64_flop"] / sec / 1e12 / FP64_PEAK_TFLOPS,
                          "valu_issue_frac": vk[key]["valu_wave_instr"] * 4 /
                          (sec * VALU_SIMDS * CLOCK_HZ)}
    # the dominant launch's roofline.  The fused update is HBM-bound: achieved = the
    # stage's algorithmic bytes (416 B per zone-update: the compulsory state traffic of a
    # fully fused stage) times the zones it updates, over its average duration, against
    # 8 TB/s; the launch's own compulsory bytes (its inputs and outputs, including the face
    # fluxes a fully fused stage would not move) are reported beside it.  A flux sweep,
    # should one be the longest launch, is bound by FP64 VALU issue and is rated by its
    # measured FP64 FLOP/s against the 78.6 TFLOP/s FP64 peak.
    sec = launch_ms[dom] * 1e-3
    if dom == "update_con2prim":
        alg = B_ALG_STAGE * nx ** 3
    else:
        alg = kb[dom]
    achieved = alg / sec / 1e9
    roofline = {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                "traffic_profile_stale": stale,
                "algorithmic_bytes_per_launch": alg,
                "launch_bytes": kb[dom], "launch_bytes_frac": kb[dom] / sec / 1e9 / HBM_PEAK_GBS,
                "avg_launch_ms": launch_ms[dom]}
    if dom.startswith("flux") and valu and dom in valu:
        tf = valu[dom]["fp64_tflops"]
        roofline = {"bound": "mfma", "pipe": "fp64 VALU", "kernel": dom, "achieved": tf,
                    "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": tf / FP64_PEAK_TFLOPS,
                    "traffic": traffic, "traffic_profile_stale": stale,
                    "valu_issue_frac": valu[dom]["valu_issue_frac"],
                    "avg_launch_ms": launch_ms[dom]}
    # the magnetization of the input's interior (b^2 / rho, comoving)
    sig = magnetization(prim_h[(slice(None),) + I])
    leg.update({
        "roofline": roofline, "valu_roofline": valu,
        "magnetization": {"sigma_percentiles_1_10_50_90_99":
                          np.percentile(sig, [1, 10, 50, 90, 99]).tolist()},
        "con2prim": {"mean_iters_warm": float(np.mean(warm_iters)),
                     # a wavefront runs its slowest lane's iterations: the mean over the
                     # update's waves (64 consecutive interior cells) of the largest count,
                     # the headroom a straggler compaction could recover
                     "mean_wave_max_iters_warm": float(np.mean(np.max(
                         np.maximum(warm_iters, 0).reshape(-1, 64), axis=1))),
                     "mean_iters_cold": float(np.mean(cold_iters)),
                     "hist_warm": np.bincount(np.maximum(warm_iters, 0).ravel()).tolist(),
                     "hist_cold": np.bincount(np.maximum(cold_iters, 0).ravel()).tolist(),
                     "fail_count": leg["fail_count"],
                     "mlp_guess": None},
        "_cons": st.cons.get(), "_mu0": mu0.get()})
    return leg


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--nx", type=int, default=NX)
    ap.add_argument("--recon", choices=("weno5z", "plm", "mp5", "constant"), default="weno5z")
    ap.add_argument("--riemann", choices=("hll", "llf"), default="hll")
    ap.add_argument("--flux-arith", choices=("fast", "exact"), default="fast",
                    help="arithmetic of the flux sweeps: fast (contraction, Newton-refined "
                         "reciprocals; fluxes within 1e-12 of the reference's rounding) or "
                         "exact (the reference's rounding, bit for bit)")
    ap.add_argument("--cpu-nx", type=int, default=NX)
    ap.add_argument("--cpu-stages", type=int, default=10)
    ap.add_argument("--cpu-nx1", type=int, default=64,
                    help="block edge of the one-thread CPU sample")
    ap.add_argument("--legs", default="plm,mp5",
                    help="further reconstructions timed on the same block (comma list)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-c3", action="store_true", help="skip the batched con2prim leg")
    ap.add_argument("--no-tune", action="store_true", help="default flux tile shapes")
    ap.add_argument("--yz-shape", default="", metavar="TI,TS",
                    help="fix the y/z flux sweep shape of the 128^3 stage (TS < 0: marching "
                         "waves of TI columns in -TS segments) instead of autotuning it")
    ap.add_argument("--c3-states", type=int, default=1 << 26)
    ap.add_argument("--c3-pairs", type=int, default=1 << 22)
    ap.add_argument("--c3-train-steps", type=int, default=200000)
    ap.add_argument("--c3-net", default="", metavar="FILE.npz",
                    help="C3's trained network: loaded from this file when it exists, else "
                         "trained and saved there")
    ap.add_argument("--no-c4", action="store_true", help="skip the uneven block pack leg")
    ap.add_argument("--c4-blocks", type=int, default=512)
    ap.add_argument("--c4-steps", type=int, default=5)
    ap.add_argument("--c4-per-block", type=int, default=1)
    ap.add_argument("--no-c4-learned", action="store_true", help="skip C4's learned-bracket run")
    ap.add_argument("--no-c5", action="store_true", help="skip the full torus mesh leg")
    ap.add_argument("--c5-steps", type=int, default=200)
    ap.add_argument("--c5-log-steps", type=int, default=50)
    ap.add_argument("--no-c5-learned", action="store_true", help="skip C5's learned-bracket run")
    ap.add_argument("--c5-pairs", type=int, default=1 << 22)
    ap.add_argument("--c5-train-steps", type=int, default=6000)
    ap.add_argument("--c5-cpu-steps", type=int, default=2,
                    help="steps of the CPU oracle mesh run compared with the GPU totals")
    args = ap.parse_args()
    # the CPU baseline's thread placement (read when the OpenMP runtime starts)
    os.environ.setdefault("OMP_PROC_BIND", "close")
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark measures one GPU; multi-GPU runs are not "
                                   "part of it (the stage would shard by meshblocks, one rank "
                                   "per GPU, with the ghost exchange between ranks)"}))
        sys.exit(2)

    from phoebus_amd import _abi as A
    from phoebus_amd.api import BlockState, Phx, block, fluid_params
    from phoebus_amd.device import DeviceArray, Event, synchronize
    from phoebus_amd.workloads import grmhd_block_prims

    def step(msg):  # progress on stderr, flushed before each GPU step
        print(f"[bench] {msg}", file=sys.stderr, flush=True)

    nx = args.nx
    rng = np.random.default_rng(SEED)
    prim_h = grmhd_block_prims(rng, nx, NG, GAMMA)
    step("HIP context")
    phx = Phx(0)
    lines = {}
    cons_h = mu0_h = None
    dt_cfl = None
    other = "exact" if args.flux_arith == "fast" else "fast"
    legs = [(args.recon, args.flux_arith), (args.recon, other)] + [
        (r, args.flux_arith) for r in args.legs.split(",") if r and r != args.recon]
    for recon, arith in legs:
        main = (recon, arith) == (args.recon, args.flux_arith)
        leg = c2_stage(phx, args, recon, prim_h, step, full=main, arith=arith)
        if main:
            main_leg = leg
            cons_h, mu0_h, dt_cfl = leg.pop("_cons"), leg.pop("_mu0"), leg["dt"]
        else:
            for k in ("_cons", "_mu0"):
                leg.pop(k, None)
        lines[f"{recon}-{arith}"] = leg
    value = main_leg["zone_updates_per_s"]
    ms_per_step = main_leg["ms_per_step"]
    out = {
        "metric": "zone-updates/s of GRMHD RK stage; con2prim mean iters, MLP-guess vs plain",
        "value": value,
        "unit": "zone-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded; see bench.py docstring)",
        "config": {"workload": f"single 3-D GRMHD meshblock {nx}^3 + {NG} ghosts, Minkowski, "
                               f"Gamma=13/9, {args.recon.upper()} + {args.riemann.upper()} + "
                               "FluxCT + divergence/RK update + robust con2prim (tol 1e-10) + "
                               "CFL, one RK stage at the input state's CFL dt",
                   "zones": nx ** 3, "recon": args.recon, "riemann": args.riemann,
                   "flux_arith": args.flux_arith,
                   "dt": dt_cfl, "cfl": 0.4,
                   "flux_tiles_yz": main_leg["flux_tiles_yz"]},
        "roofline": main_leg.pop("roofline"),
        "stage_roofline": {"bytes_per_zone_update": B_ALG_STAGE,
                           "achieved_GBs": value * B_ALG_STAGE / 1e9,
                           "frac": value * B_ALG_STAGE / 1e9 / HBM_PEAK_GBS},
        "launch_ms": main_leg["launch_ms"],
        "valu_roofline": main_leg.pop("valu_roofline"),
        "magnetization": main_leg.pop("magnetization"),
        "con2prim": main_leg.pop("con2prim"),
        "recon_legs": {r: lines[r] for r in lines},
    }
    if not args.no_c4:
        if "st" in locals():
            del st, u_out, prim_out, iters
        out["uneven_pack"] = c4_uneven_pack(
            phx, args, lambda m: print(m, file=sys.stderr, flush=True))
    if not args.no_c5:
        if "st" in locals():
            del st, u_out, prim_out, iters
        out["torus_mesh"] = c5_torus_mesh(
            phx, args, lambda m: print(m, file=sys.stderr, flush=True))
    if not args.no_c3:
        if "st" in locals():
            del st, u_out, prim_out, iters
        out["con2prim"]["mlp_guess"] = c3_learned_vs_plain(
            phx, args, lambda m: print(m, file=sys.stderr, flush=True))
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(prim_h, cons_h, mu0_h, dt_cfl, args,
                                           lambda m: print(m, file=sys.stderr, flush=True))
    print(json.dumps(out), flush=True)
    phx.close()


if __name__ == "__main__":
    main()
