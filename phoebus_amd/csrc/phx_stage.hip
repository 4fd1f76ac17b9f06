// phx_stage.hip -- one fused RK stage of the fluid path for a block (include/phx.h,
// phx_fluid_stage).  After the flux, CT and source launches, one kernel per interior cell
// forms dU/dt from the face fluxes, averages and updates the conserved state, inverts it
// (robust con2prim) and folds the cell into the CFL minimum.  U_{s+1} is written once,
// after con2prim, so it never makes a round trip through HBM between the update and the
// inversion.
#include <hip/hip_runtime.h>

#include <atomic>
#include <cstring>
#include <mutex>

#include "phx.h"
#include "phx_device.hpp"
#include "phx_internal.hpp"

namespace phx {

// The face fluxes of conserved component l at cell c: lower (m) and upper (p) face in each
// direction.  With CT (MHD, 2-D or 3-D) the magnetic components are FluxCT's rewrite
// (fluid.cpp:795-837) formed here from the corner EMFs, with the expressions and operation
// order of k_ct_flux_3d / k_ct_flux_2d, so the update gets the same bits as from the
// rewritten flux arrays without that pass; emf holds the 3 (3-D) or 1 (2-D, the z
// component) corner EMF fields of k_ct_emf_*.
template <int NDIM, bool CT, class FP>
__device__ __forceinline__ void face_fluxes(const FP &flux, const double *__restrict__ emf,
                                            int l, long c, long n, long sj, long sk,
                                            double f[3][2]) {
  if (CT && l >= 5 && NDIM == 3) {
    const double *e0 = emf, *e1 = emf + n, *e2 = emf + 2 * n;
    if (l == 5) {
      f[0][0] = f[0][1] = 0.0;
      f[1][0] = -0.5 * (e2[c] + e2[c + 1]);
      f[1][1] = -0.5 * (e2[c + sj] + e2[c + sj + 1]);
      f[2][0] = 0.5 * (e1[c] + e1[c + 1]);
      f[2][1] = 0.5 * (e1[c + sk] + e1[c + sk + 1]);
    } else if (l == 6) {
      f[0][0] = 0.5 * (e2[c] + e2[c + sj]);
      f[0][1] = 0.5 * (e2[c + 1] + e2[c + 1 + sj]);
      f[1][0] = f[1][1] = 0.0;
      f[2][0] = -0.5 * (e0[c] + e0[c + sj]);
      f[2][1] = -0.5 * (e0[c + sk] + e0[c + sk + sj]);
    } else {
      f[0][0] = -0.5 * (e1[c] + e1[c + sk]);
      f[0][1] = -0.5 * (e1[c + 1] + e1[c + 1 + sk]);
      f[1][0] = 0.5 * (e0[c] + e0[c + sk]);
      f[1][1] = 0.5 * (e0[c + sj] + e0[c + sj + sk]);
      f[2][0] = f[2][1] = 0.0;
    }
    return;
  }
  if (CT && NDIM == 2 && (l == 5 || l == 6)) {
    if (l == 5) {
      f[0][0] = f[0][1] = 0.0;
      f[1][0] = -0.5 * (emf[c] + emf[c + 1]);
      f[1][1] = -0.5 * (emf[c + sj] + emf[c + sj + 1]);
    } else {
      f[0][0] = 0.5 * (emf[c] + emf[c + sj]);
      f[0][1] = 0.5 * (emf[c + 1] + emf[c + 1 + sj]);
      f[1][0] = f[1][1] = 0.0;
    }
    return;
  }
  f[0][0] = flux[l][c];
  f[0][1] = flux[l][c + 1];
  if constexpr (NDIM >= 2) {
    f[1][0] = flux[8 + l][c];
    f[1][1] = flux[8 + l][c + sj];
  }
  if constexpr (NDIM == 3) {
    f[2][0] = flux[16 + l][c];
    f[2][1] = flux[16 + l][c + sk];
  }
}

// interior cell t of block b: returns the cell's CFL time (DBL_MAX for t beyond the block
// or when with_dt is false).  CT: take the magnetic fluxes from the corner EMFs in emf
// (face_fluxes) instead of flux.
template <bool MHD, int GEOM, bool CACHED = false, int NDIM = 3, bool CT = false, class FP,
          class CP, class PP>
__device__ __forceinline__ double update_con2prim_body(
    const DevParams &p, const BlockGeom &b, long t, const FP &flux, const double *emf,
    const CP &src, const CP &u_stage, const CP &u_base, double beta, double bdt,
    const CP &u_out, const PP &prim_out, double *__restrict__ temp,
    double *__restrict__ mu, double *__restrict__ csig, const double *__restrict__ fsig,
    double *__restrict__ fail, int32_t *__restrict__ iters, bool with_dt,
    int c2p_mode = kC2pAll, const ColdQueue &cq = ColdQueue{}, int blk = 0) {
  constexpr int NC = Layout<MHD>::ncons;
  const int ni = b.ie[0] - b.is[0] + 1, nj = b.ie[1] - b.is[1] + 1, nk = b.ie[2] - b.is[2] + 1;
  const bool active = t < (long)ni * nj * nk;
  double ldt_min = DBL_MAX;
  if (active) {
    const int i = (int)(t % ni) + b.is[0];
    const int j = (int)((t / ni) % nj) + b.is[1];
    const int k = (int)(t / ((long)ni * nj)) + b.is[2];
    const long c = ((long)k * b.nj + j) * b.ni + i;
    const long n = b.ncell;
    const long sj = b.ni, sk = (long)b.ni * b.nj;
    // parthenon FluxDivergence (uniform coordinates), SumData, AverageIndependentData,
    // UpdateIndependentData -- the operation order of the separate kernels
    const double a1 = b.dx[1] * b.dx[2], a2 = b.dx[0] * b.dx[2], a3 = b.dx[0] * b.dx[1];
    const double vol = b.dx[0] * b.dx[1] * b.dx[2];
    // every component is read and computed before anything is stored: U_s may be the same
    // container as U^n (stage 1) and U_{s+1} the same as U^n (stage 2), so a store would
    // keep the compiler from issuing the next component's loads ahead of it; with the
    // stores last, all the loads of the update can be in flight at once.  (The geometric
    // source terms exist exactly for the curved metrics.)
    constexpr bool SRC = GEOM != PHX_GEOM_MINKOWSKI;
    double cs[NC], usv[NC];
    // with beta = 1 and U_s the same container as U^n (stage 1 of the RK2 cycle) the
    // average 1 U_s + 0 U_s is U_s itself, bit for bit, for every finite or NaN component
    // (an infinite one gives NaN): such a cell's U_s is not stored back
    bool identity = beta == 1.0;
#pragma unroll
    for (int l = 0; l < NC; ++l) identity = identity && u_stage[l] == u_base[l];
    bool inf_in = false;
#pragma unroll
    for (int l = 0; l < NC; ++l) {
      double f[3][2];
      face_fluxes<NDIM, CT && MHD>(flux, emf, l, c, n, sj, sk, f);
      double du = a1 * f[0][1] - a1 * f[0][0];
      if constexpr (NDIM >= 2) du += a2 * f[1][1] - a2 * f[1][0];
      if constexpr (NDIM == 3) du += a3 * f[2][1] - a3 * f[2][0];
      double dudt = -du / vol;
      if constexpr (SRC) {
        if (l >= 1 && l <= 4) dudt = dudt + src[l][c];
      }
      const double us = u_stage[l][c];
      usv[l] = beta * us + (1 - beta) * u_base[l][c];
      inf_in = inf_in || isinf(us);
      cs[l] = usv[l] + bdt * dudt;
    }
    if (!identity || inf_in) {
#pragma unroll
      for (int l = 0; l < NC; ++l) u_stage[l][c] = usv[l];
    }
    // U_{s+1} goes out before the inversion, which re-derives it only where it used a floor
    // or ceiling (con2prim_fields); no conserved variable is live across the solve
#pragma unroll
    for (int l = 0; l < NC; ++l) u_out[l][c] = cs[l];
    // kC2pWarm: a cell with no previous root to start from goes to the learned bracket
    if (c2p_mode == kC2pWarm && enqueue_cold(cq, !(mu[c] > 0.0), blk, c, false)) return DBL_MAX;
    // ConservedToPrimitive on U_{s+1}
    double sig[3] = {0.0, 0.0, 0.0};
    int it = 0;
    double st = 0.0;
    bool solved = false;
    if (!(isnan(cs[0]) + isnan(cs[1]) + isnan(cs[4]) > 0)) {
      const double X[3] = {xc(b, 0, i), xc(b, 1, j), xc(b, 2, k)};
      st = con2prim_fields<MHD, GEOM, CACHED>(p, b, (long)j * b.ni + i, X, cs, prim_out, temp, mu,
                                              u_out, c, sig, it, solved);
    }
    // a zone the inversion left (NaN input, unbracketed root) keeps its signal speeds
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      if (d >= b.ndim) break;
      if (solved)
        csig[d * n + c] = sig[d];
      else if (with_dt)
        sig[d] = csig[d * n + c];
    }
    fail[c] = st;
    if (iters) iters[c] = it;
    // EstimateTimestepBlock with face speeds (fluid.cpp:913-928)
    if (with_dt) {
      double ldt = 0.0;
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        if (d >= b.ndim) break;
        const long off = d == 0 ? 1 : d == 1 ? sj : sk;
        const double s = smax(sig[d], smax(fsig[d * n + c], fsig[d * n + c + off]));
        ldt += s / b.dx[d];
      }
      // a NaN time (a zone whose speeds are NaN, e.g. P = 0 on the energy floor, where
      // Gamma_1 = 0/0 as in the reference) constrains nothing: the serial minimum of the
      // reference and of k_dt_partial never lets a NaN replace a number
      const double t = 1.0 / (ldt + 1.e-50);
      ldt_min = isnan(t) ? DBL_MAX : t;
    }
  }
  return ldt_min;
}

template <bool MHD, int GEOM, int NDIM>
__global__ __launch_bounds__(256) void k_update_con2prim(
    DevParams p, BlockGeom b, FluxPtrs flux, const double *__restrict__ emf, ConsPtrs src,
    ConsPtrs u_stage, ConsPtrs u_base,
    double beta, double bdt, ConsPtrs u_out, PrimPtrs prim_out, double *__restrict__ temp,
    double *__restrict__ mu, double *__restrict__ csig, const double *__restrict__ fsig,
    double *__restrict__ fail, int32_t *__restrict__ iters, double *__restrict__ partial) {
  const double v = update_con2prim_body<MHD, GEOM, false, NDIM, MHD && NDIM >= 2>(
      p, b, (long)blockIdx.x * blockDim.x + threadIdx.x, flux, emf, src, u_stage, u_base, beta,
      bdt,
      u_out, prim_out, temp, mu, csig, fsig, fail, iters, partial != nullptr);
  if (partial) {
    const double m = block_min(v);
    if (threadIdx.x == 0) partial[blockIdx.x] = m;
  }
}

// packs: the fused update over a work list of 3-D MHD blocks; dt from dt_dev when given
// (a device-resident time step, so a captured graph needs no host value).  Every FMKS
// block of a pack carries its cached geometry; the FMKS variant fits four waves per SIMD
// (128 VGPRs) without spilling.
template <int GEOM>
__global__ __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu(GEOM == PHX_GEOM_FMKS ? 4 : 1))) void k_pack_update(DevParams p, PackView v, double beta,
                                                     double dt, const double *dt_dev,
                                                     double *__restrict__ partial, int c2p_mode,
                                                     ColdQueue cq) {
  const long g = xcd_remap(blockIdx.x, gridDim.x);
  const int b = pack_find(v.start, v.n, g);
  const PackEntry &E = v.e[b];
  const double bdt = beta * (dt_dev ? dt_dev[0] : dt);
  const long n = E.bg.ncell;
  const double val = update_con2prim_body<true, GEOM, GEOM == PHX_GEOM_FMKS, 3, true>(
      p, E.bg, (g - v.start[b]) * 256 + threadIdx.x, contig(E.flux, n), E.emf, contig(E.src, n),
      contig(E.u_stage, n), contig(E.u_base, n), beta, bdt, contig(E.u_out, n),
      contig(E.prim_out, n), E.temp, E.mu, E.csig, E.fsig, E.fail, E.iters,
      partial != nullptr, c2p_mode, cq, b);
  if (partial) {
    const double m = block_min(val);
    if (threadIdx.x == 0) partial[g] = m;
  }
}

void launch_pack_update(int geom, const DevParams &dp, const PackView &v, long groups,
                        double beta, double dt, const double *dt_dev, double *partial,
                        hipStream_t s, int c2p_mode, const ColdQueue &cq, double *extra) {
  if (groups <= 0) return;
  if (c2p_mode == kC2pWarm) launch_zero_count(cq.count, s, extra);
  with_geom(geom, [&](auto G) {
    hipLaunchKernelGGL((k_pack_update<decltype(G)::value>), dim3((unsigned)groups), dim3(256), 0,
                       s, dp, v, beta, dt, dt_dev, partial, c2p_mode, cq);
  });
}

}  // namespace phx

using namespace phx;

// Device scratch of the allocation-free stage (phx_fluid_stage_noalloc): the CFL partial
// minima, one per 256 interior cells, in a file-scope device array.  Each device holds its
// own copy for the life of the process; host code reaches it only through the address
// hipGetSymbolAddress returns on that device (the array's name in host code is a
// stand-in, not a device address).
static __device__ __align__(256) double scratch[kDtPartials];

namespace {
constexpr int kMaxDevices = 64;
std::atomic<double *> g_scratch[kMaxDevices];
std::mutex g_scratch_mu;
thread_local char g_noalloc_err[256];
}  // namespace

namespace phx {
// the address of the current device's scratch, looked up once per device
int static_scratch(double **out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess) return PHX_ERR_DEVICE;
  if (dev < 0 || dev >= kMaxDevices) return PHX_ERR_UNSUPPORTED;
  double *p = g_scratch[dev].load(std::memory_order_acquire);
  if (!p) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    p = g_scratch[dev].load(std::memory_order_relaxed);
    if (!p) {
      void *q = nullptr;
      if (hipGetSymbolAddress(&q, HIP_SYMBOL(scratch)) != hipSuccess || !q)
        return PHX_ERR_DEVICE;
      p = static_cast<double *>(q);
      g_scratch[dev].store(p, std::memory_order_release);
    }
  }
  *out = p;
  return PHX_OK;
}
const char *noalloc_last_error() { return g_noalloc_err; }
}  // namespace phx

// the stage's launches; partial = the CFL partial minima buffer (at least kDtPartials)
static int fluid_stage_launches(phx_ctx *ctx, const phx_fluid_params *p, const phx_block *b,
                                const phx_block_fields *f, const phx_stage_args *a,
                                phx_stream stream, double *partial_buf) {
  int rc = check_params(ctx, p);
  if (rc) return rc;
  if ((rc = check_block(ctx, b, p->recon))) return rc;
  if (!a || !f) return ctx->fail(PHX_ERR_INVALID, "null field");
  const int np = p->mhd ? 10 : 7, nc = p->mhd ? 8 : 5;
  const BlockGeom bg = make_block_geom(b);
  const PrimPtrs prim = prim_ptrs(f, bg.ncell, np);
  const ConsPtrs cons = cons_ptrs(f, bg.ncell, nc), src_c = src_ptrs(f, bg.ncell, nc);
  const FluxPtrs flux = flux_ptrs(f, bg.ncell, nc);
  const ConsPtrs u_base = comps<8>(a->u_base, bg.ncell, nc, a->base_vars ? a->base_vars->cons : nullptr);
  const ConsPtrs u_out = comps<8>(a->u_out, bg.ncell, nc, a->out_vars ? a->out_vars->cons : nullptr);
  const PrimPtrs prim_out =
      comps<10>(a->prim_out, bg.ncell, np, a->out_vars ? a->out_vars->prim : nullptr);
  if (!complete(prim, np) || !complete(cons, nc) || !complete_flux(flux, b->ndim, nc) ||
      !f->face_sig || !f->c2p_mu || !f->cell_sig || !f->fail || !complete(u_base, nc) ||
      !complete(u_out, nc) || !complete(prim_out, np))
    return ctx->fail(PHX_ERR_INVALID, "null field");
  if (p->mhd && b->ndim > 1 && (!f->emf || b->ng < 2)) return ctx->fail(PHX_ERR_INVALID, "CT needs emf");
  const bool curved = p->geom.kind != PHX_GEOM_MINKOWSKI;
  if (curved && !complete(src_c, nc)) return ctx->fail(PHX_ERR_INVALID, "curved metric needs src scratch");
  const long n = (long)b->nx[0] * b->nx[1] * b->nx[2];
  const long nb = (n + 255) / 256;
  if (a->dt_out && nb > kDtPartials)  // refused before anything is launched
    return ctx->fail(PHX_ERR_INVALID, "block too large for the CFL reduction");
  if (a->dt_out && !partial_buf) {  // the allocation-free form looks its scratch up here
    if ((rc = static_scratch(&partial_buf)))
      return ctx->fail(rc, "no device scratch on the current device");
  }
  const DevParams dp = make_dev_params(p);
  hipStream_t s = ctx->stream_of(stream);
  static const char *kFluxName[3] = {"flux_x", "flux_y", "flux_z"};
  for (int d = 0; d < b->ndim; ++d) {
    ctx->mark_begin(kFluxName[d], s);
    launch_fluxes(d, p->recon, p->riemann, p->mhd, p->geom.kind, dp, bg, prim, flux,
                  f->face_sig, s, ctx->tile_for(p, b, d), flux_fast_of(p));
    ctx->mark_end(s);
  }
  if (p->mhd && b->ndim > 1) {  // FluxCT: the EMFs here, the B-flux rewrite in the update
    ctx->mark_begin("flux_ct", s);
    launch_ct_emf(bg, flux, f->emf, s);
    ctx->mark_end(s);
  }
  if (curved) {
    ctx->mark_begin("sources", s);
    launch_sources(p->geom.kind, p->mhd, dp, bg, prim, src_c, s);
    ctx->mark_end(s);
  }
  double *partial = a->dt_out ? partial_buf : nullptr;
  const ConsPtrs src = curved ? src_c : ConsPtrs::contiguous(nullptr, 0);
  const double bdt = a->beta * a->dt;
  ctx->mark_begin("update_con2prim", s);
#define ARGS dp, bg, flux, f->emf, src, cons, u_base, a->beta, bdt, u_out, prim_out, \
             f->temp, f->c2p_mu, f->cell_sig, f->face_sig, f->fail, a->iters, partial
  with_geom(p->geom.kind, [&](auto G) {
    with_mhd(p->mhd, [&](auto H) {
      constexpr bool MH = decltype(H)::value;
      constexpr int GM = decltype(G)::value;
      auto kernel = b->ndim == 3   ? k_update_con2prim<MH, GM, 3>
                    : b->ndim == 2 ? k_update_con2prim<MH, GM, 2>
                                   : k_update_con2prim<MH, GM, 1>;
      hipLaunchKernelGGL(kernel, dim3(nb), dim3(256), 0, s, ARGS);
    });
  });
#undef ARGS
  ctx->mark_end(s);
  if (a->dt_out) {
    ctx->mark_begin("dt_finish", s);
    launch_dt_finish(partial, (int)nb, a->cfl, a->dt_out, s);
    ctx->mark_end(s);
  }
  return ctx->check_launch();
}

extern "C" int phx_fluid_stage(phx_ctx *ctx, const phx_fluid_params *p, const phx_block *b,
                               const phx_block_fields *f, const phx_stage_args *a,
                               phx_stream stream) {
  if (!ctx) return PHX_ERR_INVALID;
  return fluid_stage_launches(ctx, p, b, f, a, stream, ctx->dt_partial);
}

extern "C" int phx_fluid_stage_noalloc(const phx_fluid_params *p, const phx_block *b,
                                       const phx_block_fields *f, const phx_stage_args *a,
                                       phx_stream stream) {
  // a context on the stack for argument checks and the message only: it owns no device
  // memory and no stream (stream NULL = the device's null stream); sweep shapes are the
  // built-in defaults
  phx_ctx local;
  const int rc = fluid_stage_launches(&local, p, b, f, a, stream, nullptr);
  if (rc) {
    std::strncpy(g_noalloc_err, local.last_error.c_str(), sizeof(g_noalloc_err) - 1);
    g_noalloc_err[sizeof(g_noalloc_err) - 1] = '\0';
  } else {
    g_noalloc_err[0] = '\0';
  }
  return rc;
}
