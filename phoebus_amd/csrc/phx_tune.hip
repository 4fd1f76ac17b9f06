// phx_tune.hip -- autotuner of the y/z flux tile shape (include/phx.h, phx_autotune).
//
// The y/z flux kernels either stage ti x ts cells of every primitive in LDS (tiles,
// ts > 0) or march one wave of ti columns by 64 / ti rows along the sweep in -ts segments
// (ts < 0).  Which shape is fastest depends on the block extent (halo overhead 2H+1 of ts
// rows per tile; one cell per segment and the last step's unused rows when marching), the
// reconstruction order (H, register pressure), the variable count (LDS per block), the
// number of waves in flight and cache-line use along i.  The tuner times every admissible
// shape of both kinds on the caller's own
// fields with HIP events and keeps the fastest per (recon, riemann, mhd, geometry, block
// size, direction) in the context; later flux and stage calls with the same key use it.
// Results are bit-identical for every shape, so tuning changes time only.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <numeric>
#include <vector>

#include "phx.h"
#include "phx_internal.hpp"

using namespace phx;

extern "C" {

int phx_autotune(phx_ctx *ctx, const phx_fluid_params *p, const phx_block *b,
                 const phx_block_fields *f, int reps, phx_stream stream) {
  if (!ctx) return PHX_ERR_INVALID;
  int rc = check_params(ctx, p);
  if (rc) return rc;
  if ((rc = check_block(ctx, b, p->recon))) return rc;
  if (!f) return ctx->fail(PHX_ERR_INVALID, "null field");
  const BlockGeom bg = make_block_geom(b);
  const int np = p->mhd ? 10 : 7, nc = p->mhd ? 8 : 5;
  const PrimPtrs prim = prim_ptrs(f, bg.ncell, np);
  const FluxPtrs flux = flux_ptrs(f, bg.ncell, nc);
  if (!complete(prim, np) || !complete_flux(flux, b->ndim, nc) || !f->face_sig)
    return ctx->fail(PHX_ERR_INVALID, "null field");
  if (reps < 1) reps = 3;
  const DevParams dp = make_dev_params(p);
  hipStream_t s = ctx->stream_of(stream);
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
    return ctx->fail(PHX_ERR_DEVICE, "event create failed");
  static const int kTi[] = {1, 2, 4, 8, 16, 32};
  static const int kTs[] = {16, 24, 32, 48, 64, 96, 128, 192, 256};
  for (int d = 1; d < b->ndim; ++d) {
    std::vector<TileShape> cand;
    for (int ti : kTi)
      for (int ts : kTs) {
        const TileShape want{ti, ts};
        const TileShape r = resolve_tile_shape(d, p->recon, p->mhd, bg, want);
        if (r.ti != ti || r.ts != ts) continue;  // not admissible as asked
        bool dup = false;
        for (const auto &c : cand) dup = dup || (c.ti == r.ti && c.ts == r.ts);
        if (!dup) cand.push_back(r);
      }
    // the full own extent (one tile per column) is always a candidate
    const TileShape full = resolve_tile_shape(d, p->recon, p->mhd, bg, TileShape{0, 1 << 20});
    for (int ti : kTi) {
      const TileShape r = resolve_tile_shape(d, p->recon, p->mhd, bg, TileShape{ti, full.ts});
      bool dup = false;
      for (const auto &c : cand) dup = dup || (c.ti == r.ti && c.ts == r.ts);
      if (!dup) cand.push_back(r);
    }
    // marching waves: ti columns by 64 / ti rows, in nseg segments of the faces.  Every
    // count up to 10 and a few above: the wave count is nseg times the column chunks, and
    // which count fills the chip's last round of waves best (four waves per SIMD, 4096 at
    // once) depends on the block's cross-section
    static const int kSeg[] = {1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 12, 16};
    for (int ti : {1, 2, 4, 8, 16})
      for (int nseg : kSeg) {
        const TileShape r = resolve_tile_shape(d, p->recon, p->mhd, bg, TileShape{ti, -nseg});
        bool dup = false;
        for (const auto &c : cand) dup = dup || (c.ti == r.ti && c.ts == r.ts);
        if (!dup) cand.push_back(r);
      }
    // ms for n launches of candidate k after one warm-up launch; < 0 when a launch failed
    auto time_cand = [&](size_t k, int n) -> float {
      const TileShape &c = cand[k];
      launch_fluxes(d, p->recon, p->riemann, p->mhd, p->geom.kind, dp, bg, prim, flux,
                    f->face_sig, s, c, flux_fast_of(p));
      (void)hipEventRecord(e0, s);
      for (int r = 0; r < n; ++r)
        launch_fluxes(d, p->recon, p->riemann, p->mhd, p->geom.kind, dp, bg, prim, flux,
                      f->face_sig, s, c, flux_fast_of(p));
      (void)hipEventRecord(e1, s);
      if (hipEventSynchronize(e1) != hipSuccess) return -1.f;
      float ms = 0.f;
      (void)hipEventElapsedTime(&ms, e0, e1);
      return ms;
    };
    auto launch_failed = [&]() {
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      const int lrc = ctx->check_launch();
      return lrc ? lrc : ctx->fail(PHX_ERR_DEVICE, "tuning launch failed");
    };
    // two passes over the candidates, each kept at its faster pass: one slow measurement
    // (a clock step, another process) does not rule a shape out
    std::vector<float> t(cand.size(), 1e30f);
    for (int pass = 0; pass < 2; ++pass)
      for (size_t k = 0; k < cand.size(); ++k) {
        const float ms = time_cand(k, reps);
        if (ms < 0.f) return launch_failed();
        if (ms < t[k]) t[k] = ms;
      }
    // the fastest few go again, three interleaved rounds of 4 x reps launches each, and are
    // ranked by their median: among dozens of candidates the fastest of two short samples
    // is often a lucky one
    std::vector<size_t> order(cand.size());
    std::iota(order.begin(), order.end(), size_t(0));
    const size_t nref = std::min<size_t>(4, cand.size());
    std::partial_sort(order.begin(), order.begin() + nref, order.end(),
                      [&](size_t a, size_t c) { return t[a] < t[c]; });
    float m[4][3];
    for (int q = 0; q < 3; ++q)
      for (size_t i = 0; i < nref; ++i)
        if ((m[i][q] = time_cand(order[i], 4 * reps)) < 0.f) return launch_failed();
    size_t best = order[0];
    float best_med = 1e30f;
    for (size_t i = 0; i < nref; ++i) {
      std::sort(m[i], m[i] + 3);
      if (m[i][1] < best_med) {
        best_med = m[i][1];
        best = order[i];
      }
    }
    ctx->tiles[phx_ctx::tile_key(p, b, d)] = cand[best];
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return ctx->check_launch();
}

int phx_tile_shape(phx_ctx *ctx, const phx_fluid_params *p, const phx_block *b, int dir,
                   int *ti, int *ts) {
  if (!ctx || !p || !b || !ti || !ts || dir < 1 || dir > 2) return PHX_ERR_INVALID;
  const TileShape r =
      resolve_tile_shape(dir, p->recon, p->mhd, make_block_geom(b), ctx->tile_for(p, b, dir));
  *ti = r.ti;
  *ts = r.ts;
  return PHX_OK;
}

int phx_set_tile_shape(phx_ctx *ctx, const phx_fluid_params *p, const phx_block *b, int dir,
                       int ti, int ts) {
  if (!ctx || !p || !b || dir < 1 || dir > 2 || ti < 0) return PHX_ERR_INVALID;
  ctx->tiles[phx_ctx::tile_key(p, b, dir)] = TileShape{ti, ts};
  return PHX_OK;
}

}  // extern "C"
