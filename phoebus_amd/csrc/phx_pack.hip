// phx_pack.hip -- the fluid stage over a pack of meshblocks (include/phx_pack.h): block
// table and work lists on the device, one launch per kernel of the stage, tile tuning and
// HIP-graph capture.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <new>
#include <numeric>
#include <vector>

#include "phx.h"
#include "phx_internal.hpp"
#include "phx_pack.h"
#include "phx_pack_impl.hpp"

using namespace phx;

namespace phx {

// tile shapes must fit every block of the pack: resolve against the smallest extent (a
// marching shape, ts < 0, keeps the fewest segments any block allows)
TileShape pack_shape(const phx_pack *k, int d, TileShape want) {
  TileShape best = resolve_tile_shape(d, k->p.recon, 1, k->geoms[0], want);
  for (const auto &g : k->geoms) {
    const TileShape t = resolve_tile_shape(d, k->p.recon, 1, g, want);
    if (want.ts < 0 ? t.ts > best.ts : t.ts < best.ts) best.ts = t.ts;
    if (t.ti < best.ti) best.ti = t.ti;
  }
  return best;
}

int build_start(phx_pack *k, int kind) {
  std::vector<long> st(k->n + 1, 0);
  for (int b = 0; b < k->n; ++b) {
    const BlockGeom &g = k->geoms[b];
    const long nf = 2L * ((long)(g.ie[1] - g.is[1] + 1) * (g.ie[2] - g.is[2] + 1) +
                          (long)(g.ie[0] - g.is[0] + 1) * (g.ie[2] - g.is[2] + 1) +
                          (long)(g.ie[0] - g.is[0] + 1) * (g.ie[1] - g.is[1] + 1));
    const long nint = (long)(g.ie[0] - g.is[0] + 1) * (g.ie[1] - g.is[1] + 1) *
                      (g.ie[2] - g.is[2] + 1);
    const long c = kind < 3    ? pack_groups_flux(kind, k->p.recon, g, k->tiles[kind])
                   : kind < 5 ? pack_groups_cells(g, kind == 3)
                   : kind == 5 ? (g.ncell + 255) / 256
                   : kind == 6 ? (nf + 255) / 256
                               : (g.ncell - nint + 255) / 256;
    st[b + 1] = st[b] + c;
  }
  if (st[k->n] > (long)UINT_MAX) return PHX_ERR_INVALID;
  k->groups[kind] = st[k->n];
  if (!k->start[kind] && hipMalloc(&k->start[kind], sizeof(long) * (k->n + 1)) != hipSuccess)
    return PHX_ERR_NOMEM;
  return hipMemcpy(k->start[kind], st.data(), sizeof(long) * (k->n + 1),
                   hipMemcpyHostToDevice) == hipSuccess
             ? PHX_OK
             : PHX_ERR_DEVICE;
}

PackView view(const phx_pack *k, int stage, int kind) {
  PackView v;
  v.e = k->entries[stage - 1];
  v.start = k->start[kind];
  v.n = k->n;
  return v;
}

void launch_stage_fluxes(phx_ctx *ctx, phx_pack *k, int st, hipStream_t s) {
  static const char *kName[3] = {"flux_x", "flux_y", "flux_z"};
  for (int d = 0; d < 3; ++d) {
    ctx->mark_begin(kName[d], s);
    launch_pack_flux(d, k->p.recon, k->p.riemann, k->p.geom.kind, k->dp, view(k, st, d),
                     k->groups[d], k->tiles[d], s, flux_fast_of(&k->p));
    ctx->mark_end(s);
  }
  // FluxCT: the corner EMFs here; the update forms the rewritten B fluxes from them
  ctx->mark_begin("flux_ct", s);
  launch_pack_ct_emf(view(k, st, 3), k->groups[3], s);
  ctx->mark_end(s);
  if (k->p.geom.kind != PHX_GEOM_MINKOWSKI) {
    ctx->mark_begin("sources", s);
    launch_pack_sources(k->p.geom.kind, k->dp, view(k, st, 4), k->groups[4], s);
    ctx->mark_end(s);
  }
}

void launch_stage(phx_ctx *ctx, phx_pack *k, const phx_pack_stage_args *a, hipStream_t s) {
  const int st = a->stage;
  launch_stage_fluxes(ctx, k, st, s);
  // with fixups on, the CFL partials are taken after them (the repaired and floored cells'
  // signal speeds), in the floors launch; with the learned inversion and no fixups, the
  // update takes those of the cells it inverts and the learned pass folds its cells into
  // one more partial (partial[groups[4]])
  const bool learned = k->mlp != nullptr;
  const bool fixups = k->fx.c2p_fixup || k->fx.floors;
  double *partial = a->dt_out ? k->partial : nullptr;
  double *extra = (learned && partial && !fixups) ? partial + k->groups[4] : nullptr;
  ctx->mark_begin(learned ? "update" : "update_con2prim", s);
  launch_pack_update(k->p.geom.kind, k->dp, view(k, st, 4), k->groups[4], st == 1 ? 1.0 : 0.5,
                     a->dt, a->dt_dev, fixups ? nullptr : partial, s,
                     learned ? kC2pWarm : kC2pAll, k->cq, extra);
  ctx->mark_end(s);
  if (learned) {  // the cells the update queued (no previous root)
    ctx->mark_begin("con2prim_learned", s);
    launch_pack_c2p_cold(k->p.geom.kind, k->dp, k->mlp, view(k, st, 4), k->cq, k->counters,
                         extra, s);
    ctx->mark_end(s);
  }
  if (fixups) {
    if (k->fx.c2p_fixup) {
      ctx->mark_begin("c2p_fixup", s);
      launch_pack_c2p_fixup(k->p.geom.kind, k->dp, k->fx, view(k, st, 4), k->groups[4], s);
      ctx->mark_end(s);
    }
    ctx->mark_begin("floors", s);
    launch_pack_floors(k->p.geom.kind, k->dp, k->fx, view(k, st, 4), k->groups[4], partial, s,
                       false, k->fq);
    ctx->mark_end(s);
  }
  if (a->dt_out) {
    ctx->mark_begin("dt_finish", s);
    launch_dt_finish(k->partial, (int)k->groups[4] + (extra ? 1 : 0), a->cfl, a->dt_out, s);
    ctx->mark_end(s);
  }
}

}  // namespace phx

namespace {

int check_stage_args(phx_ctx *ctx, const phx_pack_stage_args *a) {
  if (!a || (a->stage != 1 && a->stage != 2))
    return ctx->fail(PHX_ERR_INVALID, "stage must be 1 or 2");
  if (!a->dt_dev && !(a->dt >= 0.0)) return ctx->fail(PHX_ERR_INVALID, "invalid dt");
  return PHX_OK;
}

}  // namespace

extern "C" {

int phx_pack_create(phx_ctx *ctx, const phx_fluid_params *p, int nblocks,
                    const phx_pack_block *blocks, phx_pack **out) {
  if (!ctx || !out || !blocks || nblocks <= 0) return PHX_ERR_INVALID;
  *out = nullptr;
  int rc = check_params(ctx, p);
  if (rc) return rc;
  if (!p->mhd) return ctx->fail(PHX_ERR_UNSUPPORTED, "packs hold 3-D MHD blocks");
  phx_pack *k = new (std::nothrow) phx_pack();
  if (!k) return PHX_ERR_NOMEM;
  k->device = ctx->device;
  k->p = *p;
  k->dp = make_dev_params(p);
  k->n = nblocks;
  k->blocks.assign(blocks, blocks + nblocks);
  const bool curved = p->geom.kind != PHX_GEOM_MINKOWSKI;
  std::vector<PackEntry> e[2];
  for (int b = 0; b < nblocks; ++b) {
    const phx_pack_block &B = blocks[b];
    if ((rc = check_block(ctx, &B.block, p->recon)) || B.block.ndim != 3) {
      phx_pack_destroy(k);
      return rc ? rc : ctx->fail(PHX_ERR_UNSUPPORTED, "packs hold 3-D MHD blocks");
    }
    const phx_block_fields &f = B.f;
    if (f.vars) {
      phx_pack_destroy(k);
      return ctx->fail(PHX_ERR_UNSUPPORTED, "pack blocks take contiguous fields (no vars table)");
    }
    if (!f.prim || !f.cons || !B.cons1 || !f.flux || !f.face_sig || !f.cell_sig || !f.fail ||
        !f.c2p_mu || !f.emf || (curved && !f.src)) {
      phx_pack_destroy(k);
      return ctx->fail(PHX_ERR_INVALID, "null field in a pack block");
    }
    BlockGeom g = make_block_geom(&B.block);
    if (p->geom.kind == PHX_GEOM_FMKS) {
      // the FMKS metric does not depend on X3: one cached plane per block serves every k
      double *geo = nullptr;
      const size_t bytes = sizeof(double) * kGeoPlanes * (size_t)g.ni * g.nj;
      if (hipMalloc(&geo, bytes) != hipSuccess) {
        phx_pack_destroy(k);
        return ctx->fail(PHX_ERR_NOMEM, "geometry cache allocation failed");
      }
      k->geo.push_back(geo);
      launch_geo_fill(k->dp, g, geo, ctx->stream);
      g.geo = geo;
    }
    k->geoms.push_back(g);
    k->interior += (long)B.block.nx[0] * B.block.nx[1] * B.block.nx[2];
    for (int st = 1; st <= 2; ++st) {
      PackEntry x;
      x.bg = g;
      x.prim = f.prim;
      x.flux = f.flux;
      x.fsig = f.face_sig;
      x.emf = f.emf;
      x.src = curved ? f.src : nullptr;
      x.u_stage = st == 1 ? f.cons : B.cons1;
      x.u_base = f.cons;
      x.u_out = st == 1 ? B.cons1 : f.cons;
      x.prim_out = f.prim;
      x.temp = f.temp;
      x.mu = f.c2p_mu;
      x.csig = f.cell_sig;
      x.fail = f.fail;
      x.iters = B.iters;
      e[st - 1].push_back(x);
    }
  }
  for (int d = 0; d < 3; ++d) k->tiles[d] = pack_shape(k, d, TileShape());
  bool ok = true;
  for (int st = 0; st < 2 && ok; ++st)
    ok = hipMalloc(&k->entries[st], sizeof(PackEntry) * nblocks) == hipSuccess &&
         hipMemcpy(k->entries[st], e[st].data(), sizeof(PackEntry) * nblocks,
                   hipMemcpyHostToDevice) == hipSuccess;
  for (int kind = 0; kind < 8 && ok; ++kind) ok = build_start(k, kind) == PHX_OK;
  ok = ok && hipMalloc(&k->partial, sizeof(double) * std::max({k->groups[4] + 1, k->groups[5], 1L})) ==
                 hipSuccess;
  ok = ok && hipMalloc(&k->counters, 2 * sizeof(unsigned long long)) == hipSuccess &&
       hipMemset(k->counters, 0, 2 * sizeof(unsigned long long)) == hipSuccess;
  ok = ok && hipStreamSynchronize(ctx->stream) == hipSuccess;  // the geometry planes are filled
  if (!ok) {
    phx_pack_destroy(k);
    return ctx->fail(PHX_ERR_DEVICE, "pack set-up failed");
  }
  *out = k;
  return PHX_OK;
}

int phx_pack_destroy(phx_pack *k) {
  if (!k) return PHX_OK;
  if (k->graphs > 0) return PHX_ERR_INVALID;  // its graphs still reference the work lists
  (void)hipSetDevice(k->device);
  for (auto &e : k->entries)
    if (e) (void)hipFree(e);
  for (auto &s : k->start)
    if (s) (void)hipFree(s);
  if (k->partial) (void)hipFree(k->partial);
  if (k->counters) (void)hipFree(k->counters);
  if (k->fq.cell) (void)hipFree(k->fq.cell);
  if (k->fq.grp) (void)hipFree(k->fq.grp);
  if (k->fq.count) (void)hipFree(k->fq.count);
  if (k->cq.cell) (void)hipFree(k->cq.cell);
  if (k->cq.count) (void)hipFree(k->cq.count);
  for (double *g : k->geo) (void)hipFree(g);
  delete k;
  return PHX_OK;
}

int phx_pack_nblocks(const phx_pack *k, int *n, int64_t *interior_cells) {
  if (!k) return PHX_ERR_INVALID;
  if (n) *n = k->n;
  if (interior_cells) *interior_cells = k->interior;
  return PHX_OK;
}

int phx_pack_set_fixup(phx_ctx *ctx, phx_pack *k, const phx_fixup_params *fx) {
  if (!ctx || !k) return PHX_ERR_INVALID;
  if (!fx) {
    k->fx = DevFixup{};
    return PHX_OK;
  }
  const int rc = check_fixup(ctx, fx);
  if (rc) return rc;
  const DevFixup d = make_dev_fixup(&k->p, fx);
  if (d.floors && !k->fq.cell) {
    // room for every cell of the largest floors launch (interior or entire work list)
    const long cap = 256L * std::max(k->groups[4], k->groups[5]);
    FloorQueue q;
    q.cap = cap;
    const bool ok = hipMalloc(&q.cell, sizeof(unsigned long long) * cap) == hipSuccess &&
                    hipMalloc(&q.grp, sizeof(int) * cap) == hipSuccess &&
                    hipMalloc(&q.count, sizeof(unsigned int)) == hipSuccess;
    if (!ok) {
      if (q.cell) (void)hipFree(q.cell);
      if (q.grp) (void)hipFree(q.grp);
      if (q.count) (void)hipFree(q.count);
      return ctx->fail(PHX_ERR_NOMEM, "floors queue allocation failed");
    }
    k->fq = q;
  }
  k->fx = d;
  return PHX_OK;
}

int phx_pack_set_learned(phx_ctx *ctx, phx_pack *k, const phx_mlp *mlp) {
  if (!ctx || !k) return PHX_ERR_INVALID;
  if (mlp && !k->cq.cell) {  // the queue of cells to bracket: room for every cell of the pack
    long cells = 0;
    for (const BlockGeom &g : k->geoms) cells += g.ncell;
    ColdQueue q;
    q.cap = cells;
    if (hipMalloc(&q.cell, sizeof(unsigned long long) * cells) != hipSuccess ||
        hipMalloc(&q.count, sizeof(unsigned int)) != hipSuccess) {
      if (q.cell) (void)hipFree(q.cell);
      return ctx->fail(PHX_ERR_NOMEM, "learned-bracket queue");
    }
    k->cq = q;
  }
  k->mlp = mlp;
  return PHX_OK;
}

int phx_pack_learned_stats(phx_pack *k, int reset, int64_t *misses, int64_t *solves) {
  if (!k) return PHX_ERR_INVALID;
  unsigned long long h[2] = {0, 0};
  if (hipDeviceSynchronize() != hipSuccess ||
      hipMemcpy(h, k->counters, sizeof(h), hipMemcpyDeviceToHost) != hipSuccess)
    return PHX_ERR_DEVICE;
  if (misses) *misses = (int64_t)h[0];
  if (solves) *solves = (int64_t)h[1];
  if (reset && hipMemset(k->counters, 0, sizeof(h)) != hipSuccess) return PHX_ERR_DEVICE;
  return PHX_OK;
}

int phx_pack_stage(phx_ctx *ctx, phx_pack *k, const phx_pack_stage_args *a,
                   phx_stream stream) {
  if (!ctx || !k) return PHX_ERR_INVALID;
  int rc = check_stage_args(ctx, a);
  if (rc) return rc;
  launch_stage(ctx, k, a, ctx->stream_of(stream));
  return ctx->check_launch();
}

int phx_pack_autotune(phx_ctx *ctx, phx_pack *k, int reps, phx_stream stream) {
  if (!ctx || !k) return PHX_ERR_INVALID;
  // tuning rewrites the work lists in place; a live graph would decode them with the grid
  // size and tile shape it was captured with
  if (k->graphs > 0)
    return ctx->fail(PHX_ERR_INVALID, "pack autotune while graphs of the pack exist");
  if (reps < 1) reps = 3;
  hipStream_t s = ctx->stream_of(stream);
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
    return ctx->fail(PHX_ERR_DEVICE, "event create failed");
  static const int kTi[] = {1, 2, 4, 8, 16};
  // tiles of ti x ts, and marching waves of ti columns in -ts segments
  static const int kTs[] = {16, 24, 32, 48, 64, 96, 128, -1, -2, -3, -4, -6, -8};
  int rc = PHX_OK;
  for (int d = 1; d < 3 && rc == PHX_OK; ++d) {
    std::vector<TileShape> seen;
    for (int ti : kTi)
      for (int ts : kTs) {
        const TileShape c = pack_shape(k, d, TileShape{ti, ts});
        bool dup = false;
        for (const auto &x : seen) dup = dup || (x.ti == c.ti && x.ts == c.ts);
        if (!dup) seen.push_back(c);
      }
    // ms for n launches of candidate q after one warm-up launch; < 0 on failure (rc set)
    auto time_cand = [&](size_t q, int n) -> float {
      const TileShape &c = seen[q];
      k->tiles[d] = c;
      if ((rc = build_start(k, d)) != PHX_OK) return -1.f;
      const PackView v = view(k, 1, d);
      launch_pack_flux(d, k->p.recon, k->p.riemann, k->p.geom.kind, k->dp, v, k->groups[d], c, s,
                       flux_fast_of(&k->p));
      (void)hipEventRecord(e0, s);
      for (int r = 0; r < n; ++r)
        launch_pack_flux(d, k->p.recon, k->p.riemann, k->p.geom.kind, k->dp, v, k->groups[d],
                         c, s, flux_fast_of(&k->p));
      (void)hipEventRecord(e1, s);
      if (hipEventSynchronize(e1) != hipSuccess) {
        rc = ctx->fail(PHX_ERR_DEVICE, "tuning launch failed");
        return -1.f;
      }
      float ms = 0.f;
      (void)hipEventElapsedTime(&ms, e0, e1);
      return ms;
    };
    // as phx_autotune: two passes, each candidate kept at its faster pass, then the fastest
    // four ranked by the median of three interleaved rounds of 4 x reps launches
    std::vector<float> t(seen.size(), 1e30f);
    for (int pass = 0; pass < 2 && rc == PHX_OK; ++pass)
      for (size_t q = 0; q < seen.size() && rc == PHX_OK; ++q) {
        const float ms = time_cand(q, reps);
        if (ms >= 0.f && ms < t[q]) t[q] = ms;
      }
    std::vector<size_t> order(seen.size());
    std::iota(order.begin(), order.end(), size_t(0));
    const size_t nref = std::min<size_t>(4, seen.size());
    std::partial_sort(order.begin(), order.begin() + nref, order.end(),
                      [&](size_t a, size_t c) { return t[a] < t[c]; });
    float m[4][3] = {};
    for (int r3 = 0; r3 < 3 && rc == PHX_OK; ++r3)
      for (size_t i = 0; i < nref && rc == PHX_OK; ++i) m[i][r3] = time_cand(order[i], 4 * reps);
    size_t best = order[0];
    float best_med = 1e30f;
    for (size_t i = 0; i < nref && rc == PHX_OK; ++i) {
      std::sort(m[i], m[i] + 3);
      if (m[i][1] < best_med) {
        best_med = m[i][1];
        best = order[i];
      }
    }
    k->tiles[d] = seen[best];
    if (rc == PHX_OK) rc = build_start(k, d);
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc != PHX_OK) return rc;
  return ctx->check_launch();
}

int phx_pack_tile_shape(const phx_pack *k, int dir, int *ti, int *ts) {
  if (!k || !ti || !ts || dir < 0 || dir > 2) return PHX_ERR_INVALID;
  *ti = k->tiles[dir].ti;
  *ts = k->tiles[dir].ts;
  return PHX_OK;
}

int phx_pack_graph_create(phx_ctx *ctx, phx_pack *k, int nstages,
                          const phx_pack_stage_args *stages, phx_stream stream,
                          phx_graph **out) {
  if (!ctx || !k || !out || nstages < 1 || !stages) return PHX_ERR_INVALID;
  *out = nullptr;
  if (!stream) return ctx->fail(PHX_ERR_INVALID, "graph capture needs a non-default stream");
  for (int i = 0; i < nstages; ++i) {
    const int rc = check_stage_args(ctx, &stages[i]);
    if (rc) return rc;
  }
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  (void)hipGetLastError();
  if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess)
    return ctx->fail(PHX_ERR_DEVICE, "stream capture failed to start");
  ctx->capturing = true;
  for (int i = 0; i < nstages; ++i) launch_stage(ctx, k, &stages[i], s);
  ctx->capturing = false;
  hipGraph_t g = nullptr;
  const hipError_t ec = hipStreamEndCapture(s, &g);
  const hipError_t el = hipGetLastError();
  if (ec != hipSuccess || el != hipSuccess || !g) {
    if (g) (void)hipGraphDestroy(g);
    return ctx->fail(PHX_ERR_DEVICE, "stage capture failed");
  }
  phx_graph *G = new (std::nothrow) phx_graph();
  if (!G) {
    (void)hipGraphDestroy(g);
    return PHX_ERR_NOMEM;
  }
  G->graph = g;
  if (hipGraphInstantiate(&G->exec, g, nullptr, nullptr, 0) != hipSuccess) {
    phx_graph_destroy(G);
    return ctx->fail(PHX_ERR_DEVICE, "graph instantiation failed");
  }
  G->pack = k;
  ++k->graphs;
  *out = G;
  return PHX_OK;
}

int phx_graph_launch(phx_graph *G, phx_stream stream) {
  if (!G || !G->exec) return PHX_ERR_INVALID;
  return hipGraphLaunch(G->exec, reinterpret_cast<hipStream_t>(stream)) == hipSuccess
             ? PHX_OK
             : PHX_ERR_DEVICE;
}

int phx_graph_destroy(phx_graph *G) {
  if (!G) return PHX_OK;
  if (G->exec) (void)hipGraphExecDestroy(G->exec);
  if (G->graph) (void)hipGraphDestroy(G->graph);
  if (G->pack) --G->pack->graphs;
  delete G;
  return PHX_OK;
}

}  // extern "C"
