// phx_internal.hpp -- the library context (opaque phx_ctx of include/phx.h).
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <string>
#include <tuple>
#include <type_traits>

#include "phx.h"
#include "phx_fixup.h"

#include "phx_device.hpp"

struct phx_ctx;
struct phx_mlp;

namespace phx {

inline BlockGeom make_block_geom(const phx_block *b) {
  BlockGeom g;
  g.ndim = b->ndim;
  g.ng = b->ng;
  int n[3];
  for (int d = 0; d < 3; ++d) {
    const bool act = d < b->ndim;
    n[d] = b->nx[d] + (act ? 2 * b->ng : 0);
    g.istart[d] = act ? b->ng : 0;
    g.is[d] = g.istart[d];
    g.ie[d] = g.istart[d] + b->nx[d] - 1;
    g.xmin[d] = b->xmin[d];
    g.dx[d] = b->dx[d];
  }
  g.ni = n[0];
  g.nj = n[1];
  g.nk = n[2];
  g.ncell = (long)n[0] * n[1] * n[2];
  g.geo = nullptr;
  return g;
}

// Run f with the geometry (resp. MHD flag) as a compile-time constant:
// with_geom(kind, [&](auto G) { ... decltype(G)::value ... }).
template <int G>
using GeomTag = std::integral_constant<int, G>;
template <class F>
inline void with_geom(int kind, F &&f) {
  switch (kind) {
  case PHX_GEOM_MINKOWSKI: f(GeomTag<PHX_GEOM_MINKOWSKI>()); break;
  case PHX_GEOM_SPHERICAL_KS: f(GeomTag<PHX_GEOM_SPHERICAL_KS>()); break;
  default: f(GeomTag<PHX_GEOM_FMKS>()); break;
  }
}
template <class F>
inline void with_mhd(int mhd, F &&f) {
  if (mhd)
    f(std::true_type());
  else
    f(std::false_type());
}

// shape of the y/z flux tiles: ti consecutive i by ts cells along the sweep direction
// (0 = the built-in default); results do not depend on it
struct TileShape {
  int ti = 0, ts = 0;
};

// flux launches (phx_flux.hip): direction d of fluid.cpp:753-770 with the fused
// reconstruction; writes flux[d][m][cells] and fsig[d][cells].  fast selects the build of
// the sweeps with the fast flux arithmetic (phx_fluid_params::flux_arith).
void launch_fluxes(int d, int recon, int riemann, int mhd, int geom, const DevParams &dp,
                   const BlockGeom &bg, const PrimPtrs &prim, const FluxPtrs &flux, double *fsig,
                   hipStream_t s, TileShape shape, bool fast);
inline bool flux_fast_of(const phx_fluid_params *p) { return p->flux_arith == PHX_FLUX_ARITH_FAST; }
// the shape launch_fluxes uses for direction d (after defaults and limits)
TileShape resolve_tile_shape(int d, int recon, int mhd, const BlockGeom &bg, TileShape want);

DevParams make_dev_params(const phx_fluid_params *p);

// ---- per-component pointers of a call's fields -----------------------------------------
// component v of a contiguous field at base, unless the table has its own pointer
template <int N>
inline Comp<N> comps(const double *base, long ncell, int nv, double *const *table) {
  Comp<N> c = Comp<N>::contiguous(base, ncell, nv);
  if (table)
    for (int v = 0; v < nv && v < N; ++v)
      if (table[v]) c.p[v] = table[v];
  return c;
}
inline PrimPtrs prim_ptrs(const phx_block_fields *f, long ncell, int nprim) {
  return comps<10>(f->prim, ncell, nprim, f->vars ? f->vars->prim : nullptr);
}
inline ConsPtrs cons_ptrs(const phx_block_fields *f, long ncell, int ncons) {
  return comps<8>(f->cons, ncell, ncons, f->vars ? f->vars->cons : nullptr);
}
inline ConsPtrs src_ptrs(const phx_block_fields *f, long ncell, int ncons) {
  return comps<8>(f->src, ncell, ncons, f->vars ? f->vars->src : nullptr);
}
inline FluxPtrs flux_ptrs(const phx_block_fields *f, long ncell, int ncons) {
  FluxPtrs c = FluxPtrs::flux(f->flux, ncell, ncons);
  if (f->vars)
    for (int d = 0; d < 3; ++d)
      for (int m = 0; m < ncons; ++m)
        if (f->vars->flux[d][m]) c.p[d * 8 + m] = f->vars->flux[d][m];
  return c;
}
// every one of the first n components present
template <int N>
inline bool complete(const Comp<N> &c, int n) {
  for (int v = 0; v < n && v < N; ++v)
    if (!c.p[v]) return false;
  return true;
}
inline bool complete_flux(const FluxPtrs &c, int ndim, int ncons) {
  for (int d = 0; d < ndim; ++d)
    for (int m = 0; m < ncons; ++m)
      if (!c.p[d * 8 + m]) return false;
  return true;
}

// a cell range lo..lo+n-1 per dimension (i, j, k)
struct Range3 {
  int lo[3], n[3];
};
// validates lo..hi (inclusive; NULL = the entire block) against the block
int make_range(phx_ctx *ctx, const phx_block *b, const int lo[3], const int hi[3], Range3 &rg);

// the fixup package's switches after fixup::Initialize's normalisation (fixup.cpp:102-135)
struct DevFixup {
  int c2p_fixup, strategy, floors, mhd_floors;
  double bsqorho_max, bsqou_max;
};
DevFixup make_dev_fixup(const phx_fluid_params *p, const phx_fixup_params *fx);
int check_fixup(phx_ctx *ctx, const phx_fixup_params *fx);
int check_params(phx_ctx *ctx, const phx_fluid_params *p);
int check_block(phx_ctx *ctx, const phx_block *b, int recon);

void launch_flux_ct(const BlockGeom &bg, const FluxPtrs &flux, double *emf, hipStream_t s);
// FluxCT's corner EMFs alone (the fused updates form the rewritten B fluxes from them)
void launch_ct_emf(const BlockGeom &bg, const FluxPtrs &flux, double *emf, hipStream_t s);
void launch_sources(int geom, int mhd, const DevParams &dp, const BlockGeom &bg,
                    const PrimPtrs &prim, const ConsPtrs &src, hipStream_t s);
void launch_dt_finish(const double *partial, int np, double cfl, double *dt, hipStream_t s);

// ---- block packs (a MeshBlockPack-style work list over blocks of different sizes) -------
// One entry per block: its geometry and field pointers for the stage being run.
struct PackEntry {
  BlockGeom bg;
  const double *prim;       // reconstructed primitives (read by the flux sweeps)
  double *flux, *fsig, *emf, *src;
  double *u_stage;          // U_s (averaged in place)
  const double *u_base;     // U^n
  double *u_out;            // U_{s+1}
  double *prim_out, *temp, *mu, *csig, *fail;
  int32_t *iters;
};
// A launch over a pack: workgroup g belongs to the block b with start[b] <= g < start[b+1].
struct PackView {
  const PackEntry *e;
  const long *start;  // [n + 1] prefix sums of workgroups per block for this launch
  int n;
};

__device__ __forceinline__ int pack_find(const long *start, int n, long g) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (start[mid] <= g)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

// pack launches (phx_flux.hip, phx_kernels.hip, phx_stage.hip); 3-D MHD blocks
void launch_pack_flux(int d, int recon, int riemann, int geom, const DevParams &dp,
                      const PackView &v, long groups, TileShape shape, hipStream_t s, bool fast);
// the two builds of the sweeps (phx_flux_sweeps.hpp in phx_flux.hip and phx_flux_fast.hip)
namespace flux_exact {
void launch_fluxes(int d, int recon, int riemann, int mhd, int geom, const DevParams &dp,
                   const BlockGeom &bg, const PrimPtrs &prim, const FluxPtrs &flux, double *fsig,
                   hipStream_t s, TileShape sh);
void launch_pack_flux(int d, int recon, int riemann, int geom, const DevParams &dp,
                      const PackView &v, long groups, TileShape sh, hipStream_t s);
}  // namespace flux_exact
namespace flux_fast {
void launch_fluxes(int d, int recon, int riemann, int mhd, int geom, const DevParams &dp,
                   const BlockGeom &bg, const PrimPtrs &prim, const FluxPtrs &flux, double *fsig,
                   hipStream_t s, TileShape sh);
void launch_pack_flux(int d, int recon, int riemann, int geom, const DevParams &dp,
                      const PackView &v, long groups, TileShape sh, hipStream_t s);
}  // namespace flux_fast
void launch_pack_ct_emf(const PackView &v, long groups, hipStream_t s);
void launch_pack_sources(int geom, const DevParams &dp, const PackView &v, long groups,
                         hipStream_t s);
// The cells a plain con2prim pass leaves to the learned bracket: those with no previous
// root (mu <= 0: initial data, fresh cells).  Entries are (block << 32) | (keep << 31) |
// cell, keep = 1 for a cell whose iteration count is not to be written (an interior cell
// of an entire-block re-inversion); count is cleared (launch_zero_count) before the
// producing launch.  A cell that finds the queue full is inverted by the plain pass.
struct ColdQueue {
  unsigned long long *cell = nullptr;
  unsigned int *count = nullptr;
  long cap = 0;
};
// called by every active lane of a wave: queues this lane's cell when cold; true when the
// cell is queued (the learned pass inverts it)
__device__ __forceinline__ bool enqueue_cold(const ColdQueue &q, bool cold, int blk, long c,
                                             bool keep) {
  const unsigned long long mask = __ballot(cold);
  if (!mask) return false;
  const int lane = threadIdx.x & 63;
  const int leader = __ffsll((long long)mask) - 1;
  unsigned int base = 0;
  if (lane == leader) base = atomicAdd(q.count, (unsigned int)__popcll(mask));
  base = __shfl(base, leader, 64);
  const unsigned int slot = base + (unsigned int)__popcll(mask & ((1ull << lane) - 1ull));
  if (!cold || slot >= q.cap) return false;
  q.cell[slot] = ((unsigned long long)blk << 32) | ((unsigned long long)keep << 31) |
                 (unsigned long long)c;
  return true;
}
// clears a queue's count and, when extra is set, sets *extra = DBL_MAX
void launch_zero_count(unsigned int *count, hipStream_t s, double *extra = nullptr);
// which cells the fused update inverts: all (the plain stage), or those with a previous
// root (mu > 0) to start from, queueing the rest for the learned bracket (cq)
enum { kC2pAll = 1, kC2pWarm = 2 };
// c2p_mode == kC2pWarm: the queue is cleared first (and *extra set to DBL_MAX, the
// partial the learned pass folds its cells' CFL times into); the update's partials then
// cover the cells it inverts
void launch_pack_update(int geom, const DevParams &dp, const PackView &v, long groups,
                        double beta, double dt, const double *dt_dev, double *partial,
                        hipStream_t s, int c2p_mode = kC2pAll, const ColdQueue &cq = ColdQueue{},
                        double *extra = nullptr);
// con2prim with the learned bracket over the queued cells of a pack (phx_mlp.hip), in a
// grid-stride loop over the count the producing pass left; counters (nullable)
// accumulate [bracket misses, network-bracketed solves]; extra (nullable, interior cells
// only): the minimum of the cells' CFL times, by an atomic minimum on the bits
void launch_pack_c2p_cold(int geom, const DevParams &dp, const phx_mlp *m, const PackView &v,
                          const ColdQueue &cq, unsigned long long *counters, double *extra,
                          hipStream_t s);
// the fixups over a pack (phx_fixup.hip): con2prim failure repair on the interior cells,
// then floors and ceilings (interior cells, or the entire blocks over an entire-cell work
// list) folded with the interior CFL partial minima (partial may be null)
void launch_pack_c2p_fixup(int geom, const DevParams &dp, const DevFixup &fx, const PackView &v,
                           long groups, hipStream_t s);
// The floors run as a check pass over every cell and a repair pass over the cells it
// queues (the rare ones a floor or ceiling acts on), so the check pass is not sized for the
// repair's con2prim.  The queue holds (block << 32 | cell) and the check pass's workgroup
// (whose CFL partial the repaired cell joins; -1 for none); cap >= groups * 256.
struct FloorQueue {
  unsigned long long *cell = nullptr;
  int *grp = nullptr;
  unsigned int *count = nullptr;
  long cap = 0;
};
void launch_pack_floors(int geom, const DevParams &dp, const DevFixup &fx, const PackView &v,
                        long groups, double *partial, hipStream_t s, bool entire,
                        const FloorQueue &q);
// fills a block's cached geometry plane (FMKS; kGeoPlanes x ni x nj doubles at geo)
void launch_geo_fill(const DevParams &dp, const BlockGeom &bg, double *geo, hipStream_t s);
// groups per block for each launch kind
long pack_groups_flux(int d, int recon, const BlockGeom &bg, TileShape shape);
long pack_groups_cells(const BlockGeom &bg, bool corners);

// partial minima of the CFL reduction (one per 256-cell block of the fused stage kernel,
// so one call covers up to 16.7 M cells; the separate reduction strides beyond that)
constexpr int kDtPartials = 1 << 16;
// the allocation-free stage's device scratch (kDtPartials doubles in a file-scope device
// array, phx_stage.hip): its address on the current device, looked up once per device
int static_scratch(double **out);
// the message of the calling thread's last failed phx_fluid_stage_noalloc ("" if none)
const char *noalloc_last_error();
}  // namespace phx

struct phx_ctx {
  // optional per-launch timing: HIP events recorded around each launch of a stage
  struct TimedLaunch {
    const char *name;
    hipEvent_t start, stop;
  };
  static constexpr int kMaxTimed = 64;
  bool profiling = false;
  bool capturing = false;  // a graph capture is recording: no profiling events
  int n_timed = 0;
  TimedLaunch timed[kMaxTimed] = {};

  int device = 0;
  hipStream_t stream = nullptr;  // default stream of the context
  // autotuned y/z flux tile shapes per (recon, riemann, mhd, geom, ndim, nx, ny, nz, ng, d,
  // flux arithmetic)
  using TileKey = std::tuple<int, int, int, int, int, int, int, int, int, int, int>;
  std::map<TileKey, phx::TileShape> tiles;
  static TileKey tile_key(const phx_fluid_params *p, const phx_block *b, int d) {
    return TileKey(p->recon, p->riemann, p->mhd, p->geom.kind, b->ndim, b->nx[0], b->nx[1],
                   b->nx[2], b->ng, d, p->flux_arith);
  }
  phx::TileShape tile_for(const phx_fluid_params *p, const phx_block *b, int d) const {
    const auto it = tiles.find(tile_key(p, b, d));
    return it == tiles.end() ? phx::TileShape() : it->second;
  }
  double *dt_partial = nullptr;  // [kDtPartials]
  std::string last_error;

  hipStream_t stream_of(phx_stream s) const {
    return s ? reinterpret_cast<hipStream_t>(s) : stream;
  }
  int fail(int code, const std::string &msg) {
    last_error = msg;
    return code;
  }
  // events bracketing one launch on stream s (no-op unless profiling)
  void mark_begin(const char *name, hipStream_t s) {
    if (!profiling || capturing || n_timed >= kMaxTimed) return;
    timed[n_timed].name = name;
    (void)hipEventRecord(timed[n_timed].start, s);
  }
  void mark_end(hipStream_t s) {
    if (!profiling || capturing || n_timed >= kMaxTimed) return;
    (void)hipEventRecord(timed[n_timed].stop, s);
    ++n_timed;
  }
  int check_launch() {
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(PHX_ERR_DEVICE, hipGetErrorString(e));
    return PHX_OK;
  }
};
