// phx_mesh.hip -- a uniform mesh of equal meshblocks on one device (include/phx_mesh.h):
// the ghost-zone exchange, physical boundaries and boundary conversion as kernels over all
// blocks at once, and the RK2 driver that sequences them with the pack's stage launches.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cstring>
#include <new>
#include <vector>

#include "phx.h"
#include "phx_cells.hpp"
#include "phx_device.hpp"
#include "phx_internal.hpp"
#include "phx_mesh.h"
#include "phx_pack_impl.hpp"

namespace phx {

// the mesh as the kernels see it
struct MeshInfo {
  int nb[3];        // blocks per dimension
  int bnx[3];       // interior cells per block
  int N[3];         // mesh cells
  int ng;
  int periodic[3];  // periodic dimension
  int inflow;       // outer-x1 inflow check (FMKS)
};

__device__ __forceinline__ void block_loc(const MeshInfo &m, int gid, int lx[3]) {
  lx[0] = gid % m.nb[0];
  lx[1] = (gid / m.nb[0]) % m.nb[1];
  lx[2] = gid / (m.nb[0] * m.nb[1]);
}

// the FillGhost fields of a cell (bc_vars = primitive, fluid.cpp:159-165): primitives,
// temperature, con2prim root, fail flag and the conserved variables of the container
constexpr int kNprim = 10, kNcons = 8;
// (all 21 values are read before any is written: the source and destination fields share
// one allocation as far as the compiler can tell, so interleaved copies would each wait a
// full memory round trip)
__device__ __forceinline__ void copy_ghost_fields(const PackEntry &D, long cd, const PackEntry &S,
                                                  long cs) {
  const long nd = D.bg.ncell, ns = S.bg.ncell;
  double pr[kNprim], u[kNcons];
#pragma unroll
  for (int v = 0; v < kNprim; ++v) pr[v] = S.prim_out[v * ns + cs];
#pragma unroll
  for (int v = 0; v < kNcons; ++v) u[v] = S.u_out[v * ns + cs];
  const double t = D.temp ? S.temp[cs] : 0.0, m = S.mu[cs], f = S.fail[cs];
#pragma unroll
  for (int v = 0; v < kNprim; ++v) D.prim_out[v * nd + cd] = pr[v];
#pragma unroll
  for (int v = 0; v < kNcons; ++v) D.u_out[v * nd + cd] = u[v];
  if (D.temp) D.temp[cd] = t;
  D.mu[cd] = m;
  D.fail[cd] = f;
}

// boundary exchange between blocks of one level (parthenon's same-level ghost fill): a
// ghost cell whose mesh index lies inside the mesh, after the periodic wrap, takes the
// fields of the block that owns that index; the others are left to the physical boundaries
__global__ __launch_bounds__(256) void k_mesh_exchange(MeshInfo m, PackView v) {
  const long g = xcd_remap(blockIdx.x, gridDim.x);
  const int b = pack_find(v.start, v.n, g);
  const PackEntry &E = v.e[b];
  int idx[3];
  if (!ghost_cell(E.bg, (g - v.start[b]) * 256 + threadIdx.x, idx)) return;
  int lx[3];
  block_loc(m, b, lx);
  int src[3], sl[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    int gi = lx[d] * m.bnx[d] + idx[d] - m.ng;
    if (gi < 0 || gi >= m.N[d]) {
      if (!m.periodic[d]) return;
      gi = gi < 0 ? gi + m.N[d] : gi - m.N[d];
    }
    src[d] = gi / m.bnx[d];
    sl[d] = gi % m.bnx[d] + m.ng;
  }
  const PackEntry &S = v.e[src[0] + m.nb[0] * (src[1] + m.nb[1] * src[2])];
  const long cd = ((long)idx[2] * E.bg.nj + idx[1]) * E.bg.ni + idx[0];
  const long cs = ((long)sl[2] * S.bg.nj + sl[1]) * S.bg.ni + sl[0];
  copy_ghost_fields(E, cd, S, cs);
}

// gr_outflow with bc_vars = primitive (OutflowInnerX1 / OuterX1, phoebus_boundaries.cpp:
// 118-152, 222-262; GenericBC outflow for x2, x3, :43-116): the ghost slab of a physical
// face of dimension d copies the first (last) interior layer, all FillGhost fields, over
// the entire extent of the other dimensions
__global__ __launch_bounds__(256) void k_mesh_outflow(MeshInfo m, PackView v, int d, int inner,
                                                      int outer) {
  const long g = xcd_remap(blockIdx.x, gridDim.x);
  const int b = pack_find(v.start, v.n, g);
  int lx[3];
  block_loc(m, b, lx);
  // a block on neither physical face of d has nothing to do (the whole workgroup)
  if (!(inner && lx[d] == 0) && !(outer && lx[d] == m.nb[d] - 1)) return;
  const PackEntry &E = v.e[b];
  int idx[3];
  if (!ghost_cell(E.bg, (g - v.start[b]) * 256 + threadIdx.x, idx)) return;
  int ref[3] = {idx[0], idx[1], idx[2]};
  if (idx[d] < E.bg.is[d] && inner && lx[d] == 0)
    ref[d] = E.bg.is[d];
  else if (idx[d] > E.bg.ie[d] && outer && lx[d] == m.nb[d] - 1)
    ref[d] = E.bg.ie[d];
  else
    return;
  const long c = ((long)idx[2] * E.bg.nj + idx[1]) * E.bg.ni + idx[0];
  const long cr = ((long)ref[2] * E.bg.nj + ref[1]) * E.bg.ni + ref[0];
  copy_ghost_fields(E, c, E, cr);
}

// ConvertBoundaryConditions (phoebus_boundaries.cpp:303-448) for bc_vars = primitive: on
// the ghost slabs of every face that is not a block boundary (physical or periodic mesh
// faces), the FMKS outer-x1 inflow check (:356-380: no u^1 < 0 inflow from the outer
// boundary), then p2c into the container and the cell signal speeds
template <int GEOM>
__global__ __launch_bounds__(256) void k_mesh_convert(DevParams p, MeshInfo m, PackView v) {
  using M = typename MetricOf<GEOM>::type;
  const long g = xcd_remap(blockIdx.x, gridDim.x);
  const int b = pack_find(v.start, v.n, g);
  const PackEntry &E = v.e[b];
  const BlockGeom &bg = E.bg;
  int idx[3];
  if (!ghost_cell(bg, (g - v.start[b]) * 256 + threadIdx.x, idx)) return;
  int lx[3];
  block_loc(m, b, lx);
  bool slab = false, outer_x1 = false;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    if (idx[d] < bg.is[d] && lx[d] == 0) slab = true;
    if (idx[d] > bg.ie[d] && lx[d] == m.nb[d] - 1) {
      slab = true;
      if (d == 0) outer_x1 = true;
    }
  }
  if (!slab) return;
  const long n = bg.ncell, c = ((long)idx[2] * bg.nj + idx[1]) * bg.ni + idx[0];
  if constexpr (GEOM == PHX_GEOM_FMKS) {
    if (outer_x1 && m.inflow) {
      const M gm = metric_at<GEOM, true>(p, bg, 0, (long)idx[1] * bg.ni + idx[0],
                                         xc(bg, 0, idx[0]), xc(bg, 1, idx[1]), xc(bg, 2, idx[2]));
      double vcon[3] = {E.prim_out[n + c], E.prim_out[2 * n + c], E.prim_out[3 * n + c]};
      double W = lorentz_factor(gm, vcon);
      const double alpha = gm.alpha(), beta0 = gm.beta(0);
      const double ucon1 = vcon[0] - W * beta0 / alpha;
      if (ucon1 < 0) {
#pragma unroll
        for (int ii = 0; ii < 3; ++ii) vcon[ii] /= W;
        vcon[0] = beta0 / alpha;
        double vsq = 0.;
#pragma unroll
        for (int ii = 0; ii < 3; ++ii)
#pragma unroll
          for (int jj = 0; jj < 3; ++jj)
            if (M::nz_cov(ii + 1, jj + 1)) vsq += gm.gcov(ii + 1, jj + 1) * vcon[ii] * vcon[jj];
        W = 1. / sqrt(1. - vsq);
#pragma unroll
        for (int ii = 0; ii < 3; ++ii) E.prim_out[(1 + ii) * n + c] = W * vcon[ii];
      }
    }
  }
  prim_to_con_cell<true, GEOM, GEOM == PHX_GEOM_FMKS>(
      p, bg, idx[0], idx[1], idx[2], contig(E.prim_out, bg.ncell), contig(E.u_out, bg.ncell),
      E.csig);
}

// FixFluxes for gr_outflow (fixup.cpp:759-1062): no mass inflow through a physical face,
// F(D) <= 0 on inner faces and >= 0 on outer ones, over the face's interior extent; over
// the face-adjacent ghost layer, whose cell (is-1, j, k) stands for the inner x1 face
// stored at (is, j, k) and (ie+1, j, k) for the outer one (likewise in x2, x3)
__global__ __launch_bounds__(256) void k_mesh_fix_fluxes(MeshInfo m, PackView v,
                                                         int outflow_mask) {
  const long g = xcd_remap(blockIdx.x, gridDim.x);
  const int b = pack_find(v.start, v.n, g);
  const PackEntry &E = v.e[b];
  const BlockGeom &bg = E.bg;
  int idx[3];
  if (!face_layer_cell(bg, (g - v.start[b]) * 256 + threadIdx.x, idx)) return;
  int lx[3];
  block_loc(m, b, lx);
  const long n = bg.ncell, sd[3] = {1, bg.ni, (long)bg.ni * bg.nj};
  const long c = ((long)idx[2] * bg.nj + idx[1]) * bg.ni + idx[0];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    double *F = E.flux + (long)d * kNcons * n;  // the density flux of direction d
    if (idx[d] == bg.is[d] - 1 && (outflow_mask >> (2 * d) & 1) && lx[d] == 0)
      F[c + sd[d]] = smin(F[c + sd[d]], 0.0);
    if (idx[d] == bg.ie[d] + 1 && (outflow_mask >> (2 * d + 1) & 1) && lx[d] == m.nb[d] - 1)
      F[c] = smax(F[c], 0.0);
  }
}

// the robust con2prim over a pack's blocks on the stage's output container, over the
// cells of work-list kind 5 (entire blocks) or 6 (the face-adjacent ghost layer, what the
// c2p fixup reads); a re-inversion of the entire blocks keeps the interior's iteration
// counts (those of the stage's inversion of U_{s+1})
// (cq.cell set: the learned bracket is on, and a cell with no previous root goes to its
// queue instead)
template <int GEOM, int KIND>
__global__ __launch_bounds__(256) void k_pack_con2prim(DevParams p, PackView v, ColdQueue cq) {
  const long g = xcd_remap(blockIdx.x, gridDim.x);
  const int b = pack_find(v.start, v.n, g);
  const PackEntry &E = v.e[b];
  const BlockGeom &bg = E.bg;
  int idx[3];
  if (!cell_of_kind<KIND>(bg, (g - v.start[b]) * 256 + threadIdx.x, idx)) return;
  int32_t *it = (KIND == 5 && is_interior(bg, idx)) ? nullptr : E.iters;
  if (cq.cell) {
    const long c = ((long)idx[2] * bg.nj + idx[1]) * bg.ni + idx[0];
    if (enqueue_cold(cq, !(E.mu[c] > 0.0), b, c, it == nullptr)) return;
  }
  con2prim_cell<true, GEOM, GEOM == PHX_GEOM_FMKS>(p, bg, idx[0], idx[1], idx[2],
                                                   contig(E.u_out, bg.ncell),
                                                   contig(E.prim_out, bg.ncell),
                                                   E.temp, E.mu, E.csig, E.fail,
                                                   it);
}

}  // namespace phx

using namespace phx;

struct phx_mesh {
  phx_ctx *ctx = nullptr;
  phx_fluid_params p{};
  phx_fixup_params fxp{};
  phx_mesh_desc d{};
  MeshInfo mi{};
  int nblocks = 0;
  long interior = 0;
  std::vector<phx_block> blocks;
  std::vector<phx_pack_block> pb;
  double *arena = nullptr;
  int32_t *iters = nullptr;
  phx_pack *pack = nullptr;
  double *dt_dev = nullptr;
  double time = 0.0, dt = 0.0, tlim = 1e300;
  long ncycle = 0;
  bool has_face_speeds = false;
  int outflow_mask = 0;  // bit 2d + side: a gr_outflow face
};

namespace {

constexpr int kFields = 70;  // doubles per cell of a block (see carve)

// the per-block fields out of the arena: prim 10, cons 8, cons1 8, temp, mu, fail,
// cell_sig 3, face_sig 3, flux 24, emf 3, src 8
void carve(double *base, long nc, phx_pack_block &B) {
  double *q = base;
  auto take = [&](int nv) {
    double *r = q;
    q += (long)nv * nc;
    return r;
  };
  B.f.prim = take(10);
  B.f.cons = take(8);
  B.cons1 = take(8);
  B.f.temp = take(1);
  B.f.c2p_mu = take(1);
  B.f.fail = take(1);
  B.f.cell_sig = take(3);
  B.f.face_sig = take(3);
  B.f.flux = take(24);
  B.f.emf = take(3);
  B.f.src = take(8);
}

// the pack launches of a mesh run on its stage roles: stage 1 writes U_1 ("1" container),
// stage 2 writes U^{n+1} (base)
PackView pv(const phx_mesh *m, int st, int kind) { return view(m->pack, st, kind); }

long groups(const phx_mesh *m, int kind) { return m->pack->groups[kind]; }

void launch_exchange_and_boundaries(phx_mesh *m, int st, hipStream_t s) {
  phx_ctx *ctx = m->ctx;
  // every kernel here writes ghost cells only: a work list of them
  const PackView v = pv(m, st, 7);
  const unsigned gr = (unsigned)groups(m, 7);
  ctx->mark_begin("exchange", s);
  hipLaunchKernelGGL(k_mesh_exchange, dim3(gr), dim3(256), 0, s, m->mi, v);
  ctx->mark_end(s);
  // ApplyBoundaryConditions: faces in the order inner/outer x1, x2, x3
  for (int d = 0; d < 3; ++d) {
    const int inner = m->outflow_mask >> (2 * d) & 1, outer = m->outflow_mask >> (2 * d + 1) & 1;
    if (!inner && !outer) continue;
    ctx->mark_begin("outflow", s);
    hipLaunchKernelGGL(k_mesh_outflow, dim3(gr), dim3(256), 0, s, m->mi, v, d, inner, outer);
    ctx->mark_end(s);
  }
  ctx->mark_begin("convert", s);
  with_geom(m->p.geom.kind, [&](auto G) {
    hipLaunchKernelGGL((k_mesh_convert<decltype(G)::value>), dim3(gr), dim3(256), 0, s,
                       m->pack->dp, m->mi, v);
  });
  ctx->mark_end(s);
}

// the learned-bracket pass over the cells queued since the queue was last cleared
void launch_c2p_queued(phx_mesh *m, int st, const char *name, hipStream_t s) {
  phx_pack *k = m->pack;
  m->ctx->mark_begin(name, s);
  launch_pack_c2p_cold(m->p.geom.kind, k->dp, k->mlp, pv(m, st, 5), k->cq, k->counters, nullptr,
                       s);
  m->ctx->mark_end(s);
}

// the robust con2prim of a FillDerived over the stage's output container, on the cells of
// work-list kind 5 (entire blocks) or 6 (the face-adjacent ghost layer).  With the learned
// bracket the cells with no previous root are queued instead and the learned pass runs
// over the queue; append: the queue already holds this FillDerived's cells of the update
// (it is not cleared, and the caller runs the learned pass)
void launch_c2p(phx_mesh *m, int st, int kind, hipStream_t s, bool append = false) {
  phx_pack *k = m->pack;
  const ColdQueue cq = k->mlp ? k->cq : ColdQueue{};
  m->ctx->mark_begin(kind == 5 ? "con2prim" : "con2prim_ghost", s);
  if (cq.cell && !append) launch_zero_count(cq.count, s);
  with_geom(m->p.geom.kind, [&](auto G) {
    constexpr int GM = decltype(G)::value;
    auto kernel = kind == 5 ? k_pack_con2prim<GM, 5> : k_pack_con2prim<GM, 6>;
    hipLaunchKernelGGL(kernel, dim3((unsigned)groups(m, kind)), dim3(256), 0, s, k->dp,
                       pv(m, st, kind), cq);
  });
  m->ctx->mark_end(s);
  if (cq.cell && !append) launch_c2p_queued(m, st, "con2prim_learned", s);
}

// FillDerived + ConservedToPrimitiveFixup + ApplyFloors after the boundary exchange;
// partial: the CFL partial minima (after the last stage) or null
void launch_fill_derived_entire(phx_mesh *m, int st, double *partial, hipStream_t s) {
  phx_pack *k = m->pack;
  launch_c2p(m, st, 5, s);
  m->ctx->mark_begin("c2p_fixup", s);
  launch_pack_c2p_fixup(k->p.geom.kind, k->dp, k->fx, pv(m, st, 4), k->groups[4], s);
  m->ctx->mark_end(s);
  m->ctx->mark_begin("floors", s);
  launch_pack_floors(k->p.geom.kind, k->dp, k->fx, pv(m, st, 5), k->groups[5], partial, s, true,
                     k->fq);
  m->ctx->mark_end(s);
}

int run_stage(phx_mesh *m, int st, bool last, hipStream_t s) {
  phx_ctx *ctx = m->ctx;
  phx_pack *k = m->pack;
  launch_stage_fluxes(ctx, k, st, s);
  if (m->d.flux_fixup && m->outflow_mask) {
    ctx->mark_begin("fix_fluxes", s);
    hipLaunchKernelGGL(k_mesh_fix_fluxes, dim3((unsigned)groups(m, 6)), dim3(256), 0, s, m->mi,
                       pv(m, st, 6), m->outflow_mask);
    ctx->mark_end(s);
  }
  const bool learned = k->mlp != nullptr;
  ctx->mark_begin(learned ? "update" : "update_con2prim", s);
  launch_pack_update(k->p.geom.kind, k->dp, pv(m, st, 4), k->groups[4], st == 1 ? 1.0 : 0.5,
                     m->dt, nullptr, nullptr, s, learned ? kC2pWarm : kC2pAll, k->cq);
  ctx->mark_end(s);
  // the first FillDerived: the interior (inverted by the update) and the face-adjacent
  // ghost layer the fixup reads; with the learned bracket, the cells of both with no
  // previous root go through one learned pass
  launch_c2p(m, st, 6, s, learned);
  if (learned) launch_c2p_queued(m, st, "con2prim_learned_first", s);
  ctx->mark_begin("c2p_fixup", s);
  launch_pack_c2p_fixup(k->p.geom.kind, k->dp, k->fx, pv(m, st, 4), k->groups[4], s);
  ctx->mark_end(s);
  ctx->mark_begin("floors", s);
  launch_pack_floors(k->p.geom.kind, k->dp, k->fx, pv(m, st, 4), k->groups[4], nullptr, s,
                     false, k->fq);
  ctx->mark_end(s);
  launch_exchange_and_boundaries(m, st, s);
  launch_fill_derived_entire(m, st, last ? k->partial : nullptr, s);
  if (last) {
    ctx->mark_begin("dt_finish", s);
    launch_dt_finish(k->partial, (int)k->groups[5], m->d.cfl, m->dt_dev, s);
    ctx->mark_end(s);
  }
  return ctx->check_launch();
}

// EstimateTimestep over all blocks (fluid.cpp:899-944), first call without face speeds
double estimate_dt(phx_mesh *m, hipStream_t s) {
  double lo = DBL_MAX;
  for (int b = 0; b < m->nblocks; ++b) {
    if (phx_estimate_dt(m->ctx, &m->p, &m->blocks[b], &m->pb[b].f, m->has_face_speeds ? 1 : 0,
                        m->d.cfl, m->dt_dev, s) != PHX_OK)
      return -1.0;
    double h = -1.0;
    if (hipStreamSynchronize(s) != hipSuccess ||
        hipMemcpy(&h, m->dt_dev, sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
      return -1.0;
    lo = std::min(lo, h);
  }
  m->has_face_speeds = true;
  return lo;
}

int check_desc(phx_ctx *ctx, const phx_fluid_params *p, const phx_mesh_desc *d) {
  if (!d) return ctx->fail(PHX_ERR_INVALID, "null mesh description");
  if (!p->mhd) return ctx->fail(PHX_ERR_UNSUPPORTED, "meshes hold 3-D MHD blocks");
  for (int e = 0; e < 3; ++e) {
    if (d->nx[e] < 2 || d->block_nx[e] < 1 || d->nx[e] % d->block_nx[e] != 0)
      return ctx->fail(PHX_ERR_INVALID, "block sizes must divide the 3-D mesh");
    if (d->block_nx[e] < d->ng)
      return ctx->fail(PHX_ERR_INVALID, "blocks must be at least nghost cells wide");
    if (!(d->xmax[e] > d->xmin[e])) return ctx->fail(PHX_ERR_INVALID, "empty mesh extent");
    const int a = d->bc[2 * e], b = d->bc[2 * e + 1];
    if ((a != PHX_BC_OUTFLOW && a != PHX_BC_PERIODIC) || (b != PHX_BC_OUTFLOW && b != PHX_BC_PERIODIC))
      return ctx->fail(PHX_ERR_INVALID, "invalid boundary kind");
    if ((a == PHX_BC_PERIODIC) != (b == PHX_BC_PERIODIC))
      return ctx->fail(PHX_ERR_INVALID, "periodic boundaries come in pairs");
  }
  if (!(d->cfl > 0.0) || !(d->dt_init_fact > 0.0))
    return ctx->fail(PHX_ERR_INVALID, "cfl and dt_init_fact must be positive");
  return PHX_OK;
}

}  // namespace

extern "C" {

int phx_mesh_create(phx_ctx *ctx, const phx_fluid_params *p, const phx_fixup_params *fx,
                    const phx_mesh_desc *d, phx_mesh **out) {
  if (!ctx || !out) return PHX_ERR_INVALID;
  *out = nullptr;
  int rc = check_params(ctx, p);
  if (rc) return rc;
  if ((rc = check_desc(ctx, p, d))) return rc;
  const phx_fixup_params off = {0, PHX_FAILURE_INTERPOLATE, 0, 0, 50.0, 2500.0};
  if (fx && (rc = check_fixup(ctx, fx))) return rc;
  phx_mesh *m = new (std::nothrow) phx_mesh();
  if (!m) return PHX_ERR_NOMEM;
  m->ctx = ctx;
  m->p = *p;
  m->fxp = fx ? *fx : off;
  m->d = *d;
  MeshInfo &mi = m->mi;
  mi.ng = d->ng;
  mi.inflow = d->ox1_inflow_check && p->geom.kind == PHX_GEOM_FMKS;
  m->nblocks = 1;
  for (int e = 0; e < 3; ++e) {
    mi.bnx[e] = d->block_nx[e];
    mi.N[e] = d->nx[e];
    mi.nb[e] = d->nx[e] / d->block_nx[e];
    mi.periodic[e] = d->bc[2 * e] == PHX_BC_PERIODIC;
    m->nblocks *= mi.nb[e];
    for (int side = 0; side < 2; ++side)
      if (d->bc[2 * e + side] == PHX_BC_OUTFLOW) m->outflow_mask |= 1 << (2 * e + side);
  }
  // blocks and their fields
  const long nc = (long)(d->block_nx[0] + 2 * d->ng) * (d->block_nx[1] + 2 * d->ng) *
                  (d->block_nx[2] + 2 * d->ng);
  auto fail = [&](int code, const char *msg) {
    phx_mesh_destroy(m);
    return ctx->fail(code, msg);
  };
  if (hipMalloc(&m->arena, sizeof(double) * kFields * nc * m->nblocks) != hipSuccess ||
      hipMalloc(&m->iters, sizeof(int32_t) * nc * m->nblocks) != hipSuccess ||
      hipMalloc(&m->dt_dev, sizeof(double)) != hipSuccess)
    return fail(PHX_ERR_NOMEM, "mesh allocation failed");
  if (hipMemset(m->arena, 0, sizeof(double) * kFields * nc * m->nblocks) != hipSuccess ||
      hipMemset(m->iters, 0, sizeof(int32_t) * nc * m->nblocks) != hipSuccess)
    return fail(PHX_ERR_DEVICE, "mesh initialisation failed");
  for (int gid = 0; gid < m->nblocks; ++gid) {
    const int lx[3] = {gid % mi.nb[0], (gid / mi.nb[0]) % mi.nb[1], gid / (mi.nb[0] * mi.nb[1])};
    phx_block b{};
    b.ndim = 3;
    b.ng = d->ng;
    for (int e = 0; e < 3; ++e) {
      const double len = d->xmax[e] - d->xmin[e];
      b.nx[e] = d->block_nx[e];
      b.xmin[e] = d->xmin[e] + len * ((double)lx[e] / mi.nb[e]);
      b.dx[e] = len / d->nx[e];
    }
    if ((rc = check_block(ctx, &b, p->recon))) {
      phx_mesh_destroy(m);
      return rc;
    }
    m->blocks.push_back(b);
    phx_pack_block B{};
    B.block = b;
    carve(m->arena + (long)gid * kFields * nc, nc, B);
    B.iters = m->iters + (long)gid * nc;
    m->pb.push_back(B);
    m->interior += (long)b.nx[0] * b.nx[1] * b.nx[2];
  }
  if ((rc = phx_pack_create(ctx, p, m->nblocks, m->pb.data(), &m->pack))) {
    phx_mesh_destroy(m);
    return rc;
  }
  if ((rc = phx_pack_set_fixup(ctx, m->pack, &m->fxp))) {
    phx_mesh_destroy(m);
    return rc;
  }
  *out = m;
  return PHX_OK;
}

int phx_mesh_destroy(phx_mesh *m) {
  if (!m) return PHX_OK;
  if (m->pack) phx_pack_destroy(m->pack);
  if (m->arena) (void)hipFree(m->arena);
  if (m->iters) (void)hipFree(m->iters);
  if (m->dt_dev) (void)hipFree(m->dt_dev);
  delete m;
  return PHX_OK;
}

int phx_mesh_nblocks(const phx_mesh *m, int *nblocks, int64_t *interior_cells) {
  if (!m) return PHX_ERR_INVALID;
  if (nblocks) *nblocks = m->nblocks;
  if (interior_cells) *interior_cells = m->interior;
  return PHX_OK;
}

int phx_mesh_block(const phx_mesh *m, int gid, phx_block *b) {
  if (!m || !b || gid < 0 || gid >= m->nblocks) return PHX_ERR_INVALID;
  *b = m->blocks[gid];
  return PHX_OK;
}

static long block_cells(const phx_block &b) {
  return (long)(b.nx[0] + 2 * b.ng) * (b.nx[1] + 2 * b.ng) * (b.nx[2] + 2 * b.ng);
}

int phx_mesh_set_prim(phx_mesh *m, int gid, const double *host_prim) {
  if (!m || !host_prim || gid < 0 || gid >= m->nblocks) return PHX_ERR_INVALID;
  const size_t bytes = sizeof(double) * kNprim * block_cells(m->blocks[gid]);
  return hipMemcpy(m->pb[gid].f.prim, host_prim, bytes, hipMemcpyHostToDevice) == hipSuccess
             ? PHX_OK
             : PHX_ERR_DEVICE;
}

int phx_mesh_initialize(phx_mesh *m) {
  if (!m) return PHX_ERR_INVALID;
  phx_ctx *ctx = m->ctx;
  hipStream_t s = ctx->stream;
  // the problem generator's p2c on every block (entire), then PostInitializationCommunication:
  // floors, exchange + boundaries + conversion, FillDerived (phoebus_driver.cpp:147-202)
  for (int b = 0; b < m->nblocks; ++b) {
    int lo[3] = {0, 0, 0}, hi[3];
    const phx_block &B = m->blocks[b];
    for (int e = 0; e < 3; ++e) hi[e] = B.nx[e] + 2 * B.ng - 1;
    int rc = phx_prim_to_con(ctx, &m->p, &B, &m->pb[b].f, lo, hi, s);
    if (!rc) rc = phx_apply_floors(ctx, &m->p, &m->fxp, &B, &m->pb[b].f, lo, hi, s);
    if (rc) return rc;
  }
  // base is the output container of stage 2
  launch_exchange_and_boundaries(m, 2, s);
  launch_c2p(m, 2, 5, s);
  // the "1" container starts as a copy of base
  const long nc = block_cells(m->blocks[0]);
  for (int b = 0; b < m->nblocks; ++b)
    if (hipMemcpyAsync(m->pb[b].cons1, m->pb[b].f.cons, sizeof(double) * kNcons * nc,
                       hipMemcpyDeviceToDevice, s) != hipSuccess)
      return ctx->fail(PHX_ERR_DEVICE, "container copy failed");
  if (int rc = ctx->check_launch()) return rc;
  m->has_face_speeds = false;
  m->time = 0.0;
  m->ncycle = 0;
  const double dt0 = estimate_dt(m, s);
  if (!(dt0 > 0.0)) return ctx->fail(PHX_ERR_DEVICE, "timestep estimate failed");
  m->dt = m->d.dt_init_fact * dt0;
  return PHX_OK;
}

int phx_mesh_step(phx_mesh *m) {
  if (!m) return PHX_ERR_INVALID;
  phx_ctx *ctx = m->ctx;
  hipStream_t s = ctx->stream;
  if (m->time + m->dt > m->tlim) m->dt = m->tlim - m->time;
  for (int st = 1; st <= 2; ++st)
    if (int rc = run_stage(m, st, st == 2, s)) return rc;
  double dt_new = -1.0;
  if (hipStreamSynchronize(s) != hipSuccess ||
      hipMemcpy(&dt_new, m->dt_dev, sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
    return ctx->fail(PHX_ERR_DEVICE, "step failed");
  m->has_face_speeds = true;
  m->time += m->dt;
  m->ncycle += 1;
  if (!(dt_new > 0.0)) return ctx->fail(PHX_ERR_DEVICE, "non-positive timestep");
  m->dt = std::min(dt_new, 2.0 * m->dt);
  return PHX_OK;
}

int phx_mesh_execute(phx_mesh *m, double tlim, int max_steps) {
  if (!m) return PHX_ERR_INVALID;
  m->tlim = tlim;
  for (int n = 0; n < max_steps && m->time < m->tlim; ++n)
    if (int rc = phx_mesh_step(m)) return rc;
  return PHX_OK;
}

int phx_mesh_time(const phx_mesh *m, double *t, double *dt, int64_t *ncycle) {
  if (!m) return PHX_ERR_INVALID;
  if (t) *t = m->time;
  if (dt) *dt = m->dt;
  if (ncycle) *ncycle = m->ncycle;
  return PHX_OK;
}

int phx_mesh_get(const phx_mesh *m, int gid, int what, double *host) {
  if (!m || !host || gid < 0 || gid >= m->nblocks) return PHX_ERR_INVALID;
  const long n = block_cells(m->blocks[gid]);
  const phx_pack_block &B = m->pb[gid];
  const double *src = nullptr;
  size_t count = 0;
  switch (what) {
  case PHX_DRIVER_PRIM: src = B.f.prim; count = kNprim * n; break;
  case PHX_DRIVER_CONS: src = B.f.cons; count = kNcons * n; break;
  case PHX_DRIVER_FAIL: src = B.f.fail; count = n; break;
  case PHX_DRIVER_C2P_MU: src = B.f.c2p_mu; count = n; break;
  default: return PHX_ERR_INVALID;
  }
  if (hipStreamSynchronize(m->ctx->stream) != hipSuccess) return PHX_ERR_DEVICE;
  return hipMemcpy(host, src, sizeof(double) * count, hipMemcpyDeviceToHost) == hipSuccess
             ? PHX_OK
             : PHX_ERR_DEVICE;
}

int phx_mesh_get_iters(const phx_mesh *m, int gid, int32_t *host) {
  if (!m || !host || gid < 0 || gid >= m->nblocks) return PHX_ERR_INVALID;
  if (hipStreamSynchronize(m->ctx->stream) != hipSuccess) return PHX_ERR_DEVICE;
  return hipMemcpy(host, m->pb[gid].iters, sizeof(int32_t) * block_cells(m->blocks[gid]),
                   hipMemcpyDeviceToHost) == hipSuccess
             ? PHX_OK
             : PHX_ERR_DEVICE;
}

int phx_mesh_set_learned(phx_mesh *m, const phx_mlp *mlp) {
  if (!m) return PHX_ERR_INVALID;
  return phx_pack_set_learned(m->ctx, m->pack, mlp);
}

int phx_mesh_learned_stats(phx_mesh *m, int reset, int64_t *misses, int64_t *solves) {
  if (!m) return PHX_ERR_INVALID;
  return phx_pack_learned_stats(m->pack, reset, misses, solves);
}

int phx_mesh_autotune(phx_mesh *m, int reps) {
  if (!m) return PHX_ERR_INVALID;
  return phx_pack_autotune(m->ctx, m->pack, reps, nullptr);
}

}  // extern "C"
