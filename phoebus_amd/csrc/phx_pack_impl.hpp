// phx_pack_impl.hpp -- the pack object behind include/phx_pack.h, shared with the mesh
// driver (phx_mesh.hip), which runs its own sequence of the same pack launches.
#pragma once

#include <vector>

#include "phx.h"
#include "phx_internal.hpp"
#include "phx_pack.h"

using phx::BlockGeom;
using phx::DevFixup;
using phx::DevParams;
using phx::PackEntry;
using phx::TileShape;

struct phx_pack {
  int device = 0;
  phx_fluid_params p;
  DevParams dp;
  int n = 0;
  long interior = 0;
  std::vector<phx_pack_block> blocks;
  std::vector<BlockGeom> geoms;
  PackEntry *entries[2] = {nullptr, nullptr};  // device: stage 1 and stage 2 roles
  // work lists: 0..2 flux sweeps, 3 CT corners, 4 interior cells, 5 entire blocks,
  // 6 the ghost layer face-adjacent to the interior, 7 every ghost cell
  long *start[8] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  long groups[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  TileShape tiles[3];
  double *partial = nullptr;  // [groups[4]] CFL partial minima
  DevFixup fx{};              // c2p fixup and floors after the update (off by default)
  const phx_mlp *mlp = nullptr;          // learned con2prim bracket (null: plain solver)
  std::vector<double *> geo;             // per block: cached geometry plane (FMKS) or empty
  unsigned long long *counters = nullptr;  // device [2]: bracket misses, learned solves
  phx::ColdQueue cq;  // cells left to the learned bracket, allocated by the first set_learned
  phx::FloorQueue fq;  // the floors' repair queue, allocated when floors are set on
  // live graphs captured from this pack: they hold its work-list pointers, grid sizes and
  // tile shapes, so autotuning (which rewrites the work lists) and destruction wait for
  // them to be destroyed
  int graphs = 0;
};

struct phx_graph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  phx_pack *pack = nullptr;  // the pack it was captured from
};


namespace phx {
// tile shapes must fit every block of the pack: resolved against the smallest extent
TileShape pack_shape(const phx_pack *k, int d, TileShape want);
// (re)builds work list kind (0..5) on the device
int build_start(phx_pack *k, int kind);
// the launch view of work list kind for stage st (1, 2: which containers are U_s, U_{s+1})
PackView view(const phx_pack *k, int st, int kind);
// the flux part of a stage: the three flux sweeps, CT, geometric sources (curved metrics)
void launch_stage_fluxes(phx_ctx *ctx, phx_pack *k, int st, hipStream_t s);
// one full pack stage: fluxes, CT, sources, fused update + con2prim, fixups, CFL partials
void launch_stage(phx_ctx *ctx, phx_pack *k, const phx_pack_stage_args *a, hipStream_t s);
}  // namespace phx
