// phx_host.hpp -- host-side mirror of phoebus's fluid task interface over the C ABI.
//
// A drop-in shape for the reference's task functions: the same names and argument
// meaning (a block's data container in, TaskStatus out), with the work done by the HIP
// kernels of libphx.so through include/phx.h.  MeshBlockData here owns the device fields
// of one block the way parthenon's container does for the fluid package: conserved
// variables per RK stage container, primitives and derived fields shared ("OneCopy",
// fluid.cpp:146-151).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "phx.h"
#include "phx_fixup.h"

namespace phx {
namespace host {

enum class TaskStatus { complete, fail };

// shared (OneCopy) fields of a block
struct BlockFields {
  double *prim = nullptr, *temp = nullptr, *c2p_mu = nullptr, *fail = nullptr;
  double *cell_sig = nullptr, *face_sig = nullptr, *flux = nullptr, *emf = nullptr;
  double *src = nullptr;
  int32_t *iters = nullptr;
  double *dt = nullptr;  // one device scalar for the timestep reduction
};

class MeshBlockData {
 public:
  MeshBlockData(phx_ctx *ctx, const phx_fluid_params &p, const phx_block &b, int ncontainers);
  ~MeshBlockData();
  MeshBlockData(const MeshBlockData &) = delete;
  MeshBlockData &operator=(const MeshBlockData &) = delete;

  phx_ctx *ctx() const { return ctx_; }
  const phx_fluid_params &params() const { return p_; }
  const phx_block &block() const { return b_; }
  long ncell() const { return ncell_; }
  int nprim() const { return p_.mhd ? 10 : 7; }
  int ncons() const { return p_.mhd ? 8 : 5; }
  double *cons(int container) const { return cons_[container]; }
  const BlockFields &shared() const { return f_; }
  // the C-ABI view with the conserved variables of one container
  phx_block_fields fields(int container) const;
  // interior / entire index ranges (inclusive), ordered (i, j, k)
  void entire(int lo[3], int hi[3]) const;
  void interior(int lo[3], int hi[3]) const;
  bool ok() const { return ok_; }

 private:
  phx_ctx *ctx_;
  phx_fluid_params p_;
  phx_block b_;
  long ncell_ = 0;
  std::vector<double *> cons_;
  BlockFields f_;
  bool ok_ = false;
};

namespace fluid {
// fluid.cpp:359-449, over the entire block
TaskStatus PrimitiveToConserved(MeshBlockData *rc, int container);
// fluid.cpp:659-773
TaskStatus CalculateFluxes(MeshBlockData *rc);
// fluid.cpp:775-841
TaskStatus FluxCT(MeshBlockData *rc);
// fluid.cpp:566-657 (into the shared src field)
TaskStatus CalculateFluidSourceTerms(MeshBlockData *rc);
// fluid.cpp:470-475: robust con2prim over IndexDomain::entire of one container
TaskStatus ConservedToPrimitive(MeshBlockData *rc, int container);
// fluid.cpp:899-944; has_face_speeds mirrors the package parameter of :911-941
double EstimateTimestepBlock(MeshBlockData *rc, bool &has_face_speeds, double cfl);
}  // namespace fluid

namespace fixup {
// fixup_c2p.cpp:44-356 (fluid part, interior cells of one container)
TaskStatus ConservedToPrimitiveFixup(MeshBlockData *rc, int container, const phx_fixup_params &fx);
// fixup.cpp:285-539 (fluid part, IndexDomain::entire of one container)
TaskStatus ApplyFloors(MeshBlockData *rc, int container, const phx_fixup_params &fx);
}  // namespace fixup

namespace update {
// parthenon FluxDivergence + SumData + AverageIndependentData + UpdateIndependentData,
// phoebus_driver.cpp:560-576: containers by index
TaskStatus FluxDivergenceUpdate(MeshBlockData *rc, int stage_container, int base_container,
                                double beta, double dt, int out_container);
}  // namespace update

namespace boundaries {
enum class Kind { periodic, outflow };
// parthenon physical boundaries for conserved variables (bc_vars = conserved):
// outflow = phoebus OutflowInnerX1/OuterX1 (phoebus_boundaries.cpp:118-150), copying the
// first interior cell scaled by sqrt(gamma) ratios (1 in Minkowski); periodic = a
// single-block periodic exchange.  kinds[2*d + side].
TaskStatus ApplyBoundaryConditions(MeshBlockData *rc, int container, const Kind kinds[6]);
}  // namespace boundaries

// The fluid-only part of PhoebusDriver (phoebus_driver.cpp:121-143, 204-738) for one
// block with the LowStorageIntegrator rk2 weights (beta = 1, 1/2): per stage fluxes, CT,
// sources, update, con2prim (entire), boundaries, con2prim (entire); the timestep after
// the last stage; the first dt scaled by dt_init_fact, growth capped at 2x per step and
// the last step clipped to tlim (parthenon's EvolutionDriver conventions).  Every
// con2prim is followed by the c2p fixup and the floors (phoebus_driver.cpp:586-597,
// 664-676), which do nothing unless enabled in the fixup parameters.
class Driver {
 public:
  Driver(MeshBlockData *rc, const boundaries::Kind kinds[6], double cfl, double dt_init_fact);
  // after the primitives are set on the entire block: p2c, boundaries, con2prim, first dt
  TaskStatus Initialize();
  TaskStatus Step();          // one RK2 step of size dt()
  TaskStatus RungeKuttaStage(int stage);
  TaskStatus Execute(double tlim, int max_steps);
  double time() const { return time_; }
  double dt() const { return dt_; }
  long ncycle() const { return ncycle_; }
  void set_tlim(double t) { tlim_ = t; }
  void set_fixup(const phx_fixup_params &fx) { fx_ = fx; }

 private:
  MeshBlockData *rc_;
  boundaries::Kind kinds_[6];
  double cfl_, dt_init_fact_, time_ = 0.0, dt_ = 0.0, tlim_ = 1e300;
  long ncycle_ = 0;
  bool has_face_speeds_ = false;
  phx_fixup_params fx_ = {0, PHX_FAILURE_INTERPOLATE, 0, 0, 50.0, 2500.0};
  // FillDerived on one container followed by the fixup tasks
  TaskStatus FillDerived(int container);
};

}  // namespace host
}  // namespace phx
