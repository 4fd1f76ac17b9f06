// phx_host.hip -- host-side mirror of the fluid task interface (phx_host.hpp) and the
// C ABI of its one-block driver (include/phx_driver.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cstring>
#include <new>

#include "phx.h"
#include "phx_device.hpp"
#include "phx_driver.h"
#include "phx_host.hpp"
#include "phx_internal.hpp"

namespace phx {
namespace host {

template <class T>
static bool dalloc(T **p, size_t n) {
  if (hipMalloc(p, sizeof(T) * std::max<size_t>(n, 1)) != hipSuccess) return false;
  return hipMemset(*p, 0, sizeof(T) * std::max<size_t>(n, 1)) == hipSuccess;
}

MeshBlockData::MeshBlockData(phx_ctx *ctx, const phx_fluid_params &p, const phx_block &b,
                             int ncontainers)
    : ctx_(ctx), p_(p), b_(b) {
  ncell_ = 1;
  for (int d = 0; d < 3; ++d) ncell_ *= b.nx[d] + (d < b.ndim ? 2 * b.ng : 0);
  const long n = ncell_;
  bool ok = true;
  cons_.assign(ncontainers, nullptr);
  for (auto &c : cons_) ok = ok && dalloc(&c, ncons() * n);
  ok = ok && dalloc(&f_.prim, nprim() * n) && dalloc(&f_.temp, n) && dalloc(&f_.c2p_mu, n) &&
       dalloc(&f_.fail, n) && dalloc(&f_.cell_sig, 3 * n) && dalloc(&f_.face_sig, 3 * n) &&
       dalloc(&f_.flux, 3 * ncons() * n) && dalloc(&f_.emf, 3 * n) &&
       dalloc(&f_.src, ncons() * n) && dalloc(&f_.iters, n) && dalloc(&f_.dt, 1);
  ok_ = ok;
}

MeshBlockData::~MeshBlockData() {
  for (auto c : cons_)
    if (c) (void)hipFree(c);
  for (void *q : {(void *)f_.prim, (void *)f_.temp, (void *)f_.c2p_mu, (void *)f_.fail,
                  (void *)f_.cell_sig, (void *)f_.face_sig, (void *)f_.flux, (void *)f_.emf,
                  (void *)f_.src, (void *)f_.iters, (void *)f_.dt})
    if (q) (void)hipFree(q);
}

phx_block_fields MeshBlockData::fields(int container) const {
  phx_block_fields f{};  // contiguous fields: no per-component table
  f.prim = f_.prim;
  f.cons = cons_[container];
  f.temp = f_.temp;
  f.c2p_mu = f_.c2p_mu;
  f.fail = f_.fail;
  f.cell_sig = f_.cell_sig;
  f.face_sig = f_.face_sig;
  f.flux = f_.flux;
  f.emf = f_.emf;
  f.src = f_.src;
  return f;
}

void MeshBlockData::entire(int lo[3], int hi[3]) const {
  for (int d = 0; d < 3; ++d) {
    lo[d] = 0;
    hi[d] = b_.nx[d] + (d < b_.ndim ? 2 * b_.ng : 0) - 1;
  }
}

void MeshBlockData::interior(int lo[3], int hi[3]) const {
  for (int d = 0; d < 3; ++d) {
    lo[d] = d < b_.ndim ? b_.ng : 0;
    hi[d] = lo[d] + b_.nx[d] - 1;
  }
}

static TaskStatus st(int rc) { return rc == PHX_OK ? TaskStatus::complete : TaskStatus::fail; }

namespace fluid {
TaskStatus PrimitiveToConserved(MeshBlockData *rc, int container) {
  int lo[3], hi[3];
  rc->entire(lo, hi);
  const phx_block_fields f = rc->fields(container);
  return st(phx_prim_to_con(rc->ctx(), &rc->params(), &rc->block(), &f, lo, hi, nullptr));
}
TaskStatus CalculateFluxes(MeshBlockData *rc) {
  const phx_block_fields f = rc->fields(0);
  return st(phx_calculate_fluxes(rc->ctx(), &rc->params(), &rc->block(), &f, nullptr));
}
TaskStatus FluxCT(MeshBlockData *rc) {
  const phx_block_fields f = rc->fields(0);
  return st(phx_flux_ct(rc->ctx(), &rc->params(), &rc->block(), &f, nullptr));
}
TaskStatus CalculateFluidSourceTerms(MeshBlockData *rc) {
  const phx_block_fields f = rc->fields(0);
  return st(phx_fluid_source_terms(rc->ctx(), &rc->params(), &rc->block(), &f, nullptr));
}
TaskStatus ConservedToPrimitive(MeshBlockData *rc, int container) {
  int lo[3], hi[3];
  rc->entire(lo, hi);
  const phx_block_fields f = rc->fields(container);
  return st(phx_con2prim(rc->ctx(), &rc->params(), &rc->block(), &f, lo, hi,
                         rc->shared().iters, nullptr));
}
double EstimateTimestepBlock(MeshBlockData *rc, bool &has_face_speeds, double cfl) {
  const phx_block_fields f = rc->fields(0);
  double h = -1.0;
  const bool ok =
      phx_estimate_dt(rc->ctx(), &rc->params(), &rc->block(), &f, has_face_speeds ? 1 : 0, cfl,
                      rc->shared().dt, nullptr) == PHX_OK &&
      hipStreamSynchronize(rc->ctx()->stream) == hipSuccess &&
      hipMemcpy(&h, rc->shared().dt, sizeof(double), hipMemcpyDeviceToHost) == hipSuccess;
  has_face_speeds = true;  // fluid.cpp:940
  return ok ? h : -1.0;
}
}  // namespace fluid

namespace fixup {
TaskStatus ConservedToPrimitiveFixup(MeshBlockData *rc, int container, const phx_fixup_params &fx) {
  const phx_block_fields f = rc->fields(container);
  return st(phx_c2p_fixup(rc->ctx(), &rc->params(), &fx, &rc->block(), &f, nullptr));
}
TaskStatus ApplyFloors(MeshBlockData *rc, int container, const phx_fixup_params &fx) {
  int lo[3], hi[3];
  rc->entire(lo, hi);
  const phx_block_fields f = rc->fields(container);
  return st(phx_apply_floors(rc->ctx(), &rc->params(), &fx, &rc->block(), &f, lo, hi, nullptr));
}
}  // namespace fixup

namespace update {
TaskStatus FluxDivergenceUpdate(MeshBlockData *rc, int stage_container, int base_container,
                                double beta, double dt, int out_container) {
  const phx_block_fields f = rc->fields(0);
  const double *src = rc->params().geom.kind != PHX_GEOM_MINKOWSKI ? f.src : nullptr;
  return st(phx_flux_divergence_update(rc->ctx(), &rc->params(), &rc->block(), f.flux, src,
                                       rc->cons(stage_container), rc->cons(base_container),
                                       beta, dt, rc->cons(out_container), nullptr));
}
}  // namespace update

namespace boundaries {

// fills the ghost layers of dimension d on one side; outflow copies the first interior
// layer scaled by sqrt(gamma)(ghost) / sqrt(gamma)(ref), periodic copies the opposite
// interior layers
template <int GEOM>
__global__ void k_bc(BlockGeom b, DevParams p, int d, int side, int periodic, int nv,
                     double *__restrict__ q) {
  using M = typename MetricOf<GEOM>::type;
  const int n[3] = {b.ni, b.nj, b.nk};
  const int ng = b.ng;
  const long span0 = d == 0 ? ng : n[0];
  const long span1 = d == 1 ? ng : n[1];
  const long span2 = d == 2 ? ng : n[2];
  const long total = span0 * span1 * span2;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= total) return;
  int idx[3] = {(int)(t % span0), (int)((t / span0) % span1), (int)(t / (span0 * span1))};
  if (side == 1) idx[d] += b.ie[d] + 1;  // outer ghosts
  int ref[3] = {idx[0], idx[1], idx[2]};
  if (periodic)
    ref[d] = side == 0 ? idx[d] + (b.ie[d] - b.is[d] + 1) : idx[d] - (b.ie[d] - b.is[d] + 1);
  else
    ref[d] = side == 0 ? b.is[d] : b.ie[d];
  const long c = ((long)idx[2] * b.nj + idx[1]) * b.ni + idx[0];
  const long cr = ((long)ref[2] * b.nj + ref[1]) * b.ni + ref[0];
  double gratio = 1.0;
  if (!periodic) {
    const M g(p, xc(b, 0, idx[0]), xc(b, 1, idx[1]), xc(b, 2, idx[2]));
    const M gr(p, xc(b, 0, ref[0]), xc(b, 1, ref[1]), xc(b, 2, ref[2]));
    gratio = ratio(g.detgamma(), gr.detgamma());
  }
  for (int l = 0; l < nv; ++l) q[l * b.ncell + c] = gratio * q[l * b.ncell + cr];
}

TaskStatus ApplyBoundaryConditions(MeshBlockData *rc, int container, const Kind kinds[6]) {
  const BlockGeom bg = make_block_geom(&rc->block());
  const DevParams dp = make_dev_params(&rc->params());
  hipStream_t s = rc->ctx()->stream;
  for (int d = 0; d < rc->block().ndim; ++d) {
    const int n[3] = {bg.ni, bg.nj, bg.nk};
    long total = 1;
    for (int e = 0; e < 3; ++e) total *= e == d ? bg.ng : n[e];
    for (int side = 0; side < 2; ++side) {
      const int periodic = kinds[2 * d + side] == Kind::periodic;
      const unsigned blocks = (unsigned)((total + 255) / 256);
      with_geom(rc->params().geom.kind, [&](auto G) {
        hipLaunchKernelGGL((k_bc<decltype(G)::value>), dim3(blocks), dim3(256), 0, s, bg, dp, d,
                           side, periodic, rc->ncons(), rc->cons(container));
      });
    }
  }
  return hipGetLastError() == hipSuccess ? TaskStatus::complete : TaskStatus::fail;
}
}  // namespace boundaries

// ---- driver --------------------------------------------------------------------------------
constexpr int kBase = 0, kStage1 = 1;

Driver::Driver(MeshBlockData *rc, const boundaries::Kind kinds[6], double cfl,
               double dt_init_fact)
    : rc_(rc), cfl_(cfl), dt_init_fact_(dt_init_fact) {
  for (int i = 0; i < 6; ++i) kinds_[i] = kinds[i];
}

TaskStatus Driver::FillDerived(int container) {
  if (fluid::ConservedToPrimitive(rc_, container) != TaskStatus::complete) return TaskStatus::fail;
  if (fixup::ConservedToPrimitiveFixup(rc_, container, fx_) != TaskStatus::complete)
    return TaskStatus::fail;
  return fixup::ApplyFloors(rc_, container, fx_);
}

TaskStatus Driver::Initialize() {
  // shock_tube.cpp:121 (p2c), then PostInitializationCommunication: floors, boundaries and
  // FillDerived (phoebus_driver.cpp:147-202); the stage container starts as a copy of base
  if (fluid::PrimitiveToConserved(rc_, kBase) != TaskStatus::complete) return TaskStatus::fail;
  if (fixup::ApplyFloors(rc_, kBase, fx_) != TaskStatus::complete) return TaskStatus::fail;
  if (boundaries::ApplyBoundaryConditions(rc_, kBase, kinds_) != TaskStatus::complete)
    return TaskStatus::fail;
  if (fluid::ConservedToPrimitive(rc_, kBase) != TaskStatus::complete) return TaskStatus::fail;
  const size_t bytes = sizeof(double) * rc_->ncons() * rc_->ncell();
  if (hipMemcpyAsync(rc_->cons(kStage1), rc_->cons(kBase), bytes, hipMemcpyDeviceToDevice,
                     rc_->ctx()->stream) != hipSuccess)
    return TaskStatus::fail;
  has_face_speeds_ = false;
  dt_ = dt_init_fact_ * fluid::EstimateTimestepBlock(rc_, has_face_speeds_, cfl_);
  time_ = 0.0;
  ncycle_ = 0;
  return dt_ > 0.0 ? TaskStatus::complete : TaskStatus::fail;
}

TaskStatus Driver::RungeKuttaStage(int stage) {
  static const double beta[2] = {1.0, 0.5};
  const int sc0 = stage == 1 ? kBase : kStage1;
  const int sc1 = stage == 1 ? kStage1 : kBase;
  const bool curved = rc_->params().geom.kind != PHX_GEOM_MINKOWSKI;
  if (curved && fluid::CalculateFluidSourceTerms(rc_) != TaskStatus::complete)
    return TaskStatus::fail;
  if (fluid::CalculateFluxes(rc_) != TaskStatus::complete) return TaskStatus::fail;
  if (fluid::FluxCT(rc_) != TaskStatus::complete) return TaskStatus::fail;
  if (update::FluxDivergenceUpdate(rc_, sc0, kBase, beta[stage - 1], dt_, sc1) !=
      TaskStatus::complete)
    return TaskStatus::fail;
  if (FillDerived(sc1) != TaskStatus::complete) return TaskStatus::fail;
  if (boundaries::ApplyBoundaryConditions(rc_, sc1, kinds_) != TaskStatus::complete)
    return TaskStatus::fail;
  return FillDerived(sc1);
}

TaskStatus Driver::Step() {
  if (time_ + dt_ > tlim_) dt_ = tlim_ - time_;
  for (int stage = 1; stage <= 2; ++stage)
    if (RungeKuttaStage(stage) != TaskStatus::complete) return TaskStatus::fail;
  time_ += dt_;
  ncycle_ += 1;
  const double dt_new = fluid::EstimateTimestepBlock(rc_, has_face_speeds_, cfl_);
  if (!(dt_new > 0.0)) return TaskStatus::fail;
  dt_ = std::min(dt_new, 2.0 * dt_);
  return TaskStatus::complete;
}

TaskStatus Driver::Execute(double tlim, int max_steps) {
  tlim_ = tlim;
  for (int n = 0; n < max_steps && time_ < tlim_; ++n)
    if (Step() != TaskStatus::complete) return TaskStatus::fail;
  return TaskStatus::complete;
}

}  // namespace host
}  // namespace phx

// ================================ C ABI ==================================================
using namespace phx::host;

struct phx_driver {
  MeshBlockData *rc = nullptr;
  Driver *drv = nullptr;
};

extern "C" {

int phx_driver_create(phx_ctx *ctx, const phx_fluid_params *p, const phx_block *b,
                      const int bc[6], double cfl, double dt_init_fact, phx_driver **out) {
  if (!ctx || !p || !b || !bc || !out) return PHX_ERR_INVALID;
  if (!(cfl > 0.0) || !(dt_init_fact > 0.0))
    return ctx->fail(PHX_ERR_INVALID, "cfl and dt_init_fact must be positive");
  int rc = phx::check_params(ctx, p);
  if (rc) return rc;
  if ((rc = phx::check_block(ctx, b, p->recon))) return rc;
  boundaries::Kind kinds[6];
  for (int i = 0; i < 6; ++i) {
    if (bc[i] != PHX_BC_PERIODIC && bc[i] != PHX_BC_OUTFLOW)
      return ctx->fail(PHX_ERR_INVALID, "invalid boundary kind");
    kinds[i] = bc[i] == PHX_BC_PERIODIC ? boundaries::Kind::periodic : boundaries::Kind::outflow;
  }
  phx_driver *d = new (std::nothrow) phx_driver();
  if (!d) return PHX_ERR_NOMEM;
  d->rc = new (std::nothrow) MeshBlockData(ctx, *p, *b, 2);
  if (!d->rc || !d->rc->ok()) {
    phx_driver_destroy(d);
    return ctx->fail(PHX_ERR_NOMEM, "block allocation failed");
  }
  d->drv = new (std::nothrow) Driver(d->rc, kinds, cfl, dt_init_fact);
  if (!d->drv) {
    phx_driver_destroy(d);
    return PHX_ERR_NOMEM;
  }
  *out = d;
  return PHX_OK;
}

int phx_driver_destroy(phx_driver *d) {
  if (!d) return PHX_OK;
  delete d->drv;
  delete d->rc;
  delete d;
  return PHX_OK;
}

int phx_driver_set_fixup(phx_driver *d, const phx_fixup_params *fx) {
  if (!d || !fx) return PHX_ERR_INVALID;
  if (fx->failure_strategy != PHX_FAILURE_INTERPOLATE && fx->failure_strategy != PHX_FAILURE_FLOORS)
    return PHX_ERR_INVALID;
  d->drv->set_fixup(*fx);
  return PHX_OK;
}

int phx_driver_set_prim(phx_driver *d, const double *host_prim) {
  if (!d || !host_prim) return PHX_ERR_INVALID;
  const size_t bytes = sizeof(double) * d->rc->nprim() * d->rc->ncell();
  return hipMemcpy(d->rc->shared().prim, host_prim, bytes, hipMemcpyHostToDevice) == hipSuccess
             ? PHX_OK
             : PHX_ERR_DEVICE;
}

int phx_driver_initialize(phx_driver *d) {
  if (!d) return PHX_ERR_INVALID;
  return d->drv->Initialize() == TaskStatus::complete ? PHX_OK : PHX_ERR_DEVICE;
}

int phx_driver_step(phx_driver *d) {
  if (!d) return PHX_ERR_INVALID;
  return d->drv->Step() == TaskStatus::complete ? PHX_OK : PHX_ERR_DEVICE;
}

int phx_driver_execute(phx_driver *d, double tlim, int max_steps) {
  if (!d) return PHX_ERR_INVALID;
  return d->drv->Execute(tlim, max_steps) == TaskStatus::complete ? PHX_OK : PHX_ERR_DEVICE;
}

int phx_driver_time(const phx_driver *d, double *t, double *dt, int64_t *ncycle) {
  if (!d) return PHX_ERR_INVALID;
  if (t) *t = d->drv->time();
  if (dt) *dt = d->drv->dt();
  if (ncycle) *ncycle = d->drv->ncycle();
  return PHX_OK;
}

int phx_driver_get(const phx_driver *d, int what, double *host) {
  if (!d || !host) return PHX_ERR_INVALID;
  const long n = d->rc->ncell();
  const double *src = nullptr;
  size_t count = 0;
  switch (what) {
  case PHX_DRIVER_PRIM: src = d->rc->shared().prim; count = d->rc->nprim() * n; break;
  case PHX_DRIVER_CONS: src = d->rc->cons(0); count = d->rc->ncons() * n; break;
  case PHX_DRIVER_FAIL: src = d->rc->shared().fail; count = n; break;
  case PHX_DRIVER_C2P_MU: src = d->rc->shared().c2p_mu; count = n; break;
  default: return PHX_ERR_INVALID;
  }
  (void)hipDeviceSynchronize();
  return hipMemcpy(host, src, sizeof(double) * count, hipMemcpyDeviceToHost) == hipSuccess
             ? PHX_OK
             : PHX_ERR_DEVICE;
}

}  // extern "C"
