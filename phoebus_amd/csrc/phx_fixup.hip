// phx_fixup.hip -- the fixup package's repair of con2prim failures and its floors and
// ceilings (include/phx_fixup.h), for one block and over block packs.  Both are per-cell
// kernels: the repair reads the fail flags and primitives of the six face neighbours, the
// floors touch only their own cell (and rerun the robust con2prim where a floor was
// injected).  Most cells need neither, so the kernels are dominated by reading the fail
// flags (repair) or rho, u, v, B and the metric (floors).
#include <hip/hip_runtime.h>

#include "phx.h"
#include "phx_device.hpp"
#include "phx_fixup.h"
#include "phx_internal.hpp"
#include "phx_cells.hpp"

namespace phx {

// ConservedToPrimitiveFixupImpl (fixup_c2p.cpp:175-312), fluid part, for interior cell
// (i, j, k).  A failed neighbour inside the interior is being rewritten by this same pass:
// its weight (its fail flag) is 0 and it is skipped rather than multiplied.
template <bool MHD, int GEOM, bool CACHED = false, class PP, class CP>
__device__ __forceinline__ void c2p_fixup_cell(const DevParams &p, const DevFixup &fx,
                                               const BlockGeom &b, int i, int j, int k,
                                               const PP &prim, const CP &cons,
                                               double *temp, double *csig, const double *fail) {
  using M = typename MetricOf<GEOM>::type;
  using L = Layout<MHD>;
  const long n = b.ncell;
  const long c = ((long)k * b.nj + j) * b.ni + i;
  if (fail[c] != 0.0) return;  // FailFlags::fail == 0
  const long st[3] = {1, b.ni, (long)b.ni * b.nj};
  const int idx[3] = {i, j, k};
  double w[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  long nb[6] = {c, c, c, c, c, c};
  bool skip[6] = {true, true, true, true, true, true};
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    if (d >= b.ndim) continue;
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      const int q = 2 * d + s, ni = idx[d] + (s ? 1 : -1);
      nb[q] = s ? c + st[d] : c - st[d];
      w[q] = fail[nb[q]];
      skip[q] = w[q] == 0.0 && ni >= b.is[d] && ni <= b.ie[d];
    }
  }
  double num_valid = w[0] + w[1];  // :218-221
  if (b.ndim > 1) num_valid += w[2] + w[3];
  if (b.ndim == 3) num_valid += w[4] + w[5];
  double pr[L::nprim];
#pragma unroll
  for (int v = 0; v < L::nprim; ++v) pr[v] = prim[v][c];
  if (num_valid > 0.5 && fx.strategy == PHX_FAILURE_INTERPOLATE) {  // :225-243
    const double norm = 1.0 / num_valid;
#pragma unroll
    for (int m = 0; m < 5; ++m) {  // rho, v^1..3, u
      double t[6];
#pragma unroll
      for (int q = 0; q < 6; ++q) t[q] = skip[q] ? 0.0 : w[q] * prim[m][nb[q]];
      double acc = t[0] + t[1];
      if (b.ndim > 1) {
        acc += t[2] + t[3];
        if (b.ndim == 3) acc += t[4] + t[5];
      }
      pr[m] = norm * acc;
    }
    pr[0] = smax(pr[0], 100. * kSmall);
    pr[4] = smax(pr[4], 100. * kSmall);
  } else {  // :244-258
    pr[0] = 100. * kSmall;
    pr[4] = 100. * kSmall;
#pragma unroll
    for (int m = 0; m < 3; ++m) pr[1 + m] = 0.;
  }
  const double X[3] = {xc(b, 0, i), xc(b, 1, j), xc(b, 2, k)};
  const M g = metric_at<GEOM, CACHED>(p, b, 0, (long)j * b.ni + i, X[0], X[1], X[2]);
  const double gamma_max = p.bnd.gam_max;
  double vpcon[3] = {pr[1], pr[2], pr[3]};
  const double W = lorentz_factor(g, vpcon);
  if (W > gamma_max) {  // :269-275
    const double rescale = sqrt(gamma_max * gamma_max - 1.) / (W * W - 1.);
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      vpcon[m] *= rescale;
      pr[1 + m] = vpcon[m];
    }
  }
  const double t = eos_t_rho_sie(p, ratio(pr[4], pr[0]));  // :277-287
  pr[L::prs] = eos_p_rho_t(p, pr[0], t);
  pr[L::gm1] = ratio(eos_bmod_rho_t(p, pr[0], t), pr[L::prs]);
  double bp[3] = {0.0, 0.0, 0.0};
  if constexpr (MHD) {
#pragma unroll
    for (int m = 0; m < 3; ++m) bp[m] = pr[5 + m];
  }
  double D, S[3], bc[3], tau, sig[3];
  p2c<MHD>(pr[0], vpcon, bp, pr[4], pr[L::prs], pr[L::gm1], g, D, S, bc, tau, sig);
#pragma unroll
  for (int v = 0; v < L::nprim; ++v) prim[v][c] = pr[v];
  if (temp) temp[c] = t;
  cons[0][c] = D;
#pragma unroll
  for (int m = 0; m < 3; ++m) cons[1 + m][c] = S[m];
  cons[4][c] = tau;
  for (int d = 0; d < b.ndim; ++d) csig[d * n + c] = sig[d];
}

// the floors ApplyFloorsImpl targets for a cell (fixup.cpp:397-428): rho and u floors,
// raised by the MHD b^2 ceilings when they are on
template <bool MHD, class M>
__device__ __forceinline__ void floor_targets(const DevParams &p, const DevFixup &fx, const M &g,
                                              const double *pr, double rho_floor,
                                              double sie_floor, double &rho_floor_max,
                                              double &u_floor_max) {
  rho_floor_max = rho_floor;
  u_floor_max = rho_floor * sie_floor;
  if constexpr (MHD) {
    if (fx.mhd_floors) {  // :411-428
      double Bsq = 0.0, Bdotv = 0.0;
      const double vp[3] = {pr[1], pr[2], pr[3]};
      const double bp[3] = {pr[5], pr[6], pr[7]};
      const double W = lorentz_factor(g, vp);
      const double iW = 1.0 / W;
#pragma unroll
      for (int ii = 0; ii < 3; ++ii)
#pragma unroll
        for (int jj = 0; jj < 3; ++jj) {
          if (!M::nz_cov(ii + 1, jj + 1)) continue;
          Bsq += g.gcov(ii + 1, jj + 1) * bp[ii] * bp[jj];
          Bdotv += g.gcov(ii + 1, jj + 1) * bp[ii] * vp[jj];
        }
      const double bcon0 = W * Bdotv / g.alpha();
      const double bsq = (Bsq + g.alpha() * g.alpha() * bcon0 * bcon0) * iW * iW;
      rho_floor_max = smax(rho_floor_max, bsq / fx.bsqorho_max);
      u_floor_max = smax(u_floor_max, bsq / fx.bsqou_max);
      rho_floor_max = smax(rho_floor_max, smax(pr[4], u_floor_max) / p.bnd.sie_max);
    }
  }
}

// whether floors_cell changes cell (i, j, k): a floor (:434) or a ceiling (:501-507) acts.
// It reads only the primitives floors_cell decides on and holds none of its repair state.
template <bool MHD, int GEOM, bool CACHED = false, class PP>
__device__ __forceinline__ bool floors_needed(const DevParams &p, const DevFixup &fx,
                                              const BlockGeom &b, int i, int j, int k,
                                              const PP &prim) {
  using M = typename MetricOf<GEOM>::type;
  const long c = ((long)k * b.nj + j) * b.ni + i;
  const double X[3] = {xc(b, 0, i), xc(b, 1, j), xc(b, 2, k)};
  double rho_floor, sie_floor;
  get_floors(p.bnd, X[0], X[1], X[2], rho_floor, sie_floor);
  const M g = metric_at<GEOM, CACHED>(p, b, 0, (long)j * b.ni + i, X[0], X[1], X[2]);
  double pr[8];
#pragma unroll
  for (int v = 0; v < (MHD ? 8 : 5); ++v) pr[v] = prim[v][c];
  double rho_floor_max, u_floor_max;
  floor_targets<MHD>(p, fx, g, pr, rho_floor, sie_floor, rho_floor_max, u_floor_max);
  if (rho_floor_max - pr[0] > 0. || u_floor_max - pr[4] > 0.) return true;
  double vpcon[3] = {pr[1], pr[2], pr[3]};
  return lorentz_factor(g, vpcon) > p.bnd.gam_max || pr[4] / pr[0] > p.bnd.sie_max;
}

// ApplyFloorsImpl (fixup.cpp:379-539), fluid part, for cell (i, j, k)
template <bool MHD, int GEOM, bool CACHED = false, class PP, class CP>
__device__ __forceinline__ void floors_cell(const DevParams &p, const DevFixup &fx,
                                            const BlockGeom &b, int i, int j, int k,
                                            const PP &prim, const CP &cons,
                                            double *temp, double *mu, double *csig) {
  using M = typename MetricOf<GEOM>::type;
  using L = Layout<MHD>;
  const long n = b.ncell;
  const long c = ((long)k * b.nj + j) * b.ni + i;
  const double X[3] = {xc(b, 0, i), xc(b, 1, j), xc(b, 2, k)};
  double rho_floor, sie_floor;
  get_floors(p.bnd, X[0], X[1], X[2], rho_floor, sie_floor);
  const double gamma_max = p.bnd.gam_max, e_max = p.bnd.sie_max;
  const M g = metric_at<GEOM, CACHED>(p, b, 0, (long)j * b.ni + i, X[0], X[1], X[2]);
  double pr[L::nprim];
#pragma unroll
  for (int v = 0; v < L::nprim; ++v) pr[v] = prim[v][c];
  double rho_floor_max, u_floor_max;
  floor_targets<MHD>(p, fx, g, pr, rho_floor, sie_floor, rho_floor_max, u_floor_max);
  double drho = rho_floor_max - pr[0];
  double du = u_floor_max - pr[4];
  bool floor_applied = false;
  if (drho > 0. || du > 0.) {  // :434-440
    floor_applied = true;
    drho = smax(drho, du / sie_floor);
    du = smax(du, sie_floor * drho);
  }
  double bp[3] = {0.0, 0.0, 0.0};
  if constexpr (MHD) {
#pragma unroll
    for (int m = 0; m < 3; ++m) bp[m] = pr[5 + m];
  }
  double cs[L::ncons];
  double t = temp ? temp[c] : 0.0;
  if (floor_applied) {  // :448-498
#pragma unroll
    for (int v = 0; v < L::ncons; ++v) cs[v] = cons[v][c];
    const double vzero[3] = {0.0, 0.0, 0.0};
    const double dprs = eos_p_rho_sie(p, drho, ratio(du, drho));
    const double dgm1 = ratio(eos_bmod_rho_sie(p, drho, ratio(du, drho)), dprs);
    double dcrho, dS[3], dB[3], dtau, dsig[3];
    p2c<MHD>(drho, vzero, bp, du, dprs, dgm1, g, dcrho, dS, dB, dtau, dsig);
    cs[0] += dcrho;
#pragma unroll
    for (int m = 0; m < 3; ++m) cs[1 + m] += dS[m];
    cs[4] += dtau;
    double sig[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) sig[d] = d < b.ndim ? csig[d * n + c] : 0.0;
    double m_io = mu[c];
    int it = 0;
    const double status = con2prim_zone<MHD>(p, g, X, cs, pr, t, m_io, sig, it);
    mu[c] = m_io;
    for (int d = 0; d < b.ndim; ++d) csig[d * n + c] = sig[d];
    if (status == 0.0) {  // :473-497
      pr[0] = drho;
#pragma unroll
      for (int m = 0; m < 3; ++m) pr[1 + m] = 0.0;
      pr[4] = du;
      t = eos_t_rho_sie(p, pr[4] / pr[0]);
      double tau;
      p2c<MHD>(drho, vzero, bp, du, dprs, dgm1, g, cs[0], dS, dB, tau, dsig);
      cs[4] = tau;
#pragma unroll
      for (int m = 0; m < 3; ++m) cs[1 + m] = dS[m];
    }
  }
  double vpcon[3] = {pr[1], pr[2], pr[3]};  // :501-507
  const double W = lorentz_factor(g, vpcon);
  const bool ceiling_applied = W > gamma_max || pr[4] / pr[0] > e_max;
  if (ceiling_applied) {  // :509-526
    if (!floor_applied) {
#pragma unroll
      for (int v = 0; v < L::ncons; ++v) cs[v] = cons[v][c];
    }
    const double rescale = sqrt(gamma_max * gamma_max - 1.) / (W * W - 1.);
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      vpcon[m] *= rescale;
      pr[1 + m] = vpcon[m];
    }
    double dS[3], dB[3], tau, dsig[3];
    p2c<MHD>(pr[0], vpcon, bp, pr[4], pr[L::prs], pr[L::gm1], g, cs[0], dS, dB, tau, dsig);
    cs[4] = tau;
#pragma unroll
    for (int m = 0; m < 3; ++m) cs[1 + m] = dS[m];
  }
  if (floor_applied || ceiling_applied) {  // :528-539
    t = eos_t_rho_sie(p, ratio(pr[4], pr[0]));
    pr[L::prs] = eos_p_rho_t(p, pr[0], t);
    pr[L::gm1] = ratio(eos_bmod_rho_t(p, pr[0], t), pr[L::prs]);
#pragma unroll
    for (int v = 0; v < L::nprim; ++v) prim[v][c] = pr[v];
#pragma unroll
    for (int v = 0; v < L::ncons; ++v) cons[v][c] = cs[v];
    if (temp) temp[c] = t;
  }
}

// interior cell t of block b -> (i, j, k); false beyond the interior
__device__ __forceinline__ bool interior_cell(const BlockGeom &b, long t, int &i, int &j,
                                              int &k) {
  const int ni = b.ie[0] - b.is[0] + 1, nj = b.ie[1] - b.is[1] + 1, nk = b.ie[2] - b.is[2] + 1;
  if (t >= (long)ni * nj * nk) return false;
  i = (int)(t % ni) + b.is[0];
  j = (int)((t / ni) % nj) + b.is[1];
  k = (int)(t / ((long)ni * nj)) + b.is[2];
  return true;
}

template <bool MHD, int GEOM>
__global__ __launch_bounds__(256) void k_c2p_fixup(DevParams p, DevFixup fx, BlockGeom b,
                                                   PrimPtrs prim, ConsPtrs cons, double *temp,
                                                   double *csig, const double *fail) {
  int i, j, k;
  if (!interior_cell(b, (long)blockIdx.x * blockDim.x + threadIdx.x, i, j, k)) return;
  c2p_fixup_cell<MHD, GEOM>(p, fx, b, i, j, k, prim, cons, temp, csig, fail);
}

template <bool MHD, int GEOM>
__global__ __launch_bounds__(256) void k_apply_floors(DevParams p, DevFixup fx, BlockGeom b,
                                                      Range3 rg, PrimPtrs prim, ConsPtrs cons,
                                                      double *temp, double *mu, double *csig) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)rg.n[0] * rg.n[1] * rg.n[2]) return;
  const int i = (int)(t % rg.n[0]) + rg.lo[0];
  const int j = (int)((t / rg.n[0]) % rg.n[1]) + rg.lo[1];
  const int k = (int)(t / ((long)rg.n[0] * rg.n[1])) + rg.lo[2];
  floors_cell<MHD, GEOM>(p, fx, b, i, j, k, prim, cons, temp, mu, csig);
}

// packs (3-D MHD): the fixups on U_{s+1} and the stage's primitives
template <int GEOM>
__global__ __launch_bounds__(256) void k_pack_c2p_fixup(DevParams p, DevFixup fx, PackView v) {
  const long g = xcd_remap(blockIdx.x, gridDim.x);
  const int b = pack_find(v.start, v.n, g);
  const PackEntry &E = v.e[b];
  int i, j, k;
  if (!interior_cell(E.bg, (g - v.start[b]) * 256 + threadIdx.x, i, j, k)) return;
  c2p_fixup_cell<true, GEOM, GEOM == PHX_GEOM_FMKS>(
      p, fx, E.bg, i, j, k, contig(E.prim_out, E.bg.ncell), contig(E.u_out, E.bg.ncell),
      E.temp, E.csig, E.fail);
}

// the CFL time of interior cell (i, j, k) with face speeds (fluid.cpp:913-928)
__device__ __forceinline__ double cell_dt(const BlockGeom &bg, const PackEntry &E, int i, int j,
                                          int k) {
  const long n = bg.ncell, c = ((long)k * bg.nj + j) * bg.ni + i;
  const long off[3] = {1, bg.ni, (long)bg.ni * bg.nj};
  double ldt = 0.0;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    if (d >= bg.ndim) break;
    const double s = smax(E.csig[d * n + c], smax(E.fsig[d * n + c], E.fsig[d * n + c + off[d]]));
    ldt += s / bg.dx[d];
  }
  const double t = 1.0 / (ldt + 1.e-50);
  return isnan(t) ? DBL_MAX : t;  // a NaN time constrains nothing (as a serial minimum)
}

// the floors' check pass (on the interior cells, or on the entire blocks for a work list
// over them): a cell a floor or ceiling acts on goes to the queue, every other interior
// cell's CFL time into the workgroup's partial minimum (so the time step sees the signal
// speeds after the fixups); q.cell null: no floors, the partials only
template <int GEOM>
__global__ __launch_bounds__(256) void k_pack_floors(DevParams p, DevFixup fx, PackView v,
                                                     int entire, double *__restrict__ partial,
                                                     FloorQueue q) {
  const long g = xcd_remap(blockIdx.x, gridDim.x);
  const int b = pack_find(v.start, v.n, g);
  const PackEntry &E = v.e[b];
  const BlockGeom &bg = E.bg;
  const long t = (g - v.start[b]) * 256 + threadIdx.x;
  int i, j, k;
  bool cell, inner;
  if (entire) {
    cell = t < bg.ncell;
    i = (int)(t % bg.ni);
    j = (int)((t / bg.ni) % bg.nj);
    k = (int)(t / ((long)bg.ni * bg.nj));
    inner = cell && i >= bg.is[0] && i <= bg.ie[0] && j >= bg.is[1] && j <= bg.ie[1] &&
            k >= bg.is[2] && k <= bg.ie[2];
  } else {
    cell = inner = interior_cell(bg, t, i, j, k);
  }
  const bool defer = cell && q.cell &&
                     floors_needed<true, GEOM, GEOM == PHX_GEOM_FMKS>(
                         p, fx, bg, i, j, k, contig(E.prim_out, bg.ncell));
  // one queue slot range per wave
  const unsigned long long mask = __ballot(defer);
  if (mask) {
    const int lane = threadIdx.x & 63;
    const int leader = __ffsll((long long)mask) - 1;
    unsigned int base = 0;
    if (lane == leader) base = atomicAdd(q.count, (unsigned int)__popcll(mask));
    base = __shfl(base, leader, 64);
    const unsigned int slot = base + (unsigned int)__popcll(mask & ((1ull << lane) - 1ull));
    if (defer && slot < q.cap) {
      q.cell[slot] = ((unsigned long long)b << 32) | (unsigned long long)(((long)k * bg.nj + j) * bg.ni + i);
      q.grp[slot] = (inner && partial) ? (int)g : -1;
    }
  }
  if (partial) {
    const double m = block_min(inner && !defer ? cell_dt(bg, E, i, j, k) : DBL_MAX);
    if (threadIdx.x == 0) partial[g] = m;
  }
}

// the floors' repair pass over the queued cells (a grid-stride loop over the count the
// check pass left); a repaired interior cell's CFL time joins its partial by an atomic
// minimum on the bits (positive doubles order as their bit patterns), so the partials equal
// those of one pass over every cell
template <int GEOM>
__global__ __launch_bounds__(256) void k_pack_floors_apply(DevParams p, DevFixup fx, PackView v,
                                                           double *partial, FloorQueue q) {
  const long n = (long)min((unsigned long long)*q.count, (unsigned long long)q.cap);
  for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < n; e += (long)gridDim.x * 256) {
    const unsigned long long w = q.cell[e];
    const PackEntry &E = v.e[(int)(w >> 32)];
    const BlockGeom &bg = E.bg;
    const long c = (long)(w & 0xffffffffull);
    const int i = (int)(c % bg.ni), j = (int)((c / bg.ni) % bg.nj);
    const int k = (int)(c / ((long)bg.ni * bg.nj));
    floors_cell<true, GEOM, GEOM == PHX_GEOM_FMKS>(
        p, fx, bg, i, j, k, contig(E.prim_out, bg.ncell), contig(E.u_out, bg.ncell), E.temp,
        E.mu, E.csig);
    const int gg = q.grp[e];
    if (partial && gg >= 0)
      atomicMin(reinterpret_cast<unsigned long long *>(partial + gg),
                (unsigned long long)__double_as_longlong(cell_dt(bg, E, i, j, k)));
  }
}

__global__ void k_zero_count(unsigned int *count, double *extra) {
  if (threadIdx.x == 0) {
    *count = 0u;
    if (extra) *extra = DBL_MAX;
  }
}

void launch_zero_count(unsigned int *count, hipStream_t s, double *extra) {
  hipLaunchKernelGGL(k_zero_count, dim3(1), dim3(64), 0, s, count, extra);
}

void launch_pack_c2p_fixup(int geom, const DevParams &dp, const DevFixup &fx, const PackView &v,
                           long groups, hipStream_t s) {
  if (groups <= 0 || !fx.c2p_fixup) return;
  with_geom(geom, [&](auto G) {
    hipLaunchKernelGGL((k_pack_c2p_fixup<decltype(G)::value>), dim3((unsigned)groups),
                       dim3(256), 0, s, dp, fx, v);
  });
}

void launch_pack_floors(int geom, const DevParams &dp, const DevFixup &fx, const PackView &v,
                        long groups, double *partial, hipStream_t s, bool entire,
                        const FloorQueue &q) {
  if (groups <= 0 || (!fx.floors && !partial)) return;
  const FloorQueue qq = fx.floors && q.cell && q.cap >= groups * 256 ? q : FloorQueue{};
  if (fx.floors && !qq.cell) return;  // the owner sizes the queue: never reached
  if (qq.cell) launch_zero_count(qq.count, s);
  with_geom(geom, [&](auto G) {
    hipLaunchKernelGGL((k_pack_floors<decltype(G)::value>), dim3((unsigned)groups), dim3(256),
                       0, s, dp, fx, v, entire ? 1 : 0, partial, qq);
    if (qq.cell)
      hipLaunchKernelGGL((k_pack_floors_apply<decltype(G)::value>),
                         dim3((unsigned)(groups < 1024 ? groups : 1024)), dim3(256), 0, s, dp, fx, v,
                         partial, qq);
  });
}

DevFixup make_dev_fixup(const phx_fluid_params *p, const phx_fixup_params *fx) {
  DevFixup d{};
  if (!fx) return d;
  d.c2p_fixup = fx->enable_c2p_fixup != 0;
  d.strategy = fx->failure_strategy;
  d.mhd_floors = fx->enable_mhd_floors != 0 && p->mhd != 0;
  d.floors = fx->enable_floors != 0 || d.c2p_fixup || d.mhd_floors;
  d.bsqorho_max = fx->bsqorho_max;
  d.bsqou_max = fx->bsqou_max;
  return d;
}

int check_fixup(phx_ctx *ctx, const phx_fixup_params *fx) {
  if (!fx) return ctx->fail(PHX_ERR_INVALID, "null fixup parameters");
  if (fx->failure_strategy != PHX_FAILURE_INTERPOLATE && fx->failure_strategy != PHX_FAILURE_FLOORS)
    return ctx->fail(PHX_ERR_INVALID, "invalid failure strategy");
  if (fx->enable_mhd_floors && !(fx->bsqorho_max > 0.0 && fx->bsqou_max > 0.0))
    return ctx->fail(PHX_ERR_INVALID, "MHD ceilings must be positive");
  return PHX_OK;
}

}  // namespace phx

using namespace phx;

extern "C" int phx_c2p_fixup(phx_ctx *ctx, const phx_fluid_params *p,
                             const phx_fixup_params *fx, const phx_block *b,
                             const phx_block_fields *f, phx_stream stream) {
  if (!ctx) return PHX_ERR_INVALID;
  int rc = check_params(ctx, p);
  if (rc) return rc;
  if ((rc = check_fixup(ctx, fx))) return rc;
  if ((rc = check_block(ctx, b, PHX_RECON_CONSTANT))) return rc;
  if (!f) return ctx->fail(PHX_ERR_INVALID, "null field");
  const BlockGeom bg = make_block_geom(b);
  const int np = p->mhd ? 10 : 7, nc = p->mhd ? 8 : 5;
  const PrimPtrs prim = prim_ptrs(f, bg.ncell, np);
  const ConsPtrs cons = cons_ptrs(f, bg.ncell, nc);
  if (!complete(prim, np) || !complete(cons, nc) || !f->fail || !f->cell_sig)
    return ctx->fail(PHX_ERR_INVALID, "null field");
  const DevFixup dfx = make_dev_fixup(p, fx);
  if (!dfx.c2p_fixup) return PHX_OK;
  const DevParams dp = make_dev_params(p);
  hipStream_t s = ctx->stream_of(stream);
  const long n = (long)b->nx[0] * b->nx[1] * b->nx[2];
  with_geom(p->geom.kind, [&](auto G) {
    with_mhd(p->mhd, [&](auto H) {
      hipLaunchKernelGGL((k_c2p_fixup<decltype(H)::value, decltype(G)::value>),
                         dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, dp, dfx, bg,
                         prim, cons, f->temp, f->cell_sig, f->fail);
    });
  });
  return ctx->check_launch();
}

extern "C" int phx_apply_floors(phx_ctx *ctx, const phx_fluid_params *p,
                                const phx_fixup_params *fx, const phx_block *b,
                                const phx_block_fields *f, const int lo[3], const int hi[3],
                                phx_stream stream) {
  if (!ctx) return PHX_ERR_INVALID;
  int rc = check_params(ctx, p);
  if (rc) return rc;
  if ((rc = check_fixup(ctx, fx))) return rc;
  if ((rc = check_block(ctx, b, PHX_RECON_CONSTANT))) return rc;
  if (!f) return ctx->fail(PHX_ERR_INVALID, "null field");
  const BlockGeom bg = make_block_geom(b);
  const int np = p->mhd ? 10 : 7, nc = p->mhd ? 8 : 5;
  const PrimPtrs prim = prim_ptrs(f, bg.ncell, np);
  const ConsPtrs cons = cons_ptrs(f, bg.ncell, nc);
  if (!complete(prim, np) || !complete(cons, nc) || !f->c2p_mu || !f->cell_sig)
    return ctx->fail(PHX_ERR_INVALID, "null field");
  Range3 rg;
  if ((rc = make_range(ctx, b, lo, hi, rg))) return rc;
  const DevFixup dfx = make_dev_fixup(p, fx);
  if (!dfx.floors) return PHX_OK;
  const DevParams dp = make_dev_params(p);
  hipStream_t s = ctx->stream_of(stream);
  const long n = (long)rg.n[0] * rg.n[1] * rg.n[2];
  with_geom(p->geom.kind, [&](auto G) {
    with_mhd(p->mhd, [&](auto H) {
      hipLaunchKernelGGL((k_apply_floors<decltype(H)::value, decltype(G)::value>),
                         dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, dp, dfx, bg, rg,
                         prim, cons, f->temp, f->c2p_mu, f->cell_sig);
    });
  });
  return ctx->check_launch();
}
