// phx_flux_sweeps.hpp -- the fused reconstruction + Riemann flux kernels and their
// launchers, compiled once per flux arithmetic: phx_flux.hip includes this with
// PHX_FLUX_NS = flux_exact (IEEE division and square root, contraction off: bit-identical
// to the reference's rounding), phx_flux_fast.hip with PHX_FLUX_NS = flux_fast and
// PHX_FLUX_FAST = 1 (Newton-refined reciprocals and contraction; within ~1e-14).  The
// namespace keeps the two builds' kernels and launchers apart.
#pragma once

#include <hip/hip_runtime.h>

#include "phx.h"
#include "phx_device.hpp"
#include "phx_flux_shapes.hpp"
#include "phx_internal.hpp"

#ifndef PHX_FLUX_NS
#error "define PHX_FLUX_NS (flux_exact or flux_fast) before including phx_flux_sweeps.hpp"
#endif

namespace phx {
namespace PHX_FLUX_NS {

// Materialize a value in a register at this point.  Without it the compiler sinks the
// reconstruction arithmetic into the Riemann solver and keeps every stencil operand alive,
// which more than doubles the register footprint of the kernel.
__device__ __forceinline__ void pin(double &x) { asm volatile("" : "+v"(x)); }

template <bool MHD, int RIEMANN, int GEOM, bool CACHED, class FP>
__device__ __forceinline__ void face_out(const DevParams &p, int dir, const BlockGeom &b,
                                         const double X[3], const double *ql, const double *qr,
                                         long c, long q, const FP &flux,
                                         double *__restrict__ fsig) {
  constexpr int NC = Layout<MHD>::ncons;
  // the face's metric (x3 faces share the centre's when the metric ignores X3)
  const auto g =
      metric_at<GEOM, CACHED>(p, b, dir == 0 ? 1 : dir == 1 ? 2 : 0, q, X[0], X[1], X[2]);
  double f[NC];
  const double s = riemann_flux<MHD, RIEMANN>(p, dir, g, X, ql, qr, f);
#pragma unroll
  for (int m = 0; m < NC; ++m) flux[dir * 8 + m][c] = f[m];
  fsig[(long)dir * b.ncell + c] = s;
}


// the x sweep for logical group bx of block b; stage = LDS [NP][XB]
template <int RECON, int RIEMANN, bool MHD, int GEOM, bool CACHED = false, class PP, class FP>
__device__ __forceinline__ void flux_x_body(const DevParams &p, const BlockGeom &b, long bx,
                                            const PP &prim, const FP &flux,
                                            double *__restrict__ fsig,
                                            double (*stage)[XB]) {
  constexpr int NP = Layout<MHD>::nprim;
  constexpr int H = ReconWidth<RECON>::H;
  const int dj = b.ndim > 1, dk = b.ndim > 2;
  const int fj = b.ie[1] - b.is[1] + 1 + 2 * dj;
  const int fk = b.ie[2] - b.is[2] + 1 + 2 * dk;
  const long total = (long)b.ni * fj * fk;  // full rows of the face range
  const int t = threadIdx.x;
  long L = bx * (XB - 2 * H - 1) + t;
  const bool valid = L < total;
  if (!valid) L = total - 1;
  const int i = (int)(L % b.ni);
  const long row = L / b.ni;
  const int j = (int)(row % fj) + b.is[1] - dj;
  const int k = (int)(row / fj) + b.is[2] - dk;
  const long c = ((long)k * b.nj + j) * b.ni + i;
  // cells whose stencil lies in this block and in this row
  const bool recon_ok = t >= H && t < XB - H && i >= H && i < b.ni - H;

  // three barriers per block: primitives in, edges computed, upper edges out
#pragma unroll
  for (int v = 0; v < NP; ++v) stage[v][t] = prim[v][c];
  __syncthreads();
  double lo[NP], up[NP];
#pragma unroll
  for (int v = 0; v < NP; ++v) {
    double q[2 * H + 1];
#pragma unroll
    for (int o = 0; o < 2 * H + 1; ++o) q[o] = stage[v][recon_ok ? t + o - H : t];
    reconstruct_edges<RECON>(q, up[v], lo[v]);
    pin(up[v]);
    pin(lo[v]);
  }
  __syncthreads();
#pragma unroll
  for (int v = 0; v < NP; ++v) stage[v][t] = up[v];
  __syncthreads();
  const bool emit = valid && t >= H + 1 && t < XB - H && i >= b.is[0] - 1 && i <= b.ie[0] + 1;
  if (!emit) return;
  double ql[NP];
#pragma unroll
  for (int v = 0; v < NP; ++v) ql[v] = stage[v][t - 1];
  const double X[3] = {xf(b, 0, i), xc(b, 1, j), xc(b, 2, k)};
  face_out<MHD, RIEMANN, GEOM, CACHED>(p, 0, b, X, ql, lo, c, (long)j * b.ni + i, flux, fsig);
}

template <int RECON, int RIEMANN, bool MHD, int GEOM>
__global__ __launch_bounds__(XB) void k_flux_x(DevParams p, BlockGeom b, PrimPtrs prim,
                                                   FluxPtrs flux, double *__restrict__ fsig) {
  __shared__ double stage[Layout<MHD>::nprim][XB];  // primitives, then upper-edge states
  flux_x_body<RECON, RIEMANN, MHD, GEOM>(p, b, xcd_remap(blockIdx.x, gridDim.x), prim, flux,
                                         fsig, stage);
}

// the DIR sweep for logical tile bx of block b; lds = [NP][ts_n][ti_n] doubles
template <int DIR, int RECON, int RIEMANN, bool MHD, int GEOM, bool CACHED = false, class PP,
          class FP>
__device__ __forceinline__ void flux_tile_body(const DevParams &p, const BlockGeom &b, int bx,
                                               int ti_n, int ts_n, const PP &prim,
                                               const FP &flux,
                                               double *__restrict__ fsig, double *lds) {
  constexpr int NP = Layout<MHD>::nprim;
  constexpr int H = ReconWidth<RECON>::H;
  constexpr int T = DIR == 1 ? 2 : 1;  // the transverse dimension
  const int S = ts_n - 2 * H - 1;            // faces per tile
  const int tid = threadIdx.x;
  const int ti = tid % ti_n, ts = tid / ti_n;
  const bool live = ts < ts_n;                // threads beyond the tile only join barriers
  const int dt_ = b.ndim > T;
  const int nic = (b.ie[0] - b.is[0] + 3 + ti_n - 1) / ti_n;  // i chunks over [is-1, ie+1]
  const int nft = (b.ie[DIR] - b.is[DIR] + 3 + S - 1) / S;    // tiles over faces [s-1, e+1]
  const int ichunk = bx % nic;
  const int tile = (bx / nic) % nft;
  const int tr = bx / (nic * nft);
  const int nown = DIR == 1 ? b.nj : b.nk;
  const int i_raw = b.is[0] - 1 + ichunk * ti_n + ti;
  const int o_raw = b.is[DIR] - 2 - H + tile * S + ts;
  const int t = tr + b.is[T] - dt_;
  const int i = min(i_raw, b.ie[0] + 1);
  const int o = min(max(o_raw, 0), nown - 1);
  const int j = DIR == 1 ? o : t, k = DIR == 1 ? t : o;
  const long c = ((long)k * b.nj + j) * b.ni + i;
  const bool recon_ok =
      live && ts >= H && ts < ts_n - H && o_raw - H >= 0 && o_raw + H <= nown - 1;
  const int me = ts * ti_n + ti;

  const int plane = ts_n * ti_n;
  // three barriers per block: primitives in, edges computed, upper edges out
  if (live) {
#pragma unroll
    for (int v = 0; v < NP; ++v) lds[v * plane + me] = prim[v][c];
  }
  __syncthreads();
  double lo[NP], up[NP];
#pragma unroll
  for (int v = 0; v < NP; ++v) {
    double q[2 * H + 1];
#pragma unroll
    for (int oo = 0; oo < 2 * H + 1; ++oo)
      q[oo] = live ? lds[v * plane + (recon_ok ? me + (oo - H) * ti_n : me)] : 0.0;
    reconstruct_edges<RECON>(q, up[v], lo[v]);
    pin(up[v]);
    pin(lo[v]);
  }
  __syncthreads();
  if (live) {
#pragma unroll
    for (int v = 0; v < NP; ++v) lds[v * plane + me] = up[v];
  }
  __syncthreads();
  const bool emit = live && ts >= H + 1 && ts < ts_n - H && i_raw <= b.ie[0] + 1 &&
                    o_raw >= b.is[DIR] - 1 && o_raw <= b.ie[DIR] + 1;
  if (!emit) return;
  double ql[NP];
#pragma unroll
  for (int v = 0; v < NP; ++v) ql[v] = lds[v * plane + me - ti_n];
  double X[3];
  X[0] = xc(b, 0, i);
  X[T] = xc(b, T, t);
  X[DIR] = xf(b, DIR, o);
  face_out<MHD, RIEMANN, GEOM, CACHED>(p, DIR, b, X, ql, lo, c,
                                       (long)(DIR == 1 ? o : t) * b.ni + i, flux, fsig);
}

// the DIR sweep for logical workgroup bx of block b; lds as in march_lds_bytes.  TI > 0:
// the column count ti is that constant (the launchers pass it for the common widths), so
// the ring's row and variable strides fold into the LDS instructions' offsets instead of
// run-time multiplies; TI = 0: ti as passed
template <int DIR, int RECON, int RIEMANN, bool MHD, int GEOM, bool CACHED = false, int TI = 0,
          class PP, class FP>
__device__ __forceinline__ void flux_march_body(const DevParams &p, const BlockGeom &b, long bx,
                                                int ti_run, int nseg, const PP &prim,
                                                const FP &flux,
                                                double *__restrict__ fsig, double *lds) {
  constexpr int NP = Layout<MHD>::nprim;
  constexpr int H = ReconWidth<RECON>::H;
  constexpr int T = DIR == 1 ? 2 : 1;  // the transverse dimension
  const int ti = TI > 0 ? TI : ti_run;
  const int MR = kMarchLanes / ti, Q = MR + 2 * H, lg = __builtin_ctz((unsigned)ti);
  const int lane = threadIdx.x;
  const int dt_ = b.ndim > T;
  const int nci = b.ie[0] - b.is[0] + 3;  // i over [is-1, ie+1]
  const long ncol = (long)nci * (b.ie[T] - b.is[T] + 1 + 2 * dt_);
  const long nchunk = (ncol + ti - 1) / ti;
  const int sg = (int)(bx / nchunk);
  long col = (bx % nchunk) * ti + (lane & (ti - 1));
  const bool colv = col < ncol;
  if (!colv) col = ncol - 1;
  const int i = b.is[0] - 1 + (int)(col % nci);
  const int t = b.is[T] - dt_ + (int)(col / nci);
  const int nown = DIR == 1 ? b.nj : b.nk;
  // own-direction faces [s-1, e+1] in nseg segments; a segment's march starts one cell
  // early for the upper edge its first face needs.  The march takes K = ceil((nf + nseg) /
  // MR) steps in all; segment sg takes ks of them (the first K % nseg segments one more)
  // and covers the ks * MR - 1 faces after the previous segments' (the last one up to
  // nf), so only the last step of the last segment can have idle rows.
  const int nf = b.ie[DIR] - b.is[DIR] + 3;
  const int K = (nf + nseg + MR - 1) / MR, kq = K / nseg, kr = K % nseg;
  const int ks = kq + (sg < kr ? 1 : 0);
  const int f0 = (sg * kq + min(sg, kr)) * MR - sg;  // faces of the segments before
  const int flo = b.is[DIR] - 1 + f0;
  const int fhi = b.is[DIR] - 1 + min(f0 + ks * MR - 1, nf) - 1;
  if (ks == 0 || flo > fhi) return;  // an empty trailing segment (the whole workgroup)
  const int c0 = flo - 1;
  const int nsteps = (fhi - c0 + MR) / MR;
  const long sown = DIR == 1 ? (long)b.ni : (long)b.ni * b.nj;
  double *ring = lds;                // [NP][Q][ti]: row y in slot (y - c0 + H) % Q
  double *carry = lds + NP * Q * ti; // [NP][ti]: the upper edges of the last row
  auto load_row = [&](long base, int cl, int y, int slot) {
    const long c = base + (long)min(max(y, 0), nown - 1) * sown;
#pragma unroll
    for (int v = 0; v < NP; ++v) ring[(v * Q + slot) * ti + cl] = prim[v][c];
  };
  // the ring slot of this lane's cell, (m MR + r) % Q, stepped by MR with one wrap (MR < Q)
  int rs = lane >> lg;
  for (int m = 0; m < nsteps; ++m) {
    // opaque copies: what a step needs of the lane and its column is recomputed from them
    // each step instead of being hoisted out of the loop and held across the Riemann
    // solver (ti is a power of two: shifts and masks)
    int ln = lane, ii = i, tt = t;
    asm volatile("" : "+v"(ln), "+v"(ii), "+v"(tt));
    const int r = ln >> lg, cl = ln & (ti - 1);
    const long base = DIR == 1 ? (long)tt * b.nj * b.ni + ii : (long)tt * b.ni + ii;
    __syncthreads();  // the previous step's ring reads are done
    if (m == 0) {
      for (int y = r; y < Q; y += MR) load_row(base, cl, c0 - H + y, y);
    } else {
      // the new row m MR + 2H + r sits in slot (rs + 2H) % Q
      load_row(base, cl, c0 + m * MR + H + r, rs + 2 * H < Q ? rs + 2 * H : rs + 2 * H - Q);
    }
    __syncthreads();
    const int cc = c0 + m * MR + r;  // this lane's cell
    int sl[2 * H + 1];
#pragma unroll
    for (int o = 0; o < 2 * H + 1; ++o) sl[o] = rs + o < Q ? rs + o : rs + o - Q;
    double lo[NP], up[NP];
#pragma unroll
    for (int v = 0; v < NP; ++v) {
      double q[2 * H + 1];
#pragma unroll
      for (int o = 0; o < 2 * H + 1; ++o) q[o] = ring[(v * Q + sl[o]) * ti + cl];
      reconstruct_edges<RECON>(q, up[v], lo[v]);
      pin(up[v]);
      pin(lo[v]);
    }
    // the upper edge of the cell below: row r - 1's, or the previous step's last row
    double ql[NP];
    const int below = ln >= ti ? ln - ti : ln;
#pragma unroll
    for (int v = 0; v < NP; ++v) {
      const double sh = __shfl(up[v], below, kMarchLanes);
      ql[v] = r == 0 ? carry[v * ti + cl] : sh;
    }
    if (r == MR - 1) {
#pragma unroll
      for (int v = 0; v < NP; ++v) carry[v * ti + cl] = up[v];
    }
    if (colv && cc >= flo && cc <= fhi) {
      double X[3];
      X[0] = xc(b, 0, ii);
      X[T] = xc(b, T, tt);
      X[DIR] = xf(b, DIR, cc);
      face_out<MHD, RIEMANN, GEOM, CACHED>(p, DIR, b, X, ql, lo, base + (long)cc * sown,
                                           (long)(DIR == 1 ? cc : tt) * b.ni + ii, flux, fsig);
    }
    rs = rs + MR < Q ? rs + MR : rs + MR - Q;
  }
}

// Minkowski fits four waves per SIMD (128 VGPRs) without spilling
template <int DIR, int RECON, int RIEMANN, bool MHD, int GEOM, int TI>
__global__ __launch_bounds__(kMarchLanes)
__attribute__((amdgpu_waves_per_eu(GEOM == PHX_GEOM_MINKOWSKI ? 4 : 1))) void k_flux_march(DevParams p, BlockGeom b, int ti,
                                                            int nseg, PrimPtrs prim,
                                                            FluxPtrs flux,
                                                            double *__restrict__ fsig) {
  extern __shared__ double lds[];
  flux_march_body<DIR, RECON, RIEMANN, MHD, GEOM, false, TI>(
      p, b, xcd_remap(blockIdx.x, gridDim.x), ti, nseg, prim, flux, fsig, lds);
}

template <int DIR, int RECON, int RIEMANN, bool MHD, int GEOM>
__global__ __launch_bounds__(512) void k_flux_tile(DevParams p, BlockGeom b, int ti_n,
                                                    int ts_n, PrimPtrs prim, FluxPtrs flux,
                                                    double *__restrict__ fsig) {
  extern __shared__ double lds[];  // [NP][ts_n][ti_n]: primitives, then upper edges
  flux_tile_body<DIR, RECON, RIEMANN, MHD, GEOM>(p, b, (int)xcd_remap(blockIdx.x, gridDim.x),
                                                 ti_n, ts_n, prim, flux, fsig, lds);
}

// ---- packs: the same sweeps over a work list of blocks --------------------------------
template <int RECON, int RIEMANN, int GEOM>
__global__ __launch_bounds__(XB) void k_pack_flux_x(DevParams p, PackView v) {
  __shared__ double stage[Layout<true>::nprim][XB];
  const long g = xcd_remap(blockIdx.x, gridDim.x);
  const int b = pack_find(v.start, v.n, g);
  const PackEntry &E = v.e[b];
  flux_x_body<RECON, RIEMANN, true, GEOM, GEOM == PHX_GEOM_FMKS>(
      p, E.bg, g - v.start[b], contig(E.prim, E.bg.ncell), contig(E.flux, E.bg.ncell), E.fsig,
      stage);
}

template <int DIR, int RECON, int RIEMANN, int GEOM>
__global__ __launch_bounds__(512) void k_pack_flux_tile(DevParams p, PackView v, int ti_n,
                                                         int ts_n) {
  extern __shared__ double lds[];
  const long g = xcd_remap(blockIdx.x, gridDim.x);
  const int b = pack_find(v.start, v.n, g);
  const PackEntry &E = v.e[b];
  flux_tile_body<DIR, RECON, RIEMANN, true, GEOM, GEOM == PHX_GEOM_FMKS>(
      p, E.bg, (int)(g - v.start[b]), ti_n, ts_n, contig(E.prim, E.bg.ncell),
      contig(E.flux, E.bg.ncell), E.fsig, lds);
}

template <int DIR, int RECON, int RIEMANN, int GEOM, int TI>
__global__ __launch_bounds__(kMarchLanes)
__attribute__((amdgpu_waves_per_eu(GEOM == PHX_GEOM_MINKOWSKI ? 4 : 1))) void k_pack_flux_march(DevParams p, PackView v, int ti,
                                                                 int nseg) {
  extern __shared__ double lds[];
  const long g = xcd_remap(blockIdx.x, gridDim.x);
  const int b = pack_find(v.start, v.n, g);
  const PackEntry &E = v.e[b];
  flux_march_body<DIR, RECON, RIEMANN, true, GEOM, GEOM == PHX_GEOM_FMKS, TI>(
      p, E.bg, g - v.start[b], ti, nseg, contig(E.prim, E.bg.ncell), contig(E.flux, E.bg.ncell),
      E.fsig, lds);
}

// ---- dispatch --------------------------------------------------------------------------
// In the fast arithmetic the MHD marching kernels of the Minkowski and FMKS metrics (the
// stage's and the torus packs') are built for a compile-time column count of 8 and 16, the
// widths the autotuner keeps; every other case, and the other widths, take the run-time
// column count (fewer kernels to build)
template <bool MHD, int GEOM, class F>
static void with_march_width(int ti, F &&f) {
  if constexpr (kFluxFast && MHD && GEOM != PHX_GEOM_SPHERICAL_KS) {
    if (ti == 8) return f(std::integral_constant<int, 8>{});
    if (ti == 16) return f(std::integral_constant<int, 16>{});
  }
  f(std::integral_constant<int, 0>{});
}

template <int RECON, int RIEMANN, bool MHD, int GEOM>
static void launch_dir(int d, const DevParams &dp, const BlockGeom &bg, const PrimPtrs &prim,
                       const FluxPtrs &flux, double *fsig, hipStream_t s, TileShape shape) {
  const int dj = bg.ndim > 1, dk = bg.ndim > 2;
  if (d == 0) {
    constexpr int H = ReconWidth<RECON>::H;
    const long total = (long)bg.ni * (bg.ie[1] - bg.is[1] + 1 + 2 * dj) *
                       (bg.ie[2] - bg.is[2] + 1 + 2 * dk);
    const long stride = XB - 2 * H - 1;
    const unsigned blocks = (unsigned)((total - 2 * H - 1 + stride - 1) / stride);
    hipLaunchKernelGGL((k_flux_x<RECON, RIEMANN, MHD, GEOM>), dim3(blocks), dim3(XB), 0, s, dp,
                       bg, prim, flux, fsig);
    return;
  }
  constexpr int H = ReconWidth<RECON>::H;
  constexpr int NP = Layout<MHD>::nprim;
  const int T = d == 1 ? 2 : 1;
  const int dt_ = bg.ndim > T;
  const TileShape ts = resolve_tile_shape(d, RECON, MHD, bg, shape);
  if (ts.ts < 0) {
    const unsigned blocks = (unsigned)flux_march_groups(bg, d, ts.ti, -ts.ts);
    const size_t lds = (size_t)march_lds_bytes(NP, H, ts.ti);
    with_march_width<MHD, GEOM>(ts.ti, [&](auto W) {
      constexpr int TI = decltype(W)::value;
      if (d == 1)
        hipLaunchKernelGGL((k_flux_march<1, RECON, RIEMANN, MHD, GEOM, TI>), dim3(blocks),
                           dim3(kMarchLanes), lds, s, dp, bg, ts.ti, -ts.ts, prim, flux, fsig);
      else
        hipLaunchKernelGGL((k_flux_march<2, RECON, RIEMANN, MHD, GEOM, TI>), dim3(blocks),
                           dim3(kMarchLanes), lds, s, dp, bg, ts.ti, -ts.ts, prim, flux, fsig);
    });
    return;
  }
  const int ts_n = ts.ts, ti_n = ts.ti;
  const int S = ts_n - 2 * H - 1;
  const long nic = (bg.ie[0] - bg.is[0] + 3 + ti_n - 1) / ti_n;
  const long nft = (bg.ie[d] - bg.is[d] + 3 + S - 1) / S;
  const long ntr = bg.ie[T] - bg.is[T] + 1 + 2 * dt_;
  const unsigned blocks = (unsigned)(nic * nft * ntr);
  const int threads = ((ti_n * ts_n + 63) / 64) * 64;
  const size_t lds = sizeof(double) * (size_t)NP * ts_n * ti_n;
  if (d == 1)
    hipLaunchKernelGGL((k_flux_tile<1, RECON, RIEMANN, MHD, GEOM>), dim3(blocks), dim3(threads),
                       lds, s, dp, bg, ti_n, ts_n, prim, flux, fsig);
  else
    hipLaunchKernelGGL((k_flux_tile<2, RECON, RIEMANN, MHD, GEOM>), dim3(blocks), dim3(threads),
                       lds, s, dp, bg, ti_n, ts_n, prim, flux, fsig);
}

template <int RECON, int RIEMANN, bool MHD>
static void launch_geom(int geom, int d, const DevParams &dp, const BlockGeom &bg,
                        const PrimPtrs &prim, const FluxPtrs &flux, double *fsig, hipStream_t s,
                        TileShape sh) {
  with_geom(geom, [&](auto G) {
    launch_dir<RECON, RIEMANN, MHD, decltype(G)::value>(d, dp, bg, prim, flux, fsig, s, sh);
  });
}

template <int RECON, int RIEMANN>
static void launch_mhd(int mhd, int geom, int d, const DevParams &dp, const BlockGeom &bg,
                       const PrimPtrs &prim, const FluxPtrs &flux, double *fsig, hipStream_t s,
                       TileShape sh) {
  if (mhd)
    launch_geom<RECON, RIEMANN, true>(geom, d, dp, bg, prim, flux, fsig, s, sh);
  else
    launch_geom<RECON, RIEMANN, false>(geom, d, dp, bg, prim, flux, fsig, s, sh);
}

template <int RECON>
static void launch_riemann(int riemann, int mhd, int geom, int d, const DevParams &dp,
                           const BlockGeom &bg, const PrimPtrs &prim, const FluxPtrs &flux,
                           double *fsig, hipStream_t s, TileShape sh) {
  if (riemann == PHX_RIEMANN_HLL)
    launch_mhd<RECON, PHX_RIEMANN_HLL>(mhd, geom, d, dp, bg, prim, flux, fsig, s, sh);
  else
    launch_mhd<RECON, PHX_RIEMANN_LLF>(mhd, geom, d, dp, bg, prim, flux, fsig, s, sh);
}

// The sweeps of one reconstruction, block and pack launches.  The WENO5-Z ones are
// instantiated in a translation unit of their own (phx_flux_weno.hip,
// phx_flux_fast_weno.hip) so that the builds of the kernels run in parallel; the including
// unit defines PHX_FLUX_DISPATCH for the rest and the dispatch below.
template <int RECON>
void launch_recon(int riemann, int mhd, int geom, int d, const DevParams &dp,
                  const BlockGeom &bg, const PrimPtrs &prim, const FluxPtrs &flux, double *fsig,
                  hipStream_t s, TileShape sh);
template <int RECON>
void launch_pack_recon(int d, int riemann, int geom, const DevParams &dp, const PackView &v,
                       long groups, TileShape sh, hipStream_t s);

#ifdef PHX_FLUX_DISPATCH
extern template void launch_recon<PHX_RECON_WENO5Z>(int, int, int, int, const DevParams &,
                                                    const BlockGeom &, const PrimPtrs &,
                                                    const FluxPtrs &, double *, hipStream_t,
                                                    TileShape);
extern template void launch_pack_recon<PHX_RECON_WENO5Z>(int, int, int, const DevParams &,
                                                         const PackView &, long, TileShape,
                                                         hipStream_t);

void launch_fluxes(int d, int recon, int riemann, int mhd, int geom, const DevParams &dp,
                   const BlockGeom &bg, const PrimPtrs &prim, const FluxPtrs &flux, double *fsig,
                   hipStream_t s, TileShape sh) {
  switch (recon) {
  case PHX_RECON_CONSTANT:
    launch_recon<PHX_RECON_CONSTANT>(riemann, mhd, geom, d, dp, bg, prim, flux, fsig, s, sh);
    break;
  case PHX_RECON_LINEAR:
    launch_recon<PHX_RECON_LINEAR>(riemann, mhd, geom, d, dp, bg, prim, flux, fsig, s, sh);
    break;
  case PHX_RECON_WENO5Z:
    launch_recon<PHX_RECON_WENO5Z>(riemann, mhd, geom, d, dp, bg, prim, flux, fsig, s, sh);
    break;
  default:
    launch_recon<PHX_RECON_MP5>(riemann, mhd, geom, d, dp, bg, prim, flux, fsig, s, sh);
    break;
  }
}
#endif

template <int RECON, int RIEMANN, int GEOM>
static void pack_dir(int d, const DevParams &dp, const PackView &v, long groups, TileShape sh,
                     hipStream_t s) {
  if (groups <= 0) return;
  if (d == 0) {
    hipLaunchKernelGGL((k_pack_flux_x<RECON, RIEMANN, GEOM>), dim3((unsigned)groups), dim3(XB), 0,
                       s, dp, v);
    return;
  }
  if (sh.ts < 0) {
    const size_t lds =
        (size_t)march_lds_bytes(Layout<true>::nprim, ReconWidth<RECON>::H, sh.ti);
    with_march_width<true, GEOM>(sh.ti, [&](auto W) {
      constexpr int TI = decltype(W)::value;
      if (d == 1)
        hipLaunchKernelGGL((k_pack_flux_march<1, RECON, RIEMANN, GEOM, TI>),
                           dim3((unsigned)groups), dim3(kMarchLanes), lds, s, dp, v, sh.ti,
                           -sh.ts);
      else
        hipLaunchKernelGGL((k_pack_flux_march<2, RECON, RIEMANN, GEOM, TI>),
                           dim3((unsigned)groups), dim3(kMarchLanes), lds, s, dp, v, sh.ti,
                           -sh.ts);
    });
    return;
  }
  const int threads = ((sh.ti * sh.ts + 63) / 64) * 64;
  const size_t lds = sizeof(double) * (size_t)Layout<true>::nprim * sh.ti * sh.ts;
  if (d == 1)
    hipLaunchKernelGGL((k_pack_flux_tile<1, RECON, RIEMANN, GEOM>), dim3((unsigned)groups),
                       dim3(threads), lds, s, dp, v, sh.ti, sh.ts);
  else
    hipLaunchKernelGGL((k_pack_flux_tile<2, RECON, RIEMANN, GEOM>), dim3((unsigned)groups),
                       dim3(threads), lds, s, dp, v, sh.ti, sh.ts);
}

template <int RECON>
static void pack_recon(int d, int riemann, int geom, const DevParams &dp, const PackView &v,
                       long groups, TileShape sh, hipStream_t s) {
  with_geom(geom, [&](auto G) {
    if (riemann == PHX_RIEMANN_HLL)
      pack_dir<RECON, PHX_RIEMANN_HLL, decltype(G)::value>(d, dp, v, groups, sh, s);
    else
      pack_dir<RECON, PHX_RIEMANN_LLF, decltype(G)::value>(d, dp, v, groups, sh, s);
  });
}

template <int RECON>
void launch_recon(int riemann, int mhd, int geom, int d, const DevParams &dp,
                  const BlockGeom &bg, const PrimPtrs &prim, const FluxPtrs &flux, double *fsig,
                  hipStream_t s, TileShape sh) {
  launch_riemann<RECON>(riemann, mhd, geom, d, dp, bg, prim, flux, fsig, s, sh);
}
template <int RECON>
void launch_pack_recon(int d, int riemann, int geom, const DevParams &dp, const PackView &v,
                       long groups, TileShape sh, hipStream_t s) {
  pack_recon<RECON>(d, riemann, geom, dp, v, groups, sh, s);
}

#ifdef PHX_FLUX_DISPATCH
void launch_pack_flux(int d, int recon, int riemann, int geom, const DevParams &dp,
                      const PackView &v, long groups, TileShape sh, hipStream_t s) {
  switch (recon) {
  case PHX_RECON_CONSTANT: launch_pack_recon<PHX_RECON_CONSTANT>(d, riemann, geom, dp, v, groups, sh, s); break;
  case PHX_RECON_LINEAR: launch_pack_recon<PHX_RECON_LINEAR>(d, riemann, geom, dp, v, groups, sh, s); break;
  case PHX_RECON_WENO5Z: launch_pack_recon<PHX_RECON_WENO5Z>(d, riemann, geom, dp, v, groups, sh, s); break;
  default: launch_pack_recon<PHX_RECON_MP5>(d, riemann, geom, dp, v, groups, sh, s); break;
  }
}
#else
template void launch_recon<PHX_RECON_WENO5Z>(int, int, int, int, const DevParams &,
                                             const BlockGeom &, const PrimPtrs &,
                                             const FluxPtrs &, double *, hipStream_t, TileShape);
template void launch_pack_recon<PHX_RECON_WENO5Z>(int, int, int, const DevParams &,
                                                  const PackView &, long, TileShape,
                                                  hipStream_t);
#endif

}  // namespace PHX_FLUX_NS
}  // namespace phx
