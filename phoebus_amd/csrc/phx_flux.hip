// phx_flux.hip -- fused reconstruction + Riemann flux kernels (fluid.cpp:676-770) for gfx950.
//
// The reference reconstructs every cell once per direction into ql/qr arrays and a second
// kernel reads them back.  Here each cell is reconstructed once per direction in
// registers and the face states never reach memory:
//
//  * x (unit stride): one lane per cell along a row; primitives are staged through LDS
//    once per variable and the cell's upper-edge state is published in LDS, so lane l forms
//    face i from the upper edge of cell i-1 (lane l-1) and the lower edge of cell i (its
//    own).  Blocks overlap by 2H+1 cells.
//  * y, z (strided): lanes run along i (coalesced) and each lane marches a segment of its
//    column, sliding a register window of 2H+1 cells, carrying the previous cell's upper
//    edge from one step to the next.
//
// Results are bit-identical to reconstructing per face: the same arithmetic runs on the
// same operands.
#include <hip/hip_runtime.h>

#include "phx.h"
#include "phx_device.hpp"
#include "phx_internal.hpp"

namespace phx {

template <int RECON>
struct ReconWidth {
  static constexpr int H = RECON == PHX_RECON_CONSTANT ? 0 : RECON == PHX_RECON_LINEAR ? 1 : 2;
};

// Materialize a value in a register at this point.  Without it the compiler sinks the
// reconstruction arithmetic into the Riemann solver and keeps every stencil operand alive,
// which more than doubles the register footprint of the kernel.
__device__ __forceinline__ void pin(double &x) { asm volatile("" : "+v"(x)); }

template <bool MHD, int RIEMANN, class M>
__device__ __forceinline__ void face_out(const DevParams &p, int dir, const BlockGeom &b,
                                         const double X[3], const double *ql, const double *qr,
                                         long c, double *__restrict__ flux,
                                         double *__restrict__ fsig) {
  constexpr int NC = Layout<MHD>::ncons;
  const M g(p.a, X[0], X[1], X[2]);
  double f[NC];
  const double s = riemann_flux<MHD, RIEMANN>(p, dir, g, X, ql, qr, f);
#pragma unroll
  for (int m = 0; m < NC; ++m) flux[((long)dir * NC + m) * b.ncell + c] = f[m];
  fsig[(long)dir * b.ncell + c] = s;
}

// ---- x: one lane per cell of full rows, stencils and upper edges staged in LDS ----------
// A block of XB lanes covers XB consecutive cells of the rows it spans; lanes H+1 ..
// XB-H-1 emit the face at their cell's lower edge, so consecutive blocks overlap by 2H+1
// cells.  Each primitive is read from HBM once per lane into LDS, reconstructed from LDS,
// and the cell's upper edge is published in LDS for the next lane.
constexpr int XB = 256;

template <int RECON, int RIEMANN, bool MHD, int GEOM>
__global__ __launch_bounds__(XB) void k_flux_x(DevParams p, BlockGeom b,
                                                   const double *__restrict__ prim,
                                                   double *__restrict__ flux,
                                                   double *__restrict__ fsig) {
  using M = typename MetricOf<GEOM>::type;
  constexpr int NP = Layout<MHD>::nprim;
  constexpr int H = ReconWidth<RECON>::H;
  __shared__ double stencil[XB];
  __shared__ double upper[NP][XB];
  const int dj = b.ndim > 1, dk = b.ndim > 2;
  const int fj = b.ie[1] - b.is[1] + 1 + 2 * dj;
  const int fk = b.ie[2] - b.is[2] + 1 + 2 * dk;
  const long total = (long)b.ni * fj * fk;  // full rows of the face range
  const int t = threadIdx.x;
  long L = (long)blockIdx.x * (XB - 2 * H - 1) + t;
  const bool valid = L < total;
  if (!valid) L = total - 1;
  const int i = (int)(L % b.ni);
  const long row = L / b.ni;
  const int j = (int)(row % fj) + b.is[1] - dj;
  const int k = (int)(row / fj) + b.is[2] - dk;
  const long c = ((long)k * b.nj + j) * b.ni + i;
  // cells whose stencil lies in this block and in this row
  const bool recon_ok = t >= H && t < XB - H && i >= H && i < b.ni - H;

  double lo[NP];
#pragma unroll
  for (int v = 0; v < NP; ++v) {
    stencil[t] = prim[v * b.ncell + c];
    __syncthreads();
    double q[2 * H + 1];
#pragma unroll
    for (int o = 0; o < 2 * H + 1; ++o) q[o] = stencil[recon_ok ? t + o - H : t];
    double up;
    reconstruct_edges<RECON>(q, up, lo[v]);
    upper[v][t] = up;
    pin(lo[v]);
    __syncthreads();
  }
  const bool emit = valid && t >= H + 1 && t < XB - H && i >= b.is[0] - 1 && i <= b.ie[0] + 1;
  if (!emit) return;
  double ql[NP];
#pragma unroll
  for (int v = 0; v < NP; ++v) ql[v] = upper[v][t - 1];
  const double X[3] = {xf(b, 0, i), xc(b, 1, j), xc(b, 2, k)};
  face_out<MHD, RIEMANN, M>(p, 0, b, X, ql, lo, c, flux, fsig);
}

// ---- y, z: lanes along i, each marching a segment of its column --------------------------
template <int DIR, int RECON, int RIEMANN, bool MHD, int GEOM, int LS>
__global__ __launch_bounds__(256) void k_flux_march(DevParams p, BlockGeom b,
                                                    const double *__restrict__ prim,
                                                    double *__restrict__ flux,
                                                    double *__restrict__ fsig) {
  using M = typename MetricOf<GEOM>::type;
  constexpr int NP = Layout<MHD>::nprim;
  constexpr int H = ReconWidth<RECON>::H;
  constexpr int W = 2 * H + 1;
  // transverse dimension T: z for DIR = 1, y for DIR = 2
  constexpr int T = DIR == 1 ? 2 : 1;
  const int dt_ = b.ndim > T;
  const int fi = b.ie[0] - b.is[0] + 3;                    // i in [is-1, ie+1]
  const int ft = b.ie[T] - b.is[T] + 1 + 2 * dt_;          // t in [ts-1, te+1]
  const int nfo = b.ie[DIR] - b.is[DIR] + 3;               // own faces [s-1, e+1]
  const int nseg = (nfo + LS - 1) / LS;
  const long total = (long)fi * ft * nseg;
  const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= total) return;
  const int i = (int)(g % fi) + b.is[0] - 1;
  const long r = g / fi;
  const int t = (int)(r % ft) + b.is[T] - dt_;
  const int seg = (int)(r / ft);
  const int f0 = b.is[DIR] - 1 + seg * LS;
  const int f1 = min(f0 + LS, b.ie[DIR] + 2);  // faces [f0, f1)
  const long st = DIR == 1 ? (long)b.ni : (long)b.ni * b.nj;
  const int j0 = DIR == 1 ? 0 : t, k0 = DIR == 1 ? t : 0;
  // cell index of (own = 0) in this column
  const long base = DIR == 1 ? ((long)k0 * b.nj) * b.ni + i : (long)j0 * b.ni + i;

  // window of cells cell-H .. cell+H for the first cell f0-1
  double q[NP][W];
  const int cfirst = f0 - 1;
#pragma unroll
  for (int v = 0; v < NP; ++v) {
    const double *pv = prim + v * b.ncell + base;
#pragma unroll
    for (int o = 0; o < W; ++o) {
      const int cc = cfirst - H + o;
      q[v][o] = pv[(long)(cc < 0 ? cfirst : cc) * st];
    }
  }
  double upp[NP];
  const double xt = xc(b, T, t), xi = xc(b, 0, i);
  for (int cell = cfirst; cell < f1; ++cell) {
    double up[NP], lo[NP];
    const bool flat = cell - H < 0;  // only the unconsumed lowest face when ng < H + 2
#pragma unroll
    for (int v = 0; v < NP; ++v) {
      if (flat) {
        double qq[W];
#pragma unroll
        for (int o = 0; o < W; ++o) qq[o] = q[v][H];
        reconstruct_edges<RECON>(qq, up[v], lo[v]);
      } else {
        reconstruct_edges<RECON>(q[v], up[v], lo[v]);
      }
      pin(up[v]);
      pin(lo[v]);
    }
    if (cell >= f0) {
      double X[3];
      X[0] = xi;
      X[T] = xt;
      X[DIR] = xf(b, DIR, cell);
      face_out<MHD, RIEMANN, M>(p, DIR, b, X, upp, lo, base + (long)cell * st, flux, fsig);
    }
#pragma unroll
    for (int v = 0; v < NP; ++v) upp[v] = up[v];
    if (cell + 1 < f1) {
      const int nxt = cell + 1 + H;  // <= ie + 1 + H, inside the block for ng >= H + 1
#pragma unroll
      for (int v = 0; v < NP; ++v) {
#pragma unroll
        for (int o = 0; o < W - 1; ++o) q[v][o] = q[v][o + 1];
        q[v][W - 1] = prim[v * b.ncell + base + (long)nxt * st];
      }
    }
  }
}

// ---- dispatch --------------------------------------------------------------------------
constexpr int kSegment = 16;

template <int RECON, int RIEMANN, bool MHD, int GEOM>
static void launch_dir(int d, const DevParams &dp, const BlockGeom &bg, const double *prim,
                       double *flux, double *fsig, hipStream_t s) {
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
  const int T = d == 1 ? 2 : 1;
  const int dt_ = bg.ndim > T;
  const long fi = bg.ie[0] - bg.is[0] + 3;
  const long ft = bg.ie[T] - bg.is[T] + 1 + 2 * dt_;
  const long nseg = (bg.ie[d] - bg.is[d] + 3 + kSegment - 1) / kSegment;
  const long total = fi * ft * nseg;
  const unsigned blocks = (unsigned)((total + 255) / 256);
  if (d == 1)
    hipLaunchKernelGGL((k_flux_march<1, RECON, RIEMANN, MHD, GEOM, kSegment>), dim3(blocks),
                       dim3(256), 0, s, dp, bg, prim, flux, fsig);
  else
    hipLaunchKernelGGL((k_flux_march<2, RECON, RIEMANN, MHD, GEOM, kSegment>), dim3(blocks),
                       dim3(256), 0, s, dp, bg, prim, flux, fsig);
}

template <int RECON, int RIEMANN, bool MHD>
static void launch_geom(int geom, int d, const DevParams &dp, const BlockGeom &bg,
                        const double *prim, double *flux, double *fsig, hipStream_t s) {
  if (geom == PHX_GEOM_MINKOWSKI)
    launch_dir<RECON, RIEMANN, MHD, PHX_GEOM_MINKOWSKI>(d, dp, bg, prim, flux, fsig, s);
  else
    launch_dir<RECON, RIEMANN, MHD, PHX_GEOM_SPHERICAL_KS>(d, dp, bg, prim, flux, fsig, s);
}

template <int RECON, int RIEMANN>
static void launch_mhd(int mhd, int geom, int d, const DevParams &dp, const BlockGeom &bg,
                       const double *prim, double *flux, double *fsig, hipStream_t s) {
  if (mhd)
    launch_geom<RECON, RIEMANN, true>(geom, d, dp, bg, prim, flux, fsig, s);
  else
    launch_geom<RECON, RIEMANN, false>(geom, d, dp, bg, prim, flux, fsig, s);
}

template <int RECON>
static void launch_riemann(int riemann, int mhd, int geom, int d, const DevParams &dp,
                           const BlockGeom &bg, const double *prim, double *flux,
                           double *fsig, hipStream_t s) {
  if (riemann == PHX_RIEMANN_HLL)
    launch_mhd<RECON, PHX_RIEMANN_HLL>(mhd, geom, d, dp, bg, prim, flux, fsig, s);
  else
    launch_mhd<RECON, PHX_RIEMANN_LLF>(mhd, geom, d, dp, bg, prim, flux, fsig, s);
}

void launch_fluxes(int d, int recon, int riemann, int mhd, int geom, const DevParams &dp,
                   const BlockGeom &bg, const double *prim, double *flux, double *fsig,
                   hipStream_t s) {
  switch (recon) {
  case PHX_RECON_CONSTANT:
    launch_riemann<PHX_RECON_CONSTANT>(riemann, mhd, geom, d, dp, bg, prim, flux, fsig, s);
    break;
  case PHX_RECON_LINEAR:
    launch_riemann<PHX_RECON_LINEAR>(riemann, mhd, geom, d, dp, bg, prim, flux, fsig, s);
    break;
  case PHX_RECON_WENO5Z:
    launch_riemann<PHX_RECON_WENO5Z>(riemann, mhd, geom, d, dp, bg, prim, flux, fsig, s);
    break;
  default:
    launch_riemann<PHX_RECON_MP5>(riemann, mhd, geom, d, dp, bg, prim, flux, fsig, s);
    break;
  }
}

}  // namespace phx
