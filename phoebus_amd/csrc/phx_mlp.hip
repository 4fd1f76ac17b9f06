// phx_mlp.hip -- learned con2prim bracket (no reference counterpart): a 4-64-64-64-2
// ReLU network on BF16 MFMA with FP32 accumulation, its training kernels (forward,
// backward, Adam) and the con2prim kernel that uses it.
//
// Network layout.  Every product is formed transposed, Z^T = W . H^T, with zones as the
// MFMA N dimension: v_mfma_f32_16x16x16_bf16 leaves lane l holding rows 4(l/16)..+3 of
// column l%16 of its 16x16 output tile, which is exactly the B-operand layout the next
// layer needs (K rows 4(l/16)..+3 of column l%16).  Activations therefore stay in
// registers from layer to layer; weights are the A operand, packed on the device into
// per-lane fragments (lane l: row l%16, columns 4(l/16)..+3).  The weight-gradient
// products need the other orientation and transpose 16x16 tiles through LDS.
//
// Precision.  A plain BF16 forward pass resolves the root to a few percent, which buys a
// bracket only ~100x narrower than [0, 1].  The forward products therefore run in split
// BF16: weights and activations are carried as hi + lo BF16 pairs (lo = the BF16 rounding
// of the remainder) and each product is hi.hi + hi.lo + lo.hi on MFMA with FP32
// accumulation (about 16 significant bits, 3x the MFMA work of a plain pass).  The
// backward pass uses the hi parts only.
//
// Features (scale-free conserved invariants of the con2prim set-up): log q, log r^2,
// log b^2, log (r.b)^2 with q = tau/D, r = S/D, b = B/sqrt(D), standardized.
//
// Outputs, relative to a physical prior.  With mu = 1/(h W), the energy equation gives
// (tau + D)/D = hW - P/(rho W) + b^2 (1 + v^2)/2 - (b.v)^2/2, so z0 = 1 + q - b^2/2 =
// hW - P/(rho W) + (b^2 v^2 - (b.v)^2)/2 >= max(hW/Gamma_ad, 1): for b = 0, mu z0 lies in
// [1/Gamma_ad, 1], and log(mu z0) is a slowly varying function of the features (the
// max(., 1) only guards round-off).  o1 = log(mu z0) of
// the bracket centre, c = exp(o1)/z0; o2 = log of the bracket half-width relative to c.
// Learning the residual instead of logit(mu) leaves the network ~10x less to resolve.
// The solver evaluates the residual at the bracket ends; a sign change runs the
// reference's regula falsi inside it, anything else falls back to the plain solve on
// [0, 1] with the centre as guess.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <new>
#include <vector>

#include "phx.h"
#include "phx_device.hpp"
#include "phx_internal.hpp"
#include "phx_cells.hpp"
#include "phx_mlp.h"

namespace phx {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// parameter vector: W1[64][4] b1[64] W2[64][64] b2[64] W3[64][64] b3[64] W4[2][64] b4[2]
constexpr int kIn = 4, kHid = 64, kOut = 2;
constexpr int oW1 = 0, oB1 = oW1 + kHid * kIn, oW2 = oB1 + kHid, oB2 = oW2 + kHid * kHid;
constexpr int oW3 = oB2 + kHid, oB3 = oW3 + kHid * kHid, oW4 = oB3 + kHid;
constexpr int oB4 = oW4 + kOut * kHid, kNParam = oB4 + kOut;
static_assert(kNParam == PHX_MLP_NPARAM, "parameter count");
// compact copy of the biases (b1, b2, b3, b4), rebuilt with the packed weights
constexpr int kB1 = 0, kB2 = kHid, kB3 = 2 * kHid, kB4 = 3 * kHid, kNBias = 3 * kHid + kOut;

// Forward fragments (A = W, M = out, K = in).  Layer 1 (K = 4 inputs) runs on
// v_mfma_f32_16x16x16_bf16: 4 M-tiles, one bf16x4 (one uint64) per lane.  Layers 2-4
// (K = 64) run on gfx950's v_mfma_f32_16x16x32_bf16: per M-tile two K-chunks of 32, one
// bf16x8 (two uint64) per lane.  The K order inside chunk kc is the C layout of the
// previous layer's tiles 2kc and 2kc+1: lane group q = l/16 holds slots 8q..8q+7 =
// rows 4q..4q+3 of tile 2kc, then of tile 2kc+1, so a layer's output tiles are the next
// layer's B operand with no data movement (the weights are packed in the same order).
// Fragment f of layers 2-4: L2 f = 2 mt + kc (0..7), L3 8 + 2 mt + kc, L4 16 + kc.
// Each of the hi and lo sets: [4][64] uint64 for layer 1, then [18][64][2].
constexpr int kFwd32Frags = 8 + 8 + 2;
constexpr int kFwdSet = 4 * 64 + kFwd32Frags * 64 * 2;  // uint64 per set
constexpr int kFwdU64 = 2 * kFwdSet;                    // hi set, then lo set
// backward fragments (A = W^T: M = in, K = out) on 16x16x16
constexpr int kBwdFrags = 4 * 4 + 4 * 4 + 4 * 1;  // L2, L3, L4

struct MlpDev {
  const uint64_t *fwd;   // [kFwdU64]: the hi set, then the lo set (layout above)
  const uint64_t *bwd;   // [kBwdFrags][64]
  const float *bias;     // [kNBias] compact biases
  const float *param;    // [kNParam] fp32 master
  float fmean[4], fstd[4];
  float width_scale, width_min;
};

// FP32 -> BF16 round to nearest even: gfx950's v_cvt_pk_bf16_f32 (two values per
// instruction)
typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint32_t cvt2(float a, float b) {
  const bf16x2_t h = {(__bf16)a, (__bf16)b};
  uint32_t u;
  memcpy(&u, &h, 4);
  return u;
}
__device__ __forceinline__ uint16_t f2bf(float f) { return (uint16_t)cvt2(f, 0.f); }
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float(((uint32_t)h) << 16); }

__device__ __forceinline__ bf16x4 pack4(float a, float b, float c, float d) {
  const uint64_t u = (uint64_t)cvt2(a, b) | ((uint64_t)cvt2(c, d) << 32);
  bf16x4 v;
  memcpy(&v, &u, 8);
  return v;
}
__device__ __forceinline__ bf16x4 as_bf16x4(uint64_t u) {
  bf16x4 v;
  memcpy(&v, &u, 8);
  return v;
}
__device__ __forceinline__ f32x4 mfma(bf16x4 a, bf16x4 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
__device__ __forceinline__ f32x4 mfma32(bf16x8 a, bf16x8 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ bf16x8 as_bf16x8(uint64_t lo, uint64_t hi) {
  bf16x8 v;
  const uint64_t u[2] = {lo, hi};
  memcpy(&v, u, 16);
  return v;
}
__device__ __forceinline__ bf16x8 cat8(bf16x4 a, bf16x4 b) {
  uint64_t ua, ub;
  memcpy(&ua, &a, 8);
  memcpy(&ub, &b, 8);
  return as_bf16x8(ua, ub);
}

// ---- forward for one tile of 16 zones -------------------------------------------------
struct Split {
  bf16x4 hi, lo;
};
__device__ __forceinline__ Split split4(float a, float b, float c, float d) {
  // hi = bf16(x), lo = bf16(x - hi); the low/high halves of a packed pair widen to FP32
  // by a shift and a mask
  const uint32_t h01 = cvt2(a, b), h23 = cvt2(c, d);
  const float ra = a - __uint_as_float(h01 << 16), rb = b - __uint_as_float(h01 & 0xffff0000u);
  const float rc = c - __uint_as_float(h23 << 16), rd = d - __uint_as_float(h23 & 0xffff0000u);
  const uint64_t hi = (uint64_t)h01 | ((uint64_t)h23 << 32);
  const uint64_t lo = (uint64_t)cvt2(ra, rb) | ((uint64_t)cvt2(rc, rd) << 32);
  Split s;
  memcpy(&s.hi, &hi, 8);
  memcpy(&s.lo, &lo, 8);
  return s;
}
// split-BF16 product accumulate: W . a with W = (whi, wlo) fragments of one A tile;
// layer 1 (16x16x16, M-tile mt) and layers 2-4 (16x16x32, fragment f, B = two C tiles)
__device__ __forceinline__ f32x4 mfma3(const MlpDev &m, int mt, int lane, const Split &a,
                                      f32x4 acc, int ofs) {
  const bf16x4 wh = as_bf16x4(m.fwd[ofs + mt * 64 + lane]);
  const bf16x4 wl = as_bf16x4(m.fwd[ofs + kFwdSet + mt * 64 + lane]);
  acc = mfma(wh, a.hi, acc);
  acc = mfma(wh, a.lo, acc);
  acc = mfma(wl, a.hi, acc);
  return acc;
}
struct Split8 {
  bf16x8 hi, lo;
};
__device__ __forceinline__ Split8 split8(const Split &a, const Split &b) {
  return Split8{cat8(a.hi, b.hi), cat8(a.lo, b.lo)};
}
__device__ __forceinline__ f32x4 mfma3_32(const MlpDev &m, int f, int lane, const Split8 &a,
                                         f32x4 acc, int ofs) {
  const int q = ofs + 4 * 64 + (f * 64 + lane) * 2;
  const bf16x8 wh = as_bf16x8(m.fwd[q], m.fwd[q + 1]);
  const bf16x8 wl = as_bf16x8(m.fwd[kFwdSet + q], m.fwd[kFwdSet + q + 1]);
  acc = mfma32(wh, a.hi, acc);
  acc = mfma32(wh, a.lo, acc);
  acc = mfma32(wl, a.hi, acc);
  return acc;
}

// x: B fragment of the standardized features (lanes 0..15: features 0..3 of zone l; others
// zero), as a hi/lo pair.  h1..h3 (optional): post-activation tiles, 4 neuron tiles each.
// Returns the output tile (lanes 0..15 hold o1, o2 of zone l in elements 0, 1).  ofs is
// added to every weight and bias index; it is 0 at run time (callers may make it opaque
// to the compiler so that weight reads are not hoisted out of an enclosing loop).
__device__ __forceinline__ f32x4 mlp_tile_forward(const MlpDev &m, const Split &x, f32x4 *h1,
                                                  f32x4 *h2, f32x4 *h3, int ofs = 0) {
  const int lane = threadIdx.x & 63;
  const int r0 = 4 * (lane >> 4);
  f32x4 a[4], bcur[4];
  // layer 1: 64 x 16(4) . 16 x 16
#pragma unroll
  for (int mt = 0; mt < 4; ++mt) {
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    acc = mfma3(m, mt, lane, x, acc, ofs);
#pragma unroll
    for (int r = 0; r < 4; ++r) acc[r] = fmaxf(acc[r] + m.bias[ofs + kB1 + mt * 16 + r0 + r], 0.f);
    a[mt] = acc;
    if (h1) h1[mt] = acc;
  }
  // layers 2 and 3: 64 x 64 . 64 x 16, two K-chunks of 32 per M-tile
#pragma unroll
  for (int layer = 2; layer <= 3; ++layer) {
    const int base = layer == 2 ? 0 : 8;
    const int ob = layer == 2 ? kB2 : kB3;
    Split8 bin[2];
#pragma unroll
    for (int kc = 0; kc < 2; ++kc)
      bin[kc] = split8(split4(a[2 * kc][0], a[2 * kc][1], a[2 * kc][2], a[2 * kc][3]),
                       split4(a[2 * kc + 1][0], a[2 * kc + 1][1], a[2 * kc + 1][2],
                              a[2 * kc + 1][3]));
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
      f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int kc = 0; kc < 2; ++kc) acc = mfma3_32(m, base + 2 * mt + kc, lane, bin[kc], acc, ofs);
#pragma unroll
      for (int r = 0; r < 4; ++r) acc[r] = fmaxf(acc[r] + m.bias[ofs + ob + mt * 16 + r0 + r], 0.f);
      bcur[mt] = acc;
    }
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
      a[mt] = bcur[mt];
      if (layer == 2 && h2) h2[mt] = bcur[mt];
      if (layer == 3 && h3) h3[mt] = bcur[mt];
    }
  }
  // layer 4: 16(2) x 64 . 64 x 16
  f32x4 o = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int kc = 0; kc < 2; ++kc)
    o = mfma3_32(m, 16 + kc, lane,
                 split8(split4(a[2 * kc][0], a[2 * kc][1], a[2 * kc][2], a[2 * kc][3]),
                        split4(a[2 * kc + 1][0], a[2 * kc + 1][1], a[2 * kc + 1][2],
                               a[2 * kc + 1][3])),
                 o, ofs);
  if (r0 == 0) {
    o[0] += m.bias[ofs + kB4 + 0];
    o[1] += m.bias[ofs + kB4 + 1];
  }
  return o;
}

// standardized features of one zone from its con2prim set-up
__device__ __forceinline__ void zone_features(const MlpDev &m, const Residual &r, float f[4]) {
  const double raw[4] = {log(fmax(r.q, 0.0) + 1e-12), log(fmax(r.rsq, 0.0) + 1e-12),
                         log(fmax(r.bsq, 0.0) + 1e-12), log(fmax(r.rbsq, 0.0) + 1e-24)};
#pragma unroll
  for (int i = 0; i < 4; ++i) f[i] = ((float)raw[i] - m.fmean[i]) / m.fstd[i];
}

// the prior z0 = max(1 + q - b^2/2, 1) from standardized features (the same fp32 function
// of the features in training and in the solve)
__device__ __forceinline__ float prior_z0(const MlpDev &m, float f0, float f2) {
  const float q = expf(f0 * m.fstd[0] + m.fmean[0]) - 1e-12f;
  const float bsq = expf(f2 * m.fstd[2] + m.fmean[2]) - 1e-12f;
  return fmaxf(1.0f + q - 0.5f * bsq, 1.0f);
}

// wave-cooperative prediction for the 64 zones of this wavefront (all lanes must call it):
// returns o1, o2 of this lane's zone
__device__ __forceinline__ void mlp_predict_wave(const MlpDev &m, const float f[4], float &o1,
                                                 float &o2) {
  const int lane = threadIdx.x & 63;
  float r1 = 0.f, r2 = 0.f;
  // one tile at a time: the weight fragments are re-read per tile (through an index
  // offset the compiler cannot see is 0) rather than all 40 held in registers
#pragma unroll 1
  for (int zt = 0; zt < 4; ++zt) {
    int ofs = 0;
    asm volatile("" : "+s"(ofs));
    // B fragment: lanes 0..15 carry the features of zone zt*16 + lane
    const int src = zt * 16 + (lane & 15);
    const float g0 = __shfl(f[0], src, 64), g1 = __shfl(f[1], src, 64);
    const float g2 = __shfl(f[2], src, 64), g3 = __shfl(f[3], src, 64);
    const Split x = lane < 16 ? split4(g0, g1, g2, g3) : split4(0.f, 0.f, 0.f, 0.f);
    const f32x4 o = mlp_tile_forward(m, x, nullptr, nullptr, nullptr, ofs);
    const float t1 = __shfl(o[0], lane & 15, 64), t2 = __shfl(o[1], lane & 15, 64);
    if ((lane >> 4) == zt) {
      r1 = t1;
      r2 = t2;
    }
  }
  o1 = r1;
  o2 = r2;
}

// ---- kernels: packing, training, Adam, batch inference ------------------------------------
__global__ void k_mlp_pack(const float *__restrict__ param, uint64_t *__restrict__ fwd,
                           uint64_t *__restrict__ bwd, float *__restrict__ bias) {
  // threads: layer-1 fragments (4 x 64), layers 2-4 (18 x 64), backward (kBwdFrags x 64),
  // then the biases
  constexpr int kT1 = 4 * 64, kT32 = kFwd32Frags * 64, kTB = kBwdFrags * 64;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  auto W = [&](int layer, int o, int i) -> float {
    const int nin = layer == 1 ? kIn : kHid, nout = layer == 4 ? kOut : kHid;
    if (o >= nout || i >= nin) return 0.f;
    const int off = layer == 1 ? oW1 : layer == 2 ? oW2 : layer == 3 ? oW3 : oW4;
    return param[off + o * nin + i];
  };
  auto pack = [&](float a, float b, float c, float d) -> uint64_t {
    return (uint64_t)f2bf(a) | ((uint64_t)f2bf(b) << 16) | ((uint64_t)f2bf(c) << 32) |
           ((uint64_t)f2bf(d) << 48);
  };
  auto rest = [&](float w) { return w - bf2f(f2bf(w)); };  // the lo part of a split
  const int lane = t & 63, row = lane & 15, qg = lane >> 4;
  if (t < kT1) {  // layer 1: M-tile t / 64, K = inputs 4 qg .. 4 qg + 3 (only 0..3 exist)
    const int o = (t >> 6) * 16 + row, i = 4 * qg;
    float w4[4];
    for (int e = 0; e < 4; ++e) w4[e] = W(1, o, i + e);
    fwd[t] = pack(w4[0], w4[1], w4[2], w4[3]);
    fwd[kFwdSet + t] = pack(rest(w4[0]), rest(w4[1]), rest(w4[2]), rest(w4[3]));
    return;
  }
  if (t < kT1 + kT32) {  // layers 2-4: fragment f, K slot e -> input (2kc + e/4) 16 + 4 qg + e%4
    const int f = (t - kT1) >> 6;
    const int layer = f < 8 ? 2 : f < 16 ? 3 : 4;
    const int g = f < 8 ? f : f < 16 ? f - 8 : f - 16;
    const int mt = layer == 4 ? 0 : g >> 1, kc = layer == 4 ? g : g & 1;
    const int o = mt * 16 + row;
    float w8[8];
    for (int e = 0; e < 8; ++e) w8[e] = W(layer, o, (2 * kc + (e >> 2)) * 16 + 4 * qg + (e & 3));
    const int q = kT1 + (f * 64 + lane) * 2;
    fwd[q] = pack(w8[0], w8[1], w8[2], w8[3]);
    fwd[q + 1] = pack(w8[4], w8[5], w8[6], w8[7]);
    fwd[kFwdSet + q] = pack(rest(w8[0]), rest(w8[1]), rest(w8[2]), rest(w8[3]));
    fwd[kFwdSet + q + 1] = pack(rest(w8[4]), rest(w8[5]), rest(w8[6]), rest(w8[7]));
    return;
  }
  if (t < kT1 + kT32 + kTB) {
    const int bf = (t - kT1 - kT32) >> 6;
    const int c0 = 4 * qg;
    int layer, mt, ks;  // A = W^T: M = in (mt), K = out (ks)
    if (bf < 16) { layer = 2; mt = bf / 4; ks = bf % 4; }
    else if (bf < 32) { layer = 3; mt = (bf - 16) / 4; ks = (bf - 16) % 4; }
    else { layer = 4; mt = bf - 32; ks = 0; }
    const int i = mt * 16 + row, o = ks * 16 + c0;
    bwd[bf * 64 + lane] = pack(W(layer, o, i), W(layer, o + 1, i), W(layer, o + 2, i), W(layer, o + 3, i));
    return;
  }
  const int qb = t - (kT1 + kT32 + kTB);
  if (qb < kNBias)
    bias[qb] = qb < kB2 ? param[oB1 + qb] : qb < kB3 ? param[oB2 + qb - kB2]
               : qb < kB4 ? param[oB3 + qb - kB3] : param[oB4 + qb - kB4];
}

// transpose a 16x16 tile held in C layout (lane: column l%16, rows 4(l/16)+r) into the
// A/B-fragment layout (lane: row l%16, columns 4(l/16)+e), through LDS
__device__ __forceinline__ bf16x4 tile_transpose(uint16_t *lds, const f32x4 v) {
  const int lane = threadIdx.x & 63;
  const int col = lane & 15, r0 = 4 * (lane >> 4);
#pragma unroll
  for (int r = 0; r < 4; ++r) lds[(r0 + r) * 16 + col] = f2bf(v[r]);
  __syncthreads();
  bf16x4 out;
#pragma unroll
  for (int e = 0; e < 4; ++e) out[e] = (short)lds[col * 16 + r0 + e];
  __syncthreads();
  return out;
}

// One wavefront per block; each block walks `chunk` samples in tiles of 16, accumulates
// the loss gradient of every parameter and writes it, with its loss, to its own row of
// `partial` ([blocks][kNParam + 1]); k_mlp_reduce then sums the rows in block order, so a
// gradient does not depend on the order the blocks ran in (no float atomics).
// Loss per sample, with t = log(mu* z0): (o1 - t)^2 + lam (o2 - log(|1 - exp(t - o1)| +
// 1e-12))^2 (the second target, the log relative miss of the centre, held fixed),
// averaged over the n samples of the step.
__global__ __launch_bounds__(64) void k_mlp_train(MlpDev m, long n, int chunk,
                                                  const float *__restrict__ feat,
                                                  const float *__restrict__ target,
                                                  float lam, float *__restrict__ partial) {
  float *__restrict__ grad = partial + (long)blockIdx.x * (kNParam + 1);
  __shared__ uint16_t tl[256];
  const int lane = threadIdx.x;
  const int col = lane & 15, r0 = 4 * (lane >> 4);
  const long s0 = (long)blockIdx.x * chunk;
  const float inv_n = 1.0f / (float)n;
  // weight-gradient accumulators (C layout tiles): dW1 4 tiles (out) x 1 (in 16),
  // dW2, dW3 16 tiles, dW4 1 (out 16) x 4 (in); bias gradients per lane
  f32x4 gW1[4], gW2[16], gW3[16], gW4[4];
  float gb1[4][4] = {}, gb2[4][4] = {}, gb3[4][4] = {}, gb4[2] = {0.f, 0.f};
#pragma unroll
  for (int t = 0; t < 16; ++t) gW2[t] = gW3[t] = (f32x4){0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int t = 0; t < 4; ++t) gW1[t] = gW4[t] = (f32x4){0.f, 0.f, 0.f, 0.f};
  float lsum = 0.f;
  for (int base = 0; base < chunk; base += 16) {
    const long s = s0 + base + col;
    const bool have = s < n;
    // B fragment of the features: lanes 0..15 = features 0..3 of sample s
    Split xs = split4(0.f, 0.f, 0.f, 0.f);
    if (lane < 16 && have) xs = split4(feat[s], feat[n + s], feat[2 * n + s], feat[3 * n + s]);
    const bf16x4 x = xs.hi;
    f32x4 h1[4], h2[4], h3[4];
    f32x4 o = mlp_tile_forward(m, xs, h1, h2, h3);
    // loss gradient (lanes 0..15 own the outputs of sample s)
    f32x4 dO = {0.f, 0.f, 0.f, 0.f};
    if (lane < 16 && have) {
      const float mu = fminf(fmaxf(target[s], 1e-30f), 1.0f);
      const float t1 = logf(mu * prior_z0(m, feat[s], feat[2 * n + s]));
      const float t2 = logf(fabsf(1.0f - expf(fminf(fmaxf(t1 - o[0], -30.f), 30.f))) + 1e-12f);
      const float e1 = o[0] - t1, e2 = o[1] - t2;
      lsum += (e1 * e1 + lam * e2 * e2) * inv_n;
      dO[0] = 2.0f * e1 * inv_n;
      dO[1] = 2.0f * lam * e2 * inv_n;
    }
    // backward through layer 4: dH3^T = W4^T . dO^T
    const bf16x4 bdo = pack4(dO[0], dO[1], dO[2], dO[3]);
    f32x4 d3[4], d2[4], d1[4];
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
      f32x4 acc = {0.f, 0.f, 0.f, 0.f};
      acc = mfma(as_bf16x4(m.bwd[(32 + mt) * 64 + lane]), bdo, acc);
#pragma unroll
      for (int r = 0; r < 4; ++r) acc[r] = h3[mt][r] > 0.f ? acc[r] : 0.f;
      d3[mt] = acc;
    }
    // layer 3 and 2 backward: dH_{l-1}^T = W_l^T . dZ_l^T, masked by the activation
#pragma unroll
    for (int layer = 3; layer >= 2; --layer) {
      const int base = layer == 3 ? 16 : 0;
      const f32x4 *dz = layer == 3 ? d3 : d2;
      const f32x4 *hp = layer == 3 ? h2 : h1;
      f32x4 *dout = layer == 3 ? d2 : d1;
      bf16x4 bin[4];
#pragma unroll
      for (int ks = 0; ks < 4; ++ks) bin[ks] = pack4(dz[ks][0], dz[ks][1], dz[ks][2], dz[ks][3]);
#pragma unroll
      for (int mt = 0; mt < 4; ++mt) {
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int ks = 0; ks < 4; ++ks)
          acc = mfma(as_bf16x4(m.bwd[(base + mt * 4 + ks) * 64 + lane]), bin[ks], acc);
#pragma unroll
        for (int r = 0; r < 4; ++r) acc[r] = hp[mt][r] > 0.f ? acc[r] : 0.f;
        dout[mt] = acc;
      }
    }
    // weight gradients dW = dZ^T . H (K = the 16 samples), operands transposed via LDS
    {
      const bf16x4 a4 = tile_transpose(tl, dO);  // A: rows = outputs, K = samples
#pragma unroll
      for (int nt = 0; nt < 4; ++nt) gW4[nt] = mfma(a4, tile_transpose(tl, h3[nt]), gW4[nt]);
    }
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
      const bf16x4 a3 = tile_transpose(tl, d3[mt]);
      const bf16x4 a2 = tile_transpose(tl, d2[mt]);
      const bf16x4 a1 = tile_transpose(tl, d1[mt]);
#pragma unroll
      for (int nt = 0; nt < 4; ++nt) {
        gW3[mt * 4 + nt] = mfma(a3, tile_transpose(tl, h2[nt]), gW3[mt * 4 + nt]);
        gW2[mt * 4 + nt] = mfma(a2, tile_transpose(tl, h1[nt]), gW2[mt * 4 + nt]);
      }
      // layer 1 input tile: the features as a C-layout tile (rows = features)
      f32x4 xin = {0.f, 0.f, 0.f, 0.f};
      if (lane < 16) {
#pragma unroll
        for (int r = 0; r < 4; ++r) xin[r] = bf2f((uint16_t)x[r]);
      }
      gW1[mt] = mfma(a1, tile_transpose(tl, xin), gW1[mt]);
      // bias gradients: sum over the 16 samples of this tile (per lane first)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        gb1[mt][r] += d1[mt][r];
        gb2[mt][r] += d2[mt][r];
        gb3[mt][r] += d3[mt][r];
      }
    }
    gb4[0] += dO[0];
    gb4[1] += dO[1];
  }
  // reduce bias partials over the 16 sample columns (lanes sharing l/16)
#pragma unroll
  for (int mt = 0; mt < 4; ++mt)
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
      for (int off = 1; off < 16; off <<= 1) {
        gb1[mt][r] += __shfl_xor(gb1[mt][r], off, 64);
        gb2[mt][r] += __shfl_xor(gb2[mt][r], off, 64);
        gb3[mt][r] += __shfl_xor(gb3[mt][r], off, 64);
      }
#pragma unroll
  for (int off = 1; off < 16; off <<= 1) {
    gb4[0] += __shfl_xor(gb4[0], off, 64);
    gb4[1] += __shfl_xor(gb4[1], off, 64);
    lsum += __shfl_xor(lsum, off, 64);
  }
  // scatter: C layout of dW tiles (lane: column = in index col, rows = out r0 + r)
#pragma unroll
  for (int mt = 0; mt < 4; ++mt)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int o = mt * 16 + r0 + r;
      if (col < kIn) grad[oW1 + o * kIn + col] = gW1[mt][r];
#pragma unroll
      for (int nt = 0; nt < 4; ++nt) {
        grad[oW2 + o * kHid + nt * 16 + col] = gW2[mt * 4 + nt][r];
        grad[oW3 + o * kHid + nt * 16 + col] = gW3[mt * 4 + nt][r];
      }
      if (col == 0) {
        grad[oB1 + o] = gb1[mt][r];
        grad[oB2 + o] = gb2[mt][r];
        grad[oB3 + o] = gb3[mt][r];
      }
    }
#pragma unroll
  for (int nt = 0; nt < 4; ++nt)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int o = r0 + r;
      if (o < kOut) grad[oW4 + o * kHid + nt * 16 + col] = gW4[nt][r];
    }
  if (lane == 0) {
    grad[oB4 + 0] = gb4[0];
    grad[oB4 + 1] = gb4[1];
    grad[kNParam] = lsum;
  }
}

// grad[t] += the sum of the blocks' partial rows in block order; the loss (slot kNParam)
// goes to *loss the same way
__global__ void k_mlp_reduce(const float *__restrict__ partial, int blocks,
                             float *__restrict__ grad, float *__restrict__ loss) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t > kNParam) return;
  float acc = 0.f;
  for (int b = 0; b < blocks; ++b) acc += partial[(long)b * (kNParam + 1) + t];
  if (t < kNParam)
    grad[t] += acc;
  else
    *loss = acc;
}

__global__ void k_mlp_adam(float *__restrict__ param, float *__restrict__ grad,
                           float *__restrict__ m1, float *__restrict__ m2, float lr, float b1,
                           float b2, float eps, float c1, float c2) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= kNParam) return;
  const float g = grad[t];
  const float a = b1 * m1[t] + (1.f - b1) * g;
  const float v = b2 * m2[t] + (1.f - b2) * g * g;
  m1[t] = a;
  m2[t] = v;
  param[t] -= lr * (a / c1) / (sqrtf(v / c2) + eps);
  grad[t] = 0.f;
}

__global__ __launch_bounds__(64) void k_mlp_forward(MlpDev m, long n,
                                                    const float *__restrict__ feat,
                                                    float *__restrict__ out) {
  const int lane = threadIdx.x;
  const long s = (long)blockIdx.x * 16 + (lane & 15);
  Split x = split4(0.f, 0.f, 0.f, 0.f);
  if (lane < 16 && s < n) x = split4(feat[s], feat[n + s], feat[2 * n + s], feat[3 * n + s]);
  const f32x4 o = mlp_tile_forward(m, x, nullptr, nullptr, nullptr);
  if (lane < 16 && s < n) {
    out[s] = o[0];
    out[n + s] = o[1];
  }
}

// ---- con2prim with the learned bracket (batched states) --------------------------------
// The packed forward weights (hi and lo, 40 KB) and the biases are copied into LDS once
// per block of 256 states (an L2-resident read); every wavefront then evaluates the
// network for its 64 states from LDS.  (A persistent grid that loads them once per CU
// was tried: the loop makes the compiler keep the kernel arguments in scratch and hoist
// loop invariants into ~270 VGPRs, which costs more than the per-block copy.)
constexpr int kLearnedThreads = 256;
// training launches at most this many blocks (their partial gradients are summed in order)
constexpr int kMaxTrainBlocks = 1024;

// three waves per SIMD (the packed weights in LDS allow no more): the MFMA accumulators then
// live in VGPRs and nothing spills
template <bool MHD, int GEOM>
__global__ __launch_bounds__(kLearnedThreads)
__attribute__((amdgpu_waves_per_eu(3))) void k_con2prim_batch_learned(
    DevParams p, MlpDev m, long n, const double *__restrict__ coords, double *__restrict__ cons,
    double *__restrict__ prim, double *__restrict__ mu, double *__restrict__ fail,
    int32_t *__restrict__ iters, int32_t *__restrict__ evals, int32_t *__restrict__ fallback) {
  using M = typename MetricOf<GEOM>::type;
  using L = Layout<MHD>;
  __shared__ uint64_t wfwd[kFwdU64];
  __shared__ float wbias[kNBias];
  for (int q = threadIdx.x; q < kFwdU64; q += kLearnedThreads) wfwd[q] = m.fwd[q];
  for (int q = threadIdx.x; q < kNBias; q += kLearnedThreads) wbias[q] = m.bias[q];
  __syncthreads();
  MlpDev ml = m;
  ml.fwd = wfwd;
  ml.bias = wbias;
  {
    const long base = (long)blockIdx.x * kLearnedThreads;
    const long s = base + threadIdx.x;
    const bool have = s < n;  // every lane stays for the wave-wide network evaluation
    const long sc = have ? s : n - 1;
    // phase 1: features of this lane's state
    float f[4] = {0.f, 0.f, 0.f, 0.f};
    {
      double X[3] = {0.0, 0.0, 0.0};
      if (coords) {
        X[0] = coords[sc];
        X[1] = coords[n + sc];
      }
      const M g(p, X[0], X[1], X[2]);
      double cs[L::ncons], pr[L::nprim];
#pragma unroll
      for (int v = 0; v < L::ncons; ++v) cs[v] = cons[v * n + sc];
      C2PState z;
      if (c2p_setup<MHD>(p, g, X, cs, pr, z)) zone_features(ml, z.res, f);
    }
    // phase 2: the network for the wave's 64 states
    float o1, o2;
    mlp_predict_wave(ml, f, o1, o2);
    if (!have) return;
    // phase 3: set-up again (bit for bit the same) and the bracketed solve.  Reading the
    // inputs through pointers the compiler cannot relate to phase 1's keeps phase 1's
    // set-up from staying live across the network.
    const double *cin = cons, *xin = coords;
    asm volatile("" : "+s"(cin), "+s"(xin));
    double X[3] = {0.0, 0.0, 0.0};
    if (xin) {
      X[0] = xin[sc];
      X[1] = xin[n + sc];
    }
    const M g(p, X[0], X[1], X[2]);
    double cs[L::ncons], pr[L::nprim], sig[3];
#pragma unroll
    for (int v = 0; v < L::ncons; ++v) cs[v] = cin[v * n + sc];
#pragma unroll
    for (int v = 0; v < L::nprim; ++v) pr[v] = 0.0;
    C2PState z;
    const bool ok = c2p_setup<MHD>(p, g, X, cs, pr, z);
    double st = 0.0, root = mu[s], tt = 0.0;
    int it = 0, ev = 0, fb = 0;
    if (ok) {
      const double c = exp((double)o1) / (double)prior_z0(ml, f[0], f[2]);
      const double w = c * fmax((double)ml.width_scale * exp((double)o2), (double)ml.width_min);
      const double lo = fmax(c - w, 0.0), hi = fmin(c + w, 1.0);
      const double ylo = z.res(lo), yhi = z.res(hi);
      ev = 2;
      bool good = true;
      if (ylo * yhi <= 0.0) {
        root = regula_falsi_bracketed(z.res, lo, hi, ylo, yhi, p.tol, p.max_iter, it);
      } else {
        fb = 1;
        // the plain solve with the caller's guess when there is one, else the centre
        good = regula_falsi(z.res, 0.0, 1.0, p.tol, root > 0.0 ? root : c, p.max_iter, root, it);
        ev += 2;
        if (!good) it = -1;
      }
      if (good) {
        ev += it;
        st = c2p_unwrap<MHD>(p, g, z, root, cs, pr, tt, sig);
      }
    }
#pragma unroll
    for (int v = 0; v < 5; ++v) cons[v * n + s] = cs[v];
#pragma unroll
    for (int v = 0; v < L::nprim; ++v) prim[v * n + s] = pr[v];
    mu[s] = root;
    fail[s] = st;
    if (iters) iters[s] = it;
    if (evals) evals[s] = ev;
    if (fallback) fallback[s] = fb;
  }
}

// con2prim with the learned bracket over the cells a plain pass of a pack queued (those
// with no previous root: initial data, fresh cells), with the cell-field semantics of the
// plain pass (con2prim_cell: NaN input -> fail, every other field kept; an unbracketed
// root -> fail with iters -1) and the bracket and fallback of k_con2prim_batch_learned.
// A grid-stride loop over the queue's count, so a stage whose cells all have a previous
// root (a time-dependent run) costs one launch that reads the count and ends; a workgroup
// with no queued cell loads no weights.
// Two waves per SIMD: 243 VGPRs on Minkowski blocks, 255-256 with 14-18 spilled on the
// curved metrics, which left to themselves take 274-280 and run one wave per SIMD (the
// first stage of the 512-block torus pack: 25.4 ms against 27.4); three waves would spill
// 82 on Minkowski.
constexpr int kColdGrid = 768;  // workgroups: 3 per CU (the LDS of the weights allows 3)
template <int GEOM>
__global__ __launch_bounds__(kLearnedThreads)
__attribute__((amdgpu_waves_per_eu(2))) void
k_pack_con2prim_cold(DevParams p, MlpDev m, PackView v, ColdQueue q,
                     unsigned long long *counters, double *extra) {
  using M = typename MetricOf<GEOM>::type;
  constexpr bool CACHED = GEOM == PHX_GEOM_FMKS;  // every FMKS block of a pack has its geometry
  using L = Layout<true>;
  __shared__ uint64_t wfwd[kFwdU64];
  __shared__ float wbias[kNBias];
  const long nq = (long)min((unsigned long long)*q.count, (unsigned long long)q.cap);
  if ((long)blockIdx.x * kLearnedThreads >= nq) return;  // the whole workgroup: no cell
  for (int t = threadIdx.x; t < kFwdU64; t += kLearnedThreads) wfwd[t] = m.fwd[t];
  for (int t = threadIdx.x; t < kNBias; t += kLearnedThreads) wbias[t] = m.bias[t];
  __syncthreads();
  MlpDev ml = m;
  ml.fwd = wfwd;
  ml.bias = wbias;
  // the loop's trip count is the same for every lane of the workgroup: the network is
  // evaluated by whole wavefronts
  for (long base = (long)blockIdx.x * kLearnedThreads; base < nq;
       base += (long)gridDim.x * kLearnedThreads) {
    const long e = base + threadIdx.x;
    const bool have = e < nq;
    const unsigned long long w = have ? q.cell[e] : 0ull;
    const PackEntry &E = v.e[(int)(w >> 32)];
    const BlockGeom &bg = E.bg;
    const int ci = (int)(w & 0x7fffffffull);  // cell index (< 2^31)
    const long c = ci;
    const bool keep = (w >> 31) & 1ull;
    const int idx[3] = {ci % bg.ni, (ci / bg.ni) % bg.nj, ci / (bg.ni * bg.nj)};
    int32_t *iters = keep ? nullptr : E.iters;
    const long n = bg.ncell;
    // phase 1: features of this lane's cell
    float f[4] = {0.f, 0.f, 0.f, 0.f};
    if (have) {
      const double X[3] = {xc(bg, 0, idx[0]), xc(bg, 1, idx[1]), xc(bg, 2, idx[2])};
      const M gm = metric_at<GEOM, CACHED>(p, bg, 0, (long)idx[1] * bg.ni + idx[0], X[0], X[1], X[2]);
      double cs[L::ncons], pr[L::nprim];
#pragma unroll
      for (int t = 0; t < L::ncons; ++t) cs[t] = E.u_out[t * n + c];
      C2PState z;
      if (c2p_setup<true>(p, gm, X, cs, pr, z)) zone_features(ml, z.res, f);
    }
    // phase 2: the network for the wave's 64 cells
    float o1, o2;
    mlp_predict_wave(ml, f, o1, o2);
    if (!have) continue;
    // phase 3: the set-up again, then the bracketed solve
    double *cons = E.u_out, *prim = E.prim_out;
    double cs[L::ncons], pr[L::nprim], sig[3];
#pragma unroll
    for (int t = 0; t < L::ncons; ++t) cs[t] = cons[t * n + c];
#pragma unroll
    for (int d = 0; d < 3; ++d) sig[d] = d < bg.ndim ? E.csig[d * n + c] : 0.0;
    // the cell's CFL time with face speeds, as the update forms it (a cell the inversion
    // leaves keeps its signal speeds and still counts)
    auto fold_dt = [&]() {
      if (!extra) return;
      const long off[3] = {1, bg.ni, (long)bg.ni * bg.nj};
      double ldt = 0.0;
      for (int d = 0; d < bg.ndim; ++d) {
        const double s = smax(sig[d], smax(E.fsig[d * n + c], E.fsig[d * n + c + off[d]]));
        ldt += s / bg.dx[d];
      }
      const double t = 1.0 / (ldt + 1.e-50);
      if (!isnan(t))  // positive doubles order as their bit patterns
        atomicMin(reinterpret_cast<unsigned long long *>(extra),
                  (unsigned long long)__double_as_longlong(t));
    };
    if (isnan(cs[0]) + isnan(cs[1]) + isnan(cs[4]) > 0) {
      E.fail[c] = 0.0;
      if (iters) iters[c] = 0;
      fold_dt();
      continue;
    }
    const double X[3] = {xc(bg, 0, idx[0]), xc(bg, 1, idx[1]), xc(bg, 2, idx[2])};
    const M gm = metric_at<GEOM, CACHED>(p, bg, 0, (long)idx[1] * bg.ni + idx[0], X[0], X[1], X[2]);
#pragma unroll
    for (int t = 0; t < L::nprim; ++t) pr[t] = prim[t * n + c];
    double tt = E.temp ? E.temp[c] : 0.0, root = E.mu[c];
    C2PState z;
    c2p_setup<true>(p, gm, X, cs, pr, z);
    const double cc = exp((double)o1) / (double)prior_z0(ml, f[0], f[2]);
    const double wd = cc * fmax((double)ml.width_scale * exp((double)o2), (double)ml.width_min);
    const double lo = fmax(cc - wd, 0.0), hi = fmin(cc + wd, 1.0);
    const double ylo = z.res(lo), yhi = z.res(hi);
    int it = 0;
    bool good = true, fb = false;
    double mu_new = root;
    if (ylo * yhi <= 0.0) {
      mu_new = regula_falsi_bracketed(z.res, lo, hi, ylo, yhi, p.tol, p.max_iter, it);
    } else {  // the plain solve on [0, 1], split at the network's centre
      fb = true;
      good = regula_falsi(z.res, 0.0, 1.0, p.tol, cc, p.max_iter, mu_new, it);
    }
    double st = 0.0;
    if (good) {
      root = mu_new;
      st = c2p_unwrap<true>(p, gm, z, root, cs, pr, tt, sig);
    } else {
      it = -1;
    }
    if (counters) {  // [0] bracket misses, [1] network-bracketed solves
      if (fb) atomicAdd(&counters[0], 1ull);
      const unsigned long long act = __ballot(1);
      if ((int)(threadIdx.x & 63) == __ffsll((long long)act) - 1)
        atomicAdd(&counters[1], (unsigned long long)__popcll(act));
    }
#pragma unroll
    for (int t = 0; t < 5; ++t) cons[t * n + c] = cs[t];
#pragma unroll
    for (int t = 0; t < L::nprim; ++t) prim[t * n + c] = pr[t];
    for (int d = 0; d < bg.ndim; ++d) E.csig[d * n + c] = sig[d];
    if (E.temp) E.temp[c] = tt;
    E.mu[c] = root;
    E.fail[c] = st;
    if (iters) iters[c] = it;
    fold_dt();
  }
}

// features of batched states (training data and tests)
template <bool MHD, int GEOM>
__global__ __launch_bounds__(256) void k_c2p_features(DevParams p, MlpDev m, long n,
                                                      const double *__restrict__ coords,
                                                      const double *__restrict__ cons,
                                                      float *__restrict__ feat) {
  using M = typename MetricOf<GEOM>::type;
  using L = Layout<MHD>;
  const long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  double X[3] = {0.0, 0.0, 0.0};
  if (coords) {
    X[0] = coords[s];
    X[1] = coords[n + s];
  }
  const M g(p, X[0], X[1], X[2]);
  double cs[L::ncons], pr[L::nprim];
#pragma unroll
  for (int v = 0; v < L::ncons; ++v) cs[v] = cons[v * n + s];
  C2PState z;
  float f[4] = {0.f, 0.f, 0.f, 0.f};
  if (c2p_setup<MHD>(p, g, X, cs, pr, z)) zone_features(m, z.res, f);
  for (int i = 0; i < 4; ++i) feat[i * n + s] = f[i];
}

}  // namespace phx

using namespace phx;

// ================================ host side ==============================================
struct phx_mlp {
  int device = 0;
  float *param = nullptr, *grad = nullptr, *m1 = nullptr, *m2 = nullptr, *loss = nullptr;
  float *partial = nullptr;  // [kMaxTrainBlocks][kNParam + 1] per-block gradients
  uint64_t *fwd = nullptr, *bwd = nullptr;
  float *bias = nullptr;
  float fmean[4] = {0, 0, 0, 0}, fstd[4] = {1, 1, 1, 1};
  float width_scale = 4.0f, width_min = 1e-9f;
  long step = 0;
};

static MlpDev dev_view(const phx_mlp *m) {
  MlpDev d;
  d.fwd = m->fwd;
  d.bwd = m->bwd;
  d.bias = m->bias;
  d.param = m->param;
  for (int i = 0; i < 4; ++i) {
    d.fmean[i] = m->fmean[i];
    d.fstd[i] = m->fstd[i];
  }
  d.width_scale = m->width_scale;
  d.width_min = m->width_min;
  return d;
}

void phx::launch_pack_c2p_cold(int geom, const DevParams &dp, const phx_mlp *m,
                               const PackView &v, const ColdQueue &cq,
                               unsigned long long *counters, double *extra, hipStream_t s) {
  if (!cq.cell) return;
  const MlpDev md = dev_view(m);
  with_geom(geom, [&](auto G) {
    hipLaunchKernelGGL((k_pack_con2prim_cold<decltype(G)::value>), dim3(kColdGrid),
                       dim3(kLearnedThreads), 0, s, dp, md, v, cq, counters, extra);
  });
}

static void repack(phx_mlp *m, hipStream_t s) {
  const int threads = (4 + kFwd32Frags + kBwdFrags) * 64 + kNBias;
  hipLaunchKernelGGL(k_mlp_pack, dim3((threads + 255) / 256), dim3(256), 0, s, m->param, m->fwd,
                     m->bwd, m->bias);
}

extern "C" {

int phx_mlp_create(phx_ctx *ctx, phx_mlp **out, uint64_t seed) {
  if (!ctx || !out) return PHX_ERR_INVALID;
  phx_mlp *m = new (std::nothrow) phx_mlp();
  if (!m) return PHX_ERR_NOMEM;
  m->device = ctx->device;
  bool okm = hipMalloc(&m->param, sizeof(float) * kNParam) == hipSuccess &&
             hipMalloc(&m->grad, sizeof(float) * kNParam) == hipSuccess &&
             hipMalloc(&m->m1, sizeof(float) * kNParam) == hipSuccess &&
             hipMalloc(&m->m2, sizeof(float) * kNParam) == hipSuccess &&
             hipMalloc(&m->loss, sizeof(float)) == hipSuccess &&
             hipMalloc(&m->partial, sizeof(float) * kMaxTrainBlocks * (kNParam + 1)) == hipSuccess &&
             hipMalloc(&m->fwd, sizeof(uint64_t) * kFwdU64) == hipSuccess &&
             hipMalloc(&m->bwd, sizeof(uint64_t) * kBwdFrags * 64) == hipSuccess &&
             hipMalloc(&m->bias, sizeof(float) * kNBias) == hipSuccess;
  if (!okm) {
    phx_mlp_destroy(m);
    return PHX_ERR_NOMEM;
  }
  // He-uniform initialization from a 64-bit LCG (deterministic per seed), zero biases
  std::vector<float> w(kNParam, 0.f);
  uint64_t st = seed * 6364136223846793005ULL + 1442695040888963407ULL;
  auto uni = [&]() {
    st = st * 6364136223846793005ULL + 1442695040888963407ULL;
    return ((st >> 11) * (1.0 / 9007199254740992.0)) * 2.0 - 1.0;
  };
  auto fill = [&](int off, int nout, int nin) {
    const double lim = std::sqrt(6.0 / nin);
    for (int k = 0; k < nout * nin; ++k) w[off + k] = (float)(lim * uni());
  };
  fill(oW1, kHid, kIn);
  fill(oW2, kHid, kHid);
  fill(oW3, kHid, kHid);
  fill(oW4, kOut, kHid);
  if (phx_mlp_set_params(ctx, m, w.data()) != PHX_OK) {
    phx_mlp_destroy(m);
    return PHX_ERR_DEVICE;
  }
  *out = m;
  return PHX_OK;
}

int phx_mlp_destroy(phx_mlp *m) {
  if (!m) return PHX_OK;
  (void)hipSetDevice(m->device);
  for (void *p : {(void *)m->param, (void *)m->grad, (void *)m->m1, (void *)m->m2,
                  (void *)m->loss, (void *)m->partial, (void *)m->fwd, (void *)m->bwd,
                  (void *)m->bias})
    if (p) (void)hipFree(p);
  delete m;
  return PHX_OK;
}

int phx_mlp_set_params(phx_ctx *ctx, phx_mlp *m, const float *host_params) {
  if (!ctx || !m || !host_params) return PHX_ERR_INVALID;
  if (hipMemcpy(m->param, host_params, sizeof(float) * kNParam, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(m->grad, 0, sizeof(float) * kNParam) != hipSuccess ||
      hipMemset(m->m1, 0, sizeof(float) * kNParam) != hipSuccess ||
      hipMemset(m->m2, 0, sizeof(float) * kNParam) != hipSuccess)
    return ctx->fail(PHX_ERR_DEVICE, "mlp upload failed");
  m->step = 0;
  repack(m, ctx->stream);
  if (hipStreamSynchronize(ctx->stream) != hipSuccess) return ctx->fail(PHX_ERR_DEVICE, "mlp pack failed");
  return PHX_OK;
}

int phx_mlp_get_params(phx_ctx *ctx, const phx_mlp *m, float *host_params) {
  if (!ctx || !m || !host_params) return PHX_ERR_INVALID;
  if (hipStreamSynchronize(ctx->stream) != hipSuccess ||
      hipMemcpy(host_params, m->param, sizeof(float) * kNParam, hipMemcpyDeviceToHost) != hipSuccess)
    return ctx->fail(PHX_ERR_DEVICE, "mlp download failed");
  return PHX_OK;
}

int phx_mlp_set_normalization(phx_ctx *ctx, phx_mlp *m, const float mean[4], const float std_[4],
                              float width_scale, float width_min) {
  if (!ctx || !m || !mean || !std_) return PHX_ERR_INVALID;
  for (int i = 0; i < 4; ++i) {
    if (!(std_[i] > 0.f)) return ctx->fail(PHX_ERR_INVALID, "feature std must be positive");
    m->fmean[i] = mean[i];
    m->fstd[i] = std_[i];
  }
  if (!(width_scale > 0.f) || !(width_min >= 0.f)) return ctx->fail(PHX_ERR_INVALID, "bad width");
  m->width_scale = width_scale;
  m->width_min = width_min;
  return PHX_OK;
}

int phx_mlp_gradient(phx_ctx *ctx, phx_mlp *m, int64_t n, const float *features,
                     const float *target_mu, float lam, float *loss_out, phx_stream stream) {
  if (!ctx || !m || n <= 0 || !features || !target_mu) return PHX_ERR_INVALID;
  hipStream_t s = ctx->stream_of(stream);
  // at least 256 samples per block, at most kMaxTrainBlocks blocks (chunks of whole tiles)
  long chunk = 256;
  if ((n + chunk - 1) / chunk > kMaxTrainBlocks)
    chunk = ((n + kMaxTrainBlocks - 1) / kMaxTrainBlocks + 15) / 16 * 16;
  const long blocks = (n + chunk - 1) / chunk;
  hipLaunchKernelGGL(k_mlp_train, dim3((unsigned)blocks), dim3(64), 0, s, dev_view(m), (long)n,
                     (int)chunk, features, target_mu, lam, m->partial);
  hipLaunchKernelGGL(k_mlp_reduce, dim3((kNParam + 256) / 256), dim3(256), 0, s, m->partial,
                     (int)blocks, m->grad, m->loss);
  if (loss_out) (void)hipMemcpyAsync(loss_out, m->loss, sizeof(float), hipMemcpyDeviceToDevice, s);
  return ctx->check_launch();
}

int phx_mlp_read_gradient(phx_ctx *ctx, phx_mlp *m, float *host_grad) {
  if (!ctx || !m || !host_grad) return PHX_ERR_INVALID;
  if (hipStreamSynchronize(ctx->stream) != hipSuccess ||
      hipMemcpy(host_grad, m->grad, sizeof(float) * kNParam, hipMemcpyDeviceToHost) != hipSuccess)
    return ctx->fail(PHX_ERR_DEVICE, "gradient download failed");
  return PHX_OK;
}

int phx_mlp_zero_gradient(phx_ctx *ctx, phx_mlp *m) {
  if (!ctx || !m) return PHX_ERR_INVALID;
  if (hipMemset(m->grad, 0, sizeof(float) * kNParam) != hipSuccess)
    return ctx->fail(PHX_ERR_DEVICE, "memset failed");
  return PHX_OK;
}

int phx_mlp_adam_step(phx_ctx *ctx, phx_mlp *m, float lr, phx_stream stream) {
  if (!ctx || !m || !(lr > 0.f)) return PHX_ERR_INVALID;
  hipStream_t s = ctx->stream_of(stream);
  m->step += 1;
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.f - std::pow(b1, (float)m->step), c2 = 1.f - std::pow(b2, (float)m->step);
  hipLaunchKernelGGL(k_mlp_adam, dim3((kNParam + 255) / 256), dim3(256), 0, s, m->param, m->grad,
                     m->m1, m->m2, lr, b1, b2, 1e-8f, c1, c2);
  repack(m, s);
  return ctx->check_launch();
}

int phx_mlp_train_step(phx_ctx *ctx, phx_mlp *m, int64_t n, const float *features,
                       const float *target_mu, float lr, float lam, float *loss_out,
                       phx_stream stream) {
  int rc = phx_mlp_gradient(ctx, m, n, features, target_mu, lam, loss_out, stream);
  if (rc) return rc;
  return phx_mlp_adam_step(ctx, m, lr, stream);
}

int phx_mlp_forward(phx_ctx *ctx, const phx_mlp *m, int64_t n, const float *features,
                    float *out, phx_stream stream) {
  if (!ctx || !m || n < 0 || (n > 0 && (!features || !out))) return PHX_ERR_INVALID;
  if (n == 0) return PHX_OK;
  hipStream_t s = ctx->stream_of(stream);
  hipLaunchKernelGGL(k_mlp_forward, dim3((unsigned)((n + 15) / 16)), dim3(64), 0, s, dev_view(m),
                     (long)n, features, out);
  return ctx->check_launch();
}

int phx_con2prim_features(phx_ctx *ctx, const phx_fluid_params *p, const phx_mlp *m, int64_t n,
                          const double *coords, const double *cons, float *features,
                          phx_stream stream) {
  if (!ctx || !m) return PHX_ERR_INVALID;
  int rc = check_params(ctx, p);
  if (rc) return rc;
  if (n <= 0) return n == 0 ? PHX_OK : ctx->fail(PHX_ERR_INVALID, "negative batch");
  if (!cons || !features || (p->geom.kind != PHX_GEOM_MINKOWSKI && !coords))
    return ctx->fail(PHX_ERR_INVALID, "null field");
  const DevParams dp = make_dev_params(p);
  hipStream_t s = ctx->stream_of(stream);
  const dim3 grid((unsigned)((n + 255) / 256)), blk(256);
  const MlpDev md = dev_view(m);
  with_geom(p->geom.kind, [&](auto G) {
    with_mhd(p->mhd, [&](auto H) {
      hipLaunchKernelGGL((k_c2p_features<decltype(H)::value, decltype(G)::value>), grid, blk, 0, s, dp, md, (long)n, coords, cons, features);
    });
  });
  return ctx->check_launch();
}

int phx_con2prim_batch_learned(phx_ctx *ctx, const phx_fluid_params *p, const phx_mlp *m,
                               int64_t n, const double *coords, double *cons, double *prim,
                               double *mu, double *fail, int32_t *iters, int32_t *evals,
                               int32_t *fallback, phx_stream stream) {
  if (!ctx || !m) return PHX_ERR_INVALID;
  int rc = check_params(ctx, p);
  if (rc) return rc;
  if (n <= 0) return n == 0 ? PHX_OK : ctx->fail(PHX_ERR_INVALID, "negative batch");
  if (!cons || !prim || !mu || !fail || (p->geom.kind != PHX_GEOM_MINKOWSKI && !coords))
    return ctx->fail(PHX_ERR_INVALID, "null field");
  const DevParams dp = make_dev_params(p);
  hipStream_t s = ctx->stream_of(stream);
  const dim3 grid((unsigned)((n + kLearnedThreads - 1) / kLearnedThreads)), blk(kLearnedThreads);
  const MlpDev md = dev_view(m);
#define LARGS dp, md, (long)n, coords, cons, prim, mu, fail, iters, evals, fallback
  with_geom(p->geom.kind, [&](auto G) {
    with_mhd(p->mhd, [&](auto H) {
      hipLaunchKernelGGL((k_con2prim_batch_learned<decltype(H)::value, decltype(G)::value>), grid, blk, 0, s, LARGS);
    });
  });
#undef LARGS
  return ctx->check_launch();
}

}  // extern "C"
