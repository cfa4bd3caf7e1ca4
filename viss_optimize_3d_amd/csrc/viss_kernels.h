// Device-side building blocks shared by the forward / back-projection kernels
// (viss_nudft.hip) and the batched persistent forward (viss_batch.hip).
#pragma once

#include "viss_internal.h"

namespace viss {

// ------------------------------------------------------------ pair math
__device__ __forceinline__ float rev_fp32(double ph) {
  return (float)__builtin_amdgcn_fract(ph);  // [0,1) revolutions, exact in FP64
}

__device__ __forceinline__ void sincos_rev(float r, float& s, float& c) {
  s = __builtin_amdgcn_sinf(r);  // sin(2 pi r)
  c = __builtin_amdgcn_cosf(r);
}

__device__ __forceinline__ bool both_visible(double l, double m, double n, const VisMask& q) {
  const double da = fma(l, q.xa, fma(m, q.ya, n * q.za));
  const double db = fma(l, q.xb, fma(m, q.yb, n * q.zb));
  return (da >= q.ta) & (db >= q.tb);
}

// Conservative cone test of one visibility's two satellites against a tile.
__device__ __forceinline__ void classify(const float4 cn, const VisCls& c, bool& any_v, bool& all_v) {
  const float da = fmaf(cn.x, c.ax, fmaf(cn.y, c.ay, cn.z * c.az));
  const float db = fmaf(cn.x, c.bx, fmaf(cn.y, c.by, cn.z * c.bz));
  any_v = (da + cn.w >= c.ca - kClsEps) & (db + cn.w >= c.cb - kClsEps);
  all_v = (da - cn.w >= c.ca + kClsEps) & (db - cn.w >= c.cb + kClsEps);
}

__device__ __forceinline__ void make_mask_records(double xa, double ya, double za, double xb, double yb, double zb,
                                                  double R_moon, double orbit_r, VisMask& m, VisCls& c) {
  m.xa = xa; m.ya = ya; m.za = za;
  m.xb = xb; m.yb = yb; m.zb = zb;
  m.ta = mask_threshold(xa, ya, za, R_moon, orbit_r);
  m.tb = mask_threshold(xb, yb, zb, R_moon, orbit_r);
  const double ra = sqrt(xa * xa + ya * ya + za * za), rb = sqrt(xb * xb + yb * yb + zb * zb);
  const double ia = ra > 0 ? 1.0 / ra : 0.0, ib = rb > 0 ? 1.0 / rb : 0.0;
  c.ax = (float)(xa * ia); c.ay = (float)(ya * ia); c.az = (float)(za * ia);
  c.bx = (float)(xb * ib); c.by = (float)(yb * ib); c.bz = (float)(zb * ib);
  // a satellite at the origin sees everything with t = 0 only if s.0 >= 0: always true
  c.ca = ra > 0 ? (float)(m.ta * ia) : -2.0f;
  c.cb = rb > 0 ? (float)(m.tb * ib) : -2.0f;
}

// ------------------------------------------------------------ forward tile loop
// One visibility row per lane: phase coefficients and occultation records.
struct FwdRow {
  double u, v, w, c0;
  VisMask mk;
  VisCls cl;
  bool act;
};

template <bool MASK>
__device__ __forceinline__ FwdRow load_fwd_row(const double* __restrict__ uvw, const double* __restrict__ xyza,
                                               const double* __restrict__ xyzb, int64_t K, int64_t k, bool act,
                                               double R_moon, double orbit_r, int track) {
  FwdRow r;
  const int64_t kk = act ? k : K - 1;
  r.act = act;
  r.u = uvw[kk];
  r.v = uvw[K + kk];
  r.w = uvw[2 * K + kk];
  r.c0 = track ? -r.w : 0.0;
  r.mk = VisMask{};
  r.cl = VisCls{};
  if (MASK)
    make_mask_records(xyza[kk], xyza[K + kk], xyza[2 * K + kk], xyzb[kk], xyzb[K + kk], xyzb[2 * K + kk], R_moon,
                      orbit_r, r.mk, r.cl);
  return r;
}

// Forward view of a sky: SoA directions and brightness in sorted order, so a
// group of consecutive pixels of one tile is one wide scalar load per field.
struct SkyView {
  const double* sl;  // [3][P_pad]
  int64_t P_pad;
  const float* sI;   // [P_pad]
  const double* I64; // [P_pad]
  const float4* cone;
};

__host__ inline SkyView sky_view(const viss_sky* s) { return SkyView{s->sl, s->P_pad, s->sI, s->I64, s->cone}; }

// One tile, mixed precision: sum_p I_p (cos, sin)(2 pi phase) over its 64
// pixels, software-pipelined four pixels ahead (one scalar-load wait covers
// four pixels).  CHECK: exact FP64 occultation test per pair.
template <bool CHECK>
__device__ __forceinline__ void tile_mixed(const double* __restrict__ L, const double* __restrict__ M,
                                           const double* __restrict__ N, const float* __restrict__ Iv,
                                           const FwdRow& r, double& Sr, double& Si) {
  constexpr int G = 4;
  const double u = r.u, v = r.v, w = r.w, c0 = r.c0;
  double cl[G], cm[G], cn[G];
  float ci[G];
#pragma unroll
  for (int q = 0; q < G; ++q) {
    cl[q] = L[q];
    cm[q] = M[q];
    cn[q] = N[q];
    ci[q] = Iv[q];
  }
  float ar0 = 0.f, ai0 = 0.f, ar1 = 0.f, ai1 = 0.f;
  for (int j = 0; j < kTile; j += G) {
    const int jn = (j + G) & (kTile - 1);  // the last group reloads the first: in bounds, unused
    double nl[G], nm[G], nn[G];
    float ni[G];
#pragma unroll
    for (int q = 0; q < G; ++q) {
      nl[q] = L[jn + q];
      nm[q] = M[jn + q];
      nn[q] = N[jn + q];
      ni[q] = Iv[jn + q];
    }
#pragma unroll
    for (int q = 0; q < G; ++q) {
      float I0 = ci[q];
      if (CHECK && !both_visible(cl[q], cm[q], cn[q], r.mk)) I0 = 0.f;
      float s0, cs;
      sincos_rev(rev_fp32(fma(u, cl[q], fma(v, cm[q], fma(w, cn[q], c0)))), s0, cs);
      if (q & 1) {
        ar1 = fmaf(I0, cs, ar1);
        ai1 = fmaf(I0, s0, ai1);
      } else {
        ar0 = fmaf(I0, cs, ar0);
        ai0 = fmaf(I0, s0, ai0);
      }
    }
#pragma unroll
    for (int q = 0; q < G; ++q) {
      cl[q] = nl[q];
      cm[q] = nm[q];
      cn[q] = nn[q];
      ci[q] = ni[q];
    }
  }
  Sr += (double)ar0 + (double)ar1;
  Si += (double)ai0 + (double)ai1;
}

// Accumulates over the tiles t = list[j] (t = j when list is NULL), j in
// [j0, j1), into (Sr, Si).  All lanes of a wavefront walk the same tiles.
template <bool MASK, bool F64>
__device__ __forceinline__ void forward_tiles(const SkyView& sv, const int32_t* __restrict__ list, int64_t j0,
                                              int64_t j1, const FwdRow& r, double& Sr, double& Si) {
  const double u = r.u, v = r.v, w = r.w, c0 = r.c0;
  for (int64_t jt = j0; jt < j1; ++jt) {
    const int64_t t = list ? (int64_t)list[jt] : jt;
    bool full = true;
    if (MASK) {
      bool any_v, all_v;
      classify(sv.cone[t], r.cl, any_v, all_v);
      any_v = any_v & r.act;
      all_v = all_v | !r.act;
      if (__ballot(any_v) == 0) continue;
      full = __ballot(!all_v) == 0;
    }
    const double* __restrict__ L = sv.sl + t * kTile;
    const double* __restrict__ M = sv.sl + sv.P_pad + t * kTile;
    const double* __restrict__ N = sv.sl + 2 * sv.P_pad + t * kTile;
    if (!F64) {
      const float* __restrict__ Iv = sv.sI + t * kTile;
      if (!MASK || full)
        tile_mixed<false>(L, M, N, Iv, r, Sr, Si);
      else
        tile_mixed<true>(L, M, N, Iv, r, Sr, Si);
    } else {
      const double* __restrict__ ip = sv.I64 + t * kTile;
      for (int j = 0; j < kTile; ++j) {
        double I0 = ip[j];
        if (!full && !both_visible(L[j], M[j], N[j], r.mk)) I0 = 0.0;
        const double f = __builtin_amdgcn_fract(fma(u, L[j], fma(v, M[j], fma(w, N[j], c0))));
        double s, c;
        sincospi(2.0 * f, &s, &c);
        Sr = fma(I0, c, Sr);
        Si = fma(I0, s, Si);
      }
    }
  }
}

// V_k += Sr - i Si  (V = sum I exp(-2 pi i ph)); with herm, V_{k+K} += conj.
__device__ __forceinline__ void forward_store(double* __restrict__ V, int64_t k, int64_t K, int herm, double Sr,
                                              double Si) {
  unsafeAtomicAdd(&V[2 * k], Sr);
  unsafeAtomicAdd(&V[2 * k + 1], -Si);
  if (herm) {
    unsafeAtomicAdd(&V[2 * (k + K)], Sr);
    unsafeAtomicAdd(&V[2 * (k + K) + 1], Si);
  }
}

}  // namespace viss
