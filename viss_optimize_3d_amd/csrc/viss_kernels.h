// Device-side building blocks shared by the forward / back-projection kernels
// (viss_nudft.hip) and the batched work-table forward (viss_batch.hip).
#pragma once

#include "viss_internal.h"

namespace viss {

// Wavefront ballot of a lane predicate.  The builtin takes the predicate as a
// bool, so a comparison result that already is a lane mask is used as it is
// (HIP's __ballot(int) re-materialises it with a select and a compare).
__device__ __forceinline__ uint64_t wave_ballot(bool p) { return __builtin_amdgcn_ballot_w64(p); }

// ------------------------------------------------------------ pair math
__device__ __forceinline__ float rev_fp32(double ph) {
  return (float)__builtin_amdgcn_fract(ph);  // [0,1) revolutions, exact in FP64
}

__device__ __forceinline__ void sincos_rev(float r, float& s, float& c) {
  s = __builtin_amdgcn_sinf(r);  // sin(2 pi r)
  c = __builtin_amdgcn_cosf(r);
}

__device__ __forceinline__ bool both_visible(double l, double m, double n, const VisMask& q) {
  const double da = occ_dot(l, m, n * q.za, q.xa, q.ya);
  const double db = occ_dot(l, m, n * q.zb, q.xb, q.yb);
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

// Forward view of a sky: the sorted pixel pack, FP64 brightness and tile cones.
struct SkyView {
  const PixPack* pix;
  const double* I64;
  const float4* cone;
};

__host__ inline SkyView sky_view(const viss_sky* s) { return SkyView{s->pix, s->I64, s->cone}; }

// Accumulates sum_p I_p (cos, sin)(2 pi phase) over the tiles t = list[j]
// (t = j when list is NULL), j in [j0, j1), into (Sr, Si).  All lanes of a
// wavefront walk the same tiles; tile data are wave-uniform.
template <bool MASK, bool F64>
__device__ __forceinline__ void forward_tiles(const SkyView& sv, const int32_t* __restrict__ list, int64_t j0,
                                              int64_t j1, const FwdRow& r, double& Sr, double& Si) {
  const PixPack* __restrict__ pix = sv.pix;
  const double* __restrict__ I64 = sv.I64;
  const float4* __restrict__ cone = sv.cone;
  const double u = r.u, v = r.v, w = r.w, c0 = r.c0;
  for (int64_t jt = j0; jt < j1; ++jt) {
    const int64_t t = list ? (int64_t)list[jt] : jt;
    bool full = true;
    if (MASK) {
      bool any_v, all_v;
      classify(cone[t], r.cl, any_v, all_v);
      any_v = any_v & r.act;
      all_v = all_v | !r.act;
      if (wave_ballot(any_v) == 0) continue;
      full = wave_ballot(!all_v) == 0;
    }
    const PixPack* __restrict__ tp = pix + t * kTile;
    if (!F64) {
      float ar0 = 0.f, ai0 = 0.f, ar1 = 0.f, ai1 = 0.f;
      if (full) {
#pragma unroll 4
        for (int j = 0; j < kTile; j += 2) {
          const PixPack p0 = tp[j], p1 = tp[j + 1];
          float s0, c0f, s1, c1f;
          sincos_rev(rev_fp32(fma(u, p0.l, fma(v, p0.m, fma(w, p0.n, c0)))), s0, c0f);
          sincos_rev(rev_fp32(fma(u, p1.l, fma(v, p1.m, fma(w, p1.n, c0)))), s1, c1f);
          ar0 = fmaf(p0.I, c0f, ar0);
          ai0 = fmaf(p0.I, s0, ai0);
          ar1 = fmaf(p1.I, c1f, ar1);
          ai1 = fmaf(p1.I, s1, ai1);
        }
      } else {
#pragma unroll 2
        for (int j = 0; j < kTile; j += 2) {
          const PixPack p0 = tp[j], p1 = tp[j + 1];
          const float I0 = both_visible(p0.l, p0.m, p0.n, r.mk) ? p0.I : 0.f;
          const float I1 = both_visible(p1.l, p1.m, p1.n, r.mk) ? p1.I : 0.f;
          float s0, c0f, s1, c1f;
          sincos_rev(rev_fp32(fma(u, p0.l, fma(v, p0.m, fma(w, p0.n, c0)))), s0, c0f);
          sincos_rev(rev_fp32(fma(u, p1.l, fma(v, p1.m, fma(w, p1.n, c0)))), s1, c1f);
          ar0 = fmaf(I0, c0f, ar0);
          ai0 = fmaf(I0, s0, ai0);
          ar1 = fmaf(I1, c1f, ar1);
          ai1 = fmaf(I1, s1, ai1);
        }
      }
      Sr += (double)ar0 + (double)ar1;
      Si += (double)ai0 + (double)ai1;
    } else {
      const double* __restrict__ ip = I64 + t * kTile;
      for (int j = 0; j < kTile; ++j) {
        const PixPack p0 = tp[j];
        double I0 = ip[j];
        if (!full && !both_visible(p0.l, p0.m, p0.n, r.mk)) I0 = 0.0;
        const double f = __builtin_amdgcn_fract(fma(u, p0.l, fma(v, p0.m, fma(w, p0.n, c0))));
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

// ------------------------------------------------------------ forward, antipodal pairs
// Mixed precision, skies whose pixels come in antipodal pairs (s_b = -s_a,
// viss_sky's pair layout).  With theta = 2 pi b.s_a,
//   I_a e^{-i theta} + I_b e^{+i theta} = (I_a + I_b) cos theta - i (I_a - I_b) sin theta,
// so one FP64 phase and one sin/cos serve both pixels; the occultation test of
// b reuses the dot products of a (s_b.r = -s_a.r).  The phase-tracking term c0
// breaks the symmetry, so it is applied once per row at the end:
// V = e^{-2 pi i c0} sum_p I_p e^{-2 pi i b.s_p}.
// State of one side of a pair tile for every lane of the wavefront.
enum SideMode { kSideNone = 0, kSideFull = 1, kSideTest = 2 };

// Sums over one 64-entry pair tile, in the (Sr, Si) convention of
// forward_store: Sr = sum I cos, Si = sum I sin of each pixel's own phase.
// TA / TB: side a / b seen by no lane, wholly by every lane, or tested per
// pixel (the exact FP64 occultation test; b reuses a's dot products negated).
template <int TA, int TB>
__device__ __forceinline__ void pair_tile_forward(const PairPack* __restrict__ tp, double u, double v, double w,
                                                  const VisMask& mk, float& ar, float& ai) {
#pragma unroll 4
  for (int j = 0; j < 64; ++j) {
    const PairPack q = tp[j];
    float sn, cs;
    sincos_rev(rev_fp32(fma(u, q.l, fma(v, q.m, w * q.n))), sn, cs);
    if (TA == kSideFull && TB == kSideFull) {
      ar = fmaf(q.A, cs, ar);
      ai = fmaf(q.D, sn, ai);
    } else {
      double da = 0.0, db = 0.0;
      if (TA == kSideTest || TB == kSideTest) {
        da = occ_dot(q.l, q.m, q.n * mk.za, mk.xa, mk.ya);
        db = occ_dot(q.l, q.m, q.n * mk.zb, mk.xb, mk.yb);
      }
      const float wa = TA == kSideNone ? 0.f
                       : TA == kSideFull ? q.Ia
                                         : (((da >= mk.ta) & (db >= mk.tb)) ? q.Ia : 0.f);
      const float wb = TB == kSideNone ? 0.f
                       : TB == kSideFull ? q.Ib
                                         : (((-da >= mk.ta) & (-db >= mk.tb)) ? q.Ib : 0.f);
      if (TA != kSideNone && TB != kSideNone) {
        ar = fmaf(wa + wb, cs, ar);
        ai = fmaf(wa - wb, sn, ai);
      } else if (TA != kSideNone) {
        ar = fmaf(wa, cs, ar);
        ai = fmaf(wa, sn, ai);
      } else {
        ar = fmaf(wb, cs, ar);
        ai = fmaf(-wb, sn, ai);
      }
    }
  }
}

// Tile classification of both sides: the a pixels lie in the cone cn, the b
// pixels in (-centre, same chord).  Inactive lanes count as seeing nothing and
// everything (they never force a test).  With d = centre.r_hat, a pixel s of
// the cone has s.r_hat within d -+ w, so side a is
//   seen somewhere   when d + w >= c - eps,   i.e. d - c >= -W,
//   seen everywhere  when d - w >= c + eps,   i.e. d - c >=  W,
// and side b (-d -+ w) seen somewhere when d + c <= W, everywhere when
// d + c <= -W, with W = w + eps.  d - c comes out of the dot product's FMA
// chain; d + c = (d - c) + 2c, one more rounding, far inside eps.
__device__ __forceinline__ void classify_pair(const float4 cn, const VisCls& c, bool act, bool& anyA, bool& allA,
                                              bool& anyB, bool& allB) {
  const float W = cn.w + kClsEps;
  const float ea = fmaf(cn.x, c.ax, fmaf(cn.y, c.ay, fmaf(cn.z, c.az, -c.ca)));
  const float eb = fmaf(cn.x, c.bx, fmaf(cn.y, c.by, fmaf(cn.z, c.bz, -c.cb)));
  const float sa = ea + 2.f * c.ca, sb = eb + 2.f * c.cb;
  anyA = act & (ea >= -W) & (eb >= -W);
  allA = !act | ((ea >= W) & (eb >= W));
  anyB = act & (sa <= W) & (sb <= W);
  allB = !act | ((sa <= -W) & (sb <= -W));
}

// classify_pair as lane masks (one row per lane; actm = the lanes holding a
// row): the rows that see side a / side b somewhere (any) and everywhere
// (all).  Each comparison is balloted on its own, so it stays the lane mask
// its compare instruction writes, and the masks combine in scalar code.
struct PairSideMasks {
  uint64_t anyA, allA, anyB, allB;
};

__device__ __forceinline__ PairSideMasks classify_pair_masks(const float4 cn, const VisCls& c, uint64_t actm) {
  const float W = cn.w + kClsEps;
  const float ea = fmaf(cn.x, c.ax, fmaf(cn.y, c.ay, fmaf(cn.z, c.az, -c.ca)));
  const float eb = fmaf(cn.x, c.bx, fmaf(cn.y, c.by, fmaf(cn.z, c.bz, -c.cb)));
  const float sa = ea + 2.f * c.ca, sb = eb + 2.f * c.cb;
  PairSideMasks m;
  m.anyA = wave_ballot(ea >= -W) & wave_ballot(eb >= -W) & actm;
  m.allA = wave_ballot(ea >= W) & wave_ballot(eb >= W) & actm;
  m.anyB = wave_ballot(sa <= W) & wave_ballot(sb <= W) & actm;
  m.allB = wave_ballot(sa <= -W) & wave_ballot(sb <= -W) & actm;
  return m;
}

// The same reduced over the wavefront: side a / b seen by some row
// somewhere, by every row everywhere.
struct PairSides {
  bool seeA, wholeA, seeB, wholeB;
};

__device__ __forceinline__ PairSides classify_pair_wave(const float4 cn, const VisCls& c, uint64_t actm) {
  const PairSideMasks m = classify_pair_masks(cn, c, actm);
  return PairSides{m.anyA != 0, m.allA == actm, m.anyB != 0, m.allB == actm};
}

// Accumulates the pair tiles t = list[j] (t = j when list is NULL), j in
// [j0, j1), into (Sr, Si) without the phase-tracking term; all lanes of a
// wavefront walk the same tiles.
template <bool MASK>
__device__ __forceinline__ void forward_pair_tiles(const PairPack* __restrict__ pp, const float4* __restrict__ pcone,
                                                   const int32_t* __restrict__ list, int64_t j0, int64_t j1,
                                                   const FwdRow& r, double& Sr, double& Si) {
  const uint64_t actm = wave_ballot(r.act);
  for (int64_t jt = j0; jt < j1; ++jt) {
    const int64_t t = list ? (int64_t)list[jt] : jt;
    int ma = kSideFull, mb = kSideFull;
    if (MASK) {
      const PairSides ps = classify_pair_wave(pcone[t], r.cl, actm);
      if (!ps.seeA && !ps.seeB) continue;
      ma = !ps.seeA ? kSideNone : (ps.wholeA ? kSideFull : kSideTest);
      mb = !ps.seeB ? kSideNone : (ps.wholeB ? kSideFull : kSideTest);
    }
    const PairPack* __restrict__ tp = pp + t * 64;
    float ar = 0.f, ai = 0.f;
    switch (ma * 3 + mb) {  // wave-uniform
      case kSideFull * 3 + kSideFull: pair_tile_forward<kSideFull, kSideFull>(tp, r.u, r.v, r.w, r.mk, ar, ai); break;
      case kSideFull * 3 + kSideNone: pair_tile_forward<kSideFull, kSideNone>(tp, r.u, r.v, r.w, r.mk, ar, ai); break;
      case kSideNone * 3 + kSideFull: pair_tile_forward<kSideNone, kSideFull>(tp, r.u, r.v, r.w, r.mk, ar, ai); break;
      case kSideFull * 3 + kSideTest: pair_tile_forward<kSideFull, kSideTest>(tp, r.u, r.v, r.w, r.mk, ar, ai); break;
      case kSideTest * 3 + kSideFull: pair_tile_forward<kSideTest, kSideFull>(tp, r.u, r.v, r.w, r.mk, ar, ai); break;
      case kSideTest * 3 + kSideNone: pair_tile_forward<kSideTest, kSideNone>(tp, r.u, r.v, r.w, r.mk, ar, ai); break;
      case kSideNone * 3 + kSideTest: pair_tile_forward<kSideNone, kSideTest>(tp, r.u, r.v, r.w, r.mk, ar, ai); break;
      default: pair_tile_forward<kSideTest, kSideTest>(tp, r.u, r.v, r.w, r.mk, ar, ai); break;
    }
    Sr += (double)ar;
    Si += (double)ai;
  }
}

// ------------------------------------------------------------ forward, equatorial-mirror quads
// Quad tile at height z (viss_sky's quad layout): A = u x + v y per quad and
// C = w z per tile.  Ring z holds a = (x, y, z), b = (-x, -y, z) with phases
// A + C, -A + C; ring -z holds c = (x, y, -z), d = (-x, -y, -z) with A - C,
// -A - C.  Per ring, R = sum (wa + wb) cos A and J = sum (wa - wb) sin A, and
// the tile adds e^{-i C}(Rz - i Jz) + e^{+i C}(Rm - i Jm).
// Mode of one ring for every lane of the wavefront: nothing seen, both
// pixels seen whole, only the first (a or c), only the second (b or d), the
// first tested per pixel and the second seen whole (TF), the reverse (FT), or
// both tested.  The occultation cap is larger than a hemisphere (its cut is
// negative), so the antipode of a pixel near the cap's edge is seen: a ring
// that needs a test almost always has one side tested and the other whole.
enum RingMode { kRingNone = 0, kRingFF = 1, kRingFN = 2, kRingNF = 3, kRingTF = 4, kRingFT = 5, kRingTest = 6 };
constexpr int kRingModes = 7;

// Occultation tests of a quad in the evaluation order of the per-pixel test
// (occ_dot).  With E = occ_dot(x, y, z zs) and F = occ_dot(x, y, -z zs) for
// satellite s, the pixels' dot products are a = (x, y, z): E, b = (-x, -y, z):
// -F, c = (x, y, -z): F, d = (-x, -y, -z): -E (negation is exact), so each
// pixel's decision is the one the per-pixel test takes for those coordinates.
// zz = (z za, z zb) of the tile's ring for the lane's row.
struct QuadOcc {
  double zza, zzb;
};

// Visibility of a quad's pixels a, b, c, d for one row; only what the ring
// modes MZ (ring z: a, b) and MM (ring -z: c, d) test is evaluated.
template <int MZ, int MM>
__device__ __forceinline__ void quad_visible(double x, double y, const VisMask& mk, const QuadOcc& zz, bool& va,
                                             bool& vb, bool& vc, bool& vd) {
  constexpr bool needE = MZ == kRingTF || MZ == kRingTest || MM == kRingFT || MM == kRingTest;
  constexpr bool needF = MZ == kRingFT || MZ == kRingTest || MM == kRingTF || MM == kRingTest;
  va = vb = vc = vd = false;
  if (needE) {
    const double ea = occ_dot(x, y, zz.zza, mk.xa, mk.ya), eb = occ_dot(x, y, zz.zzb, mk.xb, mk.yb);
    va = (ea >= mk.ta) & (eb >= mk.tb);
    vd = (-ea >= mk.ta) & (-eb >= mk.tb);
  }
  if (needF) {
    const double fa = occ_dot(x, y, -zz.zza, mk.xa, mk.ya), fb = occ_dot(x, y, -zz.zzb, mk.xb, mk.yb);
    vc = (fa >= mk.ta) & (fb >= mk.tb);
    vb = (-fa >= mk.ta) & (-fb >= mk.tb);
  }
}

template <int MR>
__device__ __forceinline__ void ring_step(float w1, float w2, float sum, float dif, bool v1, bool v2, float cs,
                                          float sn, float& R, float& J) {
  if (MR == kRingFF) {
    R = fmaf(sum, cs, R);
    J = fmaf(dif, sn, J);
  } else if (MR == kRingFN) {
    R = fmaf(w1, cs, R);
    J = fmaf(w1, sn, J);
  } else if (MR == kRingNF) {
    R = fmaf(w2, cs, R);
    J = fmaf(-w2, sn, J);
  } else if (MR == kRingTF) {
    const float a = v1 ? w1 : 0.f;
    R = fmaf(a + w2, cs, R);
    J = fmaf(a - w2, sn, J);
  } else if (MR == kRingFT) {
    const float b = v2 ? w2 : 0.f;
    R = fmaf(w1 + b, cs, R);
    J = fmaf(w1 - b, sn, J);
  } else if (MR == kRingTest) {
    const float a = v1 ? w1 : 0.f, b = v2 ? w2 : 0.f;
    R = fmaf(a + b, cs, R);
    J = fmaf(a - b, sn, J);
  }
}

template <int MZ, int MM>
__device__ __forceinline__ void quad_one(const QuadPack& q, double A, const VisMask& mk, const QuadOcc& zz, float& Rz,
                                         float& Jz, float& Rm, float& Jm) {
  float sn, cs;
  sincos_rev(rev_fp32(A), sn, cs);
  bool va, vb, vc, vd;
  quad_visible<MZ, MM>(q.x, q.y, mk, zz, va, vb, vc, vd);
  ring_step<MZ>(q.Ia, q.Ib, q.Az, q.Dz, va, vb, cs, sn, Rz, Jz);
  ring_step<MM>(q.Ic, q.Id, q.Am, q.Dm, vc, vd, cs, sn, Rm, Jm);
}

// The 16 quads of a tile, each quad's record loaded (scalar loads) one quad
// ahead of its use, so that the loads of quad j + 1 are in flight while quad
// j computes: 83.8 -> 82.5 ms on the precision workload against four quads
// loaded and then computed at a time (two ahead 83.2 ms, four ahead 85.2 ms
// with 104 SGPRs and one wavefront per SIMD fewer).
// REC: the tile's a pixels sit at equal azimuth steps d on one circle (a
// HEALPix ring), so A_j = u x_j + v y_j is a sinusoid in j and
// A_{j+1} = c2 A_j - A_{j-1} with c2 = 2 cos d: one FP64 FMA a quad instead
// of a multiply and an FMA, and no (x, y) loads where no pixel is tested.
// The sky takes the recurrence only when every tile's (x, y), run through it
// in FP64, stay within 2.5e-13 of the records (k_quad_rec): a phase error
// under 2e-9 cycles at |b| = 3000 wavelengths, against the FP32 sin/cos's
// 1.25e-7.  Padding quads (zero weights) past the ring's end
// continue the recurrence: bounded for |c2| <= 2, and weighted by zero.
// z: the tile's ring height; the occultation terms z za, z zb are formed
// only for modes that test pixels.
template <int MZ, int MM, bool REC>
__device__ __forceinline__ void quad_tile_forward(const QuadPack* __restrict__ tq, double c2, double u, double v,
                                                  const VisMask& mk, double z, float& Rz, float& Jz, float& Rm,
                                                  float& Jm) {
  constexpr bool tested = MZ == kRingTF || MZ == kRingFT || MZ == kRingTest || MM == kRingTF || MM == kRingFT ||
                          MM == kRingTest;
  QuadOcc zz{0.0, 0.0};
  if (tested) {
    zz.zza = z * mk.za;
    zz.zzb = z * mk.zb;
  }
  QuadPack cur = tq[0];
  double a0 = 0.0, a1 = 0.0;
#pragma unroll
  for (int j = 0; j < kQuadTile; ++j) {
    QuadPack nxt = cur;
    if (j + 1 < kQuadTile) nxt = tq[j + 1];
    const double A = (!REC || j < 2) ? fma(u, cur.x, v * cur.y) : fma(c2, a1, -a0);
    if (REC) {
      a0 = a1;
      a1 = A;
    }
    quad_one<MZ, MM>(cur, A, mk, zz, Rz, Jz, Rm, Jm);
    cur = nxt;
  }
}

__device__ __forceinline__ int ring_mode(bool see1, bool whole1, bool see2, bool whole2) {
  const bool t1 = see1 & !whole1, t2 = see2 & !whole2;
  return (!see1 & !see2)      ? kRingNone
         : (t1 & !t2 & whole2) ? kRingTF
         : (t2 & !t1 & whole1) ? kRingFT
         : (t1 | t2)           ? kRingTest
         : (see1 & see2)       ? kRingFF
         : see1                ? kRingFN
                               : kRingNF;
}

// ring_mode of wave-uniform flags as a table: entry
// see1 | whole1 << 1 | see2 << 2 | whole2 << 3 holds the mode in 3 bits, so
// the mode costs a few scalar instructions and no branches (the chain of
// selects ring_mode compiles to in scalar code cost 1.1 ms of the forward's
// 82.7; per lane, in vector registers, the selects stay cheaper).
constexpr int ring_mode_of(int i) {
  const bool see1 = i & 1, whole1 = i & 2, see2 = i & 4, whole2 = i & 8;
  const bool t1 = see1 && !whole1, t2 = see2 && !whole2;
  return (!see1 && !see2)      ? kRingNone
         : (t1 && !t2 && whole2) ? kRingTF
         : (t2 && !t1 && whole1) ? kRingFT
         : (t1 || t2)            ? kRingTest
         : (see1 && see2)        ? kRingFF
         : see1                  ? kRingFN
                                 : kRingNF;
}
constexpr uint64_t ring_mode_table() {
  uint64_t t = 0;
  for (int i = 0; i < 16; ++i) t |= (uint64_t)ring_mode_of(i) << (3 * i);
  return t;
}
constexpr uint64_t kRingModeTable = ring_mode_table();

__device__ __forceinline__ int ring_mode_wave(bool see1, bool whole1, bool see2, bool whole2) {
  const unsigned i = (unsigned)see1 | ((unsigned)whole1 << 1) | ((unsigned)see2 << 2) | ((unsigned)whole2 << 3);
  return (int)((kRingModeTable >> (3 * i)) & 7);
}

// ring_mode of every lane as lane masks: m[i] = the rows whose ring takes
// mode i (any / all as from classify_pair_masks, all within any).
__device__ __forceinline__ void ring_mode_masks(uint64_t any1, uint64_t all1, uint64_t any2, uint64_t all2,
                                                uint64_t actm, uint64_t m[kRingModes]) {
  const uint64_t t1 = any1 & ~all1, t2 = any2 & ~all2, untested = ~(t1 | t2);
  m[kRingNone] = actm & ~any1 & ~any2;
  m[kRingTF] = t1 & ~t2 & all2;
  m[kRingFT] = t2 & ~t1 & all1;
  m[kRingTest] = (t1 | t2) & ~m[kRingTF] & ~m[kRingFT];
  m[kRingFF] = any1 & any2 & untested;
  m[kRingFN] = any1 & ~any2 & untested;
  m[kRingNF] = ~any1 & any2 & untested;
}

// Bounds of a wavefront's rows for classifying tiles once for all of them:
// the mean unit direction of each satellite over the active lanes, the
// largest distance of a lane's direction from it, and the range of the unit
// cuts.  For a unit pixel s, |s.a_k - s.a| <= |a_k - a| <= da, so a tile side
// the bounds show unseen (whole) is unseen (whole) for every lane as well.
// The margin on da covers the FP32 rounding of the bounds against that of the
// per-lane tests of classify_pair_masks.
struct WaveCls {
  float ax, ay, az, da, camin, camax;
  float bx, by, bz, db, cbmin, cbmax;
};

__device__ __forceinline__ WaveCls wave_cls(const VisCls& c, bool act) {
  // a row with a non-finite direction or cut (a NaN or infinite satellite
  // position) fails the per-lane tests; bounds built with it would be NaN and
  // settle every tile as unseen for all rows, so such a wavefront gets bounds
  // that settle nothing (every tile takes the per-lane classification)
  const bool fin = __builtin_isfinite(c.ax) & __builtin_isfinite(c.ay) & __builtin_isfinite(c.az) &
                   __builtin_isfinite(c.ca) & __builtin_isfinite(c.bx) & __builtin_isfinite(c.by) &
                   __builtin_isfinite(c.bz) & __builtin_isfinite(c.cb);
  if (wave_ballot(act & !fin) != 0)  // wave-uniform
    return WaveCls{0.f, 0.f, 0.f, __builtin_inff(), 0.f, 0.f, 0.f, 0.f, 0.f, __builtin_inff(), 0.f, 0.f};
  float v[6] = {act ? c.ax : 0.f, act ? c.ay : 0.f, act ? c.az : 0.f,
                act ? c.bx : 0.f, act ? c.by : 0.f, act ? c.bz : 0.f};
#pragma unroll
  for (int o = 32; o > 0; o >>= 1)
#pragma unroll
    for (int i = 0; i < 6; ++i) v[i] += __shfl_xor(v[i], o, 64);
  const float na = rsqrtf(fmaxf(v[0] * v[0] + v[1] * v[1] + v[2] * v[2], 1e-30f));
  const float nb = rsqrtf(fmaxf(v[3] * v[3] + v[4] * v[4] + v[5] * v[5], 1e-30f));
  WaveCls w;
  w.ax = v[0] * na; w.ay = v[1] * na; w.az = v[2] * na;
  w.bx = v[3] * nb; w.by = v[4] * nb; w.bz = v[5] * nb;
  const float ex = c.ax - w.ax, ey = c.ay - w.ay, ez = c.az - w.az;
  const float fx = c.bx - w.bx, fy = c.by - w.by, fz = c.bz - w.bz;
  // maxima over the active lanes (cuts lie in [-2, 1]: -3 is below any of them and of their negations)
  float m[6] = {act ? sqrtf(ex * ex + ey * ey + ez * ez) : 0.f, act ? sqrtf(fx * fx + fy * fy + fz * fz) : 0.f,
                act ? c.ca : -3.f, act ? -c.ca : -3.f, act ? c.cb : -3.f, act ? -c.cb : -3.f};
#pragma unroll
  for (int o = 32; o > 0; o >>= 1)
#pragma unroll
    for (int i = 0; i < 6; ++i) m[i] = fmaxf(m[i], __shfl_xor(m[i], o, 64));
  w.da = m[0] + 4e-6f;
  w.db = m[1] + 4e-6f;
  w.camax = m[2]; w.camin = -m[3];
  w.cbmax = m[4]; w.cbmin = -m[5];
  // every lane holds the same values after the butterflies: keep them in scalar registers
  const auto u = [](float x) { return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(x))); };
  return WaveCls{u(w.ax), u(w.ay), u(w.az), u(w.da), u(w.camin), u(w.camax),
                 u(w.bx), u(w.by), u(w.bz), u(w.db), u(w.cbmin), u(w.cbmax)};
}

// The ring modes forward_quad_tiles takes for a quad tile (cones c0: a and d
// pixels, c1: b and c), from the wavefront bounds, as (mz << 3) | mm; or
// kTileExact when the bounds leave a side between unseen and whole.  Where
// every side is unseen or whole by the bounds, it is so for every lane, and
// the per-lane masks give the same modes.
constexpr int kTileExact = 64;

__device__ __forceinline__ int quad_tile_code(const float4 c0, const float4 c1, const WaveCls& b) {
  bool see[2][2], whole[2][2];  // [cone][side A / side B]
  const float4 cs[2] = {c0, c1};
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const float4 cn = cs[h];
    const float W = cn.w + kClsEps;
    const float pa = fmaf(cn.x, b.ax, fmaf(cn.y, b.ay, cn.z * b.az));
    const float pb = fmaf(cn.x, b.bx, fmaf(cn.y, b.by, cn.z * b.bz));
    see[h][0] = (pa + b.da - b.camin >= -W) & (pb + b.db - b.cbmin >= -W);
    whole[h][0] = (pa - b.da - b.camax >= W) & (pb - b.db - b.cbmax >= W);
    see[h][1] = (pa - b.da + b.camin <= W) & (pb - b.db + b.cbmin <= W);
    whole[h][1] = (pa + b.da + b.camax <= -W) & (pb + b.db + b.cbmax <= -W);
  }
  const int mz = ring_mode(see[0][0], whole[0][0], see[1][0], whole[1][0]);
  const int mm = ring_mode(see[1][1], whole[1][1], see[0][1], whole[0][1]);
  return (mz >= kRingTF || mm >= kRingTF) ? kTileExact : (mz << 3) | mm;
}

// Ring modes of quad tile t from the per-lane masks (ring z: a, b; ring -z:
// c, d), or those a code of quad_tile_code carries.
__device__ __forceinline__ void quad_tile_modes(const float4* __restrict__ qcone, int64_t t, int code, const FwdRow& r,
                                                uint64_t actm, int& mz, int& mm) {
  if (code == kTileExact) {
    const PairSides ad = classify_pair_wave(qcone[2 * t], r.cl, actm);      // a and its antipode d
    const PairSides bc = classify_pair_wave(qcone[2 * t + 1], r.cl, actm);  // b and its antipode c
    mz = ring_mode_wave(ad.seeA, ad.wholeA, bc.seeA, bc.wholeA);
    mm = ring_mode_wave(bc.seeB, bc.wholeB, ad.seeB, ad.wholeB);
  } else {
    mz = code >> 3;
    mm = code & 7;
  }
}

// A run of up to kRunTiles evaluated quad tiles on one ring pair (the same
// z): each tile's ring sums (16 quads, FP32) join the run's sums Rz, Jz, Rm,
// Jm (FP32), and the ring factor (one sin/cos of w z and two complex
// products) is applied once when the run closes, straight into the FP64
// sums, instead of once a tile.  The tiles of a ring are consecutive (sorted
// by z, then azimuth); an equatorial ring of nside 256 holds 32.  On the
// precision workload (nside 256, 900,000 rows, mask) the forward took
// 77.4 -> 74.0 ms, and V's error against the all-FP64 kernels went from
// 1.95e-7 to 1.90e-7 (relative L2; max 2.30e-7 -> 1.92e-7).  Summing the
// quads of a whole run in one FP32 sum instead cost the same time but raised
// the error to 2.57e-7.
constexpr int kRunTiles = 8;
struct RingRun {
  uint64_t zbits = ~0ull;  // bits of the run's z; no FP64 z has all bits set but a NaN
  double z = 0.0;
  float Rz = 0.f, Jz = 0.f, Rm = 0.f, Jm = 0.f;
  int n = 0;
  __device__ __forceinline__ void close(const FwdRow& r, double& Sr, double& Si) {
    float sc, cc;
    sincos_rev(rev_fp32(r.w * z), sc, cc);
    Sr += (double)fmaf(cc, Rz + Rm, sc * (Jm - Jz));
    Si += (double)fmaf(sc, Rz - Rm, cc * (Jz + Jm));
  }
  __device__ __forceinline__ void open(double zt) {
    zbits = (uint64_t)__double_as_longlong(zt);
    z = zt;
    Rz = Jz = Rm = Jm = 0.f;
    n = 0;
  }
};

// Quad tile t in ring modes (mz, mm) (not both kRingNone), into the run's
// ring sums.
template <bool REC>
__device__ __forceinline__ void quad_tile_run(const QuadPack* __restrict__ qp, const double* __restrict__ qc2,
                                              int64_t t, int mz, int mm, const FwdRow& r, RingRun& rr) {
  const double c2 = REC ? qc2[t] : 0.0;
  const QuadPack* __restrict__ tq = qp + t * kQuadTile;
  float Rz = 0.f, Jz = 0.f, Rm = 0.f, Jm = 0.f;
#define VISS_QT(A, B)                                                                            \
  case A * kRingModes + B:                                                                       \
    quad_tile_forward<A, B, REC>(tq, c2, r.u, r.v, r.mk, rr.z, Rz, Jz, Rm, Jm);                 \
    break;
#define VISS_QT_ROW(A) VISS_QT(A, kRingNone) VISS_QT(A, kRingFF) VISS_QT(A, kRingFN) VISS_QT(A, kRingNF) \
                       VISS_QT(A, kRingTF) VISS_QT(A, kRingFT) VISS_QT(A, kRingTest)
  switch (mz * kRingModes + mm) {  // wave-uniform
    VISS_QT(kRingNone, kRingFF) VISS_QT(kRingNone, kRingFN) VISS_QT(kRingNone, kRingNF)
    VISS_QT(kRingNone, kRingTF) VISS_QT(kRingNone, kRingFT) VISS_QT(kRingNone, kRingTest)
    VISS_QT_ROW(kRingFF) VISS_QT_ROW(kRingFN) VISS_QT_ROW(kRingNF) VISS_QT_ROW(kRingTF) VISS_QT_ROW(kRingFT)
    VISS_QT_ROW(kRingTest)
    default: break;
  }
#undef VISS_QT_ROW
#undef VISS_QT
  rr.Rz += Rz;
  rr.Jz += Jz;
  rr.Rm += Rm;
  rr.Jm += Jm;
}

// Accumulates the quad tiles [t0, t1) into (Sr, Si) (the convention of
// forward_store, without the phase-tracking term).  With the mask the lanes
// first classify 64 tiles at a time, one each, against the wavefront's
// bounds (quad_tile_code); only the tiles the bounds leave open take the
// per-lane classification (about 34 vector instructions a tile, the bounds
// under one).  The modes, and so the results, are those of the per-lane
// classification alone.  Tiles join runs on their ring pair (RingRun).
template <bool MASK, bool REC>
__device__ __forceinline__ void forward_quad_tiles(const QuadPack* __restrict__ qp, const double* __restrict__ qz,
                                                   const double* __restrict__ qc2, const float4* __restrict__ qcone,
                                                   int64_t t0, int64_t t1, const FwdRow& r, double& Sr, double& Si) {
  const uint64_t actm = wave_ballot(r.act);
  if (MASK && actm == 0) return;  // no row: every tile would be unseen
  WaveCls wb{};
  if (MASK) wb = wave_cls(r.cl, r.act);
  RingRun rr;
  for (int64_t tb = t0; tb < t1; tb += 64) {
    const int64_t te = min(t1, tb + 64);
    int code = kTileExact;
    const int lane = lane_fresh();
    if (MASK && tb + lane < te) code = quad_tile_code(qcone[2 * (tb + lane)], qcone[2 * (tb + lane) + 1], wb);
    for (int64_t t = tb; t < te; ++t) {
      int mz = kRingFF, mm = kRingFF;
      if (MASK) {
        quad_tile_modes(qcone, t, __builtin_amdgcn_readlane(code, (int)(t - tb)), r, actm, mz, mm);
        if (mz == kRingNone && mm == kRingNone) continue;
      }
      const double z = qz[t];
      if ((uint64_t)__double_as_longlong(z) != rr.zbits || rr.n == kRunTiles) {  // wave-uniform
        if (rr.n > 0) rr.close(r, Sr, Si);
        rr.open(z);
      }
      quad_tile_run<REC>(qp, qc2, t, mz, mm, r, rr);
      ++rr.n;
    }
  }
  if (rr.n > 0) rr.close(r, Sr, Si);
}

// Applies the phase-tracking term, V e^{-2 pi i c0} = (Sr - i Si)(cos - i sin),
// and stores as forward_store.
__device__ __forceinline__ void forward_pair_store(double* __restrict__ V, int64_t k, int64_t K, int herm, double c0,
                                                   double Sr, double Si) {
  if (c0 != 0.0) {
    double sn, cs;
    sincospi(2.0 * (c0 - rint(c0)), &sn, &cs);
    const double nr = Sr * cs - Si * sn, ni = Sr * sn + Si * cs;
    Sr = nr;
    Si = ni;
  }
  forward_store(V, k, K, herm, Sr, Si);
}

}  // namespace viss
