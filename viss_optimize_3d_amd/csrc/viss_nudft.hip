// Direct-sum 3D visibility generator (forward) and back-projection / adjoint.
//
// Reference: healpix_moonback_viss (vissgen3d_noblockage.cu:76-102) runs one
// thread per visibility over every pixel with FP32 phase, acosf mask tests and
// a read-modify-write of Viss[i] in HBM per pixel; computeC
// (vissgen3d_noblockage.cu:167-182, masked vissgen3d_blockage.cu:167-192) runs
// one thread per pixel over every visibility with a dependent dcf[gs[i]]
// gather and a read-modify-write of C[idx] per pair.
//
// Here (gfx950, wave64):
//  * forward: each lane owns one visibility (u, v, w in FP64 VGPRs, FP32/FP64
//    accumulators in registers); every lane of a wavefront walks the same
//    pixel, so pixel data arrive wave-uniformly through scalar loads
//    (s_load) from the sorted sky and feed VALU ops as SGPR operands.  No
//    cross-lane traffic inside the pair loop.
//  * back-projection: each lane owns kBpPixPerLane pixels of one
//    kBpTile-pixel tile; visibilities are walked wave-uniformly (scalar
//    loads of 32-B records prepared by a prologue that folds the
//    density-compensation weight and the phase restore into y).
//  * occultation: per (lanes' visibilities, pixel tile) the tile's bounding
//    cone gives "none visible" (the wavefront skips the tile), "all
//    visible" (no per-pair test) or "partial" (exact FP64 test s.r >= t per
//    pair).  The classification is conservative, so results equal the
//    per-pair FP64 mask.
//  * precision (mixed): phase b.s + c0 formed with FP64 FMAs, reduced with
//    v_fract_f64 to [0,1) revolutions, converted to FP32 and fed to the
//    hardware v_sin_f32 / v_cos_f32 (which take revolutions); products
//    accumulate in FP32 over one tile (<= 64 terms) and the tile sums
//    combine in FP64 (blocked summation).  Partial results of different
//    pixel / visibility chunks meet through FP64 atomics in the caller's
//    accumulate-into buffer.
//  * VISS_PREC_FP64: same structure, FP64 accumulation; the forward uses
//    FP64 sincospi, the back-projection one FP64 cospi per (row, pixel) on
//    the polar form of y (x is real).
//  * sky symmetry (mixed precision): skies whose pixels pair with their
//    antipodes run the pair kernels (one phase and sin/cos per (s, -s));
//    skies whose pairs also mirror in z ring by ring (HEALPix) run the quad
//    kernels (one sin/cos per {(x, y, +-z), (-x, -y, +-z)}, the ring factor
//    exp(2 pi i w z) per tile in the forward and per 64 rows in the
//    back-projection).  The per-pixel kernels above serve every other sky.
#include <algorithm>
#include <cstdlib>
#include <mutex>

#include "viss_kernels.h"

namespace viss {

// ------------------------------------------------------------ forward
// Grid: x = ceil(K / 256) blocks of 256 visibilities (4 wavefronts), y = pixel
// chunks of `tiles_per_chunk` 64-pixel tiles.
template <bool MASK, bool F64>
__global__ __launch_bounds__(256) void k_forward(const SkyView sv, int64_t ntile, int64_t tiles_per_chunk, const double* __restrict__ uvw,
                                                 const double* __restrict__ xyza, const double* __restrict__ xyzb,
                                                 int64_t K, double R_moon, double orbit_r, int track, int herm,
                                                 double* __restrict__ V) {
  const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const FwdRow r = load_fwd_row<MASK>(uvw, xyza, xyzb, K, k, k < K, R_moon, orbit_r, track);
  const int64_t t0 = (int64_t)blockIdx.y * tiles_per_chunk;
  const int64_t t1 = min(ntile, t0 + tiles_per_chunk);
  double Sr = 0.0, Si = 0.0;
  forward_tiles<MASK, F64>(sv, nullptr, t0, t1, r, Sr, Si);
  if (r.act) forward_store(V, k, K, herm, Sr, Si);
}

// Forward over a paired sky (see forward_pair_tiles): grid as k_forward, with
// tiles of 64 antipodal pairs.
template <bool MASK>
__global__ __launch_bounds__(256) void k_forward_pairs(const PairPack* __restrict__ pp, const float4* __restrict__ pcone,
                                                       int64_t ntile, int64_t tiles_per_chunk,
                                                       const double* __restrict__ uvw, const double* __restrict__ xyza,
                                                       const double* __restrict__ xyzb, int64_t K, double R_moon,
                                                       double orbit_r, int track, int herm, double* __restrict__ V) {
  const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const FwdRow r = load_fwd_row<MASK>(uvw, xyza, xyzb, K, k, k < K, R_moon, orbit_r, track);
  const int64_t t0 = (int64_t)blockIdx.y * tiles_per_chunk;
  const int64_t t1 = min(ntile, t0 + tiles_per_chunk);
  double Sr = 0.0, Si = 0.0;
  forward_pair_tiles<MASK>(pp, pcone, nullptr, t0, t1, r, Sr, Si);
  if (r.act) forward_pair_store(V, k, K, herm, r.c0, Sr, Si);
}

// Forward over the equatorial-mirror quad layout (see forward_quad_tiles):
// grid.y chunks [0, nqchunk) walk quad tiles, the rest walk the pairs and
// singles left outside the quads.  Held to 8 wavefronts per SIMD (64 VGPRs):
// left alone the masked recurrent instance takes 66 and runs 7 (75.4 ms on
// the precision workload against 74.1 at 8).  Squeezed into 64 it spilled the
// row index and the lane index to scratch (20 B a lane, 0.35 GB of writes a
// launch); both are formed afresh where they are used (lane_fresh), and it
// fits in 63 without spilling.
template <bool MASK, bool REC>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8))) void k_forward_quads(const QuadPack* __restrict__ qp, const double* __restrict__ qz,
                                                       const double* __restrict__ qc2,
                                                       const float4* __restrict__ qcone, int64_t nqtile, int64_t qpc,
                                                       int64_t nqchunk, const PairPack* __restrict__ rp,
                                                       const float4* __restrict__ rcone, int64_t nrtile, int64_t rpc,
                                                       const double* __restrict__ uvw, const double* __restrict__ xyza,
                                                       const double* __restrict__ xyzb, int64_t K, double R_moon,
                                                       double orbit_r, int track, int herm, double* __restrict__ V) {
  const int64_t kbase = (int64_t)blockIdx.x * 256 + 64 * wave_in_block();  // wave-uniform
  const int64_t k = kbase + lane_fresh();
  const FwdRow r = load_fwd_row<MASK>(uvw, xyza, xyzb, K, k, k < K, R_moon, orbit_r, track);
  double Sr = 0.0, Si = 0.0;
  const int64_t c = blockIdx.y;
  if (c < nqchunk) {
    const int64_t t0 = c * qpc, t1 = min(nqtile, t0 + qpc);
    forward_quad_tiles<MASK, REC>(qp, qz, qc2, qcone, t0, t1, r, Sr, Si);
  } else {
    const int64_t t0 = (c - nqchunk) * rpc, t1 = min(nrtile, t0 + rpc);
    forward_pair_tiles<MASK>(rp, rcone, nullptr, t0, t1, r, Sr, Si);
  }
  if (r.act) forward_pair_store(V, kbase + lane_fresh(), K, herm, r.c0, Sr, Si);  // the row index, formed again
}

// ------------------------------------------------------------ back-projection
// Row chunks of the back-projection kernels (grid.y): by default chunk c holds
// rows [c kpc, (c + 1) kpc); a batch passes a table whose chunks never cross a
// group, with the group's image at x + g img_stride (0: one summed image).
struct RowChunks {
  const RowChunk* tab;
  int64_t img_stride;
};

// CHUNKED is a template parameter of the kernels, so that the per-row entry
// points' kernels carry none of the table's code or registers (four more
// SGPRs cost the quad back-projection one wavefront per SIMD).
template <bool CHUNKED>
__device__ __forceinline__ void chunk_rows(const RowChunks& rc, int64_t K, int64_t kpc, int64_t& k0, int64_t& k1,
                                           int64_t& xoff) {
  if (CHUNKED) {
    const RowChunk c = rc.tab[blockIdx.y];  // wave-uniform
    k0 = c.k0;
    k1 = c.k1;
    xoff = c.g * rc.img_stride;
  } else {
    k0 = (int64_t)blockIdx.y * kpc;
    k1 = min(K, k0 + kpc);
    xoff = 0;
  }
}

// Prologue: y'_k = wt_k * y_k [* exp(-2 pi i w_k)], packed with u, v, w.
__global__ void k_bp_prologue(const double* __restrict__ uvw, const double* __restrict__ xyza,
                              const double* __restrict__ xyzb, const double* __restrict__ y,
                              const double* __restrict__ wt, int64_t K, int mask, int track, int f64, double R_moon,
                              double orbit_r, VisHot* __restrict__ hot, VisHot64* __restrict__ hot64,
                              VisMask* __restrict__ vm, VisCls* __restrict__ vc) {
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < K; k += (int64_t)gridDim.x * blockDim.x) {
    const double u = uvw[k], v = uvw[K + k], w = uvw[2 * K + k];
    double yr = y[2 * k], yi = y[2 * k + 1];
    if (track) {
      double s, c;
      sincospi(-2.0 * (w - rint(w)), &s, &c);
      const double nr = yr * c - yi * s, ni = yr * s + yi * c;
      yr = nr;
      yi = ni;
    }
    if (wt) {
      yr *= wt[k];
      yi *= wt[k];
    }
    if (f64) {
      VisHot64 h;
      h.u = u; h.v = v; h.w = w; h.yr = yr; h.yi = yi;
      h.ph = atan2(yi, yr) * (0.5 / M_PI);
      h.ya = hypot(yr, yi);
      h.pad = 0.0;
      hot64[k] = h;
    } else {
      VisHot h;
      h.u = u; h.v = v; h.w = w; h.yr = (float)yr; h.yi = (float)yi;
      hot[k] = h;
    }
    if (mask) {
      VisMask m;
      VisCls c;
      make_mask_records(xyza[k], xyza[K + k], xyza[2 * K + k], xyzb[k], xyzb[K + k], xyzb[2 * K + k], R_moon,
                        orbit_r, m, c);
      vm[k] = m;
      vc[k] = c;
    }
  }
}

// Grid: x = ceil(nbtile / 4) blocks of 4 wavefronts (one kBpTile-pixel tile per
// wavefront), y = visibility chunks of `k_per_chunk` rows.
template <bool MASK, bool F64, bool CHUNKED>
__global__ __launch_bounds__(256) void k_backproject(const PixPack* __restrict__ pix, const float4* __restrict__ bcone,
                                                     int64_t nbtile, const VisHot* __restrict__ hot,
                                                     const VisHot64* __restrict__ hot64,
                                                     const VisMask* __restrict__ vm, const VisCls* __restrict__ vc,
                                                     int64_t K, int64_t k_per_chunk, double herm,
                                                     double* __restrict__ x, const RowChunks rc) {
  constexpr int PPL = kBpPixPerLane;
  const int lane = threadIdx.x & 63;
  const int64_t tile = (int64_t)blockIdx.x * 4 + wave_in_block();
  if (tile >= nbtile) return;  // wave-uniform
  double l[PPL], m[PPL], n[PPL];
  int32_t idx[PPL];
#pragma unroll
  for (int i = 0; i < PPL; ++i) {
    const PixPack q = pix[tile * kBpTile + i * 64 + lane];
    l[i] = q.l;
    m[i] = q.m;
    n[i] = q.n;
    idx[i] = q.idx;
  }
  float4 cn = make_float4(0.f, 0.f, 1.f, 2.f);
  if (MASK) cn = bcone[tile];
  int64_t k0, k1, xoff;
  chunk_rows<CHUNKED>(rc, K, k_per_chunk, k0, k1, xoff);
  double* __restrict__ xo = x + xoff;
  double acc[PPL];
#pragma unroll
  for (int i = 0; i < PPL; ++i) acc[i] = 0.0;
  for (int64_t kb = k0; kb < k1; kb += 64) {
    const int nk = (int)min((int64_t)64, k1 - kb);
    uint64_t anym, fullm;
    if (MASK) {
      bool any_v = false, all_v = false;
      if (lane < nk) classify(cn, vc[kb + lane], any_v, all_v);
      anym = wave_ballot(any_v);
      fullm = wave_ballot(all_v);
    } else {
      anym = nk == 64 ? ~0ull : ((1ull << nk) - 1ull);
      fullm = anym;
    }
    if (!F64) {
      float a[PPL];
#pragma unroll
      for (int i = 0; i < PPL; ++i) a[i] = 0.f;
      // Rows that see the whole tile, then rows that see part of it: two
      // loops rather than a branch per row, so the accumulators stay in
      // place (a per-row branch made the compiler copy them every row).
      uint64_t fm = anym & fullm;
      while (fm) {
        const int j = __builtin_ctzll(fm);
        fm &= fm - 1ull;
        const VisHot h = hot[kb + j];
#pragma unroll
        for (int i = 0; i < PPL; ++i) {
          float s, c;
          sincos_rev(rev_fp32(fma(h.u, l[i], fma(h.v, m[i], h.w * n[i]))), s, c);
          a[i] = fmaf(h.yr, c, fmaf(-h.yi, s, a[i]));
        }
      }
      uint64_t pm = MASK ? (anym & ~fullm) : 0ull;
      while (pm) {
        const int j = __builtin_ctzll(pm);
        pm &= pm - 1ull;
        const VisHot h = hot[kb + j];
        const VisMask q = vm[kb + j];
#pragma unroll
        for (int i = 0; i < PPL; ++i) {
          float s, c;
          sincos_rev(rev_fp32(fma(h.u, l[i], fma(h.v, m[i], h.w * n[i]))), s, c);
          const float t = fmaf(h.yr, c, -h.yi * s);
          a[i] += both_visible(l[i], m[i], n[i], q) ? t : 0.f;
        }
      }
#pragma unroll
      for (int i = 0; i < PPL; ++i) acc[i] += (double)a[i];
    } else {
      while (anym) {
        const int j = __builtin_ctzll(anym);
        anym &= anym - 1ull;
        const VisHot64 h = hot64[kb + j];
        const bool full = !MASK || ((fullm >> j) & 1ull);
        VisMask q{};
        if (!full) q = vm[kb + j];
#pragma unroll
        for (int i = 0; i < PPL; ++i) {
          // fract is exact, and adding ph after it rounds at 1e-16 rather
          // than at the ulp of a 10^4-cycle phase
          const double f = __builtin_amdgcn_fract(fma(h.u, l[i], fma(h.v, m[i], h.w * n[i]))) + h.ph;
          const double t = h.ya * cospi(2.0 * f);
          if (full || both_visible(l[i], m[i], n[i], q)) acc[i] += t;
        }
      }
    }
  }
#pragma unroll
  for (int i = 0; i < PPL; ++i)
    if (idx[i] >= 0) unsafeAtomicAdd(&xo[idx[i]], herm * acc[i]);
}

// ------------------------------------------------------------ back-projection, antipodal pairs
// Mixed precision, pair layout: each lane owns one antipodal pair (a, b) of a
// 64-pair tile.  Re(y e^{+i theta}) = yr cos - yi sin for a and
// Re(y e^{-i theta}) = yr cos + yi sin for b share the sin/cos.  Rows are
// classified 64 at a time (lane j tests row j) for both sides of the tile and
// run in four loops: both sides seen whole, only a, only b, and the rest with
// the exact per-pixel test.
template <bool MASK, bool CHUNKED>
__global__ __launch_bounds__(256) void k_backproject_pairs(const PairPack* __restrict__ pp,
                                                           const int2* __restrict__ pidx,
                                                           const float4* __restrict__ pcone, int64_t ntile,
                                                           const VisHot* __restrict__ hot,
                                                           const VisMask* __restrict__ vm,
                                                           const VisCls* __restrict__ vc, int64_t K,
                                                           int64_t k_per_chunk, double herm, double* __restrict__ x,
                                                           const RowChunks rc) {
  const int lane = threadIdx.x & 63;
  const int64_t tile = (int64_t)blockIdx.x * 4 + wave_in_block();
  if (tile >= ntile) return;  // wave-uniform
  const PairPack q = pp[tile * 64 + lane];
  const int2 ab = pidx[tile * 64 + lane];
  const double l = q.l, m = q.m, n = q.n;
  float4 cn = make_float4(0.f, 0.f, 1.f, 2.f);
  if (MASK) cn = pcone[tile];
  int64_t k0, k1, xoff;
  chunk_rows<CHUNKED>(rc, K, k_per_chunk, k0, k1, xoff);
  double* __restrict__ xo = x + xoff;
  double acc_a = 0.0, acc_b = 0.0;
  for (int64_t kb = k0; kb < k1; kb += 64) {
    const int nk = (int)min((int64_t)64, k1 - kb);
    uint64_t both, onlyA, onlyB, mixed;
    if (MASK) {
      bool anyA = false, allA = false, anyB = false, allB = false;
      if (lane < nk) classify_pair(cn, vc[kb + lane], true, anyA, allA, anyB, allB);
      const uint64_t sA = wave_ballot(anyA), fA = wave_ballot(allA), sB = wave_ballot(anyB), fB = wave_ballot(allB);
      both = fA & fB;
      onlyA = fA & ~sB;
      onlyB = fB & ~sA;
      mixed = (sA | sB) & ~(both | onlyA | onlyB);
    } else {
      both = nk == 64 ? ~0ull : ((1ull << nk) - 1ull);
      onlyA = onlyB = mixed = 0ull;
    }
    // Rows that see both pixels whole: t_a = P - Q, t_b = P + Q with
    // P = sum yr cos, Q = sum yi sin (two FMAs per row for both pixels).
    float sa = 0.f, sb = 0.f, P = 0.f, Q = 0.f;
    while (both) {
      const int j = __builtin_ctzll(both);
      both &= both - 1ull;
      const VisHot h = hot[kb + j];
      float sn, cs;
      sincos_rev(rev_fp32(fma(h.u, l, fma(h.v, m, h.w * n))), sn, cs);
      P = fmaf(h.yr, cs, P);
      Q = fmaf(h.yi, sn, Q);
    }
    sa += P - Q;
    sb += P + Q;
    while (onlyA) {
      const int j = __builtin_ctzll(onlyA);
      onlyA &= onlyA - 1ull;
      const VisHot h = hot[kb + j];
      float sn, cs;
      sincos_rev(rev_fp32(fma(h.u, l, fma(h.v, m, h.w * n))), sn, cs);
      sa = fmaf(h.yr, cs, fmaf(-h.yi, sn, sa));
    }
    while (onlyB) {
      const int j = __builtin_ctzll(onlyB);
      onlyB &= onlyB - 1ull;
      const VisHot h = hot[kb + j];
      float sn, cs;
      sincos_rev(rev_fp32(fma(h.u, l, fma(h.v, m, h.w * n))), sn, cs);
      sb = fmaf(h.yr, cs, fmaf(h.yi, sn, sb));
    }
    while (mixed) {
      const int j = __builtin_ctzll(mixed);
      mixed &= mixed - 1ull;
      const VisHot h = hot[kb + j];
      const VisMask mk = vm[kb + j];
      float sn, cs;
      sincos_rev(rev_fp32(fma(h.u, l, fma(h.v, m, h.w * n))), sn, cs);
      const double da = occ_dot(l, m, n * mk.za, mk.xa, mk.ya);
      const double db = occ_dot(l, m, n * mk.zb, mk.xb, mk.yb);
      const float ta = fmaf(h.yr, cs, -h.yi * sn), tb = fmaf(h.yr, cs, h.yi * sn);
      sa += ((da >= mk.ta) & (db >= mk.tb)) ? ta : 0.f;
      sb += ((-da >= mk.ta) & (-db >= mk.tb)) ? tb : 0.f;
    }
    acc_a += (double)sa;
    acc_b += (double)sb;
  }
  if (ab.x >= 0) unsafeAtomicAdd(&xo[ab.x], herm * acc_a);
  if (ab.y >= 0) unsafeAtomicAdd(&xo[ab.y], herm * acc_b);
}

// ------------------------------------------------------------ back-projection, equatorial-mirror quads
// With A = u x + v y, C = w z and y' = y e^{2 pi i C}, y'' = y e^{-2 pi i C}:
//   a = (x, y, z):    Re(y' e^{+2 pi i A}) = y'r cos A - y'i sin A
//   b = (-x, -y, z):  Re(y' e^{-2 pi i A}) = y'r cos A + y'i sin A
//   c = (x, y, -z):   the same with y''     (first form)
//   d = (-x, -y, -z): the same with y''     (second form)
// so one sin/cos of A serves the four pixels; rows whose ring shows both
// pixels whole use P = sum y'r cos, Q = sum y'i sin (a = P - Q, b = P + Q).
struct QuadBpAcc {
  float Pz, Qz, Az, Bz, Pm, Qm, Cm, Dm;
};

template <int MR>
__device__ __forceinline__ void ring_bp(float yr, float yi, float cs, float sn, bool v1, bool v2, float& P, float& Q,
                                        float& S1, float& S2) {
  if (MR == kRingFF) {
    P = fmaf(yr, cs, P);
    Q = fmaf(yi, sn, Q);
  } else if (MR == kRingFN) {
    S1 = fmaf(yr, cs, fmaf(-yi, sn, S1));
  } else if (MR == kRingNF) {
    S2 = fmaf(yr, cs, fmaf(yi, sn, S2));
  } else if (MR == kRingTF) {
    const float t1 = fmaf(yr, cs, -yi * sn);
    S1 += v1 ? t1 : 0.f;
    S2 = fmaf(yr, cs, fmaf(yi, sn, S2));
  } else if (MR == kRingFT) {
    const float t2 = fmaf(yr, cs, yi * sn);
    S1 = fmaf(yr, cs, fmaf(-yi, sn, S1));
    S2 += v2 ? t2 : 0.f;
  } else if (MR == kRingTest) {
    const float t1 = fmaf(yr, cs, -yi * sn), t2 = fmaf(yr, cs, yi * sn);
    S1 += v1 ? t1 : 0.f;
    S2 += v2 ? t2 : 0.f;
  }
}

template <int MZ, int MM>
__device__ __forceinline__ void quad_bp_rows(uint64_t msk, const VisHot* __restrict__ hot,
                                             const VisMask* __restrict__ vm, const float4* __restrict__ rot,
                                             int64_t kb, double qx, double qy, double zl, QuadBpAcc& acc) {
  while (msk) {
    const int j = __builtin_ctzll(msk);
    msk &= msk - 1ull;
    const VisHot h = hot[kb + j];
    const float4 r = rot[j];  // y' (x, y) and y'' (z, w) of this lane's ring
    float sn, cs;
    sincos_rev(rev_fp32(fma(h.u, qx, h.v * qy)), sn, cs);
    bool va = false, vb = false, vc = false, vd = false;
    if (MZ >= kRingTF || MM >= kRingTF) {  // unused tests are dropped by the compiler
      const VisMask mk = vm[kb + j];
      quad_visible<MZ, MM>(qx, qy, mk, QuadOcc{zl * mk.za, zl * mk.zb}, va, vb, vc, vd);
    }
    ring_bp<MZ>(r.x, r.y, cs, sn, va, vb, acc.Pz, acc.Qz, acc.Az, acc.Bz);
    ring_bp<MM>(r.z, r.w, cs, sn, vc, vd, acc.Pm, acc.Qm, acc.Cm, acc.Dm);
  }
}

// quad_bp_rows for a step whose 64 rows all take modes (MZ, MM), neither
// tested per pixel (the common case away from the occultation edges): the
// rows go in order, eight at a time, so their LDS reads and record loads take
// compile-time offsets, where the mask loop spends a vector add per row on the
// lane's LDS address (71.4 -> 70.4 ms with this for (FF, FF) alone, 67.3 ms
// for every such category, on the precision workload).
template <int MZ, int MM>
__device__ __forceinline__ void quad_bp_rows_all(const VisHot* __restrict__ hot, const float4* __restrict__ rot,
                                                 int64_t kb, double qx, double qy, QuadBpAcc& acc) {
  static_assert(MZ < kRingTF && MM < kRingTF, "rows without a per-pixel test only");
  const VisHot* __restrict__ hb = hot + kb;
  for (int j0 = 0; j0 < 64; j0 += 8) {
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const VisHot h = hb[j0 + i];
      const float4 r = rot[j0 + i];
      float sn, cs;
      sincos_rev(rev_fp32(fma(h.u, qx, h.v * qy)), sn, cs);
      ring_bp<MZ>(r.x, r.y, cs, sn, false, false, acc.Pz, acc.Qz, acc.Az, acc.Bz);
      ring_bp<MM>(r.z, r.w, cs, sn, false, false, acc.Pm, acc.Qm, acc.Cm, acc.Dm);
    }
  }
}

// The rows of one (ring z, ring -z) mode category.
template <int MZ, int MM>
__device__ __forceinline__ void quad_bp_cat(uint64_t msk, const VisHot* __restrict__ hot,
                                            const VisMask* __restrict__ vm, const float4* __restrict__ rot,
                                            int64_t kb, double qx, double qy, double zl, QuadBpAcc& acc) {
  if constexpr (MZ < kRingTF && MM < kRingTF) {
    if (msk == ~0ull) {
      quad_bp_rows_all<MZ, MM>(hot, rot, kb, qx, qy, acc);
      return;
    }
  }
  quad_bp_rows<MZ, MM>(msk, hot, vm, rot, kb, qx, qy, zl, acc);
}

// One wavefront per patch (4 tiles of 16 quads, one per ring of a band), one
// quad per lane (lanes 16g..16g+15 take tile patch[g]).  Per 64-row step,
// lane j loads row j, writes its y rotated by the four rings' factors to the
// wavefront's LDS slice and classifies the row against the patch's four
// sides; the rows then run in one loop per (ring z, ring -z) mode pair.
// Held to 7 wavefronts per SIMD (at most 72 VGPRs): the unrolled whole-step
// row loops need 66 VGPRs, and squeezed into 64 (8 wavefronts) they spill 12
// bytes per lane to scratch and ran 0.6 ms slower; 6 wavefronts 0.2-0.4 ms
// slower.
template <bool MASK, bool CHUNKED>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(7))) void k_backproject_quads(
    const QuadPack* __restrict__ qp, const int4* __restrict__ qi, const double* __restrict__ qz,
    const int4* __restrict__ patch, const float4* __restrict__ pcone, int64_t npatch, const VisHot* __restrict__ hot,
    const VisMask* __restrict__ vm, const VisCls* __restrict__ vc, int64_t K, int64_t k_per_chunk, double herm,
    double* __restrict__ x, const RowChunks rc) {
  __shared__ float4 rot_s[4][4][64];  // [wavefront][ring group][row]
  const int lane = threadIdx.x & 63;
  const int wv = wave_in_block();
  const int64_t p = (int64_t)blockIdx.x * 4 + wv;
  if (p >= npatch) return;  // wave-uniform; the kernel has no block barriers
  const int g = lane >> 4;
  const int4 pt = patch[p];
  const double z0 = pt.x >= 0 ? qz[pt.x] : 0.0, z1 = pt.y >= 0 ? qz[pt.y] : 0.0;
  const double z2 = pt.z >= 0 ? qz[pt.z] : 0.0, z3 = pt.w >= 0 ? qz[pt.w] : 0.0;
  const int32_t tile = g == 0 ? pt.x : (g == 1 ? pt.y : (g == 2 ? pt.z : pt.w));
  const double zl = g == 0 ? z0 : (g == 1 ? z1 : (g == 2 ? z2 : z3));
  double qx = 0.0, qy = 0.0;
  int4 idx = make_int4(-1, -1, -1, -1);
  if (tile >= 0) {
    const int64_t q = (int64_t)tile * kQuadTile + (lane & 15);
    qx = qp[q].x;
    qy = qp[q].y;
    idx = qi[q];
  }
  float4 cnA = make_float4(0.f, 0.f, 1.f, 2.f), cnB = cnA;
  if (MASK) {
    cnA = pcone[2 * p];
    cnB = pcone[2 * p + 1];
  }
  int64_t k0, k1, xoff;
  chunk_rows<CHUNKED>(rc, K, k_per_chunk, k0, k1, xoff);
  double* __restrict__ xo = x + xoff;
  double sa = 0.0, sb = 0.0, sc = 0.0, sd = 0.0;
  float4* __restrict__ rot_w = &rot_s[wv][0][0];
  QuadBpAcc acc{0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};  // FP32 sums of up to 4 steps (256 rows)
  int nstep = 0;
  const auto flush = [&]() {
    sa += (double)(acc.Pz - acc.Qz + acc.Az);
    sb += (double)(acc.Pz + acc.Qz + acc.Bz);
    sc += (double)(acc.Pm - acc.Qm + acc.Cm);
    sd += (double)(acc.Pm + acc.Qm + acc.Dm);
    acc = QuadBpAcc{0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    nstep = 0;
  };
  for (int64_t kb = k0; kb < k1; kb += 64) {
    const int nk = (int)min((int64_t)64, k1 - kb);
    const uint64_t actm = nk >= 64 ? ~0ull : ((1ull << nk) - 1ull);  // lanes holding a row
    const int64_t kr = kb + min(lane, nk - 1);  // lanes past the last row repeat it, outside actm
    const VisHot h = hot[kr];
    const double zg[4] = {z0, z1, z2, z3};
#pragma unroll
    for (int gg = 0; gg < 4; ++gg) {
      float s, c;
      sincos_rev(rev_fp32(h.w * zg[gg]), s, c);
      rot_w[gg * 64 + lane] = make_float4(fmaf(h.yr, c, -h.yi * s), fmaf(h.yr, s, h.yi * c),
                                          fmaf(h.yr, c, h.yi * s), fmaf(h.yi, c, -h.yr * s));
    }
    // rows by (ring z mode, ring -z mode): each category's rows are the AND of
    // a ring z mask and a ring -z mask; a ring z mode no row has skips its 7
    uint64_t bz[kRingModes], bm[kRingModes];
    if (MASK) {
      const VisCls cl = vc[kr];
      const PairSideMasks ad = classify_pair_masks(cnA, cl, actm);  // a, and its antipode d
      const PairSideMasks bc = classify_pair_masks(cnB, cl, actm);  // b, and its antipode c
      ring_mode_masks(ad.anyA, ad.allA, bc.anyA, bc.allA, actm, bz);
      ring_mode_masks(bc.anyB, bc.allB, ad.anyB, ad.allB, actm, bm);
    } else {
#pragma unroll
      for (int i = 0; i < kRingModes; ++i) bz[i] = bm[i] = 0ull;
      bz[kRingFF] = bm[kRingFF] = actm;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const float4* __restrict__ rot = rot_w + g * 64;
#define VISS_QB(A, B) quad_bp_cat<A, B>(bz[A] & bm[B], hot, vm, rot, kb, qx, qy, zl, acc);
#define VISS_QB_ROW(A)                                                                                   \
  if (bz[A]) {                                                                                           \
    VISS_QB(A, kRingNone) VISS_QB(A, kRingFF) VISS_QB(A, kRingFN) VISS_QB(A, kRingNF) VISS_QB(A, kRingTF) \
    VISS_QB(A, kRingFT) VISS_QB(A, kRingTest)                                                             \
  }
    if (MASK) {
      if (bz[kRingNone]) {
        VISS_QB(kRingNone, kRingFF) VISS_QB(kRingNone, kRingFN) VISS_QB(kRingNone, kRingNF)
        VISS_QB(kRingNone, kRingTF) VISS_QB(kRingNone, kRingFT) VISS_QB(kRingNone, kRingTest)
      }
      VISS_QB_ROW(kRingFF) VISS_QB_ROW(kRingFN) VISS_QB_ROW(kRingNF) VISS_QB_ROW(kRingTF) VISS_QB_ROW(kRingFT)
      VISS_QB_ROW(kRingTest)
    } else {
      VISS_QB(kRingFF, kRingFF)
    }
#undef VISS_QB_ROW
#undef VISS_QB
    if (++nstep == 4) flush();
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
  if (nstep > 0) flush();
  if (idx.x >= 0) unsafeAtomicAdd(&xo[idx.x], herm * sa);
  if (idx.y >= 0) unsafeAtomicAdd(&xo[idx.y], herm * sb);
  if (idx.z >= 0) unsafeAtomicAdd(&xo[idx.z], herm * sc);
  if (idx.w >= 0) unsafeAtomicAdd(&xo[idx.w], herm * sd);
}

// ------------------------------------------------------------ visible-pair count
__global__ __launch_bounds__(256) void k_count_visible(const PixPack* __restrict__ pix,
                                                       const float4* __restrict__ cone,
                                                       const int32_t* __restrict__ tile_valid, int64_t ntile,
                                                       int64_t tiles_per_chunk, const double* __restrict__ xyza,
                                                       const double* __restrict__ xyzb, int64_t K, double R_moon,
                                                       double orbit_r, unsigned long long* __restrict__ count) {
  const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const bool act = k < K;
  const int64_t kk = act ? k : K - 1;
  VisMask mk;
  VisCls cl;
  make_mask_records(xyza[kk], xyza[K + kk], xyza[2 * K + kk], xyzb[kk], xyzb[K + kk], xyzb[2 * K + kk], R_moon,
                    orbit_r, mk, cl);
  const int64_t t0 = (int64_t)blockIdx.y * tiles_per_chunk;
  const int64_t t1 = min(ntile, t0 + tiles_per_chunk);
  unsigned long long c = 0;
  for (int64_t t = t0; t < t1; ++t) {
    bool any_v, all_v;
    classify(cone[t], cl, any_v, all_v);
    if (!any_v) continue;
    if (all_v) {
      c += (unsigned long long)tile_valid[t];
      continue;
    }
    const PixPack* tp = pix + t * kTile;
    for (int j = 0; j < kTile; ++j) {
      const PixPack q = tp[j];
      c += (q.idx >= 0 && both_visible(q.l, q.m, q.n, mk)) ? 1ull : 0ull;
    }
  }
  if (act && c) atomicAdd(count, c);
}

// ------------------------------------------------------------ work count of the quad layout
// The work of the mixed per-row kernels over the quad layout, counted with
// their own occultation test (quad_visible, occ_dot): per row, the visible
// pixels, the quad tiles with a visible pixel, the runs of such tiles that
// share one ring factor (RingRun: up to kRunTiles consecutive tiles of one
// ring pair within a chunk of the forward's), the quads with a visible pixel
// (one phase and one sin/cos each), the rings of a quad with a visible pixel
// (one pair of multiply-adds each), and the pairs of the rest layout with a
// visible pixel (one phase and sin/cos each).  One row per lane; grid.y
// chunks of quad tiles (those of the forward, so the runs are its runs for a
// row on its own), then of rest tiles.
__global__ __launch_bounds__(256) void k_count_quad_work(
    const QuadPack* __restrict__ qp, const int4* __restrict__ qi, const double* __restrict__ qz, int64_t nqtile,
    int64_t qpc, int64_t nqchunk, const PairPack* __restrict__ rp, const int2* __restrict__ ri, int64_t nrtile,
    int64_t rpc, const double* __restrict__ xyza, const double* __restrict__ xyzb, int64_t K, double R_moon,
    double orbit_r, unsigned long long* __restrict__ counts) {
  const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const bool act = k < K;
  const int64_t kk = act ? k : K - 1;
  VisMask mk;
  VisCls cl;
  make_mask_records(xyza[kk], xyza[K + kk], xyza[2 * K + kk], xyzb[kk], xyzb[K + kk], xyzb[2 * K + kk], R_moon,
                    orbit_r, mk, cl);
  unsigned long long npix = 0, ntile = 0, nquad = 0, nring = 0, npair = 0, nrun = 0;
  const int64_t c = blockIdx.y;
  if (c < nqchunk) {
    const int64_t t1 = min(nqtile, (c + 1) * qpc);
    uint64_t zrun = ~0ull;
    int in_run = 0;
    for (int64_t t = c * qpc; t < t1; ++t) {
      const double z = qz[t];
      const QuadOcc zz{z * mk.za, z * mk.zb};
      unsigned long long tq = 0;
      for (int j = 0; j < kQuadTile; ++j) {
        const QuadPack q = qp[t * kQuadTile + j];
        const int4 id = qi[t * kQuadTile + j];
        bool va, vb, vc, vd;
        quad_visible<kRingTest, kRingTest>(q.x, q.y, mk, zz, va, vb, vc, vd);
        va &= id.x >= 0;
        vb &= id.y >= 0;
        vc &= id.z >= 0;
        vd &= id.w >= 0;
        npix += (unsigned)va + (unsigned)vb + (unsigned)vc + (unsigned)vd;
        nring += (unsigned)(va | vb) + (unsigned)(vc | vd);
        const bool any = va | vb | vc | vd;
        nquad += any;
        tq += any;
      }
      ntile += tq > 0;
      if (tq > 0) {  // as forward_quad_tiles forms its runs
        const uint64_t zb = (uint64_t)__double_as_longlong(z);
        if (zb != zrun || in_run == kRunTiles) {
          ++nrun;
          zrun = zb;
          in_run = 0;
        }
        ++in_run;
      }
    }
  } else {
    const int64_t r0 = (c - nqchunk) * rpc, r1 = min(nrtile, r0 + rpc);
    for (int64_t e = r0 * 64; e < r1 * 64; ++e) {
      const PairPack q = rp[e];
      const int2 id = ri[e];
      const double da = occ_dot(q.l, q.m, q.n * mk.za, mk.xa, mk.ya);
      const double db = occ_dot(q.l, q.m, q.n * mk.zb, mk.xb, mk.yb);
      const bool va = (id.x >= 0) & (da >= mk.ta) & (db >= mk.tb);
      const bool vb = (id.y >= 0) & (-da >= mk.ta) & (-db >= mk.tb);
      npix += (unsigned)va + (unsigned)vb;
      npair += va | vb;
    }
  }
  if (!act) npix = ntile = nquad = nring = npair = nrun = 0;
  unsigned long long v[6] = {npix, ntile, nquad, nring, npair, nrun};
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    for (int o = 32; o > 0; o >>= 1) v[i] += __shfl_xor(v[i], o, 64);
    if ((threadIdx.x & 63) == 0 && v[i]) atomicAdd(&counts[i], v[i]);
  }
}

// ------------------------------------------------------------ launch helpers
static int64_t cdiv(int64_t a, int64_t b) { return (a + b - 1) / b; }

constexpr int64_t kFwdTilesPerChunk = 256;
constexpr int64_t kBpRowsPerChunk = 4096;
constexpr int64_t kMinBlocks = 4096;  // 16 blocks of 4 wavefronts per CU

static int64_t env_int(const char* name, int64_t dflt) {
  const char* v = getenv(name);
  if (!v || !*v) return dflt;
  const long long x = atoll(v);
  return x > 0 ? (int64_t)x : dflt;
}

// Chunks of the per-row quad forward over bx blocks of rows: qpc quad tiles
// (64 pixels each) per chunk, 512 by default, which keeps the 32768 pixels of
// 256 pair tiles; rpc rest tiles (64 pairs) per chunk.  Halved while the grid
// would hold under kMinBlocks blocks; at most 32768 chunks of each kind.
static void quad_fwd_chunks(const viss_sky* sky, int64_t bx, int64_t& qpc, int64_t& rpc) {
  qpc = env_int("VISS_FWD_QTILES", 0);
  if (qpc <= 0) {
    qpc = 2 * kFwdTilesPerChunk;
    while (qpc > 32 && bx * (cdiv(sky->nqtile, qpc) + cdiv(sky->nrtile, qpc / 2)) < kMinBlocks) qpc /= 2;
  }
  qpc = std::max(qpc, cdiv(sky->nqtile, 32768));
  rpc = std::max<int64_t>(qpc / 2, cdiv(sky->nrtile, 32768));
}

static int validate_geom(const viss_mask_geometry* g, int flags) {
  if (flags & VISS_MASK) {
    VISS_REQUIRE(g != nullptr, "VISS_MASK needs a viss_mask_geometry");
    VISS_REQUIRE(g->R_moon >= 0.0, "R_moon must be >= 0");
    VISS_REQUIRE(g->orbit_radius <= 0.0 || g->orbit_radius > g->R_moon,
                 "orbit_radius must exceed R_moon when given (fixed-cut mode)");
  }
  return VISS_OK;
}

static int sky_forward_impl(viss_sky* sky, const double* I, const double* uvw, const double* xyza,
                            const double* xyzb, int64_t K, const viss_mask_geometry* geom, int flags, double* V,
                            hipStream_t s) {
  VISS_REQUIRE(K >= 0, "viss_forward: negative K");
  if (K == 0) return VISS_OK;
  VISS_REQUIRE(K <= ((int64_t)1 << 40), "viss_forward: K too large");
  VISS_REQUIRE(I && uvw && V, "viss_forward: I/uvw/V must not be NULL");
  const bool mask = flags & VISS_MASK;
  VISS_REQUIRE(!mask || (xyza && xyzb), "viss_forward: VISS_MASK needs xyz_a and xyz_b");
  int rc = validate_geom(geom, flags);
  if (rc != VISS_OK) return rc;
  const bool f64 = flags & VISS_PREC_FP64;
  rc = sky_set_brightness(sky, I, f64, s);
  if (rc != VISS_OK) return rc;
  const double R = mask ? geom->R_moon : 0.0, orb = mask ? geom->orbit_radius : 0.0;
  const int64_t bx = cdiv(K, 256);
  // 256 tiles x 64 pixels x 32 B = 512 KB per chunk (measured: 191 ms at the
  // precision workload against 242 ms with 4096-tile chunks); smaller chunks
  // when few visibility blocks would otherwise leave CUs idle.
  // mixed precision over a paired sky: 64-pair tiles, one sin/cos per pair
  const bool pairs = !f64 && sky->use_pairs;
  const int64_t ntile = pairs ? sky->nptile : sky->ntile;
  int64_t tpc = env_int("VISS_FWD_TILES", 0);
  if (tpc <= 0) {
    tpc = kFwdTilesPerChunk;
    while (tpc > 16 && bx * cdiv(ntile, tpc) < kMinBlocks) tpc /= 2;
  }
  tpc = std::max(tpc, cdiv(ntile, 65535));  // grid.y limit
  VISS_REQUIRE(bx <= INT32_MAX, "viss_forward: K too large for one launch");
  dim3 grid((unsigned)bx, (unsigned)cdiv(ntile, tpc));
  const int track = (flags & VISS_PHASE_TRACK) ? 1 : 0, herm = (flags & VISS_HERMITIAN) ? 1 : 0;
  if (!f64 && sky->use_quads) {
    int64_t qpc = 0, rpc = 0;
    quad_fwd_chunks(sky, bx, qpc, rpc);
    const int64_t nqc = cdiv(sky->nqtile, qpc), nrc = cdiv(sky->nrtile, rpc);
    dim3 qgrid((unsigned)bx, (unsigned)(nqc + nrc));
    // <MASK, REC>: the recurrent phase where the sky's tiles admit it
    auto* kern = mask ? (sky->qrec ? k_forward_quads<true, true> : k_forward_quads<true, false>)
                      : (sky->qrec ? k_forward_quads<false, true> : k_forward_quads<false, false>);
    kern<<<qgrid, 256, 0, s>>>(sky->qpack, sky->qz, sky->qc2, sky->qcone, sky->nqtile, qpc, nqc, sky->rpack,
                               sky->rcone, sky->nrtile, rpc, uvw, xyza, xyzb, K, R, orb, track, herm, V);
    VISS_LAUNCHED(k_forward_quads);
    return VISS_OK;
  }
  if (pairs) {
    if (mask)
      k_forward_pairs<true><<<grid, 256, 0, s>>>(sky->ppack, sky->pcone, ntile, tpc, uvw, xyza, xyzb, K, R, orb, track,
                                                 herm, V);
    else
      k_forward_pairs<false><<<grid, 256, 0, s>>>(sky->ppack, sky->pcone, ntile, tpc, uvw, xyza, xyzb, K, R, orb,
                                                  track, herm, V);
    VISS_LAUNCHED(k_forward_pairs);
    return VISS_OK;
  }
#define VISS_FWD(M, F)                                                                                          \
  k_forward<M, F><<<grid, 256, 0, s>>>(sky_view(sky), sky->ntile, tpc, uvw, xyza, xyzb, K, R, orb, \
                                       track, herm, V)
  if (mask) {
    if (f64) VISS_FWD(true, true); else VISS_FWD(true, false);
  } else {
    if (f64) VISS_FWD(false, true); else VISS_FWD(false, false);
  }
#undef VISS_FWD
  VISS_LAUNCHED(k_forward);
  return VISS_OK;
}

// Bytes of row records the back-projection packs for K rows (k_bp_prologue).
size_t bp_workspace_bytes(int64_t K, int flags) {
  const size_t hot = (flags & VISS_PREC_FP64) ? sizeof(VisHot64) : sizeof(VisHot);
  const size_t msk = (flags & VISS_MASK) ? sizeof(VisMask) + sizeof(VisCls) : 0;
  return (size_t)K * (hot + msk) + 256;
}

static int check_backproject_args(const double* uvw, const double* xyza, const double* xyzb, const double* y,
                                  int64_t K, const viss_mask_geometry* geom, int flags, double* x) {
  VISS_REQUIRE(K >= 0, "viss_backproject: negative K");
  if (K == 0) return VISS_OK;
  VISS_REQUIRE(K <= ((int64_t)1 << 40), "viss_backproject: K too large");
  VISS_REQUIRE(uvw && y && x, "viss_backproject: uvw/y/x must not be NULL");
  VISS_REQUIRE(!(flags & VISS_MASK) || (xyza && xyzb), "viss_backproject: VISS_MASK needs xyz_a and xyz_b");
  return validate_geom(geom, flags);
}

// The back-projection over a workspace of at least bp_workspace_bytes(K,
// flags) bytes at `ws` (256-B aligned): the prologue packs the row records
// there, then the tile kernels read them.  Arguments already checked.
// chunks / nchunks: the batch's row-chunk table (see RowChunk), or none.
int sky_backproject_ws(viss_sky* sky, const double* uvw, const double* xyza, const double* xyzb, const double* y,
                       const double* wt, int64_t K, const viss_mask_geometry* geom, int flags, double* x, char* ws,
                       hipStream_t s, const RowChunk* chunks, int64_t nchunks, int64_t img_stride) {
  if (K == 0) return VISS_OK;
  const bool mask = flags & VISS_MASK;
  const bool f64 = flags & VISS_PREC_FP64;
  const size_t hot_b = (size_t)K * (f64 ? sizeof(VisHot64) : sizeof(VisHot));
  char* base = ws;
  VisHot* hot = f64 ? nullptr : (VisHot*)base;
  VisHot64* hot64 = f64 ? (VisHot64*)base : nullptr;
  VisMask* vm = mask ? (VisMask*)(base + hot_b) : nullptr;
  VisCls* vc = mask ? (VisCls*)(base + hot_b + (size_t)K * sizeof(VisMask)) : nullptr;
  const double R = mask ? geom->R_moon : 0.0, orb = mask ? geom->orbit_radius : 0.0;
  {
    int64_t g = cdiv(K, 256);
    if (g > 65536) g = 65536;
    k_bp_prologue<<<(unsigned)g, 256, 0, s>>>(uvw, xyza, xyzb, y, wt, K, mask, (flags & VISS_PHASE_TRACK) ? 1 : 0,
                                              f64 ? 1 : 0, R, orb, hot, hot64, vm, vc);
    VISS_LAUNCHED(k_bp_prologue);
  }
  // mixed precision over a paired sky: one lane per antipodal pair
  const bool pairs = !f64 && sky->use_pairs;
  const int64_t bx = cdiv(pairs ? sky->nptile : sky->nbtile, 4);
  // 4096 rows x 64 B of records per chunk (measured: 187 ms against 210 ms
  // with 82k-row chunks)
  int64_t kpc = env_int("VISS_BP_ROWS", 0);
  if (kpc <= 0) {
    kpc = kBpRowsPerChunk;
    while (kpc > 256 && bx * cdiv(K, kpc) < kMinBlocks) kpc /= 2;
  }
  kpc = cdiv(kpc, 64) * 64;
  if (cdiv(K, kpc) > 65535) kpc = cdiv(cdiv(K, 65535), 64) * 64;
  const RowChunks rc{chunks, img_stride};
  VISS_REQUIRE(!chunks || (nchunks >= 1 && nchunks <= 65535), "backproject: %lld row chunks", (long long)nchunks);
  dim3 grid((unsigned)bx, (unsigned)(chunks ? nchunks : cdiv(K, kpc)));
  const double herm = (flags & VISS_HERMITIAN) ? 2.0 : 1.0;
#define VISS_BP(M, F)                                                                                        \
  do {                                                                                                       \
    if (chunks)                                                                                              \
      k_backproject<M, F, true><<<grid, 256, 0, s>>>(sky->pix, sky->bcone, sky->nbtile, hot, hot64, vm, vc, K, \
                                                     kpc, herm, x, rc);                                      \
    else                                                                                                     \
      k_backproject<M, F, false><<<grid, 256, 0, s>>>(sky->pix, sky->bcone, sky->nbtile, hot, hot64, vm, vc, K, \
                                                      kpc, herm, x, rc);                                     \
  } while (0)
#define VISS_BPQ(M, C)                                                                                           \
  k_backproject_quads<M, C><<<qgrid, 256, 0, s>>>(sky->qpack, sky->qidx, sky->qz, sky->qpatch, sky->qpcone,      \
                                                  sky->nqpatch, hot, vm, vc, K, kpc, herm, x, rc)
#define VISS_BPP(G, PACK, IDX, CONE, NT, KPC, M, C) \
  k_backproject_pairs<M, C><<<G, 256, 0, s>>>(PACK, IDX, CONE, NT, hot, vm, vc, K, KPC, herm, x, rc)
  if (pairs && sky->use_quads) {
    // quad patches, then the pairs and singles outside the quads
    dim3 qgrid((unsigned)cdiv(sky->nqpatch, 4), grid.y);
    if (mask) {
      if (chunks) VISS_BPQ(true, true); else VISS_BPQ(true, false);
    } else {
      if (chunks) VISS_BPQ(false, true); else VISS_BPQ(false, false);
    }
    VISS_LAUNCHED(k_backproject_quads);
    if (sky->nrtile > 0) {
      // few tiles (the equator ring of HEALPix): shorter row chunks, so that
      // about 16k wavefronts share the work rather than a few long ones
      // (a batch's chunk table keeps its own chunks)
      int64_t rpc = cdiv(cdiv(K * sky->nrtile, 16384), 64) * 64;
      rpc = std::min(std::max<int64_t>(rpc, 64), kpc);
      if (cdiv(K, rpc) > 65535) rpc = cdiv(cdiv(K, 65535), 64) * 64;
      dim3 rgrid((unsigned)cdiv(sky->nrtile, 4), (unsigned)(chunks ? nchunks : cdiv(K, rpc)));
      if (mask) {
        if (chunks) VISS_BPP(rgrid, sky->rpack, sky->ridx, sky->rcone, sky->nrtile, rpc, true, true);
        else VISS_BPP(rgrid, sky->rpack, sky->ridx, sky->rcone, sky->nrtile, rpc, true, false);
      } else {
        if (chunks) VISS_BPP(rgrid, sky->rpack, sky->ridx, sky->rcone, sky->nrtile, rpc, false, true);
        else VISS_BPP(rgrid, sky->rpack, sky->ridx, sky->rcone, sky->nrtile, rpc, false, false);
      }
      VISS_LAUNCHED(k_backproject_pairs);
    }
    return VISS_OK;
  }
  if (pairs) {
    if (mask) {
      if (chunks) VISS_BPP(grid, sky->ppack, sky->pidx, sky->pcone, sky->nptile, kpc, true, true);
      else VISS_BPP(grid, sky->ppack, sky->pidx, sky->pcone, sky->nptile, kpc, true, false);
    } else {
      if (chunks) VISS_BPP(grid, sky->ppack, sky->pidx, sky->pcone, sky->nptile, kpc, false, true);
      else VISS_BPP(grid, sky->ppack, sky->pidx, sky->pcone, sky->nptile, kpc, false, false);
    }
  } else if (mask) {
    if (f64) VISS_BP(true, true); else VISS_BP(true, false);
  } else {
    if (f64) VISS_BP(false, true); else VISS_BP(false, false);
  }
#undef VISS_BPQ
#undef VISS_BPP
#undef VISS_BP
  VISS_LAUNCHED(k_backproject);
  return VISS_OK;
}

// Workspace from the sky's growable buffer (allocates when K grows).
static int sky_backproject_impl(viss_sky* sky, const double* uvw, const double* xyza, const double* xyzb,
                                const double* y, const double* wt, int64_t K, const viss_mask_geometry* geom,
                                int flags, double* x, hipStream_t s) {
  int rc = check_backproject_args(uvw, xyza, xyzb, y, K, geom, flags, x);
  if (rc != VISS_OK || K == 0) return rc;
  rc = sky->ws.reserve(bp_workspace_bytes(K, flags), s);
  if (rc != VISS_OK) return rc;
  return sky_backproject_ws(sky, uvw, xyza, xyzb, y, wt, K, geom, flags, x, (char*)sky->ws.p, s, nullptr, 0, 0);
}

// ------------------------------------------------------------ allocation-free back-projection
// The row records of viss_sky_backproject_noalloc live in this array, one copy
// per device for the life of the process.  Its name is a host-side stand-in:
// host code uses only the device address hipGetSymbolAddress returns for it
// (kept per device in g_scratch_ptr), never the name.
static __device__ __align__(256) char scratch[VISS_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
static void* g_scratch_ptr[kMaxDevices] = {};
static std::mutex g_scratch_mu;

// Device address of `scratch` on the current device, looked up once per device
// (with that device current) and kept.  Allocates nothing.
static int scratch_ptr(char** out) {
  int dev = -1;
  VISS_HIP(hipGetDevice(&dev));
  VISS_REQUIRE(dev >= 0 && dev < kMaxDevices, "viss_sky_backproject_noalloc: device %d out of range", dev);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    VISS_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    VISS_REQUIRE(p != nullptr, "viss_sky_backproject_noalloc: no device address for the scratch array");
    g_scratch_ptr[dev] = p;
  }
  *out = (char*)g_scratch_ptr[dev];
  return VISS_OK;
}

}  // namespace viss

using namespace viss;

extern "C" {

int viss_sky_forward(viss_sky* sky, const double* I, const double* uvw, const double* xyz_a, const double* xyz_b,
                     int64_t K, const viss_mask_geometry* geom, int flags, double* V, void* stream) {
  VISS_REQUIRE(sky != nullptr, "viss_sky_forward: sky is NULL");
  int rc = check_device(sky);
  if (rc != VISS_OK) return rc;
  return sky_forward_impl(sky, I, uvw, xyz_a, xyz_b, K, geom, flags, V, as_stream(stream));
}

int viss_sky_backproject(viss_sky* sky, const double* uvw, const double* xyz_a, const double* xyz_b,
                         const double* y, const double* wt, int64_t K, const viss_mask_geometry* geom, int flags,
                         double* x, void* stream) {
  VISS_REQUIRE(sky != nullptr, "viss_sky_backproject: sky is NULL");
  int rc = check_device(sky);
  if (rc != VISS_OK) return rc;
  return sky_backproject_impl(sky, uvw, xyz_a, xyz_b, y, wt, K, geom, flags, x, as_stream(stream));
}

int viss_sky_backproject_noalloc(viss_sky* sky, const double* uvw, const double* xyz_a, const double* xyz_b,
                                 const double* y, const double* wt, int64_t K, const viss_mask_geometry* geom,
                                 int flags, double* x, void* stream) {
  VISS_REQUIRE(sky != nullptr, "viss_sky_backproject_noalloc: sky is NULL");
  int rc = check_device(sky);
  if (rc != VISS_OK) return rc;
  rc = check_backproject_args(uvw, xyz_a, xyz_b, y, K, geom, flags, x);
  if (rc != VISS_OK || K == 0) return rc;
  const size_t need = bp_workspace_bytes(K, flags);
  if (need > (size_t)VISS_NOALLOC_SCRATCH_BYTES) {
    set_error("viss_sky_backproject_noalloc: %lld rows need %zu B of row records, above the %d-B static "
              "scratch (at most %lld rows with these flags; split the rows over several calls)",
              (long long)K, need, VISS_NOALLOC_SCRATCH_BYTES,
              (long long)((VISS_NOALLOC_SCRATCH_BYTES - 256) / ((bp_workspace_bytes(1, flags) - 256))));
    return VISS_ERR_SIZE;
  }
  char* ws = nullptr;
  rc = scratch_ptr(&ws);
  if (rc != VISS_OK) return rc;
  return sky_backproject_ws(sky, uvw, xyz_a, xyz_b, y, wt, K, geom, flags, x, ws, as_stream(stream), nullptr, 0, 0);
}

int viss_forward(const double* lmn, const double* I, int64_t P, const double* uvw, const double* xyz_a,
                 const double* xyz_b, int64_t K, const viss_mask_geometry* geom, int flags, double* V,
                 void* stream) {
  VISS_REQUIRE(P >= 0 && K >= 0, "viss_forward: negative size");
  if (P == 0 || K == 0) return VISS_OK;
  VISS_REQUIRE(lmn != nullptr, "viss_forward: lmn is NULL");
  viss_sky* sky = nullptr;
  int rc = viss_sky_create(lmn, P, &sky, stream);
  if (rc != VISS_OK) return rc;
  rc = sky_forward_impl(sky, I, uvw, xyz_a, xyz_b, K, geom, flags, V, as_stream(stream));
  viss_sky_destroy(sky);
  return rc;
}

int viss_backproject(const double* lmn, int64_t P, const double* uvw, const double* xyz_a, const double* xyz_b,
                     const double* y, const double* wt, int64_t K, const viss_mask_geometry* geom, int flags,
                     double* x, void* stream) {
  VISS_REQUIRE(P >= 0 && K >= 0, "viss_backproject: negative size");
  if (P == 0 || K == 0) return VISS_OK;
  VISS_REQUIRE(lmn != nullptr, "viss_backproject: lmn is NULL");
  viss_sky* sky = nullptr;
  int rc = viss_sky_create(lmn, P, &sky, stream);
  if (rc != VISS_OK) return rc;
  rc = sky_backproject_impl(sky, uvw, xyz_a, xyz_b, y, wt, K, geom, flags, x, as_stream(stream));
  viss_sky_destroy(sky);
  return rc;
}

int viss_sky_count_work(viss_sky* sky, const double* xyz_a, const double* xyz_b, int64_t K,
                        const viss_mask_geometry* geom, int64_t* counts, void* stream) {
  VISS_REQUIRE(sky != nullptr, "viss_sky_count_work: sky is NULL");
  int rc = check_device(sky);
  if (rc != VISS_OK) return rc;
  VISS_REQUIRE(K >= 0, "viss_sky_count_work: negative K");
  VISS_REQUIRE(sky->use_quads, "viss_sky_count_work: the sky has no quad layout (its kernels are not the quad kernels)");
  if (K == 0) return VISS_OK;
  VISS_REQUIRE(xyz_a && xyz_b && counts && geom, "viss_sky_count_work: NULL argument");
  rc = validate_geom(geom, VISS_MASK);
  if (rc != VISS_OK) return rc;
  const int64_t bx = cdiv(K, 256);
  VISS_REQUIRE(bx <= INT32_MAX, "viss_sky_count_work: K too large for one launch");
  int64_t qpc = 0, rpc = 0;
  quad_fwd_chunks(sky, bx, qpc, rpc);
  const int64_t nqc = cdiv(sky->nqtile, qpc), nrc = cdiv(sky->nrtile, rpc);
  dim3 grid((unsigned)bx, (unsigned)(nqc + nrc));
  k_count_quad_work<<<grid, 256, 0, as_stream(stream)>>>(sky->qpack, sky->qidx, sky->qz, sky->nqtile, qpc, nqc,
                                                         sky->rpack, sky->ridx, sky->nrtile, rpc, xyz_a, xyz_b, K,
                                                         geom->R_moon, geom->orbit_radius,
                                                         (unsigned long long*)counts);
  VISS_LAUNCHED(k_count_quad_work);
  return VISS_OK;
}

int viss_sky_count_visible(viss_sky* sky, const double* xyz_a, const double* xyz_b, int64_t K,
                           const viss_mask_geometry* geom, int64_t* count, void* stream) {
  VISS_REQUIRE(sky != nullptr, "viss_sky_count_visible: sky is NULL");
  int rc = check_device(sky);
  if (rc != VISS_OK) return rc;
  VISS_REQUIRE(K >= 0, "viss_sky_count_visible: negative K");
  if (K == 0) return VISS_OK;
  VISS_REQUIRE(xyz_a && xyz_b && count && geom, "viss_sky_count_visible: NULL argument");
  rc = validate_geom(geom, VISS_MASK);
  if (rc != VISS_OK) return rc;
  const int64_t bx = cdiv(K, 256);
  int64_t tpc = kFwdTilesPerChunk;
  while (tpc > 16 && bx * cdiv(sky->ntile, tpc) < kMinBlocks) tpc /= 2;
  tpc = std::max(tpc, cdiv(sky->ntile, 65535));
  dim3 grid((unsigned)bx, (unsigned)cdiv(sky->ntile, tpc));
  k_count_visible<<<grid, 256, 0, as_stream(stream)>>>(sky->pix, sky->cone, sky->tile_valid, sky->ntile, tpc, xyz_a,
                                                       xyz_b, K, geom->R_moon, geom->orbit_radius,
                                                       (unsigned long long*)count);
  VISS_LAUNCHED(k_count_visible);
  return VISS_OK;
}

}  // extern "C"
