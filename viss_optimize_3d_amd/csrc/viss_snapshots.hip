// Snapshot-structured back-projection (adjoint): rows grouped as S snapshots
// of the same N satellites, every pair a<b of a snapshot one row, with
// baseline b = (r_a - r_b) / lambda.  This is the layout the reference's data
// come in (uvw rows built from satellite pairs, vissgen3d_noblockage.cu:
// 302-362) and the layout of every workload here.
//
// With G_i(p) = [s_p . r_i >= t_i] exp(+2 pi i d_i(p)), d_i = r_i . s_p / lambda,
// one row's term is  y_ab G_a conj(G_b)  = y_ab m_kp exp(+2 pi i b . s_p), so
//   x_p += Re sum_s sum_a G_a(p) H_a(p),   H_a = sum_{b>a} y_ab conj(G_b(p)).
// Per pixel and snapshot the N satellite phasors cost N FP64 dot products and
// N hardware sin/cos pairs instead of N(N-1)/2; each row is then one complex
// multiply-add.  The sum over (row, pixel) pairs is the same as the per-row
// kernel's (k_backproject): the occultation factor is exact per satellite
// (FP64 s.r >= t, the per-row test's two factors), the phase d_a - d_b equals
// b . s up to FP64 rounding, and the reductions are the same blocked FP32 ->
// FP64 sums.
#include <algorithm>
#include <cstdlib>

#include "viss_kernels.h"

namespace viss {

struct SatRec {  // one satellite of one snapshot
  double x, y, z, t;
};
static_assert(sizeof(SatRec) == 32, "SatRec is 32 B");

struct SatCls {  // unit direction and unit cut, for tile skipping
  float ax, ay, az, c;
};

// Per (snapshot, satellite): exact occultation record and classification
// record, NP >= N records per snapshot.  Records N..NP-1 of a snapshot are
// padding: a satellite at the origin that sees nothing (its phasor is 0 under
// the mask, and rows that involve it are never written).
__global__ void k_snap_sats(const double* __restrict__ pos, int64_t S, int N, int NP, double R_moon, double orbit_r,
                            int mask, SatRec* __restrict__ rec, SatCls* __restrict__ cls) {
  const int64_t SNP = S * NP;
  for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < SNP; j += (int64_t)gridDim.x * blockDim.x) {
    const int64_t sn = j / NP;
    const int k = (int)(j - sn * NP);
    if (k >= N) {
      rec[j] = SatRec{0.0, 0.0, 0.0, 1e300};
      cls[j] = SatCls{0.f, 0.f, 1.f, 4.0f};
      continue;
    }
    const int64_t i = sn * N + k;
    const double x = pos[3 * i], y = pos[3 * i + 1], z = pos[3 * i + 2];
    SatRec r;
    r.x = x;
    r.y = y;
    r.z = z;
    r.t = mask ? mask_threshold(x, y, z, R_moon, orbit_r) : -1e300;
    rec[j] = r;
    const double nr = sqrt(x * x + y * y + z * z);
    const double inv = nr > 0 ? 1.0 / nr : 0.0;
    SatCls c;
    c.ax = (float)(x * inv);
    c.ay = (float)(y * inv);
    c.az = (float)(z * inv);
    c.c = (mask && nr > 0) ? (float)(r.t * inv) : -2.0f;
    cls[j] = c;
  }
}

// Padded pair slot q in [0, NP (NP - 1) / 2) -> satellites (a, b), a < b < NP,
// in row-major order (the order of a snapshot's rows when NP = N).
__device__ __forceinline__ void pair_of(int q, int NP, int& a, int& b) {
  a = 0;
  while (q >= NP - 1 - a) {
    q -= NP - 1 - a;
    ++a;
  }
  b = a + 1 + q;
}

// Row of pair (a, b) within a snapshot of N satellites.
__device__ __forceinline__ int row_in_snapshot(int a, int b, int N) { return a * N - a * (a + 1) / 2 + (b - a - 1); }

// Per padded pair slot: y' = wt * y [* exp(-2 pi i w)], w = (z_a - z_b) / lambda,
// as FP32 complex; 0 for the slots of a padding satellite.
__global__ void k_snap_rows(const double* __restrict__ pos, const double* __restrict__ y,
                            const double* __restrict__ wt, int64_t S, int N, int NP, double inv_lambda, int track,
                            float2* __restrict__ yw) {
  const int NB = N * (N - 1) / 2, NBP = NP * (NP - 1) / 2;
  const int64_t J = S * NBP;
  for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < J; j += (int64_t)gridDim.x * blockDim.x) {
    const int64_t s = j / NBP;
    int a, b;
    pair_of((int)(j - s * NBP), NP, a, b);
    if (b >= N) {
      yw[j] = make_float2(0.f, 0.f);
      continue;
    }
    const int64_t k = s * NB + row_in_snapshot(a, b, N);
    double yr = y[2 * k], yi = y[2 * k + 1];
    if (track) {
      const double w = (pos[3 * (s * N + a) + 2] - pos[3 * (s * N + b) + 2]) * inv_lambda;
      double sn, cs;
      sincospi(-2.0 * (w - rint(w)), &sn, &cs);
      const double nr = yr * cs - yi * sn, ni = yr * sn + yi * cs;
      yr = nr;
      yi = ni;
    }
    if (wt) {
      yr *= wt[k];
      yi *= wt[k];
    }
    yw[j] = make_float2((float)yr, (float)yi);
  }
}

// Conservative cone test of the N real satellites of one snapshot against a
// tile: any_v = some satellite may see some pixel, all_v = every satellite
// sees every pixel.
template <int NMAX>
__device__ __forceinline__ void classify_snapshot(const float4 cn, const SatCls* __restrict__ c, int N, bool& any_v,
                                                  bool& all_v) {
  any_v = false;
  all_v = true;
#pragma unroll
  for (int i = 0; i < NMAX; ++i) {
    if (i < N) {
      const SatCls q = c[i];
      const float d = fmaf(cn.x, q.ax, fmaf(cn.y, q.ay, cn.z * q.az));
      any_v |= d + cn.w >= q.c - kClsEps;
      all_v &= d - cn.w >= q.c + kClsEps;
    }
  }
}

// Satellite phasors of PPL pixels: G_i = [s.r_i >= t_i] exp(+2 pi i r_i.s / lambda).
// VIS = false: every satellite sees every pixel (no test).
template <bool VIS, int NMAX, int PPL>
__device__ __forceinline__ void snapshot_phasors(const SatRec* __restrict__ R, const double (&l)[PPL],
                                                 const double (&m)[PPL], const double (&n)[PPL], double inv_lambda,
                                                 float (&gr)[NMAX][PPL], float (&gi)[NMAX][PPL]) {
#pragma unroll
  for (int i = 0; i < NMAX; ++i) {
    const SatRec q = R[i];
#pragma unroll
    for (int p = 0; p < PPL; ++p) {
      const double D = occ_dot(l[p], m[p], n[p] * q.z, q.x, q.y);
      float sn, cs;
      sincos_rev(rev_fp32(D * inv_lambda), sn, cs);
      const bool vis = !VIS || (D >= q.t);
      gr[i][p] = vis ? cs : 0.f;
      gi[i][p] = vis ? sn : 0.f;
    }
  }
}

// One snapshot's contribution to PPL pixels: Re sum_a G_a H_a, H_a = sum_{b>a} y_ab conj(G_b).
template <bool VIS, int NMAX, int PPL>
__device__ __forceinline__ void bp_snapshot(const SatRec* __restrict__ R, const float2* __restrict__ Y,
                                            const double (&l)[PPL], const double (&m)[PPL], const double (&n)[PPL],
                                            double inv_lambda, double (&acc)[PPL]) {
  float gr[NMAX][PPL], gi[NMAX][PPL];
  snapshot_phasors<VIS, NMAX, PPL>(R, l, m, n, inv_lambda, gr, gi);
  float part[PPL];
#pragma unroll
  for (int p = 0; p < PPL; ++p) part[p] = 0.f;
  int pr = 0;
#pragma unroll
  for (int a = 0; a < NMAX - 1; ++a) {
    float hr[PPL], hi[PPL];
#pragma unroll
    for (int p = 0; p < PPL; ++p) hr[p] = hi[p] = 0.f;
#pragma unroll
    for (int b = a + 1; b < NMAX; ++b) {
      const float2 yv = Y[pr++];
#pragma unroll
      for (int p = 0; p < PPL; ++p) {  // H += y conj(G_b)
        hr[p] = fmaf(yv.x, gr[b][p], fmaf(yv.y, gi[b][p], hr[p]));
        hi[p] = fmaf(yv.y, gr[b][p], fmaf(-yv.x, gi[b][p], hi[p]));
      }
    }
#pragma unroll
    for (int p = 0; p < PPL; ++p) part[p] = fmaf(gr[a][p], hr[p], fmaf(-gi[a][p], hi[p], part[p]));
  }
#pragma unroll
  for (int p = 0; p < PPL; ++p) acc[p] += (double)part[p];
}

// Grid: x = ceil(nbtile / 4) blocks of 4 wavefronts (one kBpTile-pixel tile per
// wavefront), y = snapshot chunks.  Records and rows are padded to NMAX
// satellites per snapshot (padding rows carry y = 0), so the pair loops have
// no per-satellite branches.  Snapshots are classified against the tile 64
// at a time, one per lane; those every satellite sees whole skip the
// per-pixel occultation test, those no satellite sees are skipped.  Held to 6
// wavefronts per SIMD up to NMAX 12 (the masked kernels otherwise get 81-97
// VGPRs, 4-5 wavefronts; at NMAX 16 six would spill VGPRs to scratch).
template <bool MASK, int NMAX>
__global__ __launch_bounds__(256, NMAX <= 12 ? 6 : 1) void k_backproject_snap(const PixPack* __restrict__ pix,
                                                          const float4* __restrict__ bcone, int64_t nbtile,
                                                          const SatRec* __restrict__ rec,
                                                          const SatCls* __restrict__ cls,
                                                          const float2* __restrict__ yw, int64_t S, int N,
                                                          double inv_lambda, int64_t s_per_chunk, double herm,
                                                          double* __restrict__ x) {
  constexpr int PPL = kBpPixPerLane;
  constexpr int NBP = NMAX * (NMAX - 1) / 2;
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
  const int64_t s0 = (int64_t)blockIdx.y * s_per_chunk;
  const int64_t s1 = min(S, s0 + s_per_chunk);
  double acc[PPL];
#pragma unroll
  for (int i = 0; i < PPL; ++i) acc[i] = 0.0;
  for (int64_t sb = s0; sb < s1; sb += 64) {
    const int ns = (int)min((int64_t)64, s1 - sb);
    uint64_t anym, fullm;
    if (MASK) {
      bool any_v = false, all_v = false;
      if (lane < ns) classify_snapshot<NMAX>(cn, cls + (sb + lane) * NMAX, N, any_v, all_v);
      anym = wave_ballot(any_v);
      fullm = wave_ballot(all_v);
    } else {
      anym = ns == 64 ? ~0ull : ((1ull << ns) - 1ull);
      fullm = anym;
    }
    // fully visible snapshots, then partly visible ones: two loops rather
    // than a branch per snapshot, so the accumulators stay in place
    uint64_t fm = anym & fullm;
    while (fm) {
      const int64_t s = sb + __builtin_ctzll(fm);
      fm &= fm - 1ull;
      bp_snapshot<false, NMAX, PPL>(rec + s * NMAX, yw + s * NBP, l, m, n, inv_lambda, acc);
    }
    uint64_t pm = MASK ? (anym & ~fullm) : 0ull;
    while (pm) {
      const int64_t s = sb + __builtin_ctzll(pm);
      pm &= pm - 1ull;
      bp_snapshot<true, NMAX, PPL>(rec + s * NMAX, yw + s * NBP, l, m, n, inv_lambda, acc);
    }
  }
#pragma unroll
  for (int p = 0; p < PPL; ++p)
    if (idx[p] >= 0) unsafeAtomicAdd(&x[idx[p]], herm * acc[p]);
}

// One 64-pixel tile's contribution to a snapshot's pair sums (one pixel per
// lane): (ar, ai)[pair a<b] += I conj(G_a) G_b.
template <bool VIS, int NMAX>
__device__ __forceinline__ void fwd_snapshot_tile(const PixPack& q, const SatRec* __restrict__ R, double inv_lambda,
                                                  float (&ar)[NMAX * (NMAX - 1) / 2],
                                                  float (&ai)[NMAX * (NMAX - 1) / 2]) {
  const double l[1] = {q.l}, m[1] = {q.m}, n[1] = {q.n};
  float gr[NMAX][1], gi[NMAX][1];
  snapshot_phasors<VIS, NMAX, 1>(R, l, m, n, inv_lambda, gr, gi);
  int pr = 0;
#pragma unroll
  for (int a = 0; a < NMAX - 1; ++a) {
    const float fr = q.I * gr[a][0], fi = -q.I * gi[a][0];  // I conj(G_a)
#pragma unroll
    for (int b = a + 1; b < NMAX; ++b, ++pr) {  // F_a G_b
      ar[pr] = fmaf(fr, gr[b][0], fmaf(-fi, gi[b][0], ar[pr]));
      ai[pr] = fmaf(fr, gi[b][0], fmaf(fi, gr[b][0], ai[pr]));
    }
  }
}

// Sums a snapshot's FP32 row partials over the wavefront's lanes (xor
// butterfly: every lane ends with every total); lane j keeps rows j and j + 64
// and writes them (phase-track rotation, Hermitian rows), so the rows are
// written by the wavefront in parallel.
template <int NMAX>
__device__ __forceinline__ void snapshot_store(float (&ar)[NMAX * (NMAX - 1) / 2], float (&ai)[NMAX * (NMAX - 1) / 2],
                                               int lane, int64_t s, int64_t S, int N,
                                               const SatRec* __restrict__ R, double inv_lambda, int track, int herm,
                                               double* __restrict__ V) {
  constexpr int NBM = NMAX * (NMAX - 1) / 2;
  constexpr int NSLOT = (NBM + 63) / 64;
  float myr[NSLOT], myi[NSLOT];
#pragma unroll
  for (int u = 0; u < NSLOT; ++u) myr[u] = myi[u] = 0.f;
#pragma unroll
  for (int q = 0; q < NBM; ++q) {
    float vr = ar[q], vi = ai[q];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      vr += __shfl_xor(vr, o, 64);
      vi += __shfl_xor(vi, o, 64);
    }
    if (lane == (q & 63)) {
      myr[q >> 6] = vr;
      myi[q >> 6] = vi;
    }
  }
  const int NB = N * (N - 1) / 2;
  const int64_t K = S * NB;
#pragma unroll
  for (int u = 0; u < NSLOT; ++u) {
    const int q = u * 64 + lane;
    if (q >= NBM) break;
    int a, b;
    pair_of(q, NMAX, a, b);
    if (b >= N) continue;  // a padding satellite
    const int64_t k = s * NB + row_in_snapshot(a, b, N);
    double vr = myr[u], vi = myi[u];
    if (track) {  // V e^{+2 pi i w}, w = (z_a - z_b) / lambda
      const double w = (R[a].z - R[b].z) * inv_lambda;
      double sn, cs;
      sincospi(2.0 * (w - rint(w)), &sn, &cs);
      const double nr = vr * cs - vi * sn, ni = vr * sn + vi * cs;
      vr = nr;
      vi = ni;
    }
    unsafeAtomicAdd(&V[2 * k], vr);
    unsafeAtomicAdd(&V[2 * k + 1], vi);
    if (herm) {
      unsafeAtomicAdd(&V[2 * (k + K)], vr);
      unsafeAtomicAdd(&V[2 * (k + K) + 1], -vi);
    }
  }
}

// Forward over snapshot-structured rows: V_ab(s) = sum_p I_p conj(G_a(p)) G_b(p)
// (= sum_p I_p m_kp exp(-2 pi i b . s_p)).  One wavefront per (snapshot,
// pixel range); each lane owns one pixel of every 64-pixel tile of the range
// and keeps FP32 partial sums for every pair of the snapshot in registers.
// Tiles are classified against the snapshot 64 at a time, one per lane.  At
// the end of the range the partials are summed across the wavefront and added
// into V with FP64 atomics.  Records are padded to NMAX satellites per
// snapshot: pairs with a padding satellite are accumulated and dropped, which
// keeps per-satellite branches out of the pair loop.  NMAX <= 10 holds the
// 90 FP32 partial sums in at most 168 VGPRs, 3 wavefronts per SIMD.
template <bool MASK, int NMAX>
__global__ __launch_bounds__(256, NMAX <= 10 ? 3 : 1) void k_forward_snap(
    const PixPack* __restrict__ pix, const float4* __restrict__ cone, int64_t ntile, int64_t tiles_per_range,
    const SatRec* __restrict__ rec, const SatCls* __restrict__ cls, int64_t S, int N, double inv_lambda, int track,
    int herm, double* __restrict__ V) {
  constexpr int NBM = NMAX * (NMAX - 1) / 2;
  const int lane = threadIdx.x & 63;
  const int64_t s = (int64_t)blockIdx.x * 4 + wave_in_block();
  if (s >= S) return;  // wave-uniform
  const int64_t t0 = (int64_t)blockIdx.y * tiles_per_range;
  const int64_t t1 = min(ntile, t0 + tiles_per_range);
  const SatRec* __restrict__ R = rec + s * NMAX;
  float ar[NBM], ai[NBM];
#pragma unroll
  for (int q = 0; q < NBM; ++q) ar[q] = ai[q] = 0.f;
  for (int64_t tb = t0; tb < t1; tb += 64) {
    const int nt = (int)min((int64_t)64, t1 - tb);
    uint64_t anym, fullm;
    if (MASK) {
      bool any_v = false, all_v = false;
      if (lane < nt) classify_snapshot<NMAX>(cone[tb + lane], cls + s * NMAX, N, any_v, all_v);
      anym = wave_ballot(any_v);
      fullm = wave_ballot(all_v);
    } else {
      anym = nt == 64 ? ~0ull : ((1ull << nt) - 1ull);
      fullm = anym;
    }
    uint64_t fm = anym & fullm;
    while (fm) {
      const int64_t t = tb + __builtin_ctzll(fm);
      fm &= fm - 1ull;
      fwd_snapshot_tile<false, NMAX>(pix[t * kTile + lane], R, inv_lambda, ar, ai);
    }
    uint64_t pm = MASK ? (anym & ~fullm) : 0ull;
    while (pm) {
      const int64_t t = tb + __builtin_ctzll(pm);
      pm &= pm - 1ull;
      fwd_snapshot_tile<true, NMAX>(pix[t * kTile + lane], R, inv_lambda, ar, ai);
    }
  }
  snapshot_store<NMAX>(ar, ai, lane, s, S, N, R, inv_lambda, track, herm, V);
}

// ------------------------------------------------ antipodal-pair sky (mixed)
// For a pair (a at s, b at -s), r_i.s_b = -(r_i.s_a): the b pixel's phasor is
// the conjugate of a's, g'_i = conj(g_i), masked by its own test -D_i >= t_i.
// Where every satellite sees both pixels of every pair of a tile whole, the
// two pixels share the phasors and the FP32 row work:
//   forward  V_ij += I_a conj(g_i) g_j + I_b g_i conj(g_j)
//                 = (I_a + I_b) Re F_ij + i (I_a - I_b) Im F_ij,  F_ij = conj(g_i) g_j;
//   adjoint  with P_i = sum_j y_ij Re g_j, Q_i = sum_j y_ij Im g_j (complex),
//            H_i = P_i - i Q_i for a and H'_i = P_i + i Q_i for b.
// Modes of a (snapshot, pair tile): 0 both sides seen whole, 1 only a (b
// hidden), 2 only b, 3 anything else (both sides with the per-pixel test).
enum SnapPairMode { kSnapBoth = 0, kSnapA = 1, kSnapB = 2, kSnapMasked = 3 };

template <int NMAX>
__device__ __forceinline__ void classify_snapshot_pair(const float4 cn, const SatCls* __restrict__ c, int N,
                                                       bool& anyA, bool& allA, bool& anyB, bool& allB) {
  anyA = anyB = false;
  allA = allB = true;
#pragma unroll
  for (int i = 0; i < NMAX; ++i) {
    if (i < N) {
      const SatCls q = c[i];
      const float d = fmaf(cn.x, q.ax, fmaf(cn.y, q.ay, cn.z * q.az));
      anyA |= d + cn.w >= q.c - kClsEps;
      allA &= d - cn.w >= q.c + kClsEps;
      anyB |= -d + cn.w >= q.c - kClsEps;
      allB &= -d - cn.w >= q.c + kClsEps;
    }
  }
}

// Phasors of pixel a: (gr, gi) unmasked; in kSnapMasked also the masked a-side
// phasors (ar_, ai_) and b-side phasors (br_, bi_) = [-D >= t] conj(g).
template <int MODE, int NMAX>
__device__ __forceinline__ void pair_phasors(const SatRec* __restrict__ R, double l, double m, double n,
                                             double inv_lambda, float (&gr)[NMAX], float (&gi)[NMAX],
                                             float (&ga_r)[NMAX], float (&ga_i)[NMAX], float (&gb_r)[NMAX],
                                             float (&gb_i)[NMAX]) {
#pragma unroll
  for (int i = 0; i < NMAX; ++i) {
    const SatRec q = R[i];
    const double D = occ_dot(l, m, n * q.z, q.x, q.y);
    float sn, cs;
    sincos_rev(rev_fp32(D * inv_lambda), sn, cs);
    gr[i] = cs;
    gi[i] = sn;
    if (MODE == kSnapMasked) {
      const bool va = D >= q.t, vb = -D >= q.t;
      ga_r[i] = va ? cs : 0.f;
      ga_i[i] = va ? sn : 0.f;
      gb_r[i] = vb ? cs : 0.f;
      gb_i[i] = vb ? -sn : 0.f;
    }
  }
}

// One snapshot's contribution to the pair (a, b) of this lane.
template <int MODE, int NMAX>
__device__ __forceinline__ void bp_snapshot_pair(const SatRec* __restrict__ R, const float2* __restrict__ Y, double l,
                                                 double m, double n, double inv_lambda, double& acc_a, double& acc_b) {
  float gr[NMAX], gi[NMAX], ga_r[NMAX], ga_i[NMAX], gb_r[NMAX], gb_i[NMAX];
  pair_phasors<MODE, NMAX>(R, l, m, n, inv_lambda, gr, gi, ga_r, ga_i, gb_r, gb_i);
  float pa = 0.f, pb = 0.f;
  int pr = 0;
#pragma unroll
  for (int i = 0; i < NMAX - 1; ++i) {
    if (MODE == kSnapBoth) {
      float Pr = 0.f, Pi = 0.f, Qr = 0.f, Qi = 0.f;
#pragma unroll
      for (int j = i + 1; j < NMAX; ++j) {
        const float2 y = Y[pr++];
        Pr = fmaf(y.x, gr[j], Pr);
        Pi = fmaf(y.y, gr[j], Pi);
        Qr = fmaf(y.x, gi[j], Qr);
        Qi = fmaf(y.y, gi[j], Qi);
      }
      pa = fmaf(gr[i], Pr + Qi, fmaf(-gi[i], Pi - Qr, pa));  // Re g_i (P - iQ)
      pb = fmaf(gr[i], Pr - Qi, fmaf(gi[i], Pi + Qr, pb));   // Re conj(g_i) (P + iQ)
    } else if (MODE == kSnapA) {
      float hr = 0.f, hi = 0.f;
#pragma unroll
      for (int j = i + 1; j < NMAX; ++j) {  // H += y conj(g_j)
        const float2 y = Y[pr++];
        hr = fmaf(y.x, gr[j], fmaf(y.y, gi[j], hr));
        hi = fmaf(y.y, gr[j], fmaf(-y.x, gi[j], hi));
      }
      pa = fmaf(gr[i], hr, fmaf(-gi[i], hi, pa));
    } else if (MODE == kSnapB) {
      float hr = 0.f, hi = 0.f;
#pragma unroll
      for (int j = i + 1; j < NMAX; ++j) {  // H' += y g_j
        const float2 y = Y[pr++];
        hr = fmaf(y.x, gr[j], fmaf(-y.y, gi[j], hr));
        hi = fmaf(y.y, gr[j], fmaf(y.x, gi[j], hi));
      }
      pb = fmaf(gr[i], hr, fmaf(gi[i], hi, pb));  // Re conj(g_i) H'
    } else {
      float har = 0.f, hai = 0.f, hbr = 0.f, hbi = 0.f;
#pragma unroll
      for (int j = i + 1; j < NMAX; ++j) {  // H += y conj(G_j) for each side
        const float2 y = Y[pr++];
        har = fmaf(y.x, ga_r[j], fmaf(y.y, ga_i[j], har));
        hai = fmaf(y.y, ga_r[j], fmaf(-y.x, ga_i[j], hai));
        hbr = fmaf(y.x, gb_r[j], fmaf(y.y, gb_i[j], hbr));
        hbi = fmaf(y.y, gb_r[j], fmaf(-y.x, gb_i[j], hbi));
      }
      pa = fmaf(ga_r[i], har, fmaf(-ga_i[i], hai, pa));
      pb = fmaf(gb_r[i], hbr, fmaf(-gb_i[i], hbi, pb));
    }
  }
  acc_a += (double)pa;
  acc_b += (double)pb;
}

// The adjoint over a paired sky: one wavefront per 64-pair tile, one pair per
// lane; snapshots classified 64 at a time (one per lane) against both sides of
// the tile and run in the four mode loops.
// Held to 6 wavefronts per SIMD up to NMAX 12, as k_backproject_snap: at 10
// satellites, 900,000 rows x 786,432 pixels, 71.8 ms against 75.4 ms with the
// masked kernel's own 81 VGPRs (5 wavefronts).
template <bool MASK, int NMAX>
__global__ __launch_bounds__(256, NMAX <= 12 ? 6 : 1) void k_backproject_snap_pairs(const PairPack* __restrict__ pp,
                                                                const int2* __restrict__ pidx,
                                                                const float4* __restrict__ pcone, int64_t ntile,
                                                                const SatRec* __restrict__ rec,
                                                                const SatCls* __restrict__ cls,
                                                                const float2* __restrict__ yw, int64_t S, int N,
                                                                double inv_lambda, int64_t s_per_chunk, double herm,
                                                                double* __restrict__ x) {
  constexpr int NBP = NMAX * (NMAX - 1) / 2;
  const int lane = threadIdx.x & 63;
  const int64_t tile = (int64_t)blockIdx.x * 4 + wave_in_block();
  if (tile >= ntile) return;  // wave-uniform
  const PairPack q = pp[tile * 64 + lane];
  const int2 ab = pidx[tile * 64 + lane];
  float4 cn = make_float4(0.f, 0.f, 1.f, 2.f);
  if (MASK) cn = pcone[tile];
  const int64_t s0 = (int64_t)blockIdx.y * s_per_chunk;
  const int64_t s1 = min(S, s0 + s_per_chunk);
  double acc_a = 0.0, acc_b = 0.0;
  for (int64_t sb = s0; sb < s1; sb += 64) {
    const int ns = (int)min((int64_t)64, s1 - sb);
    uint64_t both, onlyA, onlyB, masked;
    if (MASK) {
      bool anyA = false, allA = false, anyB = false, allB = false;
      if (lane < ns) classify_snapshot_pair<NMAX>(cn, cls + (sb + lane) * NMAX, N, anyA, allA, anyB, allB);
      const uint64_t sA = wave_ballot(anyA), fA = wave_ballot(allA), sB = wave_ballot(anyB), fB = wave_ballot(allB);
      both = fA & fB;
      onlyA = fA & ~sB;
      onlyB = fB & ~sA;
      masked = (sA | sB) & ~(both | onlyA | onlyB);
    } else {
      both = ns == 64 ? ~0ull : ((1ull << ns) - 1ull);
      onlyA = onlyB = masked = 0ull;
    }
    while (both) {
      const int64_t s = sb + __builtin_ctzll(both);
      both &= both - 1ull;
      bp_snapshot_pair<kSnapBoth, NMAX>(rec + s * NMAX, yw + s * NBP, q.l, q.m, q.n, inv_lambda, acc_a, acc_b);
    }
    while (onlyA) {
      const int64_t s = sb + __builtin_ctzll(onlyA);
      onlyA &= onlyA - 1ull;
      bp_snapshot_pair<kSnapA, NMAX>(rec + s * NMAX, yw + s * NBP, q.l, q.m, q.n, inv_lambda, acc_a, acc_b);
    }
    while (onlyB) {
      const int64_t s = sb + __builtin_ctzll(onlyB);
      onlyB &= onlyB - 1ull;
      bp_snapshot_pair<kSnapB, NMAX>(rec + s * NMAX, yw + s * NBP, q.l, q.m, q.n, inv_lambda, acc_a, acc_b);
    }
    while (masked) {
      const int64_t s = sb + __builtin_ctzll(masked);
      masked &= masked - 1ull;
      bp_snapshot_pair<kSnapMasked, NMAX>(rec + s * NMAX, yw + s * NBP, q.l, q.m, q.n, inv_lambda, acc_a, acc_b);
    }
  }
  if (ab.x >= 0) unsafeAtomicAdd(&x[ab.x], herm * acc_a);
  if (ab.y >= 0) unsafeAtomicAdd(&x[ab.y], herm * acc_b);
}

// One pair tile's contribution to a snapshot's row sums (one pair per lane).
template <int MODE, int NMAX>
__device__ __forceinline__ void fwd_snapshot_pair_tile(const PairPack& q, const SatRec* __restrict__ R,
                                                       double inv_lambda, float (&ar)[NMAX * (NMAX - 1) / 2],
                                                       float (&ai)[NMAX * (NMAX - 1) / 2]) {
  float gr[NMAX], gi[NMAX], ga_r[NMAX], ga_i[NMAX], gb_r[NMAX], gb_i[NMAX];
  pair_phasors<MODE, NMAX>(R, q.l, q.m, q.n, inv_lambda, gr, gi, ga_r, ga_i, gb_r, gb_i);
  int pr = 0;
#pragma unroll
  for (int i = 0; i < NMAX - 1; ++i) {
    if (MODE == kSnapBoth) {  // (I_a + I_b) Re F + i (I_a - I_b) Im F
      const float wAr = q.A * gr[i], wAi = q.A * gi[i], wDr = q.D * gr[i], wDi = q.D * gi[i];
#pragma unroll
      for (int j = i + 1; j < NMAX; ++j, ++pr) {
        ar[pr] = fmaf(wAr, gr[j], fmaf(wAi, gi[j], ar[pr]));
        ai[pr] = fmaf(wDr, gi[j], fmaf(-wDi, gr[j], ai[pr]));
      }
    } else if (MODE == kSnapA) {  // I_a conj(g_i) g_j
      const float fr = q.Ia * gr[i], fi = -q.Ia * gi[i];
#pragma unroll
      for (int j = i + 1; j < NMAX; ++j, ++pr) {
        ar[pr] = fmaf(fr, gr[j], fmaf(-fi, gi[j], ar[pr]));
        ai[pr] = fmaf(fr, gi[j], fmaf(fi, gr[j], ai[pr]));
      }
    } else if (MODE == kSnapB) {  // I_b g_i conj(g_j)
      const float fr = q.Ib * gr[i], fi = q.Ib * gi[i];
#pragma unroll
      for (int j = i + 1; j < NMAX; ++j, ++pr) {
        ar[pr] = fmaf(fr, gr[j], fmaf(fi, gi[j], ar[pr]));
        ai[pr] = fmaf(-fr, gi[j], fmaf(fi, gr[j], ai[pr]));
      }
    } else {  // I_a conj(G_i) G_j + I_b conj(G'_i) G'_j with the masked phasors of each side
      const float far_ = q.Ia * ga_r[i], fai = -q.Ia * ga_i[i];
      const float fbr = q.Ib * gb_r[i], fbi = -q.Ib * gb_i[i];
#pragma unroll
      for (int j = i + 1; j < NMAX; ++j, ++pr) {
        ar[pr] = fmaf(far_, ga_r[j], fmaf(-fai, ga_i[j], fmaf(fbr, gb_r[j], fmaf(-fbi, gb_i[j], ar[pr]))));
        ai[pr] = fmaf(far_, ga_i[j], fmaf(fai, ga_r[j], fmaf(fbr, gb_i[j], fmaf(fbi, gb_r[j], ai[pr]))));
      }
    }
  }
}

// The forward over a paired sky: one wavefront per (snapshot, range of pair
// tiles), one pair per lane; tiles classified 64 at a time (one per lane)
// against both sides.  Epilogue as k_forward_snap.
template <bool MASK, int NMAX>
__global__ __launch_bounds__(256, NMAX <= 10 ? 3 : 1) void k_forward_snap_pairs(
    const PairPack* __restrict__ pp, const float4* __restrict__ pcone, int64_t ntile, int64_t tiles_per_range,
    const SatRec* __restrict__ rec, const SatCls* __restrict__ cls, int64_t S, int N, double inv_lambda, int track,
    int herm, double* __restrict__ V) {
  constexpr int NBM = NMAX * (NMAX - 1) / 2;
  const int lane = threadIdx.x & 63;
  const int64_t s = (int64_t)blockIdx.x * 4 + wave_in_block();
  if (s >= S) return;  // wave-uniform
  const int64_t t0 = (int64_t)blockIdx.y * tiles_per_range;
  const int64_t t1 = min(ntile, t0 + tiles_per_range);
  const SatRec* __restrict__ R = rec + s * NMAX;
  float ar[NBM], ai[NBM];
#pragma unroll
  for (int k = 0; k < NBM; ++k) ar[k] = ai[k] = 0.f;
  for (int64_t tb = t0; tb < t1; tb += 64) {
    const int nt = (int)min((int64_t)64, t1 - tb);
    uint64_t both, onlyA, onlyB, masked;
    if (MASK) {
      bool anyA = false, allA = false, anyB = false, allB = false;
      if (lane < nt) classify_snapshot_pair<NMAX>(pcone[tb + lane], cls + s * NMAX, N, anyA, allA, anyB, allB);
      const uint64_t sA = wave_ballot(anyA), fA = wave_ballot(allA), sB = wave_ballot(anyB), fB = wave_ballot(allB);
      both = fA & fB;
      onlyA = fA & ~sB;
      onlyB = fB & ~sA;
      masked = (sA | sB) & ~(both | onlyA | onlyB);
    } else {
      both = nt == 64 ? ~0ull : ((1ull << nt) - 1ull);
      onlyA = onlyB = masked = 0ull;
    }
    while (both) {
      const int64_t t = tb + __builtin_ctzll(both);
      both &= both - 1ull;
      fwd_snapshot_pair_tile<kSnapBoth, NMAX>(pp[t * 64 + lane], R, inv_lambda, ar, ai);
    }
    while (onlyA) {
      const int64_t t = tb + __builtin_ctzll(onlyA);
      onlyA &= onlyA - 1ull;
      fwd_snapshot_pair_tile<kSnapA, NMAX>(pp[t * 64 + lane], R, inv_lambda, ar, ai);
    }
    while (onlyB) {
      const int64_t t = tb + __builtin_ctzll(onlyB);
      onlyB &= onlyB - 1ull;
      fwd_snapshot_pair_tile<kSnapB, NMAX>(pp[t * 64 + lane], R, inv_lambda, ar, ai);
    }
    while (masked) {
      const int64_t t = tb + __builtin_ctzll(masked);
      masked &= masked - 1ull;
      fwd_snapshot_pair_tile<kSnapMasked, NMAX>(pp[t * 64 + lane], R, inv_lambda, ar, ai);
    }
  }
  snapshot_store<NMAX>(ar, ai, lane, s, S, N, R, inv_lambda, track, herm, V);
}

static int64_t cdiv_s(int64_t a, int64_t b) { return (a + b - 1) / b; }

// Satellites per snapshot in the padded records: the kernels are instantiated
// for these counts.
static int padded_sats(int N) { return N <= 4 ? 4 : N <= 8 ? 8 : N <= 10 ? 10 : N <= 12 ? 12 : 16; }

static unsigned grid_rows(int64_t n) {
  int64_t g = cdiv_s(n, 256);
  if (g > 65536) g = 65536;
  return (unsigned)(g < 1 ? 1 : g);
}

}  // namespace viss

using namespace viss;

extern "C" int viss_sky_backproject_snapshots(viss_sky* sky, const double* sat_pos, int64_t S, int32_t N,
                                              double wavelength, const double* y, const double* wt,
                                              const viss_mask_geometry* geom, int flags, double* x, void* stream) {
  VISS_REQUIRE(sky != nullptr, "viss_sky_backproject_snapshots: sky is NULL");
  int rc = check_device(sky);
  if (rc != VISS_OK) return rc;
  VISS_REQUIRE(S >= 0 && N >= 2 && N <= 16, "viss_sky_backproject_snapshots: need S >= 0 and 2 <= N <= 16");
  VISS_REQUIRE(wavelength > 0.0, "viss_sky_backproject_snapshots: wavelength must be > 0");
  VISS_REQUIRE((flags & ~(VISS_MASK | VISS_PHASE_TRACK | VISS_HERMITIAN)) == 0,
               "viss_sky_backproject_snapshots: unsupported flags 0x%x (mixed precision only)", flags);
  if (S == 0) return VISS_OK;
  VISS_REQUIRE(sat_pos && y && x, "viss_sky_backproject_snapshots: NULL argument");
  const bool mask = flags & VISS_MASK;
  if (mask) {
    VISS_REQUIRE(geom != nullptr && geom->R_moon >= 0.0 &&
                     (geom->orbit_radius <= 0.0 || geom->orbit_radius > geom->R_moon),
                 "viss_sky_backproject_snapshots: bad geometry");
  }
  hipStream_t s = as_stream(stream);
  const int NP = padded_sats(N);  // the kernel's NMAX
  const int64_t NBP = (int64_t)NP * (NP - 1) / 2, SNP = S * NP, J = S * NBP;
  const size_t b_rec = sizeof(SatRec) * SNP, b_cls = sizeof(SatCls) * SNP, b_y = sizeof(float2) * J;
  rc = sky->ws.reserve(b_rec + b_cls + b_y + 512, s);
  if (rc != VISS_OK) return rc;
  char* base = (char*)sky->ws.p;
  SatRec* rec = (SatRec*)base;
  SatCls* cls = (SatCls*)(base + b_rec);
  float2* yw = (float2*)(base + ((b_rec + b_cls + 255) / 256) * 256);
  const double inv_lambda = 1.0 / wavelength;
  k_snap_sats<<<grid_rows(SNP), 256, 0, s>>>(sat_pos, S, N, NP, mask ? geom->R_moon : 0.0,
                                             mask ? geom->orbit_radius : 0.0, mask ? 1 : 0, rec, cls);
  VISS_LAUNCHED(k_snap_sats);
  k_snap_rows<<<grid_rows(J), 256, 0, s>>>(sat_pos, y, wt, S, N, NP, inv_lambda, (flags & VISS_PHASE_TRACK) ? 1 : 0,
                                           yw);
  VISS_LAUNCHED(k_snap_rows);
  const bool pairs = sky->use_pairs;  // one lane per antipodal pair
  const int64_t bx = cdiv_s(pairs ? sky->nptile : sky->nbtile, 4);
  // about 4096 rows per chunk, as the per-row kernel (VISS_BP_ROWS overrides it
  // for launch-shape sweeps); more chunks for small launches
  int64_t rows_per_chunk = 4096;
  if (const char* e = getenv("VISS_BP_ROWS")) {
    const long long v = atoll(e);
    if (v > 0) rows_per_chunk = v;
  }
  int64_t spc = std::max<int64_t>(1, rows_per_chunk / ((int64_t)N * (N - 1) / 2));
  while (spc > 1 && bx * cdiv_s(S, spc) < 4096) spc = (spc + 1) / 2;
  if (cdiv_s(S, spc) > 65535) spc = cdiv_s(S, 65535);
  dim3 grid((unsigned)bx, (unsigned)cdiv_s(S, spc));
  const double herm = (flags & VISS_HERMITIAN) ? 2.0 : 1.0;
#define VISS_BPS(M, NM)                                                                                              \
  do {                                                                                                               \
    if (pairs)                                                                                                       \
      k_backproject_snap_pairs<M, NM><<<grid, 256, 0, s>>>(sky->ppack, sky->pidx, sky->pcone, sky->nptile, rec, cls, \
                                                           yw, S, N, inv_lambda, spc, herm, x);                     \
    else                                                                                                             \
      k_backproject_snap<M, NM><<<grid, 256, 0, s>>>(sky->pix, sky->bcone, sky->nbtile, rec, cls, yw, S, N,          \
                                                     inv_lambda, spc, herm, x);                                      \
  } while (0)
  if (NP == 4) {
    if (mask) VISS_BPS(true, 4); else VISS_BPS(false, 4);
  } else if (NP == 8) {
    if (mask) VISS_BPS(true, 8); else VISS_BPS(false, 8);
  } else if (NP == 10) {
    if (mask) VISS_BPS(true, 10); else VISS_BPS(false, 10);
  } else if (NP == 12) {
    if (mask) VISS_BPS(true, 12); else VISS_BPS(false, 12);
  } else {
    if (mask) VISS_BPS(true, 16); else VISS_BPS(false, 16);
  }
#undef VISS_BPS
  VISS_LAUNCHED(k_backproject_snap);
  return VISS_OK;
}

extern "C" int viss_sky_forward_snapshots(viss_sky* sky, const double* I, const double* sat_pos, int64_t S, int32_t N,
                                          double wavelength, const viss_mask_geometry* geom, int flags, double* V,
                                          void* stream) {
  VISS_REQUIRE(sky != nullptr, "viss_sky_forward_snapshots: sky is NULL");
  int rc = check_device(sky);
  if (rc != VISS_OK) return rc;
  VISS_REQUIRE(S >= 0 && N >= 2 && N <= 16, "viss_sky_forward_snapshots: need S >= 0 and 2 <= N <= 16");
  VISS_REQUIRE(wavelength > 0.0, "viss_sky_forward_snapshots: wavelength must be > 0");
  VISS_REQUIRE((flags & ~(VISS_MASK | VISS_PHASE_TRACK | VISS_HERMITIAN)) == 0,
               "viss_sky_forward_snapshots: unsupported flags 0x%x (mixed precision only)", flags);
  if (S == 0) return VISS_OK;
  VISS_REQUIRE(I && sat_pos && V, "viss_sky_forward_snapshots: NULL argument");
  const bool mask = flags & VISS_MASK;
  if (mask) {
    VISS_REQUIRE(geom != nullptr && geom->R_moon >= 0.0 &&
                     (geom->orbit_radius <= 0.0 || geom->orbit_radius > geom->R_moon),
                 "viss_sky_forward_snapshots: bad geometry");
  }
  hipStream_t s = as_stream(stream);
  rc = sky_set_brightness(sky, I, false, s);
  if (rc != VISS_OK) return rc;
  const int NP = padded_sats(N);  // the kernel's NMAX
  const int64_t SNP = S * NP;
  const size_t b_rec = sizeof(SatRec) * SNP, b_cls = sizeof(SatCls) * SNP;
  rc = sky->ws.reserve(b_rec + b_cls + 256, s);
  if (rc != VISS_OK) return rc;
  SatRec* rec = (SatRec*)sky->ws.p;
  SatCls* cls = (SatCls*)((char*)sky->ws.p + b_rec);
  const double inv_lambda = 1.0 / wavelength;
  k_snap_sats<<<grid_rows(SNP), 256, 0, s>>>(sat_pos, S, N, NP, mask ? geom->R_moon : 0.0,
                                             mask ? geom->orbit_radius : 0.0, mask ? 1 : 0, rec, cls);
  VISS_LAUNCHED(k_snap_sats);
  const int64_t bx = cdiv_s(S, 4);
  // 256 tiles (16,384 pixels, 512 KB of pixel data) per range, fewer when the
  // snapshots alone would not fill the device (VISS_SNAP_TILES overrides the
  // starting length for launch-shape sweeps)
  const bool pairs = sky->use_pairs;  // one lane per antipodal pair, tiles of 64 pairs
  const int64_t ntile = pairs ? sky->nptile : sky->ntile;
  int64_t tpr = 256;
  if (const char* e = getenv("VISS_SNAP_TILES")) {
    const long long v = atoll(e);
    if (v > 0) tpr = v;
  }
  while (tpr > 16 && bx * cdiv_s(ntile, tpr) < 4096) tpr /= 2;
  tpr = std::max(tpr, cdiv_s(ntile, 65535));
  VISS_REQUIRE(bx <= INT32_MAX, "viss_sky_forward_snapshots: too many snapshots");
  dim3 grid((unsigned)bx, (unsigned)cdiv_s(ntile, tpr));
  const int track = (flags & VISS_PHASE_TRACK) ? 1 : 0, herm = (flags & VISS_HERMITIAN) ? 1 : 0;
#define VISS_FS(M, NM)                                                                                               \
  do {                                                                                                               \
    if (pairs)                                                                                                       \
      k_forward_snap_pairs<M, NM><<<grid, 256, 0, s>>>(sky->ppack, sky->pcone, sky->nptile, tpr, rec, cls, S, N,     \
                                                       inv_lambda, track, herm, V);                                  \
    else                                                                                                             \
      k_forward_snap<M, NM><<<grid, 256, 0, s>>>(sky->pix, sky->cone, sky->ntile, tpr, rec, cls, S, N, inv_lambda,   \
                                                 track, herm, V);                                                    \
  } while (0)
  if (NP == 4) {
    if (mask) VISS_FS(true, 4); else VISS_FS(false, 4);
  } else if (NP == 8) {
    if (mask) VISS_FS(true, 8); else VISS_FS(false, 8);
  } else if (NP == 10) {
    if (mask) VISS_FS(true, 10); else VISS_FS(false, 10);
  } else if (NP == 12) {
    if (mask) VISS_FS(true, 12); else VISS_FS(false, 12);
  } else {
    if (mask) VISS_FS(true, 16); else VISS_FS(false, 16);
  }
#undef VISS_FS
  VISS_LAUNCHED(k_forward_snap);
  return VISS_OK;
}
