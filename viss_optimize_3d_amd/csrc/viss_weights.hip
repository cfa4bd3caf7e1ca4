// Density-compensation weights (vissgen3d_noblockage.cu:105-164).
//
// The reference runs ceilAndScale, an unused thrust::sort, countOccurrences
// (every row compares itself with all nr shells and does a float atomicAdd:
// O(K nr)), calculateR, calculateDcf and viss_gamma_trans as six launches with
// a device sync after each.  Here:
//   k_shell_hist : gs_k (FP32, bit-exact with the reference's expression) and
//                  an LDS-privatised uint32 histogram of gs, flushed with one
//                  global atomic per non-empty bin per block      [reads 4 B, writes 4 B per row]
//   k_dcf_table  : the dcf_j table from the exact shell volumes and the counts
//   k_row_weights: the table copied into LDS per block (when it fits), then
//                  dg_k (one FP64 division), wt_k = fmin(dcf[gs_k] dg_k, 1/8)
//                  with the table entry gathered from LDS (else from L2), gs_k
//                  recomputed from bll                [reads 12 B, writes 8-16 B per row]
// All HBM-bound; the sort is dropped (its order is never used).
#include <algorithm>
#include <cstdint>

#include "viss_internal.h"

namespace viss {

#pragma clang fp contract(off)
__device__ inline int32_t shell_index(float bll) {
  const float q = (bll - 0.25f) / 0.5f;
  const float c = ceilf(q);
  // NaN and values outside int32 map to INT32_MIN (out of range; weight 0).
  if (!(c > -2147483648.0f && c < 2147483648.0f)) return INT32_MIN;
  return (int32_t)c;
}

// Shell histogram in LDS: rows outside [1, nr] add into one of 32 spare
// slots past the table (one per LDS bank, chosen by lane, so that inputs with
// many such rows do not serialise on one address), so that every row takes
// the same branch-free add; bad counts those rows other than shell 0 (the
// reference's shell index 0 is in range but uncounted).
constexpr int kHistSpare = 32;

__device__ __forceinline__ void count_lds(int32_t g, int32_t nr, uint32_t* hist, uint32_t& bad) {
  const bool in = (uint32_t)g - 1u < (uint32_t)nr;  // g in [1, nr] (unsigned: no overflow at INT32_MIN)
  bad += (uint32_t)(!in & (g != 0));
  atomicAdd(&hist[in ? g - 1 : nr + (int)(threadIdx.x & (kHistSpare - 1))], 1u);
}

__device__ __forceinline__ void count_global(int32_t g, int32_t nr, uint32_t* mb, uint32_t& bad) {
  const bool in = (uint32_t)g - 1u < (uint32_t)nr;  // g in [1, nr] (unsigned: no overflow at INT32_MIN)
  bad += (uint32_t)(!in & (g != 0));
  if (in) atomicAdd(&mb[g - 1], 1u);
}

__device__ __forceinline__ void count_bad(uint32_t bad, unsigned long long* n_bad) {
  unsigned long long b = bad;
  for (int o = 32; o > 0; o >>= 1) b += __shfl_xor(b, o, 64);
  if ((threadIdx.x & 63) == 0 && b) atomicAdd(n_bad, b);
}

// VEC: bll and gs are 16-B aligned and each thread takes 4 rows per step
// (float4 in, int4 out): the LDS histogram allows only a few blocks per CU,
// so the wider accesses are what keeps enough bytes in flight.
// LDS: the block's histogram is in LDS (nr + kHistSpare counters) and is
// flushed into mb with one global atomic per non-empty bin; otherwise every
// row adds into mb directly.
template <bool VEC, bool LDS>
__global__ __launch_bounds__(1024) void k_shell_hist(const float* __restrict__ bll, int64_t K, int32_t nr,
                                                     int32_t* __restrict__ gs, uint32_t* __restrict__ mb,
                                                     unsigned long long* __restrict__ n_bad) {
  extern __shared__ uint32_t hist[];
  if (LDS) {
    for (int j = threadIdx.x; j < nr + kHistSpare; j += blockDim.x) hist[j] = 0u;
    __syncthreads();
  }
  uint32_t bad = 0;
  const auto count = [&](int32_t g) {
    if (LDS)
      count_lds(g, nr, hist, bad);
    else
      count_global(g, nr, mb, bad);
  };
  const int64_t t0 = blockIdx.x * (int64_t)blockDim.x + threadIdx.x, stride = (int64_t)gridDim.x * blockDim.x;
  int64_t k_tail = 0;
  if (VEC) {
    const int64_t K4 = K / 4;
    const float4* __restrict__ b4 = reinterpret_cast<const float4*>(bll);
    int4* __restrict__ g4 = reinterpret_cast<int4*>(gs);
    auto one = [&](const float4 b, int64_t q) {
      const int4 g = make_int4(shell_index(b.x), shell_index(b.y), shell_index(b.z), shell_index(b.w));
      g4[q] = g;
      count(g.x);
      count(g.y);
      count(g.z);
      count(g.w);
    };
    int64_t q = t0;
    for (; q + 7 * stride < K4; q += 8 * stride) {  // eight loads in flight before the atomics
      float4 b[8];
#pragma unroll
      for (int i = 0; i < 8; ++i) b[i] = b4[q + i * stride];
#pragma unroll
      for (int i = 0; i < 8; ++i) one(b[i], q + i * stride);
    }
    for (; q < K4; q += stride) one(b4[q], q);
    k_tail = K4 * 4;
  }
  for (int64_t k = k_tail + t0; k < K; k += stride) {
    const int32_t g = shell_index(bll[k]);
    gs[k] = g;
    count(g);
  }
  if (n_bad) count_bad(bad, n_bad);
  if (LDS) {
    __syncthreads();
    for (int j = threadIdx.x; j < nr; j += blockDim.x) {
      const uint32_t c = hist[j];
      if (c) atomicAdd(&mb[j], c);
    }
  }
}

// dcf_j, j in [0, nr]: 1 / (4 pi) for j = 0, else the volume of shell jj
// (= j, or j + 1 with the blockage shift) over its count m = mb[jj - 1]; 0 past
// shell nr.  dcf_slot is the mb entry it reads, clamped into the table so that
// the load can be issued before it is known whether it is needed.
__device__ inline int64_t dcf_slot(int64_t j, int32_t nr, int shift) {
  const int64_t jj = shift ? j + 1 : j;
  return (jj >= 1 && jj <= nr) ? jj - 1 : 0;
}

__device__ inline double dcf_value(int64_t j, uint32_t m, int32_t nr, int shift) {
  const int64_t jj = shift ? j + 1 : j;
  const double R = (double)jj / 2.0 + 0.25, Rm = R - 0.5;
  const double v = (2.0 * kPi / 3.0) * (R * R * R - Rm * Rm * Rm) / (double)m;
  return j == 0 ? 1.0 / (4.0 * kPi) : (jj > nr ? 0.0 : v);
}

// The dcf table, j in [0, nr]: the entries the rows gather.
__global__ void k_dcf_table(const uint32_t* __restrict__ mb, int32_t nr, int shift, double* __restrict__ dcf) {
  for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j <= nr; j += (int64_t)gridDim.x * blockDim.x)
    dcf[j] = dcf_value(j, mb[dcf_slot(j, nr, shift)], nr, shift);
}

// dg of one row.  With gamma = asin(x), x = w / b: sin(gamma) = x and
// cos(gamma)^2 = 1 - x^2, so dg = 1.5 sqrt(|1 - 4 x^2| / (1 - x^2))
// = 1.5 sqrt(|b^2 - 4 w^2| / (b^2 - w^2)): one division and one square root,
// NaN for |w| > |b| (as asin).  b^2 is exact (b is FP32), so the denominator
// is 0 exactly when |w| = |b|, where the asin route gives cos(fl(pi/2)) =
// 6.1e-17, not 0: that value is kept, so dg and wt stay finite there as in
// the formula.  The cases the squares do not cover as x = w / b does: b = 0
// (x is infinite or NaN: NaN) and an infinite b (x = 0: 1.5, or NaN for an
// infinite or NaN w).
__device__ inline double row_dg(float b, double wk) {
  constexpr double kDgEdge = 1.5 * 1.7320508075688772 / 6.123233995736766e-17;
  const double bd = (double)b;
  const double bb = bd * bd, ww = wk * wk;
  const double q = bb - ww, a = fabs(bb - 4.0 * ww);
  double dg = q == 0.0 ? kDgEdge : 1.5 * sqrt(a / q);
  dg = bd == 0.0 ? __builtin_nan("") : dg;
  return __builtin_isinf(bd) ? (__builtin_isfinite(wk) ? 1.5 : __builtin_nan("")) : dg;
}

// wt = fmin(dcf[gs] dg, 1/8) for gs in [0, nr], else 0; dcf[gs] gathered from the table.
__device__ inline double row_wt(int32_t s, int32_t nr, double dg, const double* __restrict__ dcf) {
  const bool in = s >= 0 && s <= nr;
  const double d = dcf[in ? s : 0];
  return in ? fmin(d * dg, 0.125) : 0.0;
}

// VEC: bll 8-B, w / wt / dg 16-B aligned; rows go in pairs (one 8-B load of
// bll, one 16-B load of w, one 16-B store of wt and of dg), so the lanes of a
// wavefront load and store contiguous 512-B / 1-KB spans.  The loop is
// software-pipelined: the next two pairs' loads are issued before the current
// two are computed, so the loads stay in flight during the FP64 division and
// square root (a block holding the 128-KB table of 16,000 shells leaves room
// for one block, four wavefronts per SIMD, per CU).  The shell index is
// recomputed from bll (the same expression as k_shell_hist, so the same gs)
// instead of read back.
// LDS_TAB: the block first copies the dcf table (written by k_dcf_table) into
// LDS, 16 loads in flight per thread, and the rows gather from there.  Rows hit
// shells at random, so gathering from the global table costs a cache-line
// request per row.
template <bool VEC, bool LDS_TAB>
__global__ __launch_bounds__(1024) void k_row_weights(const float* __restrict__ bll, const double* __restrict__ w,
                                                      int64_t K, int32_t nr, const double* __restrict__ dcf_g,
                                                      double* __restrict__ dg_out, double* __restrict__ wt) {
  extern __shared__ double tab_s[];
  const double* __restrict__ dcf = dcf_g;
  const int64_t t0 = blockIdx.x * (int64_t)blockDim.x + threadIdx.x, stride = (int64_t)gridDim.x * blockDim.x;
  constexpr int F = 2;  // pairs per pipeline stage
  const int64_t K2 = VEC ? K / 2 : 0;
  const float2* __restrict__ b2 = reinterpret_cast<const float2*>(bll);
  const double2* __restrict__ w2 = reinterpret_cast<const double2*>(w);
  int64_t q = t0;
  float2 b[F];
  double2 wk[F];
  if (VEC && q + (F - 1) * stride < K2) {  // the first stage's loads go out before the table copy
#pragma unroll
    for (int i = 0; i < F; ++i) {
      b[i] = b2[q + i * stride];
      wk[i] = w2[q + i * stride];
    }
  }
  if (LDS_TAB) {
    constexpr int B = 16;
    for (int j0 = 0; j0 <= nr; j0 += B * (int)blockDim.x) {
      double d[B];
#pragma unroll
      for (int i = 0; i < B; ++i) {
        const int j = j0 + i * (int)blockDim.x + (int)threadIdx.x;
        d[i] = dcf_g[j <= nr ? j : 0];
      }
#pragma unroll
      for (int i = 0; i < B; ++i) {
        const int j = j0 + i * (int)blockDim.x + (int)threadIdx.x;
        if (j <= nr) tab_s[j] = d[i];
      }
    }
    __syncthreads();
    dcf = tab_s;
  }
  int64_t k_tail = 0;
  if (VEC) {
    double2* __restrict__ t2 = reinterpret_cast<double2*>(wt);
    double2* __restrict__ g2 = reinterpret_cast<double2*>(dg_out);
    const auto pair = [&](const float2 bb, const double2 ww, int64_t qq) {
      const double d0 = row_dg(bb.x, ww.x), d1 = row_dg(bb.y, ww.y);
      t2[qq] = make_double2(row_wt(shell_index(bb.x), nr, d0, dcf), row_wt(shell_index(bb.y), nr, d1, dcf));
      if (dg_out) g2[qq] = make_double2(d0, d1);
    };
    for (; q + (F - 1) * stride < K2; q += F * stride) {
      const int64_t qn = q + F * stride;
      const bool more = qn + (F - 1) * stride < K2;
      float2 nb[F];
      double2 nw[F];
      if (more) {
#pragma unroll
        for (int i = 0; i < F; ++i) {
          nb[i] = b2[qn + i * stride];
          nw[i] = w2[qn + i * stride];
        }
      }
#pragma unroll
      for (int i = 0; i < F; ++i) pair(b[i], wk[i], q + i * stride);
      if (more) {
#pragma unroll
        for (int i = 0; i < F; ++i) {
          b[i] = nb[i];
          wk[i] = nw[i];
        }
      }
    }
    for (; q < K2; q += stride) pair(b2[q], w2[q], q);
    k_tail = K2 * 2;
  }
  for (int64_t k = k_tail + t0; k < K; k += stride) {
    const float bk = bll[k];
    const double dgk = row_dg(bk, w[k]);
    if (dg_out) dg_out[k] = dgk;
    wt[k] = row_wt(shell_index(bk), nr, dgk, dcf);
  }
}

// ---- per-group weights of a batch: one block per chunk of one group's rows
// (RowChunk: rows [k0, k1) of group g; pad != 0 marks the group's first chunk).
// Each group has its own histogram mb + g nr and table dcf + g (nr + 1).
__global__ __launch_bounds__(1024) void k_batch_shell_hist(const float* __restrict__ bll,
                                                           const RowChunk* __restrict__ ch, int32_t nr,
                                                           int32_t* __restrict__ gs, uint32_t* __restrict__ mb,
                                                           unsigned long long* __restrict__ n_bad) {
  extern __shared__ uint32_t hist[];
  const RowChunk c = ch[blockIdx.x];
  for (int j = threadIdx.x; j < nr + kHistSpare; j += blockDim.x) hist[j] = 0u;
  __syncthreads();
  uint32_t bad = 0;
  for (int64_t k = c.k0 + threadIdx.x; k < c.k1; k += blockDim.x) {
    const int32_t g = shell_index(bll[k]);
    gs[k] = g;
    count_lds(g, nr, hist, bad);
  }
  if (n_bad) count_bad(bad, n_bad);
  __syncthreads();
  uint32_t* __restrict__ mbg = mb + c.g * (int64_t)nr;
  for (int j = threadIdx.x; j < nr; j += blockDim.x) {
    const uint32_t cnt = hist[j];
    if (cnt) atomicAdd(&mbg[j], cnt);
  }
}

// The group's dcf table in LDS (its first chunk's block writes it out), then
// the chunk's rows as k_row_weights.
__global__ __launch_bounds__(1024) void k_batch_row_weights(const float* __restrict__ bll,
                                                            const double* __restrict__ w,
                                                            const RowChunk* __restrict__ ch, int32_t nr,
                                                            const uint32_t* __restrict__ mb, int shift,
                                                            double* __restrict__ dcf, double* __restrict__ dg_out,
                                                            double* __restrict__ wt) {
  extern __shared__ double tab_s[];
  const RowChunk c = ch[blockIdx.x];
  const uint32_t* __restrict__ mbg = mb + c.g * (int64_t)nr;
  double* __restrict__ dcfg = dcf + c.g * ((int64_t)nr + 1);
  for (int j = threadIdx.x; j <= nr; j += blockDim.x) {
    const double d = dcf_value(j, mbg[dcf_slot(j, nr, shift)], nr, shift);
    tab_s[j] = d;
    if (c.pad) dcfg[j] = d;
  }
  __syncthreads();
  for (int64_t k = c.k0 + threadIdx.x; k < c.k1; k += blockDim.x) {
    const float b = bll[k];
    const double dgk = row_dg(b, w[k]);
    if (dg_out) dg_out[k] = dgk;
    wt[k] = row_wt(shell_index(b), nr, dgk, tab_s);
  }
}

#pragma clang fp contract(on)

// Per-group weights of a batch (viss_batch_weights): two launches over the
// weight chunks, arguments checked by the caller.
int batch_weights(const float* bll, const double* w, const RowChunk* wch, int64_t nwch, int32_t nr, int flags,
                  int32_t* gs, uint32_t* mb, double* dcf, double* dg, double* wt, int64_t* n_out_of_range,
                  hipStream_t s) {
  int dev = 0, lds_max = 0;
  VISS_HIP(hipGetDevice(&dev));
  VISS_HIP(hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
  const size_t tab = ((size_t)nr + 1) * sizeof(double);
  if (tab > (size_t)lds_max) {
    set_error("viss_batch_weights: %d shells need a %zu-B table per group, above the %d B of LDS of a block", nr,
              tab, lds_max);
    return VISS_ERR_SIZE;
  }
  VISS_REQUIRE(nwch >= 1 && nwch <= INT32_MAX, "viss_batch_weights: %lld weight chunks", (long long)nwch);
  k_batch_shell_hist<<<(unsigned)nwch, 1024, ((size_t)nr + kHistSpare) * sizeof(uint32_t), s>>>(
      bll, wch, nr, gs, mb, (unsigned long long*)n_out_of_range);
  VISS_LAUNCHED(k_batch_shell_hist);
  const int shift = (flags & VISS_DCF_BLOCKAGE_SHIFT) ? 1 : 0;
  k_batch_row_weights<<<(unsigned)nwch, 1024, tab, s>>>(bll, w, wch, nr, mb, shift, dcf, dg, wt);
  VISS_LAUNCHED(k_batch_row_weights);
  return VISS_OK;
}

}  // namespace viss

using namespace viss;

extern "C" int viss_weights(const float* bll, const double* w, int64_t K, int32_t nr, int flags, int32_t* gs,
                            uint32_t* mb, double* dcf, double* dg, double* wt, int64_t* n_out_of_range,
                            void* stream) {
  VISS_REQUIRE(K >= 0, "viss_weights: negative K");
  VISS_REQUIRE(nr >= 1 && nr <= (1 << 26), "viss_weights: nr %d outside [1, 2^26]", nr);
  VISS_REQUIRE(mb && dcf, "viss_weights: mb/dcf must not be NULL");
  VISS_REQUIRE(K == 0 || (bll && w && gs && wt), "viss_weights: bll/w/gs/wt must not be NULL");
  VISS_REQUIRE((flags & ~(VISS_DCF_BLOCKAGE_SHIFT)) == 0, "viss_weights: unsupported flags 0x%x", flags);
  hipStream_t s = as_stream(stream);
  int dev = 0, lds_max = 0, ncu = 0;
  VISS_HIP(hipGetDevice(&dev));
  VISS_HIP(hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
  VISS_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  if (K > 0) {
    const size_t lds = ((size_t)nr + kHistSpare) * sizeof(uint32_t);
    const bool use_lds = lds <= (size_t)lds_max;
    // half as many blocks per CU as its LDS holds tables (1..8): every block
    // ends by flushing its nr counters with global atomics on the same
    // addresses, and with two 64-KB tables per CU that flush took 12 of 75 us
    // at 4e7 rows (one per CU: 65 us)
    int lds_cu = 0;
    VISS_HIP(hipDeviceGetAttribute(&lds_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev));
    int64_t per_cu = use_lds && lds > 0 ? (int64_t)lds_cu / (2 * (int64_t)lds) : 8;
    per_cu = per_cu < 1 ? 1 : (per_cu > 8 ? 8 : per_cu);
    // 1024-thread blocks: the per-CU block count is small (the LDS table),
    // so large blocks keep enough loads in flight
    const int hb = use_lds ? 1024 : 256;
    int64_t blocks = (K + hb * 16 - 1) / (hb * 16);
    const int64_t cap = (int64_t)(ncu > 0 ? ncu : 256) * per_cu;
    if (blocks > cap) blocks = cap;
    if (blocks < 1) blocks = 1;
    const bool vec = ((uintptr_t)bll % 16 == 0) && ((uintptr_t)gs % 16 == 0);
    unsigned long long* nb = (unsigned long long*)n_out_of_range;
#define VISS_SH(V, L) k_shell_hist<V, L><<<(unsigned)blocks, hb, L ? lds : 0, s>>>(bll, K, nr, gs, mb, nb)
    if (vec) {
      if (use_lds) VISS_SH(true, true); else VISS_SH(true, false);
    } else {
      if (use_lds) VISS_SH(false, true); else VISS_SH(false, false);
    }
#undef VISS_SH
    VISS_LAUNCHED(k_shell_hist);
  }
  const int shift = (flags & VISS_DCF_BLOCKAGE_SHIFT) ? 1 : 0;
  const size_t tab = ((size_t)nr + 1) * sizeof(double);
  // the table (also for K = 0: the table of an empty histogram); the row
  // kernel copies it into LDS when it fits and the rows are many enough to pay
  // for one copy per block, else the rows gather it from L2.  Measured with
  // 1024-thread blocks: 4e7 rows over 16,012 shells, 171 us with a table in
  // LDS against 330 us from L2; 1.8e6 rows over 11,082 shells (162 rows per
  // shell), 13.5 us from L2 against 15.6 us from LDS.
  {
    int64_t g = ((int64_t)nr + 1 + 255) / 256;
    if (g > 4096) g = 4096;
    k_dcf_table<<<(unsigned)g, 256, 0, s>>>(mb, nr, shift, dcf);
    VISS_LAUNCHED(k_dcf_table);
  }
  const bool lds_tab = K > 0 && tab <= (size_t)lds_max && K >= 512 * ((int64_t)nr + 1);
  if (K > 0) {  // row weights, gathering the table
    const bool vec = ((uintptr_t)bll % 8 == 0) && ((uintptr_t)w % 16 == 0) && ((uintptr_t)wt % 16 == 0) &&
                     (dg == nullptr || (uintptr_t)dg % 16 == 0);
    // LDS path: 1024-thread blocks, as many per CU as the table copies allow;
    // else 256-thread blocks
    const int block = lds_tab ? 1024 : 256;
    int lds_cu = 0;
    VISS_HIP(hipDeviceGetAttribute(&lds_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev));
    int64_t per_cu = lds_tab ? std::max<int64_t>(1, std::min<int64_t>(2, (int64_t)lds_cu / (int64_t)tab)) : 64;
    int64_t g = ((vec ? K / 4 : K) + block - 1) / block;  // two row pairs per thread and stage
    const int64_t cap = (int64_t)(ncu > 0 ? ncu : 256) * per_cu;
    if (g > cap) g = cap;
    if (g < 1) g = 1;
#define VISS_RW(V, L) \
  k_row_weights<V, L><<<(unsigned)g, block, L ? tab : 0, s>>>(bll, w, K, nr, dcf, dg, wt)
    if (vec) {
      if (lds_tab) VISS_RW(true, true); else VISS_RW(true, false);
    } else {
      if (lds_tab) VISS_RW(false, true); else VISS_RW(false, false);
    }
#undef VISS_RW
    VISS_LAUNCHED(k_row_weights);
  }
  return VISS_OK;
}
