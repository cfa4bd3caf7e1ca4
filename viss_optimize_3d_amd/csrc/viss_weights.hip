// Density-compensation weights (vissgen3d_noblockage.cu:105-164).
//
// The reference runs ceilAndScale, an unused thrust::sort, countOccurrences
// (every row compares itself with all nr shells and does a float atomicAdd:
// O(K nr)), calculateR, calculateDcf and viss_gamma_trans as six launches with
// a device sync after each.  Here:
//   k_shell_hist : gs_k (FP32, bit-exact with the reference's expression) and
//                  an LDS-privatised uint32 histogram of gs, flushed with one
//                  global atomic per non-empty bin per block      [reads 4 B, writes 4 B per row]
//   k_row_weights: dg_k, wt_k = fmin(dcf[gs_k] dg_k, 1/8), gs_k recomputed from bll,
//                  and the dcf_j table from the exact shell volumes  [reads 12 B, writes 8-16 B per row]
// All HBM-bound; the sort is dropped (its order is never used).
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

__device__ inline void count_shell(int32_t g, int32_t nr, uint32_t* hist, uint32_t* mb, int use_lds,
                                   unsigned long long& bad) {
  if (g >= 1 && g <= nr) {
    if (use_lds)
      atomicAdd(&hist[g - 1], 1u);
    else
      atomicAdd(&mb[g - 1], 1u);
  } else if (g != 0) {
    ++bad;
  }
}

// VEC: bll and gs are 16-B aligned and each thread takes 4 rows per step
// (float4 in, int4 out): the LDS histogram allows only a few blocks per CU,
// so the wider accesses are what keeps enough bytes in flight.
template <bool VEC>
__global__ void k_shell_hist(const float* __restrict__ bll, int64_t K, int32_t nr, int32_t* __restrict__ gs,
                             uint32_t* __restrict__ mb, unsigned long long* __restrict__ n_bad, int use_lds) {
  extern __shared__ uint32_t hist[];
  if (use_lds) {
    for (int j = threadIdx.x; j < nr; j += blockDim.x) hist[j] = 0u;
    __syncthreads();
  }
  unsigned long long bad = 0;
  const int64_t t0 = blockIdx.x * (int64_t)blockDim.x + threadIdx.x, stride = (int64_t)gridDim.x * blockDim.x;
  int64_t k_tail = 0;
  if (VEC) {
    const int64_t K4 = K / 4;
    const float4* __restrict__ b4 = reinterpret_cast<const float4*>(bll);
    int4* __restrict__ g4 = reinterpret_cast<int4*>(gs);
    auto one = [&](const float4 b, int64_t q) {
      const int4 g = make_int4(shell_index(b.x), shell_index(b.y), shell_index(b.z), shell_index(b.w));
      g4[q] = g;
      count_shell(g.x, nr, hist, mb, use_lds, bad);
      count_shell(g.y, nr, hist, mb, use_lds, bad);
      count_shell(g.z, nr, hist, mb, use_lds, bad);
      count_shell(g.w, nr, hist, mb, use_lds, bad);
    };
    int64_t q = t0;
    for (; q + 3 * stride < K4; q += 4 * stride) {  // four loads in flight before the atomics
      float4 b[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) b[i] = b4[q + i * stride];
#pragma unroll
      for (int i = 0; i < 4; ++i) one(b[i], q + i * stride);
    }
    for (; q < K4; q += stride) one(b4[q], q);
    k_tail = K4 * 4;
  }
  for (int64_t k = k_tail + t0; k < K; k += stride) {
    const int32_t g = shell_index(bll[k]);
    gs[k] = g;
    count_shell(g, nr, hist, mb, use_lds, bad);
  }
  if (n_bad) {
    for (int o = 32; o > 0; o >>= 1) bad += __shfl_xor(bad, o, 64);
    if ((threadIdx.x & 63) == 0 && bad) atomicAdd(n_bad, bad);
  }
  if (use_lds) {
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

// One row's dg and wt from bll (b), w and the count of its dcf entry (m).
// gamma = asin(x), x = w / bll: sin(gamma) = x and cos(gamma)^2 = 1 - x^2 on
// [-1, 1], so dg = 1.5 sqrt(|1 - 4 x^2| / (1 - x^2)), NaN for |x| > 1 (as
// asin), and no asin or sincos is evaluated.  At |x| = 1 the asin route gives
// cos(fl(pi/2)) = 6.1e-17, not 0: that value is kept, so dg and wt stay
// finite there as in the formula.  The row's dcf entry is evaluated as the
// table's is (dcf_value), so it is the same value.
__device__ inline void row_weight(float b, double wk, int32_t s, uint32_t m, int32_t nr, int shift, double& dg,
                                  double& wt) {
  const double x = wk / (double)b;
  const double q = (1.0 - x) * (1.0 + x), a = fabs(1.0 - 4.0 * x * x);
  // q = 0 only at x = +-1, where |1 - 4 x^2| = 3
  constexpr double kDgEdge = 1.5 * 1.7320508075688772 / 6.123233995736766e-17;
  dg = q == 0.0 ? kDgEdge : 1.5 * sqrt(a / q);
  const double d = dcf_value(s, m, nr, shift);
  wt = (s < 0 || s > nr) ? 0.0 : fmin(d * dg, 0.125);
}

// The same launch writes the dcf table (entries j = thread index + multiples
// of the grid size).  VEC: bll 8-B and w / wt / dg 16-B aligned; each thread
// takes 2 rows per step.  The shell index is recomputed from bll (the same
// expression as k_shell_hist, so the same gs) instead of read back, and both
// rows' mb loads are issued before the arithmetic.
template <bool VEC>
__global__ void k_row_weights(const float* __restrict__ bll, const double* __restrict__ w, int64_t K, int32_t nr,
                              const uint32_t* __restrict__ mb, int shift, double* __restrict__ dcf,
                              double* __restrict__ dg_out, double* __restrict__ wt) {
  const int64_t t0 = blockIdx.x * (int64_t)blockDim.x + threadIdx.x, stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t j = t0; j <= nr; j += stride) dcf[j] = dcf_value(j, mb[dcf_slot(j, nr, shift)], nr, shift);
  int64_t k_tail = 0;
  if (VEC) {
    const int64_t K2 = K / 2;
    for (int64_t q = t0; q < K2; q += stride) {
      const float2 b = reinterpret_cast<const float2*>(bll)[q];
      const double2 wk = reinterpret_cast<const double2*>(w)[q];
      const int32_t s0 = shell_index(b.x), s1 = shell_index(b.y);
      const uint32_t m0 = mb[dcf_slot(s0, nr, shift)], m1 = mb[dcf_slot(s1, nr, shift)];
      double2 dg, wo;
      row_weight(b.x, wk.x, s0, m0, nr, shift, dg.x, wo.x);
      row_weight(b.y, wk.y, s1, m1, nr, shift, dg.y, wo.y);
      if (dg_out) reinterpret_cast<double2*>(dg_out)[q] = dg;
      reinterpret_cast<double2*>(wt)[q] = wo;
    }
    k_tail = K2 * 2;
  }
  for (int64_t k = k_tail + t0; k < K; k += stride) {
    const float b = bll[k];
    const int32_t sk = shell_index(b);
    double dg, wo;
    row_weight(b, w[k], sk, mb[dcf_slot(sk, nr, shift)], nr, shift, dg, wo);
    if (dg_out) dg_out[k] = dg;
    wt[k] = wo;
  }
}
#pragma clang fp contract(on)

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
    const size_t lds = (size_t)nr * sizeof(uint32_t);
    const int use_lds = lds <= (size_t)lds_max ? 1 : 0;
    // as many resident blocks as the CUs' LDS holds (2..8 per CU): each block
    // flushes its nr counters once, so no more than that
    int lds_cu = 0;
    VISS_HIP(hipDeviceGetAttribute(&lds_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev));
    int64_t per_cu = use_lds && lds > 0 ? (int64_t)lds_cu / (int64_t)lds : 8;
    per_cu = per_cu < 2 ? 2 : (per_cu > 8 ? 8 : per_cu);
    int64_t blocks = (K + 256 * 16 - 1) / (256 * 16);
    const int64_t cap = (int64_t)(ncu > 0 ? ncu : 256) * per_cu;
    if (blocks > cap) blocks = cap;
    if (blocks < 1) blocks = 1;
    const bool vec = ((uintptr_t)bll % 16 == 0) && ((uintptr_t)gs % 16 == 0);
    if (vec)
      k_shell_hist<true><<<(unsigned)blocks, 256, use_lds ? lds : 0, s>>>(bll, K, nr, gs, mb,
                                                                         (unsigned long long*)n_out_of_range, use_lds);
    else
      k_shell_hist<false><<<(unsigned)blocks, 256, use_lds ? lds : 0, s>>>(bll, K, nr, gs, mb,
                                                                          (unsigned long long*)n_out_of_range, use_lds);
    VISS_LAUNCHED(k_shell_hist);
  }
  {  // row weights and the dcf table (also for K = 0: the table of an empty histogram)
    const int shift = (flags & VISS_DCF_BLOCKAGE_SHIFT) ? 1 : 0;
    const bool vec = ((uintptr_t)bll % 8 == 0) && ((uintptr_t)w % 16 == 0) && ((uintptr_t)wt % 16 == 0) &&
                     (dg == nullptr || (uintptr_t)dg % 16 == 0);
    int64_t g = ((vec ? K / 2 : K) + 255) / 256;
    if (g > 65536) g = 65536;
    if (g < 1) g = 1;
    if (vec)
      k_row_weights<true><<<(unsigned)g, 256, 0, s>>>(bll, w, K, nr, mb, shift, dcf, dg, wt);
    else
      k_row_weights<false><<<(unsigned)g, 256, 0, s>>>(bll, w, K, nr, mb, shift, dcf, dg, wt);
    VISS_LAUNCHED(k_row_weights);
  }
  return VISS_OK;
}
