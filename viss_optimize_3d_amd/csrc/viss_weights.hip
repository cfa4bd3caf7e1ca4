// Density-compensation weights (vissgen3d_noblockage.cu:105-164).
//
// The reference runs ceilAndScale, an unused thrust::sort, countOccurrences
// (every row compares itself with all nr shells and does a float atomicAdd:
// O(K nr)), calculateR, calculateDcf and viss_gamma_trans as six launches with
// a device sync after each.  Here:
//   k_shell_hist : gs_k (FP32, bit-exact with the reference's expression) and
//                  an LDS-privatised uint32 histogram of gs, flushed with one
//                  global atomic per non-empty bin per block      [reads 4 B, writes 4 B per row]
//   k_dcf        : dcf_j from the exact shell volume               [O(nr)]
//   k_row_weights: dg_k, wt_k = fmin(dcf[gs_k] dg_k, 1/8)           [reads 16 B, writes 8-16 B per row]
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
    for (int64_t q = t0; q < K4; q += stride) {
      const float4 b = b4[q];
      const int4 g = make_int4(shell_index(b.x), shell_index(b.y), shell_index(b.z), shell_index(b.w));
      g4[q] = g;
      count_shell(g.x, nr, hist, mb, use_lds, bad);
      count_shell(g.y, nr, hist, mb, use_lds, bad);
      count_shell(g.z, nr, hist, mb, use_lds, bad);
      count_shell(g.w, nr, hist, mb, use_lds, bad);
    }
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

__global__ void k_dcf(const uint32_t* __restrict__ mb, int32_t nr, int shift, double* __restrict__ dcf) {
  for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j <= nr; j += (int64_t)gridDim.x * blockDim.x) {
    if (j == 0) {
      dcf[0] = 1.0 / (4.0 * kPi);
      continue;
    }
    const int64_t jj = shift ? j + 1 : j;
    if (jj > nr) {
      dcf[j] = 0.0;
      continue;
    }
    const double R = (double)jj / 2.0 + 0.25, Rm = R - 0.5;
    dcf[j] = (2.0 * kPi / 3.0) * (R * R * R - Rm * Rm * Rm) / (double)mb[jj - 1];
  }
}

__global__ void k_row_weights(const float* __restrict__ bll, const double* __restrict__ w, int64_t K, int32_t nr,
                              const int32_t* __restrict__ gs, const double* __restrict__ dcf,
                              double* __restrict__ dg_out, double* __restrict__ wt) {
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < K; k += (int64_t)gridDim.x * blockDim.x) {
    const double g = asin(w[k] / (double)bll[k]);
    double sg, cg;
    sincos(g, &sg, &cg);
    const double dg = 1.5 * sqrt(fabs(1.0 - 4.0 * sg * sg)) / fabs(cg);
    if (dg_out) dg_out[k] = dg;
    const int32_t s = gs[k];
    wt[k] = (s < 0 || s > nr) ? 0.0 : fmin(dcf[s] * dg, 0.125);
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
  k_dcf<<<(unsigned)((nr + 1 + 255) / 256 > 4096 ? 4096 : (nr + 1 + 255) / 256), 256, 0, s>>>(
      mb, nr, (flags & VISS_DCF_BLOCKAGE_SHIFT) ? 1 : 0, dcf);
  VISS_LAUNCHED(k_dcf);
  if (K > 0) {
    int64_t g = (K + 255) / 256;
    if (g > 65536) g = 65536;
    k_row_weights<<<(unsigned)g, 256, 0, s>>>(bll, w, K, nr, gs, dcf, dg, wt);
    VISS_LAUNCHED(k_row_weights);
  }
  return VISS_OK;
}
