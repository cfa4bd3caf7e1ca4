// Image-level utilities after the hot path: the multi-day sum the reference's
// notebooks do on the CPU (healpix3d_blockage.ipynb: total += nan_to_num(C_i)
// over 450 days; healpix3d_noblockage.ipynb the same without nan_to_num) and
// the NEST <-> RING reorder used to display a map in the other scheme.
// Both are HBM-bound elementwise passes (sum: 16 B read + 8 B written per
// pixel; reorder: 8 B read + 8 B written per pixel).
#include "viss_internal.h"

namespace viss {

__global__ void k_image_accumulate(const double* __restrict__ img, int64_t P, int nan_to_zero,
                                   double* __restrict__ total) {
  for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < P; p += (int64_t)gridDim.x * blockDim.x) {
    double v = img[p];
    if (nan_to_zero && v != v) v = 0.0;
    total[p] += v;
  }
}

__device__ inline uint32_t compact_bits(uint32_t v) {
  v &= 0x55555555u;
  v = (v | (v >> 1)) & 0x33333333u;
  v = (v | (v >> 2)) & 0x0f0f0f0fu;
  v = (v | (v >> 4)) & 0x00ff00ffu;
  v = (v | (v >> 8)) & 0x0000ffffu;
  return v;
}

// NEST index -> RING index (HEALPix face layout; tables jrll/jpll as in
// pix2ang.cu:183-184, 0-based).
__device__ inline int32_t nest2ring(int32_t ip, int32_t nside, int32_t order) {
  constexpr int8_t jrll[12] = {2, 2, 2, 2, 3, 3, 3, 3, 4, 4, 4, 4};
  constexpr int8_t jpll[12] = {1, 3, 5, 7, 0, 2, 4, 6, 1, 3, 5, 7};
  const int32_t face = ip >> (2 * order), ipf = ip & ((1 << (2 * order)) - 1);
  const int32_t ix = (int32_t)compact_bits((uint32_t)ipf), iy = (int32_t)compact_bits((uint32_t)ipf >> 1);
  const int32_t jr = jrll[face] * nside - (ix + iy) - 1;
  int32_t nr, kshift = 0, start;
  const int32_t npix = 12 * nside * nside, ncap = 2 * nside * (nside - 1);
  if (jr < nside) {
    nr = jr;
    start = 2 * jr * (jr - 1);
  } else if (jr <= 3 * nside) {
    nr = nside;
    kshift = (jr - nside) & 1;
    start = ncap + (jr - nside) * 4 * nside;
  } else {
    nr = 4 * nside - jr;
    start = npix - 2 * nr * (nr + 1);
  }
  int32_t jp = (jpll[face] * nr + (ix - iy) + 1 + kshift) / 2;
  if (jp > 4 * nside) jp -= 4 * nside;
  if (jp < 1) jp += 4 * nside;
  return start + jp - 1;
}

__global__ void k_reorder(const double* __restrict__ in, int32_t nside, int32_t order, int32_t npix, int to_ring,
                          double* __restrict__ out) {
  for (int32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < npix; p += gridDim.x * blockDim.x) {
    const int32_t r = nest2ring(p, nside, order);
    if (to_ring)
      out[r] = in[p];
    else
      out[p] = in[r];
  }
}

static unsigned grid_img(int64_t n) {
  int64_t g = (n + 255) / 256;
  if (g > 16384) g = 16384;
  return (unsigned)(g < 1 ? 1 : g);
}

}  // namespace viss

using namespace viss;

extern "C" int viss_image_accumulate(const double* img, int64_t P, int nan_to_zero, double* total, void* stream) {
  VISS_REQUIRE(P >= 0, "viss_image_accumulate: negative P");
  if (P == 0) return VISS_OK;
  VISS_REQUIRE(img && total, "viss_image_accumulate: NULL argument");
  k_image_accumulate<<<grid_img(P), 256, 0, as_stream(stream)>>>(img, P, nan_to_zero ? 1 : 0, total);
  VISS_LAUNCHED(k_image_accumulate);
  return VISS_OK;
}

extern "C" int viss_reorder(const double* in, int64_t nside, int to_ring, double* out, void* stream) {
  VISS_REQUIRE(nside >= 1 && nside <= 8192 && (nside & (nside - 1)) == 0,
               "viss_reorder: nside %lld must be a power of two in [1, 8192]", (long long)nside);
  VISS_REQUIRE(in && out && in != out, "viss_reorder: in/out must be distinct non-NULL buffers");
  int32_t order = 0;
  while ((1ll << order) < nside) ++order;
  const int64_t npix = 12 * nside * nside;
  k_reorder<<<grid_img(npix), 256, 0, as_stream(stream)>>>(in, (int32_t)nside, order, (int32_t)npix, to_ring ? 1 : 0,
                                                           out);
  VISS_LAUNCHED(k_reorder);
  return VISS_OK;
}
