// HEALPix pix2ang (NEST and RING) and the direction pre-pass.
//
// Replaces mk_pix2xy + f_pix2and_nest (pix2ang.cu:61-160).  The reference
// kernel writes twelve [npix] scratch arrays to HBM, indexes its 1-based
// MATLAB tables with +1 (so every theta is NaN) and has no bounds guard.  Here
// each thread holds its pixel's integers in registers and writes only the
// outputs (16 B/pixel for theta, phi; +8 with ring/iphi): an HBM-store-bound
// elementwise kernel.  theta is constant on a ring, so acos runs once per ring
// (4 nside - 1 of them) and the other pixels copy it, L2-resident.  The Morton
// de-interleave is done with bit masks, so the 1024-entry pix2x/pix2y lookup
// table is not needed.
//
// Bit-exactness: z = (3n^2 - nr^2) / (3n^2) and phi = (pi/2) t / nr are single
// correctly rounded FP64 operations on exact integers, identical to the host
// oracle.  theta = acos(z) uses the device library's acos (within 1 ulp of
// the host's).
#include "viss_internal.h"

namespace viss {

// All indices fit in 32 bits for nside <= 8192 (npix < 2^30), so the integer
// work uses 32-bit operations; 64-bit division would cost several times more.
__device__ inline uint32_t compact_even_bits(uint32_t v) {
  v &= 0x55555555u;
  v = (v | (v >> 1)) & 0x33333333u;
  v = (v | (v >> 2)) & 0x0f0f0f0fu;
  v = (v | (v >> 4)) & 0x00ff00ffu;
  v = (v | (v >> 8)) & 0x0000ffffu;
  return v;
}

// floor(sqrt(x)) for 0 <= x < 2^31: the FP32 root (x rounded to 24 bits, a
// root within an ulp) is within 1 of the answer, and the two loops correct it.
__device__ inline int32_t isqrt32(int32_t x) {
  int32_t s = (int32_t)__builtin_amdgcn_sqrtf((float)x);
  while ((int64_t)s * s > x) --s;
  while ((int64_t)(s + 1) * (s + 1) <= x) ++s;
  return s;
}

struct RingPos {
  int32_t jr, jp, nr, kshift;
};

// NEST: face / in-face (ix, iy) -> ring jr and in-ring jp (pix2ang.cu:99-157,
// 0-based tables jrll/jpll of pix2ang.cu:183-184).  nside = 2^order.
// The tables as arithmetic: jrll = {2,2,2,2, 3,3,3,3, 4,4,4,4} = 2 + face / 4
// and jpll = {1,3,5,7, 0,2,4,6, 1,3,5,7} = 2 (face mod 4) + 1 - (face / 4 mod 2).
__device__ inline RingPos nest_pos(int32_t ip, int32_t nside, int32_t order) {
  const int32_t face = ip >> (2 * order), ipf = ip & ((1 << (2 * order)) - 1);
  const int32_t jrll = 2 + (face >> 2), jpll = 2 * (face & 3) + 1 - ((face >> 2) & 1);
  const int32_t ix = (int32_t)compact_even_bits((uint32_t)ipf), iy = (int32_t)compact_even_bits((uint32_t)ipf >> 1);
  RingPos r;
  r.jr = jrll * nside - (ix + iy) - 1;
  if (r.jr < nside) {
    r.nr = r.jr;
    r.kshift = 0;
  } else if (r.jr <= 3 * nside) {
    r.nr = nside;
    r.kshift = (r.jr - nside) & 1;
  } else {
    r.nr = 4 * nside - r.jr;
    r.kshift = 0;
  }
  int32_t jp = (jpll * r.nr + (ix - iy) + 1 + r.kshift) / 2;
  if (jp > 4 * nside) jp -= 4 * nside;
  if (jp < 1) jp += 4 * nside;
  r.jp = jp;
  return r;
}

// RING: Gorski et al. 2005 sec. 4 (polar caps / equatorial belt).  lg4 =
// log2(4 nside) when nside is a power of two (the belt's row is then a
// shift), else -1.
__device__ inline RingPos ring_pos(int32_t p, int32_t nside, int32_t lg4) {
  const int32_t npix = 12 * nside * nside, ncap = 2 * nside * (nside - 1), nl4 = 4 * nside;
  RingPos r;
  r.kshift = 0;
  if (p < ncap) {
    const int32_t i = (1 + isqrt32(1 + 2 * p)) / 2;
    r.jr = i;
    r.nr = i;
    r.jp = p + 1 - 2 * i * (i - 1);
  } else if (p < npix - ncap) {
    const uint32_t q = (uint32_t)(p - ncap);
    const uint32_t row = lg4 >= 0 ? q >> lg4 : q / (uint32_t)nl4;
    r.jr = (int32_t)row + nside;
    r.jp = (int32_t)(q - row * (uint32_t)nl4) + 1;
    r.nr = nside;
    r.kshift = (r.jr - nside) & 1;
  } else {
    const int32_t q = npix - p;
    const int32_t i = (1 + isqrt32(2 * q - 1)) / 2;
    r.jr = nl4 - i;
    r.nr = i;
    r.jp = 4 * i + 1 - (q - 2 * i * (i - 1));
  }
  return r;
}

// Ring geometry of ring jr (1 .. 4 nside - 1): pixels in the ring and the
// phase shift of its first pixel.
__device__ inline void ring_shape(int32_t jr, int32_t nside, int32_t& nr, int32_t& kshift) {
  if (jr < nside) {
    nr = jr;
    kshift = 0;
  } else if (jr <= 3 * nside) {
    nr = nside;
    kshift = (jr - nside) & 1;
  } else {
    nr = 4 * nside - jr;
    kshift = 0;
  }
}

// Spread the low 16 bits of v to the even bit positions (inverse of
// compact_even_bits).
__device__ inline uint32_t spread_even_bits(uint32_t v) {
  v &= 0x0000ffffu;
  v = (v | (v << 8)) & 0x00ff00ffu;
  v = (v | (v << 4)) & 0x0f0f0f0fu;
  v = (v | (v << 2)) & 0x33333333u;
  v = (v | (v << 1)) & 0x55555555u;
  return v;
}

// The pixel of ring jr that holds the ring's theta: the first pixel of the
// ring in RING order; in NEST order a pixel of face 0 (rings 1 .. 2n-1), 4
// (ring 2n) or 8 (rings 2n+1 .. 4n-1) with ix + iy = jrll(face) n - jr - 1.
__device__ inline int32_t ring_holder(int32_t jr, int32_t nside, int32_t order) {
  if (order < 0) {
    const int32_t npix = 12 * nside * nside, ncap = 2 * nside * (nside - 1);
    if (jr < nside) return 2 * jr * (jr - 1);
    if (jr <= 3 * nside) return ncap + (jr - nside) * 4 * nside;
    const int32_t i = 4 * nside - jr;
    return npix - 2 * i * (i + 1);
  }
  const int32_t face = jr < 2 * nside ? 0 : (jr == 2 * nside ? 4 : 8);
  const int32_t jrll = face == 0 ? 2 : (face == 4 ? 3 : 4);
  const int32_t sum = jrll * nside - jr - 1;
  const int32_t ix = sum < nside - 1 ? sum : nside - 1, iy = sum - ix;
  return (face << (2 * order)) + (int32_t)(spread_even_bits((uint32_t)ix) | (spread_even_bits((uint32_t)iy) << 1));
}

#pragma clang fp contract(off)
__device__ inline double ring_z(int32_t jr, int32_t nr, int32_t nside) {
  const double three_ns2 = 3.0 * (double)(nside * nside);
  if (jr < nside) return (three_ns2 - (double)(nr * nr)) / three_ns2;
  if (jr <= 3 * nside) return (double)(2 * (2 * nside - jr)) / (double)(3 * nside);
  return -((three_ns2 - (double)(nr * nr)) / three_ns2);
}

// A division by a power of two is exact, so it is the same bits as the
// scaling by 2^-k (all equatorial rings when nside is a power of two).
__device__ inline double ring_phi(const RingPos& r) {
  const double a = (kPi / 2.0) * ((double)r.jp - 0.5 * (double)(r.kshift + 1));
  if ((r.nr & (r.nr - 1)) == 0) return ldexp(a, -(int)__builtin_ctz((uint32_t)r.nr));
  return a / (double)r.nr;
}

// theta is constant on a ring: acos(z) once per ring, written to the ring's
// holder pixel (ring_holder); k_pix2ang copies it to the other pixels.
__global__ void k_ring_theta(int32_t nside, int32_t order, double* __restrict__ theta) {
  const int32_t nring = 4 * nside - 1;
  for (int32_t jr = blockIdx.x * blockDim.x + threadIdx.x + 1; jr <= nring; jr += gridDim.x * blockDim.x) {
    int32_t nr, kshift;
    ring_shape(jr, nside, nr, kshift);
    theta[ring_holder(jr, nside, order)] = acos(ring_z(jr, nr, nside));
  }
}

// HOLDER: theta comes from the ring's holder pixel (k_ring_theta ran first);
// otherwise every pixel computes acos itself (small maps, where the second
// launch would cost more than the acos it saves).  Block b walks the pixels
// [b chunk, (b + 1) chunk), so a thread's pixels 256 apart mostly share a
// ring and it reloads theta only when the ring changes.
template <bool HOLDER>
__global__ void k_pix2ang(int32_t nside, int32_t order, int32_t lg4, int32_t npix, int32_t chunk,
                          double* __restrict__ theta, double* __restrict__ phi, int32_t* __restrict__ ring,
                          int32_t* __restrict__ iphi) {
  const int32_t p0 = blockIdx.x * chunk, p1 = min(npix, p0 + chunk);
  int32_t jr_held = -1;
  double th = 0.0;
  for (int32_t p = p0 + threadIdx.x; p < p1; p += blockDim.x) {
    const RingPos r = order >= 0 ? nest_pos(p, nside, order) : ring_pos(p, nside, lg4);
    if (HOLDER) {
      const int32_t h = ring_holder(r.jr, nside, order);
      if (r.jr != jr_held) {
        th = theta[h];  // written by k_ring_theta, earlier on the stream
        jr_held = r.jr;
      }
      if (p != h) theta[p] = th;
    } else {
      theta[p] = acos(ring_z(r.jr, r.nr, nside));
    }
    phi[p] = ring_phi(r);
    if (ring) ring[p] = r.jr;
    if (iphi) iphi[p] = r.jp;
  }
}

// l = sin t cos p, m = -sin t sin p, n = cos t (vissgen3d_noblockage.cu:58-70).
__global__ void k_pixel_dirs(const double* __restrict__ theta, const double* __restrict__ phi,
                             const double* __restrict__ B, int64_t P, double scale, double* __restrict__ lmn,
                             double* __restrict__ I) {
  for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < P; p += (int64_t)gridDim.x * blockDim.x) {
    double st, ct, sp, cp;
    sincos(theta[p], &st, &ct);
    sincos(phi[p], &sp, &cp);
    lmn[p] = st * cp;
    lmn[P + p] = -(st * sp);
    lmn[2 * P + p] = ct;
    if (I) I[p] = B[p] * scale;
  }
}
#pragma clang fp contract(on)

static unsigned grid_1d(int64_t n) {
  int64_t g = (n + 255) / 256;
  const int64_t cap = 256 * 64;  // 64 blocks of 256 per CU: enough to fill 256 CUs, grid-stride beyond
  if (g > cap) g = cap;
  return (unsigned)(g < 1 ? 1 : g);
}

}  // namespace viss

using namespace viss;

extern "C" int viss_pix2ang(int64_t nside, int scheme, double* theta, double* phi, int32_t* ring, int32_t* iphi,
                            void* stream) {
  VISS_REQUIRE(nside >= 1 && nside <= 8192, "viss_pix2ang: nside %lld outside [1, 8192]", (long long)nside);
  VISS_REQUIRE(scheme == VISS_SCHEME_RING || scheme == VISS_SCHEME_NEST, "viss_pix2ang: unknown scheme %d", scheme);
  VISS_REQUIRE(theta && phi, "viss_pix2ang: theta/phi must not be NULL");
  VISS_REQUIRE(scheme == VISS_SCHEME_RING || (nside & (nside - 1)) == 0,
               "viss_pix2ang: NEST needs a power-of-two nside (got %lld)", (long long)nside);
  const int64_t npix = 12 * nside * nside;
  int32_t order = -1;
  if (scheme == VISS_SCHEME_NEST) {
    order = 0;
    while ((1ll << order) < nside) ++order;
  }
  int32_t lg4 = -1;  // log2(4 nside) for a power-of-two nside
  if ((nside & (nside - 1)) == 0) {
    lg4 = 2;
    while ((1ll << (lg4 - 2)) < nside) ++lg4;
  }
  hipStream_t s = as_stream(stream);
  const unsigned grid0 = grid_1d(npix);
  const int64_t chunk = ((npix + grid0 - 1) / grid0 + 255) / 256 * 256;  // pixels per block, whole passes of 256
  const unsigned grid = (unsigned)((npix + chunk - 1) / chunk);
  if (npix >= (int64_t)1 << 22) {  // nside >= 592: one acos per ring
    k_ring_theta<<<grid_1d(4 * nside - 1), 256, 0, s>>>((int32_t)nside, order, theta);
    VISS_LAUNCHED(k_ring_theta);
    k_pix2ang<true><<<grid, 256, 0, s>>>((int32_t)nside, order, lg4, (int32_t)npix, (int32_t)chunk, theta, phi,
                                         ring, iphi);
  } else {
    k_pix2ang<false><<<grid, 256, 0, s>>>((int32_t)nside, order, lg4, (int32_t)npix, (int32_t)chunk, theta, phi,
                                          ring, iphi);
  }
  VISS_LAUNCHED(k_pix2ang);
  return VISS_OK;
}

extern "C" int viss_pixel_dirs(const double* theta, const double* phi, const double* B, int64_t P, double scale,
                               double* lmn, double* I, void* stream) {
  VISS_REQUIRE(P >= 0, "viss_pixel_dirs: negative P");
  if (P == 0) return VISS_OK;
  VISS_REQUIRE(theta && phi && lmn, "viss_pixel_dirs: theta/phi/lmn must not be NULL");
  VISS_REQUIRE((B == nullptr) == (I == nullptr), "viss_pixel_dirs: B and I must both be given or both NULL");
  k_pixel_dirs<<<grid_1d(P), 256, 0, as_stream(stream)>>>(theta, phi, B, P, scale, lmn, I);
  VISS_LAUNCHED(k_pixel_dirs);
  return VISS_OK;
}
