// Status/error plumbing and the sorted, tiled sky (viss_sky_*).
//
// The reference keeps pixels in file order and lets every visibility thread
// walk all of them (vissgen3d_noblockage.cu:86).  Here the pixel set is sorted
// once along a 3D Morton curve of the unit direction, so that consecutive
// 64-pixel tiles are spatially compact; each tile gets a bounding cone
// (centre, chord radius) that the occultation test uses to skip a tile, or to
// drop the per-pixel test, for a whole wavefront at once.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <vector>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "viss_internal.h"

namespace viss {

static thread_local char g_err[512] = "";

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

int hip_fail(hipError_t e, const char* what, const char* file, int line) {
  set_error("%s failed: %s (%d) at %s:%d", what, hipGetErrorString(e), (int)e, file, line);
  return e == hipErrorOutOfMemory ? VISS_ERR_NOMEM : VISS_ERR_HIP;
}

int DevBuf::reserve(size_t need, hipStream_t s) {
  if (need <= bytes) return VISS_OK;
  if (p) {
    // Work queued on `s` may still read the old buffer.
    VISS_HIP(hipStreamSynchronize(s));
    VISS_HIP(hipFree(p));
  }
  p = nullptr;
  bytes = 0;
  const size_t want = need + need / 4;
  VISS_HIP(hipMalloc(&p, want));
  bytes = want;
  return VISS_OK;
}

void DevBuf::release() {
  if (p) (void)hipFree(p);
  p = nullptr;
  bytes = 0;
}

// ---------------------------------------------------------------- kernels
__device__ inline uint32_t spread3(uint32_t x) {  // 10 bits -> every third bit
  x &= 0x3ffu;
  x = (x | (x << 16)) & 0x030000ffu;
  x = (x | (x << 8)) & 0x0300f00fu;
  x = (x | (x << 4)) & 0x030c30c3u;
  x = (x | (x << 2)) & 0x09249249u;
  return x;
}

__device__ inline uint32_t quant10(double c) {
  double q = (c + 1.0) * 512.0;
  q = q < 0.0 ? 0.0 : (q > 1023.0 ? 1023.0 : q);
  return (uint32_t)q;
}

__global__ void k_morton_keys(const double* __restrict__ lmn, int64_t P, uint32_t* __restrict__ key,
                              int32_t* __restrict__ val) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < P; i += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t x = quant10(lmn[i]), y = quant10(lmn[P + i]), z = quant10(lmn[2 * P + i]);
    key[i] = spread3(x) | (spread3(y) << 1) | (spread3(z) << 2);
    val[i] = (int32_t)i;
  }
}

__global__ void k_pack_sky(const double* __restrict__ lmn, int64_t P, int64_t P_pad, const int32_t* __restrict__ order,
                           PixPack* __restrict__ pix, double* __restrict__ I64) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < P_pad; i += (int64_t)gridDim.x * blockDim.x) {
    PixPack q;
    if (i < P) {
      const int32_t o = order[i];
      q.l = lmn[o];
      q.m = lmn[P + o];
      q.n = lmn[2 * P + o];
      q.idx = o;
    } else {
      q.l = q.m = q.n = 0.0;
      q.idx = -1;
    }
    q.I = 0.0f;
    pix[i] = q;
    I64[i] = 0.0;
  }
}

__device__ inline double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ inline double wave_max(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ inline int wave_isum(int v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// One wavefront per tile of `tile` pixels (tile = 64 * per_lane).
__global__ void k_tile_cones(const PixPack* __restrict__ pix, int64_t ntiles, int per_lane, float4* __restrict__ cone,
                             int32_t* __restrict__ valid_out) {
  const int lane = threadIdx.x & 63;
  const int64_t t = blockIdx.x * (int64_t)(blockDim.x >> 6) + wave_in_block();
  if (t >= ntiles) return;
  const int64_t base = t * 64 * per_lane;
  double sx = 0, sy = 0, sz = 0;
  int nvalid = 0;
  for (int i = 0; i < per_lane; ++i) {
    const PixPack q = pix[base + i * 64 + lane];
    if (q.idx >= 0) {
      sx += q.l;
      sy += q.m;
      sz += q.n;
      ++nvalid;
    }
  }
  sx = wave_sum(sx);
  sy = wave_sum(sy);
  sz = wave_sum(sz);
  nvalid = wave_isum(nvalid);
  const double nrm = sqrt(sx * sx + sy * sy + sz * sz);
  double cx = 0, cy = 0, cz = 1, chord = 2.0;
  if (nrm > 1e-6 * nvalid && nvalid > 0) {
    cx = sx / nrm;
    cy = sy / nrm;
    cz = sz / nrm;
    double mx = 0;
    for (int i = 0; i < per_lane; ++i) {
      const PixPack q = pix[base + i * 64 + lane];
      if (q.idx >= 0) {
        const double dx = q.l - cx, dy = q.m - cy, dz = q.n - cz;
        mx = fmax(mx, sqrt(dx * dx + dy * dy + dz * dz));
      }
    }
    chord = wave_max(mx) * (1.0 + 1e-6) + 1e-6;
  }
  if (lane == 0) {
    cone[t] = make_float4((float)cx, (float)cy, (float)cz, (float)chord);
    if (valid_out) valid_out[t] = nvalid;
  }
}

__global__ void k_scatter_brightness(PixPack* __restrict__ pix, int64_t P_pad, const double* __restrict__ I,
                                     double* __restrict__ I64, int want64) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < P_pad; i += (int64_t)gridDim.x * blockDim.x) {
    const int32_t o = pix[i].idx;
    const double v = o >= 0 ? I[o] : 0.0;
    pix[i].I = (float)v;
    if (want64) I64[i] = v;
  }
}

// ------------------------------------------------ antipodal-pair layout
// HEALPix (and any point-symmetric pixelisation) pairs every pixel with its
// antipode.  Pairs are found by sorting a 63-bit Morton key of each pixel's
// canonical direction (its upper-hemisphere representative): a pixel and its
// antipode share the key, so they end up adjacent.  A run of exactly two equal
// keys, one pixel from each hemisphere, whose canonical directions agree to
// 1e-12 is a pair; every other pixel is a single.
constexpr double kHemiEps = 1e-9;    // |n|, |m| below this count as zero (equator)
constexpr double kPairTol = 1e-12;   // max |c_a - c_b| of a pair's canonical directions

// Written with selects only: the same test as nested early returns plus a
// conditional negation in canonical() compiled (hipcc, gfx950) to code that
// dropped the negation for lanes with |n| <= kHemiEps.
__device__ inline bool upper_half(double l, double m, double n) {
  const bool by_n = fabs(n) > kHemiEps, by_m = fabs(m) > kHemiEps;
  return by_n ? (n > 0.0) : (by_m ? (m > 0.0) : (l >= 0.0));
}

__device__ inline uint64_t spread3_21(uint32_t x) {  // 21 bits -> every third bit of 63
  uint64_t v = x & 0x1fffffu;
  v = (v | (v << 32)) & 0x1f00000000ffffull;
  v = (v | (v << 16)) & 0x1f0000ff0000ffull;
  v = (v | (v << 8)) & 0x100f00f00f00f00full;
  v = (v | (v << 4)) & 0x10c30c30c30c30c3ull;
  v = (v | (v << 2)) & 0x1249249249249249ull;
  return v;
}

// Nearest of 2^21 levels on [-1, 1]: cell edges fall between dyadic values, so
// coordinates a pixelisation places exactly (0, +-1/2, ...) and their +-1e-17
// rounding twins land in the same cell.
__device__ inline uint32_t quant21(double c) {
  double q = (c + 1.0) * 1048576.0 + 0.5;
  q = q < 0.0 ? 0.0 : (q > 2097151.0 ? 2097151.0 : q);
  return (uint32_t)q;
}

__device__ inline void canonical(const double* __restrict__ lmn, int64_t P, int64_t i, double& x, double& y,
                                 double& z, bool& up) {
  const double l = lmn[i], m = lmn[P + i], n = lmn[2 * P + i];
  up = upper_half(l, m, n);
  const double sg = up ? 1.0 : -1.0;
  x = sg * l;
  y = sg * m;
  z = sg * n;
}

__global__ void k_pair_keys(const double* __restrict__ lmn, int64_t P, uint64_t* __restrict__ key,
                            int32_t* __restrict__ val) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < P; i += (int64_t)gridDim.x * blockDim.x) {
    double x, y, z;
    bool up;
    canonical(lmn, P, i, x, y, z, up);
    key[i] = spread3_21(quant21(x)) | (spread3_21(quant21(y)) << 1) | (spread3_21(quant21(z)) << 2);
    val[i] = (int32_t)i;
  }
}

// Sorted position i starts a pair (i, i+1).
__device__ inline bool pair_starts(const uint64_t* __restrict__ key, const int32_t* __restrict__ val,
                                   const double* __restrict__ lmn, int64_t P, int64_t i) {
  if (i < 0 || i + 1 >= P) return false;
  // every condition evaluated and combined with &, no early returns (see upper_half)
  const uint64_t k = key[i];
  const bool before = i > 0 ? key[i - 1] == k : false;
  const bool after = i + 2 < P ? key[i + 2] == k : false;
  double xa, ya, za, xb, yb, zb;
  bool ua, ub;
  canonical(lmn, P, val[i], xa, ya, za, ua);
  canonical(lmn, P, val[i + 1], xb, yb, zb, ub);
  return (key[i + 1] == k) & !before & !after & (ua != ub) & (fabs(xa - xb) <= kPairTol) &
         (fabs(ya - yb) <= kPairTol) & (fabs(za - zb) <= kPairTol);
}

__global__ void k_pair_flags(const uint64_t* __restrict__ key, const int32_t* __restrict__ val,
                             const double* __restrict__ lmn, int64_t P, uint8_t* __restrict__ is_pair,
                             uint8_t* __restrict__ is_single) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < P; i += (int64_t)gridDim.x * blockDim.x) {
    const bool st = pair_starts(key, val, lmn, P, i);
    is_pair[i] = st ? 1 : 0;
    is_single[i] = (!st && !pair_starts(key, val, lmn, P, i - 1)) ? 1 : 0;
  }
}

// Entries [0, npair): pairs; [npair_pad, npair_pad + nsingle): singles; the
// rest padding (no pixel).
__global__ void k_pack_pairs(const double* __restrict__ lmn, int64_t P, const int32_t* __restrict__ val,
                             const int32_t* __restrict__ pair_pos, int64_t npair, int64_t npair_pad,
                             const int32_t* __restrict__ single_pos, int64_t nsingle, int64_t n,
                             PairPack* __restrict__ pp, int2* __restrict__ pidx) {
  for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < n; j += (int64_t)gridDim.x * blockDim.x) {
    int32_t a = -1, b = -1;
    if (j < npair) {
      const int32_t i = pair_pos[j], p0 = val[i], p1 = val[i + 1];
      const bool up0 = upper_half(lmn[p0], lmn[P + p0], lmn[2 * P + p0]);
      a = up0 ? p0 : p1;
      b = up0 ? p1 : p0;
    } else if (j >= npair_pad && j < npair_pad + nsingle) {
      a = val[single_pos[j - npair_pad]];
    }
    PairPack q;
    q.l = a >= 0 ? lmn[a] : 0.0;
    q.m = a >= 0 ? lmn[P + a] : 0.0;
    q.n = a >= 0 ? lmn[2 * P + a] : 0.0;
    q.A = q.D = q.Ia = q.Ib = 0.f;
    pp[j] = q;
    pidx[j] = make_int2(a, b);
  }
}

// One wavefront per 64-entry tile: bounding cone of its a pixels.
__global__ void k_pair_cones(const PairPack* __restrict__ pp, const int2* __restrict__ pidx, int64_t ntiles,
                             float4* __restrict__ cone) {
  const int lane = threadIdx.x & 63;
  const int64_t t = blockIdx.x * (int64_t)(blockDim.x >> 6) + wave_in_block();
  if (t >= ntiles) return;
  const PairPack q = pp[t * 64 + lane];
  const bool valid = pidx[t * 64 + lane].x >= 0;
  double sx = valid ? q.l : 0.0, sy = valid ? q.m : 0.0, sz = valid ? q.n : 0.0;
  sx = wave_sum(sx);
  sy = wave_sum(sy);
  sz = wave_sum(sz);
  const int nvalid = wave_isum(valid ? 1 : 0);
  const double nrm = sqrt(sx * sx + sy * sy + sz * sz);
  double cx = 0, cy = 0, cz = 1, chord = 2.0;
  if (nrm > 1e-6 * nvalid && nvalid > 0) {
    cx = sx / nrm;
    cy = sy / nrm;
    cz = sz / nrm;
    double mx = 0;
    if (valid) {
      const double dx = q.l - cx, dy = q.m - cy, dz = q.n - cz;
      mx = sqrt(dx * dx + dy * dy + dz * dz);
    }
    chord = wave_max(mx) * (1.0 + 1e-6) + 1e-6;
  }
  if (lane == 0) cone[t] = make_float4((float)cx, (float)cy, (float)cz, (float)chord);
}

__global__ void k_scatter_pairs(PairPack* __restrict__ pp, const int2* __restrict__ pidx, int64_t n,
                                const double* __restrict__ I) {
  for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < n; j += (int64_t)gridDim.x * blockDim.x) {
    const int2 ab = pidx[j];
    const double ia = ab.x >= 0 ? I[ab.x] : 0.0, ib = ab.y >= 0 ? I[ab.y] : 0.0;
    pp[j].A = (float)(ia + ib);
    pp[j].D = (float)(ia - ib);
    pp[j].Ia = (float)ia;
    pp[j].Ib = (float)ib;
  }
}

static unsigned grid_for(int64_t n, int block) {
  int64_t g = (n + block - 1) / block;
  if (g > 65536) g = 65536;
  if (g < 1) g = 1;
  return (unsigned)g;
}

int check_device(const viss_sky* sky) {
  int dev = -1;
  VISS_HIP(hipGetDevice(&dev));
  if (dev != sky->device) {
    set_error("viss_sky was created on device %d but device %d is current", sky->device, dev);
    return VISS_ERR_INVALID;
  }
  return VISS_OK;
}

// Temporary device buffers of a build step, freed on the stream when the step
// returns (early error returns included).
class StreamBufs {
 public:
  explicit StreamBufs(hipStream_t s) : s_(s) {}
  StreamBufs(const StreamBufs&) = delete;
  StreamBufs& operator=(const StreamBufs&) = delete;
  ~StreamBufs() {
    for (void* p : ptrs_) (void)hipFreeAsync(p, s_);
  }
  template <class T>
  hipError_t alloc(T** p, size_t bytes) {
    void* q = nullptr;
    const hipError_t e = hipMallocAsync(&q, bytes, s_);
    if (e == hipSuccess) {
      ptrs_.push_back(q);
      *p = static_cast<T*>(q);
    }
    return e;
  }

 private:
  hipStream_t s_;
  std::vector<void*> ptrs_;
};

// Antipodal-pair layout (see k_pair_keys).  Synchronises once, to size the
// arrays from the pair count.
static int sky_build_pairs(viss_sky* sky, const double* lmn, int64_t P, hipStream_t s) {
  uint64_t *k_in = nullptr, *k_out = nullptr;
  int32_t *v_in = nullptr, *v_out = nullptr, *pair_pos = nullptr, *single_pos = nullptr, *counts = nullptr;
  uint8_t *is_pair = nullptr, *is_single = nullptr;
  void* tmp = nullptr;
  size_t tmp_bytes = 0, sel_bytes = 0;
  StreamBufs bufs(s);
  VISS_HIP(bufs.alloc(&k_in, sizeof(uint64_t) * P));
  VISS_HIP(bufs.alloc(&k_out, sizeof(uint64_t) * P));
  VISS_HIP(bufs.alloc(&v_in, sizeof(int32_t) * P));
  VISS_HIP(bufs.alloc(&v_out, sizeof(int32_t) * P));
  VISS_HIP(bufs.alloc(&pair_pos, sizeof(int32_t) * P));
  VISS_HIP(bufs.alloc(&single_pos, sizeof(int32_t) * P));
  VISS_HIP(bufs.alloc(&counts, sizeof(int32_t) * 2));
  VISS_HIP(bufs.alloc(&is_pair, P));
  VISS_HIP(bufs.alloc(&is_single, P));
  k_pair_keys<<<grid_for(P, 256), 256, 0, s>>>(lmn, P, k_in, v_in);
  VISS_LAUNCHED(k_pair_keys);
  VISS_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, k_in, k_out, v_in, v_out, (int)P, 0, 63, s));
  hipcub::CountingInputIterator<int32_t> iota(0);
  VISS_HIP(hipcub::DeviceSelect::Flagged(nullptr, sel_bytes, iota, is_pair, pair_pos, counts, (int)P, s));
  tmp_bytes = std::max(tmp_bytes, sel_bytes);
  VISS_HIP(bufs.alloc(&tmp, tmp_bytes));
  VISS_HIP(hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, k_in, k_out, v_in, v_out, (int)P, 0, 63, s));
  k_pair_flags<<<grid_for(P, 256), 256, 0, s>>>(k_out, v_out, lmn, P, is_pair, is_single);
  VISS_LAUNCHED(k_pair_flags);
  size_t b = tmp_bytes;
  VISS_HIP(hipcub::DeviceSelect::Flagged(tmp, b, iota, is_pair, pair_pos, counts, (int)P, s));
  b = tmp_bytes;
  VISS_HIP(hipcub::DeviceSelect::Flagged(tmp, b, iota, is_single, single_pos, counts + 1, (int)P, s));
  int32_t hc[2] = {0, 0};
  VISS_HIP(hipMemcpyAsync(hc, counts, sizeof(hc), hipMemcpyDeviceToHost, s));
  VISS_HIP(hipStreamSynchronize(s));
  sky->npair = hc[0];
  sky->nsingle = hc[1];
  int rc = VISS_OK;
  if (2 * sky->npair + sky->nsingle != P) {
    set_error("viss_sky: pair detection lost pixels (%lld pairs, %lld singles, %lld pixels)",
              (long long)sky->npair, (long long)sky->nsingle, (long long)P);
    rc = VISS_ERR_HIP;
  } else {
    const int64_t npair_pad = (sky->npair + 63) / 64 * 64, nsingle_pad = (sky->nsingle + 63) / 64 * 64;
    sky->nptile = (npair_pad + nsingle_pad) / 64;
    sky->use_pairs = 2 * sky->npair >= P / 2;
    const int64_t n = sky->nptile * 64;
    if (n > 0) {
      VISS_HIP(hipMalloc((void**)&sky->ppack, sizeof(PairPack) * n));
      VISS_HIP(hipMalloc((void**)&sky->pidx, sizeof(int2) * n));
      VISS_HIP(hipMalloc((void**)&sky->pcone, sizeof(float4) * sky->nptile));
      k_pack_pairs<<<grid_for(n, 256), 256, 0, s>>>(lmn, P, v_out, pair_pos, sky->npair, npair_pad, single_pos,
                                                     sky->nsingle, n, sky->ppack, sky->pidx);
      VISS_LAUNCHED(k_pack_pairs);
      k_pair_cones<<<(unsigned)((sky->nptile + 3) / 4), 256, 0, s>>>(sky->ppack, sky->pidx, sky->nptile,
                                                                      sky->pcone);
      VISS_LAUNCHED(k_pair_cones);
    }
  }
  return rc;
}

// ------------------------------------------------ equatorial-mirror quads
// HEALPix is also symmetric under z -> -z, ring by ring.  Two antipodal pairs
// whose a pixels (upper hemisphere) are (x, y, z) and (-x, -y, z) form a quad
// {a, b, c, d} = {(x, y, z), (-x, -y, z), (x, y, -z), (-x, -y, -z)}.  They are
// found like the pairs: a Morton key of the a pixel's direction with (x, y)
// replaced by its upper-half-plane representative; a run of exactly two equal
// keys, opposite half-planes, z and (x, y) agreeing to 1e-12 is a quad.

// Upper half-plane in (x, y), written with selects (see upper_half).
__device__ inline bool xy_upper(double l, double m) {
  return fabs(m) > kHemiEps ? (m > 0.0) : (l > 0.0);
}

__global__ void k_quad_keys(const PairPack* __restrict__ pp, int64_t npair, uint64_t* __restrict__ key,
                            int32_t* __restrict__ val) {
  for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < npair; j += (int64_t)gridDim.x * blockDim.x) {
    const double l = pp[j].l, m = pp[j].m, n = pp[j].n;
    const double sg = xy_upper(l, m) ? 1.0 : -1.0;
    const uint64_t k = spread3_21(quant21(sg * l)) | (spread3_21(quant21(sg * m)) << 1) |
                       (spread3_21(quant21(n)) << 2);
    key[j] = n > kHemiEps ? k : ~0ull;  // equator pairs never form quads
    val[j] = (int32_t)j;
  }
}

// Sorted position i starts a quad (pairs val[i], val[i+1]).
__device__ inline bool quad_starts(const uint64_t* __restrict__ key, const int32_t* __restrict__ val,
                                   const PairPack* __restrict__ pp, int64_t npair, int64_t i) {
  if (i < 0 || i + 1 >= npair) return false;
  const uint64_t k = key[i];
  const bool before = i > 0 ? key[i - 1] == k : false;
  const bool after = i + 2 < npair ? key[i + 2] == k : false;
  const PairPack p = pp[val[i]], q = pp[val[i + 1]];
  const bool up_p = xy_upper(p.l, p.m), up_q = xy_upper(q.l, q.m);
  // z agrees to 1e-12, not bitwise: the second pair's a pixel is the antipode
  // of a south-ring pixel, whose cos(pi - theta) can differ from -cos(theta)
  // by an ulp.  Tiles take the first pair's z (a north-ring pixel, the same
  // bits along a ring).
  return (k != ~0ull) & (key[i + 1] == k) & !before & !after & (up_p != up_q) & (fabs(p.n - q.n) <= kPairTol) &
         (fabs(p.l + q.l) <= kPairTol) & (fabs(p.m + q.m) <= kPairTol);
}

__global__ void k_quad_flags(const uint64_t* __restrict__ key, const int32_t* __restrict__ val,
                             const PairPack* __restrict__ pp, int64_t npair, uint8_t* __restrict__ is_quad) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < npair; i += (int64_t)gridDim.x * blockDim.x)
    is_quad[i] = quad_starts(key, val, pp, npair, i) ? 1 : 0;
}

// Quad q (sorted position start[q]): members P1 (a pixel in the upper
// half-plane) and P2; sort keys azimuth then z (both >= 0, so their IEEE bits
// order like the values).
__global__ void k_quad_members(const int32_t* __restrict__ start, const int32_t* __restrict__ val,
                               const PairPack* __restrict__ pp, int64_t nquad, int2* __restrict__ mem,
                               uint64_t* __restrict__ phi_key, uint64_t* __restrict__ z_key,
                               int32_t* __restrict__ qid, uint8_t* __restrict__ in_quad) {
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < nquad; q += (int64_t)gridDim.x * blockDim.x) {
    const int32_t i = start[q], j0 = val[i], j1 = val[i + 1];
    const bool first = xy_upper(pp[j0].l, pp[j0].m);
    const int32_t p1 = first ? j0 : j1, p2 = first ? j1 : j0;
    mem[q] = make_int2(p1, p2);
    const double phi = atan2(pp[p1].m, pp[p1].l) + 0.0;  // [0, pi]; + 0.0 maps -0 to +0
    phi_key[q] = __double_as_longlong(phi < 0.0 ? 0.0 : phi);
    z_key[q] = __double_as_longlong(pp[p1].n);
    qid[q] = (int32_t)q;
    in_quad[p1] = 1;
    in_quad[p2] = 1;
  }
}

__global__ void k_gather_u64(const uint64_t* __restrict__ src, const int32_t* __restrict__ idx, int64_t n,
                             uint64_t* __restrict__ dst) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = src[idx[i]];
}

__global__ void k_new_ring(const uint64_t* __restrict__ zs, int64_t n, int32_t* __restrict__ flag) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    flag[i] = (i == 0 || zs[i] != zs[i - 1]) ? 1 : 0;
}

// Tiles of ring r: ceil(count / kQuadTile); rings beyond nring get none.
__global__ void k_ring_tiles(const int32_t* __restrict__ rstart, const int32_t* __restrict__ nring_d, int64_t n,
                             int32_t* __restrict__ ntiles) {
  const int32_t nring = *nring_d;
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
    const int64_t end = r + 1 < nring ? rstart[r + 1] : n;
    ntiles[r] = r < nring ? (int32_t)((end - rstart[r] + kQuadTile - 1) / kQuadTile) : 0;
  }
}

__global__ void k_quad_clear(QuadPack* __restrict__ qp, int4* __restrict__ qi, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    QuadPack z;
    z.x = z.y = 0.0;
    z.Az = z.Dz = z.Am = z.Dm = z.Ia = z.Ib = z.Ic = z.Id = 0.f;
    qp[i] = z;
    qi[i] = make_int4(-1, -1, -1, -1);
  }
}

// Sorted quad i (ring ring_of[i], position i - rstart) -> its tile slot.
__global__ void k_quad_pack(const int32_t* __restrict__ order, const int2* __restrict__ mem,
                            const int32_t* __restrict__ ring_incl, const int32_t* __restrict__ rstart,
                            const int32_t* __restrict__ toff, const PairPack* __restrict__ pp,
                            const int2* __restrict__ pidx, int64_t nquad, QuadPack* __restrict__ qp,
                            int4* __restrict__ qi, double* __restrict__ qz, int2* __restrict__ tile_pos) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nquad; i += (int64_t)gridDim.x * blockDim.x) {
    const int2 m = mem[order[i]];
    const int32_t r = ring_incl[i] - 1;
    const int64_t pos = i - rstart[r];
    const int64_t t = toff[r] + pos / kQuadTile, slot = t * kQuadTile + pos % kQuadTile;
    const PairPack p1 = pp[m.x];
    qp[slot].x = p1.l;
    qp[slot].y = p1.m;
    const int2 i1 = pidx[m.x], i2 = pidx[m.y];
    qi[slot] = make_int4(i1.x, i2.x, i2.y, i1.y);  // a = P1.a, b = P2.a, c = P2.b, d = P1.b
    if (pos % kQuadTile == 0) {
      qz[t] = p1.n;
      tile_pos[t] = make_int2(r, (int32_t)(pos / kQuadTile));
    }
  }
}

// Phase recurrence of each quad tile (quad_tile_forward): the a pixels of a
// HEALPix ring sit at equal azimuth steps d on one circle, so (x_j, y_j), and
// with them A_j = u x_j + v y_j, obey q_{j+1} = c2 q_j - q_{j-1}, c2 = 2 cos d.
// qc2[t] = 2 (p0 . p1) / |p0|^2 from the tile's first two quads (the quads of a
// tile are a prefix of its slots, sorted by azimuth); a tile whose (x, y), run
// through the recurrence in FP64 as the kernels run it, leave the records by
// more than kRecTol anywhere counts in *bad, and the sky then keeps the direct
// phase for every tile.  A HEALPix sky stays within 9e-14 (nside 64-8192).
constexpr double kRecTol = 2.5e-13;
__global__ void k_quad_rec(const QuadPack* __restrict__ qp, const int4* __restrict__ qi, int64_t ntiles,
                           double* __restrict__ qc2, int32_t* __restrict__ bad) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < ntiles; t += (int64_t)gridDim.x * blockDim.x) {
    const QuadPack* __restrict__ tq = qp + t * kQuadTile;
    int n = 0;
    for (int j = 0; j < kQuadTile; ++j) n += qi[t * kQuadTile + j].x >= 0 ? 1 : 0;
    double c2 = 0.0;
    bool ok = true;
    if (n >= 3) {
      const double x0 = tq[0].x, y0 = tq[0].y, x1 = tq[1].x, y1 = tq[1].y;
      c2 = 2.0 * (x0 * x1 + y0 * y1) / (x0 * x0 + y0 * y0);
      ok = fabs(c2) <= 2.0;  // false for NaN too
      double px = x0, py = y0, cx = x1, cy = y1;
      for (int j = 2; j < n; ++j) {
        const double nx = fma(c2, cx, -px), ny = fma(c2, cy, -py);
        ok = ok & (fabs(nx - tq[j].x) <= kRecTol) & (fabs(ny - tq[j].y) <= kRecTol);
        px = cx;
        py = cy;
        cx = nx;
        cy = ny;
      }
    }
    qc2[t] = ok ? c2 : 0.0;
    if (!ok) atomicAdd(bad, 1);
  }
}

// Cones of the a pixels (x, y, z) and b pixels (-x, -y, z) of each tile; the
// d and c pixels are their antipodes (the negated cones).
__global__ void k_quad_cones(const QuadPack* __restrict__ qp, const int4* __restrict__ qi,
                             const double* __restrict__ qz, int64_t ntiles, float4* __restrict__ cone) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < ntiles; t += (int64_t)gridDim.x * blockDim.x) {
    const double z = qz[t];
    double sx = 0.0, sy = 0.0;
    int cnt = 0;
    for (int j = 0; j < kQuadTile; ++j) {
      const bool v = qi[t * kQuadTile + j].x >= 0;
      sx += v ? qp[t * kQuadTile + j].x : 0.0;
      sy += v ? qp[t * kQuadTile + j].y : 0.0;
      cnt += v ? 1 : 0;
    }
    for (int side = 0; side < 2; ++side) {
      const double sg = side == 0 ? 1.0 : -1.0;
      const double cx0 = sg * sx, cy0 = sg * sy, cz0 = cnt * z;
      const double nrm = sqrt(cx0 * cx0 + cy0 * cy0 + cz0 * cz0);
      double cx = 0.0, cy = 0.0, cz = 1.0, chord = 2.0;
      if (cnt > 0 && nrm > 1e-6 * cnt) {
        cx = cx0 / nrm;
        cy = cy0 / nrm;
        cz = cz0 / nrm;
        double mx = 0.0;
        for (int j = 0; j < kQuadTile; ++j) {
          const bool v = qi[t * kQuadTile + j].x >= 0;
          const double dx = sg * qp[t * kQuadTile + j].x - cx, dy = sg * qp[t * kQuadTile + j].y - cy, dz = z - cz;
          mx = v ? fmax(mx, sqrt(dx * dx + dy * dy + dz * dz)) : mx;
        }
        chord = mx * (1.0 + 1e-6) + 1e-6;
      }
      cone[2 * t + side] = make_float4((float)cx, (float)cy, (float)cz, (float)chord);
    }
  }
}

// Cones of the a pixels and b pixels of each back-projection patch (4 tiles).
__global__ void k_patch_cones(const QuadPack* __restrict__ qp, const int4* __restrict__ qi,
                              const double* __restrict__ qz, const int4* __restrict__ patch, int64_t npatch,
                              float4* __restrict__ cone) {
  for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < npatch; p += (int64_t)gridDim.x * blockDim.x) {
    const int4 pt = patch[p];
    const int32_t tl[4] = {pt.x, pt.y, pt.z, pt.w};
    for (int side = 0; side < 2; ++side) {
      const double sg = side == 0 ? 1.0 : -1.0;
      double sx = 0.0, sy = 0.0, sz = 0.0;
      int cnt = 0;
      for (int g = 0; g < 4; ++g) {
        if (tl[g] < 0) continue;
        const double z = qz[tl[g]];
        for (int j = 0; j < kQuadTile; ++j) {
          const int64_t q = (int64_t)tl[g] * kQuadTile + j;
          const bool v = qi[q].x >= 0;
          sx += v ? sg * qp[q].x : 0.0;
          sy += v ? sg * qp[q].y : 0.0;
          sz += v ? z : 0.0;
          cnt += v ? 1 : 0;
        }
      }
      const double nrm = sqrt(sx * sx + sy * sy + sz * sz);
      double cx = 0.0, cy = 0.0, cz = 1.0, chord = 2.0;
      if (cnt > 0 && nrm > 1e-6 * cnt) {
        cx = sx / nrm;
        cy = sy / nrm;
        cz = sz / nrm;
        double mx = 0.0;
        for (int g = 0; g < 4; ++g) {
          if (tl[g] < 0) continue;
          const double z = qz[tl[g]];
          for (int j = 0; j < kQuadTile; ++j) {
            const int64_t q = (int64_t)tl[g] * kQuadTile + j;
            const double dx = sg * qp[q].x - cx, dy = sg * qp[q].y - cy, dz = z - cz;
            mx = qi[q].x >= 0 ? fmax(mx, sqrt(dx * dx + dy * dy + dz * dz)) : mx;
          }
        }
        chord = mx * (1.0 + 1e-6) + 1e-6;
      }
      cone[2 * p + side] = make_float4((float)cx, (float)cy, (float)cz, (float)chord);
    }
  }
}

__global__ void k_scatter_quads(QuadPack* __restrict__ qp, const int4* __restrict__ qi, int64_t n,
                                const double* __restrict__ I) {
  for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < n; j += (int64_t)gridDim.x * blockDim.x) {
    const int4 q = qi[j];
    const double a = q.x >= 0 ? I[q.x] : 0.0, b = q.y >= 0 ? I[q.y] : 0.0;
    const double c = q.z >= 0 ? I[q.z] : 0.0, d = q.w >= 0 ? I[q.w] : 0.0;
    qp[j].Az = (float)(a + b);
    qp[j].Dz = (float)(a - b);
    qp[j].Am = (float)(c + d);
    qp[j].Dm = (float)(c - d);
    qp[j].Ia = (float)a;
    qp[j].Ib = (float)b;
    qp[j].Ic = (float)c;
    qp[j].Id = (float)d;
  }
}

// Rest layout: entry t takes main-layout entry src[t] (pairs not in a quad,
// then the singles), padding where src is -1.
__global__ void k_rest_gather(const int32_t* __restrict__ rest_pairs, int64_t nrp, int64_t nrp_pad,
                              int64_t single0, int64_t nsingle, int64_t n, const PairPack* __restrict__ pp,
                              const int2* __restrict__ pidx, PairPack* __restrict__ rp, int2* __restrict__ ri) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t src = t < nrp ? (int64_t)rest_pairs[t]
                                : ((t >= nrp_pad && t < nrp_pad + nsingle) ? single0 + (t - nrp_pad) : -1);
    PairPack z;
    z.l = z.m = z.n = 0.0;
    z.A = z.D = z.Ia = z.Ib = 0.f;
    rp[t] = src >= 0 ? pp[src] : z;
    ri[t] = src >= 0 ? pidx[src] : make_int2(-1, -1);
  }
}

__global__ void k_rest_flags(const uint8_t* __restrict__ in_quad, int64_t npair, uint8_t* __restrict__ rest) {
  for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < npair; j += (int64_t)gridDim.x * blockDim.x)
    rest[j] = in_quad[j] ? 0 : 1;
}

// Quad layout on top of the pair layout.  Synchronises twice (quad count,
// tile count).  Leaves use_quads false when the quads cover under half of
// the pixels or their tiles are under 2/3 full.
static int sky_build_quads(viss_sky* sky, hipStream_t s) {
  const int64_t np = sky->npair;
  if (!sky->use_pairs || np < 2) return VISS_OK;
  uint64_t *k_in = nullptr, *k_out = nullptr, *phi_key = nullptr, *z_key = nullptr, *k2 = nullptr;
  int32_t *v_in = nullptr, *v_out = nullptr, *start = nullptr, *counts = nullptr, *qid = nullptr, *order = nullptr,
          *order2 = nullptr, *flag = nullptr, *ring_incl = nullptr, *rstart = nullptr, *ntl = nullptr,
          *toff = nullptr, *rest_pairs = nullptr;
  int2* mem = nullptr;
  uint8_t *is_quad = nullptr, *in_quad = nullptr, *rest = nullptr;
  void* tmp = nullptr;
  size_t tmp_bytes = 0, b = 0;
  hipcub::CountingInputIterator<int32_t> iota(0);
  int rc = VISS_OK;
  StreamBufs bufs(s);
  VISS_HIP(bufs.alloc(&k_in, sizeof(uint64_t) * np));
  VISS_HIP(bufs.alloc(&k_out, sizeof(uint64_t) * np));
  VISS_HIP(bufs.alloc(&phi_key, sizeof(uint64_t) * np));
  VISS_HIP(bufs.alloc(&z_key, sizeof(uint64_t) * np));
  VISS_HIP(bufs.alloc(&k2, sizeof(uint64_t) * np));
  VISS_HIP(bufs.alloc(&v_in, sizeof(int32_t) * np));
  VISS_HIP(bufs.alloc(&v_out, sizeof(int32_t) * np));
  VISS_HIP(bufs.alloc(&start, sizeof(int32_t) * np));
  VISS_HIP(bufs.alloc(&counts, sizeof(int32_t) * 4));
  VISS_HIP(bufs.alloc(&qid, sizeof(int32_t) * np));
  VISS_HIP(bufs.alloc(&order, sizeof(int32_t) * np));
  VISS_HIP(bufs.alloc(&order2, sizeof(int32_t) * np));
  VISS_HIP(bufs.alloc(&flag, sizeof(int32_t) * np));
  VISS_HIP(bufs.alloc(&ring_incl, sizeof(int32_t) * np));
  VISS_HIP(bufs.alloc(&rstart, sizeof(int32_t) * np));
  VISS_HIP(bufs.alloc(&ntl, sizeof(int32_t) * np));
  VISS_HIP(bufs.alloc(&toff, sizeof(int32_t) * np));
  VISS_HIP(bufs.alloc(&rest_pairs, sizeof(int32_t) * np));
  VISS_HIP(bufs.alloc(&mem, sizeof(int2) * np));
  VISS_HIP(bufs.alloc(&is_quad, np));
  VISS_HIP(bufs.alloc(&in_quad, np));
  VISS_HIP(bufs.alloc(&rest, np));
  // scratch large enough for every primitive below
  {
    size_t t1 = 0, t2 = 0, t3 = 0, t4 = 0, t5 = 0;
    VISS_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, t1, k_in, k_out, v_in, v_out, (int)np, 0, 64, s));
    VISS_HIP(hipcub::DeviceSelect::Flagged(nullptr, t2, iota, is_quad, start, counts, (int)np, s));
    VISS_HIP(hipcub::DeviceScan::InclusiveSum(nullptr, t3, flag, ring_incl, (int)np, s));
    VISS_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, t4, ntl, toff, (int)np, s));
    VISS_HIP(hipcub::DeviceSelect::Flagged(nullptr, t5, iota, flag, rstart, counts + 1, (int)np, s));
    tmp_bytes = std::max(std::max(std::max(t1, t2), std::max(t3, t4)), t5);
  }
  VISS_HIP(bufs.alloc(&tmp, tmp_bytes));
  VISS_HIP(hipMemsetAsync(in_quad, 0, np, s));
  k_quad_keys<<<grid_for(np, 256), 256, 0, s>>>(sky->ppack, np, k_in, v_in);
  VISS_LAUNCHED(k_quad_keys);
  b = tmp_bytes;
  VISS_HIP(hipcub::DeviceRadixSort::SortPairs(tmp, b, k_in, k_out, v_in, v_out, (int)np, 0, 64, s));
  k_quad_flags<<<grid_for(np, 256), 256, 0, s>>>(k_out, v_out, sky->ppack, np, is_quad);
  VISS_LAUNCHED(k_quad_flags);
  b = tmp_bytes;
  VISS_HIP(hipcub::DeviceSelect::Flagged(tmp, b, iota, is_quad, start, counts, (int)np, s));
  int32_t hq = 0;
  VISS_HIP(hipMemcpyAsync(&hq, counts, sizeof(hq), hipMemcpyDeviceToHost, s));
  VISS_HIP(hipStreamSynchronize(s));
  const int64_t nq = hq;
  sky->nquad = nq;
  if (nq > 0 && 4 * nq >= sky->P / 2) {
    k_quad_members<<<grid_for(nq, 256), 256, 0, s>>>(start, v_out, sky->ppack, nq, mem, phi_key, z_key, qid,
                                                     in_quad);
    VISS_LAUNCHED(k_quad_members);
    // azimuth, then z (stable): sorted by (z, azimuth)
    b = tmp_bytes;
    VISS_HIP(hipcub::DeviceRadixSort::SortPairs(tmp, b, phi_key, k_out, qid, order2, (int)nq, 0, 64, s));
    k_gather_u64<<<grid_for(nq, 256), 256, 0, s>>>(z_key, order2, nq, k2);
    VISS_LAUNCHED(k_gather_u64);
    b = tmp_bytes;
    VISS_HIP(hipcub::DeviceRadixSort::SortPairs(tmp, b, k2, k_in, order2, order, (int)nq, 0, 64, s));
    k_new_ring<<<grid_for(nq, 256), 256, 0, s>>>(k_in, nq, flag);
    VISS_LAUNCHED(k_new_ring);
    b = tmp_bytes;
    VISS_HIP(hipcub::DeviceScan::InclusiveSum(tmp, b, flag, ring_incl, (int)nq, s));
    b = tmp_bytes;
    VISS_HIP(hipcub::DeviceSelect::Flagged(tmp, b, iota, flag, rstart, counts + 1, (int)nq, s));
    k_ring_tiles<<<grid_for(nq, 256), 256, 0, s>>>(rstart, counts + 1, nq, ntl);
    VISS_LAUNCHED(k_ring_tiles);
    b = tmp_bytes;
    VISS_HIP(hipcub::DeviceScan::ExclusiveSum(tmp, b, ntl, toff, (int)nq, s));
    int32_t last[2] = {0, 0};
    VISS_HIP(hipMemcpyAsync(&last[0], toff + nq - 1, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    VISS_HIP(hipMemcpyAsync(&last[1], ntl + nq - 1, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    VISS_HIP(hipStreamSynchronize(s));
    const int64_t nqt = (int64_t)last[0] + last[1];
    if (3 * nq >= 2 * kQuadTile * nqt) {
      sky->nqtile = nqt;
      const int64_t nslot = nqt * kQuadTile;
      VISS_HIP(hipMalloc((void**)&sky->qpack, sizeof(QuadPack) * nslot));
      VISS_HIP(hipMalloc((void**)&sky->qidx, sizeof(int4) * nslot));
      VISS_HIP(hipMalloc((void**)&sky->qz, sizeof(double) * nqt));
      VISS_HIP(hipMalloc((void**)&sky->qc2, sizeof(double) * nqt));
      VISS_HIP(hipMalloc((void**)&sky->qcone, sizeof(float4) * 2 * nqt));
      k_quad_clear<<<grid_for(nslot, 256), 256, 0, s>>>(sky->qpack, sky->qidx, nslot);
      VISS_LAUNCHED(k_quad_clear);
      int2* tile_pos = nullptr;
      VISS_HIP(bufs.alloc(&tile_pos, sizeof(int2) * nqt));
      k_quad_pack<<<grid_for(nq, 256), 256, 0, s>>>(order, mem, ring_incl, rstart, toff, sky->ppack, sky->pidx, nq,
                                                    sky->qpack, sky->qidx, sky->qz, tile_pos);
      VISS_LAUNCHED(k_quad_pack);
      k_quad_cones<<<grid_for(nqt, 256), 256, 0, s>>>(sky->qpack, sky->qidx, sky->qz, nqt, sky->qcone);
      VISS_LAUNCHED(k_quad_cones);
      VISS_HIP(hipMemsetAsync(counts + 3, 0, sizeof(int32_t), s));
      k_quad_rec<<<grid_for(nqt, 256), 256, 0, s>>>(sky->qpack, sky->qidx, nqt, sky->qc2, counts + 3);
      VISS_LAUNCHED(k_quad_rec);
      // Back-projection patches: bands of 4 consecutive rings (in z order);
      // patch j of a band takes azimuth tile j of each ring, so a wavefront's
      // 64 quads stay compact.
      std::vector<int2> hpos((size_t)nqt);
      VISS_HIP(hipMemcpyAsync(hpos.data(), tile_pos, sizeof(int2) * nqt, hipMemcpyDeviceToHost, s));
      VISS_HIP(hipStreamSynchronize(s));
      std::vector<int4> patches;
      {
        int64_t t = 0;
        while (t < nqt) {
          const int32_t band = hpos[(size_t)t].x / 4;
          int64_t e = t;
          while (e < nqt && hpos[(size_t)e].x / 4 == band) ++e;
          int32_t nphi = 0;
          for (int64_t i = t; i < e; ++i) nphi = std::max(nphi, hpos[(size_t)i].y + 1);
          const size_t p0 = patches.size();
          patches.resize(p0 + (size_t)nphi, make_int4(-1, -1, -1, -1));
          for (int64_t i = t; i < e; ++i) {
            int4& pt = patches[p0 + (size_t)hpos[(size_t)i].y];
            const int g = hpos[(size_t)i].x % 4;
            (g == 0 ? pt.x : g == 1 ? pt.y : g == 2 ? pt.z : pt.w) = (int32_t)i;
          }
          t = e;
        }
      }
      sky->nqpatch = (int64_t)patches.size();
      VISS_HIP(hipMalloc((void**)&sky->qpatch, sizeof(int4) * sky->nqpatch));
      VISS_HIP(hipMalloc((void**)&sky->qpcone, sizeof(float4) * 2 * sky->nqpatch));
      VISS_HIP(hipMemcpyAsync(sky->qpatch, patches.data(), sizeof(int4) * sky->nqpatch, hipMemcpyHostToDevice, s));
      k_patch_cones<<<grid_for(sky->nqpatch, 256), 256, 0, s>>>(sky->qpack, sky->qidx, sky->qz, sky->qpatch,
                                                                sky->nqpatch, sky->qpcone);
      VISS_LAUNCHED(k_patch_cones);
      VISS_HIP(hipStreamSynchronize(s));  // the host patch table goes out of scope
      // rest: pairs outside quads, then the singles of the pair layout
      k_rest_flags<<<grid_for(np, 256), 256, 0, s>>>(in_quad, np, rest);
      VISS_LAUNCHED(k_rest_flags);
      b = tmp_bytes;
      VISS_HIP(hipcub::DeviceSelect::Flagged(tmp, b, iota, rest, rest_pairs, counts + 2, (int)np, s));
      int32_t hr[2] = {0, 0};  // rest pairs, tiles off the phase recurrence
      VISS_HIP(hipMemcpyAsync(hr, counts + 2, sizeof(hr), hipMemcpyDeviceToHost, s));
      VISS_HIP(hipStreamSynchronize(s));
      // VISS_QUAD_RECURRENCE=0 keeps the direct phase (A/B runs and the test
      // that compares the two phases on one sky)
      const char* rec_env = getenv("VISS_QUAD_RECURRENCE");
      sky->qrec = hr[1] == 0 && !(rec_env && rec_env[0] == '0');
      const int64_t nrp = hr[0], nrp_pad = (nrp + 63) / 64 * 64, ns_pad = (sky->nsingle + 63) / 64 * 64;
      const int64_t single0 = (sky->npair + 63) / 64 * 64;
      sky->nrtile = (nrp_pad + ns_pad) / 64;
      if (nrp != np - 2 * nq) {
        set_error("viss_sky: quad layout lost pairs (%lld quads, %lld rest pairs, %lld pairs)", (long long)nq,
                  (long long)nrp, (long long)np);
        rc = VISS_ERR_HIP;
      } else {
        if (sky->nrtile > 0) {
          const int64_t n = sky->nrtile * 64;
          VISS_HIP(hipMalloc((void**)&sky->rpack, sizeof(PairPack) * n));
          VISS_HIP(hipMalloc((void**)&sky->ridx, sizeof(int2) * n));
          VISS_HIP(hipMalloc((void**)&sky->rcone, sizeof(float4) * sky->nrtile));
          k_rest_gather<<<grid_for(n, 256), 256, 0, s>>>(rest_pairs, nrp, nrp_pad, single0, sky->nsingle, n,
                                                         sky->ppack, sky->pidx, sky->rpack, sky->ridx);
          VISS_LAUNCHED(k_rest_gather);
          k_pair_cones<<<(unsigned)((sky->nrtile + 3) / 4), 256, 0, s>>>(sky->rpack, sky->ridx, sky->nrtile,
                                                                          sky->rcone);
          VISS_LAUNCHED(k_pair_cones);
        }
        sky->use_quads = true;
      }
    }
  }
  return rc;
}

int sky_build(viss_sky* sky, const double* lmn, int64_t P, hipStream_t s) {
  VISS_REQUIRE(P > 0 && P <= (int64_t)INT32_MAX - kPadTo, "viss_sky: pixel count %lld out of range", (long long)P);
  VISS_HIP(hipGetDevice(&sky->device));
  sky->P = P;
  sky->P_pad = (P + kPadTo - 1) / kPadTo * kPadTo;
  sky->ntile = sky->P_pad / kTile;
  sky->nbtile = sky->P_pad / kBpTile;
  const int64_t Pp = sky->P_pad;

  uint32_t *k_in = nullptr, *k_out = nullptr;
  int32_t *v_in = nullptr, *v_out = nullptr;
  void* tmp = nullptr;
  size_t tmp_bytes = 0;
  StreamBufs bufs(s);
  VISS_HIP(bufs.alloc(&k_in, sizeof(uint32_t) * P));
  VISS_HIP(bufs.alloc(&k_out, sizeof(uint32_t) * P));
  VISS_HIP(bufs.alloc(&v_in, sizeof(int32_t) * P));
  VISS_HIP(bufs.alloc(&v_out, sizeof(int32_t) * P));
  k_morton_keys<<<grid_for(P, 256), 256, 0, s>>>(lmn, P, k_in, v_in);
  VISS_LAUNCHED(k_morton_keys);
  VISS_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, k_in, k_out, v_in, v_out, (int)P, 0, 30, s));
  VISS_HIP(bufs.alloc(&tmp, tmp_bytes));
  VISS_HIP(hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, k_in, k_out, v_in, v_out, (int)P, 0, 30, s));

  VISS_HIP(hipMalloc((void**)&sky->pix, sizeof(PixPack) * Pp));
  VISS_HIP(hipMalloc((void**)&sky->I64, sizeof(double) * Pp));
  VISS_HIP(hipMalloc((void**)&sky->cone, sizeof(float4) * sky->ntile));
  VISS_HIP(hipMalloc((void**)&sky->bcone, sizeof(float4) * sky->nbtile));
  VISS_HIP(hipMalloc((void**)&sky->tile_valid, sizeof(int32_t) * sky->ntile));
  k_pack_sky<<<grid_for(Pp, 256), 256, 0, s>>>(lmn, P, Pp, v_out, sky->pix, sky->I64);
  VISS_LAUNCHED(k_pack_sky);
  k_tile_cones<<<(unsigned)((sky->ntile + 3) / 4), 256, 0, s>>>(sky->pix, sky->ntile, kTile / 64, sky->cone,
                                                                  sky->tile_valid);
  VISS_LAUNCHED(k_tile_cones);
  k_tile_cones<<<(unsigned)((sky->nbtile + 3) / 4), 256, 0, s>>>(sky->pix, sky->nbtile, kBpTile / 64, sky->bcone,
                                                                   nullptr);
  VISS_LAUNCHED(k_tile_cones);
  const int rc = sky_build_pairs(sky, lmn, P, s);
  if (rc != VISS_OK) return rc;
  return sky_build_quads(sky, s);
}

void sky_free(viss_sky* sky) {
  // Plain hipFree synchronises with outstanding work on the buffers.
  if (sky->pix) (void)hipFree(sky->pix);
  if (sky->I64) (void)hipFree(sky->I64);
  if (sky->cone) (void)hipFree(sky->cone);
  if (sky->bcone) (void)hipFree(sky->bcone);
  if (sky->tile_valid) (void)hipFree(sky->tile_valid);
  if (sky->ppack) (void)hipFree(sky->ppack);
  if (sky->pidx) (void)hipFree(sky->pidx);
  if (sky->pcone) (void)hipFree(sky->pcone);
  sky->ppack = nullptr;
  sky->pidx = nullptr;
  sky->pcone = nullptr;
  for (void* p : {(void*)sky->qpack, (void*)sky->qidx, (void*)sky->qz, (void*)sky->qc2, (void*)sky->qcone,
                  (void*)sky->qpatch, (void*)sky->qpcone, (void*)sky->rpack, (void*)sky->ridx, (void*)sky->rcone})
    if (p) (void)hipFree(p);
  sky->qpatch = nullptr;
  sky->qpcone = nullptr;
  sky->qpack = nullptr;
  sky->qidx = nullptr;
  sky->qz = nullptr;
  sky->qc2 = nullptr;
  sky->qrec = false;
  sky->qcone = nullptr;
  sky->rpack = nullptr;
  sky->ridx = nullptr;
  sky->rcone = nullptr;
  sky->ws.release();
  sky->pix = nullptr;
  sky->I64 = nullptr;
  sky->cone = sky->bcone = nullptr;
  sky->tile_valid = nullptr;
}

int sky_set_brightness(viss_sky* sky, const double* I, bool want64, hipStream_t s) {
  k_scatter_brightness<<<grid_for(sky->P_pad, 256), 256, 0, s>>>(sky->pix, sky->P_pad, I, sky->I64, want64 ? 1 : 0);
  VISS_LAUNCHED(k_scatter_brightness);
  if (sky->nptile > 0) {
    k_scatter_pairs<<<grid_for(sky->nptile * 64, 256), 256, 0, s>>>(sky->ppack, sky->pidx, sky->nptile * 64, I);
    VISS_LAUNCHED(k_scatter_pairs);
  }
  if (sky->use_quads) {
    const int64_t nslot = sky->nqtile * kQuadTile;
    k_scatter_quads<<<grid_for(nslot, 256), 256, 0, s>>>(sky->qpack, sky->qidx, nslot, I);
    VISS_LAUNCHED(k_scatter_quads);
    if (sky->nrtile > 0) {
      k_scatter_pairs<<<grid_for(sky->nrtile * 64, 256), 256, 0, s>>>(sky->rpack, sky->ridx, sky->nrtile * 64, I);
      VISS_LAUNCHED(k_scatter_pairs);
    }
  }
  return VISS_OK;
}

}  // namespace viss

using namespace viss;

extern "C" {

int viss_version(void) { return VISS_VERSION; }

const char* viss_status_string(int status) {
  switch (status) {
    case VISS_OK: return "ok";
    case VISS_ERR_INVALID: return "invalid argument";
    case VISS_ERR_HIP: return "HIP error";
    case VISS_ERR_SIZE: return "size out of range";
    case VISS_ERR_NOMEM: return "out of device memory";
    default: return "unknown status";
  }
}

const char* viss_last_error(void) { return g_err; }

int viss_sky_create(const double* lmn, int64_t P, viss_sky** out, void* stream) {
  VISS_REQUIRE(out != nullptr, "viss_sky_create: out is NULL");
  *out = nullptr;
  VISS_REQUIRE(lmn != nullptr, "viss_sky_create: lmn is NULL");
  viss_sky* sky = new (std::nothrow) viss_sky();
  if (!sky) {
    set_error("viss_sky_create: host allocation failed");
    return VISS_ERR_NOMEM;
  }
  const int rc = sky_build(sky, lmn, P, as_stream(stream));
  if (rc != VISS_OK) {
    sky_free(sky);
    delete sky;
    return rc;
  }
  *out = sky;
  return VISS_OK;
}

int viss_sky_destroy(viss_sky* sky) {
  if (!sky) return VISS_OK;
  sky_free(sky);
  delete sky;
  return VISS_OK;
}

int64_t viss_sky_npix(const viss_sky* sky) { return sky ? sky->P : -1; }

int viss_sky_pairs(const viss_sky* sky, int64_t* npair, int64_t* nsingle, int* uses_pairs) {
  VISS_REQUIRE(sky != nullptr, "viss_sky_pairs: sky is NULL");
  if (npair) *npair = sky->npair;
  if (nsingle) *nsingle = sky->nsingle;
  if (uses_pairs) *uses_pairs = sky->use_pairs ? 1 : 0;
  return VISS_OK;
}

int viss_sky_quads(const viss_sky* sky, int64_t* nquad, int64_t* nquad_tiles, int* uses_quads) {
  VISS_REQUIRE(sky != nullptr, "viss_sky_quads: sky is NULL");
  if (nquad) *nquad = sky->nquad;
  if (nquad_tiles) *nquad_tiles = sky->nqtile;
  if (uses_quads) *uses_quads = !sky->use_quads ? 0 : sky->qrec ? 2 : 1;
  return VISS_OK;
}

}  // extern "C"
