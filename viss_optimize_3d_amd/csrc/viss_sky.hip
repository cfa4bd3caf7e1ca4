// Status/error plumbing and the sorted, tiled sky (viss_sky_*).
//
// The reference keeps pixels in file order and lets every visibility thread
// walk all of them (vissgen3d_noblockage.cu:86).  Here the pixel set is sorted
// once along a 3D Morton curve of the unit direction, so that consecutive
// 64-pixel tiles are spatially compact; each tile gets a bounding cone
// (centre, chord radius) that the occultation test uses to skip a tile, or to
// drop the per-pixel test, for a whole wavefront at once.
#include <hipcub/hipcub.hpp>

#include <cstdarg>
#include <cstdio>
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
  VISS_HIP(hipMallocAsync((void**)&k_in, sizeof(uint32_t) * P, s));
  VISS_HIP(hipMallocAsync((void**)&k_out, sizeof(uint32_t) * P, s));
  VISS_HIP(hipMallocAsync((void**)&v_in, sizeof(int32_t) * P, s));
  VISS_HIP(hipMallocAsync((void**)&v_out, sizeof(int32_t) * P, s));
  k_morton_keys<<<grid_for(P, 256), 256, 0, s>>>(lmn, P, k_in, v_in);
  VISS_LAUNCHED(k_morton_keys);
  VISS_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, k_in, k_out, v_in, v_out, (int)P, 0, 30, s));
  VISS_HIP(hipMallocAsync(&tmp, tmp_bytes, s));
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
  VISS_HIP(hipFreeAsync(tmp, s));
  VISS_HIP(hipFreeAsync(k_in, s));
  VISS_HIP(hipFreeAsync(k_out, s));
  VISS_HIP(hipFreeAsync(v_in, s));
  VISS_HIP(hipFreeAsync(v_out, s));
  return VISS_OK;
}

void sky_free(viss_sky* sky) {
  // Plain hipFree synchronises with outstanding work on the buffers.
  if (sky->pix) (void)hipFree(sky->pix);
  if (sky->I64) (void)hipFree(sky->I64);
  if (sky->cone) (void)hipFree(sky->cone);
  if (sky->bcone) (void)hipFree(sky->bcone);
  if (sky->tile_valid) (void)hipFree(sky->tile_valid);
  sky->ws.release();
  sky->pix = nullptr;
  sky->I64 = nullptr;
  sky->cone = sky->bcone = nullptr;
  sky->tile_valid = nullptr;
}

int sky_set_brightness(viss_sky* sky, const double* I, bool want64, hipStream_t s) {
  k_scatter_brightness<<<grid_for(sky->P_pad, 256), 256, 0, s>>>(sky->pix, sky->P_pad, I, sky->I64, want64 ? 1 : 0);
  VISS_LAUNCHED(k_scatter_brightness);
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

}  // extern "C"
