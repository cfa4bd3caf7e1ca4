// Internal declarations shared by the HIP translation units of libviss_hip.so.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "viss.h"

namespace viss {

constexpr double kPi = 3.14159265358979323846264338327950288;

// Thread-local error message (viss_last_error) and helpers returning a status.
void set_error(const char* fmt, ...) __attribute__((format(printf, 1, 2)));
int hip_fail(hipError_t e, const char* what, const char* file, int line);

#define VISS_HIP(call)                                                          \
  do {                                                                          \
    hipError_t viss_e_ = (call);                                                \
    if (viss_e_ != hipSuccess) return ::viss::hip_fail(viss_e_, #call, __FILE__, __LINE__); \
  } while (0)

#define VISS_REQUIRE(cond, ...)          \
  do {                                   \
    if (!(cond)) {                       \
      ::viss::set_error(__VA_ARGS__);    \
      return VISS_ERR_INVALID;           \
    }                                    \
  } while (0)

// Launch-error check after a kernel launch.
#define VISS_LAUNCHED(name) VISS_HIP(hipGetLastError())

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// ---------------------------------------------------------------- packing
// One pixel of the sorted sky: direction (FP64), FP32 brightness slot written
// per forward call, and the caller's index (-1 for padding).
struct PixPack {
  double l, m, n;
  float I;
  int32_t idx;
};
static_assert(sizeof(PixPack) == 32, "PixPack is 32 B");

// Per-visibility data read wave-uniformly by the back-projection.
// An antipodal pixel pair (a, b) with s_b = -s_a, or a single pixel a (Ib = 0).
// b.s_b = -b.s_a, so one phase and one sin/cos serve both pixels.
struct PairPack {
  double l, m, n;  // direction of pixel a (pixel b is at -(l, m, n))
  float A, D;      // Ia + Ib and Ia - Ib (both pixels seen)
  float Ia, Ib;    // brightness of a and b (one side seen, or the per-pixel mask)
};
static_assert(sizeof(PairPack) == 40, "PairPack is 40 B");

// Quad of the equatorial-mirror layout: pixels a = (x, y, z), b = (-x, -y, z)
// on ring z and c = (x, y, -z), d = (-x, -y, -z) on ring -z (z is the tile's).
// With A = u x + v y and C = w z their phases are A + C, -A + C, A - C, -A - C,
// so one sin/cos of A serves all four; Az/Dz = Ia +- Ib, Am/Dm = Ic +- Id.
struct QuadPack {
  double x, y;
  float Az, Dz, Am, Dm;
  float Ia, Ib, Ic, Id;
};
static_assert(sizeof(QuadPack) == 48, "QuadPack is 48 B");
constexpr int kQuadTile = 16;  // quads per tile, all on one ring pair

struct VisHot {  // mixed precision: phase in FP64, weighted visibility in FP32
  double u, v, w;
  float yr, yi;
};
static_assert(sizeof(VisHot) == 32, "VisHot is 32 B");

// All-FP64 path.  x is real, so a row adds Re(y e^{2 pi i t}) =
// ya cos(2 pi (t + ph)) with y = ya e^{2 pi i ph}: one FP64 cospi per
// (row, pixel) instead of sincospi.
struct VisHot64 {
  double u, v, w, yr, yi;
  double ph;  // arg(y) / (2 pi), revolutions in (-1/2, 1/2]
  double ya;  // |y|
  double pad;
};
static_assert(sizeof(VisHot64) == 64, "VisHot64 is 64 B");

struct VisMask {  // exact occultation test: s.r_a >= ta && s.r_b >= tb (FP64)
  double xa, ya, za, ta, xb, yb, zb, tb;
};
static_assert(sizeof(VisMask) == 64, "VisMask is 64 B");

struct VisCls {  // tile classification: unit vectors and unit cuts (FP32)
  float ax, ay, az, ca, bx, by, bz, cb;
};
static_assert(sizeof(VisCls) == 32, "VisCls is 32 B");

// Tile classification margin (FP32 rounding of unit vectors, cuts and cone
// data is below 1e-6; the margin keeps every "all visible" / "none visible"
// decision consistent with the exact FP64 per-pixel test).
constexpr float kClsEps = 1e-5f;

constexpr int kTile = 64;       // forward tile (pixels)
#ifndef VISS_BP_PIX_PER_LANE
#define VISS_BP_PIX_PER_LANE 2
#endif
constexpr int kBpPixPerLane = VISS_BP_PIX_PER_LANE;  // back-projection pixels per lane (measured best of 1, 2, 4, 8)
constexpr int kBpTile = 64 * kBpPixPerLane;
constexpr int kPadTo = kBpTile;  // sky padding granularity

// Index of the calling wavefront in its block.  threadIdx.x >> 6 is the same
// on every lane of a wavefront, but the compiler cannot prove it;
// readfirstlane makes it (and every address derived from it) uniform, so
// per-wavefront data is fetched with scalar loads into SGPRs instead of
// per-lane vector loads that would each hold a copy in VGPRs.
__device__ __forceinline__ int wave_in_block() { return __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)); }

// The lane's index in its wavefront, formed afresh (two VALU instructions)
// where it is used.  The asm is volatile so that the compiler neither hoists
// it nor merges it with another copy: a lane index computed once and kept
// live across a kernel is one more VGPR, and in a kernel held to 64 VGPRs it
// is the one spilled to scratch.
__device__ __forceinline__ int lane_fresh() {
  int l;
  asm volatile("v_mbcnt_lo_u32_b32 %0, -1, 0\n\tv_mbcnt_hi_u32_b32 %0, -1, %0" : "=v"(l));
  return l;
}

// Occultation threshold t(r): s visible from r iff s.r >= t
// (vissgen3d_noblockage.cu:88-94,240-243 in dot-product form).  Written with
// explicit FMAs so that no compiler contraction changes its rounding: the
// oracle evaluates the same operations, so both get the same t bit for bit.
__host__ __device__ inline double mask_threshold(double x, double y, double z, double R_moon, double orbit_radius) {
  const double r2 = fma(x, x, fma(y, y, z * z));
  if (orbit_radius > 0.0) {
    const double q = R_moon / orbit_radius;
    return -sqrt(fma(-q, q, 1.0)) * sqrt(r2);
  }
  const double d = fma(-R_moon, R_moon, r2);
  return d > 0.0 ? -sqrt(d) : 0.0;
}

// The occultation test's dot product s.r in the one evaluation order every
// kernel and the oracle use: fma(sx, rx, fma(sy, ry, sz * rz)).  Kernels that
// serve the antipode -s of a pixel negate it (exact), and the quad kernels
// pass sz * rz of the ring (z) or of its mirror ring (-z, whose product is the
// exact negation).
__host__ __device__ inline double occ_dot(double sx, double sy, double szrz, double rx, double ry) {
  return fma(sx, rx, fma(sy, ry, szrz));
}

// Device buffers growable in stream order.
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  int reserve(size_t need, hipStream_t s);
  void release();
};

}  // namespace viss

struct viss_sky {
  int device = -1;
  int64_t P = 0;       // pixels
  int64_t P_pad = 0;   // padded to kPadTo
  int64_t ntile = 0;   // P_pad / kTile
  int64_t nbtile = 0;  // P_pad / kBpTile
  viss::PixPack* pix = nullptr;  // [P_pad] sorted
  double* I64 = nullptr;         // [P_pad] FP64 brightness (all-FP64 forward)
  float4* cone = nullptr;        // [ntile] bounding cones of kTile-pixel tiles
  float4* bcone = nullptr;       // [nbtile] bounding cones of kBpTile-pixel tiles
  int32_t* tile_valid = nullptr; // [ntile] number of real pixels per tile
  // Antipodal-pair layout (mixed-precision per-row kernels): pairs sorted along
  // the Morton curve of pixel a, then the unpaired pixels, each section padded
  // to whole 64-entry tiles.
  int64_t npair = 0;             // antipodal pairs found
  int64_t nsingle = 0;           // pixels without a partner
  int64_t nptile = 0;            // 64-entry tiles of the pair layout
  bool use_pairs = false;        // at least half of the pixels are paired
  viss::PairPack* ppack = nullptr;  // [nptile * 64]
  int2* pidx = nullptr;             // [nptile * 64] caller indices of (a, b); -1 if absent
  float4* pcone = nullptr;          // [nptile] bounding cones of the a pixels of each tile
  // Equatorial-mirror quads (mixed-precision per-row forward): two antipodal
  // pairs whose a pixels are (x, y, z) and (-x, -y, z).  Tiles of kQuadTile
  // quads sharing one z, sorted by z then azimuth; the pairs and singles left
  // over form a second pair layout (rest).
  int64_t nquad = 0;                // quads found
  int64_t nqtile = 0;               // kQuadTile-quad tiles
  bool use_quads = false;           // quads cover half the pixels, tiles at least 2/3 full
  viss::QuadPack* qpack = nullptr;  // [nqtile * kQuadTile]
  int4* qidx = nullptr;             // [nqtile * kQuadTile] caller indices of (a, b, c, d); -1 if absent
  double* qz = nullptr;             // [nqtile] z of each tile
  double* qc2 = nullptr;            // [nqtile] 2 cos of the azimuth step of each tile's quads
  bool qrec = false;                // every tile's phase follows the recurrence over qc2
  float4* qcone = nullptr;          // [2 nqtile] cones of the a and b pixels of each tile
  int64_t nqpatch = 0;              // back-projection patches: 4 tiles (64 quads), one per ring of a band
  int4* qpatch = nullptr;           // [nqpatch] tile of each 16-lane group; -1 if absent
  float4* qpcone = nullptr;         // [2 nqpatch] cones of the a and b pixels of each patch
  int64_t nrtile = 0;               // 64-entry tiles of the rest layout
  viss::PairPack* rpack = nullptr;  // [nrtile * 64]
  int2* ridx = nullptr;             // [nrtile * 64]
  float4* rcone = nullptr;          // [nrtile]
  viss::DevBuf ws;               // per-call visibility packs
};

namespace viss {
// A chunk of back-projection rows [k0, k1) of group g (see sky_backproject_ws).
struct RowChunk {
  int64_t k0, k1, g, pad;
};
static_assert(sizeof(RowChunk) == 32, "RowChunk is 32 B");

int sky_build(viss_sky* sky, const double* lmn, int64_t P, hipStream_t s);
// Bytes of the row records the back-projection packs for K rows with these flags.
size_t bp_workspace_bytes(int64_t K, int flags);
// The back-projection over a workspace of bp_workspace_bytes(K, flags) bytes at
// ws (256-B aligned), arguments already checked.  chunks (device, nchunks
// entries) replaces the fixed row chunks: block row c takes rows
// [chunks[c].k0, chunks[c].k1) and adds into x + chunks[c].g * img_stride.
// Per-group weights of a batch over weight chunks wch (device; pad != 0 on a
// group's first chunk): histograms mb [G nr], tables dcf [G (nr + 1)].
int batch_weights(const float* bll, const double* w, const RowChunk* wch, int64_t nwch, int32_t nr, int flags,
                  int32_t* gs, uint32_t* mb, double* dcf, double* dg, double* wt, int64_t* n_out_of_range,
                  hipStream_t s);
int sky_backproject_ws(viss_sky* sky, const double* uvw, const double* xyza, const double* xyzb, const double* y,
                       const double* wt, int64_t K, const viss_mask_geometry* geom, int flags, double* x, char* ws,
                       hipStream_t s, const RowChunk* chunks, int64_t nchunks, int64_t img_stride);
void sky_free(viss_sky* sky);
int sky_set_brightness(viss_sky* sky, const double* I, bool want64, hipStream_t s);
int check_device(const viss_sky* sky);
}  // namespace viss
