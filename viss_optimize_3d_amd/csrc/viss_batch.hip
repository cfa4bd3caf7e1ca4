// Batched forward over many visibility groups (snapshot groups x frequencies)
// that share one sky, in ONE launch over a work table, replayable as a hipGraph.
//
// The reference runs one day per loop iteration, one launch per kernel, with a
// device sync after each (vissgen3d_noblockage.cu:270-557).  With occultation
// each group only sees part of the sky, unevenly: here every group gets a
// compacted list of the 64-pixel tiles that any of its satellites can see
// (a conservative cone test of the tile against the bounding cone of the
// group's satellite directions), and a work table of (group, 64 rows, range
// of its tile list) items is covered by one launch, one wavefront per item.
// The per-item tile-range length is chosen per (visible-tile-count,
// row-count) bucket by an autotuner that times the
// candidates on the device (viss_batch_autotune), and the table can be saved
// and reloaded.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <new>
#include <vector>

#include "viss_kernels.h"

namespace viss {

struct WorkItem {
  int64_t row0;
  int32_t g, nrows, j0, j1;
  int32_t pad0, pad1;
};
static_assert(sizeof(WorkItem) == 32, "WorkItem is 32 B");

// Bounding cone of the satellite directions of each group and the lowest unit
// cut among them.  One block per group.
__global__ __launch_bounds__(256) void k_group_bounds(const double* __restrict__ xyza, const double* __restrict__ xyzb,
                                                      int64_t Ktot, const int64_t* __restrict__ offs, double R_moon,
                                                      double orbit_r, float4* __restrict__ gcone,
                                                      float* __restrict__ gcut) {
  __shared__ double red[4][4];
  __shared__ double dir[4];
  const int g = blockIdx.x, lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int64_t o0 = offs[g], o1 = offs[g + 1];
  double sx = 0, sy = 0, sz = 0, cmin = 2.0;
  for (int64_t k = o0 + threadIdx.x; k < o1; k += 256) {
    VisMask m;
    VisCls c;
    make_mask_records(xyza[k], xyza[Ktot + k], xyza[2 * Ktot + k], xyzb[k], xyzb[Ktot + k], xyzb[2 * Ktot + k],
                      R_moon, orbit_r, m, c);
    sx += (double)c.ax + (double)c.bx;
    sy += (double)c.ay + (double)c.by;
    sz += (double)c.az + (double)c.bz;
    cmin = fmin(cmin, fmin((double)c.ca, (double)c.cb));
  }
  for (int o = 32; o > 0; o >>= 1) {
    sx += __shfl_xor(sx, o, 64);
    sy += __shfl_xor(sy, o, 64);
    sz += __shfl_xor(sz, o, 64);
    cmin = fmin(cmin, __shfl_xor(cmin, o, 64));
  }
  if (lane == 0) {
    red[wv][0] = sx;
    red[wv][1] = sy;
    red[wv][2] = sz;
    red[wv][3] = cmin;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double X = 0, Y = 0, Z = 0, C = 2.0;
    for (int i = 0; i < 4; ++i) {
      X += red[i][0];
      Y += red[i][1];
      Z += red[i][2];
      C = fmin(C, red[i][3]);
    }
    const double nrm = sqrt(X * X + Y * Y + Z * Z);
    dir[0] = nrm > 1e-9 ? X / nrm : 0.0;
    dir[1] = nrm > 1e-9 ? Y / nrm : 0.0;
    dir[2] = nrm > 1e-9 ? Z / nrm : 1.0;
    dir[3] = nrm > 1e-9 ? C : -3.0;  // degenerate group: everything possibly visible
  }
  __syncthreads();
  double mx = 0.0;
  for (int64_t k = o0 + threadIdx.x; k < o1; k += 256) {
    VisMask m;
    VisCls c;
    make_mask_records(xyza[k], xyza[Ktot + k], xyza[2 * Ktot + k], xyzb[k], xyzb[Ktot + k], xyzb[2 * Ktot + k],
                      R_moon, orbit_r, m, c);
    const double ax = c.ax - dir[0], ay = c.ay - dir[1], az = c.az - dir[2];
    const double bx = c.bx - dir[0], by = c.by - dir[1], bz = c.bz - dir[2];
    mx = fmax(mx, fmax(sqrt(ax * ax + ay * ay + az * az), sqrt(bx * bx + by * by + bz * bz)));
  }
  for (int o = 32; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o, 64));
  __syncthreads();
  if (lane == 0) red[wv][0] = mx;
  __syncthreads();
  if (threadIdx.x == 0) {
    double M = 0;
    for (int i = 0; i < 4; ++i) M = fmax(M, red[i][0]);
    gcone[g] = make_float4((float)dir[0], (float)dir[1], (float)dir[2], (float)(M * (1.0 + 1e-6) + 1e-6));
    gcut[g] = (float)dir[3];
  }
}

// Compacted list of tiles some satellite of the group may see.  One block per group.
// both_sides: tiles of antipodal pairs, kept when the a pixels (cone c) or the
// b pixels (cone -c) may be seen.
__global__ __launch_bounds__(256) void k_group_tiles(const float4* __restrict__ cone, int64_t ntile, int both_sides,
                                                     const float4* __restrict__ gcone, const float* __restrict__ gcut,
                                                     int32_t* __restrict__ tiles, int32_t* __restrict__ count) {
  __shared__ int wcount[4];
  const int g = blockIdx.x, lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const float4 gc = gcone[g];
  const float cut = gcut[g];
  int32_t* out = tiles + (int64_t)g * ntile;
  int running = 0;
  for (int64_t base = 0; base < ntile; base += 256) {
    const int64_t t = base + threadIdx.x;
    bool keep = false;
    if (t < ntile) {
      const float4 c = cone[t];
      const float d = fmaf(c.x, gc.x, fmaf(c.y, gc.y, c.z * gc.z));
      keep = d + c.w + gc.w >= cut - kClsEps;
      if (both_sides) keep |= -d + c.w + gc.w >= cut - kClsEps;
    }
    const uint64_t bal = wave_ballot(keep);
    if (lane == 0) wcount[wv] = __popcll(bal);
    __syncthreads();
    int before = running;
    for (int i = 0; i < wv; ++i) before += wcount[i];
    const int within = __popcll(bal & ((1ull << lane) - 1ull));
    if (keep) out[before + within] = (int32_t)t;
    const int total = wcount[0] + wcount[1] + wcount[2] + wcount[3];
    __syncthreads();
    running += total;
  }
  if (threadIdx.x == 0) count[g] = running;
}

// Tiles of the quad layout, then the rest tiles, as one index space: two
// cones per tile (quad tiles: a and b pixels; rest tiles: the a cone twice),
// each standing for itself and its antipode.
__global__ void k_batch_cones(const float4* __restrict__ qcone, int64_t nqtile, const float4* __restrict__ rcone,
                              int64_t nrtile, float4* __restrict__ out) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < nqtile + nrtile;
       t += (int64_t)gridDim.x * blockDim.x) {
    const float4 c1 = t < nqtile ? qcone[2 * t] : rcone[t - nqtile];
    const float4 c2 = t < nqtile ? qcone[2 * t + 1] : rcone[t - nqtile];
    out[2 * t] = c1;
    out[2 * t + 1] = c2;
  }
}

// Compacted list of the tiles (two cones each, both sides) some satellite of
// the group may see.  One block per group.
__global__ __launch_bounds__(256) void k_group_tiles2(const float4* __restrict__ cone2, int64_t ntile,
                                                      const float4* __restrict__ gcone, const float* __restrict__ gcut,
                                                      int32_t* __restrict__ tiles, int32_t* __restrict__ count) {
  __shared__ int wcount[4];
  const int g = blockIdx.x, lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const float4 gc = gcone[g];
  const float cut = gcut[g];
  int32_t* out = tiles + (int64_t)g * ntile;
  int running = 0;
  for (int64_t base = 0; base < ntile; base += 256) {
    const int64_t t = base + threadIdx.x;
    bool keep = false;
    if (t < ntile) {
      for (int h = 0; h < 2; ++h) {
        const float4 c = cone2[2 * t + h];
        const float d = fmaf(c.x, gc.x, fmaf(c.y, gc.y, c.z * gc.z));
        keep |= (d + c.w + gc.w >= cut - kClsEps) | (-d + c.w + gc.w >= cut - kClsEps);
      }
    }
    const uint64_t bal = wave_ballot(keep);
    if (lane == 0) wcount[wv] = __popcll(bal);
    __syncthreads();
    int before = running;
    for (int i = 0; i < wv; ++i) before += wcount[i];
    const int within = __popcll(bal & ((1ull << lane) - 1ull));
    if (keep) out[before + within] = (int32_t)t;
    const int total = wcount[0] + wcount[1] + wcount[2] + wcount[3];
    __syncthreads();
    running += total;
  }
  if (threadIdx.x == 0) count[g] = running;
}

// One launch over the whole work table, one wavefront per item (up to 64
// rows, one per lane).  The hardware dispatcher is the work queue.  A
// persistent loop pulling items through an atomic counter measured 2.6x
// slower: there the FP64 atomics into V that end one item precede the loads
// of the next, and the compiler emitted per-lane vector loads for the sky
// data instead of scalar loads (SGPR operands).
template <bool MASK, bool F64>
__global__ __launch_bounds__(256) void k_forward_batch(const SkyView sv, const int32_t* __restrict__ tiles,
                                                       int64_t ntile, const WorkItem* __restrict__ items,
                                                       int64_t n_items, const double* __restrict__ uvw,
                                                       const double* __restrict__ xyza,
                                                       const double* __restrict__ xyzb, int64_t Ktot, double R_moon,
                                                       double orbit_r, int track, double* __restrict__ V) {
  const int lane = threadIdx.x & 63;
  const int64_t it = (int64_t)blockIdx.x * 4 + wave_in_block();
  if (it >= n_items) return;  // wave-uniform
  const WorkItem w = items[it];
  const int64_t k = w.row0 + lane;
  const FwdRow r = load_fwd_row<MASK>(uvw, xyza, xyzb, Ktot, k, lane < w.nrows, R_moon, orbit_r, track);
  double Sr = 0.0, Si = 0.0;
  forward_tiles<MASK, F64>(sv, tiles + (int64_t)w.g * ntile, w.j0, w.j1, r, Sr, Si);
  if (r.act) forward_store(V, k, Ktot, 0, Sr, Si);
}

// The same work table over a paired sky (mixed precision): items are 64 rows x
// a range of the group's list of 64-pair tiles.
template <bool MASK>
__global__ __launch_bounds__(256) void k_forward_batch_pairs(const PairPack* __restrict__ pp,
                                                             const float4* __restrict__ pcone,
                                                             const int32_t* __restrict__ tiles, int64_t ntile,
                                                             const WorkItem* __restrict__ items, int64_t n_items,
                                                             const double* __restrict__ uvw,
                                                             const double* __restrict__ xyza,
                                                             const double* __restrict__ xyzb, int64_t Ktot,
                                                             double R_moon, double orbit_r, int track,
                                                             double* __restrict__ V) {
  const int lane = threadIdx.x & 63;
  const int64_t it = (int64_t)blockIdx.x * 4 + wave_in_block();
  if (it >= n_items) return;  // wave-uniform
  const WorkItem w = items[it];
  const int64_t k = w.row0 + lane;
  const FwdRow r = load_fwd_row<MASK>(uvw, xyza, xyzb, Ktot, k, lane < w.nrows, R_moon, orbit_r, track);
  double Sr = 0.0, Si = 0.0;
  forward_pair_tiles<MASK>(pp, pcone, tiles + (int64_t)w.g * ntile, w.j0, w.j1, r, Sr, Si);
  if (r.act) forward_pair_store(V, k, Ktot, 0, r.c0, Sr, Si);
}

// The same work table over the quad layout (mixed precision): list entries
// below nqtile are 16-quad tiles, the others tiles of the rest layout.  Both
// tile loops are inlined here, and left to itself the register allocator
// gives the kernel over 100 VGPRs (4 wavefronts per SIMD).  Held to 7
// wavefronts it fits in 66 without spilling (8 would spill to scratch): at
// 4.48e6 rows x 196,608 pixels, 102.9 ms per pass against 128 ms at 4.
// Quad tiles join ring-pair runs as in forward_quad_tiles, but a run is cut
// by list position: it holds the evaluated tiles of one ring pair among
// positions [8 m, 8 m + 8).  Items start at multiples of 8 (build_items), so
// no run crosses an item and the FP32 sums are the same for every work
// table: results differ between tables (autotuned or not) by the order of
// the FP64 sums only.
template <bool MASK, bool REC>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(7))) void k_forward_batch_quads(const QuadPack* __restrict__ qp,
                                                             const double* __restrict__ qz,
                                                             const double* __restrict__ qc2,
                                                             const float4* __restrict__ qcone, int64_t nqtile,
                                                             const PairPack* __restrict__ rp,
                                                             const float4* __restrict__ rcone,
                                                             const int32_t* __restrict__ tiles, int64_t ntile,
                                                             const WorkItem* __restrict__ items, int64_t n_items,
                                                             const double* __restrict__ uvw,
                                                             const double* __restrict__ xyza,
                                                             const double* __restrict__ xyzb, int64_t Ktot,
                                                             double R_moon, double orbit_r, int track,
                                                             double* __restrict__ V) {
  const int lane = threadIdx.x & 63;
  const int64_t it = (int64_t)blockIdx.x * 4 + wave_in_block();
  if (it >= n_items) return;  // wave-uniform
  const WorkItem w = items[it];
  const int64_t k = w.row0 + lane;
  const FwdRow r = load_fwd_row<MASK>(uvw, xyza, xyzb, Ktot, k, lane < w.nrows, R_moon, orbit_r, track);
  const int32_t* __restrict__ list = tiles + (int64_t)w.g * ntile;
  const uint64_t actm = wave_ballot(r.act);
  if (MASK && actm == 0) return;  // wave-uniform: no row in this item
  WaveCls wb{};
  if (MASK) wb = wave_cls(r.cl, r.act);
  double Sr = 0.0, Si = 0.0;
  RingRun rr;
  int64_t run_blk = -1;  // list position / 8 of the open run
  // list entries 64 at a time: each lane classifies one quad tile against the
  // item's bounds (as forward_quad_tiles), then the wavefront walks them.
  for (int64_t jb = w.j0; jb < w.j1; jb += 64) {
    const int64_t je = min(w.j1, jb + 64);
    int code = kTileExact;
    if (MASK && jb + lane < je) {
      const int64_t e = list[jb + lane];
      if (e < nqtile) code = quad_tile_code(qcone[2 * e], qcone[2 * e + 1], wb);
    }
    for (int64_t j = jb; j < je; ++j) {
      const int64_t t = list[j];  // wave-uniform
      if (t < nqtile) {
        int mz = kRingFF, mm = kRingFF;
        if (MASK) {
          quad_tile_modes(qcone, t, __builtin_amdgcn_readlane(code, (int)(j - jb)), r, actm, mz, mm);
          if (mz == kRingNone && mm == kRingNone) continue;
        }
        const double z = qz[t];
        if ((uint64_t)__double_as_longlong(z) != rr.zbits || (j >> 3) != run_blk) {  // wave-uniform
          if (rr.n > 0) rr.close(r, Sr, Si);
          rr.open(z);
          run_blk = j >> 3;
        }
        quad_tile_run<REC>(qp, qc2, t, mz, mm, r, rr);
        ++rr.n;
      } else {
        forward_pair_tiles<MASK>(rp, rcone, nullptr, t - nqtile, t - nqtile + 1, r, Sr, Si);
      }
    }
  }
  if (rr.n > 0) rr.close(r, Sr, Si);
  if (r.act) forward_pair_store(V, k, Ktot, 0, r.c0, Sr, Si);
}

}  // namespace viss

using namespace viss;

struct viss_batch {
  viss_sky* sky = nullptr;
  int device = -1;
  int64_t G = 0, Ktot = 0;
  std::vector<int64_t> offs;
  const double *uvw = nullptr, *xa = nullptr, *xb = nullptr;
  double R = 0, orb = 0;
  int flags = 0;
  bool pairs = false;   // mixed precision over a paired sky: tiles of 64 antipodal pairs
  bool quads = false;   // mixed precision over the quad layout: quad tiles, then rest tiles
  float4* tcone2 = nullptr;  // [2 ntile] cones of the quad layout's tiles (quads)
  int64_t ntile = 0;    // tiles of the layout in use (pairs: sky->nptile, else sky->ntile)
  int64_t* d_offs = nullptr;
  float4* gcone = nullptr;
  float* gcut = nullptr;
  int32_t* tiles = nullptr;
  int32_t* d_count = nullptr;
  std::vector<int32_t> ntiles_g;
  std::map<int, int> tune;  // bucket -> tiles per item
  WorkItem* items = nullptr;
  int64_t n_items = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  const double* g_I = nullptr;
  double* g_V = nullptr;
  hipStream_t g_stream = nullptr;
  // back-projection and weights over the groups' rows
  viss::RowChunk* bp_chunks = nullptr;  // row chunks within groups (back-projection)
  int64_t n_bp_chunks = 0;
  viss::RowChunk* w_chunks = nullptr;   // row chunks within groups, one at least per group (weights)
  int64_t n_w_chunks = 0;
  viss::DevBuf bp_ws;                   // row records of the back-projection (sized on first use)
  hipGraph_t bp_graph = nullptr;
  hipGraphExec_t bp_exec = nullptr;
  const double *bg_y = nullptr, *bg_wt = nullptr;
  double* bg_x = nullptr;
  int bg_flags = -1, bg_per_group = -1;
  hipStream_t bg_stream = nullptr;
};

namespace viss {

static int bucket_of(int64_t nt, int64_t rows) {
  int a = 0, b = 0;
  while ((1ll << (a + 1)) <= nt) ++a;
  while ((1ll << (b + 1)) <= rows) ++b;
  return a * 64 + b;
}

static int default_tpc(int64_t nt) {
  int64_t t = nt / 8;
  if (t < 16) t = 16;
  if (t > 256) t = 256;
  return (int)t;
}

static void build_items(const viss_batch* b, const std::vector<int>& groups, std::vector<WorkItem>& out,
                        const std::map<int, int>& tune, int force_tpc) {
  out.clear();
  for (int g : groups) {
    const int64_t rows = b->offs[g + 1] - b->offs[g];
    const int64_t nt = b->ntiles_g[g];
    if (rows <= 0 || nt <= 0) continue;
    int tpc = force_tpc;
    if (tpc <= 0) {
      auto it = tune.find(bucket_of(nt, rows));
      tpc = it != tune.end() ? it->second : default_tpc(nt);
    }
    tpc = (int)std::min<int64_t>(((int64_t)tpc + 7) / 8 * 8, 1 << 30);  // items start at multiples of 8
    for (int64_t r0 = 0; r0 < rows; r0 += 64)
      for (int64_t j = 0; j < nt; j += tpc) {
        WorkItem w;
        w.row0 = b->offs[g] + r0;
        w.g = g;
        w.nrows = (int32_t)std::min<int64_t>(64, rows - r0);
        w.j0 = (int32_t)j;
        w.j1 = (int32_t)std::min<int64_t>(nt, j + tpc);
        w.pad0 = w.pad1 = 0;
        out.push_back(w);
      }
  }
  // Items that run at the same time should read the same tiles: order by the
  // position in the (tile-sorted) lists, then by group: blocks are dispatched
  // in index order, so the launch sweeps the sky once and items resident at
  // the same time share L2.
  std::stable_sort(out.begin(), out.end(), [](const WorkItem& a, const WorkItem& c) {
    if (a.j0 != c.j0) return a.j0 < c.j0;
    return a.g < c.g;
  });
}

static int upload_items(viss_batch* b, const std::vector<WorkItem>& v, hipStream_t s) {
  if (b->items) VISS_HIP(hipFree(b->items));
  b->items = nullptr;
  VISS_REQUIRE(v.size() < (size_t)INT32_MAX / 2, "viss_batch: work table too large (%zu items)", v.size());
  b->n_items = (int64_t)v.size();
  if (!v.empty()) {
    VISS_HIP(hipMalloc((void**)&b->items, v.size() * sizeof(WorkItem)));
    VISS_HIP(hipMemcpyAsync(b->items, v.data(), v.size() * sizeof(WorkItem), hipMemcpyHostToDevice, s));
    VISS_HIP(hipStreamSynchronize(s));
  }
  // invalidate a captured graph (it holds the old table)
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
  return VISS_OK;
}

static int launch_batch(viss_batch* b, const WorkItem* items, int64_t n_items, double* V, hipStream_t s) {
  viss_sky* sky = b->sky;
  if (n_items == 0) return VISS_OK;
  const int64_t blocks = (n_items + 3) / 4;
  VISS_REQUIRE(blocks <= INT32_MAX, "viss_batch: work table too large");
  const unsigned grid = (unsigned)blocks;
  const int track = (b->flags & VISS_PHASE_TRACK) ? 1 : 0;
  const bool mask = b->flags & VISS_MASK, f64 = b->flags & VISS_PREC_FP64;
  if (b->quads) {
    auto* kern = mask ? (sky->qrec ? k_forward_batch_quads<true, true> : k_forward_batch_quads<true, false>)
                      : (sky->qrec ? k_forward_batch_quads<false, true> : k_forward_batch_quads<false, false>);
    kern<<<grid, 256, 0, s>>>(sky->qpack, sky->qz, sky->qc2, sky->qcone, sky->nqtile, sky->rpack, sky->rcone,
                              b->tiles, b->ntile, items, n_items, b->uvw, b->xa, b->xb, b->Ktot, b->R, b->orb, track,
                              V);
    VISS_LAUNCHED(k_forward_batch_quads);
    return VISS_OK;
  }
  if (b->pairs) {
    if (mask)
      k_forward_batch_pairs<true><<<grid, 256, 0, s>>>(sky->ppack, sky->pcone, b->tiles, b->ntile, items, n_items,
                                                       b->uvw, b->xa, b->xb, b->Ktot, b->R, b->orb, track, V);
    else
      k_forward_batch_pairs<false><<<grid, 256, 0, s>>>(sky->ppack, sky->pcone, b->tiles, b->ntile, items, n_items,
                                                        b->uvw, b->xa, b->xb, b->Ktot, b->R, b->orb, track, V);
    VISS_LAUNCHED(k_forward_batch_pairs);
    return VISS_OK;
  }
#define VISS_FB(M, F)                                                                                             \
  k_forward_batch<M, F><<<grid, 256, 0, s>>>(sky_view(sky), b->tiles, b->ntile, items, n_items, b->uvw, b->xa, \
                                             b->xb, b->Ktot, b->R, b->orb, track, V)
  if (mask) {
    if (f64) VISS_FB(true, true); else VISS_FB(true, false);
  } else {
    if (f64) VISS_FB(false, true); else VISS_FB(false, false);
  }
#undef VISS_FB
  VISS_LAUNCHED(k_forward_batch);
  return VISS_OK;
}

static void batch_free(viss_batch* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  if (b->bp_exec) (void)hipGraphExecDestroy(b->bp_exec);
  if (b->bp_graph) (void)hipGraphDestroy(b->bp_graph);
  if (b->bp_chunks) (void)hipFree(b->bp_chunks);
  if (b->w_chunks) (void)hipFree(b->w_chunks);
  b->bp_ws.release();
  if (b->d_offs) (void)hipFree(b->d_offs);
  if (b->gcone) (void)hipFree(b->gcone);
  if (b->gcut) (void)hipFree(b->gcut);
  if (b->tiles) (void)hipFree(b->tiles);
  if (b->tcone2) (void)hipFree(b->tcone2);
  if (b->d_count) (void)hipFree(b->d_count);
  if (b->items) (void)hipFree(b->items);
}

// Row chunks that never cross a group: back-projection chunks of at most cr
// rows (cr doubled until the table fits a grid's y extent), and weight chunks
// of at most 32768 rows with at least one (possibly empty) chunk per group.
static int build_chunks(viss_batch* b, hipStream_t s) {
  const int64_t G = b->G;
  std::vector<RowChunk> bp, wc;
  for (int64_t cr = 4096;; cr *= 2) {
    bp.clear();
    for (int64_t g = 0; g < G; ++g)
      for (int64_t k0 = b->offs[g]; k0 < b->offs[g + 1]; k0 += cr)
        bp.push_back(RowChunk{k0, std::min(b->offs[g + 1], k0 + cr), g, 0});
    if ((int64_t)bp.size() <= 65535 || cr > ((int64_t)1 << 40)) break;
  }
  VISS_REQUIRE((int64_t)bp.size() <= 65535, "viss_batch: %zu back-projection row chunks", bp.size());
  constexpr int64_t kWeightChunk = 32768;
  for (int64_t g = 0; g < G; ++g) {
    int64_t k0 = b->offs[g];
    do {
      const int64_t k1 = std::min(b->offs[g + 1], k0 + kWeightChunk);
      wc.push_back(RowChunk{k0, k1, g, k0 == b->offs[g] ? 1 : 0});
      k0 = k1;
    } while (k0 < b->offs[g + 1]);
  }
  b->n_bp_chunks = (int64_t)bp.size();
  b->n_w_chunks = (int64_t)wc.size();
  if (!bp.empty()) {
    VISS_HIP(hipMalloc((void**)&b->bp_chunks, bp.size() * sizeof(RowChunk)));
    VISS_HIP(hipMemcpyAsync(b->bp_chunks, bp.data(), bp.size() * sizeof(RowChunk), hipMemcpyHostToDevice, s));
  }
  VISS_HIP(hipMalloc((void**)&b->w_chunks, wc.size() * sizeof(RowChunk)));
  VISS_HIP(hipMemcpyAsync(b->w_chunks, wc.data(), wc.size() * sizeof(RowChunk), hipMemcpyHostToDevice, s));
  VISS_HIP(hipStreamSynchronize(s));  // the host tables go out of scope
  return VISS_OK;
}

static int batch_build(viss_batch* b, hipStream_t s) {
  viss_sky* sky = b->sky;
  const int64_t G = b->G;
  {
    const int rc = build_chunks(b, s);
    if (rc != VISS_OK) return rc;
  }
  VISS_HIP(hipMalloc((void**)&b->d_offs, sizeof(int64_t) * (G + 1)));
  VISS_HIP(hipMalloc((void**)&b->gcone, sizeof(float4) * G));
  VISS_HIP(hipMalloc((void**)&b->gcut, sizeof(float) * G));
  b->pairs = !(b->flags & VISS_PREC_FP64) && sky->use_pairs;
  b->quads = b->pairs && sky->use_quads;
  b->ntile = b->quads ? sky->nqtile + sky->nrtile : (b->pairs ? sky->nptile : sky->ntile);
  if (b->quads) {
    VISS_HIP(hipMalloc((void**)&b->tcone2, sizeof(float4) * 2 * b->ntile));
    k_batch_cones<<<(unsigned)std::min<int64_t>((b->ntile + 255) / 256, 65536), 256, 0, s>>>(
        sky->qcone, sky->nqtile, sky->rcone, sky->nrtile, b->tcone2);
    VISS_LAUNCHED(k_batch_cones);
  }
  VISS_HIP(hipMalloc((void**)&b->tiles, sizeof(int32_t) * G * b->ntile));
  VISS_HIP(hipMalloc((void**)&b->d_count, sizeof(int32_t) * G));
  VISS_HIP(hipMemcpyAsync(b->d_offs, b->offs.data(), sizeof(int64_t) * (G + 1), hipMemcpyHostToDevice, s));
  b->ntiles_g.assign(G, (int32_t)b->ntile);
  std::vector<int32_t> all;
  if (b->flags & VISS_MASK) {
    k_group_bounds<<<(unsigned)G, 256, 0, s>>>(b->xa, b->xb, b->Ktot, b->d_offs, b->R, b->orb, b->gcone, b->gcut);
    VISS_LAUNCHED(k_group_bounds);
    if (b->quads) {
      k_group_tiles2<<<(unsigned)G, 256, 0, s>>>(b->tcone2, b->ntile, b->gcone, b->gcut, b->tiles, b->d_count);
      VISS_LAUNCHED(k_group_tiles2);
    } else {
      k_group_tiles<<<(unsigned)G, 256, 0, s>>>(b->pairs ? sky->pcone : sky->cone, b->ntile, b->pairs ? 1 : 0,
                                                b->gcone, b->gcut, b->tiles, b->d_count);
      VISS_LAUNCHED(k_group_tiles);
    }
    VISS_HIP(hipMemcpyAsync(b->ntiles_g.data(), b->d_count, sizeof(int32_t) * G, hipMemcpyDeviceToHost, s));
  } else {
    all.resize(b->ntile);
    for (int64_t t = 0; t < b->ntile; ++t) all[t] = (int32_t)t;
    for (int64_t g = 0; g < G; ++g)
      VISS_HIP(hipMemcpyAsync(b->tiles + g * b->ntile, all.data(), sizeof(int32_t) * b->ntile,
                              hipMemcpyHostToDevice, s));
  }
  VISS_HIP(hipStreamSynchronize(s));
  std::vector<int> groups(G);
  for (int64_t g = 0; g < G; ++g) groups[g] = (int)g;
  std::vector<WorkItem> v;
  build_items(b, groups, v, b->tune, 0);
  return upload_items(b, v, s);
}

}  // namespace viss

extern "C" {

int viss_batch_create(viss_sky* sky, int64_t n_groups, const int64_t* row_offsets, const double* uvw,
                      const double* xyz_a, const double* xyz_b, const viss_mask_geometry* geom, int flags,
                      viss_batch** out, void* stream) {
  VISS_REQUIRE(out != nullptr, "viss_batch_create: out is NULL");
  *out = nullptr;
  VISS_REQUIRE(sky != nullptr && row_offsets != nullptr, "viss_batch_create: NULL argument");
  VISS_REQUIRE(n_groups >= 1 && n_groups <= (1 << 20), "viss_batch_create: n_groups %lld out of range",
               (long long)n_groups);
  VISS_REQUIRE((flags & ~(VISS_MASK | VISS_PHASE_TRACK | VISS_PREC_FP64)) == 0,
               "viss_batch_create: unsupported flags 0x%x", flags);
  int rc = check_device(sky);
  if (rc != VISS_OK) return rc;
  VISS_REQUIRE(row_offsets[0] == 0, "viss_batch_create: row_offsets[0] must be 0");
  for (int64_t g = 0; g < n_groups; ++g)
    VISS_REQUIRE(row_offsets[g + 1] >= row_offsets[g], "viss_batch_create: row_offsets must be non-decreasing");
  const int64_t Ktot = row_offsets[n_groups];
  VISS_REQUIRE(Ktot <= ((int64_t)1 << 40), "viss_batch_create: too many rows");
  VISS_REQUIRE(Ktot == 0 || uvw != nullptr, "viss_batch_create: uvw is NULL");
  if (flags & VISS_MASK) {
    VISS_REQUIRE(geom && (Ktot == 0 || (xyz_a && xyz_b)),
                 "viss_batch_create: VISS_MASK needs geometry and satellite positions");
    VISS_REQUIRE(geom->R_moon >= 0.0 && (geom->orbit_radius <= 0.0 || geom->orbit_radius > geom->R_moon),
                 "viss_batch_create: bad geometry");
  }
  viss_batch* b = new (std::nothrow) viss_batch();
  if (!b) {
    set_error("viss_batch_create: host allocation failed");
    return VISS_ERR_NOMEM;
  }
  b->sky = sky;
  b->device = sky->device;
  b->G = n_groups;
  b->Ktot = Ktot;
  b->offs.assign(row_offsets, row_offsets + n_groups + 1);
  b->uvw = uvw;
  b->xa = xyz_a;
  b->xb = xyz_b;
  b->R = (flags & VISS_MASK) ? geom->R_moon : 0.0;
  b->orb = (flags & VISS_MASK) ? geom->orbit_radius : 0.0;
  b->flags = flags;
  rc = batch_build(b, as_stream(stream));
  if (rc != VISS_OK) {
    batch_free(b);
    delete b;
    return rc;
  }
  *out = b;
  return VISS_OK;
}

int viss_batch_destroy(viss_batch* b) {
  if (!b) return VISS_OK;
  batch_free(b);
  delete b;
  return VISS_OK;
}

int viss_batch_forward(viss_batch* b, const double* I, double* V, int use_graph, void* stream) {
  VISS_REQUIRE(b != nullptr && I != nullptr, "viss_batch_forward: NULL argument");
  int rc = check_device(b->sky);
  if (rc != VISS_OK) return rc;
  if (b->Ktot == 0) return VISS_OK;  // no row: nothing to add (V may be NULL)
  VISS_REQUIRE(V != nullptr, "viss_batch_forward: V is NULL");
  hipStream_t s = as_stream(stream);
  VISS_REQUIRE(!use_graph || s != nullptr,
               "viss_batch_forward: graph mode needs a non-default stream (capture is not allowed on stream 0)");
  if (!use_graph) {
    rc = sky_set_brightness(b->sky, I, (b->flags & VISS_PREC_FP64) != 0, s);
    if (rc != VISS_OK) return rc;
    return launch_batch(b, b->items, b->n_items, V, s);
  }
  if (!b->exec || b->g_I != I || b->g_V != V || b->g_stream != s) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    VISS_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    rc = sky_set_brightness(b->sky, I, (b->flags & VISS_PREC_FP64) != 0, s);
    if (rc == VISS_OK) rc = launch_batch(b, b->items, b->n_items, V, s);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(s, &g);
    if (rc != VISS_OK) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture", __FILE__, __LINE__);
    b->graph = g;
    VISS_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    b->g_I = I;
    b->g_V = V;
    b->g_stream = s;
  }
  VISS_HIP(hipGraphLaunch(b->exec, s));
  return VISS_OK;
}

int viss_batch_autotune(viss_batch* b, const double* I, double* V_scratch, int reps, void* stream) {
  VISS_REQUIRE(b != nullptr && I != nullptr, "viss_batch_autotune: NULL argument");
  VISS_REQUIRE(reps >= 1 && reps <= 100, "viss_batch_autotune: reps %d outside [1, 100]", reps);
  int rc = check_device(b->sky);
  if (rc != VISS_OK) return rc;
  if (b->Ktot == 0) return VISS_OK;  // no row, no bucket to tune
  VISS_REQUIRE(V_scratch != nullptr, "viss_batch_autotune: V_scratch is NULL");
  hipStream_t s = as_stream(stream);
  rc = sky_set_brightness(b->sky, I, (b->flags & VISS_PREC_FP64) != 0, s);
  if (rc != VISS_OK) return rc;
  std::map<int, std::vector<int>> buckets;
  for (int64_t g = 0; g < b->G; ++g) {
    const int64_t rows = b->offs[g + 1] - b->offs[g];
    if (rows > 0 && b->ntiles_g[g] > 0) buckets[bucket_of(b->ntiles_g[g], rows)].push_back((int)g);
  }
  static const int cands[] = {8, 16, 32, 64, 128, 256, 1 << 20};
  // timing events and the candidate work table, released on every return path
  struct Scratch {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    WorkItem* tmp = nullptr;
    size_t n = 0;
    ~Scratch() {
      if (tmp) (void)hipFree(tmp);
      if (e0) (void)hipEventDestroy(e0);
      if (e1) (void)hipEventDestroy(e1);
    }
  } sc;
  VISS_HIP(hipEventCreate(&sc.e0));
  VISS_HIP(hipEventCreate(&sc.e1));
  for (auto& kv : buckets) {
    float best = 1e30f;
    int best_tpc = default_tpc(b->ntiles_g[kv.second[0]]);
    for (int tpc : cands) {
      std::vector<WorkItem> v;
      build_items(b, kv.second, v, b->tune, tpc);
      if (v.size() > sc.n) {
        if (sc.tmp) VISS_HIP(hipFree(sc.tmp));
        sc.tmp = nullptr;
        VISS_HIP(hipMalloc((void**)&sc.tmp, v.size() * sizeof(WorkItem)));
        sc.n = v.size();
      }
      VISS_HIP(hipMemcpyAsync(sc.tmp, v.data(), v.size() * sizeof(WorkItem), hipMemcpyHostToDevice, s));
      rc = launch_batch(b, sc.tmp, (int64_t)v.size(), V_scratch, s);  // warm
      if (rc != VISS_OK) return rc;
      VISS_HIP(hipEventRecord(sc.e0, s));
      for (int r = 0; r < reps && rc == VISS_OK; ++r) rc = launch_batch(b, sc.tmp, (int64_t)v.size(), V_scratch, s);
      if (rc != VISS_OK) return rc;
      VISS_HIP(hipEventRecord(sc.e1, s));
      VISS_HIP(hipEventSynchronize(sc.e1));
      float ms = 0;
      VISS_HIP(hipEventElapsedTime(&ms, sc.e0, sc.e1));
      if (ms < best) {
        best = ms;
        best_tpc = tpc;
      }
    }
    b->tune[kv.first] = best_tpc;
  }
  std::vector<int> groups(b->G);
  for (int64_t g = 0; g < b->G; ++g) groups[g] = (int)g;
  std::vector<WorkItem> v;
  build_items(b, groups, v, b->tune, 0);
  return upload_items(b, v, s);
}

int viss_batch_tuning_save(const viss_batch* b, const char* path) {
  VISS_REQUIRE(b != nullptr && path != nullptr, "viss_batch_tuning_save: NULL argument");
  FILE* f = fopen(path, "w");
  VISS_REQUIRE(f != nullptr, "viss_batch_tuning_save: cannot open %s", path);
  fprintf(f, "# viss batch tuning: log2(visible tiles) log2(rows) tiles_per_item\n");
  for (auto& kv : b->tune) fprintf(f, "%d %d %d\n", kv.first / 64, kv.first % 64, kv.second);
  fclose(f);
  return VISS_OK;
}

int viss_batch_tuning_load(viss_batch* b, const char* path, void* stream) {
  VISS_REQUIRE(b != nullptr && path != nullptr, "viss_batch_tuning_load: NULL argument");
  FILE* f = fopen(path, "r");
  VISS_REQUIRE(f != nullptr, "viss_batch_tuning_load: cannot open %s", path);
  char line[256];
  std::map<int, int> t;
  while (fgets(line, sizeof line, f)) {
    int a, c, tpc;
    if (line[0] == '#') continue;
    if (sscanf(line, "%d %d %d", &a, &c, &tpc) == 3 && a >= 0 && a < 64 && c >= 0 && c < 64 && tpc >= 1)
      t[a * 64 + c] = tpc;
  }
  fclose(f);
  b->tune = t;
  std::vector<int> groups(b->G);
  for (int64_t g = 0; g < b->G; ++g) groups[g] = (int)g;
  std::vector<WorkItem> v;
  build_items(b, groups, v, b->tune, 0);
  return upload_items(b, v, as_stream(stream));
}

int viss_batch_weights(viss_batch* b, const float* bll, int32_t nr, int flags, int32_t* gs, uint32_t* mb,
                       double* dcf, double* dg, double* wt, int64_t* n_out_of_range, void* stream) {
  VISS_REQUIRE(b != nullptr, "viss_batch_weights: batch is NULL");
  int rc = check_device(b->sky);
  if (rc != VISS_OK) return rc;
  VISS_REQUIRE(nr >= 1 && nr <= (1 << 26), "viss_batch_weights: nr %d outside [1, 2^26]", nr);
  VISS_REQUIRE(mb && dcf, "viss_batch_weights: mb/dcf must not be NULL");
  VISS_REQUIRE(b->Ktot == 0 || (bll && gs && wt), "viss_batch_weights: bll/gs/wt must not be NULL");
  VISS_REQUIRE((flags & ~(VISS_DCF_BLOCKAGE_SHIFT)) == 0, "viss_batch_weights: unsupported flags 0x%x", flags);
  return batch_weights(bll, b->uvw + 2 * b->Ktot, b->w_chunks, b->n_w_chunks, nr, flags, gs, mb, dcf, dg, wt,
                       n_out_of_range, as_stream(stream));
}

// Prologue and tile kernels of the back-projection over every group's rows,
// on the batch's row chunks.
static int launch_batch_bp(viss_batch* b, const double* y, const double* wt, int flags, double* x, int per_group,
                           hipStream_t s) {
  const int f = (b->flags & (VISS_MASK | VISS_PREC_FP64)) | flags;
  const viss_mask_geometry geom{b->R, b->orb};
  return sky_backproject_ws(b->sky, b->uvw, b->xa, b->xb, y, wt, b->Ktot, &geom, f, x, (char*)b->bp_ws.p, s,
                            b->bp_chunks, b->n_bp_chunks, per_group ? b->sky->P : 0);
}

int viss_batch_backproject(viss_batch* b, const double* y, const double* wt, int flags, double* x, int per_group,
                           int use_graph, void* stream) {
  VISS_REQUIRE(b != nullptr && x != nullptr, "viss_batch_backproject: NULL argument");
  int rc = check_device(b->sky);
  if (rc != VISS_OK) return rc;
  VISS_REQUIRE((flags & ~(VISS_PHASE_TRACK | VISS_HERMITIAN)) == 0, "viss_batch_backproject: unsupported flags 0x%x",
               flags);
  if (b->Ktot == 0) return VISS_OK;  // no row: x unchanged (y may be NULL)
  VISS_REQUIRE(y != nullptr, "viss_batch_backproject: y is NULL");
  hipStream_t s = as_stream(stream);
  VISS_REQUIRE(!use_graph || s != nullptr,
               "viss_batch_backproject: graph mode needs a non-default stream (capture is not allowed on stream 0)");
  rc = b->bp_ws.reserve(bp_workspace_bytes(b->Ktot, (b->flags & (VISS_MASK | VISS_PREC_FP64)) | flags), s);
  if (rc != VISS_OK) return rc;
  if (!use_graph) return launch_batch_bp(b, y, wt, flags, x, per_group, s);
  if (!b->bp_exec || b->bg_y != y || b->bg_wt != wt || b->bg_x != x || b->bg_flags != flags ||
      b->bg_per_group != (per_group ? 1 : 0) || b->bg_stream != s) {
    if (b->bp_exec) (void)hipGraphExecDestroy(b->bp_exec);
    if (b->bp_graph) (void)hipGraphDestroy(b->bp_graph);
    b->bp_exec = nullptr;
    b->bp_graph = nullptr;
    VISS_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    rc = launch_batch_bp(b, y, wt, flags, x, per_group, s);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(s, &g);
    if (rc != VISS_OK) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture", __FILE__, __LINE__);
    b->bp_graph = g;
    VISS_HIP(hipGraphInstantiate(&b->bp_exec, b->bp_graph, nullptr, nullptr, 0));
    b->bg_y = y;
    b->bg_wt = wt;
    b->bg_x = x;
    b->bg_flags = flags;
    b->bg_per_group = per_group ? 1 : 0;
    b->bg_stream = s;
  }
  VISS_HIP(hipGraphLaunch(b->bp_exec, s));
  return VISS_OK;
}

int viss_batch_stats(const viss_batch* b, int64_t* visible_tiles, int64_t* total_tiles, int64_t* n_items,
                     int64_t* grid_blocks) {
  VISS_REQUIRE(b != nullptr, "viss_batch_stats: NULL batch");
  int64_t vt = 0;
  for (int32_t c : b->ntiles_g) vt += c;
  if (visible_tiles) *visible_tiles = vt;
  if (total_tiles) *total_tiles = b->G * b->ntile;
  if (n_items) *n_items = b->n_items;
  if (grid_blocks) *grid_blocks = (b->n_items + 3) / 4;
  return VISS_OK;
}

}  // extern "C"
