/*
 * viss.h -- C ABI of the MI355X (gfx950) lunar-orbit interferometer hot path:
 * HEALPix pix2ang, the direct-sum 3D visibility generator (forward), the
 * density-compensation weights and the back-projection / adjoint.
 *
 * Conventions (all entry points):
 *   - every array argument is a DEVICE pointer on the current HIP device,
 *     owned by the caller; `stream` is a hipStream_t (NULL = default stream);
 *   - calls are asynchronous on `stream`.  These calls synchronise or
 *     allocate device memory; every other call does neither (it only
 *     launches kernels on `stream`):
 *       viss_sky_create        allocates the sky and synchronises the stream
 *                              (up to 6 times) to size its layouts;
 *       viss_sky_destroy, viss_batch_destroy   free (hipFree waits for the device);
 *       viss_forward, viss_backproject         create and destroy a temporary sky per call;
 *       viss_sky_backproject, viss_sky_forward_snapshots,
 *       viss_sky_backproject_snapshots         when the sky's row-record buffer must grow:
 *                              the first such call on a sky and any later call that needs
 *                              more bytes than any before (more rows, FP64 or masked records):
 *                              hipStreamSynchronize, hipFree, hipMalloc; otherwise neither;
 *       viss_batch_create, viss_batch_autotune, viss_batch_tuning_load
 *                              allocate the work table and synchronise the stream;
 *       viss_batch_forward / viss_batch_backproject with use_graph: on the
 *                              first call and when their arguments or the stream change,
 *                              they capture and instantiate a graph (host-side objects,
 *                              no device memory, no synchronisation);
 *       viss_batch_backproject the first call on a batch (its row-record buffer);
 *     viss_sky_backproject_noalloc never allocates or synchronises;
 *   - a sequence without such calls (viss_sky_forward, viss_sky_backproject
 *     after its buffer is sized, viss_weights, viss_pix2ang, ...) can be
 *     captured into a HIP graph;
 *   - the return value is a status (VISS_OK = 0, negative on error); the
 *     message of the last failure on the calling thread is viss_last_error();
 *   - "SoA 3xN" means three consecutive blocks of N doubles (x[0..N), y, z);
 *   - complex arrays are interleaved (re, im) doubles (std::complex<double>);
 *   - output kernels ACCUMULATE (+=) into the caller's zeroed buffers, as the
 *     reference's thrust::device_vector outputs do.
 *
 * The reference exposes no library: its boundary is the pair of executables
 * pix2ang.cu / vissgen3d_*.cu and the raw-pointer kernels they launch.  Each
 * entry point below names the reference kernel(s) and launch site it replaces
 * (paths relative to the reference repository root).
 */
#ifndef VISS_H
#define VISS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VISS_VERSION 100

/* status codes */
#define VISS_OK 0
#define VISS_ERR_INVALID (-1) /* bad argument (null pointer, size, flag) */
#define VISS_ERR_HIP (-2)     /* HIP runtime / launch error */
#define VISS_ERR_SIZE (-3)    /* size overflow or above a stated limit */
#define VISS_ERR_NOMEM (-4)   /* device allocation failed */

/* flags */
#define VISS_MASK 1               /* lunar occultation mask m_kp */
#define VISS_PHASE_TRACK 2        /* forward: output phase u l + v m + w (n-1); backproject: input is phase-tracked */
#define VISS_DCF_BLOCKAGE_SHIFT 4 /* weights: vissgen3d_blockage.cu:138 indexing of dcf */
#define VISS_PREC_FP64 8          /* all-FP64 kernels (default: FP64 phase, FP32 sincos/accumulate, FP64 combine) */
#define VISS_HERMITIAN 16         /* rows k+K are (-b_k, conj): forward writes them, backproject folds them in */

/* pix2ang schemes */
#define VISS_SCHEME_RING 0
#define VISS_SCHEME_NEST 1

/* Occultation geometry.  Pixel direction s is visible from satellite r iff
 *   s . r >= t(r),  t(r) = cos(Phi) |r|,  Phi = pi - asin(R_moon / orbit_radius)   if orbit_radius > 0
 *                   t(r) = -sqrt(|r|^2 - R_moon^2)                                    otherwise.
 * orbit_radius > 0 is the reference's fixed cut (vissgen3d_noblockage.cu:240-243,
 * R = 1737.1 km, R + h = 2037.1 km); <= 0 uses each satellite's own |r|. */
typedef struct viss_mask_geometry {
  double R_moon;
  double orbit_radius;
} viss_mask_geometry;

int viss_version(void);
const char* viss_status_string(int status);
const char* viss_last_error(void);

/* HEALPix pix2ang for all 12 nside^2 pixels.  Replaces mk_pix2xy +
 * f_pix2and_nest (pix2ang.cu:61-160), launched by vissGen (pix2ang.cu:
 * 190-264); adds the RING scheme.  theta, phi: [P] FP64 (colatitude in
 * [0,pi], longitude in [0,2pi)); ring, iphi: [P] int32 1-based ring number
 * from the north pole and 1-based index within the ring (nullable).
 * 1 <= nside <= 8192. */
int viss_pix2ang(int64_t nside, int scheme, double* theta, double* phi, int32_t* ring, int32_t* iphi,
                 void* stream);

/* Direction cosines and scaled brightness, healpix_moonback_pre
 * (vissgen3d_noblockage.cu:52-73, launch :404-415):
 *   l = sin(theta)cos(phi), m = -sin(theta)sin(phi), n = cos(theta), I = B * scale.
 * lmn: SoA 3xP FP64 output; B / I nullable (both or neither). */
int viss_pixel_dirs(const double* theta, const double* phi, const double* B, int64_t P, double scale, double* lmn,
                    double* I, void* stream);

/* Forward direct sum, healpix_moonback_viss (vissgen3d_noblockage.cu:76-102,
 * launch :418-440):
 *   V_k += sum_p [mask] I_p exp(-2 pi i (u_k l_p + v_k m_p + w_k n_p [- w_k with VISS_PHASE_TRACK]))
 * lmn: SoA 3xP; I: [P]; uvw: SoA 3xK [wavelengths]; xyz_a/xyz_b: SoA 3xK [m]
 * (only read with VISS_MASK); V: complex [K] ([2K] with VISS_HERMITIAN, rows
 * k+K receiving conj(V_k), vissgen3d_noblockage.cu:100).  geom is read with
 * VISS_MASK.  Builds and frees a temporary sky (see viss_sky_create): allocates
 * and synchronises. */
int viss_forward(const double* lmn, const double* I, int64_t P, const double* uvw, const double* xyz_a,
                 const double* xyz_b, int64_t K, const viss_mask_geometry* geom, int flags, double* V, void* stream);

/* Back-projection / adjoint, computeC (vissgen3d_noblockage.cu:167-182,
 * masked form vissgen3d_blockage.cu:167-192, launch :511-527):
 *   x_p += herm * Re sum_k [mask] wt_k y'_k exp(+2 pi i (u_k l_p + v_k m_p + w_k n_p))
 * y'_k = y_k, or y_k exp(-2 pi i w_k) with VISS_PHASE_TRACK (the phase restore
 * of viss_gamma_trans, vissgen3d_noblockage.cu:149-151); wt: [K] FP64 or
 * NULL (unit weights: the adjoint A^H of viss_forward); herm = 2 with
 * VISS_HERMITIAN (the K rows are the first half of a Hermitian set), else 1.
 * y: complex [K]; x: [P] FP64.  Builds and frees a temporary sky: allocates and
 * synchronises (viss_sky_backproject on a reusable sky does not, once sized). */
int viss_backproject(const double* lmn, int64_t P, const double* uvw, const double* xyz_a, const double* xyz_b,
                     const double* y, const double* wt, int64_t K, const viss_mask_geometry* geom, int flags,
                     double* x, void* stream);

/* Density-compensation weights, vissgen3d_noblockage.cu:105-164 (ceilAndScale,
 * thrust::sort, countOccurrences, calculateR, calculateDcf, viss_gamma_trans;
 * launches :448-507):
 *   gs_k = ceil((bll_k - 1/4) / (1/2))  (FP32, as the reference)   -> gs [K] int32
 *   mb_j = #{k : gs_k = j}, j = 1..nr                               -> mb [nr] uint32 (mb[j-1])
 *   dcf_0 = 1/(4 pi), dcf_j = (2 pi/3)(R^3 - (R-1/2)^3)/mb_j, R = j/2 + 1/4  -> dcf [nr+1] FP64
 *   dg_k = 1.5 sqrt|1 - 4 sin^2 gamma| / |cos gamma|, gamma = asin(w/bll)   -> dg [K] FP64 (nullable)
 *   wt_k = fmin(dcf[gs_k] dg_k, 1/8)  (NaN -> 1/8)                  -> wt [K] FP64
 * Rows with gs outside [0, nr] (out-of-bounds reads in the reference) get
 * wt = 0 and are counted into *n_out_of_range (device int64, accumulated,
 * nullable).  VISS_DCF_BLOCKAGE_SHIFT: dcf_j uses shell j+1
 * (vissgen3d_blockage.cu:138) and dcf_nr = 0.  mb must be zeroed by the caller
 * (it accumulates).  1 <= nr <= 2^26. */
int viss_weights(const float* bll, const double* w, int64_t K, int32_t nr, int flags, int32_t* gs, uint32_t* mb,
                 double* dcf, double* dg, double* wt, int64_t* n_out_of_range, void* stream);

/* ---- reusable sky: the pixel set sorted into spatially compact 64-pixel
 * tiles with bounding cones, packed for the kernels.  Building it costs a
 * sort; viss_forward / viss_backproject build a temporary one per call.  A
 * sky is bound to the device current at creation; calls that share one sky
 * must be ordered (same stream or synchronised).
 *
 * Creation also pairs every pixel with its antipode (s_b = -s_a to 1e-12),
 * as HEALPix pixels are: the mixed-precision per-row forward and
 * back-projection then evaluate one phase and one sin/cos per pair.  Skies
 * with fewer than half of their pixels paired use the per-pixel kernels.
 * Pairs whose first pixels are (x, y, z) and (-x, -y, z) form quads
 * {(+-x, +-y, +-z)} (HEALPix is symmetric under z -> -z ring by ring): the
 * mixed-precision per-row forward then evaluates one sin/cos per quad and
 * applies each ring's factor exp(-2 pi i w z) once per tile of 16 quads.
 * Creation allocates the sky's arrays and synchronises the stream up to 6
 * times (to size the layouts); destruction frees them. */
typedef struct viss_sky viss_sky;

int viss_sky_create(const double* lmn, int64_t P, viss_sky** sky, void* stream);
int viss_sky_destroy(viss_sky* sky);
int64_t viss_sky_npix(const viss_sky* sky);
/* Antipodal pairs and unpaired pixels found at creation (host pointers,
 * nullable); *uses_pairs = 1 when the mixed per-row kernels use the pairs. */
int viss_sky_pairs(const viss_sky* sky, int64_t* npair, int64_t* nsingle, int* uses_pairs);
/* Equatorial-mirror quads found at creation, their 16-quad tiles and
 * whether the mixed per-row forward uses them (host pointers, nullable):
 * *uses_quads = 0 no, 1 yes, 2 yes with the recurrent phase (every tile's
 * quads at equal azimuth steps on one circle, as on a HEALPix ring: one FP64
 * FMA a quad for u x + v y).  Layout query only: no counterpart in the
 * reference. */
int viss_sky_quads(const viss_sky* sky, int64_t* nquad, int64_t* nquad_tiles, int* uses_quads);

int viss_sky_forward(viss_sky* sky, const double* I, const double* uvw, const double* xyz_a, const double* xyz_b,
                     int64_t K, const viss_mask_geometry* geom, int flags, double* V, void* stream);
/* viss_sky_backproject packs per-row records (32 B, 64 B with VISS_PREC_FP64,
 * plus 96 B with VISS_MASK) into a buffer the sky owns; the call that first
 * needs more bytes than the buffer holds synchronises `stream`, frees the old
 * buffer and allocates a larger one (25% headroom).  The row records of one
 * call are read by that call's kernels only, so calls on one sky must be
 * ordered (same stream). */
int viss_sky_backproject(viss_sky* sky, const double* uvw, const double* xyz_a, const double* xyz_b,
                         const double* y, const double* wt, int64_t K, const viss_mask_geometry* geom, int flags,
                         double* x, void* stream);

/* Allocation-free viss_sky_backproject.  No counterpart in the reference (its
 * kernels take caller-owned buffers only): this is for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same arguments and result as
 * viss_sky_backproject, whose row records (32 B per row, 64 B with
 * VISS_PREC_FP64, plus 96 B with VISS_MASK) go to a sky-owned buffer that
 * grows on demand; here they go to a static device array of
 * VISS_NOALLOC_SCRATCH_BYTES per device, held until the process ends.  The
 * call allocates and frees nothing and never synchronises (its first call on a
 * device looks up the array's address, hipGetSymbolAddress).  K above what
 * the array holds (262,142 mixed-precision rows with VISS_MASK, 209,713 with
 * VISS_PREC_FP64 | VISS_MASK; split larger row sets over calls, the result
 * accumulates) returns VISS_ERR_SIZE.  Every call on one device shares that
 * device's array: calls that overlap on one device (different streams,
 * threads or skies without ordering) race. */
#define VISS_NOALLOC_SCRATCH_BYTES (32 << 20)
int viss_sky_backproject_noalloc(viss_sky* sky, const double* uvw, const double* xyz_a, const double* xyz_b,
                                 const double* y, const double* wt, int64_t K, const viss_mask_geometry* geom,
                                 int flags, double* x, void* stream);

/* Number of (k, p) pairs with mask = 1 (accumulated into *count, device
 * int64): the evaluated-pair count used for throughput accounting. */
int viss_sky_count_visible(viss_sky* sky, const double* xyz_a, const double* xyz_b, int64_t K,
                           const viss_mask_geometry* geom, int64_t* count, void* stream);

/* Work of the mixed-precision per-row kernels over a sky's quad layout for K
 * masked rows, counted with the kernels' own occultation test; accumulated
 * into counts (device int64 [6]):
 *   [0] visible (row, pixel) pairs (= viss_sky_count_visible),
 *   [1] (row, 16-quad tile) with a visible pixel,
 *   [2] (row, quad) with a visible pixel: one phase and one sin/cos each,
 *   [3] (row, quad, ring z or -z) with a visible pixel: one pair of FMAs each,
 *   [4] (row, pair outside the quads) with a visible pixel: one phase and sin/cos,
 *   [5] (row, run of up to 8 consecutive such tiles on one ring pair, within
 *       one of the forward's tile chunks): one ring factor each.
 * The measurement of the kernels' algorithmic work (no counterpart in the
 * reference).  VISS_ERR_INVALID for a sky without the quad layout. */
int viss_sky_count_work(viss_sky* sky, const double* xyz_a, const double* xyz_b, int64_t K,
                        const viss_mask_geometry* geom, int64_t* counts, void* stream);

/* ---- batched forward: many visibility groups (snapshot groups x
 * frequencies) over one sky in ONE launch.  Replaces the
 * reference's per-day loop of forward launches (vissgen3d_noblockage.cu:
 * 270-440) for in-memory batches.  Rows of group g are rows
 * [row_offsets[g], row_offsets[g+1]) of the SoA 3 x K_total arrays uvw /
 * xyz_a / xyz_b (device; they must stay valid for the batch's lifetime);
 * row_offsets is a HOST array of n_groups + 1 entries starting at 0.  Each
 * group gets a compacted list of the sky tiles its satellites can see;
 * the lists are cut into work items (64 rows x a tile range sized per
 * bucket of list length and group rows) and one wavefront runs each item.  flags: VISS_MASK,
 * VISS_PHASE_TRACK, VISS_PREC_FP64.  The batch keeps a pointer to `sky`.
 * Groups may be empty; with no row at all (K_total = 0) the row arrays, and
 * the row-sized arguments of the calls below (V, y, wt, bll, gs, dg, V_scratch),
 * may be NULL, and the calls add nothing. */
typedef struct viss_batch viss_batch;

int viss_batch_create(viss_sky* sky, int64_t n_groups, const int64_t* row_offsets, const double* uvw,
                      const double* xyz_a, const double* xyz_b, const viss_mask_geometry* geom, int flags,
                      viss_batch** batch, void* stream);
int viss_batch_destroy(viss_batch* batch);
/* V[k] += forward for every row k of every group (complex [K_total]).
 * use_graph != 0: the brightness scatter and the batch kernel are
 * captured once into a hipGraph (re-captured if I, V or the stream
 * change) and replayed; needs a non-default stream. */
int viss_batch_forward(viss_batch* batch, const double* I, double* V, int use_graph, void* stream);
/* Times candidate tile-range lengths per (log2 visible tiles, log2 rows)
 * bucket on the device (reps launches each, accumulating into V_scratch
 * [K_total] complex) and keeps the fastest; synchronous. */
int viss_batch_autotune(viss_batch* batch, const double* I, double* V_scratch, int reps, void* stream);
int viss_batch_tuning_save(const viss_batch* batch, const char* path);
int viss_batch_tuning_load(viss_batch* batch, const char* path, void* stream);
/* Density-compensation weights of every group, each group counted on its own
 * (the reference computes them per day, vissgen3d_noblockage.cu:448-507;
 * viss_weights per group), in two launches: one shell histogram per group
 * (LDS-privatised, one block per chunk of a group's rows) and the rows'
 * weights, each block building its group's dcf table in LDS.  bll: [K_total]
 * FP32 baseline lengths; w is the batch's uvw row 2.  gs, dg (nullable), wt:
 * [K_total]; mb: [G nr] uint32, group g's counts at mb + g nr (zeroed by the
 * caller: it accumulates); dcf: [G (nr + 1)], group g's table at
 * dcf + g (nr + 1); n_out_of_range: total over groups (device int64,
 * accumulated, nullable).  One nr for every group (the largest shell any row
 * falls in, or the reference's ceil(2 bl_max / lambda) of the highest
 * frequency): a group's weights do not depend on nr beyond its own largest
 * shell, except with VISS_DCF_BLOCKAGE_SHIFT (dcf_nr = 0).  VISS_ERR_SIZE when
 * a group's table, (nr + 1) doubles, exceeds one block's LDS (nr <= 20,479 on
 * gfx950).  Launches kernels only. */
int viss_batch_weights(viss_batch* batch, const float* bll, int32_t nr, int flags, int32_t* gs, uint32_t* mb,
                       double* dcf, double* dg, double* wt, int64_t* n_out_of_range, void* stream);
/* Back-projection of every group's rows, computeC per day
 * (vissgen3d_noblockage.cu:511-527) for all groups in one prologue and one
 * tile-kernel launch: y: complex [K_total], wt: [K_total] FP64 (e.g. from
 * viss_batch_weights) or NULL (the adjoint).  per_group == 0: x [P] receives
 * the sum over groups (the notebooks' multi-day total); per_group != 0: x is
 * [G][P], group g's image at x + g P.  flags: VISS_PHASE_TRACK,
 * VISS_HERMITIAN (as viss_sky_backproject); occultation and precision are the
 * batch's.  The row records go to a buffer the batch owns: the first call (and
 * one with larger records) synchronises and allocates it, later calls launch
 * kernels only.  use_graph != 0: the launches are captured once into a
 * hipGraph (re-captured when y, wt, x, flags, per_group or the stream change)
 * and replayed; needs a non-default stream. */
int viss_batch_backproject(viss_batch* batch, const double* y, const double* wt, int flags, double* x, int per_group,
                           int use_graph, void* stream);
/* Sizes of the batch: tiles kept over all groups, tiles before compaction,
 * work items, blocks of the batch launch (4 items per block).  Every pointer is a host pointer
 * (nullable). */
int viss_batch_stats(const viss_batch* batch, int64_t* visible_tiles, int64_t* total_tiles, int64_t* n_items,
                     int64_t* grid_blocks);

/* ---- snapshot-structured rows: S snapshots of the same N satellites
 * (2 <= N <= 16), each snapshot's N(N-1)/2 pairs a<b in row-major
 * upper-triangle order (a = 0..N-2, b = a+1..N-1), rows snapshot-major;
 * baseline b = (r_a - r_b) / wavelength.  sat_pos: [S][N][3] FP64 metres.
 * Same sum as viss_sky_backproject over those rows (the reference's rows are
 * built this way, vissgen3d_noblockage.cu:302-362), computed with one phasor
 * per (satellite, pixel) instead of one per (row, pixel).  y, wt: [S*N(N-1)/2]
 * (complex / FP64, wt nullable).  flags: VISS_MASK, VISS_PHASE_TRACK,
 * VISS_HERMITIAN; mixed precision only. */
/* Forward of snapshot-structured rows: V[s*N(N-1)/2 + pair] += sum_p [mask] I_p exp(-2 pi i b.s_p)
 * (the same sum as viss_sky_forward over those rows); V: complex [S*N(N-1)/2]
 * ([2x] with VISS_HERMITIAN). */
int viss_sky_forward_snapshots(viss_sky* sky, const double* I, const double* sat_pos, int64_t S, int32_t N,
                               double wavelength, const viss_mask_geometry* geom, int flags, double* V, void* stream);
int viss_sky_backproject_snapshots(viss_sky* sky, const double* sat_pos, int64_t S, int32_t N, double wavelength,
                                   const double* y, const double* wt, const viss_mask_geometry* geom, int flags,
                                   double* x, void* stream);

/* ---- image utilities after the hot path ----
 * Multi-day aggregation, the notebooks' sum over days
 * (healpix3d_blockage.ipynb: total += np.nan_to_num(C_i);
 * healpix3d_noblockage.ipynb: total += C_i):
 *   total[p] += (nan_to_zero && isnan(img[p])) ? 0 : img[p]. */
int viss_image_accumulate(const double* img, int64_t P, int nan_to_zero, double* total, void* stream);
/* HEALPix scheme reorder of a map of 12 nside^2 FP64 values (nside a power of
 * two): to_ring != 0 maps a NEST-ordered map to RING order, else RING to NEST
 * (the notebooks display NEST maps, hp.mollview(nest=True)). in != out. */
int viss_reorder(const double* in, int64_t nside, int to_ring, double* out, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* VISS_H */
