/*
 * viss_oracle.c -- FP64 CPU restatement of the reference's hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP library in
 * viss_optimize_3d_amd/csrc.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load it; the product path never calls it.
 *
 * Every function cites the reference file:line whose formula it restates
 * (paths relative to the reference repository root).  Deliberate departures
 * from the reference (indexing bugs, FP32 phase, acos mask) are listed at each
 * function; the FP32 "as the reference computes it" forward is kept separately
 * (oracle_forward_ref32) as the CPU baseline of the reference formulation.
 *
 * Pinning (what checks each function, tests/test_oracle_*.py):
 *   pix2ang_nest / pix2ang_ring : nside=1 closed-form table, ring populations,
 *       z values per ring, NEST<->RING multiset equality through an
 *       independently written nest2ring, antipodal symmetry, sum_p s_p = 0,
 *       an independently written ang2pix round trip.
 *   pixel_dirs  : unit norm, closed forms at the poles / equator.
 *   forward     : single point source closed form, b=0 sum, Hermitian symmetry,
 *       linearity, mask all-ones when the cut is below -1, mpmath brute force.
 *   weights     : integer closed forms (bincount), exact shell volumes, dg
 *       closed form, fmin NaN semantics.
 *   backproject : adjoint dot-product test <Ax,y> = <x,A^H y>, point-source
 *       closed form, mpmath brute force.
 *   mask_flips  : its dot side equals count_visible, its acos side the count
 *       of forward_ref32 with unit brightness and zero baselines.
 *   mask_threshold / visible : the same expression in exact rational
 *       arithmetic on pairs placed on the cut, where the evaluation order
 *       decides (test_oracle_cut.py).
 *
 * Build: oracle/Makefile (gcc -O2 -fopenmp -ffp-contract=off).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define VISS_MASK 1
#define VISS_PHASE_TRACK 2
#define VISS_DCF_BLOCKAGE_SHIFT 4
#define VISS_HERMITIAN 16

static const double PI = 3.14159265358979323846264338327950288;

/* Exact integer square root (floor). */
static int64_t isqrt64(int64_t x) {
  if (x <= 0) return 0;
  int64_t s = (int64_t)sqrt((double)x);
  while (s * s > x) --s;
  while ((s + 1) * (s + 1) <= x) ++s;
  return s;
}

/* 0-based Morton de-interleave of the low 2*nbits bits: even bits -> x, odd
 * bits -> y.  Same table mk_pix2xy builds 10 bits at a time
 * (pix2ang.cu:61-88); here done bit by bit for any nside <= 2^29. */
static void deinterleave(int64_t v, int64_t* x, int64_t* y) {
  int64_t ix = 0, iy = 0;
  for (int b = 0; b < 31; ++b) {
    ix |= ((v >> (2 * b)) & 1) << b;
    iy |= ((v >> (2 * b + 1)) & 1) << b;
  }
  *x = ix;
  *y = iy;
}

/* NEST pix2ang, 0-based restatement of f_pix2and_nest (pix2ang.cu:92-160) with
 * the +1 indexing of pix2x/jrll/jpll removed (pix2ang.cu:109-110,115-116,123,
 * 151), a bounds-safe pixel loop and FP64 arithmetic.  Constants jrll/jpll are
 * pix2ang.cu:183-184; fact1/fact2 pix2ang.cu:205-206.
 * Outputs ring (jr, 1-based from the north pole) and iphi (jp, 1-based in ring)
 * so that integer parity can be checked bit-exactly. */
int oracle_pix2ang_nest(int64_t nside, double* theta, double* phi, int32_t* ring, int32_t* iphi) {
  static const int jrll[12] = {2, 2, 2, 2, 3, 3, 3, 3, 4, 4, 4, 4};
  static const int jpll[12] = {1, 3, 5, 7, 0, 2, 4, 6, 1, 3, 5, 7};
  if (nside < 1 || nside > (1 << 13)) return -1;
  if (nside & (nside - 1)) return -1; /* NEST numbers the nside^2 pixels of a face in bit order */
  const int64_t npface = nside * nside, npix = 12 * npface, nl4 = 4 * nside;
  const double three_ns2 = 3.0 * (double)npface;
#pragma omp parallel for schedule(static)
  for (int64_t ip = 0; ip < npix; ++ip) {
    const int64_t face = ip / npface, ipf = ip % npface;
    int64_t ix, iy;
    deinterleave(ipf, &ix, &iy);
    const int64_t jrt = ix + iy, jpt = ix - iy;
    const int64_t jr = jrll[face] * nside - jrt - 1;
    int64_t nr, kshift;
    double z;
    if (jr < nside) { /* north polar cap */
      nr = jr;
      z = ((double)(three_ns2 - (double)(nr * nr))) / three_ns2;
      kshift = 0;
    } else if (jr <= 3 * nside) { /* equatorial belt */
      nr = nside;
      z = (double)(2 * (2 * nside - jr)) / (double)(3 * nside);
      kshift = (jr - nside) & 1;
    } else { /* south polar cap */
      nr = nl4 - jr;
      z = -((double)(three_ns2 - (double)(nr * nr))) / three_ns2;
      kshift = 0;
    }
    int64_t jp = (jpll[face] * nr + jpt + 1 + kshift) / 2;
    if (jp > nl4) jp -= nl4;
    if (jp < 1) jp += nl4;
    theta[ip] = acos(z);
    phi[ip] = (PI / 2.0) * ((double)jp - 0.5 * (double)(kshift + 1)) / (double)nr;
    if (ring) ring[ip] = (int32_t)jr;
    if (iphi) iphi[ip] = (int32_t)jp;
  }
  return 0;
}

/* RING pix2ang (not in the reference; the HEALPix standard, Gorski et al. 2005,
 * ApJ 622:759, sec. 4).  Uses the same z / phi expressions as the NEST routine
 * above so that the two schemes give bit-identical (theta, phi) for the same
 * (ring, iphi). */
int oracle_pix2ang_ring(int64_t nside, double* theta, double* phi, int32_t* ring, int32_t* iphi) {
  if (nside < 1 || nside > (1 << 13)) return -1;
  const int64_t npface = nside * nside, npix = 12 * npface, nl4 = 4 * nside;
  const int64_t ncap = 2 * nside * (nside - 1);
  const double three_ns2 = 3.0 * (double)npface;
#pragma omp parallel for schedule(static)
  for (int64_t p = 0; p < npix; ++p) {
    int64_t jr, jp, nr, kshift = 0;
    double z;
    if (p < ncap) {
      const int64_t i = (1 + isqrt64(1 + 2 * p)) / 2;
      jp = p + 1 - 2 * i * (i - 1);
      jr = i;
      nr = i;
      z = ((double)(three_ns2 - (double)(nr * nr))) / three_ns2;
    } else if (p < npix - ncap) {
      const int64_t q = p - ncap;
      jr = q / nl4 + nside;
      jp = q % nl4 + 1;
      nr = nside;
      kshift = (jr - nside) & 1;
      z = (double)(2 * (2 * nside - jr)) / (double)(3 * nside);
    } else {
      const int64_t q = npix - p;
      const int64_t i = (1 + isqrt64(2 * q - 1)) / 2;
      jp = 4 * i + 1 - (q - 2 * i * (i - 1));
      jr = nl4 - i;
      nr = i;
      z = -((double)(three_ns2 - (double)(nr * nr))) / three_ns2;
    }
    theta[p] = acos(z);
    phi[p] = (PI / 2.0) * ((double)jp - 0.5 * (double)(kshift + 1)) / (double)nr;
    if (ring) ring[p] = (int32_t)jr;
    if (iphi) iphi[p] = (int32_t)jp;
  }
  return 0;
}

/* Direction cosines and scaled brightness, healpix_moonback_pre
 * (vissgen3d_noblockage.cu:52-73): lat = pi/2 - theta, phi' = -wrap(phi), so
 * l = sin(theta)cos(phi), m = -sin(theta)sin(phi), n = cos(theta), and
 * I = B * scale with scale = 4 pi / npix (vissgen3d_noblockage.cu:233). */
void oracle_pixel_dirs(const double* theta, const double* phi, const double* B, int64_t P, double scale,
                       double* l, double* m, double* n, double* I) {
#pragma omp parallel for schedule(static)
  for (int64_t p = 0; p < P; ++p) {
    const double st = sin(theta[p]), ct = cos(theta[p]);
    l[p] = st * cos(phi[p]);
    m[p] = -(st * sin(phi[p]));
    n[p] = ct;
    if (I) I[p] = B[p] * scale;
  }
}

/* Per-satellite occultation threshold t such that pixel direction s is
 * visible from r iff s.r >= t.  The reference tests acosf(s.r/|r|) <= Phi with
 * Phi = pi - asin(R/(R+h)) (vissgen3d_noblockage.cu:88-94,240-243), i.e.
 * s.r >= cos(Phi)|r|.  orbit_radius > 0 reproduces that fixed cut
 * (orbit_radius = R+h); orbit_radius <= 0 uses each satellite's own |r|:
 * t = -sqrt(|r|^2 - R^2). */
double oracle_mask_threshold(double x, double y, double z, double R_moon, double orbit_radius) {
  /* Explicit fma: the decision at the cut is an exact comparison, so t and
   * s.r are evaluated in one fixed order (the GPU kernels use the same). */
  const double r2 = fma(x, x, fma(y, y, z * z));
  if (orbit_radius > 0.0) {
    const double q = R_moon / orbit_radius;
    const double c = -sqrt(fma(-q, q, 1.0));
    return c * sqrt(r2);
  }
  const double d = fma(-R_moon, R_moon, r2);
  return d > 0.0 ? -sqrt(d) : 0.0; /* satellite inside the moon: sees the upper half-space only */
}

/* s.r >= t, with s.r = fma(l, x, fma(m, y, n * z)). */
static inline int visible(double l, double m, double n, double x, double y, double z, double t) {
  return fma(l, x, fma(m, y, n * z)) >= t;
}

/* Forward direct sum, healpix_moonback_viss (vissgen3d_noblockage.cu:76-102):
 *   V_k = sum_p [mask] I_p exp(-2 pi i phase_kp)
 * phase = u l + v m + w n (full 3D phase), or u l + v m + w (n-1) with
 * VISS_PHASE_TRACK (the reference's output, vissgen3d_noblockage.cu:95).
 * Mask (VISS_MASK) in dot-product form with per-row satellite positions
 * xyz_a / xyz_b (SoA 3xK).  With VISS_HERMITIAN the caller's buffers hold 2K
 * rows and row k+K receives conj(V_k) (vissgen3d_noblockage.cu:100).
 * Phase reduced mod 1 before the trig call so the FP64 sin/cos see [-1/2,1/2].
 */
void oracle_forward(const double* l, const double* m, const double* n, const double* I, int64_t P,
                    const double* u, const double* v, const double* w, const double* xyz_a, const double* xyz_b,
                    int64_t K, int flags, double R_moon, double orbit_radius, double* V_re, double* V_im) {
  const int masked = flags & VISS_MASK;
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t k = 0; k < K; ++k) {
    double ta = 0, tb = 0, xa = 0, ya = 0, za = 0, xb = 0, yb = 0, zb = 0;
    if (masked) {
      xa = xyz_a[k]; ya = xyz_a[K + k]; za = xyz_a[2 * K + k];
      xb = xyz_b[k]; yb = xyz_b[K + k]; zb = xyz_b[2 * K + k];
      ta = oracle_mask_threshold(xa, ya, za, R_moon, orbit_radius);
      tb = oracle_mask_threshold(xb, yb, zb, R_moon, orbit_radius);
    }
    const double c0 = (flags & VISS_PHASE_TRACK) ? -w[k] : 0.0;
    double sr = 0.0, si = 0.0;
    for (int64_t p = 0; p < P; ++p) {
      if (masked && !(visible(l[p], m[p], n[p], xa, ya, za, ta) && visible(l[p], m[p], n[p], xb, yb, zb, tb)))
        continue;
      const double ph = ((u[k] * l[p] + v[k] * m[p]) + w[k] * n[p]) + c0;
      const double f = ph - nearbyint(ph);
      const double a = 2.0 * PI * f;
      sr += I[p] * cos(a);
      si -= I[p] * sin(a);
    }
    V_re[k] = sr;
    V_im[k] = si;
    if (flags & VISS_HERMITIAN) {
      V_re[K + k] = sr;
      V_im[K + k] = -si;
    }
  }
}

/* The reference's own FP32 formulation of the forward, kept as the CPU
 * baseline of "the reference formulation": FP32 phase with (n-1) tracking,
 * acosf mask against the fixed cut Phi (vissgen3d_noblockage.cu:86-99). */
void oracle_forward_ref32(const float* l, const float* m, const float* n, const float* B, int64_t P,
                          const float* u, const float* v, const float* w, const float* xyz_a, const float* xyz_b,
                          int64_t K, float phi_cut, float* V_re, float* V_im) {
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t k = 0; k < K; ++k) {
    const float xa = xyz_a[k], ya = xyz_a[K + k], za = xyz_a[2 * K + k];
    const float xb = xyz_b[k], yb = xyz_b[K + k], zb = xyz_b[2 * K + k];
    const float na = sqrtf(xa * xa + ya * ya + za * za), nb = sqrtf(xb * xb + yb * yb + zb * zb);
    float sr = 0.f, si = 0.f;
    for (int64_t p = 0; p < P; ++p) {
      const float b1 = acosf((l[p] * xa + m[p] * ya + n[p] * za) / na);
      const float b2 = acosf((l[p] * xb + m[p] * yb + n[p] * zb) / nb);
      if (b1 <= phi_cut && b2 <= phi_cut) {
        const float ph = u[k] * l[p] + v[k] * m[p] + w[k] * (n[p] - 1.0f);
        const float a = -2.0f * (float)PI * ph;
        sr += B[p] * cosf(a);
        si += B[p] * sinf(a);
      }
    }
    V_re[k] = sr;
    V_im[k] = si;
  }
}

/* Occultation decisions of the reference's FP32 form (vissgen3d_noblockage.cu:
 * 86-94: beta_i = acosf(s.r_i / |r_i|) <= Phi for both satellites, every
 * operand FP32) against the FP64 dot form s.r_i >= t_i (oracle_mask_threshold
 * with R_moon / orbit_radius) on every (row, pixel) pair.  out[0]: pairs only
 * the acos form keeps; out[1]: pairs only the dot form keeps; out[2]: pairs
 * both keep. */
void oracle_mask_flips(const double* l, const double* m, const double* n, int64_t P, const double* xyz_a,
                       const double* xyz_b, int64_t K, float phi_cut, double R_moon, double orbit_radius,
                       int64_t* out) {
  int64_t acos_only = 0, dot_only = 0, both = 0;
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : acos_only, dot_only, both)
  for (int64_t k = 0; k < K; ++k) {
    const double xa = xyz_a[k], ya = xyz_a[K + k], za = xyz_a[2 * K + k];
    const double xb = xyz_b[k], yb = xyz_b[K + k], zb = xyz_b[2 * K + k];
    const double ta = oracle_mask_threshold(xa, ya, za, R_moon, orbit_radius);
    const double tb = oracle_mask_threshold(xb, yb, zb, R_moon, orbit_radius);
    const float fxa = (float)xa, fya = (float)ya, fza = (float)za, fxb = (float)xb, fyb = (float)yb, fzb = (float)zb;
    const float na = sqrtf(fxa * fxa + fya * fya + fza * fza), nb = sqrtf(fxb * fxb + fyb * fyb + fzb * fzb);
    for (int64_t p = 0; p < P; ++p) {
      const float fl = (float)l[p], fm = (float)m[p], fn = (float)n[p];
      const int ref = acosf((fl * fxa + fm * fya + fn * fza) / na) <= phi_cut &&
                      acosf((fl * fxb + fm * fyb + fn * fzb) / nb) <= phi_cut;
      const int dot = visible(l[p], m[p], n[p], xa, ya, za, ta) && visible(l[p], m[p], n[p], xb, yb, zb, tb);
      acos_only += ref && !dot;
      dot_only += dot && !ref;
      both += ref && dot;
    }
  }
  out[0] = acos_only;
  out[1] = dot_only;
  out[2] = both;
}

/* Shell index of one baseline length, ceilAndScale (vissgen3d_noblockage.cu:
 * 105-111): gs = (int)ceil((bll - 1/4) / (1/2)), evaluated in FP32 exactly as
 * the reference does so the integer result is bit-exact. */
int32_t oracle_shell_index(float bll) {
  volatile float d = bll - 0.25f;
  volatile float q = d / 0.5f;
  const float c = ceilf(q);
  /* NaN and values outside int32 (undefined casts in the reference) map to
   * INT32_MIN: out of range, weight 0. */
  if (!(c > -2147483648.0f && c < 2147483648.0f)) return INT32_MIN;
  return (int32_t)c;
}

/* Density-compensation weights, vissgen3d_noblockage.cu:105-164:
 *   gs_k  = shell index (ceilAndScale, :105-111)
 *   mb_j  = #{k : gs_k = j}, j = 1..nr (countOccurrences, :114-123; integer
 *           counts instead of float atomics)
 *   dcf_0 = 1/(4 pi); dcf_j = (2 pi / 3)(R^3 - (R - 1/2)^3) / mb with
 *           R = R_j = j/2 + 1/4, mb = mb_j (calculateR/calculateDcf :125-141).
 *           VISS_DCF_BLOCKAGE_SHIFT reproduces vissgen3d_blockage.cu:138
 *           (R_{j+1}, mb_{j+1}); its out-of-range entry j = nr is defined as 0.
 *   gamma = asin(w/bll); dg = 1.5 |sqrt(1 - 4 sin^2 gamma)| / |cos gamma|
 *           (viss_gamma_trans :153-162)
 *   wt_k  = fmin(dcf[gs_k] dg_k, 1/8) (computeC :176-177, fminf semantics:
 *           NaN -> 1/8).  Rows with gs outside [0, nr] (the reference reads
 *           out of bounds there) get wt = 0 and are counted in the return value.
 * mb has nr entries (mb[j-1] counts shell j); dcf has nr+1 entries. */
int64_t oracle_weights(const float* bll, const double* w, int64_t K, int32_t nr, int flags, int32_t* gs, uint32_t* mb,
                       double* dcf, double* dg, double* wt) {
  memset(mb, 0, sizeof(uint32_t) * (size_t)nr);
  for (int64_t k = 0; k < K; ++k) {
    gs[k] = oracle_shell_index(bll[k]);
    if (gs[k] >= 1 && gs[k] <= nr) mb[gs[k] - 1] += 1;
  }
  dcf[0] = 1.0 / (4.0 * PI);
  for (int32_t j = 1; j <= nr; ++j) {
    int32_t jj = (flags & VISS_DCF_BLOCKAGE_SHIFT) ? j + 1 : j;
    if (jj > nr) {
      dcf[j] = 0.0;
      continue;
    }
    const double R = (double)jj / 2.0 + 0.25, Rm = R - 0.5;
    dcf[j] = (2.0 * PI / 3.0) * (R * R * R - Rm * Rm * Rm) / (double)mb[jj - 1];
  }
  int64_t bad = 0;
  for (int64_t k = 0; k < K; ++k) {
    const double g = asin(w[k] / (double)bll[k]);
    const double sg = sin(g), cg = cos(g);
    dg[k] = 1.5 * sqrt(fabs(1.0 - 4.0 * sg * sg)) / fabs(cg);
    if (gs[k] < 0 || gs[k] > nr) {
      wt[k] = 0.0;
      ++bad;
      continue;
    }
    wt[k] = fmin(dcf[gs[k]] * dg[k], 0.125);
  }
  return bad;
}

/* Back-projection / adjoint, computeC (vissgen3d_noblockage.cu:167-182) and
 * its masked form (vissgen3d_blockage.cu:167-192):
 *   x_p = Re sum_k [mask] wt_k y_k exp(+2 pi i (u l + v m + w n))
 * y is the full-phase visibility (VISS_PHASE_TRACK: y is the reference's
 * phase-tracked V and is first multiplied by exp(-2 pi i w), viss_gamma_trans
 * :149-151).  wt == NULL means unit weights (the adjoint A^H of the forward).
 * VISS_HERMITIAN: the K rows are the first half of a Hermitian set whose
 * second half is (-b, conj y); their sum is 2 Re of the first half's. */
void oracle_backproject(const double* l, const double* m, const double* n, int64_t P, const double* u, const double* v,
                        const double* w, const double* xyz_a, const double* xyz_b, const double* y_re,
                        const double* y_im, const double* wt, int64_t K, int flags, double R_moon,
                        double orbit_radius, double* x) {
  const int masked = flags & VISS_MASK;
  double* yr = (double*)malloc(sizeof(double) * (size_t)K);
  double* yi = (double*)malloc(sizeof(double) * (size_t)K);
  double* ta = (double*)malloc(sizeof(double) * (size_t)K);
  double* tb = (double*)malloc(sizeof(double) * (size_t)K);
  for (int64_t k = 0; k < K; ++k) {
    double a = y_re[k], b = y_im[k];
    if (flags & VISS_PHASE_TRACK) {
      const double f = w[k] - nearbyint(w[k]);
      const double c = cos(2.0 * PI * f), s = -sin(2.0 * PI * f);
      const double na = a * c - b * s, nb = a * s + b * c;
      a = na;
      b = nb;
    }
    const double g = wt ? wt[k] : 1.0;
    yr[k] = a * g;
    yi[k] = b * g;
    if (masked) {
      ta[k] = oracle_mask_threshold(xyz_a[k], xyz_a[K + k], xyz_a[2 * K + k], R_moon, orbit_radius);
      tb[k] = oracle_mask_threshold(xyz_b[k], xyz_b[K + k], xyz_b[2 * K + k], R_moon, orbit_radius);
    }
  }
  const double herm = (flags & VISS_HERMITIAN) ? 2.0 : 1.0;
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t p = 0; p < P; ++p) {
    double acc = 0.0;
    for (int64_t k = 0; k < K; ++k) {
      if (masked && !(visible(l[p], m[p], n[p], xyz_a[k], xyz_a[K + k], xyz_a[2 * K + k], ta[k]) &&
                      visible(l[p], m[p], n[p], xyz_b[k], xyz_b[K + k], xyz_b[2 * K + k], tb[k])))
        continue;
      const double ph = (u[k] * l[p] + v[k] * m[p]) + w[k] * n[p];
      const double f = ph - nearbyint(ph);
      const double a = 2.0 * PI * f;
      acc += yr[k] * cos(a) - yi[k] * sin(a);
    }
    x[p] = herm * acc;
  }
  free(yr);
  free(yi);
  free(ta);
  free(tb);
}

/* Number of (k, p) pairs with mask = 1 (the evaluated-pair count used for
 * throughput accounting). */
int64_t oracle_count_visible(const double* l, const double* m, const double* n, int64_t P, const double* xyz_a,
                             const double* xyz_b, int64_t K, double R_moon, double orbit_radius) {
  int64_t total = 0;
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : total)
  for (int64_t k = 0; k < K; ++k) {
    const double xa = xyz_a[k], ya = xyz_a[K + k], za = xyz_a[2 * K + k];
    const double xb = xyz_b[k], yb = xyz_b[K + k], zb = xyz_b[2 * K + k];
    const double ta = oracle_mask_threshold(xa, ya, za, R_moon, orbit_radius);
    const double tb = oracle_mask_threshold(xb, yb, zb, R_moon, orbit_radius);
    int64_t c = 0;
    for (int64_t p = 0; p < P; ++p)
      c += visible(l[p], m[p], n[p], xa, ya, za, ta) && visible(l[p], m[p], n[p], xb, yb, zb, tb);
    total += c;
  }
  return total;
}
