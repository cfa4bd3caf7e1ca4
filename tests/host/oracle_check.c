/* Drives every function of oracle/viss_oracle.c on small and edge-case inputs,
 * built with -fsanitize=address,undefined by tests/test_host_sanitize.py: an
 * out-of-bounds access, a use of uninitialised or freed memory, an undefined
 * shift, overflow or cast aborts the run.  It also checks a few facts each
 * call must satisfy (finite outputs, counts that add up), so that the calls
 * are made with meaningful arguments.  Exit status 0 when all pass. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

int oracle_pix2ang_nest(int64_t nside, double* theta, double* phi, int32_t* ring, int32_t* iphi);
int oracle_pix2ang_ring(int64_t nside, double* theta, double* phi, int32_t* ring, int32_t* iphi);
void oracle_pixel_dirs(const double* theta, const double* phi, const double* B, int64_t P, double scale, double* l,
                       double* m, double* n, double* I);
double oracle_mask_threshold(double x, double y, double z, double R_moon, double orbit_radius);
void oracle_forward(const double* l, const double* m, const double* n, const double* I, int64_t P, const double* u,
                    const double* v, const double* w, const double* xyz_a, const double* xyz_b, int64_t K, int flags,
                    double R_moon, double orbit_radius, double* V_re, double* V_im);
void oracle_forward_ref32(const float* l, const float* m, const float* n, const float* B, int64_t P, const float* u,
                          const float* v, const float* w, const float* xyz_a, const float* xyz_b, int64_t K,
                          float phi_cut, float* V_re, float* V_im);
void oracle_mask_flips(const double* l, const double* m, const double* n, int64_t P, const double* xyz_a,
                       const double* xyz_b, int64_t K, float phi_cut, double R_moon, double orbit_radius,
                       int64_t* out);
int32_t oracle_shell_index(float bll);
int64_t oracle_weights(const float* bll, const double* w, int64_t K, int32_t nr, int flags, int32_t* gs, uint32_t* mb,
                       double* dcf, double* dg, double* wt);
void oracle_backproject(const double* l, const double* m, const double* n, int64_t P, const double* u, const double* v,
                        const double* w, const double* xyz_a, const double* xyz_b, const double* y_re,
                        const double* y_im, const double* wt, int64_t K, int flags, double R_moon,
                        double orbit_radius, double* x);
int64_t oracle_count_visible(const double* l, const double* m, const double* n, int64_t P, const double* xyz_a,
                             const double* xyz_b, int64_t K, double R_moon, double orbit_radius);

static int failures = 0;
#define CHECK(c, ...)                  \
  do {                                 \
    if (!(c)) {                        \
      printf("FAIL: " __VA_ARGS__);    \
      printf("\n");                    \
      ++failures;                      \
    }                                  \
  } while (0)

static double urand(uint64_t* s) {
  *s = *s * 6364136223846793005ull + 1442695040888963407ull;
  return (double)(*s >> 11) * (1.0 / 9007199254740992.0);
}

static void* xalloc(size_t n) {
  void* p = malloc(n ? n : 1);
  if (!p) {
    printf("FAIL: out of memory\n");
    exit(2);
  }
  return p;
}

int main(void) {
  uint64_t seed = 7;
  /* pix2ang: every nside from 1 to 8 and two larger powers of two, both schemes, nullable outputs */
  const int64_t nsides[] = {1, 2, 3, 4, 5, 7, 8, 16, 64};
  for (size_t i = 0; i < sizeof nsides / sizeof nsides[0]; ++i) {
    const int64_t ns = nsides[i], P = 12 * ns * ns;
    double* th = xalloc(sizeof(double) * P);
    double* ph = xalloc(sizeof(double) * P);
    int32_t* ring = xalloc(sizeof(int32_t) * P);
    int32_t* iphi = xalloc(sizeof(int32_t) * P);
    for (int scheme = 0; scheme < 2; ++scheme) {
      const int rc = scheme ? oracle_pix2ang_nest(ns, th, ph, ring, iphi) : oracle_pix2ang_ring(ns, th, ph, ring, iphi);
      if (scheme && (ns & (ns - 1))) { /* NEST needs a power-of-two nside */
        CHECK(rc == -1, "NEST nside %lld rejected", (long long)ns);
        continue;
      }
      CHECK(rc == 0, "pix2ang nside %lld scheme %d rc %d", (long long)ns, scheme, rc);
      for (int64_t p = 0; p < P; ++p)
        CHECK(th[p] >= 0 && th[p] <= M_PI && ph[p] >= 0 && ph[p] < 2 * M_PI && ring[p] >= 1 &&
                  ring[p] <= 4 * ns - 1 && iphi[p] >= 1,
              "pix2ang nside %lld scheme %d pixel %lld", (long long)ns, scheme, (long long)p);
      CHECK((scheme ? oracle_pix2ang_nest(ns, th, ph, NULL, NULL) : oracle_pix2ang_ring(ns, th, ph, NULL, NULL)) == 0,
            "pix2ang without ring / iphi");
    }
    free(th);
    free(ph);
    free(ring);
    free(iphi);
  }
  CHECK(oracle_pix2ang_ring(0, NULL, NULL, NULL, NULL) == -1, "nside 0 rejected");
  CHECK(oracle_pix2ang_nest(1 << 14, NULL, NULL, NULL, NULL) == -1, "nside 2^14 rejected");

  /* a small sky and rows: forward / backproject / counts with every flag combination */
  const int64_t ns = 4, P = 12 * ns * ns, K = 37;
  double *th = xalloc(8 * P), *ph = xalloc(8 * P), *B = xalloc(8 * P), *l = xalloc(8 * P), *m = xalloc(8 * P),
         *n = xalloc(8 * P), *I = xalloc(8 * P);
  oracle_pix2ang_ring(ns, th, ph, NULL, NULL);
  for (int64_t p = 0; p < P; ++p) B[p] = 1.0 + urand(&seed);
  oracle_pixel_dirs(th, ph, B, P, 4 * M_PI / P, l, m, n, I);
  oracle_pixel_dirs(th, ph, NULL, P, 1.0, l, m, n, NULL);
  for (int64_t p = 0; p < P; ++p) CHECK(fabs(l[p] * l[p] + m[p] * m[p] + n[p] * n[p] - 1) < 1e-14, "unit direction");
  double *u = xalloc(8 * K), *v = xalloc(8 * K), *w = xalloc(8 * K), *xa = xalloc(24 * K), *xb = xalloc(24 * K);
  for (int64_t k = 0; k < K; ++k) {
    u[k] = 3000 * (urand(&seed) - 0.5);
    v[k] = 3000 * (urand(&seed) - 0.5);
    w[k] = 3000 * (urand(&seed) - 0.5);
    for (int c = 0; c < 3; ++c) {
      xa[c * K + k] = 2.0e6 * (urand(&seed) - 0.5);
      xb[c * K + k] = 2.0e6 * (urand(&seed) - 0.5);
    }
  }
  xa[0] = xa[K] = xa[2 * K] = 0.0; /* a satellite at the origin (inside the moon) */
  double *Vr = xalloc(16 * K), *Vi = xalloc(16 * K), *yr = xalloc(8 * K), *yi = xalloc(8 * K), *wt = xalloc(8 * K),
         *x = xalloc(8 * P);
  for (int64_t k = 0; k < K; ++k) {
    yr[k] = urand(&seed) - 0.5;
    yi[k] = urand(&seed) - 0.5;
    wt[k] = urand(&seed);
  }
  for (int flags = 0; flags < 32; ++flags) {
    if (flags & 4) continue; /* VISS_DCF_BLOCKAGE_SHIFT is a weights flag */
    for (int cut = 0; cut < 2; ++cut) {
      const double orb = cut ? 2037.1e3 : 0.0;
      oracle_forward(l, m, n, I, P, u, v, w, xa, xb, K, flags, 1737.4e3, orb, Vr, Vi);
      const int64_t rows = (flags & 16) ? 2 * K : K;
      for (int64_t k = 0; k < rows; ++k) CHECK(isfinite(Vr[k]) && isfinite(Vi[k]), "forward flags %d row %lld", flags, (long long)k);
      oracle_backproject(l, m, n, P, u, v, w, xa, xb, yr, yi, (flags & 1) ? wt : NULL, K, flags, 1737.4e3, orb, x);
      for (int64_t p = 0; p < P; ++p) CHECK(isfinite(x[p]), "backproject flags %d pixel %lld", flags, (long long)p);
    }
  }
  const int64_t cnt = oracle_count_visible(l, m, n, P, xa, xb, K, 1737.4e3, 0.0);
  CHECK(cnt > 0 && cnt <= P * K, "count_visible %lld", (long long)cnt);
  int64_t flips[3] = {-1, -1, -1};
  oracle_mask_flips(l, m, n, P, xa, xb, K, 2.1204f, 1737.1e3, 2037.1e3, flips);
  CHECK(flips[0] >= 0 && flips[1] >= 0 && flips[2] >= 0 && flips[0] + flips[1] + flips[2] <= P * K, "mask_flips");
  {
    float *fl = xalloc(4 * P), *fm = xalloc(4 * P), *fn = xalloc(4 * P), *fB = xalloc(4 * P), *fu = xalloc(4 * K),
          *fv = xalloc(4 * K), *fw = xalloc(4 * K), *fa = xalloc(12 * K), *fb = xalloc(12 * K), *fr = xalloc(4 * K),
          *fi = xalloc(4 * K);
    for (int64_t p = 0; p < P; ++p) {
      fl[p] = (float)l[p];
      fm[p] = (float)m[p];
      fn[p] = (float)n[p];
      fB[p] = (float)B[p];
    }
    for (int64_t k = 0; k < K; ++k) {
      fu[k] = (float)u[k];
      fv[k] = (float)v[k];
      fw[k] = (float)w[k];
    }
    for (int64_t i = 0; i < 3 * K; ++i) {
      fa[i] = (float)xa[i];
      fb[i] = (float)xb[i];
    }
    oracle_forward_ref32(fl, fm, fn, fB, P, fu, fv, fw, fa, fb, K, 2.1204f, fr, fi);
    for (int64_t k = 1; k < K; ++k) CHECK(isfinite(fr[k]) && isfinite(fi[k]), "forward_ref32 row %lld", (long long)k);
    free(fl); free(fm); free(fn); free(fB); free(fu); free(fv); free(fw); free(fa); free(fb); free(fr); free(fi);
  }
  /* empty inputs */
  oracle_forward(l, m, n, I, P, u, v, w, xa, xb, 0, 1, 1737.4e3, 0.0, Vr, Vi);
  oracle_backproject(l, m, n, 0, u, v, w, xa, xb, yr, yi, NULL, K, 1, 1737.4e3, 0.0, x);
  CHECK(oracle_count_visible(l, m, n, 0, xa, xb, K, 1737.4e3, 0.0) == 0, "count_visible of no pixels");

  /* weights: shell indices at and beyond the table, NaN, infinities, negative lengths, |w| > bll */
  const float special[] = {0.0f, 0.25f, 0.2500001f, 0.75f, 1.0f, 1e3f, 3.5e4f, 1e30f, -1.0f, -1e30f, INFINITY,
                           -INFINITY, NAN, 2147483520.0f, 2147483648.0f, 1e-30f};
  const int64_t KW = 64;
  float* bll = xalloc(4 * KW);
  double *ww = xalloc(8 * KW), *dg = xalloc(8 * KW), *wtw = xalloc(8 * KW);
  int32_t* gs = xalloc(4 * KW);
  for (int64_t k = 0; k < KW; ++k) {
    bll[k] = k < 16 ? special[k] : (float)(100.0 * urand(&seed));
    ww[k] = (urand(&seed) - 0.5) * 3.0 * (double)(k < 16 ? 1.0f : bll[k]);
  }
  const int32_t nrs[] = {1, 2, 50, 200};
  for (size_t i = 0; i < sizeof nrs / sizeof nrs[0]; ++i) {
    const int32_t nr = nrs[i];
    uint32_t* mb = xalloc(4 * (size_t)nr);
    double* dcf = xalloc(8 * ((size_t)nr + 1));
    for (int shift = 0; shift < 2; ++shift) {
      const int64_t bad = oracle_weights(bll, ww, KW, nr, shift ? 4 : 0, gs, mb, dcf, dg, wtw);
      int64_t in = 0, outside = 0, counted = 0;
      for (int64_t k = 0; k < KW; ++k) {
        if (gs[k] >= 1 && gs[k] <= nr) ++in;
        if (gs[k] < 0 || gs[k] > nr) ++outside;
        CHECK(gs[k] == oracle_shell_index(bll[k]), "shell index row %lld", (long long)k);
        CHECK(wtw[k] == wtw[k], "weight of row %lld is a number", (long long)k);
      }
      for (int32_t j = 0; j < nr; ++j) counted += mb[j];
      CHECK(counted == in, "histogram counts %lld of %lld rows in range", (long long)counted, (long long)in);
      CHECK(bad == outside, "out-of-range rows %lld vs %lld", (long long)bad, (long long)outside);
    }
    free(mb);
    free(dcf);
  }
  CHECK(oracle_weights(bll, ww, 0, 5, 0, gs, (uint32_t[5]){0}, (double[6]){0}, dg, wtw) == 0, "weights of no rows");
  CHECK(isfinite(oracle_mask_threshold(0, 0, 0, 1737.4e3, 0.0)), "threshold at the origin");

  free(th); free(ph); free(B); free(l); free(m); free(n); free(I); free(u); free(v); free(w); free(xa); free(xb);
  free(Vr); free(Vi); free(yr); free(yi); free(wt); free(x); free(bll); free(ww); free(dg); free(wtw); free(gs);
  printf(failures ? "FAILED (%d)\n" : "OK\n", failures);
  return failures != 0;
}
