// vissgen3d -- drop-in for the reference's vissgen3d_{noblockage,blockage}[_10m]
// executables (vissgen3d_noblockage.cu:186-570 and its three siblings).
//
// Reference contract (no arguments; everything hard-coded per file):
//   in : <dir>/B.txt (npix, then npix brightness values), theta_heal.txt,
//        phi_heal.txt (npix values each) (vissgen3d_noblockage.cu:207-230);
//        per day p = 1..days: uvw{p}day1M.txt (3 header tokens, rows u v w),
//        xyza{p}day1M.txt / xyzb{p}day1M.txt (3 header tokens, rows x y z),
//        bll{p}day1M.txt (1 header token, one value per row) (:302-378);
//   out: <outdir>/C{p}day1M.txt, npix lines of Re(C) (:536-556).
// The 10 MHz variants read B_10Mhz.txt / theta_heal_10Mhz.txt /
// phi_heal_10Mhz.txt from ./earth_10Mhz/ and run one day
// (vissgen3d_noblockage_10m.cu:49,199,208-210,543,569).
//
// Per day the pipeline is the reference's, through the C ABI:
//   forward  (masked in every variant, phase tracked, Hermitian half,
//             vissgen3d_noblockage.cu:400-440)      -> viss_sky_forward
//   weights  (:448-507)                              -> viss_weights
//   imaging  (:511-527; masked in the blockage variants) -> viss_sky_backproject
// with the phase restore of viss_gamma_trans folded into the imaging call.
// When the second half of the rows is the negated first half (the layout the
// reference's conj trick assumes) the imaging folds the two halves into one
// pass (2 Re of the first half), which is exact.
//
// Documented departures (each selectable back where it is well defined):
//   * rows whose shell index exceeds nr (out-of-bounds dcf reads in the
//     reference) extend nr instead (--no-extend-nr: weight 0);
//   * dcf uses the noblockage indexing in every variant
//     (--dcf-blockage-shift: vissgen3d_blockage.cu:138);
//   * all arithmetic on inputs is FP64 (phase) / mixed (--fp64: all FP64);
//   * --binary also writes each image as raw FP64 (C{p}day1M.bin) for fast
//     aggregation (vissimage_sum --binary);
//   * --exact-healpix (opt-in) replaces angle files that are a HEALPix map's
//     angles at reduced precision by that map's exact FP64 angles.
#include <hip/hip_runtime.h>
#include <sys/stat.h>
#include <sys/time.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <future>
#include <string>
#include <vector>

#include "textio.h"
#include "viss.h"

#define HIPCHK(x)                                                                   \
  do {                                                                              \
    hipError_t e_ = (x);                                                            \
    if (e_ != hipSuccess) {                                                         \
      fprintf(stderr, "vissgen3d: %s failed: %s\n", #x, hipGetErrorString(e_));     \
      return 1;                                                                     \
    }                                                                               \
  } while (0)
#define VISSCHK(x)                                                                  \
  do {                                                                              \
    int rc_ = (x);                                                                  \
    if (rc_ != VISS_OK) {                                                           \
      fprintf(stderr, "vissgen3d: %s failed (%d): %s\n", #x, rc_, viss_last_error()); \
      return 1;                                                                     \
    }                                                                               \
  } while (0)

namespace {

struct Options {
  std::string variant = "noblockage";
  std::string in_dir, out_dir, sky_name, theta_name, phi_name;
  double freq = 0;
  int days = -1, first_day = 1;
  bool lambda_c = false;
  double R_moon = 1737.1e3, altitude = 300e3;  // vissgen3d_noblockage.cu:240-241
  bool per_sat_cut = false;
  double bl_max = 100e3;  // :248
  bool dcf_shift = false, extend_nr = true, fp64 = false, resume = false, binary = false, timing = false;
  bool exact_healpix = false;
  int image_mask = -1;  // -1: per variant
  int precision = 0;
  int device = 0;
};

int usage() {
  fprintf(stderr,
          "usage: vissgen3d [--variant noblockage|blockage|noblockage_10m|blockage_10m] [--in DIR] [--out DIR]\n"
          "                 [--sky FILE] [--theta FILE] [--phi FILE] [--freq HZ] [--days N] [--first-day D]\n"
          "                 [--lambda-c] [--R-moon M] [--altitude M] [--per-satellite-cut] [--bl-max M]\n"
          "                 [--dcf-blockage-shift] [--no-extend-nr] [--image-mask on|off] [--fp64]\n"
          "                 [--precision P] [--binary] [--resume] [--timing] [--exact-healpix] [--device D]\n");
  return 2;
}

bool file_exists(const std::string& p) {
  struct stat st;
  return stat(p.c_str(), &st) == 0;
}

// Numbers of a per-day file after `skip` header tokens, grouped by `width`.
// `text` is the caller's buffer, kept from file to file and day to day so that
// the reads after the first go to pages already mapped.
size_t read_rows(const std::string& path, size_t skip, size_t width, std::string& text, vissio::HostVec& out) {
  out.clear();
  if (!vissio::slurp(path, text)) return 0;
  vissio::parse_numbers_into(text.data(), text.data() + text.size(), skip, out);
  const size_t rows = out.size() / width;
  out.resize(rows * width);
  return rows;
}

// Same FP32 expression as the library's shell index (ceilAndScale,
// vissgen3d_noblockage.cu:105-111), for sizing nr on the host.
int shell_index(float bll) {
  volatile float d = bll - 0.25f;
  volatile float q = d / 0.5f;
  const float c = ceilf(q);
  if (!(c > -2147483648.0f && c < 2147483648.0f)) return INT32_MIN;
  return (int)c;
}

template <typename T>
int dev_alloc(T** p, size_t n) {
  if (*p) HIPCHK(hipFree(*p));
  *p = nullptr;
  if (n) HIPCHK(hipMalloc((void**)p, n * sizeof(T)));
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  Options o;
  for (int i = 1; i < argc; ++i) {
    const std::string a = argv[i];
    auto next = [&](void) -> const char* { return i + 1 < argc ? argv[++i] : nullptr; };
    const char* v = nullptr;
    if (a == "--variant" && (v = next())) o.variant = v;
    else if (a == "--in" && (v = next())) o.in_dir = v;
    else if (a == "--out" && (v = next())) o.out_dir = v;
    else if (a == "--sky" && (v = next())) o.sky_name = v;
    else if (a == "--theta" && (v = next())) o.theta_name = v;
    else if (a == "--phi" && (v = next())) o.phi_name = v;
    else if (a == "--freq" && (v = next())) o.freq = atof(v);
    else if (a == "--days" && (v = next())) o.days = atoi(v);
    else if (a == "--first-day" && (v = next())) o.first_day = atoi(v);
    else if (a == "--lambda-c") o.lambda_c = true;
    else if (a == "--R-moon" && (v = next())) o.R_moon = atof(v);
    else if (a == "--altitude" && (v = next())) o.altitude = atof(v);
    else if (a == "--per-satellite-cut") o.per_sat_cut = true;
    else if (a == "--bl-max" && (v = next())) o.bl_max = atof(v);
    else if (a == "--dcf-blockage-shift") o.dcf_shift = true;
    else if (a == "--no-extend-nr") o.extend_nr = false;
    else if (a == "--image-mask" && (v = next())) o.image_mask = !strcmp(v, "on") ? 1 : (!strcmp(v, "off") ? 0 : -2);
    else if (a == "--fp64") o.fp64 = true;
    else if (a == "--precision" && (v = next())) o.precision = atoi(v);
    else if (a == "--resume") o.resume = true;
    else if (a == "--binary") o.binary = true;
    else if (a == "--timing") o.timing = true;
    else if (a == "--exact-healpix") o.exact_healpix = true;
    else if (a == "--device" && (v = next())) o.device = atoi(v);
    else return usage();
  }
  const bool blockage = o.variant.rfind("blockage", 0) == 0;
  const bool ten = o.variant.size() > 4 && o.variant.substr(o.variant.size() - 4) == "_10m";
  if (o.variant != "noblockage" && o.variant != "blockage" && o.variant != "noblockage_10m" &&
      o.variant != "blockage_10m")
    return usage();
  if (o.image_mask == -2) return usage();
  // per-variant constants of the four reference files
  if (o.in_dir.empty()) o.in_dir = ten ? "./earth_10Mhz/" : "./earth_1Mhz/";
  if (o.in_dir.back() != '/') o.in_dir += '/';
  if (o.out_dir.empty()) o.out_dir = std::string("3d") + (blockage ? "blockage" : "noblockage") + (ten ? "10M" : "");
  if (o.out_dir.back() != '/') o.out_dir += '/';
  if (o.sky_name.empty()) o.sky_name = ten ? "B_10Mhz.txt" : "B.txt";
  if (o.theta_name.empty()) o.theta_name = ten ? "theta_heal_10Mhz.txt" : "theta_heal.txt";
  if (o.phi_name.empty()) o.phi_name = ten ? "phi_heal_10Mhz.txt" : "phi_heal.txt";
  if (o.freq <= 0) o.freq = ten ? 1e7 : 1e6;
  if (o.days < 0) o.days = ten ? 1 : 450;
  const bool image_mask = o.image_mask == -1 ? blockage : (o.image_mask == 1);

  struct timeval t_start, t_end;
  gettimeofday(&t_start, nullptr);
  HIPCHK(hipSetDevice(o.device));
  printf("variant: %s\nfrequency: %g\ndays: %d\n", o.variant.c_str(), o.freq, o.days);

  // ---- sky: B, theta, phi (vissgen3d_noblockage.cu:207-230)
  std::string text;
  if (!vissio::slurp(o.in_dir + o.sky_name, text)) {
    fprintf(stderr, "vissgen3d: cannot open %s\n", (o.in_dir + o.sky_name).c_str());
    return 1;
  }
  std::vector<double> Bv = vissio::parse_numbers(text);
  if (Bv.empty() || Bv[0] < 1) {
    fprintf(stderr, "vissgen3d: %s has no npix header\n", o.sky_name.c_str());
    return 1;
  }
  const int64_t P = (int64_t)Bv[0];
  std::vector<double> B(P, 0.0), th(P, 0.0), ph(P, 0.0);
  for (int64_t i = 0; i < P && i + 1 < (int64_t)Bv.size(); ++i) B[i] = Bv[i + 1];
  {
    std::string tt, pt;
    if (!vissio::slurp(o.in_dir + o.theta_name, tt) || !vissio::slurp(o.in_dir + o.phi_name, pt)) {
      fprintf(stderr, "vissgen3d: cannot open theta/phi files in %s\n", o.in_dir.c_str());
      return 1;
    }
    std::vector<double> a = vissio::parse_numbers(tt), b = vissio::parse_numbers(pt);
    for (int64_t i = 0; i < P; ++i) {
      th[i] = i < (int64_t)a.size() ? a[i] : 0.0;
      ph[i] = i < (int64_t)b.size() ? b[i] : 0.0;
    }
  }
  printf("npix: %lld\n", (long long)P);
  const double scale = 4.0 * M_PI / (double)P;  // vissgen3d_noblockage.cu:233
  const float lamda_f = o.lambda_c ? (float)(299792458.0 / o.freq) : (float)(3e8 / o.freq);  // :249
  const float nr_f = ceilf((float)o.bl_max / lamda_f * 2.0f);                                // :250
  const int nr_ref = (int)nr_f;
  printf("lamda: %g\nnr: %d\n", (double)lamda_f, nr_ref);
  viss_mask_geometry geom;
  geom.R_moon = o.R_moon;
  geom.orbit_radius = o.per_sat_cut ? 0.0 : o.R_moon + o.altitude;

  double *d_th = nullptr, *d_ph = nullptr, *d_B = nullptr, *d_lmn = nullptr, *d_I = nullptr, *d_C = nullptr;
  if (dev_alloc(&d_th, P) || dev_alloc(&d_ph, P) || dev_alloc(&d_B, P) || dev_alloc(&d_lmn, 3 * P) ||
      dev_alloc(&d_I, P) || dev_alloc(&d_C, P))
    return 1;
  if (o.exact_healpix) {
    // Opt-in departure from the files: when they are HEALPix angles written
    // at reduced precision (the reference's 6 significant digits move an
    // angle by at most 5e-6 rad), use the exact FP64 angles of the same map,
    // whose symmetries give the quad kernels.  RING and, for a power-of-two
    // nside, NEST order are tried; a file that is not such a map is an error.
    const int64_t nside = (int64_t)llround(std::sqrt((double)P / 12.0));
    if (nside < 1 || 12 * nside * nside != P) {
      fprintf(stderr, "vissgen3d: --exact-healpix: npix %lld is not 12 nside^2\n", (long long)P);
      return 1;
    }
    const double tol = 2e-5;  // rad; the pixel spacing is 1.2e-4 rad even at nside 8192
    std::vector<double> et(P), ep(P), bt, bp;
    double best = INFINITY;
    int best_scheme = -1;
    for (int scheme : {VISS_SCHEME_RING, VISS_SCHEME_NEST}) {
      if (scheme == VISS_SCHEME_NEST && (nside & (nside - 1)) != 0) continue;
      VISSCHK(viss_pix2ang(nside, scheme, d_th, d_ph, nullptr, nullptr, nullptr));
      HIPCHK(hipMemcpy(et.data(), d_th, P * sizeof(double), hipMemcpyDeviceToHost));
      HIPCHK(hipMemcpy(ep.data(), d_ph, P * sizeof(double), hipMemcpyDeviceToHost));
      double dev = 0.0;
      for (int64_t i = 0; i < P; ++i) {
        const double dt = std::fabs(th[i] - et[i]);
        double dp = std::fabs(ph[i] - ep[i]);
        dp = std::min(dp, 2.0 * M_PI - dp);  // phi near 0 and near 2 pi are neighbours
        const double d = std::max(dt, dp);
        if (!(d <= dev)) dev = d == d ? d : INFINITY;  // NaN in a file: no match
      }
      if (dev < best) {
        best = dev;
        best_scheme = scheme;
        bt = et;
        bp = ep;
      }
    }
    if (!(best <= tol)) {
      fprintf(stderr,
              "vissgen3d: --exact-healpix: the angle files are not HEALPix RING or NEST angles of nside %lld "
              "to within %g rad (closest: %g rad)\n",
              (long long)nside, tol, best);
      return 1;
    }
    th.swap(bt);
    ph.swap(bp);
    printf("exact HEALPix angles: %s, nside %lld (the files deviate by at most %.3g rad)\n",
           best_scheme == VISS_SCHEME_NEST ? "NEST" : "RING", (long long)nside, best);
  }
  HIPCHK(hipMemcpy(d_th, th.data(), P * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(d_ph, ph.data(), P * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(d_B, B.data(), P * sizeof(double), hipMemcpyHostToDevice));
  VISSCHK(viss_pixel_dirs(d_th, d_ph, d_B, P, scale, d_lmn, d_I, nullptr));
  viss_sky* sky = nullptr;
  VISSCHK(viss_sky_create(d_lmn, P, &sky, nullptr));
  {
    // The symmetric layouts need the directions' symmetries to hold to 1e-12:
    // angle files at the reference's 6 significant digits break them and run
    // the per-pixel kernels (write them with pix2ang --precision 17 instead).
    int64_t npair = 0, nsingle = 0, nquad = 0, nqt = 0;
    int use_pairs = 0, use_quads = 0;
    VISSCHK(viss_sky_pairs(sky, &npair, &nsingle, &use_pairs));
    VISSCHK(viss_sky_quads(sky, &nquad, &nqt, &use_quads));
    printf("sky: %lld pixels, %lld antipodal pairs, %lld mirror quads: %s kernels%s\n", (long long)P,
           (long long)npair, (long long)nquad, use_quads ? "quad" : (use_pairs ? "pair" : "per-pixel"),
           use_quads == 2 ? " (recurrent phase)" : "");
  }

  double *d_uvw = nullptr, *d_xa = nullptr, *d_xb = nullptr, *d_uvwf = nullptr, *d_xaf = nullptr, *d_xbf = nullptr;
  double *d_w = nullptr, *d_V = nullptr, *d_dcf = nullptr, *d_wt = nullptr;
  float* d_bll = nullptr;
  int32_t* d_gs = nullptr;
  uint32_t* d_mb = nullptr;
  int64_t* d_bad = nullptr;
  size_t cap_rows = 0, cap_nr = 0;
  if (dev_alloc(&d_bad, 1)) return 1;
  std::vector<double> hC(P);
  mkdir(o.out_dir.c_str(), 0755);

  // One day's input, read and parsed on a host thread while the previous day
  // runs on the GPU and is written (the text parse is the largest host cost
  // of a day).
  struct DayInput {
    vissio::HostVec uvw_r, xa_r, xb_r, bll_r;
    size_t n_uvw = 0, n_xa = 0, n_xb = 0, n_bll = 0;
    double s_read = 0.0;
  };
  using clk = std::chrono::steady_clock;
  const auto since = [](clk::time_point a) { return std::chrono::duration<double>(clk::now() - a).count(); };
  // one read runs at a time (the next day's starts after this day's input is
  // taken), so the days share one text buffer
  std::string day_text;
  const auto read_day = [&o, &since, &day_text](int day) {
    const clk::time_point t0 = clk::now();
    const std::string tag = std::to_string(day) + "day1M.txt";
    DayInput in;
    in.n_uvw = read_rows(o.in_dir + "uvw" + tag, 3, 3, day_text, in.uvw_r);
    in.n_xa = read_rows(o.in_dir + "xyza" + tag, 3, 3, day_text, in.xa_r);
    in.n_xb = read_rows(o.in_dir + "xyzb" + tag, 3, 3, day_text, in.xb_r);
    in.n_bll = read_rows(o.in_dir + "bll" + tag, 1, 1, day_text, in.bll_r);
    in.s_read = since(t0);
    return in;
  };
  std::vector<int> todo;
  for (int day = o.first_day; day < o.first_day + o.days; ++day) {
    const std::string out_path = o.out_dir + "C" + std::to_string(day) + "day1M.txt";
    if (o.resume && file_exists(out_path))
      printf("day %d: %s exists, skipped\n", day, out_path.c_str());
    else
      todo.push_back(day);
  }
  std::future<DayInput> pending;
  if (!todo.empty()) pending = std::async(std::launch::async, read_day, todo[0]);

  for (size_t i_day = 0; i_day < todo.size(); ++i_day) {
    const int day = todo[i_day];
    const std::string tag = std::to_string(day) + "day1M.txt";
    const std::string out_path = o.out_dir + "C" + tag;
    const clk::time_point t_wait = clk::now();
    DayInput in = pending.get();
    const double s_wait = since(t_wait);
    if (i_day + 1 < todo.size()) pending = std::async(std::launch::async, read_day, todo[i_day + 1]);
    const vissio::HostVec& uvw_r = in.uvw_r;
    const vissio::HostVec& xa_r = in.xa_r;
    const vissio::HostVec& xb_r = in.xb_r;
    const vissio::HostVec& bll_r = in.bll_r;
    const size_t n_uvw = in.n_uvw, n_xa = in.n_xa, n_xb = in.n_xb, n_bll = in.n_bll;
    printf("day %d: uvw_index %zu xyz1_index %zu xyz2_index %zu bll_index %zu\n", day, n_uvw, n_xa, n_xb, n_bll);
    if (n_xa != n_uvw || n_xb != n_uvw || n_bll != n_uvw)
      fprintf(stderr, "vissgen3d: day %d: row counts differ; missing rows read as 0\n", day);
    const double s_read = in.s_read;
    const clk::time_point t_gpu = clk::now();
    double s_prep = 0.0, s_h2d = 0.0, s_kernels = 0.0;
    const size_t rows = n_uvw;
    const size_t amount = rows / 2;  // vissgen3d_noblockage.cu:400 (integer division)
    std::fill(hC.begin(), hC.end(), 0.0);
    if (rows > 0) {
      // SoA host arrays: all rows, and the first `amount` rows for the forward.
      // The loops over rows run on the host's threads, and every element is
      // written once (missing rows as 0), so no array is zero-filled first.
      const int64_t nrow = (int64_t)rows, namt = (int64_t)amount;
      vissio::HostVec uvw(3 * rows), xa(3 * rows), xb(3 * rows), w(rows);
      std::vector<float, vissio::NoInitAlloc<float>> bll(rows);
#pragma omp parallel for schedule(static)
      for (int64_t k = 0; k < nrow; ++k) {
        for (int c = 0; c < 3; ++c) {
          uvw[c * rows + k] = uvw_r[3 * k + c];
          xa[c * rows + k] = (size_t)k < n_xa ? xa_r[3 * k + c] : 0.0;
          xb[c * rows + k] = (size_t)k < n_xb ? xb_r[3 * k + c] : 0.0;
        }
        w[k] = uvw_r[3 * k + 2];
        bll[k] = (size_t)k < n_bll ? (float)bll_r[k] : 0.0f;
      }
      vissio::HostVec uvwf(3 * amount), xaf(3 * amount), xbf(3 * amount);
#pragma omp parallel for schedule(static)
      for (int64_t k = 0; k < namt; ++k)
        for (int c = 0; c < 3; ++c) {
          uvwf[c * amount + k] = uvw[c * rows + k];
          xaf[c * amount + k] = xa[c * rows + k];
          xbf[c * amount + k] = xb[c * rows + k];
        }
      // shell count: the reference's nr, extended to the largest shell index present
      int nr = nr_ref;
      int64_t n_over = 0;
      const bool extend = o.extend_nr;
#pragma omp parallel for schedule(static) reduction(+ : n_over) reduction(max : nr)
      for (int64_t k = 0; k < nrow; ++k) {
        const int g = shell_index(bll[k]);
        if (g > nr_ref) ++n_over;
        if (extend && g > nr) nr = g;
      }
      if (n_over)
        printf("day %d: %lld rows beyond shell %d (%s)\n", day, (long long)n_over, nr_ref,
               o.extend_nr ? "nr extended" : "weight 0");
      // Hermitian fold check: row k + amount = (-b_k, same bll, satellites swapped or equal)
      bool fold = rows == 2 * amount && amount > 0;
      if (fold) {
        int64_t n_bad = 0;
#pragma omp parallel for schedule(static) reduction(+ : n_bad)
        for (int64_t k = 0; k < namt; ++k) {
          const size_t j = (size_t)k + amount;
          bool ok = bll[j] == bll[k];
          for (int c = 0; c < 3; ++c) ok = ok && uvw[c * rows + j] == -uvw[c * rows + k];
          if (image_mask && ok) {
            bool same = true, swapped = true;
            for (int c = 0; c < 3; ++c) {
              same = same && xa[c * rows + j] == xa[c * rows + k] && xb[c * rows + j] == xb[c * rows + k];
              swapped = swapped && xa[c * rows + j] == xb[c * rows + k] && xb[c * rows + j] == xa[c * rows + k];
            }
            ok = same || swapped;
          }
          n_bad += ok ? 0 : 1;
        }
        fold = n_bad == 0;
      }
      s_prep = since(t_gpu);
      const clk::time_point t_h2d = clk::now();
      if (rows > cap_rows) {
        cap_rows = rows;
        if (dev_alloc(&d_uvw, 3 * rows) || dev_alloc(&d_xa, 3 * rows) || dev_alloc(&d_xb, 3 * rows) ||
            dev_alloc(&d_uvwf, 3 * rows) || dev_alloc(&d_xaf, 3 * rows) || dev_alloc(&d_xbf, 3 * rows) ||
            dev_alloc(&d_w, rows) || dev_alloc(&d_V, 2 * rows) || dev_alloc(&d_wt, rows) || dev_alloc(&d_bll, rows) ||
            dev_alloc(&d_gs, rows))
          return 1;
      }
      if ((size_t)nr > cap_nr) {
        cap_nr = nr;
        if (dev_alloc(&d_mb, nr) || dev_alloc(&d_dcf, nr + 1)) return 1;
      }
      HIPCHK(hipMemcpy(d_uvw, uvw.data(), 3 * rows * sizeof(double), hipMemcpyHostToDevice));
      HIPCHK(hipMemcpy(d_xa, xa.data(), 3 * rows * sizeof(double), hipMemcpyHostToDevice));
      HIPCHK(hipMemcpy(d_xb, xb.data(), 3 * rows * sizeof(double), hipMemcpyHostToDevice));
      if (amount) {
        HIPCHK(hipMemcpy(d_uvwf, uvwf.data(), 3 * amount * sizeof(double), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(d_xaf, xaf.data(), 3 * amount * sizeof(double), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(d_xbf, xbf.data(), 3 * amount * sizeof(double), hipMemcpyHostToDevice));
      }
      HIPCHK(hipMemcpy(d_w, w.data(), rows * sizeof(double), hipMemcpyHostToDevice));
      HIPCHK(hipMemcpy(d_bll, bll.data(), rows * sizeof(float), hipMemcpyHostToDevice));
      s_h2d = since(t_h2d);
      const clk::time_point t_kernels = clk::now();
      HIPCHK(hipMemset(d_V, 0, 2 * rows * sizeof(double)));
      HIPCHK(hipMemset(d_mb, 0, nr * sizeof(uint32_t)));
      HIPCHK(hipMemset(d_bad, 0, sizeof(int64_t)));
      HIPCHK(hipMemset(d_C, 0, P * sizeof(double)));
      const int prec = o.fp64 ? VISS_PREC_FP64 : 0;
      // forward: masked in every variant, phase tracked, conj into the second half
      if (amount)
        VISSCHK(viss_sky_forward(sky, d_I, d_uvwf, d_xaf, d_xbf, (int64_t)amount, &geom,
                                 VISS_MASK | VISS_PHASE_TRACK | VISS_HERMITIAN | prec, d_V, nullptr));
      VISSCHK(viss_weights(d_bll, d_w, (int64_t)rows, nr, o.dcf_shift ? VISS_DCF_BLOCKAGE_SHIFT : 0, d_gs, d_mb,
                           d_dcf, nullptr, d_wt, d_bad, nullptr));
      const int img = VISS_PHASE_TRACK | (image_mask ? VISS_MASK : 0) | prec;
      if (fold)
        VISSCHK(viss_sky_backproject(sky, d_uvwf, d_xaf, d_xbf, d_V, d_wt, (int64_t)amount, &geom, img | VISS_HERMITIAN,
                                     d_C, nullptr));
      else
        VISSCHK(viss_sky_backproject(sky, d_uvw, d_xa, d_xb, d_V, d_wt, (int64_t)rows, &geom, img, d_C, nullptr));
      HIPCHK(hipDeviceSynchronize());
      s_kernels = since(t_kernels);
      HIPCHK(hipMemcpy(hC.data(), d_C, P * sizeof(double), hipMemcpyDeviceToHost));
      printf("day %d: forward %zu rows, imaging %s\n", day, amount, fold ? "folded Hermitian halves" : "all rows");
    }
    const double s_gpu = since(t_gpu);
    const clk::time_point t_write = clk::now();
    if (!vissio::write_column(out_path, hC.data(), P, o.precision)) {
      fprintf(stderr, "vissgen3d: cannot write %s\n", out_path.c_str());
      return 1;
    }
    if (o.binary) {  // raw little-endian FP64 sidecar, C{p}day1M.bin
      const std::string bin_path = o.out_dir + "C" + std::to_string(day) + "day1M.bin";
      FILE* f = fopen(bin_path.c_str(), "wb");
      if (!f || fwrite(hC.data(), sizeof(double), P, f) != (size_t)P) {
        fprintf(stderr, "vissgen3d: cannot write %s\n", bin_path.c_str());
        if (f) fclose(f);
        return 1;
      }
      fclose(f);
    }
    if (o.timing)
      printf("day %d: read %.3f s (4 text files, on a host thread; waited %.3f s for it), host prep + gpu %.3f s "
             "(host loops %.3f, copies in %.3f, kernels %.3f), write %.3f s\n",
             day, s_read, s_wait, s_gpu, s_prep, s_h2d, s_kernels, since(t_write));
    printf("day %d: saved %s\n", day, out_path.c_str());
  }
  viss_sky_destroy(sky);
  gettimeofday(&t_end, nullptr);
  printf("total time: %gs\n", (t_end.tv_sec - t_start.tv_sec) + (t_end.tv_usec - t_start.tv_usec) * 1e-6);
  return 0;
}
