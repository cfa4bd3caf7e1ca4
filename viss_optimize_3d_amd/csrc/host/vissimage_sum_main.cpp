// vissimage_sum -- the multi-day aggregation the reference's viewer notebooks
// do on the CPU (healpix3d_noblockage.ipynb / healpix3d_blockage.ipynb:
// total = loadtxt(C1); for i in 2..450: total += [nan_to_num](loadtxt(Ci))),
// as an executable over the C ABI: per-day images are read (text, or the
// binary sidecars vissgen3d --binary writes), summed on the GPU with
// viss_image_accumulate, optionally reordered NEST -> RING (viss_reorder), and
// written once.
//
//   vissimage_sum --dir 3dblockage [--days 450] [--first-day 1] [--nan-to-zero]
//                 [--binary] [--to-ring NSIDE] [--out total.txt] [--precision P] [--device D]
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "textio.h"
#include "viss.h"

#define HIPCHK(x)                                                                      \
  do {                                                                                 \
    hipError_t e_ = (x);                                                               \
    if (e_ != hipSuccess) {                                                            \
      fprintf(stderr, "vissimage_sum: %s failed: %s\n", #x, hipGetErrorString(e_));    \
      return 1;                                                                        \
    }                                                                                  \
  } while (0)

static int usage() {
  fprintf(stderr,
          "usage: vissimage_sum --dir DIR [--days N] [--first-day D] [--nan-to-zero] [--binary]\n"
          "                     [--to-ring NSIDE] [--out FILE] [--precision P] [--device D]\n");
  return 2;
}

static bool read_binary(const std::string& path, std::vector<double>& v) {
  FILE* f = fopen(path.c_str(), "rb");
  if (!f) return false;
  fseek(f, 0, SEEK_END);
  const long n = ftell(f);
  fseek(f, 0, SEEK_SET);
  v.resize((size_t)n / sizeof(double));
  const size_t got = fread(v.data(), sizeof(double), v.size(), f);
  fclose(f);
  return got == v.size();
}

int main(int argc, char** argv) {
  std::string dir, out = "total.txt";
  int days = 450, first = 1, precision = 0, device = 0;
  long long to_ring = 0;
  bool nan0 = false, binary = false;
  for (int i = 1; i < argc; ++i) {
    const std::string a = argv[i];
    auto next = [&](void) -> const char* { return i + 1 < argc ? argv[++i] : nullptr; };
    const char* v = nullptr;
    if (a == "--dir" && (v = next())) dir = v;
    else if (a == "--days" && (v = next())) days = atoi(v);
    else if (a == "--first-day" && (v = next())) first = atoi(v);
    else if (a == "--nan-to-zero") nan0 = true;
    else if (a == "--binary") binary = true;
    else if (a == "--to-ring" && (v = next())) to_ring = atoll(v);
    else if (a == "--out" && (v = next())) out = v;
    else if (a == "--precision" && (v = next())) precision = atoi(v);
    else if (a == "--device" && (v = next())) device = atoi(v);
    else return usage();
  }
  if (dir.empty() || days < 1) return usage();
  if (dir.back() != '/') dir += '/';
  HIPCHK(hipSetDevice(device));
  std::vector<double> img;
  double *d_img = nullptr, *d_total = nullptr;
  int64_t P = -1;
  int used = 0;
  for (int d = first; d < first + days; ++d) {
    const std::string base = dir + "C" + std::to_string(d) + "day1M";
    img.clear();
    if (binary) {
      if (!read_binary(base + ".bin", img)) {
        fprintf(stderr, "vissimage_sum: missing %s.bin, skipped\n", base.c_str());
        continue;
      }
    } else {
      std::string text;
      if (!vissio::slurp(base + ".txt", text)) {
        fprintf(stderr, "vissimage_sum: missing %s.txt, skipped\n", base.c_str());
        continue;
      }
      img = vissio::parse_numbers(text);  // strtod reads "nan" / "-nan" as written by ostream
    }
    if (P < 0) {
      P = (int64_t)img.size();
      HIPCHK(hipMalloc((void**)&d_img, sizeof(double) * P));
      HIPCHK(hipMalloc((void**)&d_total, sizeof(double) * P));
      HIPCHK(hipMemset(d_total, 0, sizeof(double) * P));
    }
    if ((int64_t)img.size() != P) {
      fprintf(stderr, "vissimage_sum: %s has %zu values, expected %lld\n", base.c_str(), img.size(), (long long)P);
      return 1;
    }
    HIPCHK(hipMemcpy(d_img, img.data(), sizeof(double) * P, hipMemcpyHostToDevice));
    if (viss_image_accumulate(d_img, P, nan0 ? 1 : 0, d_total, nullptr) != VISS_OK) {
      fprintf(stderr, "vissimage_sum: %s\n", viss_last_error());
      return 1;
    }
    ++used;
  }
  if (P < 0) {
    fprintf(stderr, "vissimage_sum: no image found in %s\n", dir.c_str());
    return 1;
  }
  if (to_ring > 0) {
    if (12 * to_ring * to_ring != P) {
      fprintf(stderr, "vissimage_sum: --to-ring %lld does not match %lld pixels\n", to_ring, (long long)P);
      return 1;
    }
    if (viss_reorder(d_total, to_ring, 1, d_img, nullptr) != VISS_OK) {
      fprintf(stderr, "vissimage_sum: %s\n", viss_last_error());
      return 1;
    }
    std::swap(d_img, d_total);
  }
  std::vector<double> total(P);
  HIPCHK(hipMemcpy(total.data(), d_total, sizeof(double) * P, hipMemcpyDeviceToHost));
  HIPCHK(hipFree(d_img));
  HIPCHK(hipFree(d_total));
  if (!vissio::write_column(out, total.data(), P, precision)) {
    fprintf(stderr, "vissimage_sum: cannot write %s\n", out.c_str());
    return 1;
  }
  printf("summed %d images of %lld pixels into %s\n", used, (long long)P, out.c_str());
  return 0;
}
