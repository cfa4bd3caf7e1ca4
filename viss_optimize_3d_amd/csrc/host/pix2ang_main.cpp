// pix2ang -- drop-in for the reference's pix2ang executable (pix2ang.cu:163-276).
//
// Reference contract: no arguments; reads ./B.txt whose first token is npix
// (pix2ang.cu:166-176), sets nside = round(sqrt(npix/12)) (:178), runs NEST
// pix2ang on the GPU and writes ./theta_heal.txt and ./phi_heal.txt, one float
// per line at ostream's default precision (:42-58, :266-267).
//
// Same defaults here; the GPU work goes through the C ABI (viss_pix2ang).
// Options (none needed for the drop-in):
//   --in FILE        sky file whose first token is npix   (default B.txt)
//   --theta FILE     theta output                          (default theta_heal.txt)
//   --phi FILE       phi output                            (default phi_heal.txt)
//   --nside N        take nside from the flag instead of the sky file
//   --scheme nest|ring                                     (default nest, as the reference)
//   --precision P    write doubles with P digits instead of the reference's float text
//   --device D       HIP device                            (default 0)
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "textio.h"
#include "viss.h"

static int usage() {
  fprintf(stderr,
          "usage: pix2ang [--in B.txt] [--theta theta_heal.txt] [--phi phi_heal.txt] [--nside N]\n"
          "               [--scheme nest|ring] [--precision P] [--device D]\n");
  return 2;
}

#define HIPCHK(x)                                                                  \
  do {                                                                             \
    hipError_t e_ = (x);                                                           \
    if (e_ != hipSuccess) {                                                        \
      fprintf(stderr, "pix2ang: %s failed: %s\n", #x, hipGetErrorString(e_));      \
      return 1;                                                                    \
    }                                                                              \
  } while (0)

int main(int argc, char** argv) {
  std::string in = "B.txt", out_t = "theta_heal.txt", out_p = "phi_heal.txt";
  long long nside = -1;
  int scheme = VISS_SCHEME_NEST, precision = 0, device = 0;
  for (int i = 1; i < argc; ++i) {
    const std::string a = argv[i];
    auto next = [&](void) -> const char* { return i + 1 < argc ? argv[++i] : nullptr; };
    const char* v = nullptr;
    if (a == "--in" && (v = next())) in = v;
    else if (a == "--theta" && (v = next())) out_t = v;
    else if (a == "--phi" && (v = next())) out_p = v;
    else if (a == "--nside" && (v = next())) nside = atoll(v);
    else if (a == "--scheme" && (v = next())) {
      if (!strcmp(v, "nest")) scheme = VISS_SCHEME_NEST;
      else if (!strcmp(v, "ring")) scheme = VISS_SCHEME_RING;
      else return usage();
    } else if (a == "--precision" && (v = next())) precision = atoi(v);
    else if (a == "--device" && (v = next())) device = atoi(v);
    else return usage();
  }
  if (nside <= 0) {
    std::string text;
    long long npix = 0;
    if (vissio::slurp(in, text)) {
      std::vector<double> first = vissio::parse_numbers(text.substr(0, 64));
      if (!first.empty()) npix = (long long)first[0];
    }
    std::printf("npix: %lld\n", npix);
    nside = llround(std::sqrt((double)npix / 12.0));
    if (nside < 1) {
      fprintf(stderr, "pix2ang: cannot read npix from %s\n", in.c_str());
      return 1;
    }
  }
  std::printf("nside: %lld\n", nside);
  HIPCHK(hipSetDevice(device));
  const long long P = 12 * nside * nside;
  double *th = nullptr, *ph = nullptr;
  HIPCHK(hipMalloc((void**)&th, sizeof(double) * P));
  HIPCHK(hipMalloc((void**)&ph, sizeof(double) * P));
  const int rc = viss_pix2ang(nside, scheme, th, ph, nullptr, nullptr, nullptr);
  if (rc != VISS_OK) {
    fprintf(stderr, "pix2ang: viss_pix2ang: %s\n", viss_last_error());
    return 1;
  }
  std::vector<double> ht(P), hp(P);
  HIPCHK(hipMemcpy(ht.data(), th, sizeof(double) * P, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(hp.data(), ph, sizeof(double) * P, hipMemcpyDeviceToHost));
  HIPCHK(hipFree(th));
  HIPCHK(hipFree(ph));
  if (!vissio::write_column(out_t, ht.data(), P, precision) || !vissio::write_column(out_p, hp.data(), P, precision)) {
    fprintf(stderr, "pix2ang: cannot write %s / %s\n", out_t.c_str(), out_p.c_str());
    return 1;
  }
  return 0;
}
