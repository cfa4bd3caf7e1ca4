// Checks vissio::parse_numbers and vissio::write_column (the drivers' text
// I/O) against strtod and snprintf: parsed values bit-identical, written bytes
// identical, on random values at the precisions the reference's files use and
// on edge-case tokens.  Exit status 0 when everything matches.
#include <cinttypes>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "textio.h"

static std::vector<double> parse_ref(const std::string& text) {
  std::vector<double> v;
  const char* p = text.c_str();
  const char* end = p + text.size();
  while (p < end) {
    while (p < end && vissio::is_space(*p)) ++p;
    if (p >= end) break;
    char* q = nullptr;
    const double d = strtod(p, &q);
    if (q == p) break;
    v.push_back(d);
    p = q;
  }
  return v;
}

static bool same_bits(double a, double b) {
  uint64_t x, y;
  memcpy(&x, &a, 8);
  memcpy(&y, &b, 8);
  return x == y || (a != a && b != b);
}

static int check_parse(const std::string& text, const char* what) {
  const std::vector<double> a = vissio::parse_numbers(text), b = parse_ref(text);
  if (a.size() != b.size()) {
    printf("FAIL parse %s: %zu numbers vs %zu from strtod\n", what, a.size(), b.size());
    return 1;
  }
  for (size_t i = 0; i < a.size(); ++i)
    if (!same_bits(a[i], b[i])) {
      printf("FAIL parse %s: #%zu %.17g vs strtod %.17g\n", what, i, a[i], b[i]);
      return 1;
    }
  printf("parse %s: %zu numbers identical\n", what, a.size());
  return 0;
}

int main() {
  std::mt19937_64 rng(12345);
  std::uniform_real_distribution<double> u(0.0, 1.0);
  int bad = 0;
  // random values at several magnitudes and precisions, mixed separators
  std::vector<double> vals;
  for (int i = 0; i < 400000; ++i) {
    const double mag = std::pow(10.0, (int)(u(rng) * 40.0) - 20);
    double x = (u(rng) - 0.5) * 2.0 * mag;
    if (i % 7 == 0) x = (double)(float)x;
    vals.push_back(x);
  }
  std::string text;
  char buf[64];
  const int precs[] = {6, 9, 17};
  const char* seps[] = {" ", "\n", "\t", "  ", "\r\n"};
  for (size_t i = 0; i < vals.size(); ++i) {
    snprintf(buf, sizeof buf, "%.*g", precs[i % 3], vals[i]);
    text += buf;
    text += seps[i % 5];
  }
  bad |= check_parse(text, "random (multi-threaded)");
  bad |= check_parse(text.substr(0, 4000), "random (one thread)");
  // edge tokens, then a token that is not a number (parsing stops there)
  const char* edge =
      "+1.5 -0 0 00001 .5 5. +.25 1e+06 1E-06 -1.7976931348623157e308 4.9406564584124654e-324 "
      "2.2250738585072014e-308 1e-400 1e400 inf -inf nan 0x1p3 123456789012345678901234567890 "
      "0.00000000000000000000012345 9007199254740993 9007199254740992 1e22 1e23 1e-22 1e-23 "
      "3.14159265358979323846264338327950288 -0.000123457 1.5abc 7 8";
  bad |= check_parse(edge, "edge tokens");
  // tokens past the short-mantissa fast path: signs, specials, hex, overflow,
  // the two sides of the smallest subnormal's rounding, and "+-1" (not a number)
  bad |= check_parse("+inf -nan +0x1p3 -0x1.8p-2 1.7976931348623159e308 2.4703282292062327e-324 "
                     "2.4703282292062326e-324 12345678901234567.5e-3 -98765432109876543e+5 +12345678901234567 "
                     "+-1 9",
                     "long mantissas and specials");
  bad |= check_parse("1 2 3 x 4 5", "stop at a word");
  // malformed and truncated tokens, huge exponents and digit strings, no trailing newline
  bad |= check_parse("1e", "exponent without digits");
  bad |= check_parse("2 1e+", "exponent sign without digits");
  bad |= check_parse("3 -", "a lone sign");
  bad |= check_parse("4 .", "a lone point");
  bad |= check_parse("5 +.e5", "point and exponent without digits");
  bad |= check_parse("6 1.5e-", "truncated negative exponent");
  bad |= check_parse("7 1e999999999999999999 -1e-999999999999999999 8", "huge exponents");
  bad |= check_parse("9 0e123456789 0.0000e-99999", "zero with huge exponents");
  bad |= check_parse(std::string(1000, '7') + " " + "0." + std::string(800, '0') + "1 " + std::string(400, '9') +
                         "e-390",
                     "long digit strings");
  bad |= check_parse("nan(123) nan(abc) -nan() infinity -INFINITY 1", "nan payloads and spelled infinity");
  bad |= check_parse("1 2 3", "no trailing newline");
  bad |= check_parse("\n\n-0.5", "leading blank lines");
  // fuzz: random strings over the characters numbers are made of
  {
    const char alpha[] = "0123456789+-.eEinfaxp \n\t";
    std::uniform_int_distribution<int> pick(0, (int)sizeof(alpha) - 2), len(0, 40);
    int fuzz_bad = 0;
    for (int t = 0; t < 20000 && !fuzz_bad; ++t) {
      std::string f;
      const int nt = 1 + t % 7;
      for (int i = 0; i < nt; ++i) {
        const int L = len(rng);
        for (int j = 0; j < L; ++j) f += alpha[pick(rng)];
        f += (t & 1) ? " " : "\n";
      }
      if (t % 3 == 0 && !f.empty()) f.pop_back();  // no trailing separator
      const std::vector<double> a = vissio::parse_numbers(f), b = parse_ref(f);
      bool same = a.size() == b.size();
      for (size_t i = 0; same && i < a.size(); ++i) same = same_bits(a[i], b[i]);
      if (!same) {
        printf("FAIL parse fuzz #%d: \"%s\": %zu vs %zu numbers\n", t, f.c_str(), a.size(), b.size());
        fuzz_bad = 1;
      }
    }
    if (!fuzz_bad) printf("parse fuzz: 20000 random strings identical to strtod\n");
    bad |= fuzz_bad;
  }
  // a large buffer split over threads where a split point falls inside long tokens
  {
    std::string big;
    for (int i = 0; i < 200000; ++i) big += (i % 13 == 0) ? std::string(300, '1') + "e-290 " : "12.5 ";
    bad |= check_parse(big, "long tokens across thread splits");
  }
  bad |= check_parse("", "empty");
  bad |= check_parse("   \n\t  ", "whitespace only");
  // the whole exponent range (random finite bit patterns, subnormals among
  // them) at 15 to 19 significant digits (the Eisel-Lemire path) and 20 (past it)
  {
    std::string t;
    for (int i = 0; i < 300000; ++i) {
      uint64_t b = rng();
      if ((b >> 52 & 0x7FF) == 0x7FF) b &= ~(1ull << 62);  // no inf / nan
      if (i % 10 == 0) b &= (1ull << 63) | ((1ull << 52) - 1);  // subnormal
      double x;
      memcpy(&x, &b, 8);
      snprintf(buf, sizeof buf, "%.*g ", 15 + i % 6, x);
      t += buf;
    }
    bad |= check_parse(t, "full exponent range, 15-20 digits");
  }
  // ties: decimal values exactly halfway between two doubles, at most 19
  // digits (round half to even), and their neighbours one unit in the last
  // printed digit away
  {
    std::string t;
    for (int i = 0; i < 40000; ++i) {
      const int k = i % 5;  // n in [2^(53-k), 2^(54-k)): ulp 2^-k, midpoint n + 2^-(k+1)
      const uint64_t n = (1ull << (53 - k)) + (rng() >> (11 + k));
      static const char* frac[] = {".5", ".25", ".125", ".0625", ".03125"};
      if (k == 0) {  // integers past 2^53: ulp 2, midpoint odd
        const uint64_t m = (1ull << 53) + 2 * (rng() >> 12) + 1;
        snprintf(buf, sizeof buf, "%" PRIu64 " %" PRIu64 " %" PRIu64 " ", m, m - 1, m + 1);
      } else {
        snprintf(buf, sizeof buf, "%" PRIu64 "%s ", n - 1, frac[k - 1]);
      }
      t += buf;
    }
    t += "9007199254740993 9007199254740995 4503599627370496.5 4503599627370497.5 "
         "2.2250738585072011e-308 2.2250738585072012e-308 2.2250738585072014e-308 "
         "4.9406564584124654e-324 2.4703282292062327e-324 2.4703282292062328e-324 "
         "1.7976931348623157e308 1.7976931348623158e308 1.7976931348623159e308 "
         "9999999999999999999e-343 9999999999999999999e288 1e-342 1e-343 1e308 1e309 "
         "-0 0e400 -0e-400 1844674407370955161e1 18446744073709551615 ";
    bad |= check_parse(t, "halfway and boundary values");
  }
  // parse_numbers_into: the numbers after `skip`, into a vector that is not
  // zero-filled (multi-threaded and one thread)
  for (const size_t skip : {(size_t)0, (size_t)3, (size_t)1000000})
    for (const size_t len : {text.size(), (size_t)4000}) {
      const std::string sub = text.substr(0, len);
      const std::vector<double> ref = parse_ref(sub);
      vissio::HostVec got;
      vissio::parse_numbers_into(sub.data(), sub.data() + sub.size(), skip, got);
      bool same = got.size() == (ref.size() > skip ? ref.size() - skip : 0);
      for (size_t i = 0; same && i < got.size(); ++i) same = same_bits(got[i], ref[i + skip]);
      if (!same) {
        printf("FAIL parse_numbers_into skip %zu on %zu bytes\n", skip, len);
        bad = 1;
      } else {
        printf("parse_numbers_into skip %zu on %zu bytes: %zu numbers identical\n", skip, len, got.size());
      }
    }
  // slurp of a file read by several threads (32 MB and more) into a reused buffer
  {
    const char* path = "textio_check_big.tmp";
    std::string want;
    want.reserve(48u << 20);
    while (want.size() < (40u << 20)) want += text;
    FILE* f = fopen(path, "wb");
    fwrite(want.data(), 1, want.size(), f);
    fclose(f);
    std::string got(1000, 'x');
    const bool ok1 = vissio::slurp(path, got) && got == want;
    const bool ok2 = vissio::slurp(path, got) && got == want;  // again into the same buffer
    std::remove(path);
    if (!ok1 || !ok2) {
      printf("FAIL slurp of %zu bytes\n", want.size());
      bad = 1;
    } else {
      printf("slurp: %zu bytes identical, twice\n", want.size());
    }
  }
  // writer: bytes as snprintf("%.6g") of the float value, or "%.*g" of the double
  for (int prec : {0, 9, 17}) {
    const char* path = "textio_check.tmp";
    std::vector<double> w(vals.begin(), vals.begin() + 300000);
    w.push_back(0.0);
    w.push_back(-0.0);
    w.push_back(1.0 / 0.0);
    w.push_back(-1.0 / 0.0);
    w.push_back(1e-45);
    if (!vissio::write_column(path, w.data(), w.size(), prec)) {
      printf("FAIL write: cannot write %s\n", path);
      return 1;
    }
    std::string got;
    vissio::slurp(path, got);
    std::remove(path);
    std::string want;
    for (double x : w) {
      if (prec > 0)
        snprintf(buf, sizeof buf, "%.*g\n", prec, x);
      else
        snprintf(buf, sizeof buf, "%.6g\n", (double)(float)x);
      want += buf;
    }
    if (got != want) {
      size_t i = 0;
      while (i < got.size() && i < want.size() && got[i] == want[i]) ++i;
      printf("FAIL write precision %d: bytes differ at %zu (%zu vs %zu bytes)\n", prec, i, got.size(), want.size());
      bad = 1;
    } else {
      printf("write precision %d: %zu lines, %zu bytes identical\n", prec, w.size(), got.size());
    }
  }
  printf(bad ? "FAILED\n" : "OK\n");
  return bad;
}
