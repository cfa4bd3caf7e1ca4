// Text I/O in the reference's on-disk layouts (whitespace-separated numbers,
// outputs one value per line at ostream's default precision).
//
// Parsing and formatting run on several host threads: at 4e7 rows a day the
// text files hold ~1e9 numbers, and one thread of strtod / ostream would
// dominate a day's wall clock.
#pragma once

#include <algorithm>
#include <cerrno>
#include <charconv>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

namespace vissio {

// Read a whole file; returns false if it cannot be opened.
inline bool slurp(const std::string& path, std::string& out) {
  std::ifstream f(path, std::ios::binary);
  if (!f.is_open()) return false;
  f.seekg(0, std::ios::end);
  const std::streamoff n = f.tellg();
  f.seekg(0, std::ios::beg);
  if (n < 0) {
    std::ostringstream ss;
    ss << f.rdbuf();
    out = ss.str();
    return true;
  }
  out.resize((size_t)n);
  if (n > 0) f.read(&out[0], n);
  return (bool)f || f.eof();
}

inline int io_threads() {
  const unsigned hc = std::thread::hardware_concurrency();
  return (int)std::max(1u, std::min(hc, 16u));
}

inline bool is_space(char c) { return c == ' ' || c == '\n' || c == '\t' || c == '\r'; }

// One token starting at p (not whitespace).  Decimal tokens with at most 19
// significant digits whose value is m * 10^e with m < 2^53 and |e| <= 22 are
// converted with one IEEE multiplication or division of exact operands, so
// they are correctly rounded, as strtod's result is (Clinger's fast path).
// Every other token goes to strtod, which also defines where it ends.
// Returns false when the token is not a number.
inline bool parse_token(const char* p, const char* end, double& out, const char*& next) {
  static const double kPow10[23] = {1e0,  1e1,  1e2,  1e3,  1e4,  1e5,  1e6,  1e7,  1e8,  1e9,  1e10, 1e11,
                                    1e12, 1e13, 1e14, 1e15, 1e16, 1e17, 1e18, 1e19, 1e20, 1e21, 1e22};
  const char* s = p;
  bool neg = false;
  if (s < end && (*s == '-' || *s == '+')) {
    neg = *s == '-';
    ++s;
  }
  uint64_t m = 0;
  int nd = 0, frac = 0;
  bool any = false, fast = true;
  while (s < end && *s >= '0' && *s <= '9') {
    if (nd < 19) {
      m = m * 10 + (uint64_t)(*s - '0');
      if (m) ++nd;
    } else {
      fast = false;
    }
    any = true;
    ++s;
  }
  if (s < end && *s == '.') {
    ++s;
    while (s < end && *s >= '0' && *s <= '9') {
      if (nd < 19) {
        m = m * 10 + (uint64_t)(*s - '0');
        if (m) ++nd;
        ++frac;
      } else {
        fast = false;
      }
      any = true;
      ++s;
    }
  }
  int e = 0;
  if (any && s < end && (*s == 'e' || *s == 'E')) {
    const char* t = s + 1;
    bool eneg = false;
    if (t < end && (*t == '-' || *t == '+')) {
      eneg = *t == '-';
      ++t;
    }
    if (t < end && *t >= '0' && *t <= '9') {
      int ev = 0;
      while (t < end && *t >= '0' && *t <= '9') {
        if (ev < 100000) ev = ev * 10 + (*t - '0');
        ++t;
      }
      e = eneg ? -ev : ev;
      s = t;
    }
  }
  const int e10 = e - frac;
  const bool ends = s >= end || is_space(*s);
  if (any && fast && ends && m < (1ull << 53) && e10 >= -22 && e10 <= 22) {
    double v = (double)m;
    v = e10 >= 0 ? v * kPow10[e10] : v / kPow10[-e10];
    out = neg ? -v : v;
    next = s;
    return true;
  }
  // std::from_chars is correctly rounded too (17-digit text, whose
  // mantissas exceed 2^53, lands here).  It takes no '+' sign and no hex, and
  // reports out-of-range values as errors: its result is used only when it
  // read the whole token, else strtod decides.
  {
    const char* b = (p < end && *p == '+') ? p + 1 : p;
    double v;
    const std::from_chars_result r = std::from_chars(b, end, v);
    if (r.ec == std::errc() && r.ptr > b && (r.ptr >= end || is_space(*r.ptr)) && (b == p || *b != '-')) {
      out = v;
      next = r.ptr;
      return true;
    }
  }
  // strtod needs a terminated string: copy the token (bounded by whitespace)
  const char* q = p;
  while (q < end && !is_space(*q)) ++q;
  std::string tok(p, q);
  char* stop = nullptr;
  errno = 0;
  const double d = strtod(tok.c_str(), &stop);
  if (stop == tok.c_str()) return false;
  out = d;
  next = p + (stop - tok.c_str());
  return true;
}

// Numbers in [p, end) until the first token that is not a number.  Returns
// false if it stopped early.
inline bool parse_range(const char* p, const char* end, std::vector<double>& v) {
  while (p < end) {
    while (p < end && is_space(*p)) ++p;
    if (p >= end) break;
    double d;
    const char* next = p;
    if (!parse_token(p, end, d, next)) return false;
    v.push_back(d);
    p = next;
  }
  return true;
}

// Parse whitespace-separated numbers (the grammar operator>> and strtod
// accept for the files the reference reads).  Stops at the first token that
// is not a number.  Large texts are split at whitespace into one range per
// thread; the ranges are joined in order.
inline std::vector<double> parse_numbers(const std::string& text) {
  std::vector<double> v;
  const char* b = text.data();
  const char* e = b + text.size();
  const int nt = text.size() < (size_t)(4 << 20) ? 1 : io_threads();
  if (nt == 1) {
    v.reserve(text.size() / 8);
    parse_range(b, e, v);
    return v;
  }
  std::vector<const char*> cut(nt + 1);
  cut[0] = b;
  cut[nt] = e;
  for (int i = 1; i < nt; ++i) {
    const char* c = b + text.size() * i / nt;
    c = std::max(c, cut[i - 1]);
    while (c < e && !is_space(*c)) ++c;
    cut[i] = c;
  }
  std::vector<std::vector<double>> part(nt);
  std::vector<char> ok(nt, 1);
  std::vector<std::thread> th;
  for (int i = 0; i < nt; ++i)
    th.emplace_back([&, i] {
      part[i].reserve((size_t)(cut[i + 1] - cut[i]) / 8);
      ok[i] = parse_range(cut[i], cut[i + 1], part[i]) ? 1 : 0;
    });
  for (auto& t : th) t.join();
  size_t n = 0;
  for (int i = 0; i < nt; ++i) {
    n += part[i].size();
    if (!ok[i]) break;
  }
  v.reserve(n);
  for (int i = 0; i < nt; ++i) {
    v.insert(v.end(), part[i].begin(), part[i].end());
    if (!ok[i]) break;
  }
  return v;
}

// x[i] followed by '\n', as `ostream << value << '\n'` writes it:
// precision <= 0 is the reference's `file << (float)value` (6 significant
// digits, general format); otherwise the double with that many digits.
// std::to_chars with a precision formats as printf("%.*g") does.
inline void format_line(double x, int precision, std::string& out) {
  char buf[64];
  std::to_chars_result r;
  if (precision > 0)
    r = std::to_chars(buf, buf + sizeof(buf) - 1, x, std::chars_format::general, precision);
  else
    r = std::to_chars(buf, buf + sizeof(buf) - 1, (float)x, std::chars_format::general, 6);
  *r.ptr++ = '\n';
  out.append(buf, r.ptr);
}

// Write one value per line (formatting split over threads, written in order).
inline bool write_column(const std::string& path, const double* x, size_t n, int precision) {
  std::ofstream f(path, std::ios::binary);
  if (!f.is_open()) return false;
  const size_t chunk = 1 << 20;
  const int nt = n < (1 << 18) ? 1 : io_threads();
  std::vector<std::string> bufs(nt);
  for (size_t base = 0; base < n; base += chunk * nt) {
    std::vector<std::thread> th;
    for (int i = 0; i < nt; ++i) {
      const size_t s0 = base + (size_t)i * chunk, s1 = std::min(n, s0 + chunk);
      bufs[i].clear();
      if (s0 >= s1) continue;
      auto work = [&, i, s0, s1] {
        bufs[i].reserve((s1 - s0) * 14);
        for (size_t k = s0; k < s1; ++k) format_line(x[k], precision, bufs[i]);
      };
      if (nt == 1)
        work();
      else
        th.emplace_back(work);
    }
    for (auto& t : th) t.join();
    for (int i = 0; i < nt; ++i) f.write(bufs[i].data(), (std::streamsize)bufs[i].size());
  }
  return (bool)f;
}

}  // namespace vissio
