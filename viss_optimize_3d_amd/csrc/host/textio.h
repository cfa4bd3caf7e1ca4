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
#include <cstring>
#include <fstream>
#include <memory>
#include <sstream>
#include <string>
#include <thread>
#include <utility>
#include <vector>

#include <fcntl.h>
#include <sys/stat.h>
#include <unistd.h>

#include "pow5_table.h"

namespace vissio {

inline int io_threads() {
  const unsigned hc = std::thread::hardware_concurrency();
  return (int)std::max(1u, std::min(hc, 16u));
}

// Read a whole file into out; returns false if it cannot be opened.  out's
// storage is reused (resize from its current size), so a caller that passes
// the same string for files of similar size pays the page faults once.
// Files of 32 MB and more are read by io_threads() threads, each a pread of
// its own slice (one thread copies from the page cache at a few GB/s).
inline bool slurp(const std::string& path, std::string& out) {
  const int fd = ::open(path.c_str(), O_RDONLY);
  if (fd < 0) return false;
  struct stat st;
  if (::fstat(fd, &st) != 0 || !S_ISREG(st.st_mode)) {  // not a regular file: read it as a stream
    ::close(fd);
    std::ifstream f(path, std::ios::binary);
    if (!f.is_open()) return false;
    std::ostringstream ss;
    ss << f.rdbuf();
    out = ss.str();
    return true;
  }
  const size_t n = (size_t)st.st_size;
  out.resize(n);
  const auto read_slice = [&](size_t off, size_t len) {
    while (len > 0) {
      const ssize_t r = ::pread(fd, &out[off], len, (off_t)off);
      if (r <= 0) return r == 0 ? off : (size_t)-1;  // end of file (file shrank), or an error
      off += (size_t)r;
      len -= (size_t)r;
    }
    return off;
  };
  bool ok = true;
  const int nt = n < ((size_t)32 << 20) ? 1 : io_threads();
  if (nt == 1) {
    const size_t got = read_slice(0, n);
    ok = got != (size_t)-1;
    if (ok) out.resize(got);
  } else {
    std::vector<size_t> got(nt);
    std::vector<std::thread> th;
    for (int i = 0; i < nt; ++i) {
      const size_t a = n * i / nt, b = n * (i + 1) / nt;
      th.emplace_back([&, i, a, b] { got[i] = read_slice(a, b - a); });
    }
    for (auto& t : th) t.join();
    for (int i = 0; i < nt && ok; ++i) ok = got[i] == n * (i + 1) / nt;
  }
  ::close(fd);
  return ok;
}

inline bool is_space(char c) { return c == ' ' || c == '\n' || c == '\t' || c == '\r'; }

// w * 10^q (w: the exact decimal mantissa, at most 19 digits) correctly
// rounded to the nearest double, ties to even, by the Eisel-Lemire algorithm
// (D. Lemire, "Number Parsing at a Gigabyte per Second", Software: Practice and
// Experience 51(8), 2021): the top 64 bits of w times a 128-bit truncation of
// 5^q (pow5_table.h) give the 54 leading bits of the value, and a second
// product settles the cases the first leaves open.  Returns false in the rare
// case the 128-bit product cannot decide the rounding; the caller then parses
// the token another way.
inline bool eisel_lemire(uint64_t w, int q, bool neg, double& out) {
  uint64_t bits = 0;
  if (w == 0 || q < kPow5QMin) {
    bits = 0;
  } else if (q > kPow5QMax) {
    bits = 0x7FFull << 52;
  } else {
    const int lz = __builtin_clzll(w);
    w <<= lz;
    const uint64_t* t = kPow5[q - kPow5QMin];
    const unsigned __int128 p1 = (unsigned __int128)w * t[0];
    uint64_t hi = (uint64_t)(p1 >> 64), lo = (uint64_t)p1;
    if ((hi & 0x1FF) == 0x1FF) {  // the 9 bits below the 55 kept may carry: add the second word's product
      const uint64_t h2 = (uint64_t)(((unsigned __int128)w * t[1]) >> 64);
      lo += h2;
      hi += lo < h2;
    }
    if (lo == ~0ull && (q < -27 || q > 55)) return false;
    const int upper = (int)(hi >> 63);
    const int shift = upper + 9;  // keep 55 bits: 53 of the double, one to round, one for `upper`
    uint64_t mant = hi >> shift;
    int e2 = (int)(((152170 + 65536) * (int64_t)q) >> 16) + 63 + upper - lz + 1023;
    if (e2 <= 0) {  // subnormal (or zero)
      if (1 - e2 >= 64) {
        mant = 0;
        e2 = 0;
      } else {
        mant >>= 1 - e2;
        mant += mant & 1;
        mant >>= 1;
        e2 = mant < (1ull << 52) ? 0 : 1;
      }
      bits = (mant & ((1ull << 52) - 1)) | ((uint64_t)e2 << 52);
    } else {
      // halfway between two doubles (possible only where 5^q fits in 64 bits): round to even
      if (lo <= 1 && q >= -4 && q <= 23 && (mant & 3) == 1 && (mant << shift) == hi) mant &= ~1ull;
      mant += mant & 1;
      mant >>= 1;
      if (mant >= (2ull << 52)) {
        mant = 1ull << 52;
        ++e2;
      }
      mant &= ~(1ull << 52);
      bits = e2 >= 0x7FF ? (0x7FFull << 52) : (mant | ((uint64_t)e2 << 52));
    }
  }
  bits |= (uint64_t)neg << 63;
  std::memcpy(&out, &bits, sizeof out);
  return true;
}

// One token starting at p (not whitespace).  Decimal tokens with at most 19
// significant digits whose value is m * 10^e with m < 2^53 and |e| <= 22 are
// converted with one IEEE multiplication or division of exact operands, so
// they are correctly rounded, as strtod's result is (Clinger's fast path);
// other decimal tokens of at most 19 significant digits by eisel_lemire, also
// correctly rounded.  Every other token goes to std::from_chars or strtod,
// which also defines where it ends.  Returns false when the token is not a
// number.
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
  if (any && fast && ends) {
    if (m < (1ull << 53) && e10 >= -22 && e10 <= 22) {
      double v = (double)m;
      v = e10 >= 0 ? v * kPow10[e10] : v / kPow10[-e10];
      out = neg ? -v : v;
      next = s;
      return true;
    }
    // 17-digit text (mantissas past 2^53) and large or small exponents
    if (eisel_lemire(m, e10, neg, out)) {
      next = s;
      return true;
    }
  }
  // std::from_chars is correctly rounded too (mantissas of more than 19
  // digits and the rare undecided Eisel-Lemire cases land here).  It takes no '+' sign and no hex, and
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

// Allocator whose resize leaves new elements uninitialised (no zero fill):
// a vector that is about to be overwritten in parallel then takes its page
// faults in the threads that write it, not in one serial memset.
template <class T>
struct NoInitAlloc : std::allocator<T> {
  template <class U>
  struct rebind {
    using other = NoInitAlloc<U>;
  };
  NoInitAlloc() = default;
  template <class U>
  NoInitAlloc(const NoInitAlloc<U>&) noexcept {}
  template <class U>
  void construct(U* p) noexcept {
    ::new ((void*)p) U;
  }
  template <class U, class... A>
  void construct(U* p, A&&... a) {
    ::new ((void*)p) U(std::forward<A>(a)...);
  }
};
using HostVec = std::vector<double, NoInitAlloc<double>>;

// Parse whitespace-separated numbers (the grammar operator>> and strtod
// accept for the files the reference reads) in [b, e) into out, dropping the
// first `skip`.  Stops at the first token that is not a number; returns false
// if it stopped early.  Large texts are split at whitespace into one range per
// thread; each thread parses its range, and after the counts are known each
// copies its numbers to their place in out.
template <class Vec>
inline bool parse_numbers_into(const char* b, const char* e, size_t skip, Vec& out) {
  const size_t len = (size_t)(e - b);
  const int nt = len < (size_t)(4 << 20) ? 1 : io_threads();
  if (nt == 1) {
    std::vector<double> v;
    v.reserve(len / 8);
    const bool ok = parse_range(b, e, v);
    const size_t k = std::min(skip, v.size());
    out.assign(v.begin() + k, v.end());
    return ok;
  }
  std::vector<const char*> cut(nt + 1);
  cut[0] = b;
  cut[nt] = e;
  for (int i = 1; i < nt; ++i) {
    const char* c = b + len * i / nt;
    c = std::max(c, cut[i - 1]);
    while (c < e && !is_space(*c)) ++c;
    cut[i] = c;
  }
  std::vector<std::vector<double>> part(nt);
  std::vector<char> ok(nt, 1);
  {
    std::vector<std::thread> th;
    for (int i = 0; i < nt; ++i)
      th.emplace_back([&, i] {
        part[i].reserve((size_t)(cut[i + 1] - cut[i]) / 8);
        ok[i] = parse_range(cut[i], cut[i + 1], part[i]) ? 1 : 0;
      });
    for (auto& t : th) t.join();
  }
  // the parts up to and including the first that stopped early
  int last = nt - 1;
  for (int i = 0; i < nt; ++i)
    if (!ok[i]) {
      last = i;
      break;
    }
  std::vector<size_t> dst(last + 2, 0);  // position of part i's first kept number in out
  std::vector<size_t> src(last + 1, 0);  // its first kept number within part i
  size_t before = 0;                     // numbers of the parts before i
  for (int i = 0; i <= last; ++i) {
    src[i] = skip > before ? std::min(skip - before, part[i].size()) : 0;
    dst[i + 1] = dst[i] + (part[i].size() - src[i]);
    before += part[i].size();
  }
  out.resize(dst[last + 1]);
  {
    std::vector<std::thread> th;
    for (int i = 0; i <= last; ++i)
      th.emplace_back([&, i] {
        const size_t n = part[i].size() - src[i];
        if (n) std::memcpy(out.data() + dst[i], part[i].data() + src[i], n * sizeof(double));
        std::vector<double>().swap(part[i]);  // free in the thread that filled it
      });
    for (auto& t : th) t.join();
  }
  return last == nt - 1 && ok[last];
}

inline std::vector<double> parse_numbers(const std::string& text) {
  std::vector<double> v;
  parse_numbers_into(text.data(), text.data() + text.size(), 0, v);
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
