// Text I/O in the reference's on-disk layouts (whitespace-separated numbers,
// outputs one value per line at ostream's default precision).
#pragma once

#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

namespace vissio {

// Read a whole file; returns false if it cannot be opened.
inline bool slurp(const std::string& path, std::string& out) {
  std::ifstream f(path, std::ios::binary);
  if (!f.is_open()) return false;
  std::ostringstream ss;
  ss << f.rdbuf();
  out = ss.str();
  return true;
}

// Parse whitespace-separated numbers with strtod (the same grammar as
// operator>> for the files the reference reads).  Stops at the first token
// that is not a number.
inline std::vector<double> parse_numbers(const std::string& text) {
  std::vector<double> v;
  v.reserve(text.size() / 8);
  const char* p = text.c_str();
  const char* end = p + text.size();
  while (p < end) {
    while (p < end && (*p == ' ' || *p == '\n' || *p == '\t' || *p == '\r')) ++p;
    if (p >= end) break;
    char* q = nullptr;
    errno = 0;
    const double d = strtod(p, &q);
    if (q == p) break;
    v.push_back(d);
    p = q;
  }
  return v;
}

// Write one value per line.  precision <= 0 reproduces the reference's
// `file << (float)value << std::endl` (ostream default: 6 significant digits,
// general format); otherwise the double is written with that many digits.
inline bool write_column(const std::string& path, const double* x, size_t n, int precision) {
  std::ofstream f(path);
  if (!f.is_open()) return false;
  if (precision > 0) f.precision(precision);
  std::string buf;
  buf.reserve(1 << 20);
  std::ostringstream line;
  if (precision > 0) line.precision(precision);
  for (size_t i = 0; i < n; ++i) {
    line.str("");
    if (precision > 0)
      line << x[i] << '\n';
    else
      line << (float)x[i] << '\n';
    buf += line.str();
    if (buf.size() > (1 << 20) - 64) {
      f << buf;
      buf.clear();
    }
  }
  f << buf;
  return (bool)f;
}

}  // namespace vissio
