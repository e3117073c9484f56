// Output writers of a sweep: MAT v4 `Zc.mat` (fasterhenry-cli/src/mat.rs:31-95)
// and the SPICE subcircuit stamp (fasterhenry-cli/src/spice.rs:37-126).
// Host code; the formats are byte-for-byte those of the reference.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "fh_host.hpp"

namespace fh {
namespace {

void put_i32(std::vector<uint8_t> &out, int32_t v) {
  uint8_t b[4];
  std::memcpy(b, &v, 4);  // little-endian host (x86-64)
  out.insert(out.end(), b, b + 4);
}

void put_f64(std::vector<uint8_t> &out, double v) {
  uint8_t b[8];
  std::memcpy(b, &v, 8);
  out.insert(out.end(), b, b + 8);
}

// One MAT v4 record: mopt 0 (double, full, little-endian), m, n, imagf,
// namlen, the name's bytes (no terminator), then the real block and, for a
// complex matrix, the imaginary block, column-major.
void matrix(std::vector<uint8_t> &out, const std::string &name, int32_t m, int32_t n, bool complex_,
            const std::vector<double> &re, const std::vector<double> &im) {
  put_i32(out, 0);
  put_i32(out, m);
  put_i32(out, n);
  put_i32(out, complex_ ? 1 : 0);
  put_i32(out, static_cast<int32_t>(name.size()));
  out.insert(out.end(), name.begin(), name.end());
  for (double v : re) put_f64(out, v);
  if (complex_)
    for (double v : im) put_f64(out, v);
}

// Rust's `{:.17e}`: 17 digits after the point, exponent without sign for
// positive values and without leading zeros ("1.50000000000000000e-9").
std::string rust_exp17(double v) {
  if (std::isnan(v)) return "NaN";
  char buf[64];
  std::snprintf(buf, sizeof buf, "%.17e", v);
  std::string s(buf);
  const size_t e = s.find('e');
  if (e == std::string::npos) return s;  // inf / nan
  std::string mant = s.substr(0, e), ex = s.substr(e + 1);
  bool neg = false;
  size_t i = 0;
  if (i < ex.size() && (ex[i] == '+' || ex[i] == '-')) {
    neg = ex[i] == '-';
    ++i;
  }
  while (i + 1 < ex.size() && ex[i] == '0') ++i;
  return mant + "e" + (neg ? "-" : "") + ex.substr(i);
}

}  // namespace
}  // namespace fh

using namespace fh;

extern "C" {

fh_status fh_zc_mat(const double *freqs, size_t K, const double *z, size_t p, uint8_t *out, size_t capacity,
                    size_t *size) {
  clear_error();
  if (!size || (K && !freqs) || (K && p && !z)) return set_error(FH_E_INVALID_ARGUMENT, "null argument");
  if (K > INT32_MAX || p > INT32_MAX) return set_error(FH_E_INVALID_ARGUMENT, "sweep too large for MAT v4");
  std::vector<uint8_t> bytes;
  matrix(bytes, "freqs", 1, static_cast<int32_t>(K), false, std::vector<double>(freqs, freqs + K), {});
  for (size_t k = 0; k < K; ++k) {
    std::vector<double> re, im;
    re.reserve(p * p);
    im.reserve(p * p);
    for (size_t col = 0; col < p; ++col)
      for (size_t row = 0; row < p; ++row) {
        re.push_back(z[2 * ((k * p + row) * p + col)]);
        im.push_back(z[2 * ((k * p + row) * p + col) + 1]);
      }
    matrix(bytes, "Zc_" + std::to_string(k + 1), static_cast<int32_t>(p), static_cast<int32_t>(p), true, re, im);
  }
  *size = bytes.size();
  if (out) {
    if (capacity < bytes.size()) return set_error(FH_E_INVALID_ARGUMENT, "capacity too small");
    std::memcpy(out, bytes.data(), bytes.size());
  }
  return FH_OK;
}

fh_status fh_spice_subckt(const double *freqs, size_t K, const double *z, size_t p, size_t index,
                          const char *name, char *out, size_t capacity, size_t *length) {
  clear_error();
  if (!length || !name || !freqs || (p && !z)) return set_error(FH_E_INVALID_ARGUMENT, "null argument");
  if (index >= K) return set_error(FH_E_INVALID_ARGUMENT, "frequency index out of range");
  const double f = freqs[index];
  const double omega = 6.283185307179586 * f;  // TAU * f
  const double *zk = z + 2 * index * p * p;
  auto resistance = [&](size_t i, size_t j) { return zk[2 * (i * p + j)]; };
  auto inductance = [&](size_t i, size_t j) { return zk[2 * (i * p + j) + 1] / omega; };
  const std::string nm(name);
  std::string s;
  s += "* fasterhenry SPICE export: " + nm + " at " + rust_exp17(f) + " Hz\n";
  s += "* port current is positive into p<i>; Z from the R (H sources) and L (coupled inductors) stamps\n";
  s += ".subckt " + nm + " ";
  for (size_t i = 0; i < p; ++i)
    s += (i ? " p" : "p") + std::to_string(i + 1) + " n" + std::to_string(i + 1);
  s += "\n";
  // p<i> -- L<i> -- H_i_1 -- ... -- vsense<i> -- n<i>
  for (size_t i = 0; i < p; ++i) {
    const double l = f > 0.0 ? inductance(i, i) : 0.0;
    s += "L" + std::to_string(i + 1) + " p" + std::to_string(i + 1) + " a" + std::to_string(i) + "_0 " +
         rust_exp17(l) + "\n";
    std::string node = "a" + std::to_string(i) + "_0";
    for (size_t j = 0; j < p; ++j) {
      const double r = resistance(i, j);
      if (r != 0.0) {
        const std::string next = "a" + std::to_string(i) + "_" + std::to_string(j + 1);
        s += "H" + std::to_string(i + 1) + "_" + std::to_string(j + 1) + " " + node + " " + next + " vsense" +
             std::to_string(j + 1) + " " + rust_exp17(r) + "\n";
        node = next;
      }
    }
    s += "vsense" + std::to_string(i + 1) + " " + node + " n" + std::to_string(i + 1) + " 0\n";
  }
  if (f > 0.0)
    for (size_t i = 0; i < p; ++i)
      for (size_t j = i + 1; j < p; ++j) {
        const double den = std::sqrt(inductance(i, i) * inductance(j, j));
        double c = inductance(i, j) / den;
        const double lim = 1.0 - 1e-12;
        if (c < -lim) c = -lim;  // f64::clamp(-(1 - 1e-12), 1 - 1e-12); NaN stays NaN
        if (c > lim) c = lim;
        s += "K" + std::to_string(i + 1) + "_" + std::to_string(j + 1) + " L" + std::to_string(i + 1) + " L" +
             std::to_string(j + 1) + " " + rust_exp17(c) + "\n";
      }
  s += ".ends " + nm + "\n";
  *length = s.size();
  if (out) {
    if (capacity < s.size() + 1) return set_error(FH_E_INVALID_ARGUMENT, "capacity too small");
    std::memcpy(out, s.c_str(), s.size() + 1);
  }
  return FH_OK;
}

fh_status fh_subckt_name_from_path(const char *path, char *out, size_t capacity, size_t *length) {
  clear_error();
  if (!path || !length) return set_error(FH_E_INVALID_ARGUMENT, "null argument");
  // Path::file_stem: the last component (".." and "/" have none), without
  // its last extension unless the name starts with the only dot.
  std::string p(path);
  for (;;) {  // trailing separators and "." components do not name the file
    if (p.size() > 1 && p.back() == '/') p.pop_back();
    else if (p.size() > 2 && p.compare(p.size() - 2, 2, "/.") == 0) p.resize(p.size() - 2);
    else break;
  }
  const size_t slash = p.rfind('/');
  std::string file = slash == std::string::npos ? p : p.substr(slash + 1);
  if (file == "." || file == ".." || file == "/") file.clear();
  std::string stem = file;
  const size_t dot = file.rfind('.');
  if (dot != std::string::npos && dot > 0) stem = file.substr(0, dot);
  // every character outside ASCII [A-Za-z0-9_] -> one '_'; a UTF-8
  // sequence is one character
  std::string name;
  for (size_t i = 0; i < stem.size();) {
    const unsigned char c = static_cast<unsigned char>(stem[i]);
    size_t len = 1;
    if (c >= 0xF0) len = 4;
    else if (c >= 0xE0) len = 3;
    else if (c >= 0xC0) len = 2;
    const bool alnum = (c >= '0' && c <= '9') || (c >= 'a' && c <= 'z') || (c >= 'A' && c <= 'Z');
    if (alnum || c == '_') name += static_cast<char>(c);
    else name += '_';
    i += len;
  }
  if (name.empty()) name = "fasterhenry";
  *length = name.size();
  if (out) {
    if (capacity < name.size() + 1) return set_error(FH_E_INVALID_ARGUMENT, "capacity too small");
    std::memcpy(out, name.c_str(), name.size() + 1);
  }
  return FH_OK;
}

}  // extern "C"
