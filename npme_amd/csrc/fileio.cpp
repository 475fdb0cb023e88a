// On-disk formats of the NPME command-line tools: the charge/coordinate
// file (@Coord plus @ChargeReal or @ChargeComplex) and the potential file
// (@Vreal or @Vcomplex), with the field widths, values per line and reading
// rules of src/ReadPrint.cpp:234-739, and the compareV error metric of
// src/SupportFunctions.cpp:88-305.
//
// Files are read into memory whole and parsed with strtod (a 1M-charge
// coordinate file is ~40 MB); writers format into one buffer and write it
// once.
#include <algorithm>
#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "npme_internal.h"

namespace npme {

namespace {

[[noreturn]] void io_error(const std::string &m) { throw Error{NPME_ERR_IO, m}; }

std::string slurp(const char *path)
{
  FILE *f = std::fopen(path, "rb");
  if (!f) io_error(std::string("Error opening ") + path);
  std::string s;
  char buf[1 << 16];
  size_t k;
  while ((k = std::fread(buf, 1, sizeof(buf), f)) > 0) s.append(buf, k);
  std::fclose(f);
  return s;
}

// stream-like cursor: tokens are separated by any white space, as with
// operator>> on an istream; skip_line() moves past the next '\n'
struct Cursor {
  const std::string &s;
  const char *file;
  size_t i = 0;
  Cursor(const std::string &str, const char *f) : s(str), file(f) {}
  bool eof() const { return i >= s.size(); }
  void skip_ws()
  {
    while (i < s.size() && (s[i] == ' ' || s[i] == '\t' || s[i] == '\n' || s[i] == '\r' ||
                            s[i] == '\v' || s[i] == '\f'))
      i++;
  }
  std::string token()
  {
    skip_ws();
    const size_t b = i;
    while (i < s.size() && !std::isspace((unsigned char)s[i])) i++;
    return s.substr(b, i - b);
  }
  double number(const char *what)
  {
    skip_ws();
    if (eof()) io_error(std::string("Error reading ") + file + ": unexpected end of file while reading " + what);
    const char *p = s.c_str() + i;
    char *end = nullptr;
    errno = 0;
    const double v = std::strtod(p, &end);
    if (end == p) io_error(std::string("Error reading ") + file + ": bad number while reading " + what);
    i += (size_t)(end - p);
    return v;
  }
  void skip_line()
  {
    while (i < s.size() && s[i] != '\n') i++;
    if (i < s.size()) i++;
  }
  // the line starting at the cursor, which moves past it
  std::string line()
  {
    const size_t b = i;
    while (i < s.size() && s[i] != '\n') i++;
    std::string l = s.substr(b, i - b);
    if (i < s.size()) i++;
    return l;
  }
  // moves past the first line whose first token is flag
  bool go_to_line(const char *flag)
  {
    while (!eof()) {
      const std::string t = token();
      skip_line();
      if (t == flag) return true;
    }
    return false;
  }
};

std::vector<std::string> words(const std::string &l)
{
  std::vector<std::string> w;
  size_t i = 0;
  while (i < l.size()) {
    while (i < l.size() && std::isspace((unsigned char)l[i])) i++;
    const size_t b = i;
    while (i < l.size() && !std::isspace((unsigned char)l[i])) i++;
    if (i > b) w.push_back(l.substr(b, i - b));
  }
  return w;
}

size_t parse_count(const std::string &t, const char *file, const char *section)
{
  char *end = nullptr;
  const unsigned long long v = std::strtoull(t.c_str(), &end, 10);
  if (t.empty() || *end != '\0')
    io_error(std::string("Error reading ") + file + ": bad nCharge in " + section);
  return (size_t)v;
}

double *copy_out(const std::vector<double> &v)
{
  double *p = static_cast<double *>(std::malloc(std::max<size_t>(v.size(), 1) * sizeof(double)));
  if (!p) throw std::bad_alloc();
  if (!v.empty()) std::memcpy(p, v.data(), v.size() * sizeof(double));
  return p;
}

struct Writer {
  std::string buf;
  void put(const char *s) { buf += s; }
  template <typename... A> void fmt(const char *f, A... a)
  {
    char tmp[128];
    const int k = std::snprintf(tmp, sizeof(tmp), f, a...);
    buf.append(tmp, (size_t)k);
  }
  void save(const char *path) const
  {
    FILE *f = std::fopen(path, "wb");
    if (!f) io_error(std::string("Error opening ") + path);
    const size_t k = std::fwrite(buf.data(), 1, buf.size(), f);
    const int rc = std::fclose(f);
    if (k != buf.size() || rc != 0) io_error(std::string("Error writing ") + path);
  }
};

}  // namespace

}  // namespace npme

using namespace npme;

extern "C" {

// NPME_ReadSingleBoxChargeCoord (src/ReadPrint.cpp:282-440): the first line
// whose first token is @Coord, then "nCharge n" and n lines of which the
// first three numbers are x y z; then the first line starting with
// @ChargeReal or @ChargeComplex, the line "nCharge n", and the n (2n for
// complex, interleaved re, im) values, five per line
int npme_read_coord_charge(const char *path, size_t *nCharge, int *isChargeReal, double **coord,
                           double **charge)
{
  return guarded([&] {
    if (!path || !nCharge || !isChargeReal || !coord || !charge)
      throw Error{NPME_ERR_ARG, "npme_read_coord_charge: NULL argument"};
    *coord = *charge = nullptr;
    const std::string s = slurp(path);
    Cursor c(s, path);
    if (!c.go_to_line("@Coord")) io_error(std::string("Error reading ") + path + ": no @Coord line");
    if (c.token() != "nCharge")
      io_error(std::string("Error reading ") + path + ": while reading @Coord, expecting: nCharge XX");
    const size_t n = parse_count(c.token(), path, "@Coord");
    c.skip_line();
    std::vector<double> xyz(3 * n);
    for (size_t i = 0; i < n; i++) {
      for (int p = 0; p < 3; p++) xyz[3 * i + p] = c.number("coordinates");
      c.skip_line();
    }
    int real = -1;
    while (!c.eof()) {
      const std::string l = c.line();
      if (l.empty() || l[0] != '@') continue;
      const std::string t = words(l)[0];
      if (t == "@ChargeReal") real = 1;
      else if (t == "@ChargeComplex") real = 0;
      if (real >= 0) break;
    }
    if (real < 0)
      io_error(std::string("Error reading ") + path + ": did not find @ChargeReal or @ChargeComplex flags");
    const std::vector<std::string> w = words(c.line());
    if (w.size() < 2 || w[0] != "nCharge")
      io_error(std::string("Error reading ") + path +
               ": after @ChargeReal/@ChargeComplex, expecting: nCharge XX");
    const size_t nq = parse_count(w[1], path, "the charge section");
    if (nq != n)
      io_error(std::string("Error reading ") + path + ": nCharge of the charges (" +
               std::to_string(nq) + ") differs from that of @Coord (" + std::to_string(n) + ")");
    const size_t tot = (real ? 1 : 2) * n;
    std::vector<double> q(tot);
    for (size_t k = 0; k < tot; k++) {
      q[k] = c.number("charges");
      if (k % 5 == 4 || k + 1 == tot) c.skip_line();
    }
    *coord = copy_out(xyz);
    *charge = copy_out(q);
    *nCharge = n;
    *isChargeReal = real;
  });
}

// NPME_PrintSingleBoxChargeCoord (src/ReadPrint.cpp:234-280)
int npme_write_coord_charge(const char *path, size_t nCharge, int isChargeReal, const double *coord,
                            const double *charge)
{
  return guarded([&] {
    if (!path || (nCharge > 0 && (!coord || !charge)))
      throw Error{NPME_ERR_ARG, "npme_write_coord_charge: NULL argument"};
    Writer w;
    w.buf.reserve(nCharge * (isChargeReal ? 52 : 67) + 128);
    w.put("@Coord\n");
    w.fmt("  nCharge %zu\n", nCharge);
    for (size_t i = 0; i < nCharge; i++)
      w.fmt("    %10.6f %10.6f %10.6f\n", coord[3 * i], coord[3 * i + 1], coord[3 * i + 2]);
    w.put("\n\n");
    w.put(isChargeReal ? "@ChargeReal\n" : "@ChargeComplex\n");
    w.fmt("  nCharge %zu", nCharge);
    const size_t tot = (isChargeReal ? 1 : 2) * nCharge;
    for (size_t k = 0; k < tot; k++) {
      if (k % 5 == 0) w.put("\n    ");
      w.fmt("%14.6e ", charge[k]);
    }
    w.put("\n\n");
    w.save(path);
  });
}

// NPME_PrintSingleBox_V (src/ReadPrint.cpp:487-560): four values per line,
// %18.10e (or %24.16e)
int npme_write_v(const char *path, size_t nCharge, int isChargeReal, const double *V, int highPrecision)
{
  return guarded([&] {
    if (!path || (nCharge > 0 && !V)) throw Error{NPME_ERR_ARG, "npme_write_v: NULL argument"};
    const char *f = highPrecision ? "%24.16e " : "%18.10e ";
    Writer w;
    const size_t tot = (isChargeReal ? 4 : 8) * nCharge;
    w.buf.reserve(tot * (highPrecision ? 25 : 19) + tot / 4 * 5 + 64);
    w.put(isChargeReal ? "@Vreal\n" : "@Vcomplex\n");
    w.fmt("  nCharge %zu", nCharge);
    for (size_t k = 0; k < tot; k++) {
      if (k % 4 == 0) w.put("\n    ");
      w.fmt(f, V[k]);
    }
    w.put("\n\n");
    w.save(path);
  });
}

// NPME_ReadSingleBox_V (src/ReadPrint.cpp:565-739): the first line starting
// with @Vreal or @Vcomplex, then "nCharge n" and 4n (8n) values, four per
// line
int npme_read_v(const char *path, size_t *nCharge, int *isChargeReal, double **V)
{
  return guarded([&] {
    if (!path || !nCharge || !isChargeReal || !V) throw Error{NPME_ERR_ARG, "npme_read_v: NULL argument"};
    *V = nullptr;
    const std::string s = slurp(path);
    Cursor c(s, path);
    int real = -1;
    while (!c.eof()) {
      const std::string l = c.line();
      if (l.empty() || l[0] != '@') continue;
      const std::string t = words(l)[0];
      if (t == "@Vreal") real = 1;
      else if (t == "@Vcomplex") real = 0;
      if (real >= 0) break;
    }
    if (real < 0) io_error(std::string("Error reading ") + path + ": did not find @Vreal or @Vcomplex flags");
    if (c.token() != "nCharge")
      io_error(std::string("Error reading ") + path + ": expecting: nCharge XX");
    const size_t n = parse_count(c.token(), path, real ? "@Vreal" : "@Vcomplex");
    c.skip_line();
    const size_t tot = (real ? 4 : 8) * n;
    std::vector<double> v(tot);
    for (size_t k = 0; k < tot; k++) {
      v[k] = c.number("potentials");
      if (k % 4 == 3) c.skip_line();
    }
    *V = copy_out(v);
    *nCharge = n;
    *isChargeReal = real;
  });
}

void npme_free(void *p) { std::free(p); }

// NPME_CompareV1Arrays and NPME_CalcV1AvgVecMag
// (src/SupportFunctions.cpp:88-121,176-305): out[0] = sqrt(mean |dV|^2),
// out[1] = sqrt(mean over the 3n gradient entries |d dV|^2),
// out[2] = mean |Vref|, out[3] = mean over i and p = 0, 1, 2 of
// |Vref[4i + p]| (the reference's average, which includes V and leaves out
// dV/dz); complex entries use the modulus
int npme_compare_v(size_t nCharge, int isChargeReal, const double *V, const double *Vref, double out[4])
{
  return guarded([&] {
    if (!V || !Vref || !out || nCharge == 0) throw Error{NPME_ERR_ARG, "npme_compare_v: bad argument"};
    const int w = isChargeReal ? 1 : 2;
    auto mag = [&](const double *a, size_t k) {
      return w == 1 ? std::fabs(a[k]) : std::hypot(a[2 * k], a[2 * k + 1]);
    };
    auto diff = [&](size_t k) {
      if (w == 1) return std::fabs(V[k] - Vref[k]);
      return std::hypot(V[2 * k] - Vref[2 * k], V[2 * k + 1] - Vref[2 * k + 1]);
    };
    double eV = 0, edV = 0, mV = 0, mdV = 0;
    for (size_t i = 0; i < nCharge; i++) {
      for (int p = 0; p < 4; p++) {
        const double d = diff(4 * i + p);
        (p == 0 ? eV : edV) += d * d;
      }
      mV += mag(Vref, 4 * i);
      for (int p = 0; p < 3; p++) mdV += mag(Vref, 4 * i + p);
    }
    out[0] = std::sqrt(eV / (double)nCharge);
    out[1] = std::sqrt(edV / (3.0 * (double)nCharge));
    out[2] = mV / (double)nCharge;
    out[3] = mdV / (3.0 * (double)nCharge);
  });
}

}  // extern "C"
