// Host-side frame I/O: fixed-column PDB parsing and row formatting of the
// reference's text outputs, byte-identical to what the reference writes.
//
//   PDB records   CPET/utils/io.py:574-636 (parse_pdb): ATOM/HETATM lines,
//                 serial [6:11], name [12:16], resName [17:20], chain [21],
//                 resSeq [22:26], x/y/z [30:38] [38:46] [46:54], charge
//                 [55:64]; fields are stripped and parsed as Python's int()
//                 and float() do (decimal text -> correctly rounded double),
//                 except that digit-group underscores ("1_000"), which
//                 Python accepts, are rejected.
//   .dat rows     CPET/utils/io.py:97-101 np.savetxt(fmt="%.3f")
//   .top rows     CPET/source/CPET.py:123 np.savetxt default fmt "%.18e"
//                 Values are formatted from the double value of the stored
//                 float16/float32/float64, columns joined by ' ', rows by '\n'.

#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/cpet_hip.h"

namespace cpet {
void set_error(const char* fmt, ...);
}

namespace {

bool read_file(const char* path, std::string& out) {
  FILE* f = std::fopen(path, "rb");
  if (!f) return false;
  std::fseek(f, 0, SEEK_END);
  const long n = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  out.resize(n > 0 ? (size_t)n : 0);
  const size_t got = n > 0 ? std::fread(&out[0], 1, (size_t)n, f) : 0;
  std::fclose(f);
  return got == out.size();
}

// Calls fn(line_start, line_len) for every line of text.
template <typename Fn>
void for_lines(const std::string& s, Fn fn) {
  size_t i = 0;
  while (i < s.size()) {
    size_t j = s.find('\n', i);
    if (j == std::string::npos) j = s.size();
    size_t e = j;
    if (e > i && s[e - 1] == '\r') --e;  // text-mode reading turns CRLF into LF
    fn(s.data() + i, e - i);
    i = j + 1;
  }
}

bool is_record(const char* l, size_t n) {
  return (n >= 4 && std::memcmp(l, "ATOM", 4) == 0) || (n >= 6 && std::memcmp(l, "HETATM", 6) == 0);
}

bool is_ascii(const char* l, size_t n) {
  for (size_t i = 0; i < n; ++i)
    if ((unsigned char)l[i] >= 0x80) return false;
  return true;
}

// line[a:b] with Python slicing (clipped to the line), then strip(); the
// result is a view into the line.
struct Field {
  const char* p;
  size_t n;
};

// str.strip()'s ASCII whitespace: \t \n \v \f \r, \x1c-\x1f and space.
inline bool py_space(unsigned char c) { return c == ' ' || (c >= 9 && c <= 13) || (c >= 0x1c && c <= 0x1f); }

Field field(const char* l, size_t n, size_t a, size_t b) {
  if (a >= n) return {l, 0};
  if (b > n) b = n;
  size_t s = a, e = b;
  while (s < e && py_space((unsigned char)l[s])) ++s;
  while (e > s && py_space((unsigned char)l[e - 1])) --e;
  return {l + s, e - s};
}

// Copies a field into a NUL-terminated buffer; false if empty or too long.
bool terminated(Field f, char* buf, size_t cap) {
  if (f.n == 0 || f.n >= cap) return false;
  std::memcpy(buf, f.p, f.n);
  buf[f.n] = '\0';
  return true;
}

bool parse_int(Field f, int32_t* v) {
  char buf[16];
  if (!terminated(f, buf, sizeof(buf))) return false;
  char* end = nullptr;
  errno = 0;
  const long x = std::strtol(buf, &end, 10);
  if (errno || *end != '\0' || x < INT32_MIN || x > INT32_MAX) return false;
  *v = (int32_t)x;
  return true;
}

// Python's float() accepts decimal literals, inf/infinity and nan; strtod also
// takes hexadecimal ("0x1p3"), which float() rejects.
bool parse_float(Field f, double* v) {
  char buf[64];  // "%.18e" values are 24-25 characters
  if (!terminated(f, buf, sizeof(buf)) || std::strpbrk(buf, "xX")) return false;
  char* end = nullptr;
  const double x = std::strtod(buf, &end);  // correctly rounded (glibc)
  if (*end != '\0') return false;
  *v = x;
  return true;
}

void copy_fixed(char* dst, size_t width, Field f) {
  std::memset(dst, 0, width);
  std::memcpy(dst, f.p, f.n < width ? f.n : width);
}

double half_to_double(uint16_t h) {
  // IEEE 754 binary16 -> value
  const int sign = (h >> 15) & 1, exp = (h >> 10) & 0x1f, man = h & 0x3ff;
  double v;
  if (exp == 0)
    v = std::ldexp((double)man, -24);
  else if (exp == 31)
    v = man ? NAN : INFINITY;
  else
    v = std::ldexp((double)(man | 0x400), exp - 25);
  return sign ? -v : v;
}

double value_at(const void* data, int dtype, size_t i) {
  switch (dtype) {
    case 0: return (double)((const float*)data)[i];
    case 1: return ((const double*)data)[i];
    default: return half_to_double(((const uint16_t*)data)[i]);
  }
}

// "%.3f" of a float32 or float16 value v without printf: v has at most 24
// significant bits, so v * 1000 is exact in double (34 bits), and rounding it
// to an integer with ties to even (the default rounding mode) gives exactly
// the digits printf's correctly rounded conversion prints.  Valid while
// |v * 1000| < 2^53; the sign is printed for every negative v, zero
// included, as printf does ("-0.000").
inline bool fast_fixed3(double v, char* out, int* len) {
  const double t = v * 1000.0;
  if (!(std::fabs(t) < 9007199254740992.0)) return false;
  const uint64_t k = (uint64_t)std::fabs(std::nearbyint(t));
  char tmp[24];
  int n = 0;
  uint64_t ip = k / 1000;
  const unsigned fr = (unsigned)(k % 1000);
  tmp[n++] = (char)('0' + fr % 10);
  tmp[n++] = (char)('0' + (fr / 10) % 10);
  tmp[n++] = (char)('0' + fr / 100);
  tmp[n++] = '.';
  do {
    tmp[n++] = (char)('0' + ip % 10);
    ip /= 10;
  } while (ip);
  int m = 0;
  if (std::signbit(v)) out[m++] = '-';
  while (n) out[m++] = tmp[--n];
  *len = m;
  return true;
}

// Python's "%.3f" / "%.18e" formatting of a float is C's, except that Python
// spells every NaN "nan" (C prints "-nan" for a negative one).
void format_rows(const void* data, int dtype, int64_t r0, int64_t r1, int cols, int fmt,
                 std::string& out) {
  // "%.3f" of the largest double is 309 integer digits + sign + ".000"
  char buf[352];
  const char* f = fmt == 0 ? "%.3f" : "%.18e";
  for (int64_t r = r0; r < r1; ++r) {
    for (int c = 0; c < cols; ++c) {
      const double v = value_at(data, dtype, (size_t)r * cols + c);
      int n;
      if (fmt == 0 && dtype != 1 && fast_fixed3(v, buf, &n)) {
      } else if (std::isnan(v))
        n = std::snprintf(buf, sizeof(buf), "nan");
      else
        n = std::snprintf(buf, sizeof(buf), f, v);
      out.append(buf, n < (int)sizeof(buf) ? (size_t)n : sizeof(buf) - 1);
      out.push_back(c + 1 < cols ? ' ' : '\n');
    }
  }
}

}  // namespace

using cpet::set_error;

extern "C" {

int cpet_pdb_scan(const char* path, int* n_atoms) {
  if (!path || !n_atoms) {
    set_error("cpet_pdb_scan: NULL argument");
    return CPET_ERR_ARG;
  }
  std::string s;
  if (!read_file(path, s)) {
    set_error("cannot read %s", path);
    return CPET_ERR_ARG;
  }
  int n = 0;
  for_lines(s, [&](const char* l, size_t len) { n += is_record(l, len) ? 1 : 0; });
  *n_atoms = n;
  return CPET_OK;
}

int cpet_pdb_parse(const char* path, int n_atoms, int get_charges, double* xyz, double* q,
                   int32_t* atom_number, int32_t* resnum, char* atom_type, char* resname,
                   char* chain) {
  if (!path || n_atoms < 0 || !xyz || !atom_number || !resnum || !atom_type || !resname ||
      !chain || (get_charges && !q)) {
    set_error("cpet_pdb_parse: bad arguments");
    return CPET_ERR_ARG;
  }
  std::string s;
  if (!read_file(path, s)) {
    set_error("cannot read %s", path);
    return CPET_ERR_ARG;
  }
  int i = 0, lineno = 0, rc = CPET_OK;
  for_lines(s, [&](const char* l, size_t n) {
    ++lineno;
    if (rc != CPET_OK || !is_record(l, n)) return;
    if (!is_ascii(l, n)) {
      set_error("non-ASCII bytes in %s line %d", path, lineno);
      rc = CPET_ERR_ARG;
      return;
    }
    if (i >= n_atoms) {
      set_error("%s has more than %d atom records", path, n_atoms);
      rc = CPET_ERR_ARG;
      return;
    }
    double v[4];
    if (!parse_int(field(l, n, 6, 11), &atom_number[i]) ||
        !parse_int(field(l, n, 22, 26), &resnum[i]) || !parse_float(field(l, n, 30, 38), &v[0]) ||
        !parse_float(field(l, n, 38, 46), &v[1]) || !parse_float(field(l, n, 46, 54), &v[2]) ||
        (get_charges && !parse_float(field(l, n, 55, 64), &v[3]))) {
      set_error("%s line %d: invalid literal in a fixed-width field", path, lineno);
      rc = CPET_ERR_ARG;
      return;
    }
    xyz[3 * i] = v[0];
    xyz[3 * i + 1] = v[1];
    xyz[3 * i + 2] = v[2];
    if (get_charges) q[i] = v[3];
    copy_fixed(atom_type + 4 * (size_t)i, 4, field(l, n, 12, 16));
    copy_fixed(resname + 3 * (size_t)i, 3, field(l, n, 17, 20));
    chain[i] = n > 21 ? l[21] : '\0';
    ++i;
  });
  if (rc == CPET_OK && i != n_atoms) {
    set_error("%s has %d atom records, expected %d", path, i, n_atoms);
    rc = CPET_ERR_ARG;
  }
  return rc;
}

// PQR records as parse_pqr reads them (CPET/utils/io.py:478-571): tokens of
// line.split(); a first token longer than 6 characters ("HETATM12345") holds
// the serial from character 6 and shifts the other tokens left by one; the
// residue number is token 5 (4 when shifted), or, when that token is longer
// than 4 characters, token 4 (3) without its first character; residue
// numbers above 999 move the fixed columns right by digits - 3 (the
// reference's int(log10(n) - 2)); x/y/z are columns [30:38] [38:46] [46:54]
// and the charge [54:61], shifted.  Token fields are copied NUL-padded into
// `width`-byte slots.  Any record Python would reject, or that this parser
// does not read the way Python would (non-ASCII bytes, a token longer than
// width - 1, digit-group underscores), returns CPET_ERR_ARG naming the line,
// so the caller can fall back to the line-by-line restatement.
int cpet_pqr_parse(const char* path, int n_atoms, int width, double* xyz, double* q,
                   int32_t* atom_number, int32_t* resnum, char* atom_type, char* resname) {
  if (!path || n_atoms < 0 || width < 2 || !xyz || !q || !atom_number || !resnum ||
      !atom_type || !resname) {
    set_error("cpet_pqr_parse: bad arguments");
    return CPET_ERR_ARG;
  }
  std::string s;
  if (!read_file(path, s)) {
    set_error("cannot read %s", path);
    return CPET_ERR_ARG;
  }
  int i = 0, lineno = 0, rc = CPET_OK;
  for_lines(s, [&](const char* l, size_t n) {
    ++lineno;
    if (rc != CPET_OK || !is_record(l, n)) return;
    auto fail = [&](const char* why) {
      set_error("%s line %d: %s", path, lineno, why);
      rc = CPET_ERR_ARG;
    };
    if (!is_ascii(l, n)) return fail("non-ASCII bytes");
    if (i >= n_atoms) return fail("more atom records than counted");
    Field tok[8];
    int nt = 0;
    for (size_t k = 0; k < n && nt < 8;) {  // the first tokens of line.split()
      while (k < n && py_space((unsigned char)l[k])) ++k;
      size_t e = k;
      while (e < n && !py_space((unsigned char)l[e])) ++e;
      if (e > k) tok[nt++] = Field{l + k, e - k};
      k = e;
    }
    const bool merged = nt > 0 && tok[0].n > 6;
    const int res_ind = merged ? 4 : 5;
    if (nt <= res_ind) return fail("too few fields");
    const Field serial = merged ? Field{tok[0].p + 6, tok[0].n - 6} : tok[1];
    const Field aname = merged ? tok[1] : tok[2];
    const Field rname = merged ? tok[2] : tok[3];
    const Field rtok = tok[res_ind].n > 4 ? Field{tok[res_ind - 1].p + 1, tok[res_ind - 1].n - 1}
                                          : tok[res_ind];
    int32_t rv = 0;
    if (!parse_int(serial, &atom_number[i]) || !parse_int(rtok, &rv))
      return fail("invalid serial or residue number");
    if (aname.n >= (size_t)width || rname.n >= (size_t)width) return fail("token too long");
    int shift = 0;
    if (rv > 999) {
      int digits = 0;
      for (int64_t t = rv; t > 0; t /= 10) ++digits;
      shift = digits - 3;
    }
    double v[4];
    if (!parse_float(field(l, n, 30 + shift, 38 + shift), &v[0]) ||
        !parse_float(field(l, n, 38 + shift, 46 + shift), &v[1]) ||
        !parse_float(field(l, n, 46 + shift, 54 + shift), &v[2]) ||
        !parse_float(field(l, n, 54 + shift, 61 + shift), &v[3]))
      return fail("charge or coordinates are not usable numbers");
    xyz[3 * i] = v[0];
    xyz[3 * i + 1] = v[1];
    xyz[3 * i + 2] = v[2];
    q[i] = v[3];
    resnum[i] = rv;
    copy_fixed(atom_type + (size_t)width * i, (size_t)width, aname);
    copy_fixed(resname + (size_t)width * i, (size_t)width, rname);
    ++i;
  });
  if (rc == CPET_OK && i != n_atoms) {
    set_error("%s has %d atom records, expected %d", path, i, n_atoms);
    rc = CPET_ERR_ARG;
  }
  return rc;
}

int cpet_write_rows(const char* path, const char* header, const void* data, int dtype,
                    int64_t rows, int cols, int fmt, int n_threads) {
  if (!path || !data || rows < 0 || cols <= 0 || dtype < 0 || dtype > 2 || fmt < 0 || fmt > 1) {
    set_error("cpet_write_rows: bad arguments");
    return CPET_ERR_ARG;
  }
  int nt = n_threads > 0 ? n_threads : (int)std::thread::hardware_concurrency();
  if (nt < 1) nt = 1;
  if (rows < 4096) nt = 1;
  std::vector<std::string> parts(nt);
  std::vector<std::thread> th;
  for (int t = 0; t < nt; ++t) {
    const int64_t r0 = rows * t / nt, r1 = rows * (t + 1) / nt;
    th.emplace_back([&, t, r0, r1] {
      parts[t].reserve((size_t)(r1 - r0) * cols * (fmt == 0 ? 8 : 26));
      format_rows(data, dtype, r0, r1, cols, fmt, parts[t]);
    });
  }
  for (auto& x : th) x.join();
  FILE* f = std::fopen(path, "wb");
  if (!f) {
    set_error("cannot open %s for writing", path);
    return CPET_ERR_ARG;
  }
  bool ok = true;
  if (header) ok = std::fwrite(header, 1, std::strlen(header), f) == std::strlen(header);
  for (auto& p : parts) ok = ok && std::fwrite(p.data(), 1, p.size(), f) == p.size();
  ok = (std::fclose(f) == 0) && ok;
  if (!ok) {
    set_error("write to %s failed", path);
    return CPET_ERR_ARG;
  }
  return CPET_OK;
}

int cpet_read_rows(const char* path, int cols, double* out, int64_t cap, int64_t* n_rows) {
  if (!path || cols <= 0 || cap < 0 || !n_rows) {
    set_error("cpet_read_rows: bad arguments");
    return CPET_ERR_ARG;
  }
  std::string s;
  if (!read_file(path, s)) {
    set_error("cannot read %s", path);
    return CPET_ERR_ARG;
  }
  int64_t n = 0;
  int lineno = 0, rc = CPET_OK;
  for_lines(s, [&](const char* l, size_t len) {
    ++lineno;
    if (rc != CPET_OK || (len > 0 && l[0] == '#')) return;
    if (out) {
      if (n >= cap) {
        set_error("%s has more than %lld rows", path, (long long)cap);
        rc = CPET_ERR_ARG;
        return;
      }
      size_t k = 0;
      for (int c = 0; c < cols; ++c) {
        while (k < len && py_space((unsigned char)l[k])) ++k;
        size_t e = k;
        while (e < len && !py_space((unsigned char)l[e])) ++e;
        if (e == k || !parse_float(Field{l + k, e - k}, &out[n * cols + c])) {
          set_error("%s line %d: expected %d numbers", path, lineno, cols);
          rc = CPET_ERR_ARG;
          return;
        }
        k = e;
      }
    }
    ++n;
  });
  *n_rows = n;
  return rc;
}

}  // extern "C"
