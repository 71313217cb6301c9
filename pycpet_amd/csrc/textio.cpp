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

#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cctype>
#include <charconv>
#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <string_view>
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

// A read-only view of a whole file: mapped when it is a regular file (no
// copy of the text), else read into memory.
struct TextView {
  const char* data = nullptr;
  size_t size = 0;
  void* map = nullptr;
  std::string copy;
  bool open(const char* path) {
    const int fd = ::open(path, O_RDONLY);
    if (fd < 0) return false;
    struct stat st;
    if (::fstat(fd, &st) == 0 && S_ISREG(st.st_mode) && st.st_size > 0) {
      void* m = ::mmap(nullptr, (size_t)st.st_size, PROT_READ, MAP_PRIVATE, fd, 0);
      if (m != MAP_FAILED) {
        ::madvise(m, (size_t)st.st_size, MADV_SEQUENTIAL);
        map = m;
        data = (const char*)m;
        size = (size_t)st.st_size;
        ::close(fd);
        return true;
      }
    }
    ::close(fd);
    if (!read_file(path, copy)) return false;
    data = copy.data();
    size = copy.size();
    return true;
  }
  ~TextView() {
    if (map) ::munmap(map, size);
  }
};

// A text cut into up to 16 pieces at line boundaries (whole lines, in file
// order) for per-line work on threads; small texts stay in one piece.
struct Pieces {
  std::string_view s;
  std::vector<size_t> cut;
  int n = 1;
  explicit Pieces(std::string_view text) : s(text) {
    const int hw = (int)std::thread::hardware_concurrency();
    n = std::max(1, std::min({hw, 16, (int)(s.size() >> 20) + 1}));
    cut.assign(n + 1, s.size());
    cut[0] = 0;
    for (int t = 1; t < n; ++t) {
      const size_t p = std::max(cut[t - 1], s.size() * t / n);
      const size_t j = p < s.size() ? s.find('\n', p) : std::string_view::npos;
      cut[t] = j == std::string_view::npos ? s.size() : j + 1;
    }
  }
  // fn(line, len) for every line of piece t; CRLF reads as LF, as text-mode
  // reading turns CRLF into LF
  template <typename Fn>
  void lines(int t, Fn fn) const {
    size_t i = cut[t];
    while (i < cut[t + 1]) {
      size_t j = s.find('\n', i);
      if (j == std::string_view::npos || j >= cut[t + 1]) j = cut[t + 1];
      size_t e = j;
      if (e > i && s[e - 1] == '\r') --e;
      fn(s.data() + i, e - i);
      i = j + 1;
    }
  }
  // work(t) for every piece, piece 0 on the calling thread
  template <typename Work>
  void run(Work work) const {
    std::vector<std::thread> th;
    for (int t = 1; t < n; ++t) th.emplace_back(work, t);
    work(0);
    for (auto& x : th) x.join();
  }
  // per-piece line and record counts (records: lines where rec(l, len))
  template <typename Rec>
  void count(Rec rec, std::vector<int64_t>& lines_, std::vector<int64_t>& recs) const {
    lines_.assign(n, 0);
    recs.assign(n, 0);
    run([&](int t) {
      int64_t nl = 0, nr = 0;  // counted locally: neighbouring entries share a cache line
      lines(t, [&](const char* l, size_t len) {
        ++nl;
        nr += rec(l, len) ? 1 : 0;
      });
      lines_[t] = nl;
      recs[t] = nr;
    });
  }
};

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

// Python's float() accepts decimal literals with an optional sign,
// inf/infinity and nan.  std::from_chars reads the same decimal forms,
// correctly rounded, without a leading '+' and without hexadecimal; a value
// beyond the double range, which from_chars reports as out of range, goes
// to strtod, which rounds it to +-inf or a (sub)normal or zero as float()
// does.
bool parse_float(Field f, double* v) {
  const char* b = f.p;
  const char* e = f.p + f.n;
  if (b == e) return false;
  if (*b == '+') {
    ++b;
    if (b == e || *b == '+' || *b == '-') return false;
  }
  double x = 0.0;
  const auto r = std::from_chars(b, e, x);
  if (r.ptr != e) return false;
  if (r.ec == std::errc()) {
    *v = x;
    return true;
  }
  if (r.ec != std::errc::result_out_of_range) return false;
  char buf[64];
  if (!terminated(f, buf, sizeof(buf))) return false;
  char* end = nullptr;
  x = std::strtod(buf, &end);  // correctly rounded (glibc)
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
// spells every NaN "nan" (C prints "-nan" for a negative one).  std::to_chars
// with an explicit precision is specified as printf's conversion with that
// precision (correctly rounded, "inf"/"-inf", "-0.000"), at about a fifth of
// snprintf's cost per value.
void format_rows(const void* data, int dtype, int64_t r0, int64_t r1, int cols, int fmt,
                 std::string& out) {
  // "%.3f" of the largest double is 309 integer digits + sign + ".000"
  char buf[352];
  const auto form = fmt == 0 ? std::chars_format::fixed : std::chars_format::scientific;
  const int prec = fmt == 0 ? 3 : 18;
  for (int64_t r = r0; r < r1; ++r) {
    for (int c = 0; c < cols; ++c) {
      const double v = value_at(data, dtype, (size_t)r * cols + c);
      int n;
      if (fmt == 0 && dtype != 1 && fast_fixed3(v, buf, &n)) {
      } else if (std::isnan(v)) {
        std::memcpy(buf, "nan", 3);
        n = 3;
      } else {
        n = (int)(std::to_chars(buf, buf + sizeof(buf), v, form, prec).ptr - buf);
      }
      out.append(buf, (size_t)n);
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
  TextView tv;
  if (!tv.open(path)) {
    set_error("cannot read %s", path);
    return CPET_ERR_ARG;
  }
  const Pieces pc(std::string_view(tv.data, tv.size));
  std::vector<int64_t> lines, recs;
  pc.count(is_record, lines, recs);
  int64_t n = 0;
  for (int64_t r : recs) n += r;
  if (n > INT32_MAX) {
    set_error("%s has more than %d atom records", path, INT32_MAX);
    return CPET_ERR_ARG;
  }
  *n_atoms = (int)n;
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
  TextView tv;
  if (!tv.open(path)) {
    set_error("cannot read %s", path);
    return CPET_ERR_ARG;
  }
  // counts per piece, then every piece parses its records from its first
  // record index on; the earliest failing line is reported
  const Pieces pc(std::string_view(tv.data, tv.size));
  std::vector<int64_t> lines, recs;
  pc.count(is_record, lines, recs);
  int64_t total = 0;
  for (int64_t r : recs) total += r;
  if (total > n_atoms) {
    set_error("%s has more than %d atom records", path, n_atoms);
    return CPET_ERR_ARG;
  }
  if (total != n_atoms) {
    set_error("%s has %lld atom records, expected %d", path, (long long)total, n_atoms);
    return CPET_ERR_ARG;
  }
  std::vector<int64_t> rec0(pc.n, 0), line0(pc.n, 0), bad_line(pc.n, -1);
  std::vector<int> bad_kind(pc.n, 0);  // 1: non-ASCII, 2: invalid literal
  for (int t = 1; t < pc.n; ++t) {
    rec0[t] = rec0[t - 1] + recs[t - 1];
    line0[t] = line0[t - 1] + lines[t - 1];
  }
  pc.run([&](int t) {
    int64_t i = rec0[t], k = 0;
    pc.lines(t, [&](const char* l, size_t n) {
      ++k;
      if (bad_line[t] >= 0 || !is_record(l, n)) return;
      if (!is_ascii(l, n)) {
        bad_line[t] = line0[t] + k;
        bad_kind[t] = 1;
        return;
      }
      double v[4];
      if (!parse_int(field(l, n, 6, 11), &atom_number[i]) ||
          !parse_int(field(l, n, 22, 26), &resnum[i]) ||
          !parse_float(field(l, n, 30, 38), &v[0]) || !parse_float(field(l, n, 38, 46), &v[1]) ||
          !parse_float(field(l, n, 46, 54), &v[2]) ||
          (get_charges && !parse_float(field(l, n, 55, 64), &v[3]))) {
        bad_line[t] = line0[t] + k;
        bad_kind[t] = 2;
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
  });
  for (int t = 0; t < pc.n; ++t) {
    if (bad_line[t] < 0) continue;
    if (bad_kind[t] == 1)
      set_error("non-ASCII bytes in %s line %lld", path, (long long)bad_line[t]);
    else
      set_error("%s line %lld: invalid literal in a fixed-width field", path,
                (long long)bad_line[t]);
    return CPET_ERR_ARG;
  }
  return CPET_OK;
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
  TextView tv;
  if (!tv.open(path)) {
    set_error("cannot read %s", path);
    return CPET_ERR_ARG;
  }
  const Pieces pc(std::string_view(tv.data, tv.size));
  std::vector<int64_t> lines, recs;
  pc.count(is_record, lines, recs);
  int64_t total = 0;
  for (int64_t r : recs) total += r;
  if (total != n_atoms) {
    set_error("%s has %lld atom records, expected %d", path, (long long)total, n_atoms);
    return CPET_ERR_ARG;
  }
  std::vector<int64_t> rec0(pc.n, 0), line0(pc.n, 0), bad_line(pc.n, -1);
  std::vector<const char*> bad_why(pc.n, nullptr);
  for (int t = 1; t < pc.n; ++t) {
    rec0[t] = rec0[t - 1] + recs[t - 1];
    line0[t] = line0[t - 1] + lines[t - 1];
  }
  pc.run([&](int t) {
    int64_t i = rec0[t], k = 0;
    pc.lines(t, [&](const char* l, size_t n) {
      ++k;
      if (bad_line[t] >= 0 || !is_record(l, n)) return;
      auto fail = [&](const char* why) {
        bad_line[t] = line0[t] + k;
        bad_why[t] = why;
      };
      if (!is_ascii(l, n)) return fail("non-ASCII bytes");
      Field tok[8];
      int nt = 0;
      for (size_t a = 0; a < n && nt < 8;) {  // the first tokens of line.split()
        while (a < n && py_space((unsigned char)l[a])) ++a;
        size_t e = a;
        while (e < n && !py_space((unsigned char)l[e])) ++e;
        if (e > a) tok[nt++] = Field{l + a, e - a};
        a = e;
      }
      const bool merged = nt > 0 && tok[0].n > 6;
      const int res_ind = merged ? 4 : 5;
      if (nt <= res_ind) return fail("too few fields");
      const Field serial = merged ? Field{tok[0].p + 6, tok[0].n - 6} : tok[1];
      const Field aname = merged ? tok[1] : tok[2];
      const Field rname = merged ? tok[2] : tok[3];
      const Field rtok = tok[res_ind].n > 4
                             ? Field{tok[res_ind - 1].p + 1, tok[res_ind - 1].n - 1}
                             : tok[res_ind];
      int32_t rv = 0;
      if (!parse_int(serial, &atom_number[i]) || !parse_int(rtok, &rv))
        return fail("invalid serial or residue number");
      if (aname.n >= (size_t)width || rname.n >= (size_t)width) return fail("token too long");
      int shift = 0;
      if (rv > 999) {
        int digits = 0;
        for (int64_t d = rv; d > 0; d /= 10) ++digits;
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
  });
  for (int t = 0; t < pc.n; ++t)
    if (bad_line[t] >= 0) {
      set_error("%s line %lld: %s", path, (long long)bad_line[t], bad_why[t]);
      return CPET_ERR_ARG;
    }
  return CPET_OK;
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
  TextView tv;
  if (!tv.open(path)) {
    set_error("cannot read %s", path);
    return CPET_ERR_ARG;
  }
  // rows are the non-comment lines; pieces count them, then parse from their
  // first row on, and the earliest failing line is reported
  const Pieces pc(std::string_view(tv.data, tv.size));
  std::vector<int64_t> lines, rows;
  pc.count([](const char* l, size_t len) { return !(len > 0 && l[0] == '#'); }, lines, rows);
  int64_t total = 0;
  for (int64_t r : rows) total += r;
  *n_rows = total;
  if (!out) return CPET_OK;
  if (total > cap) {
    set_error("%s has more than %lld rows", path, (long long)cap);
    return CPET_ERR_ARG;
  }
  std::vector<int64_t> row0(pc.n, 0), line0(pc.n, 0), bad(pc.n, -1);
  for (int t = 1; t < pc.n; ++t) {
    row0[t] = row0[t - 1] + rows[t - 1];
    line0[t] = line0[t - 1] + lines[t - 1];
  }
  pc.run([&](int t) {
    int64_t r = row0[t], k = 0;
    pc.lines(t, [&](const char* l, size_t len) {
      ++k;
      if (bad[t] >= 0 || (len > 0 && l[0] == '#')) return;
      size_t a = 0;
      for (int c = 0; c < cols; ++c) {
        while (a < len && py_space((unsigned char)l[a])) ++a;
        size_t e = a;
        while (e < len && !py_space((unsigned char)l[e])) ++e;
        if (e == a || !parse_float(Field{l + a, e - a}, &out[r * cols + c])) {
          bad[t] = line0[t] + k;
          return;
        }
        a = e;
      }
      ++r;
    });
  });
  for (int t = 0; t < pc.n; ++t)
    if (bad[t] >= 0) {
      set_error("%s line %lld: expected %d numbers", path, (long long)bad[t], cols);
      return CPET_ERR_ARG;
    }
  return CPET_OK;
}

}  // extern "C"
