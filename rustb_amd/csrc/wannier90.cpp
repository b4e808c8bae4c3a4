// Wannier90 input files to model storage: Wannier90::from_hr
// (wannier90.rs:49-1228) for Model<SPIN, 3, NoRMatrix>.
//
// Files, from the prefix {path}{seedname}:
//   _hr.dat        (required) nsta, n_R, Wigner-Seitz degeneracies, then
//                  n_R * nsta^2 lines "R1 R2 R3 m n Re Im", n outer, m inner
//                  (the element is <m,0|H|n,R>, stored as ham[R][m][n] after
//                  division by the degeneracy; wannier90.rs:129-235);
//   .win           (required) unit_cell_cart, spinors, projections,
//                  atoms_cart / atoms_frac (wannier90.rs:252-559);
//   _centres.xyz   (optional) Wannier centres (orbital positions) and atoms
//                  (wannier90.rs:565-711); without it the orbitals sit on the
//                  .win atoms (wannier90.rs:712-782);
//   _wsvec.dat     (optional) Wigner-Seitz shifts that redistribute each
//                  hopping over equivalent R (wannier90.rs:904-1133).
// R = 0 becomes storage row 0 (the last R = 0 block of the file wins, as the
// reference overwrites it) with the Fermi level subtracted from its diagonal;
// the other R follow in file order.
//
// The _hr.dat body is the bulk of the input (n_R n^2 lines, ~1 GB for
// n = 128, n_R = 1331).  It is memory-mapped and parsed by host threads over
// equal byte ranges cut at line ends; a newline count per range gives each
// range's first line number, so every thread writes its elements straight
// into their storage slots.  Errors are those of the reference, reported for
// the first offending line in its order (lowest line; the R fields of a block
// before its matrix fields).  Numbers follow Rust's grammar for isize/usize/
// f64 and are rounded correctly (exact decimal fast path, strtod otherwise).
// Where the reference panics (a file shorter than its header says, an index
// beyond nsta), the loader returns RTB_ERR_FILE_PARSE instead.
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "rtb_internal.hpp"

namespace {

constexpr double kBohrToAngstrom = 0.52917721067;  // wannier90.rs:9

// Atomic species accepted by AtomType::from_str (atom_struct.rs:485-580):
// the elements H (Z = 1) to Ra (Z = 88), case-sensitive.
const char* const kElements[] = {
    "H",  "He", "Li", "Be", "B",  "C",  "N",  "O",  "F",  "Ne", "Na", "Mg", "Al", "Si", "P",  "S",  "Cl", "Ar",
    "K",  "Ca", "Sc", "Ti", "V",  "Cr", "Mn", "Fe", "Co", "Ni", "Cu", "Zn", "Ga", "Ge", "As", "Se", "Br", "Kr",
    "Rb", "Sr", "Y",  "Zr", "Nb", "Mo", "Tc", "Ru", "Rh", "Pd", "Ag", "Cd", "In", "Sn", "Sb", "Te", "I",  "Xe",
    "Cs", "Ba", "La", "Ce", "Pr", "Nd", "Pm", "Sm", "Eu", "Gd", "Tb", "Dy", "Ho", "Er", "Tm", "Yb", "Lu", "Hf",
    "Ta", "W",  "Re", "Os", "Ir", "Pt", "Au", "Hg", "Tl", "Pb", "Bi", "Po", "At", "Rn", "Fr", "Ra"};
constexpr int kNumElements = sizeof(kElements) / sizeof(kElements[0]);

int element_number(const std::string& s) {
  for (int z = 0; z < kNumElements; ++z)
    if (s == kElements[z]) return z + 1;
  return 0;
}

// OrbProj codes in the declaration order of the enum (atom_struct.rs:15-77).
enum OrbCode {
  kS, kPx, kPy, kPz, kDxy, kDyz, kDxz, kDz2, kDx2y2, kFz3, kFxz2, kFyz2, kFzx2y2, kFxyz, kFxx23y2, kFy3x2y2,
  kSp1, kSp2, kSp2_1, kSp2_2, kSp2_3, kSp3_1, kSp3_2, kSp3_3, kSp3_4, kSp3d_1, kSp3d_2, kSp3d_3, kSp3d_4,
  kSp3d_5, kSp3d2_1, kSp3d2_2, kSp3d2_3, kSp3d2_4, kSp3d2_5, kSp3d2_6, kNumOrbCodes
};
const char* const kOrbNames[kNumOrbCodes] = {
    "s",     "px",    "py",    "pz",    "dxy",    "dyz",    "dxz",    "dz2",    "dx2y2",  "fz3",
    "fxz2",  "fyz2",  "fzx2y2", "fxyz", "fxx23y2", "fy3x2y2", "sp_1",  "sp_2",   "sp2_1",  "sp2_2",
    "sp2_3", "sp3_1", "sp3_2", "sp3_3", "sp3_4",  "sp3d_1", "sp3d_2", "sp3d_3", "sp3d_4", "sp3d_5",
    "sp3d2_1", "sp3d2_2", "sp3d2_3", "sp3d2_4", "sp3d2_5", "sp3d2_6"};

// A projection item of the .win projections block and the orbitals it
// expands to, in the order the reference lists them (wannier90.rs:355-420).
bool projection_orbitals(const std::string& item, std::vector<int>& out) {
  struct Entry {
    const char* name;
    std::vector<int> orbs;
  };
  static const Entry table[] = {
      {"s", {kS}},
      {"p", {kPz, kPx, kPy}},
      {"d", {kDz2, kDxz, kDyz, kDx2y2, kDxy}},
      {"f", {kFz3, kFxz2, kFyz2, kFzx2y2, kFxyz, kFxx23y2, kFy3x2y2}},
      {"sp3", {kSp3_1, kSp3_2, kSp3_3, kSp3_4}},
      {"sp2", {kSp2_1, kSp2_2, kSp2_3}},
      {"sp", {kSp1, kSp2}},
      {"sp3d", {kSp3d_1, kSp3d_2, kSp3d_3, kSp3d_4, kSp3d_5}},
      {"sp3d2", {kSp3d2_1, kSp3d2_2, kSp3d2_3, kSp3d2_4, kSp3d2_5, kSp3d2_6}},
      {"px", {kPx}},
      {"py", {kPy}},
      {"pz", {kPz}},
      {"dxy", {kDxy}},
      {"dxz", {kDxz}},
      {"dyz", {kDyz}},
      {"dz2", {kDz2}},
      {"dx2-y2", {kDx2y2}},
  };
  for (const auto& e : table)
    if (item == e.name) {
      out.insert(out.end(), e.orbs.begin(), e.orbs.end());
      return true;
    }
  return false;
}

// ---- errors ----
struct Error {
  int code = RTB_OK;
  std::string msg;
};

int file_parse(Error& err, const std::string& file, const std::string& message) {
  err.code = RTB_ERR_FILE_PARSE;
  err.msg = "Failed to parse data from file '" + file + "': " + message;
  return err.code;
}

std::string os_error_text(int e) { return std::string(std::strerror(e)) + " (os error " + std::to_string(e) + ")"; }

int out_of_memory(Error& err) {
  err.code = RTB_ERR_UNSUPPORTED;
  err.msg = "out of host memory for the Hamiltonian";
  return err.code;
}

// ---- text primitives (Rust str semantics, ASCII whitespace) ----
inline bool is_ws(char c) { return c == ' ' || (c >= '\t' && c <= '\r'); }

struct Tok {
  const char* p;
  const char* e;
  std::string str() const { return std::string(p, e); }
  size_t size() const { return static_cast<size_t>(e - p); }
};

// split_whitespace().next() over [p, e); false at the end of the line.
inline bool next_token(const char*& p, const char* e, Tok& t) {
  while (p < e && is_ws(*p)) ++p;
  if (p == e) return false;
  t.p = p;
  while (p < e && !is_ws(*p)) ++p;
  t.e = p;
  return true;
}

Tok trim(const char* p, const char* e) {
  while (p < e && is_ws(*p)) ++p;
  while (e > p && is_ws(e[-1])) --e;
  return {p, e};
}

// <usize as FromStr> / <isize as FromStr>: an optional sign ('-' only for
// signed), decimal digits, no overflow.  Returns "" or the ParseIntError text.
const char* parse_int(Tok t, bool is_signed, int64_t& out_s, uint64_t& out_u) {
  const char* p = t.p;
  if (p == t.e) return "cannot parse integer from empty string";
  bool neg = false;
  if (*p == '+' || *p == '-') {
    if (*p == '-') {
      if (!is_signed) return "invalid digit found in string";
      neg = true;
    }
    ++p;
    if (p == t.e) return "invalid digit found in string";
  }
  uint64_t v = 0;
  const uint64_t limit = is_signed ? (neg ? (1ull << 63) : (1ull << 63) - 1) : ~0ull;
  for (; p < t.e; ++p) {
    if (*p < '0' || *p > '9') return "invalid digit found in string";
    const uint64_t d = static_cast<uint64_t>(*p - '0');
    if (v > (limit - d) / 10) return neg ? "number too small to fit in target type" : "number too large to fit in target type";
    v = v * 10 + d;
  }
  out_u = v;
  out_s = neg ? static_cast<int64_t>(0 - v) : static_cast<int64_t>(v);
  return "";
}

inline bool ieq(const char* p, const char* e, const char* lit) {
  const size_t n = std::strlen(lit);
  if (static_cast<size_t>(e - p) != n) return false;
  for (size_t i = 0; i < n; ++i)
    if ((p[i] | 0x20) != lit[i]) return false;
  return true;
}

// <f64 as FromStr>: [+-]? ( inf | infinity | nan  (any case)
//                        | digits [. digits*] | . digits ) ( [eE] [+-]? digits )?
// Correctly rounded: at most 19 digits form an exact integer m; with
// m <= 2^53 and a power 10^q, |q| <= 22, the value is m * 10^q or m / 10^-q in
// one IEEE operation on exact operands (Clinger's fast path); anything else
// goes through strtod, which rounds correctly.  Returns false for a string
// Rust rejects ("invalid float literal").
bool parse_f64(Tok t, double& out) {
  static const double kPow10[23] = {1e0,  1e1,  1e2,  1e3,  1e4,  1e5,  1e6,  1e7,  1e8,  1e9,  1e10, 1e11,
                                    1e12, 1e13, 1e14, 1e15, 1e16, 1e17, 1e18, 1e19, 1e20, 1e21, 1e22};
  const char* p = t.p;
  const char* e = t.e;
  if (p == e) return false;
  bool neg = false;
  if (*p == '+' || *p == '-') {
    neg = *p == '-';
    ++p;
  }
  if (p < e && ((*p | 0x20) == 'i' || (*p | 0x20) == 'n')) {
    if (ieq(p, e, "inf") || ieq(p, e, "infinity")) {
      out = neg ? -HUGE_VAL : HUGE_VAL;
      return true;
    }
    if (ieq(p, e, "nan")) {
      out = neg ? -NAN : NAN;
      return true;
    }
    return false;
  }
  uint64_t m = 0;  // wraps beyond 19 digits; only used when it did not
  const char* d0 = p;
  while (p < e && static_cast<unsigned>(*p - '0') < 10u) m = m * 10 + static_cast<unsigned>(*p++ - '0');
  const size_t ni = static_cast<size_t>(p - d0);
  size_t nf = 0;
  if (p < e && *p == '.') {
    const char* f0 = ++p;
    while (p < e && static_cast<unsigned>(*p - '0') < 10u) m = m * 10 + static_cast<unsigned>(*p++ - '0');
    nf = static_cast<size_t>(p - f0);
  }
  if (ni + nf == 0) return false;
  int64_t q = -static_cast<int64_t>(nf);
  if (p < e) {
    if ((*p | 0x20) != 'e') return false;
    ++p;
    bool eneg = false;
    if (p < e && (*p == '+' || *p == '-')) {
      eneg = *p == '-';
      ++p;
    }
    if (p == e) return false;
    int64_t x = 0;
    for (; p < e; ++p) {
      if (static_cast<unsigned>(*p - '0') >= 10u) return false;
      if (x < 100000000) x = x * 10 + (*p - '0');
    }
    q += eneg ? -x : x;
  }
  if (ni + nf <= 19) {
    if (m == 0) {
      out = neg ? -0.0 : 0.0;
      return true;
    }
    if (m <= (1ull << 53) && q >= -22 && q <= 22) {
      const double v = static_cast<double>(m);
      out = q >= 0 ? v * kPow10[q] : v / kPow10[-q];
      if (neg) out = -out;
      return true;
    }
  }
  std::string s(t.p, t.e);
  out = std::strtod(s.c_str(), nullptr);
  return true;
}

// Standard UTF-8 validation (what Rust's lines() requires of the text).
bool valid_utf8(const unsigned char* p, const unsigned char* e) {
  while (p < e) {
    const unsigned char c = *p;
    if (c < 0x80) {
      ++p;
      continue;
    }
    int n;
    uint32_t cp;
    if ((c & 0xE0) == 0xC0) {
      n = 1;
      cp = c & 0x1F;
    } else if ((c & 0xF0) == 0xE0) {
      n = 2;
      cp = c & 0x0F;
    } else if ((c & 0xF8) == 0xF0) {
      n = 3;
      cp = c & 0x07;
    } else {
      return false;
    }
    if (e - p <= n) return false;
    for (int i = 1; i <= n; ++i) {
      if ((p[i] & 0xC0) != 0x80) return false;
      cp = (cp << 6) | (p[i] & 0x3F);
    }
    if ((n == 1 && cp < 0x80) || (n == 2 && cp < 0x800) || (n == 3 && cp < 0x10000) || cp > 0x10FFFF ||
        (cp >= 0xD800 && cp <= 0xDFFF))
      return false;
    p += n + 1;
  }
  return true;
}

// ---- whole small files as lines ----
int read_lines(const std::string& path, std::vector<std::string>& lines, Error& err, bool& opened) {
  opened = false;
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) {
    err.code = RTB_ERR_FILE_OPEN;
    err.msg = os_error_text(errno);
    return err.code;
  }
  opened = true;
  std::string data;
  char buf[1 << 16];
  size_t got;
  while ((got = std::fread(buf, 1, sizeof buf, f)) > 0) data.append(buf, got);
  std::fclose(f);
  if (!valid_utf8(reinterpret_cast<const unsigned char*>(data.data()),
                  reinterpret_cast<const unsigned char*>(data.data()) + data.size())) {
    err.code = RTB_ERR_IO;
    err.msg = "I/O error: stream did not contain valid UTF-8";
    return err.code;
  }
  size_t b = 0;
  while (b < data.size()) {
    size_t nl = data.find('\n', b);
    if (nl == std::string::npos) {
      lines.emplace_back(data, b);
      break;
    }
    size_t end = nl;
    if (end > b && data[end - 1] == '\r') --end;
    lines.emplace_back(data, b, end - b);
    b = nl + 1;
  }
  return RTB_OK;
}

std::vector<Tok> tokens_of(const std::string& s) {
  std::vector<Tok> out;
  const char* p = s.data();
  const char* e = p + s.size();
  Tok t;
  while (next_token(p, e, t)) out.push_back(t);
  return out;
}

std::string lower(std::string s) {
  for (auto& c : s)
    if (c >= 'A' && c <= 'Z') c = static_cast<char>(c - 'A' + 'a');
  return s;
}

// win_data_line (wannier90.rs:11-17): text before the first '!', then before
// the first '#', trimmed.
std::string win_data_line(const std::string& line) {
  size_t cut = line.find('!');
  std::string s = cut == std::string::npos ? line : line.substr(0, cut);
  cut = s.find('#');
  if (cut != std::string::npos) s.resize(cut);
  Tok t = trim(s.data(), s.data() + s.size());
  return t.str();
}

bool length_unit_scale(const std::string& token, double& scale) {
  const std::string l = lower(token);
  if (l == "ang" || l == "angstrom" || l == "angstroms") {
    scale = 1.0;
    return true;
  }
  if (l == "bohr") {
    scale = kBohrToAngstrom;
    return true;
  }
  return false;
}

// 3 x 3 inverse by LU with partial pivoting (an exact zero pivot is the
// singular case the reference's LAPACK inverse reports).
bool inverse3(const double* a, double* inv) {
  double lu[9];
  std::memcpy(lu, a, sizeof lu);
  int piv[3] = {0, 1, 2};
  for (int k = 0; k < 3; ++k) {
    int p = k;
    for (int i = k + 1; i < 3; ++i)
      if (std::fabs(lu[i * 3 + k]) > std::fabs(lu[p * 3 + k])) p = i;
    if (lu[p * 3 + k] == 0.0) return false;
    if (p != k) {
      for (int j = 0; j < 3; ++j) std::swap(lu[k * 3 + j], lu[p * 3 + j]);
      std::swap(piv[k], piv[p]);
    }
    for (int i = k + 1; i < 3; ++i) {
      lu[i * 3 + k] /= lu[k * 3 + k];
      for (int j = k + 1; j < 3; ++j) lu[i * 3 + j] -= lu[i * 3 + k] * lu[k * 3 + j];
    }
  }
  for (int c = 0; c < 3; ++c) {
    double x[3];
    for (int i = 0; i < 3; ++i) {
      double s = piv[i] == c ? 1.0 : 0.0;
      for (int j = 0; j < i; ++j) s -= lu[i * 3 + j] * x[j];
      x[i] = s;
    }
    for (int i = 2; i >= 0; --i) {
      double s = x[i];
      for (int j = i + 1; j < 3; ++j) s -= lu[i * 3 + j] * x[j];
      x[i] = s / lu[i * 3 + i];
    }
    for (int i = 0; i < 3; ++i) inv[i * 3 + c] = x[i];
  }
  return true;
}

// row vector times 3 x 3 matrix
void row_times(const double* v, const double* m, double* out) {
  for (int c = 0; c < 3; ++c) out[c] = v[0] * m[c] + v[1] * m[3 + c] + v[2] * m[6 + c];
}

struct RKey {
  int64_t r[3];
  bool operator==(const RKey& o) const { return r[0] == o.r[0] && r[1] == o.r[1] && r[2] == o.r[2]; }
};
struct RHash {
  size_t operator()(const RKey& k) const {
    uint64_t h = 0x9e3779b97f4a7c15ull;
    for (int d = 0; d < 3; ++d) {
      h ^= static_cast<uint64_t>(k.r[d]) * 0xbf58476d1ce4e5b9ull;
      h = (h << 27) | (h >> 37);
    }
    return static_cast<size_t>(h);
  }
};

// ---- memory-mapped file ----
struct Mapped {
  const char* data = nullptr;
  size_t size = 0;
  void* map = nullptr;
  ~Mapped() {
    if (map) munmap(map, size);
  }
};

int map_file(const std::string& path, Mapped& m, Error& err) {
  int fd = ::open(path.c_str(), O_RDONLY);
  if (fd < 0) {
    err.code = RTB_ERR_FILE_OPEN;
    err.msg = os_error_text(errno);
    return err.code;
  }
  struct stat st;
  if (fstat(fd, &st) != 0) {
    const int e = errno;
    ::close(fd);
    err.code = RTB_ERR_IO;
    err.msg = "I/O error: " + os_error_text(e);
    return err.code;
  }
  m.size = static_cast<size_t>(st.st_size);
  if (m.size > 0) {
    void* p = mmap(nullptr, m.size, PROT_READ, MAP_PRIVATE, fd, 0);
    if (p == MAP_FAILED) {
      const int e = errno;
      ::close(fd);
      err.code = RTB_ERR_IO;
      err.msg = "I/O error: " + os_error_text(e);
      return err.code;
    }
    m.map = p;
    m.data = static_cast<const char*>(p);
  }
  ::close(fd);
  return RTB_OK;
}

}  // namespace

struct rtb_w90 {
  int spin = 0, norb = 0, nsta = 0, nR = 0;
  double lat[9] = {};
  std::vector<double> orb;      // [norb][3] fractional
  std::vector<int> orb_proj;    // [norb] OrbProj codes
  struct Atom {
    int species;
    double pos[3];
    int first, count;
  };
  std::vector<Atom> atoms;
  std::vector<int64_t> hamR;    // [nR][3], row 0 = R = 0
  double* ham = nullptr;        // [nR][nsta][nsta] complex
  std::vector<uint64_t> weights;  // the _hr.dat degeneracies, file order
  double* rmat = nullptr;       // [nR][3][nsta][nsta] complex (with _r.dat)
  ~rtb_w90() {
    std::free(ham);
    std::free(rmat);
  }
};

namespace {

// One parse thread's range of a body and its first error (key: the
// reference's order, line-major, then field group).
struct RangeErr {
  uint64_t key = ~0ull;
  std::string msg;
};

// The body of a line-oriented file split for host threads: T byte ranges
// [cut[t], cut[t+1]) that start at line starts, the global line number of
// each range's first line (a newline count per range; n_line is the number
// of the line at b0), and whether any range is not UTF-8.
struct LineRanges {
  int T = 1;
  std::vector<size_t> cut;
  std::vector<uint64_t> first_line;
  uint64_t total_lines = 0;
  bool bad_utf8 = false;
};

template <typename F>
void run_ranges(int T, F&& fn) {
  std::vector<std::thread> th;
  for (int t = 1; t < T; ++t) th.emplace_back(fn, t);
  fn(0);
  for (auto& x : th) x.join();
}

LineRanges split_lines(const char* buf, size_t S, size_t b0, uint64_t n_line, int n_threads) {
  LineRanges lr;
  int T = n_threads > 0 ? n_threads : std::min(16, static_cast<int>(std::thread::hardware_concurrency()));
  T = std::max(1, std::min(T, 64));
  const size_t body = S - b0;
  if (body < static_cast<size_t>(T) * 4096) T = 1;
  lr.T = T;
  lr.cut.assign(T + 1, 0);
  lr.cut[0] = b0;
  lr.cut[T] = S;
  for (int t = 1; t < T; ++t) {
    size_t c = b0 + body / T * t;
    const char* nl = c < S ? static_cast<const char*>(std::memchr(buf + c, '\n', S - c)) : nullptr;
    c = nl ? static_cast<size_t>(nl - buf) + 1 : S;
    lr.cut[t] = std::max(c, lr.cut[t - 1]);
  }
  std::vector<uint64_t> nlines(T, 0);
  std::vector<char> bad(T, 0);
  run_ranges(T, [&](int t) {
    const char* p = buf + lr.cut[t];
    const char* e = buf + lr.cut[t + 1];
    // branch-free so it vectorises: newlines and the OR of all bytes
    uint64_t c = 0;
    unsigned char hi = 0;
    const size_t len = static_cast<size_t>(e - p);
    const unsigned char* uq = reinterpret_cast<const unsigned char*>(p);
    for (size_t j = 0; j < len; ++j) {
      c += uq[j] == '\n';
      hi |= uq[j];
    }
    if (t == T - 1 && e > p && e[-1] != '\n') ++c;  // last line without '\n'
    nlines[t] = c;
    if ((hi & 0x80) &&
        !valid_utf8(reinterpret_cast<const unsigned char*>(p), reinterpret_cast<const unsigned char*>(e)))
      bad[t] = 1;
  });
  lr.first_line.assign(T + 1, 0);
  lr.first_line[0] = n_line;
  for (int t = 0; t < T; ++t) {
    lr.first_line[t + 1] = lr.first_line[t] + nlines[t];
    lr.bad_utf8 = lr.bad_utf8 || bad[t];
  }
  lr.total_lines = lr.first_line[T];
  return lr;
}

int parse_hr(const std::string& file, int n_threads, double zero_energy, rtb_w90& out, Error& err) {
  Mapped mf;
  {
    Error e;
    if (map_file(file, mf, e) != RTB_OK) {
      if (e.code == RTB_ERR_FILE_OPEN) {
        err.code = RTB_ERR_FILE_OPEN;
        err.msg = "Failed to create file '" + file + "': Unable to open HR file: " + e.msg;
      } else {
        err = e;
      }
      return err.code;
    }
  }
  const char* buf = mf.data;
  const size_t S = mf.size;

  // header lines, serially: 0 comment, 1 nsta, 2 n_R, 3.. degeneracies up to
  // the first line holding a '.' (wannier90.rs:91-125)
  struct Line {
    const char* p;
    const char* e;
  };
  size_t pos = 0;
  auto next_line = [&](Line& l) -> bool {
    if (pos >= S) return false;
    const char* nl = static_cast<const char*>(std::memchr(buf + pos, '\n', S - pos));
    const size_t end = nl ? static_cast<size_t>(nl - buf) : S;
    l.p = buf + pos;
    l.e = buf + end;
    if (nl && l.e > l.p && l.e[-1] == '\r') --l.e;
    pos = nl ? end + 1 : S;
    return true;
  };
  Line l0, l1, l2;
  if (!next_line(l0) || !next_line(l1) || !next_line(l2))
    return file_parse(err, file, "Truncated _hr.dat: missing the orbital or R-vector count line");
  // every line must be UTF-8 before anything is parsed (the reference reads
  // the whole file through lines() first), so header errors are reported
  // after the UTF-8 checks, in the reference's order: nsta, n_R, degeneracies
  uint64_t u = 0;
  int64_t s;
  std::string header_err;
  const char* pe = parse_int(trim(l1.p, l1.e), false, s, u);
  if (*pe) header_err = std::string("Failed to parse nsta: ") + pe;
  const uint64_t nsta = *pe ? 0 : u;
  pe = parse_int(trim(l2.p, l2.e), false, s, u);
  if (*pe && header_err.empty()) header_err = std::string("Failed to parse n_R: ") + pe;
  const uint64_t nR = *pe ? 0 : u;

  // line numbers and byte offsets: body from the first line with a '.'
  std::vector<uint64_t> weights;
  size_t n_line = 0, body_off = 0;
  size_t line_no = 3;
  std::string weight_err;
  {
    Line l;
    size_t at = pos;
    while (next_line(l)) {
      if (std::memchr(l.p, '.', static_cast<size_t>(l.e - l.p))) {
        n_line = line_no;
        body_off = at;
        break;
      }
      const char* p = l.p;
      Tok t;
      while (weight_err.empty() && next_token(p, l.e, t)) {
        const char* we = parse_int(t, false, s, u);
        if (*we)
          weight_err = std::string("Failed to parse weight: ") + we;
        else
          weights.push_back(u);
      }
      ++line_no;
      at = pos;
    }
    if (n_line == 0) body_off = 0;  // no '.' anywhere: the reference indexes from line 0
  }
  // UTF-8 of everything before the body (the body ranges check their own)
  const size_t scan_from = n_line ? body_off : S;
  if (!valid_utf8(reinterpret_cast<const unsigned char*>(buf),
                  reinterpret_cast<const unsigned char*>(buf) + std::min(scan_from, S))) {
    err.code = RTB_ERR_IO;
    err.msg = "I/O error: stream did not contain valid UTF-8";
    return err.code;
  }

  const size_t b0 = n_line ? body_off : 0;
  const LineRanges lr = split_lines(buf, S, b0, n_line, n_threads);
  if (lr.bad_utf8) {
    err.code = RTB_ERR_IO;
    err.msg = "I/O error: stream did not contain valid UTF-8";
    return err.code;
  }
  const int T = lr.T;
  const std::vector<size_t>& cut = lr.cut;
  const std::vector<uint64_t>& first_line = lr.first_line;
  if (!header_err.empty()) return file_parse(err, file, header_err);
  if (!weight_err.empty()) return file_parse(err, file, weight_err);
  const uint64_t total_lines = lr.total_lines;

  const uint64_t n2 = nsta * nsta;
  if (nsta > 0 && (nsta > (1u << 20) || nR > (1ull << 40) / std::max<uint64_t>(n2, 1)))
    return file_parse(err, file, "Truncated _hr.dat: the header sizes exceed the file");
  const uint64_t need = nR * n2;  // body lines read by the reference
  const uint64_t last_needed = nsta == 0 ? (nR ? n_line + 1 : 0) : n_line + need;
  if (last_needed > total_lines)
    return file_parse(err, file, "Truncated _hr.dat: header declares " + std::to_string(nR) + " R vectors of " +
                                     std::to_string(nsta) + "^2 elements, but the file has " +
                                     std::to_string(total_lines) + " lines");
  if (nsta > 0 && weights.size() < nR)
    return file_parse(err, file, "Truncated _hr.dat: " + std::to_string(weights.size()) +
                                     " degeneracy weights for " + std::to_string(nR) + " R vectors");

  // storage: slot 0 is R = 0; speculate that the single R = 0 block sits in
  // the middle of the list (Wannier90's R order), fix up otherwise
  const size_t n = static_cast<size_t>(nsta);
  std::vector<int64_t> R(nR * 3, 0);
  const size_t zguess = nR ? (nR - 1) / 2 : 0;
  auto slot_guess = [&](size_t ir) -> size_t { return ir == zguess ? 0 : (ir < zguess ? ir + 1 : ir); };
  const size_t bytes = std::max<size_t>(nR, 1) * n * n * 2 * sizeof(double);
  double* ham = static_cast<double*>(std::malloc(std::max<size_t>(bytes, 16)));
  if (!ham) return out_of_memory(err);

  std::vector<RangeErr> rerr(T);
  auto parse_range = [&](int t) {
    const char* p = buf + cut[t];
    const char* e = buf + cut[t + 1];
    uint64_t L = first_line[t];
    const uint64_t stop = n_line + need;
    RangeErr& re = rerr[t];
    auto set = [&](uint64_t key, std::string m) {
      if (key < re.key) {
        re.key = key;
        re.msg = std::move(m);
      }
    };
    while (p < e && L < stop) {
      const char* nl = static_cast<const char*>(std::memchr(p, '\n', static_cast<size_t>(e - p)));
      const char* le = nl ? nl : e;
      const uint64_t rel = L - n_line;
      const uint64_t ir = rel / n2, w = rel % n2;
      const uint64_t ii = w / n, jj = w % n;  // ind_i (outer), ind_j (inner)
      const char* q = p;
      Tok tk;
      if (w == 0) {
        for (int d = 0; d < 3; ++d) {
          if (!next_token(q, le, tk)) {
            set(2 * L, "Missing R vector component");
            break;
          }
          int64_t sv;
          uint64_t uv;
          const char* me = parse_int(tk, true, sv, uv);
          if (*me) {
            set(2 * L, std::string("Failed to parse R vector: ") + me);
            break;
          }
          R[ir * 3 + d] = sv;
        }
        q = p;
      }
      int k = 0;
      while (k < 5 && next_token(q, le, tk)) ++k;
      double re_v = 0.0, im_v = 0.0;
      if (k < 5 || !next_token(q, le, tk)) {
        set(2 * L + 1, "Missing Hamiltonian real part");
      } else if (!parse_f64(tk, re_v)) {
        set(2 * L + 1, "Failed to parse Hamiltonian real part: invalid float literal");
      } else if (!next_token(q, le, tk)) {
        set(2 * L + 1, "Missing Hamiltonian imaginary part");
      } else if (!parse_f64(tk, im_v)) {
        set(2 * L + 1, "Failed to parse Hamiltonian imaginary part: invalid float literal");
      }
      const double wd = static_cast<double>(weights[ir]);
      double* dst = ham + ((slot_guess(ir) * n + jj) * n + ii) * 2;
      dst[0] = re_v / wd;
      dst[1] = im_v / wd;
      p = nl ? nl + 1 : e;
      ++L;
    }
  };
  if (nsta == 0) {
    // no matrix lines: every block reads its R from line n_line
    Line l;
    pos = b0;
    if (nR && next_line(l)) {
      const char* q = l.p;
      Tok tk;
      for (int d = 0; d < 3; ++d) {
        if (!next_token(q, l.e, tk)) {
          std::free(ham);
          return file_parse(err, file, "Missing R vector component");
        }
        const char* me = parse_int(tk, true, s, u);
        if (*me) {
          std::free(ham);
          return file_parse(err, file, std::string("Failed to parse R vector: ") + me);
        }
        for (size_t ir = 0; ir < nR; ++ir) R[ir * 3 + d] = s;
      }
    }
  } else {
    run_ranges(T, parse_range);
    const RangeErr* first = nullptr;
    for (const auto& r : rerr)
      if (r.key != ~0ull && (!first || r.key < first->key)) first = &r;
    if (first) {
      std::free(ham);
      return file_parse(err, file, first->msg);
    }
  }

  // final order: [last R = 0 block or zeros] + nonzero-R blocks in file order
  size_t z = SIZE_MAX, nz0 = 0;
  for (size_t ir = 0; ir < nR; ++ir)
    if (R[ir * 3] == 0 && R[ir * 3 + 1] == 0 && R[ir * 3 + 2] == 0) {
      z = ir;
      ++nz0;
    }
  const size_t nfinal = 1 + nR - nz0;
  const size_t blk = n * n * 2;
  if (!(nz0 == 1 && z == zguess)) {
    double* fixed = static_cast<double*>(std::malloc(std::max<size_t>(nfinal * blk * sizeof(double), 16)));
    if (!fixed) {
      std::free(ham);
      return out_of_memory(err);
    }
    if (z == SIZE_MAX)
      std::memset(fixed, 0, blk * sizeof(double));
    else if (nsta)
      std::memcpy(fixed, ham + slot_guess(z) * blk, blk * sizeof(double));
    size_t sl = 1;
    for (size_t ir = 0; ir < nR; ++ir) {
      if (R[ir * 3] == 0 && R[ir * 3 + 1] == 0 && R[ir * 3 + 2] == 0) continue;
      if (nsta) std::memcpy(fixed + sl * blk, ham + slot_guess(ir) * blk, blk * sizeof(double));
      ++sl;
    }
    std::free(ham);
    ham = fixed;
  }
  out.hamR.assign(nfinal * 3, 0);
  {
    size_t sl = 1;
    for (size_t ir = 0; ir < nR; ++ir) {
      if (R[ir * 3] == 0 && R[ir * 3 + 1] == 0 && R[ir * 3 + 2] == 0) continue;
      for (int d = 0; d < 3; ++d) out.hamR[sl * 3 + d] = R[ir * 3 + d];
      ++sl;
    }
  }
  // Fermi level (wannier90.rs:247-249)
  for (size_t i = 0; i < n; ++i) ham[(i * n + i) * 2] -= zero_energy;
  out.ham = ham;
  out.nsta = static_cast<int>(nsta);
  out.nR = static_cast<int>(nfinal);
  out.weights = std::move(weights);
  return RTB_OK;
}

// seedname_r.dat (wannier90.rs:805-902): header, num_wann, n_R, then n_R
// blocks of nsta^2 lines "R1 R2 R3 m n x_re x_im y_re y_im z_re z_im" (n
// outer, m inner; no degeneracy lines), each block's R looked up in hamR and
// its values divided by the _hr.dat degeneracy of the same block number:
// rmat[index(R)][d][m][n].  Parsed like the _hr.dat body: threads over line
// ranges into a file-order buffer, then each block copied to its R's row.
int parse_rdat(const std::string& file, const std::string& xyz_file, int n_threads, rtb_w90& m, Error& err) {
  Mapped mf;
  {
    Error e;
    if (map_file(file, mf, e) != RTB_OK) {
      if (e.code == RTB_ERR_FILE_OPEN) {
        err.code = RTB_ERR_FILE_OPEN;
        err.msg = "Failed to create file '" + file + "': R::HAS_RMATRIX=true but _r.dat file not found";
      } else {
        err = e;
      }
      return err.code;
    }
  }
  const char* buf = mf.data;
  const size_t S = mf.size;
  const size_t n = static_cast<size_t>(m.nsta), n2 = n * n;
  m.rmat = static_cast<double*>(std::calloc(std::max<size_t>(static_cast<size_t>(m.nR) * 3 * n2 * 2, 2), sizeof(double)));
  if (!m.rmat) return out_of_memory(err);
  // the first three lines, then the body from line 3
  size_t pos = 0, b0 = S;
  const char* l2p = nullptr;
  const char* l2e = nullptr;
  for (int ln = 0; ln < 3 && pos <= S; ++ln) {
    const char* nl = pos < S ? static_cast<const char*>(std::memchr(buf + pos, '\n', S - pos)) : nullptr;
    const size_t end = nl ? static_cast<size_t>(nl - buf) : S;
    if (ln == 2) {
      l2p = buf + pos;
      l2e = buf + end;
    }
    pos = nl ? end + 1 : S;
    if (ln == 2) b0 = pos;
    if (!nl && ln < 2) {
      pos = S + 1;
      break;
    }
  }
  const LineRanges lr = split_lines(buf, S, std::min(b0, S), 3, n_threads);
  if (lr.bad_utf8 || !valid_utf8(reinterpret_cast<const unsigned char*>(buf),
                                 reinterpret_cast<const unsigned char*>(buf) + std::min(b0, S)))
    return file_parse(err, xyz_file, "Failed to read line: stream did not contain valid UTF-8");
  if (!l2p) return file_parse(err, file, "Truncated _r.dat: missing the R-vector count line");
  uint64_t u;
  int64_t sv;
  const char* pe = parse_int(trim(l2p, l2e), false, sv, u);
  if (*pe) return file_parse(err, file, std::string("Failed to parse n_R: ") + pe);
  const uint64_t nRr = u;
  if (n == 0) return RTB_OK;
  if (nRr > (1ull << 40) / n2 || 3 + nRr * n2 > lr.total_lines)
    return file_parse(err, file, "Truncated _r.dat: header declares " + std::to_string(nRr) + " R vectors of " +
                                     std::to_string(n) + "^2 elements, but the file has " +
                                     std::to_string(lr.total_lines) + " lines");
  if (m.weights.size() < nRr)
    return file_parse(err, file, "Truncated _r.dat: " + std::to_string(nRr) + " R vectors but " +
                                     std::to_string(m.weights.size()) + " _hr.dat degeneracy weights");
  std::unordered_map<RKey, size_t, RHash> index;
  for (int r = 0; r < m.nR; ++r)
    index.emplace(RKey{{m.hamR[r * 3], m.hamR[r * 3 + 1], m.hamR[r * 3 + 2]}}, static_cast<size_t>(r));  // find_R
  double* tmp = static_cast<double*>(std::malloc(std::max<size_t>(nRr * 3 * n2 * 2, 2) * sizeof(double)));
  if (!tmp) return out_of_memory(err);
  std::vector<int64_t> slot(nRr, -1);
  std::vector<RangeErr> rerr(lr.T);
  const uint64_t stop = 3 + nRr * n2;
  run_ranges(lr.T, [&](int t) {
    const char* p = buf + lr.cut[t];
    const char* e = buf + lr.cut[t + 1];
    uint64_t L = lr.first_line[t];
    RangeErr& re = rerr[t];
    auto set = [&](uint64_t key, std::string msg) {
      if (key < re.key) {
        re.key = key;
        re.msg = std::move(msg);
      }
    };
    while (p < e && L < stop) {
      const char* nl = static_cast<const char*>(std::memchr(p, '\n', static_cast<size_t>(e - p)));
      const char* le = nl ? nl : e;
      const uint64_t rel = L - 3, ir = rel / n2, w = rel % n2;
      const uint64_t ii = w / n, jj = w % n;  // ind_i (outer), ind_j (inner)
      const char* q = p;
      Tok tk;
      if (w == 0) {  // R, then its row in hamR (3 L + 0, 3 L + 1)
        int64_t R[3];
        bool ok = true;
        for (int d = 0; d < 3 && ok; ++d) {
          if (!next_token(q, le, tk)) {
            set(3 * L, "Missing R vector component");
            ok = false;
          } else {
            int64_t s2;
            uint64_t u2;
            const char* me = parse_int(tk, true, s2, u2);
            if (*me) {
              set(3 * L, std::string("Failed to parse R vector: ") + me);
              ok = false;
            }
            R[d] = s2;
          }
        }
        if (ok) {
          auto it = index.find(RKey{{R[0], R[1], R[2]}});
          if (it == index.end())
            set(3 * L + 1, "R vector [" + std::to_string(R[0]) + ", " + std::to_string(R[1]) + ", " +
                               std::to_string(R[2]) + "] not found in Hamiltonian");
          else
            slot[ir] = static_cast<int64_t>(it->second);
        }
        q = p;
      }
      int k = 0;
      while (k < 5 && next_token(q, le, tk)) ++k;
      const double wd = static_cast<double>(m.weights[ir]);
      for (int d = 0; d < 3; ++d) {
        double v[2] = {0.0, 0.0};
        for (int c = 0; c < 2; ++c) {
          const char* part = c == 0 ? "real" : "imaginary";
          if (k < 5 || !next_token(q, le, tk)) {
            set(3 * L + 2, std::string("Missing R matrix ") + part + " part");
            k = 0;
            break;
          }
          if (!parse_f64(tk, v[c])) {
            set(3 * L + 2, std::string("Failed to parse R matrix ") + part + " part: invalid float literal");
            break;
          }
        }
        double* dst = tmp + (((ir * 3 + d) * n + jj) * n + ii) * 2;
        dst[0] = v[0] / wd;
        dst[1] = v[1] / wd;
      }
      p = nl ? nl + 1 : e;
      ++L;
    }
  });
  const RangeErr* first = nullptr;
  for (const auto& r : rerr)
    if (r.key != ~0ull && (!first || r.key < first->key)) first = &r;
  if (first) {
    std::free(tmp);
    return file_parse(err, file, first->msg);
  }
  for (uint64_t ir = 0; ir < nRr; ++ir)  // later blocks of one R overwrite earlier ones
    std::memcpy(m.rmat + static_cast<size_t>(slot[ir]) * 3 * n2 * 2, tmp + ir * 3 * n2 * 2, 3 * n2 * 2 * sizeof(double));
  std::free(tmp);
  return RTB_OK;
}

// The r-matrix made Hermitian across +-R (wannier90.rs:1134-1160): for rows
// r = 0 .. nR - 2 in order (the reference's loop stops before the last row),
// rmat[r][d][i][j] = (rmat[r][d][i][j] + conj(rmat[-R][d][j][i])) / 2 and then
// rmat[-R][d][j][i] = conj(rmat[r][d][i][j]), element by element.
int hermitize_rmatrix(rtb_w90& m, Error& err) {
  const size_t n = static_cast<size_t>(m.nsta), n2 = n * n;
  std::unordered_map<RKey, size_t, RHash> index;
  for (int r = 0; r < m.nR; ++r) index.emplace(RKey{{m.hamR[r * 3], m.hamR[r * 3 + 1], m.hamR[r * 3 + 2]}}, r);
  for (int r = 0; r + 1 < m.nR; ++r) {
    auto it = index.find(RKey{{-m.hamR[r * 3], -m.hamR[r * 3 + 1], -m.hamR[r * 3 + 2]}});
    if (it == index.end()) {
      err.code = RTB_ERR_MISSING_CONJUGATE;
      err.msg = "Missing Hermitian conjugate for R vector: [" + std::to_string(m.hamR[r * 3]) + ", " +
                std::to_string(m.hamR[r * 3 + 1]) + ", " + std::to_string(m.hamR[r * 3 + 2]) + "]";
      return err.code;
    }
    for (size_t i = 0; i < n; ++i)
      for (size_t j = 0; j < n; ++j)
        for (int d = 0; d < 3; ++d) {
          double* a = m.rmat + ((static_cast<size_t>(r) * 3 + d) * n2 + i * n + j) * 2;
          double* b = m.rmat + ((it->second * 3 + d) * n2 + j * n + i) * 2;
          a[0] = (a[0] + b[0]) / 2.0;
          a[1] = (a[1] + -b[1]) / 2.0;
          b[0] = a[0];
          b[1] = -a[1];
        }
  }
  return RTB_OK;
}

struct WinData {
  double lat[9] = {};
  bool spin = false;
  std::vector<int> proj_species;            // per projection line
  std::vector<std::vector<int>> proj_orbs;  // per projection line
  std::vector<std::string> atom_name;       // atoms_cart / atoms_frac
  std::vector<double> atom_pos;             // Cartesian [natom][3]
};

int parse_win(const std::string& file, WinData& w, Error& err) {
  std::vector<std::string> lines;
  bool opened;
  if (read_lines(file, lines, err, opened) != RTB_OK) {
    if (!opened) err.msg = "Failed to create file '" + file + "': Unable to open win file: " + err.msg;
    return err.code;
  }
  size_t i = 0;
  auto next = [&](std::string& out) -> bool {
    if (i >= lines.size()) return false;
    out = lines[i++];
    return true;
  };
  auto coord = [&](const Tok& t, const char* what, double& v) -> bool {
    if (!parse_f64(t, v)) {
      file_parse(err, file, std::string("Failed to parse ") + what + ": invalid float literal");
      return false;
    }
    return true;
  };
  std::string raw;
  while (next(raw)) {
    const std::string key = lower(win_data_line(raw));
    if (key.find("begin unit_cell_cart") != std::string::npos) {
      double unit = 1.0;
      std::vector<double> rows;
      while (rows.size() < 9) {
        if (!next(raw)) return file_parse(err, file, "Missing lattice vector line");
        const std::string line = win_data_line(raw);
        if (line.empty()) continue;
        auto tk = tokens_of(line);
        if (rows.empty() && tk.size() == 1) {
          if (!length_unit_scale(tk[0].str(), unit))
            return file_parse(err, file, "Unknown unit_cell_cart unit '" + tk[0].str() + "'");
          continue;
        }
        if (tk.size() != 3)
          return file_parse(err, file, "A unit_cell_cart row must contain 3 numbers, found " +
                                           std::to_string(tk.size()) + " in '" + line + "'");
        for (int d = 0; d < 3; ++d) {
          double v;
          if (!coord(tk[d], "lattice vector", v)) return err.code;
          rows.push_back(v * unit);
        }
      }
      std::copy(rows.begin(), rows.end(), w.lat);
    } else if (key.find("spinors") != std::string::npos && key.find('t') != std::string::npos) {
      w.spin = true;
    } else if (key.find("begin projections") != std::string::npos) {
      for (;;) {
        if (!next(raw)) return file_parse(err, file, "Unexpected end of file");
        const std::string line = win_data_line(raw);
        if (line.empty()) continue;
        if (lower(line).find("end projections") != std::string::npos) break;
        std::vector<std::string> prj;
        size_t b = 0;
        for (size_t c = 0; c <= line.size(); ++c)
          if (c == line.size() || line[c] == ',' || line[c] == ';' || line[c] == ':') {
            prj.push_back(trim(line.data() + b, line.data() + c).str());
            b = c + 1;
          }
        if (prj.size() < 2 || prj[0].empty())
          return file_parse(err, file, "Malformed projection line '" + line + "': expected species:orbital");
        std::vector<int> orbs;
        for (size_t k = 1; k < prj.size(); ++k)
          if (!projection_orbitals(prj[k], orbs)) {
            err.code = RTB_ERR_ORBITAL_PROJECTION;
            err.msg = "Invalid orbital projection string: 'Unrecognized projection '" + prj[k] +
                      "' in seedname.win'";
            return err.code;
          }
        const int zs = element_number(prj[0]);
        if (!zs) return file_parse(err, file, "Unknown atomic species '" + prj[0] + "' in begin projections");
        w.proj_orbs.push_back(std::move(orbs));
        w.proj_species.push_back(zs);
      }
    } else if (key.find("begin atoms_cart") != std::string::npos) {
      double unit = 1.0;
      bool first = true;
      for (;;) {
        if (!next(raw)) return file_parse(err, file, "Unexpected end of file");
        const std::string line = win_data_line(raw);
        if (line.empty()) continue;
        if (lower(line).find("end atoms_cart") != std::string::npos) break;
        auto tk = tokens_of(line);
        if (first && tk.size() == 1) {
          if (!length_unit_scale(tk[0].str(), unit))
            return file_parse(err, file, "Unknown atoms_cart unit '" + tk[0].str() + "'");
          first = false;
          continue;
        }
        first = false;
        if (tk.size() != 4)
          return file_parse(err, file, "An atoms_cart row must contain a species and 3 coordinates, found " +
                                           std::to_string(tk.size()) + " fields in '" + line + "'");
        w.atom_name.push_back(tk[0].str());
        for (int d = 0; d < 3; ++d) {
          double v;
          if (!coord(tk[d + 1], "Cartesian atom position", v)) return err.code;
          w.atom_pos.push_back(v * unit);
        }
      }
    } else if (key.find("begin atoms_frac") != std::string::npos) {
      for (;;) {
        if (!next(raw)) return file_parse(err, file, "Unexpected end of file");
        const std::string line = win_data_line(raw);
        if (line.empty()) continue;
        if (lower(line).find("end atoms_frac") != std::string::npos) break;
        auto tk = tokens_of(line);
        if (tk.size() != 4)
          return file_parse(err, file, "An atoms_frac row must contain a species and 3 coordinates, found " +
                                           std::to_string(tk.size()) + " fields in '" + line + "'");
        w.atom_name.push_back(tk[0].str());
        double f[3], c[3];
        for (int d = 0; d < 3; ++d)
          if (!coord(tk[d + 1], "fractional atom position", f[d])) return err.code;
        row_times(f, w.lat, c);
        w.atom_pos.insert(w.atom_pos.end(), c, c + 3);
      }
    }
  }
  return RTB_OK;
}

int linalg_singular(Error& err) {
  err.code = RTB_ERR_LINALG;
  err.msg = "Linear algebra operation failed: the lattice matrix is singular";
  return err.code;
}

// _centres.xyz branch (wannier90.rs:565-711) or the .win atoms fallback
// (wannier90.rs:712-782): orbital positions, projections and atoms.
int place_orbitals(const std::string& xyz_file, const std::string& win_file, const WinData& w, int nsta,
                   rtb_w90& out, Error& err) {
  std::vector<std::string> lines;
  bool opened;
  Error e;
  double inv[9];
  const int spin = w.spin ? 1 : 0;
  if (read_lines(xyz_file, lines, e, opened) == RTB_OK || opened) {
    if (e.code != RTB_OK) return file_parse(err, xyz_file, "Failed to read line: stream did not contain valid UTF-8");
    if (lines.size() < 2) return file_parse(err, xyz_file, "_centres.xyz must contain a count and comment line");
    uint64_t u;
    int64_t s;
    const char* pe = parse_int(trim(lines[0].data(), lines[0].data() + lines[0].size()), false, s, u);
    if (*pe) return file_parse(err, xyz_file, std::string("Invalid _centres.xyz entry count: ") + pe);
    const uint64_t declared = u, available = lines.size() - 2;
    if (declared < static_cast<uint64_t>(nsta) || available < declared)
      return file_parse(err, xyz_file, "Truncated _centres.xyz: header declares " + std::to_string(declared) +
                                           " entries, HR requires " + std::to_string(nsta) +
                                           " Wannier-centre entries, but only " + std::to_string(available) +
                                           " data lines are present");
    const int norb = spin ? nsta / 2 : nsta;
    std::vector<double> cart(static_cast<size_t>(norb) * 3);
    for (int i = 0; i < norb; ++i) {
      auto tk = tokens_of(lines[i + 2]);
      if (tk.size() < 4)
        return file_parse(err, xyz_file, "Malformed Wannier-centre row " + std::to_string(i + 1) +
                                             ": expected a label and 3 coordinates");
      for (int d = 0; d < 3; ++d)
        if (!parse_f64(tk[d + 1], cart[i * 3 + d]))
          return file_parse(err, xyz_file, "Failed to parse Wannier-centre position: invalid float literal");
    }
    if (!inverse3(w.lat, inv)) return linalg_singular(err);
    out.orb.assign(static_cast<size_t>(norb) * 3, 0.0);
    for (int i = 0; i < norb; ++i) row_times(&cart[i * 3], inv, &out.orb[i * 3]);
    const uint64_t natom = declared - static_cast<uint64_t>(nsta);
    std::vector<double> apos(natom * 3);
    std::vector<int> aname(natom);
    for (uint64_t i = 0; i < natom; ++i) {
      auto tk = tokens_of(lines[i + 2 + nsta]);
      if (tk.size() < 4)
        return file_parse(err, xyz_file, "Malformed atom row " + std::to_string(i + 1) +
                                             " in _centres.xyz: expected a species and 3 coordinates");
      for (int d = 0; d < 3; ++d)
        if (!parse_f64(tk[d + 1], apos[i * 3 + d]))
          return file_parse(err, xyz_file, "Failed to parse atom position: invalid float literal");
      aname[i] = element_number(tk[0].str());
      if (!aname[i])
        return file_parse(err, xyz_file, "Unknown atomic species '" + tk[0].str() +
                                             "' in _centres.xyz; species must be listed in the win file's "
                                             "begin projections block");
    }
    std::vector<int> dropped;
    for (uint64_t i = 0; i < natom; ++i) {
      const int first = static_cast<int>(out.orb_proj.size());
      for (size_t j = 0; j < w.proj_species.size(); ++j)
        if (w.proj_species[j] == aname[i])
          out.orb_proj.insert(out.orb_proj.end(), w.proj_orbs[j].begin(), w.proj_orbs[j].end());
      const int count = static_cast<int>(out.orb_proj.size()) - first;
      if (count == 0) {
        if (std::find(dropped.begin(), dropped.end(), aname[i]) == dropped.end()) {
          std::fprintf(stderr,
                       "warning: dropping '%s' atoms from _centres.xyz (no orbitals in the projections block)\n",
                       kElements[aname[i] - 1]);
          dropped.push_back(aname[i]);
        }
        continue;
      }
      rtb_w90::Atom a{aname[i], {0, 0, 0}, first, count};
      row_times(&apos[i * 3], inv, a.pos);
      out.atoms.push_back(a);
    }
    if (static_cast<int>(out.orb_proj.size()) != norb) {
      std::string detail = std::to_string(norb) + " orbitals declared in the projections block, but " +
                           (static_cast<int>(out.orb_proj.size()) < norb ? "only " : "") +
                           std::to_string(out.orb_proj.size()) +
                           (static_cast<int>(out.orb_proj.size()) < norb ? " could be" : " were") +
                           " assigned to atoms";
      return file_parse(err, xyz_file, "species mismatch between _centres.xyz and the win projections block: " +
                                           detail);
    }
    out.norb = norb;
    return RTB_OK;
  }
  // no _centres.xyz: every projection line of a species gives its atoms
  // orbitals at the atom's position
  if (!inverse3(w.lat, inv)) return linalg_singular(err);
  std::vector<int> dropped;
  const size_t natom = w.atom_name.size();
  for (size_t i = 0; i < natom; ++i) {
    const int zs = element_number(w.atom_name[i]);
    if (!zs)
      return file_parse(err, win_file,
                        "Unknown atomic species '" + w.atom_name[i] + "' in begin atoms_frac block");
    double frac[3];
    row_times(&w.atom_pos[i * 3], inv, frac);
    const int first = static_cast<int>(out.orb_proj.size());
    for (size_t j = 0; j < w.proj_species.size(); ++j)
      if (w.proj_species[j] == zs) {
        out.orb_proj.insert(out.orb_proj.end(), w.proj_orbs[j].begin(), w.proj_orbs[j].end());
        for (size_t k = 0; k < w.proj_orbs[j].size(); ++k) out.orb.insert(out.orb.end(), frac, frac + 3);
      }
    const int count = static_cast<int>(out.orb_proj.size()) - first;
    if (count == 0) {
      if (std::find(dropped.begin(), dropped.end(), zs) == dropped.end()) {
        std::fprintf(stderr,
                     "warning: dropping '%s' atoms from begin atoms_frac/atoms_cart (no orbitals in the "
                     "projections block)\n",
                     kElements[zs - 1]);
        dropped.push_back(zs);
      }
      continue;
    }
    rtb_w90::Atom a{zs, {frac[0], frac[1], frac[2]}, first, count};
    out.atoms.push_back(a);
  }
  const int expected = spin ? nsta / 2 : nsta;
  const int got = static_cast<int>(out.orb_proj.size());
  if (got != expected) {
    std::string detail = std::to_string(expected) + " orbitals declared in the HR file, but " +
                         (got < expected ? "only " : "") + std::to_string(got) +
                         (got < expected ? " could be" : " were") + " assigned to atoms";
    return file_parse(err, win_file,
                      "species mismatch between begin atoms_frac and the projections block: " + detail);
  }
  out.norb = got;
  return RTB_OK;
}

// _wsvec.dat (wannier90.rs:1042-1133): each (R, i, j) hopping is divided by
// its shift count and added at R + shift for every listed shift; new R are
// appended in order of first appearance after R = 0.
int apply_wsvec(const std::string& file, const std::string& xyz_file, rtb_w90& m, Error& err) {
  std::vector<std::string> lines;
  bool opened;
  Error e;
  if (read_lines(file, lines, e, opened) != RTB_OK) {
    if (!opened) return RTB_OK;  // optional file
    // the reference names the _centres.xyz path in this error
    return file_parse(err, xyz_file, "Failed to read line: stream did not contain valid UTF-8");
  }
  const size_t n = static_cast<size_t>(m.nsta), blk = n * n * 2;
  std::unordered_map<RKey, size_t, RHash> old_idx, new_idx;
  for (int r = 0; r < m.nR; ++r) {
    RKey k{{m.hamR[r * 3], m.hamR[r * 3 + 1], m.hamR[r * 3 + 2]}};
    old_idx.emplace(k, static_cast<size_t>(r));  // first match, as find_R
  }
  std::vector<int64_t> nR_list = {0, 0, 0};
  std::vector<double> nham(blk, 0.0);
  const bool have_r = m.rmat != nullptr;  // wannier90.rs:929-1041: the r-matrix moves with the hopping
  std::vector<double> nrmat(have_r ? 3 * blk : 0, 0.0);
  new_idx.emplace(RKey{{0, 0, 0}}, 0);
  if (lines.empty()) return file_parse(err, file, "Truncated _wsvec.dat: empty file");
  size_t i = 0;
  uint64_t u;
  int64_t s;
  auto line_at = [&](size_t k, const std::string*& l) -> bool {
    if (k >= lines.size()) {
      file_parse(err, file, "Truncated _wsvec.dat: missing line " + std::to_string(k + 1));
      return false;
    }
    l = &lines[k];
    return true;
  };
  while (i < lines.size() - 1) {
    ++i;
    const std::string* l;
    if (!line_at(i, l)) return err.code;
    auto tk = tokens_of(*l);
    int64_t R[3];
    for (int d = 0; d < 3; ++d) {
      if (static_cast<size_t>(d) >= tk.size()) return file_parse(err, file, "Missing R vector component");
      const char* pe = parse_int(tk[d], true, s, u);
      if (*pe) return file_parse(err, file, std::string("Failed to parse R vector: ") + pe);
      R[d] = s;
    }
    uint64_t oi[2];
    for (int d = 0; d < 2; ++d) {
      if (static_cast<size_t>(3 + d) >= tk.size()) return file_parse(err, file, "Missing orbital index");
      const char* pe = parse_int(tk[3 + d], false, s, u);
      if (*pe) return file_parse(err, file, std::string("Failed to parse orbital index: ") + pe);
      if (u == 0 || u > n) return file_parse(err, file, "orbital index " + std::to_string(u) + " out of range");
      oi[d] = u - 1;
    }
    ++i;
    if (!line_at(i, l)) return err.code;
    tk = tokens_of(*l);
    if (tk.empty()) return file_parse(err, file, "Missing weight value");
    {
      const char* pe = parse_int(tk[0], false, s, u);
      if (*pe) return file_parse(err, file, std::string("Failed to parse weight: ") + pe);
    }
    const uint64_t weight = u;
    auto it = old_idx.find(RKey{{R[0], R[1], R[2]}});
    if (it == old_idx.end())
      return file_parse(err, file, "R vector [" + std::to_string(R[0]) + ", " + std::to_string(R[1]) + ", " +
                                       std::to_string(R[2]) + "] not found in Hamiltonian");
    const double* h = m.ham + (it->second * n * n + oi[0] * n + oi[1]) * 2;
    const double wd = static_cast<double>(weight);
    const double hre = h[0] / wd, him = h[1] / wd;
    double rre[3] = {0, 0, 0}, rim[3] = {0, 0, 0};
    if (have_r)
      for (int d = 0; d < 3; ++d) {
        const double* x = m.rmat + ((it->second * 3 + d) * n * n + oi[0] * n + oi[1]) * 2;
        rre[d] = x[0] / wd;
        rim[d] = x[1] / wd;
      }
    for (uint64_t k = 0; k < weight; ++k) {
      ++i;
      if (!line_at(i, l)) return err.code;
      tk = tokens_of(*l);
      int64_t sh[3];
      for (int d = 0; d < 3; ++d) {
        if (static_cast<size_t>(d) >= tk.size() || *parse_int(tk[d], true, s, u))
          return file_parse(err, file, "malformed shift line " + std::to_string(i + 1));
        sh[d] = s;
      }
      RKey nk{{R[0] + sh[0], R[1] + sh[1], R[2] + sh[2]}};
      auto f = new_idx.find(nk);
      size_t slot;
      if (f == new_idx.end()) {
        slot = nR_list.size() / 3;
        new_idx.emplace(nk, slot);
        nR_list.insert(nR_list.end(), nk.r, nk.r + 3);
        nham.resize(nham.size() + blk, 0.0);
        if (have_r) nrmat.resize(nrmat.size() + 3 * blk, 0.0);
      } else {
        slot = f->second;
      }
      double* d = nham.data() + (slot * n * n + oi[0] * n + oi[1]) * 2;
      if (f == new_idx.end() && !have_r) {  // a new block takes the value (keeps -0.0);
        d[0] = hre;                           // with _r.dat the reference adds to zero
        d[1] = him;
      } else {
        d[0] += hre;
        d[1] += him;
      }
      if (have_r)
        for (int c = 0; c < 3; ++c) {
          double* x = nrmat.data() + ((slot * 3 + c) * n * n + oi[0] * n + oi[1]) * 2;
          x[0] += rre[c];  // the reference adds here for new and old blocks alike
          x[1] += rim[c];
        }
    }
  }
  double* fresh = static_cast<double*>(std::malloc(std::max<size_t>(nham.size() * sizeof(double), 16)));
  if (!fresh) return out_of_memory(err);
  std::memcpy(fresh, nham.data(), nham.size() * sizeof(double));
  std::free(m.ham);
  m.ham = fresh;
  if (have_r) {
    double* fr = static_cast<double*>(std::malloc(std::max<size_t>(nrmat.size() * sizeof(double), 16)));
    if (!fr) return out_of_memory(err);
    std::memcpy(fr, nrmat.data(), nrmat.size() * sizeof(double));
    std::free(m.rmat);
    m.rmat = fr;
  }
  m.hamR = std::move(nR_list);
  m.nR = static_cast<int>(m.hamR.size() / 3);
  return RTB_OK;
}

// Model::validate (model.rs:564-667) for the loaded arrays.
int validate(const rtb_w90& m, Error& err) {
  auto invariant = [&](const char* name, const std::string& msg) {
    err.code = RTB_ERR_INVALID_MODEL;
    err.msg = std::string("invalid model invariant '") + name + "': " + msg;
    return err.code;
  };
  if (m.norb == 0) {
    err.code = RTB_ERR_NO_ORBITALS;
    err.msg = "a model must contain at least one orbital";
    return err.code;
  }
  for (double v : m.lat)
    if (!std::isfinite(v)) return invariant("finite_geometry", "lattice and orbital positions must be finite");
  for (double v : m.orb)
    if (!std::isfinite(v)) return invariant("finite_geometry", "lattice and orbital positions must be finite");
  for (size_t a = 0; a < m.atoms.size(); ++a)
    for (double v : m.atoms[a].pos)
      if (!std::isfinite(v))
        return invariant("atomic_position", "atom " + std::to_string(a) + " must have 3 finite coordinates");
  for (size_t a = 0; a < m.atoms.size(); ++a) {
    const auto& at = m.atoms[a];
    for (int o = at.first; o < at.first + at.count; ++o)
      for (int d = 0; d < 3; ++d) {
        double delta = std::fmod(std::fabs(m.orb[o * 3 + d] - at.pos[d]), 1.0);
        if (delta < 0) delta += 1.0;
        const double dist = std::min(delta, 1.0 - delta);
        if (dist > 0.1) {
          char buf[64];
          std::snprintf(buf, sizeof buf, "%.3f", dist);
          return invariant("orbital_atom_position",
                           "orbital " + std::to_string(o) + " is " + buf +
                               " (fractional, modulo a lattice vector) away from its parent atom " +
                               std::to_string(a) +
                               ", exceeding the tolerance 0.1; orbitals must sit at their parent atom's position");
        }
      }
  }
  const int nsta_model = m.spin ? 2 * m.norb : m.norb;
  if (nsta_model != m.nsta)
    return invariant("hamiltonian_shape", "expected (" + std::to_string(m.nR) + ", " + std::to_string(nsta_model) +
                                              ", " + std::to_string(nsta_model) + "), found (" +
                                              std::to_string(m.nR) + ", " + std::to_string(m.nsta) + ", " +
                                              std::to_string(m.nsta) + ")");
  return RTB_OK;
}

}  // namespace

extern "C" {

static int w90_read(const char* path, const char* seedname, double zero_energy, int spin, int n_threads,
                    bool with_r, rtb_w90** out) {
  if (!path || !seedname || !out) return rtb::fail(RTB_ERR_INVALID_ARGUMENT, "null argument");
  *out = nullptr;
  if (spin < -1 || spin > 1) return rtb::fail(RTB_ERR_INVALID_ARGUMENT, "spin must be -1, 0 or 1");
  const std::string prefix = std::string(path) + seedname;
  auto* m = new rtb_w90();
  Error err;
  auto bail = [&]() {
    delete m;
    rtb::set_error("%s", err.msg.c_str());
    return err.code;
  };
  if (parse_hr(prefix + "_hr.dat", n_threads, zero_energy, *m, err) != RTB_OK) return bail();
  WinData w;
  if (parse_win(prefix + ".win", w, err) != RTB_OK) return bail();
  if (spin >= 0 && (w.spin ? 1 : 0) != spin) {
    err.code = RTB_ERR_SPIN_MISMATCH;
    err.msg = std::string("Unexpected error: Spin mismatch: Wannier90 .win file has spin=") +
              (w.spin ? "true" : "false") + " but Model was constructed with SPIN=" + (spin ? "true" : "false");
    return bail();
  }
  m->spin = w.spin ? 1 : 0;
  std::copy(w.lat, w.lat + 9, m->lat);
  if (place_orbitals(prefix + "_centres.xyz", prefix + ".win", w, m->nsta, *m, err) != RTB_OK) return bail();
  if (with_r && parse_rdat(prefix + "_r.dat", prefix + "_centres.xyz", n_threads, *m, err) != RTB_OK) return bail();
  if (apply_wsvec(prefix + "_wsvec.dat", prefix + "_centres.xyz", *m, err) != RTB_OK) return bail();
  if (with_r && hermitize_rmatrix(*m, err) != RTB_OK) return bail();
  if (validate(*m, err) != RTB_OK) return bail();
  *out = m;
  return RTB_OK;
}

int rtb_w90_read(const char* path, const char* seedname, double zero_energy, int spin, int n_threads,
                 rtb_w90** out) {
  return w90_read(path, seedname, zero_energy, spin, n_threads, false, out);
}

int rtb_w90_read_with_rmatrix(const char* path, const char* seedname, double zero_energy, int spin, int n_threads,
                              rtb_w90** out) {
  return w90_read(path, seedname, zero_energy, spin, n_threads, true, out);
}

int rtb_w90_export_rmatrix(const rtb_w90* m, double* rmatrix) {
  if (!m || !rmatrix) return rtb::fail(RTB_ERR_INVALID_ARGUMENT, "null argument");
  if (!m->rmat) return rtb::fail(RTB_ERR_INVALID_ARGUMENT, "the model was read without _r.dat");
  std::memcpy(rmatrix, m->rmat, static_cast<size_t>(m->nR) * 3 * m->nsta * m->nsta * 2 * sizeof(double));
  return RTB_OK;
}

void rtb_w90_destroy(rtb_w90* m) { delete m; }

int rtb_w90_info(const rtb_w90* m, int* spin, int* norb, int* nsta, int* n_R, int* natom) {
  if (!m) return rtb::fail(RTB_ERR_INVALID_ARGUMENT, "null model");
  if (spin) *spin = m->spin;
  if (norb) *norb = m->norb;
  if (nsta) *nsta = m->nsta;
  if (n_R) *n_R = m->nR;
  if (natom) *natom = static_cast<int>(m->atoms.size());
  return RTB_OK;
}

int rtb_w90_export(const rtb_w90* m, int64_t* hamR, double* ham, double* orb, double* lat, int* orb_proj) {
  if (!m) return rtb::fail(RTB_ERR_INVALID_ARGUMENT, "null model");
  if (hamR) std::memcpy(hamR, m->hamR.data(), m->hamR.size() * sizeof(int64_t));
  if (ham) {
    const size_t total = static_cast<size_t>(m->nR) * m->nsta * m->nsta * 2;
    // large copies split over threads (a 1331 x 128 x 128 model is 349 MB)
    const int T = total > (1u << 24) ? 8 : 1;
    std::vector<std::thread> th;
    for (int t = 0; t < T; ++t) {
      const size_t a = total / T * t, b = t == T - 1 ? total : total / T * (t + 1);
      th.emplace_back([=] { std::memcpy(ham + a, m->ham + a, (b - a) * sizeof(double)); });
    }
    for (auto& x : th) x.join();
  }
  if (orb) std::memcpy(orb, m->orb.data(), m->orb.size() * sizeof(double));
  if (lat) std::memcpy(lat, m->lat, sizeof m->lat);
  if (orb_proj) std::copy(m->orb_proj.begin(), m->orb_proj.end(), orb_proj);
  return RTB_OK;
}

int rtb_w90_atom(const rtb_w90* m, int i, int* species, double* pos, int* first_orb, int* n_orb) {
  if (!m) return rtb::fail(RTB_ERR_INVALID_ARGUMENT, "null model");
  if (i < 0 || i >= static_cast<int>(m->atoms.size()))
    return rtb::fail(RTB_ERR_INVALID_ARGUMENT, "atom index out of range");
  const auto& a = m->atoms[i];
  if (species) *species = a.species;
  if (pos) std::memcpy(pos, a.pos, sizeof a.pos);
  if (first_orb) *first_orb = a.first;
  if (n_orb) *n_orb = a.count;
  return RTB_OK;
}

const char* rtb_element_symbol(int z) { return z >= 1 && z <= kNumElements ? kElements[z - 1] : ""; }

const char* rtb_orbital_projection_name(int code) {
  return code >= 0 && code < kNumOrbCodes ? kOrbNames[code] : "";
}

int rtb_w90_model(const rtb_w90* m, rtb_model** out) {
  if (!m || !out) return rtb::fail(RTB_ERR_INVALID_ARGUMENT, "null argument");
  RTB_TRY(rtb_model_create(3, m->spin, m->norb, m->nR, m->hamR.data(), m->ham, m->orb.data(), m->lat, out));
  if (m->rmat) {
    const int rc = rtb_model_set_rmatrix(*out, m->rmat);
    if (rc != RTB_OK) {
      rtb_model_destroy(*out);
      *out = nullptr;
      return rc;
    }
  }
  return RTB_OK;
}

}  // extern "C"
