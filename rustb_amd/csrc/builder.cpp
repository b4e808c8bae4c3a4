// Host-side model builder and seeded synthetic workloads.
//
// The builder mirrors the hopping API of the reference (model_build.rs):
// storage is `ham[iR][i][j] = <i,0|H|j,R>` with row 0 at R = 0, a new R appends
// the pair (R, -R), and every write to R also writes the Hermitian partner at
// -R.  Unlike the reference's linear `find_R` scan (model_utils.rs:18-34), R
// lookup is a hash map, so building a 1331-R model is linear in its size.
#include <cmath>
#include <complex>
#include <cstring>
#include <map>
#include <unordered_map>

#include "rtb_internal.hpp"

namespace {

using cd = std::complex<double>;

struct RKey {
  int64_t r[3];
  bool operator==(const RKey& o) const { return r[0] == o.r[0] && r[1] == o.r[1] && r[2] == o.r[2]; }
};
struct RHash {
  size_t operator()(const RKey& k) const {
    uint64_t h = 1469598103934665603ull;
    for (int d = 0; d < 3; ++d) {
      h ^= static_cast<uint64_t>(k.r[d]) + 0x9e3779b97f4a7c15ull + (h << 6) + (h >> 2);
    }
    return static_cast<size_t>(h);
  }
};

}  // namespace

struct rtb_builder {
  int dim = 0, spin = 0, norb = 0, nsta = 0;
  std::vector<double> lat, orb;
  std::vector<int64_t> hamR;      // [nR][dim]
  std::vector<cd> ham;            // [nR][n][n]
  std::unordered_map<RKey, int, RHash> index;

  RKey key(const int64_t* R) const {
    RKey k{{0, 0, 0}};
    for (int d = 0; d < dim; ++d) k.r[d] = R[d];
    return k;
  }
  int find(const int64_t* R) const {
    auto it = index.find(key(R));
    return it == index.end() ? -1 : it->second;
  }
  int nR() const { return static_cast<int>(hamR.size() / dim); }
  cd* block(int iR) { return ham.data() + static_cast<size_t>(iR) * nsta * nsta; }
  int push(const int64_t* R) {
    int iR = nR();
    for (int d = 0; d < dim; ++d) hamR.push_back(R[d]);
    ham.resize(ham.size() + static_cast<size_t>(nsta) * nsta, cd(0.0, 0.0));
    index[key(R)] = iR;
    return iR;
  }
};

namespace {

// update_hamiltonian! / add_hamiltonian! (model_build.rs:71-145).
void decorate(rtb_builder* b, cd* blk, cd tmp, int i, int j, int pauli, bool acc) {
  const int n = b->nsta, no = b->norb;
  auto put = [&](int r, int c, cd v) {
    if (acc) {
      blk[r * n + c] += v;
    } else {
      blk[r * n + c] = v;
    }
  };
  if (!b->spin) {
    put(i, j, tmp);
    return;
  }
  const cd I(0.0, 1.0);
  switch (pauli) {
    case RTB_SPIN_NONE:
      put(i, j, tmp);
      put(i + no, j + no, tmp);
      break;
    case RTB_SPIN_X:
      put(i + no, j, tmp);
      put(i, j + no, tmp);
      break;
    case RTB_SPIN_Y:
      put(i + no, j, tmp * I);
      put(i, j + no, -tmp * I);
      break;
    default:  // RTB_SPIN_Z
      put(i, j, tmp);
      put(i + no, j + no, -tmp);
      break;
  }
}

int hop(rtb_builder* b, double re, double im, int i, int j, const int64_t* R, int pauli, bool acc) {
  if (!b || !R) return rtb::fail(RTB_ERR_INVALID_ARGUMENT, "null builder or R");
  if (i < 0 || j < 0 || i >= b->norb || j >= b->norb)
    return rtb::fail(RTB_ERR_DIMENSION_MISMATCH, "ind_i and ind_j must be less than norb");
  if (pauli < RTB_SPIN_NONE || pauli > RTB_SPIN_Z)
    return rtb::fail(RTB_ERR_INVALID_ARGUMENT, "pauli must be RTB_SPIN_NONE..RTB_SPIN_Z");
  bool zeroR = true;
  for (int d = 0; d < b->dim; ++d) zeroR = zeroR && R[d] == 0;
  if (i == j && zeroR && im != 0.0)  // model_build.rs:505
    return rtb::fail(RTB_ERR_ONSITE_NOT_REAL, "the onsite hopping must be real");
  if (!b->spin) pauli = RTB_SPIN_NONE;
  const cd tmp(re, im);
  int64_t mR[3] = {0, 0, 0};
  for (int d = 0; d < b->dim; ++d) mR[d] = -R[d];
  int idx = b->find(R);
  if (idx >= 0) {
    int inv = b->find(mR);
    decorate(b, b->block(idx), tmp, i, j, pauli, acc);
    if (idx != 0 || i != j) decorate(b, b->block(inv), std::conj(tmp), j, i, pauli, acc);
  } else {
    // a new R uses set semantics for both blocks (model_build.rs:545-556)
    int a = b->push(R);
    decorate(b, b->block(a), tmp, i, j, pauli, false);
    int c = b->push(mR);
    decorate(b, b->block(c), std::conj(tmp), j, i, pauli, false);
  }
  return RTB_OK;
}

// ---- SplitMix64-seeded xoshiro256** with Box-Muller normals ----
struct Rng {
  uint64_t s[4];
  explicit Rng(uint64_t seed) {
    uint64_t x = seed;
    for (int i = 0; i < 4; ++i) {
      x += 0x9e3779b97f4a7c15ull;
      uint64_t z = x;
      z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
      z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
      s[i] = z ^ (z >> 31);
    }
  }
  static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  uint64_t next() {
    uint64_t r = rotl(s[1] * 5, 7) * 9;
    uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return r;
  }
  double uniform() { return static_cast<double>(next() >> 11) * 0x1.0p-53; }  // [0,1)
  double uniform(double a, double b) { return a + (b - a) * uniform(); }
  // standard complex normal CN(0,1): (N(0,1) + i N(0,1)) / sqrt(2)
  cd cnormal() {
    double u1 = (static_cast<double>(next() >> 11) + 1.0) * 0x1.0p-53;  // (0,1]
    double u2 = uniform();
    double r = std::sqrt(-2.0 * std::log(u1));
    return cd(r * std::cos(2.0 * M_PI * u2), r * std::sin(2.0 * M_PI * u2)) * M_SQRT1_2;
  }
  double normal() { return cnormal().real() * M_SQRT2; }
  uint64_t below(uint64_t n) { return next() % n; }
};

// canonical half: first nonzero component positive
bool canonical(const int64_t* R, int dim) {
  for (int d = 0; d < dim; ++d) {
    if (R[d] > 0) return true;
    if (R[d] < 0) return false;
  }
  return false;
}

}  // namespace

extern "C" {

int rtb_builder_new(int dim, int spin, int norb, const double* lat, const double* orb,
                    rtb_builder** out) {
  if (!out || !lat || !orb) return rtb::fail(RTB_ERR_INVALID_ARGUMENT, "null argument");
  *out = nullptr;
  if (dim < 1 || dim > 3) return rtb::fail(RTB_ERR_INVALID_MODEL, "dim must be 1, 2 or 3");
  if (norb < 1) return rtb::fail(RTB_ERR_INVALID_MODEL, "the model has no orbitals");
  for (int i = 0; i < dim * dim; ++i)
    if (!std::isfinite(lat[i])) return rtb::fail(RTB_ERR_INVALID_MODEL, "lattice must be finite");
  for (int i = 0; i < norb * dim; ++i)
    if (!std::isfinite(orb[i])) return rtb::fail(RTB_ERR_INVALID_MODEL, "orbitals must be finite");
  auto* b = new rtb_builder();
  b->dim = dim;
  b->spin = spin ? 1 : 0;
  b->norb = norb;
  b->nsta = spin ? 2 * norb : norb;
  b->lat.assign(lat, lat + dim * dim);
  b->orb.assign(orb, orb + norb * dim);
  int64_t zero[3] = {0, 0, 0};
  b->push(zero);  // model_build.rs:244-245
  *out = b;
  return RTB_OK;
}

void rtb_builder_destroy(rtb_builder* b) { delete b; }

int rtb_builder_set_hop(rtb_builder* b, double re, double im, int i, int j, const int64_t* R,
                        int pauli) {
  return hop(b, re, im, i, j, R, pauli, false);
}

int rtb_builder_add_hop(rtb_builder* b, double re, double im, int i, int j, const int64_t* R,
                        int pauli) {
  return hop(b, re, im, i, j, R, pauli, true);
}

int rtb_builder_add_element(rtb_builder* b, double re, double im, int i, int j, const int64_t* R) {
  if (!b || !R) return rtb::fail(RTB_ERR_INVALID_ARGUMENT, "null builder or R");
  if (i < 0 || j < 0 || i >= b->nsta || j >= b->nsta)
    return rtb::fail(RTB_ERR_DIMENSION_MISMATCH, "orbital indices must be less than nsta");
  bool zeroR = true;
  for (int d = 0; d < b->dim; ++d) zeroR = zeroR && R[d] == 0;
  if (i == j && zeroR && im != 0.0)
    return rtb::fail(RTB_ERR_ONSITE_NOT_REAL, "the onsite hopping must be real");
  const cd tmp(re, im);
  int64_t mR[3] = {0, 0, 0};
  for (int d = 0; d < b->dim; ++d) mR[d] = -R[d];
  const int n = b->nsta;
  int idx = b->find(R);
  if (idx >= 0) {
    int inv = b->find(mR);
    b->block(idx)[i * n + j] = tmp;
    if (idx != 0 || i != j) b->block(inv)[j * n + i] = std::conj(tmp);
  } else {
    int a = b->push(R);
    b->block(a)[i * n + j] = tmp;
    int c = b->push(mR);
    b->block(c)[j * n + i] = std::conj(tmp);
  }
  return RTB_OK;
}

int rtb_builder_set_onsite(rtb_builder* b, const double* values, int pauli) {
  if (!b || !values) return rtb::fail(RTB_ERR_INVALID_ARGUMENT, "null argument");
  int64_t zero[3] = {0, 0, 0};
  for (int i = 0; i < b->norb; ++i) RTB_TRY(hop(b, values[i], 0.0, i, i, zero, pauli, false));
  return RTB_OK;
}

int rtb_builder_add_onsite(rtb_builder* b, const double* values, int pauli) {
  if (!b || !values) return rtb::fail(RTB_ERR_INVALID_ARGUMENT, "null argument");
  int64_t zero[3] = {0, 0, 0};
  for (int i = 0; i < b->norb; ++i) RTB_TRY(hop(b, values[i], 0.0, i, i, zero, pauli, true));
  return RTB_OK;
}

int rtb_builder_n_R(const rtb_builder* b) { return b ? b->nR() : RTB_ERR_INVALID_ARGUMENT; }
int rtb_builder_nsta(const rtb_builder* b) { return b ? b->nsta : RTB_ERR_INVALID_ARGUMENT; }

int rtb_builder_export(const rtb_builder* b, int64_t* hamR, double* ham) {
  if (!b || !hamR || !ham) return rtb::fail(RTB_ERR_INVALID_ARGUMENT, "null argument");
  std::memcpy(hamR, b->hamR.data(), b->hamR.size() * sizeof(int64_t));
  std::memcpy(ham, b->ham.data(), b->ham.size() * sizeof(cd));
  return RTB_OK;
}

int rtb_synth_model(const rtb_synth_spec* sp, int* n_R, int* nsta, int64_t* hamR, double* ham,
                    double* orb, double* lat) {
  if (!sp || !n_R || !nsta) return rtb::fail(RTB_ERR_INVALID_ARGUMENT, "null argument");
  const int dim = sp->dim;
  if (dim < 1 || dim > 3) return rtb::fail(RTB_ERR_INVALID_MODEL, "dim must be 1..3");
  const int norb = sp->kind == 4 ? sp->norb * sp->layers : sp->norb;
  const int n = sp->spin ? 2 * norb : norb;
  const int side = 2 * sp->rmax + 1;
  int total = 1;
  for (int d = 0; d < dim; ++d) total *= side;
  // canonical R list in lexicographic order (first axis slowest)
  std::vector<int64_t> canon;
  for (int flat = 0; flat < total; ++flat) {
    int64_t R[3] = {0, 0, 0};
    int rem = flat;
    for (int d = dim - 1; d >= 0; --d) {
      R[d] = rem % side - sp->rmax;
      rem /= side;
    }
    if (canonical(R, dim)) canon.insert(canon.end(), R, R + dim);
  }
  const int ncanon = static_cast<int>(canon.size() / dim);
  Rng rng(sp->seed);
  // kind 3 keeps a random subset of (n_R_target - 1) / 2 canonical vectors
  std::vector<int> keep;
  if (sp->kind == 3) {
    int want = (sp->n_R_target - 1) / 2;
    if (want < 0 || want > ncanon)
      return rtb::fail(RTB_ERR_INVALID_ARGUMENT, "n_R_target outside 1..(2*ncanon+1)");
    std::vector<int> perm(ncanon);
    for (int i = 0; i < ncanon; ++i) perm[i] = i;
    for (int i = 0; i < want; ++i) {  // partial Fisher-Yates
      int j = i + static_cast<int>(rng.below(static_cast<uint64_t>(ncanon - i)));
      std::swap(perm[i], perm[j]);
    }
    keep.assign(perm.begin(), perm.begin() + want);
    std::sort(keep.begin(), keep.end());
  } else {
    for (int i = 0; i < ncanon; ++i) keep.push_back(i);
  }
  const int nRtot = 1 + 2 * static_cast<int>(keep.size());
  *n_R = nRtot;
  *nsta = n;
  if (!ham) return RTB_OK;  // size query
  if (!hamR || !orb || !lat) return rtb::fail(RTB_ERR_INVALID_ARGUMENT, "null output array");

  for (int i = 0; i < dim * dim; ++i) lat[i] = (i % (dim + 1) == 0) ? 1.0 : 0.0;
  for (int i = 0; i < norb * dim; ++i) orb[i] = rng.uniform();
  std::memset(ham, 0, sizeof(double) * 2 * static_cast<size_t>(nRtot) * n * n);
  std::memset(hamR, 0, sizeof(int64_t) * static_cast<size_t>(nRtot) * dim);
  auto H = [&](int iR, int r, int c) -> cd& {
    return reinterpret_cast<cd*>(ham)[(static_cast<size_t>(iR) * n + r) * n + c];
  };
  // R = 0 block: real onsite diagonal, Hermitian upper triangle
  for (int i = 0; i < n; ++i) H(0, i, i) = cd(rng.uniform(-1.0, 1.0), 0.0);
  if (sp->kind == 4) {
    const int L = sp->layers, no = sp->norb;
    for (int i = 0; i < n; ++i) {
      for (int j = i + 1; j < n; ++j) {
        int li = (i % norb) / no, lj = (j % norb) / no;
        bool same_spin = (i < norb) == (j < norb);
        cd v(0.0, 0.0);
        if (li == lj) {
          v = rng.cnormal() * 0.1;  // SOC-like onsite N(0, 0.1)
        } else if (same_spin) {
          int dl = li > lj ? li - lj : lj - li;
          v = rng.cnormal() * (0.3 * std::pow(0.5, dl - 1));
        }
        (void)L;
        H(0, i, j) = v;
        H(0, j, i) = std::conj(v);
      }
    }
  } else {
    for (int i = 0; i < n; ++i) {
      for (int j = i + 1; j < n; ++j) {
        cd v = rng.cnormal() * sp->scale;
        H(0, i, j) = v;
        H(0, j, i) = std::conj(v);
      }
    }
  }
  int iR = 1;
  for (int c : keep) {
    const int64_t* R = &canon[static_cast<size_t>(c) * dim];
    double r2 = 0.0;
    for (int d = 0; d < dim; ++d) r2 += static_cast<double>(R[d] * R[d]);
    const double amp = sp->scale * std::exp(-std::sqrt(r2) / sp->decay);
    for (int d = 0; d < dim; ++d) {
      hamR[iR * dim + d] = R[d];
      hamR[(iR + 1) * dim + d] = -R[d];
    }
    for (int i = 0; i < n; ++i) {
      for (int j = 0; j < n; ++j) {
        cd v(0.0, 0.0);
        if (sp->kind == 4) {
          // in-plane hopping inside one layer and one spin block
          int li = (i % norb) / sp->norb, lj = (j % norb) / sp->norb;
          bool same_spin = (i < norb) == (j < norb);
          if (li == lj && same_spin) v = rng.cnormal() * amp;
        } else {
          v = rng.cnormal() * amp;
        }
        H(iR, i, j) = v;
        H(iR + 1, j, i) = std::conj(v);
      }
    }
    iR += 2;
  }
  return RTB_OK;
}

}  // extern "C"
