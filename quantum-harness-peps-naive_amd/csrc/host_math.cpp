// host_math.cpp -- see host_math.h.
#include "host_math.h"

#include <algorithm>

namespace peps {

bool checked_add(u128 a, u128 b, u128* out) { return !__builtin_add_overflow(a, b, out); }
bool checked_mul(u128 a, u128 b, u128* out) { return !__builtin_mul_overflow(a, b, out); }

static unsigned in_key(const Legs& l) {
  return l.column_in | (l.row_in << 1) | (l.diag_dr_in << 2) | (l.diag_dl_in << 3);
}

static unsigned all_key(const Legs& l) {
  return l.column_in | (l.column_out << 1) | (l.row_in << 2) | (l.row_out << 3) |
         (l.diag_dr_in << 4) | (l.diag_dr_out << 5) | (l.diag_dl_in << 6) | (l.diag_dl_out << 7);
}

// Eq. (16) of the paper: an empty site passes every signal through; an
// occupied site (alpha = 1) needs all four incoming signals 0 and emits 1 on
// every outgoing leg.
std::vector<TensorEntry> site_tensor_b() {
  std::vector<TensorEntry> b;
  for (unsigned sig = 0; sig < 16; ++sig) {
    uint8_t c = sig & 1, r = (sig >> 1) & 1, d1 = (sig >> 2) & 1, d2 = (sig >> 3) & 1;
    b.push_back(TensorEntry{0, Legs{c, c, r, r, d1, d1, d2, d2}, 1});
  }
  b.push_back(TensorEntry{1, Legs{0, 1, 0, 1, 0, 1, 0, 1}, 1});
  return b;
}

std::vector<TensorEntry> site_tensor_c(const std::vector<TensorEntry>& b) {
  std::vector<TensorEntry> c;
  for (const TensorEntry& e : b) {
    auto it = std::find_if(c.begin(), c.end(),
                           [&](const TensorEntry& x) { return all_key(x.legs) == all_key(e.legs); });
    if (it == c.end())
      c.push_back(TensorEntry{0, e.legs, e.value});
    else
      it->value += e.value;
  }
  std::sort(c.begin(), c.end(), [](const TensorEntry& x, const TensorEntry& y) {
    return all_key(x.legs) < all_key(y.legs);
  });
  return c;
}

bool compile_row_operator(const std::vector<TensorEntry>& c, Legs* occupied, std::string* err) {
  bool seen[16] = {false};
  int n_occ = 0;
  for (const TensorEntry& e : c) {
    const Legs& l = e.legs;
    const bool pass = l.column_in == l.column_out && l.row_in == l.row_out &&
                      l.diag_dr_in == l.diag_dr_out && l.diag_dl_in == l.diag_dl_out;
    const bool occ = l.column_in == 0 && l.column_out == 1 && l.row_in == 0 && l.row_out == 1 &&
                     l.diag_dr_in == 0 && l.diag_dr_out == 1 && l.diag_dl_in == 0 &&
                     l.diag_dl_out == 1;
    if (e.value != 1) {
      *err = "compiled row operator requires unit Sec. VI coefficients";
      return false;
    }
    if (pass) {
      if (seen[in_key(l)]) {
        const unsigned k = in_key(l);
        *err = "duplicate C pass-through signature ";
        for (int b = 3; b >= 0; --b) err->push_back((char)('0' + ((k >> b) & 1u)));
        return false;
      }
      seen[in_key(l)] = true;
    } else if (occ) {
      if (++n_occ > 1) {
        *err = "multiple occupied C entries";
        return false;
      }
      *occupied = l;
    } else {
      *err = "C contains an entry outside the Sec. VI row automaton";
      return false;
    }
  }
  for (bool s : seen)
    if (!s) {
      *err = "C is missing a pass-through signature";
      return false;
    }
  if (n_occ != 1) {
    *err = "C is missing the occupied entry";
    return false;
  }
  return true;
}

static const uint64_t kPrimes[4] = {4294967291ull, 4294967279ull, 4294967231ull, 4294967197ull};

const uint64_t* crt_primes() { return kPrimes; }

static bool is_prime(uint64_t v) {
  if (v < 2) return false;
  if (v % 2 == 0) return v == 2;
  for (uint64_t d = 3; d * d <= v; d += 2)
    if (v % d == 0) return false;
  return true;
}

bool factorial(unsigned n, u128* out) {
  u128 f = 1;
  for (unsigned k = 2; k <= n; ++k)
    if (!checked_mul(f, k, &f)) return false;
  *out = f;
  return true;
}

bool crt_plan(unsigned n, std::vector<uint64_t>* primes, u128* product, u128* fact,
              std::string* err) {
  if (n > 34) {
    *err = "wide CRT backend supports N <= 34 because 35! exceeds u128";
    return false;
  }
  if (!factorial(n, fact)) {
    *err = "N! exactness bound does not fit u128";
    return false;
  }
  primes->clear();
  u128 m = 1;
  for (uint64_t p : kPrimes) {
    if (!is_prime(p)) {
      *err = "CRT modulus failed deterministic primality check";
      return false;
    }
    primes->push_back(p);
    if (!checked_mul(m, p, &m)) {
      *err = "CRT modulus product overflow";
      return false;
    }
    if (m > *fact) {
      *product = m;
      return true;
    }
  }
  *err = "available certified CRT product does not exceed N! bound";
  return false;
}

static uint64_t pow_mod(uint64_t b, uint64_t e, uint64_t m) {
  uint64_t r = 1;
  b %= m;
  while (e) {
    if (e & 1) r = (uint64_t)(((u128)r * b) % m);
    b = (uint64_t)(((u128)b * b) % m);
    e >>= 1;
  }
  return r;
}

// Incremental CRT: x_{k+1} = x_k + M_k * ((r_k - x_k) * M_k^{-1} mod p_k),
// with the inverse from Fermat's little theorem.
bool crt_reconstruct(const uint64_t* residues, const uint64_t* primes, unsigned count, u128* out,
                     std::string* err) {
  if (count == 0) {
    *err = "CRT residue/modulus length mismatch";
    return false;
  }
  u128 x = 0, m = 1;
  for (unsigned i = 0; i < count; ++i) {
    const uint64_t p = primes[i];
    const uint64_t r = residues[i] % p;
    const uint64_t xm = (uint64_t)(x % p);
    const uint64_t delta = r >= xm ? r - xm : r + p - xm;
    const uint64_t inv = pow_mod((uint64_t)(m % p), p - 2, p);
    const uint64_t step = (uint64_t)(((u128)delta * inv) % p);
    u128 t;
    if (!checked_mul(m, step, &t) || !checked_add(x, t, &x)) {
      *err = "CRT reconstruction overflow";
      return false;
    }
    if (!checked_mul(m, p, &m)) {
      *err = "CRT reconstruction modulus overflow";
      return false;
    }
  }
  *out = x;
  return true;
}

bool known_count(unsigned n, u128* out) {
  static const char* kQ[28] = {"1",
                               "1",
                               "0",
                               "0",
                               "2",
                               "10",
                               "4",
                               "40",
                               "92",
                               "352",
                               "724",
                               "2680",
                               "14200",
                               "73712",
                               "365596",
                               "2279184",
                               "14772512",
                               "95815104",
                               "666090624",
                               "4968057848",
                               "39029188884",
                               "314666222712",
                               "2691008701644",
                               "24233937684440",
                               "227514171973736",
                               "2207893435808352",
                               "22317699616364044",
                               "234907967154122528"};
  if (n >= 28) return false;
  u128 v = 0;
  for (const char* c = kQ[n]; *c; ++c) v = v * 10 + (unsigned)(*c - '0');
  *out = v;
  return true;
}

std::string u128_str(u128 v) {
  if (v == 0) return "0";
  std::string s;
  while (v) {
    s.push_back((char)('0' + (int)(v % 10)));
    v /= 10;
  }
  std::reverse(s.begin(), s.end());
  return s;
}

}  // namespace peps
