// sha256.cpp -- SHA-256 as specified in FIPS 180-4.  The round constants are
// the first 32 fraction bits of the cube roots of the first 64 primes and the
// initial hash the first 32 fraction bits of the square roots of the first 8
// primes (FIPS 180-4, 4.2.2 and 5.3.3); both are computed here by exact
// integer roots rather than tabulated.
#include "sha256.h"

#include <algorithm>
#include <cstring>

namespace peps {

namespace {

using u128 = unsigned __int128;

// floor(p^(1/k) * 2^32) mod 2^32 for k = 2 or 3, by bisection on integers.
uint32_t root_fraction(uint64_t p, int k) {
  const u128 target = (u128)p << (32 * k);
  uint64_t lo = 0, hi = (uint64_t)1 << 42;  // p < 2^10, so the root is below 2^(10/k + 32)
  while (hi - lo > 1) {
    const uint64_t mid = lo + (hi - lo) / 2;
    u128 v = 1;
    for (int i = 0; i < k; ++i) v *= mid;
    if (v <= target)
      lo = mid;
    else
      hi = mid;
  }
  return (uint32_t)lo;
}

struct Constants {
  uint32_t k[64];
  uint32_t h0[8];
  Constants() {
    int found = 0;
    for (uint64_t c = 2; found < 64; ++c) {
      bool prime = true;
      for (uint64_t d = 2; d * d <= c; ++d)
        if (c % d == 0) prime = false;
      if (!prime) continue;
      if (found < 8) h0[found] = root_fraction(c, 2);
      k[found++] = root_fraction(c, 3);
    }
  }
};

const Constants& constants() {
  static const Constants c;
  return c;
}

inline uint32_t rotr(uint32_t x, int n) { return (x >> n) | (x << (32 - n)); }

}  // namespace

Sha256::Sha256() { std::memcpy(h_, constants().h0, sizeof(h_)); }

void Sha256::block(const uint8_t* p) {
  const uint32_t* K = constants().k;
  uint32_t w[64];
  for (int i = 0; i < 16; ++i)
    w[i] = ((uint32_t)p[4 * i] << 24) | ((uint32_t)p[4 * i + 1] << 16) | ((uint32_t)p[4 * i + 2] << 8) |
           (uint32_t)p[4 * i + 3];
  for (int i = 16; i < 64; ++i) {
    const uint32_t s0 = rotr(w[i - 15], 7) ^ rotr(w[i - 15], 18) ^ (w[i - 15] >> 3);
    const uint32_t s1 = rotr(w[i - 2], 17) ^ rotr(w[i - 2], 19) ^ (w[i - 2] >> 10);
    w[i] = w[i - 16] + s0 + w[i - 7] + s1;
  }
  uint32_t a = h_[0], b = h_[1], c = h_[2], d = h_[3], e = h_[4], f = h_[5], g = h_[6], h = h_[7];
  for (int i = 0; i < 64; ++i) {
    const uint32_t S1 = rotr(e, 6) ^ rotr(e, 11) ^ rotr(e, 25);
    const uint32_t ch = (e & f) ^ (~e & g);
    const uint32_t t1 = h + S1 + ch + K[i] + w[i];
    const uint32_t S0 = rotr(a, 2) ^ rotr(a, 13) ^ rotr(a, 22);
    const uint32_t maj = (a & b) ^ (a & c) ^ (b & c);
    const uint32_t t2 = S0 + maj;
    h = g;
    g = f;
    f = e;
    e = d + t1;
    d = c;
    c = b;
    b = a;
    a = t1 + t2;
  }
  h_[0] += a;
  h_[1] += b;
  h_[2] += c;
  h_[3] += d;
  h_[4] += e;
  h_[5] += f;
  h_[6] += g;
  h_[7] += h;
}

void Sha256::update(const void* data, size_t len) {
  const uint8_t* p = (const uint8_t*)data;
  bits_ += (uint64_t)len * 8;
  while (len) {
    const size_t take = std::min<size_t>(64 - fill_, len);
    std::memcpy(buf_ + fill_, p, take);
    fill_ += take;
    p += take;
    len -= take;
    if (fill_ == 64) {
      block(buf_);
      fill_ = 0;
    }
  }
}

void Sha256::finish(uint8_t out[32]) {
  const uint64_t bits = bits_;
  const uint8_t one = 0x80, zero = 0;
  update(&one, 1);
  while (fill_ != 56) update(&zero, 1);
  uint8_t len[8];
  for (int i = 0; i < 8; ++i) len[i] = (uint8_t)(bits >> (56 - 8 * i));
  update(len, 8);
  for (int i = 0; i < 8; ++i) {
    out[4 * i] = (uint8_t)(h_[i] >> 24);
    out[4 * i + 1] = (uint8_t)(h_[i] >> 16);
    out[4 * i + 2] = (uint8_t)(h_[i] >> 8);
    out[4 * i + 3] = (uint8_t)h_[i];
  }
}

}  // namespace peps
