// sha256.h -- SHA-256 (FIPS 180-4) for layer trace digests.
#pragma once

#include <stddef.h>
#include <stdint.h>

namespace peps {

class Sha256 {
 public:
  Sha256();
  void update(const void* data, size_t len);
  void finish(uint8_t out[32]);

 private:
  void block(const uint8_t* p);
  uint32_t h_[8];
  uint8_t buf_[64];
  size_t fill_ = 0;
  uint64_t bits_ = 0;
};

}  // namespace peps
