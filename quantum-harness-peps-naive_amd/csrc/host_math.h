// host_math.h -- host-side exact arithmetic and the local tensors of the
// product library: the Sec. VI site tensors B and C, the fail-closed
// compilation of C into the occupied entry, the CRT plan and the Garner
// reconstruction into a checked unsigned __int128.
#pragma once

#include <stdint.h>

#include <string>
#include <vector>

namespace peps {

using u128 = unsigned __int128;

struct Legs {
  uint8_t column_in, column_out, row_in, row_out, diag_dr_in, diag_dr_out, diag_dl_in, diag_dl_out;
};

struct TensorEntry {
  uint8_t alpha;  // physical index (B only)
  Legs legs;
  uint64_t value;
};

// SiteTensorB::sec_vi (reference src/lib.rs:43-83).
std::vector<TensorEntry> site_tensor_b();
// SiteTensorC::from_b (src/lib.rs:102-131): sum over alpha, sorted by legs key.
std::vector<TensorEntry> site_tensor_c(const std::vector<TensorEntry>& b);
// CompiledRowOperator::compile (src/lib.rs:437-485).  Returns false and sets
// *err unless C is exactly 16 identity pass-throughs plus one unit 0->1
// all-channel emission; on success *occupied is that entry.
bool compile_row_operator(const std::vector<TensorEntry>& c, Legs* occupied, std::string* err);

bool checked_add(u128 a, u128 b, u128* out);
bool checked_mul(u128 a, u128 b, u128* out);

// wide_crt_plan (src/lib.rs:3859-3886): shortest prefix of the reference's
// 32-bit primes whose product exceeds N!.
bool crt_plan(unsigned n, std::vector<uint64_t>* primes, u128* product, u128* factorial,
              std::string* err);
// The reference's prime list (CRT_PRIME_CANDIDATES, src/lib.rs:3767).
const uint64_t* crt_primes();
// reconstruct_crt (src/lib.rs:4134-4163).
bool crt_reconstruct(const uint64_t* residues, const uint64_t* primes, unsigned count, u128* out,
                     std::string* err);
bool factorial(unsigned n, u128* out);
// known_count (src/lib.rs:5752-5784).
bool known_count(unsigned n, u128* out);

std::string u128_str(u128 v);

}  // namespace peps
