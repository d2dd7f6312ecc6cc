// fidelity.cpp -- tensor-fidelity entry points of the C ABI: the explicit C
// tensor and its fail-closed compilation, the literal sitewise contraction
// (contract_rows_sitewise, src/lib.rs:5629-5750) on the GPU, and the layer
// trace certificate (canonical layers hashed with SHA-256).
#include <algorithm>
#include <chrono>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/peps_mi355x.h"
#include "engine.h"
#include "frontend.h"
#include "host_math.h"
#include "sha256.h"

using peps::abi;
using peps::fail;
using peps::u128;

namespace {

peps::TensorEntry from_abi(const peps_tensor_entry& e) {
  return peps::TensorEntry{0,
                           peps::Legs{e.column_in, e.column_out, e.row_in, e.row_out, e.diag_dr_in,
                                      e.diag_dr_out, e.diag_dl_in, e.diag_dl_out},
                           e.value};
}

peps_tensor_entry to_abi(const peps::TensorEntry& e) {
  peps_tensor_entry o;
  std::memset(&o, 0, sizeof(o));
  o.column_in = e.legs.column_in;
  o.column_out = e.legs.column_out;
  o.row_in = e.legs.row_in;
  o.row_out = e.legs.row_out;
  o.diag_dr_in = e.legs.diag_dr_in;
  o.diag_dr_out = e.legs.diag_dr_out;
  o.diag_dl_in = e.legs.diag_dl_in;
  o.diag_dl_out = e.legs.diag_dl_out;
  o.value = e.value;
  return o;
}

// The tensor as the sitewise kernels read it: entries grouped by the
// reference's input_key (src/lib.rs:157-162), in the given order within a
// group (SiteTensorC::from_b keeps legs_key order, src/lib.rs:113-124).
bool site_table(const std::vector<peps::TensorEntry>& c, const std::vector<uint64_t>& primes,
                peps::SiteTable* t, std::string* err) {
  if (c.size() > (size_t)peps::kSiteMaxEntries) {
    *err = "C has more than " + std::to_string(peps::kSiteMaxEntries) + " entries";
    return false;
  }
  std::memset(t, 0, sizeof(*t));
  unsigned pos = 0;
  for (unsigned in = 0; in < 16; ++in) {
    t->first[in] = (uint8_t)pos;
    for (const peps::TensorEntry& e : c) {
      const peps::Legs& l = e.legs;
      for (uint8_t v : {l.column_in, l.column_out, l.row_in, l.row_out, l.diag_dr_in, l.diag_dr_out,
                        l.diag_dl_in, l.diag_dl_out})
        if (v > 1) {
          *err = "C entry with a non-binary virtual leg";
          return false;
        }
      const unsigned key = l.column_in | (l.row_in << 1) | (l.diag_dr_in << 2) | (l.diag_dl_in << 3);
      if (key != in) continue;
      t->out[pos] = (uint8_t)(l.column_out | (l.row_out << 1) | (l.diag_dr_out << 2) | (l.diag_dl_out << 3));
      for (size_t q = 0; q < primes.size(); ++q) t->value[pos][q] = (uint32_t)(e.value % primes[q]);
      ++pos;
    }
    t->count[in] = (uint8_t)(pos - t->first[in]);
  }
  return true;
}

int literal(uint32_t n, const std::vector<peps::TensorEntry>& c, bool bound_check, const peps_options* opts,
            peps_contraction_result* out, char* err, size_t err_cap) {
  if (!out) return fail(PEPS_E_INVALID, "null result pointer", err, err_cap);
  std::memset(out, 0, sizeof(*out));
  out->n = n;
  if (n > 42) return fail(PEPS_E_INVALID, "the packed u128 virtual-boundary backend supports N <= 42", err, err_cap);
  if (n == 0) {  // src/lib.rs:5640-5653
    out->count = abi(1);
    out->peak_states = 1;
    out->peak_rss_bytes = peps::rss_bytes();
    return PEPS_OK;
  }
  std::string e;
  int rc = peps::check_board(n, &e);
  if (rc) return fail(rc, e, err, err_cap);
  peps::Resolved r;
  rc = peps::resolve_options(n, opts, &r, &e);
  if (rc) return fail(rc, e, err, err_cap);
  peps::SiteTable table;
  if (!site_table(c, r.primes, &table, &e)) return fail(PEPS_E_INVALID, e, err, err_cap);
  peps::RunSpec spec;
  spec.n = n;
  spec.seeds.push_back(peps::Seed{0, 1});
  spec.site = &table;
  spec.materialize = true;
  spec.bound_check = bound_check;
  peps::RunOut ro;
  const auto t0 = std::chrono::steady_clock::now();
  rc = peps::run_engine(r, spec, &ro, &e);
  if (rc) return fail(rc, e, err, err_cap);
  const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  peps::fill_result(n, ro, r.primes, el, out,
                    r.site_counters ? peps::TensorCounters::kSitewiseLinear : peps::TensorCounters::kSitewise,
                    c.size());
  return PEPS_OK;
}

struct CanonicalState {
  uint64_t cols, dr, dl;
  u128 weight;
  bool operator<(const CanonicalState& o) const {
    if (cols != o.cols) return cols < o.cols;
    if (dr != o.dr) return dr < o.dr;
    if (dl != o.dl) return dl < o.dl;
    return weight < o.weight;
  }
};

}  // namespace

extern "C" {

int peps_site_tensor_c(peps_tensor_entry* out, uint32_t cap, uint32_t* count) {
  const std::vector<peps::TensorEntry> c = peps::site_tensor_c(peps::site_tensor_b());
  if (count) *count = (uint32_t)c.size();
  for (uint32_t i = 0; out && i < cap && i < c.size(); ++i) out[i] = to_abi(c[i]);
  return PEPS_OK;
}

int peps_compile_row_operator(const peps_tensor_entry* c, uint32_t count, peps_tensor_entry* occupied,
                              char* err, size_t err_cap) {
  if (!c && count) return fail(PEPS_E_INVALID, "null tensor", err, err_cap);
  std::vector<peps::TensorEntry> t;
  for (uint32_t i = 0; i < count; ++i) t.push_back(from_abi(c[i]));
  peps::Legs occ;
  std::string e;
  if (!peps::compile_row_operator(t, &occ, &e)) return fail(PEPS_E_INVALID, e, err, err_cap);
  if (occupied) *occupied = to_abi(peps::TensorEntry{0, occ, 1});
  return PEPS_OK;
}

int peps_contract_rows_sitewise(uint32_t n, const peps_options* opts, peps_contraction_result* out,
                                char* err, size_t err_cap) {
  return literal(n, peps::site_tensor_c(peps::site_tensor_b()), true, opts, out, err, err_cap);
}

int peps_contract_rows_with_tensor(uint32_t n, const peps_tensor_entry* c, uint32_t count,
                                   const peps_options* opts, peps_contraction_result* out, char* err,
                                   size_t err_cap) {
  if (!c && count) return fail(PEPS_E_INVALID, "null tensor", err, err_cap);
  std::vector<peps::TensorEntry> t;
  for (uint32_t i = 0; i < count; ++i) t.push_back(from_abi(c[i]));
  return literal(n, t, false, opts, out, err, err_cap);
}

int peps_trace_layers(uint32_t n, uint32_t row_operator, uint32_t symmetry, const peps_tensor_entry* c,
                      uint32_t c_count, const peps_options* opts, peps_layer_digest* out, uint32_t cap,
                      uint32_t* rows, char* err, size_t err_cap) {
  if (rows) *rows = 0;
  if (row_operator > PEPS_ROW_OPERATOR_SITEWISE) return fail(PEPS_E_INVALID, "unknown row operator", err, err_cap);
  if (symmetry > PEPS_SYMMETRY_TOP_ROW_ORBITS) return fail(PEPS_E_INVALID, "unknown symmetry mode", err, err_cap);
  if (row_operator == PEPS_ROW_OPERATOR_SITEWISE && symmetry != PEPS_SYMMETRY_NONE)
    return fail(PEPS_E_INVALID, "the literal sitewise row transfer runs without row-0 orbit slicing", err,
                err_cap);
  if (!c && c_count) return fail(PEPS_E_INVALID, "null tensor", err, err_cap);
  std::string e;
  int rc = peps::check_board(n, &e);
  if (rc) return fail(rc, e, err, err_cap);
  if (n == 0) return PEPS_OK;
  std::vector<peps::TensorEntry> tensor;
  if (c)
    for (uint32_t i = 0; i < c_count; ++i) tensor.push_back(from_abi(c[i]));
  else
    tensor = peps::site_tensor_c(peps::site_tensor_b());
  peps::Resolved r;
  rc = peps::resolve_options(n, opts, &r, &e);
  if (rc) return fail(rc, e, err, err_cap);
  peps::SiteTable table;
  peps::RunSpec spec;
  spec.n = n;
  spec.seeds.push_back(peps::Seed{0, 1});
  spec.materialize = true;
  spec.orbit_first_row = symmetry == PEPS_SYMMETRY_TOP_ROW_ORBITS;
  if (row_operator == PEPS_ROW_OPERATOR_COMPILED) {
    // Fail closed: the compiled relation is only valid for exactly Sec. VI's C.
    peps::Legs occ;
    if (!peps::compile_row_operator(tensor, &occ, &e)) return fail(PEPS_E_INVALID, e, err, err_cap);
  } else {
    if (!site_table(tensor, r.primes, &table, &e)) return fail(PEPS_E_INVALID, e, err, err_cap);
    spec.site = &table;
    spec.bound_check = c == nullptr;
  }
  const uint64_t bm = ((uint64_t)1 << n) - 1;
  std::vector<peps_layer_digest> digests;
  const std::vector<uint64_t> primes = r.primes;
  spec.on_layer = [&](unsigned row, const std::vector<uint64_t>& keys, const std::vector<uint32_t>& res,
                      std::string* lerr) -> int {
    const size_t S = keys.size(), L = primes.size();
    std::vector<CanonicalState> st(S);
    std::vector<uint64_t> rr(L);
    u128 sum = 0;
    for (size_t i = 0; i < S; ++i) {
      for (size_t l = 0; l < L; ++l) rr[l] = res[l * S + i];
      u128 w;
      if (!peps::crt_reconstruct(rr.data(), primes.data(), (unsigned)L, &w, lerr)) return PEPS_E_OVERFLOW;
      st[i] = CanonicalState{keys[i] & bm, (keys[i] >> n) & bm, (keys[i] >> (2 * n)) & bm, w};
      sum += w;
    }
    std::sort(st.begin(), st.end());
    peps::Sha256 h;
    std::vector<uint8_t> buf;
    buf.reserve(40 * 4096);
    auto put = [&](uint64_t v) {
      for (int b = 0; b < 8; ++b) buf.push_back((uint8_t)(v >> (8 * b)));
    };
    for (const CanonicalState& s : st) {
      put(s.cols);
      put(s.dr);
      put(s.dl);
      put((uint64_t)s.weight);
      put((uint64_t)(s.weight >> 64));
      if (buf.size() >= 40 * 4096) {
        h.update(buf.data(), buf.size());
        buf.clear();
      }
    }
    h.update(buf.data(), buf.size());
    peps_layer_digest d;
    std::memset(&d, 0, sizeof(d));
    d.row = row;
    d.support = S;
    d.weight_sum = abi(sum);
    h.finish(d.sha256);
    digests.push_back(d);
    return 0;
  };
  peps::RunOut ro;
  rc = peps::run_engine(r, spec, &ro, &e);
  if (rc) return fail(rc, e, err, err_cap);
  for (size_t i = 0; i < digests.size() && i < ro.layers.size(); ++i) digests[i].transitions = abi(ro.layers[i].completed);
  if (rows) *rows = (uint32_t)digests.size();
  for (uint32_t i = 0; out && i < cap && i < digests.size(); ++i) out[i] = digests[i];
  return PEPS_OK;
}

int peps_sha256(const void* data, size_t len, uint8_t out[32]) {
  if (!out || (!data && len)) return PEPS_E_INVALID;
  peps::Sha256 h;
  h.update(data, len);
  h.finish(out);
  return PEPS_OK;
}

}  // extern "C"
