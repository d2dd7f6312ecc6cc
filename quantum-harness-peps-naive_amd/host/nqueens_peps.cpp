// nqueens_peps.cpp -- command-line driver over libpeps_mi355x.so with the
// reference CLI's grammar and output lines (src/main.rs:1-174):
//   nqueens-peps solve N [--layers]
//   nqueens-peps bench MAX_N [--min N] [--repeats R] [--csv]
// plus two forms the reference CLI does not have:
//   nqueens-peps batched N [--depth D]      (prefix-sliced form)
//   nqueens-peps prefix N C0 [C1 ...]       (completions of a placed prefix)
// Errors print "error: <message>" on stderr and exit with status 1.
#include <algorithm>
#include <cinttypes>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/peps_mi355x.h"

namespace {

using u128 = unsigned __int128;

struct Error {
  std::string msg;
};

u128 val(peps_u128 v) { return ((u128)v.hi << 64) | v.lo; }

std::string str(u128 v) {
  if (v == 0) return "0";
  std::string s;
  while (v) {
    s.push_back((char)('0' + (int)(v % 10)));
    v /= 10;
  }
  std::reverse(s.begin(), s.end());
  return s;
}

// Rust's usize::from_str: optional '+', then at least one decimal digit.
size_t parse_usize(const char* value, const std::string& label) {
  if (!value) throw Error{"missing " + label};
  const char* p = value;
  if (*p == '+') ++p;
  if (!*p) throw Error{"invalid " + label};
  size_t v = 0;
  for (; *p; ++p) {
    if (*p < '0' || *p > '9') throw Error{"invalid " + label};
    size_t d = (size_t)(*p - '0');
    if (v > (SIZE_MAX - d) / 10) throw Error{"invalid " + label};
    v = v * 10 + d;
  }
  return v;
}

void usage() {
  std::fprintf(stderr,
               "Usage:\n  nqueens-peps-naive solve N [--layers]\n  "
               "nqueens-peps-naive bench MAX_N [--min N] [--repeats R] [--csv]\n");
}

peps_contraction_result contract(size_t n) {
  if (n > 0xffffffffu) throw Error{"the packed u128 virtual-boundary backend supports N <= 42"};
  peps_contraction_result r;
  char err[1024] = {0};
  peps_options o;
  std::memset(&o, 0, sizeof(o));
  o.collect_layers = 1;
  if (peps_contract_rows((uint32_t)n, &o, &r, err, sizeof(err)) != PEPS_OK) throw Error{err};
  return r;
}

bool known(size_t n, u128* out) {
  peps_u128 v;
  if (n > 0xffffffffu || peps_known_count((uint32_t)n, &v) != PEPS_OK) return false;
  *out = val(v);
  return true;
}

// src/main.rs:25-70
void solve(int argc, char** argv, int i) {
  size_t n = parse_usize(i < argc ? argv[i] : nullptr, "N");
  bool layers = false;
  for (int k = i + 1; k < argc; ++k)
    if (std::strcmp(argv[k], "--layers") == 0) layers = true;
  peps_contraction_result r = contract(n);
  u128 expected;
  const bool has = known(n, &expected);
  const bool verified = has && expected == val(r.count);
  std::printf(
      "N=%zu Q(N)=%s elapsed_s=%.6f peak_rss_bytes=%" PRIu64 " peak_states=%" PRIu64
      " tensor_entries_examined=%s tensor_entries_matched=%s row_operator_candidates=%s "
      "row_operator_matched=%s verified=%s\n",
      n, str(val(r.count)).c_str(), r.elapsed_s, r.peak_rss_bytes, r.peak_states,
      str(val(r.tensor_entries_examined)).c_str(), str(val(r.tensor_entries_matched)).c_str(),
      str(val(r.row_operator_candidates)).c_str(), str(val(r.row_operator_matched)).c_str(),
      verified ? "true" : "false");
  if (layers) {
    std::printf(
        "row,input_states,tensor_entries_examined,tensor_entries_matched,row_operator_candidates,"
        "row_operator_matched,completed_row_terms,output_states,output_weight,elapsed_s,"
        "peak_rss_bytes\n");
    for (uint32_t k = 0; k < r.layer_count; ++k) {
      const peps_layer_metric& l = r.layers[k];
      std::printf("%u,%" PRIu64 ",%s,%s,%s,%s,%s,%" PRIu64 ",%s,%.6f,%" PRIu64 "\n", l.row + 1,
                  l.input_states, str(val(l.tensor_entries_examined)).c_str(),
                  str(val(l.tensor_entries_matched)).c_str(),
                  str(val(l.row_operator_candidates)).c_str(),
                  str(val(l.row_operator_matched)).c_str(), str(val(l.completed_row_terms)).c_str(),
                  l.output_states, str(val(l.output_weight)).c_str(), l.elapsed_s, l.peak_rss_bytes);
    }
  }
  if (has && !verified) throw Error{"known-count verification failed for N=" + std::to_string(n)};
}

// src/main.rs:72-152
void bench(int argc, char** argv, int i) {
  size_t max_n = parse_usize(i < argc ? argv[i] : nullptr, "MAX_N");
  size_t min_n = 1, repeats = 3;
  bool csv = false;
  for (int k = i + 1; k < argc;) {
    std::string a = argv[k];
    if (a == "--min") {
      min_n = parse_usize(k + 1 < argc ? argv[k + 1] : nullptr, "--min value");
      k += 2;
    } else if (a == "--repeats") {
      repeats = parse_usize(k + 1 < argc ? argv[k + 1] : nullptr, "--repeats value");
      k += 2;
    } else if (a == "--csv") {
      csv = true;
      k += 1;
    } else {
      throw Error{"unknown option: " + a};
    }
  }
  if (repeats == 0) throw Error{"--repeats must be positive"};
  if (min_n > max_n) throw Error{"--min must not exceed MAX_N"};
  if (csv)
    std::printf(
        "N,count,known_count,verified,median_elapsed_s,min_elapsed_s,peak_rss_bytes,peak_states,"
        "tensor_entries_examined,tensor_entries_matched,row_operator_candidates,"
        "row_operator_matched,repeats\n");
  for (size_t n = min_n; n <= max_n; ++n) {
    std::vector<double> el;
    peps_contraction_result last;
    for (size_t r = 0; r < repeats; ++r) {
      last = contract(n);
      el.push_back(last.elapsed_s);
    }
    const double minimum = *std::min_element(el.begin(), el.end());
    std::sort(el.begin(), el.end());
    const double med = el[el.size() / 2];
    u128 k;
    const bool has = known(n, &k);
    const bool verified = has && k == val(last.count);
    if (has && !verified) throw Error{"known-count verification failed for N=" + std::to_string(n)};
    if (csv) {
      std::printf("%zu,%s,%s,%s,%.9f,%.9f,%" PRIu64 ",%" PRIu64 ",%s,%s,%s,%s,%zu\n", n,
                  str(val(last.count)).c_str(), has ? str(k).c_str() : "",
                  verified ? "true" : "false", med, minimum, last.peak_rss_bytes, last.peak_states,
                  str(val(last.tensor_entries_examined)).c_str(),
                  str(val(last.tensor_entries_matched)).c_str(),
                  str(val(last.row_operator_candidates)).c_str(),
                  str(val(last.row_operator_matched)).c_str(), repeats);
    } else {
      std::printf("N=%2zu Q(N)=%-10s median=%.6fs peak_rss=%.1fMiB peak_states=%-8" PRIu64
                  " verified=%s\n",
                  n, str(val(last.count)).c_str(), med,
                  (double)last.peak_rss_bytes / (1024.0 * 1024.0), last.peak_states,
                  verified ? "true" : "false");
    }
    std::fflush(stdout);
  }
}

void progress(void*, uint64_t done, uint64_t total) {
  std::fprintf(stderr, "batched: %llu/%llu tasks\n", (unsigned long long)done, (unsigned long long)total);
  std::fflush(stderr);
}

void batched(int argc, char** argv, int i) {
  peps_set_progress(progress, nullptr);
  size_t n = parse_usize(i < argc ? argv[i] : nullptr, "N");
  size_t depth = 2;
  for (int k = i + 1; k < argc; ++k)
    if (std::strcmp(argv[k], "--depth") == 0) depth = parse_usize(k + 1 < argc ? argv[k + 1] : nullptr, "--depth value");
  peps_contraction_result r;
  peps_options o;
  std::memset(&o, 0, sizeof(o));
  uint32_t tasks = 0;
  char err[1024] = {0};
  if (peps_count_batched((uint32_t)n, (uint32_t)depth, &o, &r, &tasks, nullptr, 0, err, sizeof(err)) != PEPS_OK)
    throw Error{err};
  u128 k;
  const bool has = known(n, &k);
  std::printf("N=%zu Q(N)=%s tasks=%u elapsed_s=%.6f kernel_s=%.6f peak_states=%" PRIu64
              " row_operator_matched=%s verified=%s\n",
              n, str(val(r.count)).c_str(), tasks, r.elapsed_s, r.kernel_s, r.peak_states,
              str(val(r.row_operator_matched)).c_str(),
              has && k == val(r.count) ? "true" : "false");
  if (has && k != val(r.count)) throw Error{"known-count verification failed for N=" + std::to_string(n)};
}

void prefix(int argc, char** argv, int i) {
  size_t n = parse_usize(i < argc ? argv[i] : nullptr, "N");
  std::vector<uint32_t> cols;
  for (int k = i + 1; k < argc; ++k) cols.push_back((uint32_t)parse_usize(argv[k], "column"));
  peps_u128 c;
  peps_options o;
  std::memset(&o, 0, sizeof(o));
  char err[1024] = {0};
  if (peps_count_prefix((uint32_t)n, cols.data(), (uint32_t)cols.size(), &o, &c, err, sizeof(err)) != PEPS_OK)
    throw Error{err};
  std::printf("N=%zu prefix_len=%zu count=%s\n", n, cols.size(), str(val(c)).c_str());
}

}  // namespace

int main(int argc, char** argv) {
  try {
    if (argc < 2) {
      usage();
      throw Error{"missing or invalid command"};
    }
    std::string cmd = argv[1];
    if (cmd == "solve")
      solve(argc, argv, 2);
    else if (cmd == "bench")
      bench(argc, argv, 2);
    else if (cmd == "batched")
      batched(argc, argv, 2);
    else if (cmd == "prefix")
      prefix(argc, argv, 2);
    else {
      usage();
      throw Error{"missing or invalid command"};
    }
  } catch (const Error& e) {
    std::fprintf(stderr, "error: %s\n", e.msg.c_str());
    return 1;
  }
  return 0;
}
