// layer_census.cpp -- TEST INFRASTRUCTURE ONLY (part of the CPU oracle).
//
// Two independent CPU routes to the per-row figures of the reference's
// default row loop (contract_rows_sort_reduce_with_modes, src/lib.rs:5287-5415,
// counters :5370-5390) at board sizes where the sort-reduce restatement
// (peps_oracle.cpp) needs more memory than a test can spend:
//
//  * depth_counts: the number of non-attacking r-row prefixes for every r,
//    each prefix weighted by the reference's row-0 vertical-reflection orbit
//    weight (src/lib.rs:721-749) when orbits are on.  Every row's
//    `output_weight` is the sum of the weights of its merged states, i.e. of
//    all the (r + 1)-row prefixes, and the last one is the count.  A
//    bitboard DFS (the search of src/dfs_bitmask.rs:67-109 with a per-depth
//    tally) counts them.
//
//  * frontier_census: the reference's per-row state counts
//    (`input_states`, `row_operator_matched`, `completed_row_terms`,
//    `output_states`) without weights, by a breadth-first pass over the
//    boundaries grouped by their column set.  A boundary after row r is
//    (columns, diag_dr, diag_dl) with popcount(columns) = r + 1; its
//    successors through the compiled operator (src/lib.rs:594-638) have
//    column set columns | {c}.  So the successors with column set T come
//    from the groups T - {c}, c in T, and are merged within their group
//    (sort and unique), one group at a time: memory is two layers of 4-byte
//    keys (the low word of diag_dr | diag_dl << n, filed by its high bits)
//    and the merges are small.
//
// Both are pinned by the tests against the published per-layer supports and
// totals (benchmarks/e35_frontier_audit_layers.csv, e12, e32/e36, n14_layers)
// and against the sort-reduce restatement where that runs.
#include <omp.h>

#include <algorithm>
#include <atomic>
#include <cstring>
#include <exception>
#include <string>
#include <thread>
#include <vector>

#include "peps_oracle.hpp"

namespace oracle {

namespace {

// Nodes below (cols, dl, dr) at depths depth+1 .. n, unweighted, added into t.
void tally_dfs(uint64_t mask, uint64_t cols, uint64_t dl, uint64_t dr, unsigned depth, unsigned n,
               uint64_t* t) {
  uint64_t avail = mask & ~(cols | dl | dr);
  t[depth + 1] += (uint64_t)__builtin_popcountll(avail);
  if (depth + 1 == n) return;
  while (avail) {
    const uint64_t b = avail & (~avail + 1);
    avail ^= b;
    tally_dfs(mask, cols | b, ((dl | b) << 1) & mask, (dr | b) >> 1, depth + 1, n, t);
  }
}

struct Seed {
  uint64_t cols, dl, dr;
  unsigned depth;
  u128 weight;
};

}  // namespace

namespace {

// Adds to out[d] the weighted nodes at depths seed.depth + 1 .. n below
// every seed (out[0 .. seed depth] are the caller's), splitting the seeds a
// few rows deeper so every thread has work.
void tally_seeds(unsigned n, std::vector<Seed> seeds, unsigned threads, std::vector<u128>* out) {
  const uint64_t mask = ((uint64_t)1 << n) - 1;
  const size_t target = (size_t)std::max(1u, threads) * 64;
  while (seeds.size() < target && !seeds.empty() && seeds[0].depth + 1 < n) {
    std::vector<Seed> next;
    for (const Seed& s : seeds) {
      uint64_t avail = mask & ~(s.cols | s.dl | s.dr);
      while (avail) {
        const uint64_t b = avail & (~avail + 1);
        avail ^= b;
        next.push_back(Seed{s.cols | b, ((s.dl | b) << 1) & mask, (s.dr | b) >> 1, s.depth + 1, s.weight});
        (*out)[s.depth + 1] += s.weight;
      }
    }
    seeds.swap(next);
  }
  if (seeds.empty() || seeds[0].depth == n) return;
  const unsigned workers = std::max(1u, std::min<unsigned>(threads, (unsigned)seeds.size()));
  std::vector<std::vector<u128>> part(workers, std::vector<u128>(n + 1, 0));
  std::atomic<size_t> next{0};
  auto run = [&](unsigned w) {
    std::vector<uint64_t> t(n + 2);
    for (;;) {
      const size_t i = next.fetch_add(1);
      if (i >= seeds.size()) break;
      const Seed& s = seeds[i];
      std::fill(t.begin(), t.end(), 0);
      tally_dfs(mask, s.cols, s.dl, s.dr, s.depth, n, t.data());
      for (unsigned d = s.depth + 1; d <= n; ++d) part[w][d] += (u128)t[d] * s.weight;
    }
  };
  std::vector<std::thread> pool;
  for (unsigned w = 1; w < workers; ++w) pool.emplace_back(run, w);
  run(0);
  for (auto& th : pool) th.join();
  for (unsigned w = 0; w < workers; ++w)
    for (unsigned d = 0; d <= n; ++d) (*out)[d] += part[w][d];
}

}  // namespace

std::vector<u128> depth_counts(unsigned n, bool orbits, unsigned threads) {
  if (n > 27) throw OracleError("depth counts support N <= 27");
  std::vector<u128> out(n + 1, 0);
  out[0] = 1;
  if (n == 0) return out;
  const uint64_t mask = ((uint64_t)1 << n) - 1;
  // Row 0: the orbit representatives c <= n - 1 - c, weight 2 for a mirror
  // pair and 1 for the centre of an odd board (src/lib.rs:721-732).
  std::vector<Seed> seeds;
  for (unsigned c = 0; c < n; ++c) {
    u128 w = 1;
    if (orbits) {
      if (c > n - 1 - c) continue;
      w = c == n - 1 - c ? 1 : 2;
    }
    const uint64_t b = (uint64_t)1 << c;
    seeds.push_back(Seed{b, (b << 1) & mask, b >> 1, 1, w});
    out[1] += w;
  }
  tally_seeds(n, std::move(seeds), threads, &out);
  return out;
}

// Completions of the prefix cols[0..k) by the same multi-threaded DFS (the
// search of src/dfs_bitmask.rs:67-109 from the prefix's attack masks); 0 for
// an attacking prefix.
u128 count_prefix_parallel(unsigned n, const uint32_t* cols, unsigned k, unsigned threads) {
  if (n > 27) throw OracleError("prefix counts support N <= 27");
  if (k > n) throw OracleError("prefix longer than the board");
  const uint64_t mask = n ? ((uint64_t)1 << n) - 1 : 0;
  uint64_t c = 0, dl = 0, dr = 0;
  for (unsigned i = 0; i < k; ++i) {
    if (cols[i] >= n) throw OracleError("prefix column outside the board");
    const uint64_t b = (uint64_t)1 << cols[i];
    if ((c | dl | dr) & b) return 0;
    c |= b;
    dl = ((dl | b) << 1) & mask;
    dr = (dr | b) >> 1;
  }
  if (k == n) return 1;
  std::vector<u128> out(n + 1, 0);
  tally_seeds(n, {Seed{c, dl, dr, k, 1}}, threads, &out);
  return out[n];
}

namespace {

// Colex rank of a k-subset of {0..n-1} given as a bit mask: sum over its
// members p_0 < p_1 < ... of C(p_i, i + 1).  Masks of one popcount ordered by
// value have consecutive ranks.
struct Ranker {
  uint64_t binom[64][64];
  explicit Ranker(unsigned n) {
    std::memset(binom, 0, sizeof(binom));
    for (unsigned a = 0; a <= n; ++a) {
      binom[a][0] = 1;
      for (unsigned b = 1; b <= a; ++b) binom[a][b] = binom[a - 1][b - 1] + (b <= a - 1 ? binom[a - 1][b] : 0);
    }
  }
  uint64_t rank(uint64_t m) const {
    uint64_t r = 0;
    unsigned i = 0;
    while (m) {
      const unsigned p = (unsigned)__builtin_ctzll(m);
      m &= m - 1;
      ++i;
      r += binom[p][i];
    }
    return r;
  }
};

}  // namespace

// One column-set group of a layer: its keys (diag_dr | diag_dl << n, below
// 2^(2n) <= 2^42) sorted, stored as 32-bit low words in one array per value
// of the bits above bit 32, so a state costs 4 bytes (an N = 18 layer of
// 4.3e9 states fits in 17 GB).
struct Group {
  std::vector<std::vector<uint32_t>> by_high;
  template <typename F>
  void for_each(F&& f) const {
    for (size_t h = 0; h < by_high.size(); ++h)
      for (uint32_t lo : by_high[h]) f(((uint64_t)h << 32) | lo);
  }
  void assign(const std::vector<uint64_t>& sorted) {
    by_high.clear();
    for (uint64_t k : sorted) {
      const size_t h = (size_t)(k >> 32);
      if (by_high.size() <= h) by_high.resize(h + 1);
      by_high[h].push_back((uint32_t)k);
    }
    for (auto& v : by_high) v.shrink_to_fit();
  }
};

std::vector<CensusRow> frontier_census(unsigned n, bool orbits, unsigned threads) {
  if (n == 0 || n > 21) throw OracleError("frontier census supports 1 <= N <= 21");
  const uint64_t mask = ((uint64_t)1 << n) - 1;
  const Ranker rk(n);
  std::vector<CensusRow> rows;
  // Layer after row 0, grouped by column set (a single column c: rank c).
  std::vector<Group> cur(n);
  CensusRow r0{0, 1, n, 0, 0};
  for (unsigned c = 0; c < n; ++c) {
    if (orbits && c > n - 1 - c) continue;
    const uint64_t b = (uint64_t)1 << c;
    cur[c].assign({((b << 1) & mask) | ((b >> 1) << n)});
    r0.completed += 1;
    r0.output_states += 1;
  }
  rows.push_back(r0);
  uint64_t S = r0.output_states;
  omp_set_num_threads((int)std::max(1u, threads));
  for (unsigned row = 1; row < n; ++row) {
    const unsigned k = row + 1;  // columns set after this row
    // Child column sets in value order (Gosper's hack): index = colex rank.
    std::vector<uint64_t> masks;
    masks.reserve(rk.binom[n][k]);
    for (uint64_t m = ((uint64_t)1 << k) - 1; m <= mask;) {
      masks.push_back(m);
      const uint64_t c = m & (~m + 1), r = m + c;
      m = (((r ^ m) >> 2) / c) | r;
    }
    std::vector<Group> nxt(masks.size());
    uint64_t completed = 0, out_states = 0;
#pragma omp parallel reduction(+ : completed, out_states)
    {
      std::vector<uint64_t> buf;
#pragma omp for schedule(dynamic, 8)
      for (size_t i = 0; i < masks.size(); ++i) {
        const uint64_t T = masks[i];
        buf.clear();
        for (uint64_t rest = T; rest;) {
          const uint64_t b = rest & (~rest + 1);
          rest ^= b;
          cur[rk.rank(T ^ b)].for_each([&](uint64_t key) {
            const uint64_t dr = key & mask, dl = key >> n;
            if ((dr | dl) & b) return;
            buf.push_back((((dr | b) << 1) & mask) | (((dl | b) >> 1) << n));
          });
        }
        completed += buf.size();
        if (buf.empty()) continue;
        std::sort(buf.begin(), buf.end());
        buf.erase(std::unique(buf.begin(), buf.end()), buf.end());
        out_states += buf.size();
        nxt[i].assign(buf);
      }
    }
    rows.push_back(CensusRow{row, S, completed, completed, out_states});
    S = out_states;
    cur.swap(nxt);
  }
  return rows;
}

}  // namespace oracle

extern "C" {

int oracle_count_prefix_parallel(uint32_t n, const uint32_t* cols, uint32_t k, uint32_t threads,
                                 peps_u128* out, char* err, size_t cap) {
  try {
    const oracle::u128 v = oracle::count_prefix_parallel(n, cols, k, threads);
    *out = peps_u128{(uint64_t)v, (uint64_t)(v >> 64)};
    return 0;
  } catch (const std::exception& e) {
    if (err && cap) {
      std::strncpy(err, e.what(), cap - 1);
      err[cap - 1] = 0;
    }
    return 1;
  }
}

int oracle_depth_counts(uint32_t n, uint32_t orbits, uint32_t threads, peps_u128* out, char* err,
                        size_t cap) {
  try {
    const std::vector<oracle::u128> v = oracle::depth_counts(n, orbits != 0, threads);
    for (size_t d = 0; d < v.size(); ++d) out[d] = peps_u128{(uint64_t)v[d], (uint64_t)(v[d] >> 64)};
    return 0;
  } catch (const std::exception& e) {
    if (err && cap) {
      std::strncpy(err, e.what(), cap - 1);
      err[cap - 1] = 0;
    }
    return 1;
  }
}

// rows: n x 5 u64 -- row, input_states, row_operator_matched,
// completed_row_terms, output_states.
int oracle_frontier_census(uint32_t n, uint32_t orbits, uint32_t threads, uint64_t* rows, char* err,
                           size_t cap) {
  try {
    const std::vector<oracle::CensusRow> v = oracle::frontier_census(n, orbits != 0, threads);
    for (size_t i = 0; i < v.size(); ++i) {
      rows[5 * i + 0] = v[i].row;
      rows[5 * i + 1] = v[i].input_states;
      rows[5 * i + 2] = v[i].matched;
      rows[5 * i + 3] = v[i].completed;
      rows[5 * i + 4] = v[i].output_states;
    }
    return 0;
  } catch (const std::exception& e) {
    if (err && cap) {
      std::strncpy(err, e.what(), cap - 1);
      err[cap - 1] = 0;
    }
    return 1;
  }
}

}  // extern "C"
