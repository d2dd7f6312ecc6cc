// Host-only check of split_retry (csrc/split_retry.h): a memory shortage
// splits the task range and retries the halves; any other failure, an
// internal consistency failure in particular, ends the call at once.
#include <cstdio>
#include <string>
#include <utility>
#include <vector>

#include "split_retry.h"

static int failures = 0;
#define CHECK(c)                                               \
  do {                                                         \
    if (!(c)) {                                                \
      std::printf("FAIL line %d: %s\n", __LINE__, #c);         \
      ++failures;                                              \
    }                                                          \
  } while (0)

int main() {
  using Calls = std::vector<std::pair<size_t, size_t>>;
  // 1. Ranges above 3 tasks do not fit: 8 tasks end as halves and quarters, in order.
  {
    Calls calls, done, split;
    auto run = [&](size_t a, size_t b) {
      calls.emplace_back(a, b);
      if (b - a > 3) return (int)PEPS_E_MEMORY;
      done.emplace_back(a, b);
      return (int)PEPS_OK;
    };
    auto exceeded = [&](size_t a, size_t b) { split.emplace_back(a, b); };
    CHECK(peps::split_retry(0, 8, run, exceeded) == PEPS_OK);
    CHECK((done == Calls{{0, 2}, {2, 4}, {4, 6}, {6, 8}}));
    CHECK((split == Calls{{0, 8}, {0, 4}, {4, 8}}));
    CHECK(calls.size() == 7);
  }
  // 2. An internal failure is returned as is: one call, no split.
  {
    Calls calls, split;
    auto run = [&](size_t a, size_t b) {
      calls.emplace_back(a, b);
      return (int)PEPS_E_INTERNAL;
    };
    auto exceeded = [&](size_t a, size_t b) { split.emplace_back(a, b); };
    CHECK(peps::split_retry(0, 8, run, exceeded) == PEPS_E_INTERNAL);
    CHECK(calls.size() == 1 && split.empty());
  }
  // 3. Device and overflow failures likewise.
  for (int code : {(int)PEPS_E_DEVICE, (int)PEPS_E_OVERFLOW, (int)PEPS_E_INVALID}) {
    int n = 0;
    CHECK(peps::split_retry(3, 9, [&](size_t, size_t) { ++n; return code; }, nullptr) == code);
    CHECK(n == 1);
  }
  // 4. A single task that does not fit fails with PEPS_E_MEMORY.
  {
    int n = 0;
    CHECK(peps::split_retry(5, 6, [&](size_t, size_t) { ++n; return (int)PEPS_E_MEMORY; }, nullptr) ==
          PEPS_E_MEMORY);
    CHECK(n == 1);
  }
  // 5. An internal failure after a split stops the remaining halves.
  {
    Calls calls;
    auto run = [&](size_t a, size_t b) {
      calls.emplace_back(a, b);
      if (b - a > 2) return (int)PEPS_E_MEMORY;
      return a == 0 ? (int)PEPS_E_INTERNAL : (int)PEPS_OK;
    };
    CHECK(peps::split_retry(0, 4, run, nullptr) == PEPS_E_INTERNAL);
    CHECK((calls == Calls{{0, 4}, {0, 2}}));
  }
  // 6. An empty range runs once (the caller decides what an empty group means).
  {
    int n = 0;
    CHECK(peps::split_retry(2, 2, [&](size_t, size_t) { ++n; return (int)PEPS_OK; }, nullptr) == PEPS_OK);
    CHECK(n == 1);
  }
  std::printf(failures ? "split_retry: %d failures\n" : "split_retry: ok\n", failures);
  return failures ? 1 : 0;
}
