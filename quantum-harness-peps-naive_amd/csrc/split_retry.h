// split_retry.h -- run a range of prefix tasks, halving it when it does not
// fit the HBM budget.
//
// The batched form and the request server launch task groups whose frontier
// size is projected, not known.  A group the arena cannot hold comes back as
// PEPS_E_MEMORY (returned only when the arena refuses an allocation); such a
// group is split in two and both halves are run.  Every other failure --
// including PEPS_E_INTERNAL, a consistency check of the device pipeline --
// ends the call with its own code, so a sizing bug is never hidden behind a
// smaller launch.  Host-only: no HIP types, so it is unit-tested on the CPU.
#pragma once

#include <cstddef>
#include <functional>
#include <utility>
#include <vector>

#include "../../include/peps_mi355x.h"

namespace peps {

// Runs run(a, b) over [g0, g1).  A PEPS_E_MEMORY result of a range of two or
// more tasks calls exceeded(a, b) (may be empty) and runs [a, mid) then
// [mid, b).  Returns PEPS_OK, or the first other failure (a single task's
// PEPS_E_MEMORY included).
inline int split_retry(size_t g0, size_t g1, const std::function<int(size_t, size_t)>& run,
                       const std::function<void(size_t, size_t)>& exceeded) {
  std::vector<std::pair<size_t, size_t>> work{{g0, g1}};
  while (!work.empty()) {
    const auto [a, b] = work.back();
    work.pop_back();
    const int rc = run(a, b);
    if (rc == PEPS_E_MEMORY && b - a > 1) {
      if (exceeded) exceeded(a, b);
      const size_t mid = a + (b - a) / 2;
      work.emplace_back(mid, b);
      work.emplace_back(a, mid);
      continue;
    }
    if (rc != PEPS_OK) return rc;
  }
  return PEPS_OK;
}

}  // namespace peps
