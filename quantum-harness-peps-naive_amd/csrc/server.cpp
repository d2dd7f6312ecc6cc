// server.cpp -- the request server of the C ABI (peps_server_*).
//
// Requests are (N, prefix) counts.  Each is split into prefix tasks at a
// common start row D = max(k, min(2, N-1)) so that requests of one N (and one
// D) can share a launch.  Within a launch every request is one slice: the
// tasks of a request share its slice id, so their frontiers merge exactly as
// in that request's own count, while different requests never merge, and
// every request's final sum comes back separately.  One worker thread per
// server picks a class, fills a launch from every active request of that
// class, runs it on the engine (whose frontiers stay in HBM), credits the
// slice counts to their requests and repeats; new requests join at the next
// launch of their class.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/peps_mi355x.h"
#include "frontend.h"
#include "split_retry.h"

using peps::fail;
using peps::u128;

namespace {

struct Request {
  uint64_t id = 0;
  unsigned n = 0, k = 0, depth = 0;
  std::vector<peps::Task> tasks;
  std::vector<double> proxy;
  size_t next = 0;       // first task not yet launched
  size_t finished = 0;   // tasks whose counts are in
  double remaining = 0;  // proxy sum of the tasks not yet launched
  u128 count = 0;
  int status = PEPS_E_PENDING;
  std::string error;
  double t_submit = 0, t_start = -1, t_finish = -1;
  uint32_t launches = 0, max_batch = 0;
};

// Tasks [a, b) of a launch as one row pipeline with one slice per owning
// request (slice ids in order of first appearance); seeds of one request
// that reach the same boundary are merged.  *slice_owner[s] is slice s's
// request.
peps::RunSpec request_slices_spec(unsigned n, unsigned depth, const std::vector<peps::Task>& tasks,
                                  const std::vector<Request*>& owner, size_t a, size_t b,
                                  std::vector<Request*>* slice_owner) {
  peps::RunSpec spec;
  spec.n = n;
  spec.first_row = depth;
  spec.orbit_first_row = depth == 0;
  spec.per_slice = true;
  slice_owner->clear();
  std::unordered_map<Request*, uint64_t> sid;
  std::unordered_map<uint64_t, size_t> at;  // key | slice << 3n -> seed index
  for (size_t i = a; i < b; ++i) {
    auto ins = sid.emplace(owner[i], slice_owner->size());
    if (ins.second) slice_owner->push_back(owner[i]);
    const uint64_t key = tasks[i].key | (ins.first->second << (3 * n));
    auto seed = at.emplace(key, spec.seeds.size());
    if (seed.second)
      spec.seeds.push_back(peps::Seed{key, tasks[i].weight});
    else
      spec.seeds[seed.first->second].weight += tasks[i].weight;
  }
  unsigned bits = 0;
  while (((size_t)1 << bits) < slice_owner->size()) ++bits;
  spec.slice_bits = bits;
  return spec;
}

}  // namespace

struct peps_server {
  peps_options opts{};
  uint32_t policy = PEPS_POLICY_SHORTEST_FIRST;
  uint32_t max_pending = 0;
  std::chrono::steady_clock::time_point t0;

  std::mutex mu;
  std::condition_variable work_cv, done_cv;
  std::map<uint64_t, Request> requests;  // node-based: stable addresses
  std::deque<uint64_t> runnable;         // requests with unlaunched tasks, in submit order
  std::map<unsigned, peps::Resolved> resolved;  // per N
  peps::Projection projection;
  uint64_t next_id = 1;
  uint64_t pending = 0, submitted = 0, completed = 0, failed = 0, launches = 0, tasks_run = 0, retries = 0;
  double busy_s = 0;
  double sec_per_work = 0;  // learned launch seconds per unit of projected work
  bool stop = false;
  std::thread worker;

  double now() const {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }

  // Called with `mu` held.
  void finish(Request& r, int status, const std::string& error) {
    r.status = status;
    r.error = error;
    r.t_finish = now();
    --pending;
    if (status == PEPS_OK)
      ++completed;
    else
      ++failed;
    done_cv.notify_all();
  }

  // The request whose class is served next (mu held, runnable non-empty):
  // the oldest (FIFO), the one with the least projected remaining work
  // (shortest first), or the one with the highest response ratio
  // (wait + service) / service, service = projected work at the learned
  // seconds per unit of work -- short requests still go first, and a long
  // one's priority grows with its wait, so it cannot starve.
  Request* leader() {
    Request* best = nullptr;
    double best_key = 0;
    const double t = now();
    for (uint64_t id : runnable) {
      Request& r = requests.at(id);
      if (policy == PEPS_POLICY_FIFO) return &r;
      const double ratio = projection.ratio(r.n, r.depth);
      const double work = r.remaining * (ratio > 0 ? ratio : 1.0);
      double key = work;
      if (policy == PEPS_POLICY_RESPONSE_RATIO && sec_per_work > 0) {
        const double svc = std::max(work * sec_per_work, 1e-4);
        key = -(t - r.t_submit + svc) / svc;
      }
      if (!best || key < best_key) {
        best = &r;
        best_key = key;
      }
    }
    return best;
  }

  void run_worker() {
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
      work_cv.wait(lk, [&] { return stop || !runnable.empty(); });
      if (runnable.empty()) break;  // stop requested and everything launched
      Request* lead = leader();
      const unsigned n = lead->n, depth = lead->depth;
      const peps::Resolved& res = resolved.at(n);
      // Fill one launch: the leader's tasks first, then the other requests of
      // the class in submit order, while the projected candidates fit.
      std::string e;
      peps::Engine* eng = peps::Engine::get((int)res.device, &e);
      const double ratio = projection.ratio(n, depth);
      const double buffer = eng ? (double)eng->candidate_capacity((int)res.primes.size(), res.records) : 0.0;
      const double cap = peps::kAdmitFraction * buffer;
      // The leader's own tasks form one merged frontier, as in its plain
      // count, which the engine runs up to the whole candidate buffer (its
      // partition levels use exact offsets where the fixed-capacity margin
      // is missing).  When all of its remaining tasks are projected to fit
      // kLeadFraction of the buffer they go into this launch together, since
      // splitting them would stop them merging; a request too large for one
      // launch is packed within the margin, like the requests that join it.
      constexpr double kLeadFraction = 0.8;
      const double lead_rest = ratio > 0 ? lead->remaining * ratio : 0.0;
      const double lead_cap = ratio > 0 && lead_rest <= kLeadFraction * buffer ? buffer : cap;
      const size_t slices = peps::max_slices(n, res.slice_group);
      const size_t limit = ratio > 0 ? slices : std::min<size_t>(slices, 8);
      std::vector<peps::Task> batch;
      std::vector<Request*> owner;
      std::vector<double> bproxy;
      double load = 0;
      std::vector<Request*> members{lead};
      for (uint64_t id : runnable) {
        Request& r = requests.at(id);
        if (&r != lead && r.n == n && r.depth == depth) members.push_back(&r);
      }
      for (Request* r : members) {
        const double admit = r == lead ? lead_cap : cap;
        while (r->next < r->tasks.size() && batch.size() < limit) {
          const double add = r->proxy[r->next] * ratio;
          if (!batch.empty() && ratio > 0 && admit > 0 && load + add > admit) break;
          load += add;
          batch.push_back(r->tasks[r->next]);
          owner.push_back(r);
          bproxy.push_back(r->proxy[r->next]);
          r->remaining -= r->proxy[r->next];
          ++r->next;
        }
        if (batch.size() >= limit || (ratio > 0 && cap > 0 && load >= cap)) break;
      }
      runnable.erase(std::remove_if(runnable.begin(), runnable.end(),
                                    [&](uint64_t id) {
                                      const Request& r = requests.at(id);
                                      return r.next >= r.tasks.size();
                                    }),
                     runnable.end());
      std::vector<Request*> distinct = owner;
      std::sort(distinct.begin(), distinct.end());
      distinct.erase(std::unique(distinct.begin(), distinct.end()), distinct.end());
      const double t_launch = now();
      for (Request* r : distinct) {
        if (r->t_start < 0) r->t_start = t_launch;
        r->max_batch = std::max<uint32_t>(r->max_batch, (uint32_t)distinct.size());
      }
      lk.unlock();

      // Per-request counts of each sub-launch, credited after the lock is
      // retaken.  A launch that does not fit the HBM budget is split in two
      // and retried.
      std::vector<std::pair<Request*, u128>> credit;
      std::vector<std::pair<size_t, size_t>> ranges;
      std::vector<peps::RunOut> outs;
      int rc = PEPS_OK;
      {
        auto run = [&](size_t a, size_t b) -> int {
          std::vector<Request*> slice_owner;
          peps::RunSpec spec = request_slices_spec(n, depth, batch, owner, a, b, &slice_owner);
          peps::RunOut ro;
          std::string se;
          const int code = peps::run_engine(res, spec, &ro, &se);
          if (code != PEPS_OK) {
            e = se;
            return code;
          }
          for (size_t sl = 0; sl < slice_owner.size(); ++sl) credit.emplace_back(slice_owner[sl], ro.slice_counts[sl]);
          ranges.emplace_back(a, b);
          ro.slice_counts.clear();
          outs.push_back(std::move(ro));
          return PEPS_OK;
        };
        // The projection under-estimated this class: learn from the failure
        // so the next launches are admitted smaller.
        auto exceeded = [&](size_t a, size_t b) {
          double psum = 0;
          for (size_t i = a; i < b; ++i) psum += bproxy[i];
          std::lock_guard<std::mutex> g(mu);
          projection.exceeded(n, depth, psum, cap / peps::kAdmitFraction);
          ++retries;
        };
        rc = peps::split_retry(0, batch.size(), run, exceeded);
      }
      const double t_end = now();

      lk.lock();
      busy_s += t_end - t_launch;
      if (rc == PEPS_OK && ratio > 0) {
        double work = 0;
        for (double x : bproxy) work += x * ratio;
        if (work > 0) {
          const double spw = (t_end - t_launch) / work;
          sec_per_work = sec_per_work > 0 ? 0.8 * sec_per_work + 0.2 * spw : spw;
        }
      }
      launches += ranges.size();
      tasks_run += batch.size();
      for (size_t j = 0; j < ranges.size(); ++j) {
        double psum = 0;
        for (size_t i = ranges[j].first; i < ranges[j].second; ++i) psum += bproxy[i];
        projection.learn(n, depth, psum, outs[j]);
        std::vector<Request*> seen;
        for (size_t i = ranges[j].first; i < ranges[j].second; ++i)
          if (std::find(seen.begin(), seen.end(), owner[i]) == seen.end()) seen.push_back(owner[i]);
        for (Request* r : seen) ++r->launches;
      }
      if (rc != PEPS_OK) {
        for (Request* r : distinct) {
          if (r->status != PEPS_E_PENDING) continue;
          runnable.erase(std::remove(runnable.begin(), runnable.end(), r->id), runnable.end());
          finish(*r, rc, e);
        }
        continue;
      }
      for (const auto& [r, c] : credit) {
        if (r->status != PEPS_E_PENDING) continue;
        if (!peps::checked_add(r->count, c, &r->count)) {
          runnable.erase(std::remove(runnable.begin(), runnable.end(), r->id), runnable.end());
          finish(*r, PEPS_E_OVERFLOW, "count overflow");
        }
      }
      for (size_t i = 0; i < batch.size(); ++i) {
        Request* r = owner[i];
        if (r->status != PEPS_E_PENDING) continue;
        if (++r->finished == r->tasks.size()) {
          u128 fact;
          peps::factorial(r->n, &fact);
          if (r->count > fact)
            finish(*r, PEPS_E_OVERFLOW, "count exceeds N!");
          else
            finish(*r, PEPS_OK, "");
        }
      }
    }
  }

  void fill(const Request& r, peps_request_result* out) const {
    std::memset(out, 0, sizeof(*out));
    out->ticket = r.id;
    out->n = r.n;
    out->prefix_len = r.k;
    out->status = r.status;
    out->tasks = (uint32_t)r.tasks.size();
    out->launches = r.launches;
    out->max_batch = r.max_batch;
    out->count = peps_u128{(uint64_t)r.count, (uint64_t)(r.count >> 64)};
    out->submit_s = r.t_submit;
    out->start_s = r.t_start;
    out->finish_s = r.t_finish;
    const size_t m = std::min(sizeof(out->error) - 1, r.error.size());
    std::memcpy(out->error, r.error.data(), m);
  }
};

extern "C" {

int peps_server_create(const peps_options* opts, uint32_t policy, uint32_t max_pending,
                       peps_server** out, char* err, size_t err_cap) {
  if (!out) return fail(PEPS_E_INVALID, "null server pointer", err, err_cap);
  *out = nullptr;
  if (policy > PEPS_POLICY_RESPONSE_RATIO) return fail(PEPS_E_INVALID, "unknown scheduling policy", err, err_cap);
  std::string e;
  peps::Resolved probe;
  peps_options o{};
  if (opts) o = *opts;
  o.crt_lanes = 0;  // the lane plan depends on N: checked per request
  int rc = peps::resolve_options(1, &o, &probe, &e);
  if (rc) return fail(rc, e, err, err_cap);
  if (!peps::Engine::get((int)probe.device, &e)) return fail(PEPS_E_DEVICE, e, err, err_cap);
  std::unique_ptr<peps_server> s(new peps_server());
  if (opts) s->opts = *opts;
  s->policy = policy;
  s->max_pending = max_pending;
  s->t0 = std::chrono::steady_clock::now();
  peps_server* raw = s.get();
  s->worker = std::thread([raw] { raw->run_worker(); });
  *out = s.release();
  return PEPS_OK;
}

int peps_server_submit(peps_server* s, uint32_t n, const uint32_t* cols, uint32_t k,
                       uint64_t* ticket, char* err, size_t err_cap) {
  if (!s || !ticket) return fail(PEPS_E_INVALID, "null server or ticket pointer", err, err_cap);
  std::string e;
  if (k > n) return fail(PEPS_E_INVALID, "prefix longer than the board", err, err_cap);
  if (k && !cols) return fail(PEPS_E_INVALID, "null prefix", err, err_cap);
  for (uint32_t i = 0; i < k; ++i)
    if (cols[i] >= n) return fail(PEPS_E_INVALID, "prefix column outside the board", err, err_cap);
  int rc = peps::check_board(n, &e);
  if (rc) return fail(rc, e, err, err_cap);
  peps::Resolved res;
  rc = peps::resolve_options(n, &s->opts, &res, &e);
  if (rc) return fail(rc, e, err, err_cap);

  Request r;
  r.n = n;
  r.k = k;
  uint64_t key = 0;
  bool trivial = false;
  if (n == 0) {  // src/lib.rs:5295-5307
    r.count = 1;
    trivial = true;
  } else if (!peps::prefix_state(n, cols, k, &key)) {
    trivial = true;  // an attacking prefix has no completions
  } else if (k == n) {
    r.count = 1;
    trivial = true;
  } else {
    r.depth = std::max(k, std::min(2u, n - 1));
    r.tasks = k == 0 ? peps::make_tasks(n, r.depth)
                     : peps::expand_task(n, peps::Task{key, 1}, k, r.depth);
    trivial = r.tasks.empty();
    for (const peps::Task& t : r.tasks) {
      r.proxy.push_back(s->projection.proxy(n, r.depth, t.key));
      r.remaining += r.proxy.back();
    }
  }

  std::lock_guard<std::mutex> lk(s->mu);
  if (s->stop) return fail(PEPS_E_INVALID, "server is shutting down", err, err_cap);
  if (s->max_pending && s->pending >= s->max_pending)
    return fail(PEPS_E_BUSY, "request queue is full", err, err_cap);
  r.id = s->next_id++;
  r.t_submit = s->now();
  *ticket = r.id;
  ++s->submitted;
  ++s->pending;
  s->resolved.emplace(n, res);
  Request& stored = s->requests.emplace(r.id, std::move(r)).first->second;
  if (trivial) {
    stored.t_start = stored.t_submit;
    s->finish(stored, PEPS_OK, "");
  } else {
    s->runnable.push_back(stored.id);
    s->work_cv.notify_one();
  }
  return PEPS_OK;
}

int peps_server_poll(peps_server* s, uint64_t ticket, peps_request_result* out) {
  if (!s || !out) return PEPS_E_INVALID;
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->requests.find(ticket);
  if (it == s->requests.end()) return PEPS_E_INVALID;
  s->fill(it->second, out);
  return it->second.status;
}

int peps_server_wait(peps_server* s, uint64_t ticket, double timeout_s, peps_request_result* out) {
  if (!s || !out) return PEPS_E_INVALID;
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->requests.find(ticket);
  if (it == s->requests.end()) return PEPS_E_INVALID;
  Request& r = it->second;
  auto ready = [&] { return r.status != PEPS_E_PENDING; };
  if (timeout_s < 0)
    s->done_cv.wait(lk, ready);
  else
    s->done_cv.wait_for(lk, std::chrono::duration<double>(timeout_s), ready);
  s->fill(r, out);
  return r.status;
}

int peps_server_stats_get(peps_server* s, peps_server_stats* out) {
  if (!s || !out) return PEPS_E_INVALID;
  std::lock_guard<std::mutex> lk(s->mu);
  out->submitted = s->submitted;
  out->completed = s->completed;
  out->failed = s->failed;
  out->pending = s->pending;
  out->launches = s->launches;
  out->tasks = s->tasks_run;
  out->retries = s->retries;
  out->busy_s = s->busy_s;
  out->uptime_s = s->now();
  return PEPS_OK;
}

void peps_server_destroy(peps_server* s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->work_cv.notify_all();
  if (s->worker.joinable()) s->worker.join();
  delete s;
}

}  // extern "C"
