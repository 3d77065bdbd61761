// Request server for incremental simplification (continuous batching).
//
// No counterpart in the reference, which simplifies one formula per call and
// frees its device state afterwards (src/gpu/simplify.cu:43-55).  Here each
// session is a context whose formula stays resident in device memory; a
// request appends clauses to it, replaces its assumed variables and
// simplifies it (pfs_append, pfs_assume, pfs_simplify).  Requests carry a
// release time; workers (one host thread each, every context on its own
// stream) take the earliest released request whose session is idle and
// whose session fits the device-memory budget next to the sessions that are
// running, so requests that arrive while others run are admitted as workers
// and memory free up, and their device work overlaps.  A session runs its
// requests one at a time in submission order, so every result equals the
// one the session would get alone.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/pfsimp.h"

namespace {

struct Request {
  pfs_ctx* ctx = nullptr;
  double release_ms = 0;
  uint32_t maxvar = 0;
  std::vector<uint64_t> off;
  std::vector<uint32_t> lits, assumed;
  // results
  bool taken = false, done = false;
  int32_t result = PFS_RESULT_UNSOLVED;
  pfs_status status = PFS_OK;
  double start_ms = 0, end_ms = 0;
};

}  // namespace

struct pfs_server {
  uint32_t nworkers = 1;
  uint64_t budget = 0;
  std::vector<std::unique_ptr<Request>> reqs;
  std::unordered_map<pfs_ctx*, bool> busy;      // session running a request
  std::unordered_map<pfs_ctx*, uint64_t> used;  // device bytes of running sessions
  uint64_t running_bytes = 0;
  size_t first_open = 0;  // every request before it is taken
  std::mutex m;
  std::condition_variable cv;
  std::vector<std::thread> workers;
  bool started = false, stopping = false;
  std::chrono::steady_clock::time_point t0;

  double now_ms() const {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  }
  static uint64_t ctx_bytes(pfs_ctx* c) {
    uint64_t u = 0, p = 0;
    pfs_memory(c, &u, &p);
    return u;
  }
  // the earliest released request whose session is idle, is the session's
  // oldest open request, and fits the memory budget; -1 if none.  *wake is
  // set to the next release time in the future.
  long pick(double now, double* wake) {
    long best = -1;
    *wake = -1;
    std::unordered_map<pfs_ctx*, bool> seen;  // sessions with an older open request
    while (first_open < reqs.size() && reqs[first_open]->taken) first_open++;
    for (size_t i = first_open; i < reqs.size(); i++) {
      Request& r = *reqs[i];
      if (r.taken) continue;
      const bool older = seen.count(r.ctx) != 0;
      seen[r.ctx] = true;
      if (older || busy[r.ctx]) continue;
      if (r.release_ms > now) {
        if (*wake < 0 || r.release_ms < *wake) *wake = r.release_ms;
        continue;
      }
      if (budget && running_bytes) {
        // growth per new pool word: the two clause blocks, and the two words
        // of the occurrence-table build's (literal, handle) pairs
        const uint64_t need = ctx_bytes(r.ctx) + 16 * (r.lits.size() + 3 * r.off.size());
        if (running_bytes + need > budget) continue;
      }
      if (best < 0 || r.release_ms < reqs[best]->release_ms) best = (long)i;
    }
    return best;
  }
  void work() {
    std::unique_lock<std::mutex> lk(m);
    while (true) {
      if (stopping) return;
      double wake;
      const long i = pick(now_ms(), &wake);
      if (i < 0) {
        if (wake >= 0)
          cv.wait_until(lk, t0 + std::chrono::duration_cast<std::chrono::steady_clock::duration>(
                                     std::chrono::duration<double, std::milli>(wake)));
        else
          cv.wait(lk);
        continue;
      }
      Request& r = *reqs[i];
      r.taken = true;
      busy[r.ctx] = true;
      const uint64_t b = ctx_bytes(r.ctx);
      used[r.ctx] = b;
      running_bytes += b;
      lk.unlock();
      r.start_ms = now_ms();
      pfs_status st = PFS_OK;
      if (!r.off.empty() && r.off.size() > 1)
        st = pfs_append(r.ctx, r.maxvar, r.off.size() - 1, r.off.data(), r.lits.data());
      else if (r.maxvar)
        st = pfs_append(r.ctx, r.maxvar, 0, nullptr, nullptr);
      if (st == PFS_OK) st = pfs_assume(r.ctx, r.assumed.data(), (uint32_t)r.assumed.size());
      if (st == PFS_OK) st = pfs_simplify(r.ctx, &r.result);
      r.end_ms = now_ms();
      lk.lock();
      r.status = st;
      r.done = true;
      busy[r.ctx] = false;
      running_bytes -= used[r.ctx];
      used[r.ctx] = 0;
      cv.notify_all();
    }
  }
};

extern "C" {

pfs_status pfs_server_create(uint32_t workers, uint64_t mem_budget, pfs_server** out) {
  if (!out) return PFS_BAD_INPUT;
  pfs_server* s = new pfs_server();
  s->nworkers = std::max<uint32_t>(1, workers);
  s->budget = mem_budget;
  s->t0 = std::chrono::steady_clock::now();
  *out = s;
  return PFS_OK;
}

pfs_status pfs_server_submit(pfs_server* s, pfs_ctx* session, double release_ms, uint32_t maxvar, uint64_t ncls,
                             const uint64_t* offsets, const uint32_t* lits, const uint32_t* assumed,
                             uint32_t nassumed, uint32_t* ticket) {
  if (!s || !session || (ncls && (!offsets || !lits)) || (nassumed && !assumed)) return PFS_BAD_INPUT;
  auto r = std::make_unique<Request>();
  r->ctx = session;
  r->release_ms = release_ms;
  r->maxvar = maxvar;
  if (ncls) {
    r->off.assign(offsets, offsets + ncls + 1);
    const uint64_t base = offsets[0];
    for (auto& o : r->off) o -= base;
    r->lits.assign(lits + base, lits + offsets[ncls]);
  }
  r->assumed.assign(assumed, assumed + nassumed);
  std::lock_guard<std::mutex> lk(s->m);
  if (ticket) *ticket = (uint32_t)s->reqs.size();
  s->reqs.push_back(std::move(r));
  s->cv.notify_all();
  return PFS_OK;
}

pfs_status pfs_server_start(pfs_server* s) {
  if (!s) return PFS_BAD_INPUT;
  std::lock_guard<std::mutex> lk(s->m);
  if (s->started) return PFS_BAD_STATE;
  s->started = true;
  s->t0 = std::chrono::steady_clock::now();
  for (uint32_t w = 0; w < s->nworkers; w++) s->workers.emplace_back([s] { s->work(); });
  return PFS_OK;
}

pfs_status pfs_server_wait(pfs_server* s, uint32_t ticket, int32_t* result, double* release_ms, double* start_ms,
                           double* end_ms) {
  if (!s) return PFS_BAD_INPUT;
  std::unique_lock<std::mutex> lk(s->m);
  if (ticket >= s->reqs.size()) return PFS_BAD_INPUT;
  if (!s->started) return PFS_BAD_STATE;
  Request& r = *s->reqs[ticket];
  s->cv.wait(lk, [&] { return r.done; });
  if (result) *result = r.result;
  if (release_ms) *release_ms = r.release_ms;
  if (start_ms) *start_ms = r.start_ms;
  if (end_ms) *end_ms = r.end_ms;
  return r.status;
}

void pfs_server_destroy(pfs_server* s) {
  if (!s) return;
  {
    std::unique_lock<std::mutex> lk(s->m);
    if (s->started) {
      // finish what was submitted, then stop the workers
      s->cv.wait(lk, [&] {
        for (auto& r : s->reqs)
          if (!r->done) return false;
        return true;
      });
    }
    s->stopping = true;
    s->cv.notify_all();
  }
  for (auto& t : s->workers) t.join();
  delete s;
}

}  // extern "C"
