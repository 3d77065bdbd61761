// Host build of the device arena's allocator (parafrost_amd/csrc/pfs_arena.cpp)
// for the CPU tests.  TEST INFRASTRUCTURE ONLY.
//
// The device runtime calls the arena makes are stubbed with host memory
// (slabs from malloc; copies queued per stream and run when a stream is
// waited on, in a random interleaving with the other streams' queued copies,
// as streams run concurrently on the device), so the placement, coalescing,
// in-place growth, stream-ordered release and compaction logic run
// unchanged.  A randomized driver fills every live block with a pattern
// derived from its id and checks, after every operation, that each block's
// slot still points at its own pattern, that blocks never overlap and that
// the accounting adds up.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <random>
#include <set>
#include <vector>

#include "../../parafrost_amd/csrc/pfs_arena.hpp"

namespace pfs {
struct Copy {
  void* dst;
  const void* src;
  size_t bytes;
};
struct Stream {
  int id;
  std::deque<Copy> q;  // queued copies, in stream order
};
struct Marker {
  Stream* st;
};
static std::vector<Stream*> g_streams;
static std::mt19937_64 g_sched(1);
static size_t g_dev_limit = ~(size_t)0, g_dev_used = 0;
void* dev_alloc(size_t bytes) {
  if (g_dev_used + bytes > g_dev_limit) throw std::bad_alloc();
  void* p = aligned_alloc(256, (bytes + 255) & ~(size_t)255);
  if (!p) throw std::bad_alloc();
  g_dev_used += bytes;
  return p;
}
void dev_free(void* p) { free(p); }
void d2d(void* dst, const void* src, size_t bytes, Stream* st) { st->q.push_back(Copy{dst, src, bytes}); }
// runs queued copies of randomly chosen streams until st has none left
bool stream_wait(Stream* st) {
  while (!st->q.empty()) {
    std::vector<Stream*> busy;
    for (Stream* s : g_streams)
      if (!s->q.empty()) busy.push_back(s);
    Stream* s = busy[g_sched() % busy.size()];
    const Copy c = s->q.front();
    s->q.pop_front();
    memmove(c.dst, c.src, c.bytes);
  }
  return true;
}
Marker* marker_record(Stream* st) { return new Marker{st}; }
bool marker_done(Marker* m) { return m->st->q.empty(); }
void marker_wait(Marker* m) { stream_wait(m->st); }
void marker_free(Marker* m) { delete m; }
void dev_mem_info(size_t* f, size_t* t) { *f = *t = 1ull << 30; }
}  // namespace pfs

#include "../../parafrost_amd/csrc/pfs_arena.cpp"

using namespace pfs;

struct Live {
  void* p;      // the slot the arena keeps current
  size_t size;
  uint32_t id;
  int client;
  bool stable;
};

static bool check(std::map<uint32_t, Live*>& live, DeviceArena& a, size_t* client_used, int nclients) {
  std::vector<std::pair<char*, char*>> spans;
  size_t per[8] = {0};
  for (auto& kv : live) {
    Live* l = kv.second;
    const uint32_t* w = (const uint32_t*)l->p;
    for (size_t k = 0; k < l->size / 4; k++)
      if (w[k] != (l->id * 2654435761u) + (uint32_t)k) {
        printf("block %u corrupted at word %zu\n", l->id, k);
        return false;
      }
    spans.push_back({(char*)l->p, (char*)l->p + l->size});
    per[l->client] += (l->size + 255) & ~(size_t)255;
  }
  std::sort(spans.begin(), spans.end());
  for (size_t i = 1; i < spans.size(); i++)
    if (spans[i].first < spans[i - 1].second) {
      printf("blocks overlap\n");
      return false;
    }
  for (int c = 0; c < nclients; c++)
    if (per[c] != client_used[c]) {
      printf("client %d accounting %zu != %zu\n", c, per[c], client_used[c]);
      return false;
    }
  return true;
}

int main(int argc, char** argv) {
  const int seeds = argc > 1 ? atoi(argv[1]) : 50;
  const int ops = argc > 2 ? atoi(argv[2]) : 3000;
  uint64_t compactions = 0, moved = 0, fails = 0, grows = 0;
  for (int seed = 0; seed < seeds; seed++) {
    std::mt19937_64 rng(seed);
    const size_t cap = (size_t)(1 + rng() % 8) << 18;  // 256 KiB .. 2 MiB
    DeviceArena* a = DeviceArena::create_private(0, cap);
    const int nclients = 4;
    Stream sts[nclients];
    ArenaClient cl[nclients];
    g_streams.clear();
    for (int c = 0; c < nclients; c++) {
      sts[c].id = c;
      g_streams.push_back(&sts[c]);
      cl[c].arena = a;
      cl[c].st = &sts[c];
    }
    std::map<uint32_t, Live*> live;
    uint32_t next_id = 1;
    for (int op = 0; op < ops; op++) {
      const int c = (int)(rng() % nclients);
      const int kind = (int)(rng() % 10);
      if (kind < 5 || live.empty()) {  // allocate
        Live* l = new Live();
        l->size = ((rng() % 2 ? 1 + rng() % 2048 : 1 + rng() % (cap / 6)) + 3) & ~(size_t)3;
        l->id = next_id++;
        l->client = c;
        l->stable = rng() % 4 == 0;
        try {
          a->alloc(&cl[c], l->size, l->stable ? R_STABLE : R_DYNAMIC, &l->p);
        } catch (std::bad_alloc&) {
          fails++;
          delete l;
          continue;
        }
        uint32_t* w = (uint32_t*)l->p;
        for (size_t k = 0; k < l->size / 4; k++) w[k] = (l->id * 2654435761u) + (uint32_t)k;
        live[l->id] = l;
      } else if (kind < 8) {  // release a random block of client c
        auto it = live.begin();
        std::advance(it, rng() % live.size());
        Live* l = it->second;
        const bool idle = rng() % 4 == 0;  // the owner waited for its stream: freed at once
        if (idle) stream_wait(&sts[l->client]);
        a->release(&cl[l->client], l->p, idle);
        live.erase(it);
        delete l;
      } else if (kind == 8) {  // grow a block in place (or not)
        auto it = live.begin();
        std::advance(it, rng() % live.size());
        Live* l = it->second;
        const size_t nb = (l->size + 4 + rng() % 16384) & ~(size_t)3;
        if (a->grow(&cl[l->client], l->p, nb)) {
          uint32_t* w = (uint32_t*)l->p;
          for (size_t k = l->size / 4; k < nb / 4; k++) w[k] = (l->id * 2654435761u) + (uint32_t)k;
          l->size = nb;
          grows++;
        }
      } else if (rng() % 4 == 0) {  // client c's pending releases returned (its stream is going away)
        a->drain(&sts[c]);
      } else {  // compaction with client c at a safe point (every client idle here)
        moved += a->compact(&cl[c], true);
        compactions++;
      }
      size_t used[nclients];
      for (int k = 0; k < nclients; k++) used[k] = cl[k].used;
      if (!check(live, *a, used, nclients)) {
        printf("seed %d op %d failed\n", seed, op);
        return 1;
      }
    }
    for (auto& kv : live) {
      a->release(&cl[kv.second->client], kv.second->p);
      delete kv.second;
    }
    a->compact(&cl[0], true);  // also returns every pending release to the holes
    const ArenaStats s = a->stats();
    if (s.used != 0 || s.holes != 1 || s.largest_hole != s.capacity) {
      printf("seed %d: arena not empty at the end (used %llu, holes %llu)\n", seed, (unsigned long long)s.used,
             (unsigned long long)s.holes);
      return 1;
    }
    compactions += s.compactions;
    delete a;
  }
  printf("{\"seeds\": %d, \"compactions\": %llu, \"moved_bytes\": %llu, \"alloc_fails\": %llu, \"grows\": %llu}\n",
         seeds, (unsigned long long)compactions, (unsigned long long)moved, (unsigned long long)fails,
         (unsigned long long)grows);
  return 0;
}
