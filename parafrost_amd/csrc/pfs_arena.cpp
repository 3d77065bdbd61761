// Device arena (see pfs_arena.hpp).
#include "pfs_arena.hpp"

#include <algorithm>
#include <cstdlib>
#include <new>
#include <string>

namespace pfs {

static constexpr size_t ALIGN = 256;  // every block starts on a 256-byte boundary
static size_t round_up(size_t x) { return (x + ALIGN - 1) & ~(ALIGN - 1); }
static constexpr uint64_t GiB = 1ull << 30, MiB = 1ull << 20;
// a moved block whose old and new ranges overlap is copied in pieces of the
// distance it moves (each piece's target was already copied out); more
// pieces than this and the block stays where it is
static constexpr size_t MAX_PIECES = 256;

struct ArenaRegistry {
  std::mutex m;
  std::map<int, DeviceArena*> arenas;                       // the default arena of each device
  std::map<int, std::pair<uint64_t, uint64_t>> config;      // first slab, limit
};
static ArenaRegistry& registry() {
  static ArenaRegistry* r = new ArenaRegistry();  // lives until the process ends, like its pools
  return *r;
}

DeviceArena& DeviceArena::get(int device) {
  ArenaRegistry& R = registry();
  std::lock_guard<std::mutex> lk(R.m);
  auto it = R.arenas.find(device);
  if (it != R.arenas.end()) return *it->second;
  DeviceArena* a = new DeviceArena(device);
  auto c = R.config.find(device);
  if (c != R.config.end()) {
    a->first_slab_ = c->second.first;
    a->limit_ = c->second.second;
  } else if (const char* e = getenv("PFS_ARENA_MB")) {
    a->first_slab_ = strtoull(e, nullptr, 10) * MiB;
  }
  R.arenas[device] = a;
  return *a;
}

bool DeviceArena::configure(int device, uint64_t first_slab, uint64_t limit) {
  ArenaRegistry& R = registry();
  std::lock_guard<std::mutex> lk(R.m);
  if (R.arenas.count(device)) return false;
  R.config[device] = {first_slab, limit};
  return true;
}

DeviceArena* DeviceArena::create_private(int device, uint64_t capacity) {
  DeviceArena* a = new DeviceArena(device);
  a->first_slab_ = a->limit_ = round_up(capacity);
  return a;
}

DeviceArena::~DeviceArena() {
  reclaim(true);
  for (Slab& s : slabs_) dev_free(s.base);
}

bool DeviceArena::add_slab(size_t bytes) {
  bytes = round_up(bytes);
  if (limit_ && capacity_ + bytes > limit_) return false;
  void* p = nullptr;
  try {
    p = dev_alloc(bytes);
  } catch (std::bad_alloc&) {
    return false;
  }
  Slab s;
  s.base = (char*)p;
  s.size = bytes;
  s.holes[0] = bytes;
  slabs_.push_back(std::move(s));
  capacity_ += bytes;
  slab_allocs_++;
  return true;
}

// dynamic blocks: the lowest hole that fits; stable blocks: the top of the
// highest hole that fits
bool DeviceArena::place(ArenaClient* c, size_t bytes, Region r, void** slot, void** out) {
  for (size_t si = 0; si < slabs_.size(); si++) {
    Slab& s = slabs_[si];
    auto fit = s.holes.end();
    if (r == R_DYNAMIC) {
      for (auto it = s.holes.begin(); it != s.holes.end(); ++it)
        if (it->second >= bytes) {
          fit = it;
          break;
        }
    } else {
      for (auto it = s.holes.rbegin(); it != s.holes.rend(); ++it)
        if (it->second >= bytes) {
          fit = std::prev(it.base());
          break;
        }
    }
    if (fit == s.holes.end()) continue;
    const size_t hoff = fit->first, hsize = fit->second;
    s.holes.erase(fit);
    size_t off;
    if (r == R_DYNAMIC) {
      off = hoff;
      if (hsize > bytes) s.holes[hoff + bytes] = hsize - bytes;
    } else {
      off = hoff + hsize - bytes;
      if (hsize > bytes) s.holes[hoff] = hsize - bytes;
    }
    s.blocks[off] = Block{bytes, r, c, slot};
    used_ += bytes;
    peak_ = std::max(peak_, used_);
    c->used += bytes;
    *out = s.base + off;
    if (slot) *slot = *out;
    return true;
  }
  return false;
}

static void give_range(std::map<size_t, size_t>& holes, size_t off, size_t bytes);

void DeviceArena::free_block(size_t si, size_t off) {
  Slab& s = slabs_[si];
  auto b = s.blocks.find(off);
  const size_t bytes = b->second.size;
  s.blocks.erase(b);
  give_range(s.holes, off, bytes);
}

bool DeviceArena::find(void* p, size_t* si, size_t* off) {
  for (size_t i = 0; i < slabs_.size(); i++) {
    const char* b = slabs_[i].base;
    if ((char*)p >= b && (char*)p < b + slabs_[i].size) {
      *si = i;
      *off = (size_t)((char*)p - b);
      return slabs_[i].blocks.count(*off) != 0;
    }
  }
  return false;
}

// released blocks whose owner's stream has passed the release go back to
// the holes (all of them, waiting, with 'wait')
void DeviceArena::reclaim(bool wait) {
  size_t j = 0;
  for (size_t i = 0; i < pending_.size(); i++) {
    Pending& p = pending_[i];
    const bool done = wait ? (marker_wait(p.marker), true) : marker_done(p.marker);
    if (done) {
      marker_free(p.marker);
      free_block(p.slab, p.off);
    } else {
      pending_[j++] = p;
    }
  }
  pending_.resize(j);
}

void* DeviceArena::alloc(ArenaClient* c, size_t bytes, Region r, void** slot) {
  std::lock_guard<std::mutex> lk(m_);
  bytes = round_up(std::max<size_t>(bytes, 1));
  if (slabs_.empty()) {
    // the first slab: the configured size, else a quarter of the device's
    // free memory (at most 64 GiB), halved while the runtime refuses it; a
    // private arena's one slab has exactly its capacity
    uint64_t want = first_slab_;
    if (!want) {
      size_t fr = 0, tot = 0;
      dev_mem_info(&fr, &tot);
      want = std::min<uint64_t>(64 * GiB, std::max<uint64_t>(fr / 4, 256 * MiB));
    }
    if (limit_ && first_slab_ == limit_) {
      if (!add_slab(want)) throw std::bad_alloc();
    } else {
      want = std::max<uint64_t>(want, bytes);
      while (!add_slab(want)) {
        if (want <= std::max<uint64_t>(bytes, 64 * MiB)) throw std::bad_alloc();
        want = std::max<uint64_t>(want / 2, bytes);
      }
    }
  }
  void* out = nullptr;
  reclaim(false);
  if (place(c, bytes, r, slot, &out)) return out;
  if (!pending_.empty()) {
    reclaim(true);
    if (place(c, bytes, r, slot, &out)) return out;
  }
  // no hole fits: compact the blocks of idle contexts, then grow the pool
  want_compact_ = true;
  if (compact_locked(c, false) && place(c, bytes, r, slot, &out)) return out;
  if (add_slab(std::max<uint64_t>(bytes, std::min<uint64_t>(first_slab_ ? first_slab_ : 16 * GiB, 16 * GiB))) &&
      place(c, bytes, r, slot, &out))
    return out;
  if (add_slab(bytes) && place(c, bytes, r, slot, &out)) return out;
  throw std::bad_alloc();
}

bool DeviceArena::grow(ArenaClient* c, void* p, size_t bytes) {
  std::lock_guard<std::mutex> lk(m_);
  size_t si, off;
  if (!find(p, &si, &off)) return false;
  Slab& s = slabs_[si];
  Block& b = s.blocks[off];
  bytes = round_up(bytes);
  if (bytes <= b.size) return true;
  auto h = s.holes.find(off + b.size);
  if (h == s.holes.end() || b.size + h->second < bytes) return false;
  const size_t add = bytes - b.size, rest = h->second - add, at = h->first + add;
  s.holes.erase(h);
  if (rest) s.holes[at] = rest;
  b.size = bytes;
  used_ += add;
  peak_ = std::max(peak_, used_);
  c->used += add;
  return true;
}

void DeviceArena::release(ArenaClient* c, void* p, bool idle) {
  if (!p) return;
  std::lock_guard<std::mutex> lk(m_);
  size_t si, off;
  if (!find(p, &si, &off)) return;
  Block& b = slabs_[si].blocks[off];
  if (!b.owner) return;  // already released
  used_ -= b.size;
  c->used -= b.size;
  if (idle) {
    free_block(si, off);
    return;
  }
  // the owner's queued work may still read the block: it returns to the
  // holes once the stream has passed this point
  b.owner = nullptr;
  b.slot = nullptr;
  pending_.push_back(Pending{marker_record(c->st), c->st, si, off});
}

void DeviceArena::drain(Stream* st) {
  std::lock_guard<std::mutex> lk(m_);
  size_t j = 0;
  for (size_t i = 0; i < pending_.size(); i++) {
    Pending& p = pending_[i];
    if (p.st == st) {
      marker_wait(p.marker);
      marker_free(p.marker);
      free_block(p.slab, p.off);
    } else {
      pending_[j++] = p;
    }
  }
  pending_.resize(j);
}

void DeviceArena::retarget(void* p, void** slot) {
  if (!p) return;
  std::lock_guard<std::mutex> lk(m_);
  size_t si, off;
  if (find(p, &si, &off) && slabs_[si].blocks[off].owner) slabs_[si].blocks[off].slot = slot;
}

uint64_t DeviceArena::compact(ArenaClient* self, bool self_movable) {
  std::lock_guard<std::mutex> lk(m_);
  return compact_locked(self, self_movable);
}

// takes [off, off + bytes) out of the hole that contains it
static void take_range(std::map<size_t, size_t>& holes, size_t off, size_t bytes) {
  auto h = std::prev(holes.upper_bound(off));
  const size_t ho = h->first, hs = h->second;
  holes.erase(h);
  if (off > ho) holes[ho] = off - ho;
  if (off + bytes < ho + hs) holes[off + bytes] = ho + hs - off - bytes;
}

// returns [off, off + bytes) to the holes, merged with its neighbours
static void give_range(std::map<size_t, size_t>& holes, size_t off, size_t bytes) {
  size_t lo = off, hi = off + bytes;
  auto nx = holes.lower_bound(off);
  if (nx != holes.end() && nx->first == hi) {
    hi += nx->second;
    nx = holes.erase(nx);
  }
  if (nx != holes.begin()) {
    auto pv = std::prev(nx);
    if (pv->first + pv->second == lo) {
      lo = pv->first;
      holes.erase(pv);
    }
  }
  holes[lo] = hi - lo;
}

// copies a block from 'from' to 'to' on one stream; ranges that overlap are
// copied in pieces of the distance moved, in the order that reads every
// piece before it is overwritten (ascending for a move down, descending for
// a move up).  False (nothing copied) when that takes more than MAX_PIECES.
static bool copy_block(char* base, size_t from, size_t to, size_t bytes, Stream* st) {
  const size_t gap = from > to ? from - to : to - from;
  if (gap < bytes && (bytes + gap - 1) / gap > MAX_PIECES) return false;
  if (gap >= bytes) {
    d2d(base + to, base + from, bytes, st);
  } else if (to < from) {
    for (size_t done = 0; done < bytes; done += gap)
      d2d(base + to + done, base + from + done, std::min(gap, bytes - done), st);
  } else {
    for (size_t end = bytes; end > 0; end -= std::min(gap, end)) {
      const size_t lo = end - std::min(gap, end);
      d2d(base + to + lo, base + from + lo, end - lo, st);
    }
  }
  return true;
}

// Compaction.  Blocks of idle clients move, the caller's only when it holds
// no device address; every other block is an obstacle that stays put.
//   dynamic blocks, lowest first: into the lowest hole below them that
//     holds them whole, else down over the hole right below them;
//   stable blocks, highest first: into the top of the highest hole above
//     them that holds them whole, else up over the hole right above them.
// Every copy runs on one stream, in this order, so a block may land in the
// space another block left earlier in the same compaction; the owners'
// streams are idle, and the copies are finished before any of them runs
// again.
uint64_t DeviceArena::compact_locked(ArenaClient* self, bool self_movable) {
  reclaim(true);
  want_compact_ = false;
  std::vector<ArenaClient*> locked;
  std::map<ArenaClient*, bool> movable;
  for (Slab& s : slabs_)
    for (auto& kv : s.blocks) {
      ArenaClient* o = kv.second.owner;
      if (!o || movable.count(o)) continue;
      if (o == self) {
        movable[o] = self_movable;
      } else if (o->busy.try_lock()) {
        locked.push_back(o);
        movable[o] = stream_wait(o->st);  // an idle stream; a failed one keeps its blocks
      } else {
        movable[o] = false;
      }
    }
  Stream* cs = self ? self->st : nullptr;
  uint64_t moved = 0;
  for (Slab& s : slabs_) {
    auto relocate = [&](size_t from, size_t to) {
      Block b = s.blocks[from];
      if (!cs) cs = b.owner->st;
      if (!copy_block(s.base, from, to, b.size, cs)) return;
      s.blocks.erase(from);
      give_range(s.holes, from, b.size);
      take_range(s.holes, to, b.size);
      s.blocks[to] = b;
      *b.slot = s.base + to;
      moved += b.size;
      moved_blocks_++;
    };
    auto can_move = [&](const Block& b, Region r) { return b.owner && b.slot && b.region == r && movable[b.owner]; };
    std::vector<size_t> order;
    for (auto& kv : s.blocks)
      if (can_move(kv.second, R_DYNAMIC)) order.push_back(kv.first);
    for (size_t off : order) {  // ascending
      const size_t size = s.blocks[off].size;
      size_t to = off;
      for (auto& h : s.holes) {
        if (h.first >= off) break;
        if (h.second >= size && h.first + size <= off) {
          to = h.first;
          break;
        }
        if (h.first + h.second == off) to = h.first;  // the hole right below: slide
      }
      if (to != off) relocate(off, to);
    }
    order.clear();
    for (auto& kv : s.blocks)
      if (can_move(kv.second, R_STABLE)) order.push_back(kv.first);
    for (auto it = order.rbegin(); it != order.rend(); ++it) {  // descending
      const size_t off = *it, size = s.blocks[off].size;
      size_t to = off;
      for (auto h = s.holes.rbegin(); h != s.holes.rend(); ++h) {
        const size_t hend = h->first + h->second;
        if (hend <= off + size) break;
        if (h->second >= size && h->first >= off + size) {
          to = hend - size;
          break;
        }
        if (h->first == off + size) to = hend - size;  // the hole right above: slide
      }
      if (to != off) relocate(off, to);
    }
  }
  if (moved && cs) stream_wait(cs);  // the moves are done before any owner runs again
  for (ArenaClient* o : locked) o->busy.unlock();
  compactions_++;
  moved_bytes_ += moved;
  return moved;
}

ArenaStats DeviceArena::stats() {
  std::lock_guard<std::mutex> lk(m_);
  ArenaStats st;
  st.capacity = capacity_;
  st.used = used_;
  st.peak = peak_;
  st.slabs = slabs_.size();
  for (Slab& s : slabs_)
    for (auto& h : s.holes) {
      st.holes++;
      st.largest_hole = std::max<uint64_t>(st.largest_hole, h.second);
    }
  st.compactions = compactions_;
  st.moved_blocks = moved_blocks_;
  st.moved_bytes = moved_bytes_;
  st.slab_allocs = slab_allocs_;
  st.pending = pending_.size();
  return st;
}

}  // namespace pfs
