// Device arena: one pool of device memory per device, shared by every
// context on it, sub-allocated on the host (the role of the reference's
// cuArena pool, src/gpu/memory.cu:99-166, behind cuMM).
//
// The pool is reserved once (hipMalloc of a slab; further slabs only when a
// request cannot fit after compaction), so the simplify rounds, serving
// requests and batch runs allocate and free without the runtime: no
// hipMalloc, and no hipFree, whose implicit device-wide synchronisation would
// stall every other context's stream.
//
// Two regions per slab, as in the reference:
//   stable  -- per-variable arrays, placed from the top of a slab;
//              compaction moves them up;
//   dynamic -- clause blocks, occurrence tables and temporaries, placed from
//              the bottom; compaction moves them down.
// A block names the pointer that holds its address (its slot); compaction
// rewrites that pointer.  A block is moved only while its context cannot be
// using the address: the context is between calls (its client lock is free
// and its stream is synchronised) or is the caller at the start of a call.
// Releases are stream-ordered: a freed block returns to the pool when the
// owner's stream has passed the release (a timing-free event), so no context
// can be handed memory that queued work of another still reads.
#pragma once

#include <cstddef>
#include <cstdint>
#include <map>
#include <mutex>
#include <vector>

#include "pfs_launch.hpp"

namespace pfs {

enum Region { R_STABLE = 0, R_DYNAMIC = 1 };

class DeviceArena;

// A context's handle: its stream (copies of its moved blocks run there) and
// the lock it holds while it works on the device.
struct ArenaClient {
  DeviceArena* arena = nullptr;
  Stream* st = nullptr;
  std::mutex busy;
  size_t used = 0;  // bytes of live blocks
};

struct ArenaStats {
  uint64_t capacity = 0, used = 0, peak = 0, largest_hole = 0, holes = 0, slabs = 0;
  uint64_t compactions = 0, moved_blocks = 0, moved_bytes = 0, slab_allocs = 0, pending = 0;
};

class DeviceArena {
 public:
  // the default arena of a device, created on first use (its first slab is
  // reserved by the first allocation)
  static DeviceArena& get(int device);
  // the default arena's first slab size and the most all its slabs together
  // may reach (0: no limit); false once the arena exists
  static bool configure(int device, uint64_t first_slab, uint64_t limit);
  // an arena of one slab of exactly 'capacity' bytes that never grows
  // (admission by a fixed device budget); destroyed with delete
  static DeviceArena* create_private(int device, uint64_t capacity);
  ~DeviceArena();

  // a block of 'bytes' for client c; *slot receives the address and is kept
  // up to date when compaction moves the block.  Throws std::bad_alloc.
  void* alloc(ArenaClient* c, size_t bytes, Region r, void** slot);
  // grows the block at p in place to 'bytes' when the space after it is free
  bool grow(ArenaClient* c, void* p, size_t bytes);
  // returns a block; 'idle': the owner's stream has no queued work (the
  // block is free at once, without a stream marker)
  void release(ArenaClient* c, void* p, bool idle = false);
  // returns every pending release made on stream st (idle) to the holes, so
  // that no marker outlives the stream it was recorded on
  void drain(Stream* st);
  // the pointer that holds the block at p is now *slot (the block changed
  // hands between two buffer records)
  void retarget(void* p, void** slot);
  // moves the blocks of idle clients (and of 'self' when self_movable: the
  // caller at a point where it holds no device address) over the holes,
  // dynamic ones down and stable ones up.  Returns the bytes moved.
  uint64_t compact(ArenaClient* self, bool self_movable);
  // an allocation found no hole although the pool may hold enough in total
  bool wants_compaction() const { return want_compact_; }
  ArenaStats stats();
  int device() const { return device_; }

 private:
  struct Block {
    size_t size;
    Region region;
    ArenaClient* owner;  // null once released
    void** slot;         // null once released
  };
  struct Slab {
    char* base = nullptr;
    size_t size = 0;
    std::map<size_t, Block> blocks;  // by offset
    std::map<size_t, size_t> holes;  // offset -> size, coalesced
  };
  struct Pending {
    Marker* marker;  // the owner's stream position at the release
    Stream* st;      // that stream
    size_t slab, off;
  };
  explicit DeviceArena(int device) : device_(device) {}
  bool add_slab(size_t bytes);
  bool place(ArenaClient* c, size_t bytes, Region r, void** slot, void** out);
  void free_block(size_t slab, size_t off);
  void reclaim(bool wait);
  bool find(void* p, size_t* slab, size_t* off);
  uint64_t compact_locked(ArenaClient* self, bool self_movable);

  int device_;
  std::mutex m_;
  std::vector<Slab> slabs_;
  std::vector<Pending> pending_;
  uint64_t used_ = 0, peak_ = 0, capacity_ = 0;
  uint64_t first_slab_ = 0, limit_ = 0;
  uint64_t compactions_ = 0, moved_blocks_ = 0, moved_bytes_ = 0, slab_allocs_ = 0;
  bool want_compact_ = false;
};

}  // namespace pfs
