// HIP kernels of the simplifier for CDNA4 (gfx950, wave64).
//
// Work mapping:
//   * clause-parallel stages (BCP clean-up, counting, compaction) use one
//     thread per clause;
//   * the occurrence table is built by literal ranges: 1024-thread
//     workgroups per tile of clauses count and partition the occurrences
//     by range in LDS, then one workgroup per range counts, scans and
//     places its lists in LDS (no global atomic per occurrence);
//   * list-parallel stages (BCP visits, list deletions) use one workgroup
//     per occurrence list with the threads striding the list;
//   * per-variable stages (SUB, BVE, BCE, ERE, LCVE election) use one
//     thread per elected variable in 64-thread workgroups, so that one
//     wavefront finishes as soon as its own variables are done;
//   * scans and the stable radix sort use 256-thread workgroups with
//     wave64 shuffles for the in-wave prefix and LDS for the wave totals.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <new>
#include <vector>
#include <stdexcept>
#include <string>

#include "pfs_launch.hpp"

namespace pfs {

#define PFS_HIP(call)                                                                      \
  do {                                                                                     \
    hipError_t e_ = (call);                                                                \
    if (e_ != hipSuccess)                                                                  \
      throw std::runtime_error(std::string("HIP error: ") + hipGetErrorString(e_) + " at " + \
                               __FILE__ + ":" + std::to_string(__LINE__));                 \
  } while (0)

struct KRec {
  const char* name;
  hipEvent_t a, b;
};
// Small device-to-host reads (counts, flags, a few records) land in pinned
// staging memory: a copy into pageable memory makes the runtime stage it
// itself, which adds tens of microseconds to every read-and-wait.  The
// copies to their destinations happen in stream_sync, which every caller
// runs before it looks at what it read.
constexpr size_t PIN_BYTES = 64 << 10, PIN_SMALL = 4096;
struct PinCopy {
  void* dst;
  size_t off, bytes;
};
struct Stream {
  hipStream_t s;
  int device;
  bool prof = false;        // record events around every kernel launch
  std::vector<KRec> recs;   // pending event pairs
  unsigned char* pin = nullptr;  // PIN_BYTES of pinned host memory
  size_t pin_used = 0;
  std::vector<PinCopy> pend;     // staged reads not yet copied out
};
static void pin_flush(Stream* st) {
  for (const PinCopy& c : st->pend) memcpy(c.dst, st->pin + c.off, c.bytes);
  st->pend.clear();
  st->pin_used = 0;
}
struct Event {
  hipEvent_t e;
};

static inline unsigned grid_for(size_t n, unsigned block) {
  return (unsigned)((n + block - 1) / block);
}
#define LAUNCH_CHECK() PFS_HIP(hipGetLastError())
// launch on the context stream; with profiling on, bracket it with events
#define KLAUNCH(kern, grid, block, st, ...) KLAUNCH_LDS(kern, grid, block, 0, st, __VA_ARGS__)
// the same with 'lds' bytes of dynamic LDS reserved per workgroup
#define KLAUNCH_LDS(kern, grid, block, lds, st, ...)          \
  do {                                                         \
    hipEvent_t e0_ = nullptr;                                  \
    if ((st)->prof) {                                          \
      PFS_HIP(hipEventCreate(&e0_));                           \
      PFS_HIP(hipEventRecord(e0_, (st)->s));                   \
    }                                                          \
    kern<<<(grid), (block), (lds), (st)->s>>>(__VA_ARGS__);    \
    LAUNCH_CHECK();                                            \
    if ((st)->prof) {                                          \
      hipEvent_t e1_;                                          \
      PFS_HIP(hipEventCreate(&e1_));                           \
      PFS_HIP(hipEventRecord(e1_, (st)->s));                   \
      (st)->recs.push_back(KRec{#kern, e0_, e1_});             \
    }                                                          \
  } while (0)

// ---------------------------------------------------------------- runtime
Stream* stream_create(int device) {
  PFS_HIP(hipSetDevice(device));
  Stream* st = new Stream();
  st->device = device;
  PFS_HIP(hipStreamCreateWithFlags(&st->s, hipStreamNonBlocking));
  if (hipHostMalloc((void**)&st->pin, PIN_BYTES, hipHostMallocDefault) != hipSuccess) st->pin = nullptr;
  return st;
}
void prof_enable(Stream* st, bool on) { st->prof = on; }
void stream_bind(Stream* st) { PFS_HIP(hipSetDevice(st->device)); }
void prof_collect(Stream* st, std::vector<KernelTime>& out) {
  PFS_HIP(hipStreamSynchronize(st->s));
  pin_flush(st);
  std::map<std::string, size_t> idx;
  for (size_t i = 0; i < out.size(); i++) idx[out[i].name] = i;
  for (const KRec& r : st->recs) {
    float ms = 0;
    PFS_HIP(hipEventElapsedTime(&ms, r.a, r.b));
    (void)hipEventDestroy(r.a);
    (void)hipEventDestroy(r.b);
    auto it = idx.find(r.name);
    if (it == idx.end()) {
      idx[r.name] = out.size();
      out.push_back(KernelTime{r.name, 0.0, 0});
      it = idx.find(r.name);
    }
    out[it->second].ms += ms;
    out[it->second].launches++;
  }
  st->recs.clear();
}
void stream_destroy(Stream* st) {
  if (!st) return;
  for (const KRec& r : st->recs) {
    (void)hipEventDestroy(r.a);
    (void)hipEventDestroy(r.b);
  }
  (void)hipStreamSynchronize(st->s);
  // staged reads still pending here belong to a call that ended with an
  // error: their destinations may be gone, so they are dropped
  st->pend.clear();
  if (st->pin) (void)hipHostFree(st->pin);
  (void)hipStreamDestroy(st->s);
  delete st;
}
void stream_sync(Stream* st) {
  const hipError_t e = hipStreamSynchronize(st->s);
  if (e != hipSuccess) {
    // the staged reads did not complete: drop them (their destinations are
    // locals of frames the exception unwinds), then report
    st->pend.clear();
    st->pin_used = 0;
    throw std::runtime_error(std::string("HIP error: ") + hipGetErrorString(e) + " at " + __FILE__ + ":" +
                             std::to_string(__LINE__));
  }
  pin_flush(st);
}
bool stream_quiesce(Stream* st) {
  if (!st) return true;
  const bool ok = hipStreamSynchronize(st->s) == hipSuccess;
  if (!ok) (void)hipGetLastError();
  st->pend.clear();
  st->pin_used = 0;
  return ok;
}
void* dev_alloc(size_t bytes) {
  void* p = nullptr;
  if (bytes == 0) bytes = 256;
  if (hipMalloc(&p, bytes) != hipSuccess) {
    (void)hipGetLastError();
    throw std::bad_alloc();
  }
  return p;
}
void dev_free(void* p) {
  if (p) (void)hipFree(p);
}
// small host-to-device copies go through the same pinned area (the host
// data is copied there first; the area is reused only after stream_sync)
void h2d(void* dst, const void* src, size_t bytes, Stream* st) {
  if (!bytes) return;
  if (st->pin && bytes <= PIN_SMALL) {
    if (st->pin_used + bytes > PIN_BYTES) stream_sync(st);
    memcpy(st->pin + st->pin_used, src, bytes);
    PFS_HIP(hipMemcpyAsync(dst, st->pin + st->pin_used, bytes, hipMemcpyHostToDevice, st->s));
    st->pin_used += (bytes + 15) & ~(size_t)15;
    return;
  }
  PFS_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st->s));
}
void d2h(void* dst, const void* src, size_t bytes, Stream* st) {
  if (!bytes) return;
  if (st->pin && bytes <= PIN_SMALL) {
    if (st->pin_used + bytes > PIN_BYTES) stream_sync(st);
    PFS_HIP(hipMemcpyAsync(st->pin + st->pin_used, src, bytes, hipMemcpyDeviceToHost, st->s));
    st->pend.push_back(PinCopy{dst, st->pin_used, bytes});
    st->pin_used += (bytes + 15) & ~(size_t)15;
    return;
  }
  PFS_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, st->s));
}
void d2d(void* dst, const void* src, size_t bytes, Stream* st) {
  if (bytes) PFS_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, st->s));
}
// 16 bytes per thread: one 16-byte store where the address allows it
__global__ void fill_u32_k(u32* d, u32 v, size_t n) {
  const size_t i = ((size_t)blockIdx.x * blockDim.x + threadIdx.x) * 4;
  if (i + 4 <= n && ((uintptr_t)(d + i) & 15u) == 0) {
    *(uint4*)(d + i) = make_uint4(v, v, v, v);
    return;
  }
  for (size_t k = i; k < n && k < i + 4; k++) d[k] = v;
}
__global__ void fill_u8_k(u8* d, u8 v, size_t n) {
  const size_t i = ((size_t)blockIdx.x * blockDim.x + threadIdx.x) * 16;
  if (i + 16 <= n && ((uintptr_t)(d + i) & 15u) == 0) {
    const u32 w = v * 0x01010101u;
    *(uint4*)(d + i) = make_uint4(w, w, w, w);
    return;
  }
  for (size_t k = i; k < n && k < i + 16; k++) d[k] = v;
}
void fill_u32(u32* dst, u32 value, size_t n, Stream* st) {
  if (!n) return;
  KLAUNCH(fill_u32_k, grid_for((n + 3) / 4, 256), 256, st, dst, value, n);
}
void fill_u8(u8* dst, u8 value, size_t n, Stream* st) {
  if (!n) return;
  KLAUNCH(fill_u8_k, grid_for((n + 15) / 16, 256), 256, st, dst, value, n);
}
struct Marker {
  hipEvent_t e;
};
Marker* marker_record(Stream* st) {
  Marker* m = new Marker();
  PFS_HIP(hipEventCreateWithFlags(&m->e, hipEventDisableTiming));
  PFS_HIP(hipEventRecord(m->e, st->s));
  return m;
}
bool marker_done(Marker* m) {
  const hipError_t e = hipEventQuery(m->e);
  if (e == hipErrorNotReady) return false;
  if (e != hipSuccess) (void)hipGetLastError();
  return true;
}
void marker_wait(Marker* m) {
  if (hipEventSynchronize(m->e) != hipSuccess) (void)hipGetLastError();
}
void marker_free(Marker* m) {
  if (!m) return;
  (void)hipEventDestroy(m->e);
  delete m;
}
bool stream_wait(Stream* st) {
  if (hipStreamSynchronize(st->s) == hipSuccess) return true;
  (void)hipGetLastError();
  return false;
}
void dev_mem_info(size_t* fr, size_t* tot) {
  if (hipMemGetInfo(fr, tot) != hipSuccess) {
    (void)hipGetLastError();
    *fr = *tot = 0;
  }
}
Event* event_record(Stream* st) {
  Event* e = new Event();
  PFS_HIP(hipEventCreate(&e->e));
  PFS_HIP(hipEventRecord(e->e, st->s));
  return e;
}
float event_ms(Event* a, Event* b) {
  PFS_HIP(hipEventSynchronize(b->e));
  float ms = 0;
  PFS_HIP(hipEventElapsedTime(&ms, a->e, b->e));
  return ms;
}
void event_free(Event* e) {
  if (!e) return;
  (void)hipEventDestroy(e->e);
  delete e;
}
const char* device_name(int device) {
  static hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) return "unknown";
  return prop.name;
}
int device_count() {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

// ---------------------------------------------------------------- block scan
constexpr int ST = 256;          // threads per scan workgroup
constexpr int SIPT = 8;          // items per thread
constexpr int STILE = ST * SIPT; // items per workgroup
constexpr int NW = ST / 64;      // wavefronts per workgroup

// exclusive scan of one value per thread over the workgroup; *total gets
// the workgroup sum.  'lw' holds NW words of LDS.
template <int W>  // wavefronts per workgroup
__device__ inline u32 block_excl_scan_w(u32 v, u32* lw, u32* total) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  u32 x = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const u32 y = __shfl_up(x, d, 64);
    if (lane >= d) x += y;
  }
  if (lane == 63) lw[wid] = x;
  __syncthreads();
  u32 pre = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < W; w++) {
    const u32 s = lw[w];
    pre += (w < wid) ? s : 0u;
    tot += s;
  }
  __syncthreads();
  *total = tot;
  return pre + x - v;
}
__device__ inline u32 block_excl_scan(u32 v, u32* lw, u32* total) { return block_excl_scan_w<NW>(v, lw, total); }

// order-free compaction: each thread holds the keep bits of SIPT
// consecutive items; the workgroup reserves its slots with one atomic and
// returns each thread's first slot
__device__ inline u32 block_reserve(u32 bits, u32* lw, u32* counter) {
  __shared__ u32 base;
  u32 tot;
  const u32 pre = block_excl_scan(__popc(bits), lw, &tot);
  if (threadIdx.x == 0) base = tot ? atomicAdd(counter, tot) : 0u;
  __syncthreads();
  return base + pre;
}
// grid-stride reductions: a bounded number of workgroups, one atomic per
// wavefront, so a single result word never sees more than a few thousand
// atomics
constexpr unsigned RED_BLOCKS = 1024;
__device__ inline u32 wave_max(u32 x) {
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) x = max(x, (u32)__shfl_xor(x, d, 64));
  return x;
}
__device__ inline unsigned long long wave_sum64(unsigned long long x) {
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) x += __shfl_xor(x, d, 64);
  return x;
}
// Long scans: each workgroup covers SCAN_SUB tiles of STILE items (one
// partial sum per workgroup, so the single-workgroup pass over the partials
// stays short); a thread's SIPT consecutive items move as two 16-byte
// loads and stores when the arrays are 16-byte aligned (VEC).
constexpr int SCAN_SUB = 4;
constexpr size_t SBLK = (size_t)STILE * SCAN_SUB;
template <bool VEC>
__device__ inline void load_items(const u32* in, size_t b, size_t n, u32 (&v)[SIPT]) {
  if (VEC && b + SIPT <= n) {
    const uint4 x = *(const uint4*)(in + b), y = *(const uint4*)(in + b + 4);
    v[0] = x.x, v[1] = x.y, v[2] = x.z, v[3] = x.w, v[4] = y.x, v[5] = y.y, v[6] = y.z, v[7] = y.w;
    return;
  }
#pragma unroll
  for (int i = 0; i < SIPT; i++) v[i] = (b + i < n) ? in[b + i] : 0u;
}
template <bool VEC>
__device__ inline void store_items(u32* out, size_t b, size_t n, const u32 (&v)[SIPT]) {
  if (VEC && b + SIPT <= n) {
    *(uint4*)(out + b) = make_uint4(v[0], v[1], v[2], v[3]);
    *(uint4*)(out + b + 4) = make_uint4(v[4], v[5], v[6], v[7]);
    return;
  }
#pragma unroll
  for (int i = 0; i < SIPT; i++)
    if (b + i < n) out[b + i] = v[i];
}
static_assert(SIPT == 8, "two 16-byte moves per thread");
template <bool VEC>
__global__ void scan_reduce_k(const u32* in, size_t n, u32* part) {
  __shared__ u32 lw[NW];
  u32 s = 0;
#pragma unroll
  for (int q = 0; q < SCAN_SUB; q++) {
    u32 v[SIPT];
    load_items<VEC>(in, (size_t)blockIdx.x * SBLK + (size_t)q * STILE + (size_t)threadIdx.x * SIPT, n, v);
#pragma unroll
    for (int i = 0; i < SIPT; i++) s += v[i];
  }
  u32 tot;
  block_excl_scan(s, lw, &tot);
  if (threadIdx.x == 0) part[blockIdx.x] = tot;
}
// exclusive scan of the workgroup partials by one workgroup
__global__ void scan_top_k(u32* part, size_t nb, u32* total) {
  __shared__ u32 lw[NW];
  u32 carry = 0;
  for (size_t base = 0; base < nb; base += STILE) {
    const size_t b = base + (size_t)threadIdx.x * SIPT;
    u32 v[SIPT];
    u32 s = 0;
#pragma unroll
    for (int i = 0; i < SIPT; i++) {
      v[i] = (b + i < nb) ? part[b + i] : 0u;
      s += v[i];
    }
    u32 tot;
    u32 pre = block_excl_scan(s, lw, &tot) + carry;
#pragma unroll
    for (int i = 0; i < SIPT; i++) {
      if (b + i < nb) part[b + i] = pre;
      pre += v[i];
    }
    carry += tot;
  }
  if (threadIdx.x == 0 && total) *total = carry;
}
// each thread reads its items before the tile's barriers and writes them
// after, so in == out is fine
template <bool VEC>
__global__ void scan_down_k(const u32* in, u32* out, size_t n, const u32* part) {
  __shared__ u32 lw[NW];
  u32 carry = part[blockIdx.x];
#pragma unroll 1
  for (int q = 0; q < SCAN_SUB; q++) {
    const size_t base = (size_t)blockIdx.x * SBLK + (size_t)q * STILE + (size_t)threadIdx.x * SIPT;
    u32 v[SIPT];
    load_items<VEC>(in, base, n, v);
    u32 s = 0;
#pragma unroll
    for (int i = 0; i < SIPT; i++) s += v[i];
    u32 tot;
    u32 pre = block_excl_scan(s, lw, &tot) + carry;
#pragma unroll
    for (int i = 0; i < SIPT; i++) {
      const u32 x = v[i];
      v[i] = pre;
      pre += x;
    }
    store_items<VEC>(out, base, n, v);
    carry += tot;
  }
}
// short scans in one launch: one workgroup of 1024 threads walks the input
// in tiles of 8192 items with a running carry (each thread reads its items
// before the tile's barriers and writes them after, so in == out is fine)
constexpr int S1_T = 1024;
constexpr size_t S1_MAX = 4 * (size_t)S1_T * SIPT;  // at most four tiles
__global__ void __launch_bounds__(S1_T) scan_one_k(const u32* in, u32* out, size_t n, u32* total) {
  __shared__ u32 lw[S1_T / 64];
  u32 carry = 0;
  for (size_t base = 0; base < n; base += (size_t)S1_T * SIPT) {
    const size_t b = base + (size_t)threadIdx.x * SIPT;
    u32 v[SIPT];
    u32 s = 0;
#pragma unroll
    for (int i = 0; i < SIPT; i++) {
      v[i] = (b + i < n) ? in[b + i] : 0u;
      s += v[i];
    }
    u32 tot;
    u32 pre = block_excl_scan_w<S1_T / 64>(s, lw, &tot) + carry;
#pragma unroll
    for (int i = 0; i < SIPT; i++) {
      if (b + i < n) out[b + i] = pre;
      pre += v[i];
    }
    carry += tot;
  }
  if (threadIdx.x == 0 && total) *total = carry;
}
size_t scan_tmp_bytes(size_t n) { return (grid_for(n, SBLK) + 64) * sizeof(u32); }
void scan_u32(const u32* in, u32* out, size_t n, u32* total, void* tmp, Stream* st) {
  if (n == 0) {
    if (total) fill_u32(total, 0, 1, st);
    return;
  }
  if (n <= S1_MAX) {
    KLAUNCH(scan_one_k, 1, S1_T, st, in, out, n, total);
    return;
  }
  const unsigned nb = grid_for(n, SBLK);
  u32* part = (u32*)tmp;
  if ((((uintptr_t)in | (uintptr_t)out) & 15) == 0) {
    KLAUNCH(scan_reduce_k<true>, nb, ST, st, in, n, part);
    KLAUNCH(scan_top_k, 1, ST, st, part, nb, total);
    KLAUNCH(scan_down_k<true>, nb, ST, st, in, out, n, part);
  } else {
    KLAUNCH(scan_reduce_k<false>, nb, ST, st, in, n, part);
    KLAUNCH(scan_top_k, 1, ST, st, part, nb, total);
    KLAUNCH(scan_down_k<false>, nb, ST, st, in, out, n, part);
  }
}

// ---------------------------------------------------------------- stable radix sort
// LSD passes of up to 8 bits.  A pass counts each tile's digits (rs_count_k),
// scans the digit-major counts, then every tile sorts itself by the digit
// in LDS with stable one-bit splits and writes each digit's run to its
// place (rs_scatter_k), so the stores go out in runs rather than one
// scattered word per item.
constexpr int RS_T = 256, RS_IPT = 8, RS_TILE = RS_T * RS_IPT;  // threads, items per thread, items per tile
constexpr int RB = 8, RBUCK = 1 << RB;
static_assert(RS_T == RBUCK, "one digit bin per thread");

__global__ void __launch_bounds__(RS_T) rs_count_k(const u32* keys, size_t n, int shift, u32 mask, u32* counts,
                                                  u32 nb) {
  __shared__ u32 h[RBUCK];
  h[threadIdx.x] = 0;  // RS_T == RBUCK
  __syncthreads();
  const size_t base = (size_t)blockIdx.x * RS_TILE;
#pragma unroll
  for (int i = 0; i < RS_IPT; i++) {
    const size_t j = base + (size_t)i * RS_T + threadIdx.x;
    if (j < n) atomicAdd(&h[(keys[j] >> shift) & mask], 1u);
  }
  __syncthreads();
  counts[(size_t)threadIdx.x * nb + blockIdx.x] = h[threadIdx.x];
}
__global__ void __launch_bounds__(RS_T) rs_scatter_k(const u32* kin, const u32* vin, u32* kout, u32* vout,
                                                    size_t n, int shift, int width, const u32* offs, u32 nb) {
  __shared__ u32 sk[2][RS_TILE], sv[2][RS_TILE];
  __shared__ u32 start[RBUCK];
  __shared__ u32 lw[RS_T / 64];
  const u32 mask = (1u << width) - 1u;
  const size_t base = (size_t)blockIdx.x * RS_TILE;
  const u32 cnt = (u32)min((size_t)RS_TILE, n - base);
  // the tile in input order; slots past the input take the largest digit
  // and stay behind every item (they come last and the splits are stable)
#pragma unroll
  for (int i = 0; i < RS_IPT; i++) {
    const u32 j = i * RS_T + threadIdx.x;
    sk[0][j] = j < cnt ? kin[base + j] : (mask << shift);
    sv[0][j] = j < cnt ? vin[base + j] : 0u;
  }
  start[threadIdx.x] = 0;
  __syncthreads();
  // one stable split per digit bit: thread t moves items t*IPT .. t*IPT+IPT-1
  int b = 0;
  for (int bit = 0; bit < width; bit++, b ^= 1) {
    const u32 j0 = threadIdx.x * RS_IPT;
    u32 k[RS_IPT], v[RS_IPT], z = 0;
#pragma unroll
    for (int i = 0; i < RS_IPT; i++) {
      k[i] = sk[b][j0 + i];
      v[i] = sv[b][j0 + i];
      z += ((k[i] >> (shift + bit)) & 1u) ^ 1u;
    }
    u32 zt;
    u32 zp = block_excl_scan_w<RS_T / 64>(z, lw, &zt);
    u32 op = j0 - zp;  // ones before this thread's items
#pragma unroll
    for (int i = 0; i < RS_IPT; i++) {
      const u32 dst = ((k[i] >> (shift + bit)) & 1u) ? zt + op++ : zp++;
      sk[b ^ 1][dst] = k[i];
      sv[b ^ 1][dst] = v[i];
    }
    __syncthreads();
  }
  // where each digit's run starts in the sorted tile
#pragma unroll
  for (int i = 0; i < RS_IPT; i++) {
    const u32 j = i * RS_T + threadIdx.x;
    if (j < cnt) atomicAdd(&start[(sk[b][j] >> shift) & mask], 1u);
  }
  __syncthreads();
  u32 tot;
  const u32 c = start[threadIdx.x];
  const u32 pre = block_excl_scan_w<RS_T / 64>(c, lw, &tot);
  start[threadIdx.x] = pre;
  __syncthreads();
#pragma unroll
  for (int i = 0; i < RS_IPT; i++) {
    const u32 j = i * RS_T + threadIdx.x;
    if (j < cnt) {
      const u32 key = sk[b][j], d = (key >> shift) & mask;
      const u32 dst = offs[(size_t)d * nb + blockIdx.x] + (j - start[d]);
      kout[dst] = key;
      vout[dst] = sv[b][j];
    }
  }
}
size_t radix_tmp_bytes(size_t n) {
  const size_t nb = grid_for(n ? n : 1, RS_TILE);
  return 2 * n * sizeof(u32) + RBUCK * nb * sizeof(u32) + scan_tmp_bytes(RBUCK * nb) + 256;
}
// sorts by the low 'bits' bits of the keys (higher bits are ignored)
void radix_sort_pairs(u32* keys, u32* vals, size_t n, int bits, void* tmp, Stream* st) {
  if (n < 2 || bits <= 0) return;
  bits = std::min(bits, 32);
  const u32 nb = grid_for(n, RS_TILE);
  u32* k2 = (u32*)tmp;
  u32* v2 = k2 + n;
  u32* counts = v2 + n;
  void* stmp = counts + (size_t)RBUCK * nb;
  u32 *ka = keys, *va = vals, *kb = k2, *vb = v2;
  int passes = 0;
  for (int shift = 0; shift < bits; shift += RB) {
    const int width = std::min(RB, bits - shift);
    const u32 mask = (1u << width) - 1u;
    KLAUNCH(rs_count_k, nb, RS_T, st, ka, n, shift, mask, counts, nb);
    scan_u32(counts, counts, (size_t)RBUCK * nb, nullptr, stmp, st);
    KLAUNCH(rs_scatter_k, nb, RS_T, st, ka, va, kb, vb, n, shift, width, counts, nb);
    u32* t = ka; ka = kb; kb = t;
    t = va; va = vb; vb = t;
    passes++;
  }
  if (passes & 1) {
    d2d(keys, ka, n * sizeof(u32), st);
    d2d(vals, va, n * sizeof(u32), st);
  }
}

// Stable sort of the clause order by the list key (size, first literal, last
// literal, signature): the reference's aggresivesort of the result
// (src/cpu/simplify.cpp:263 with CNF_CMP_KEY, simplify.h:29-42; the GPU
// build's stable sort with OLIST_CMP, src/gpu/cnf.cu:232-233).  Four stable
// radix passes over the handles, least significant key first.
__global__ void clause_key_k(const u32* pool, const u32* handles, u32 n, int which, u32* keys) {
  const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const u32 c = handles[i];
  const u32 size = hdr_size(pool[c]);
  u32 k;
  if (which == 0) k = pool[c + 1];                                // signature
  else if (which == 1) k = size ? pool[c + CL_HDR + size - 1] : 0;  // last literal
  else if (which == 2) k = size ? pool[c + CL_HDR] : 0;             // first literal
  else k = size;
  keys[i] = k;
}
void k_aggr_sort(const DB& db, u32* keys, void* tmp, Stream* st) {
  const u32 n = db.ncls;
  if (n < 2) return;
  for (int which = 0; which < 4; which++) {
    KLAUNCH(clause_key_k, grid_for(n, 256), 256, st, db.lits, db.c_off, n, which, keys);
    radix_sort_pairs(keys, db.c_off, n, which == 3 ? 24 : 32, tmp, st);
  }
}

// ---------------------------------------------------------------- OT
// clause-parallel kernels run over clause indices and work on the handles
// c_off[i]
__global__ void hist_k(DB db, u32* hist) {
  const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < db.ncls) hist_clause(db, db.c_off[i], hist);
}
__global__ void scatter_k(DB db, u32* cursor) {
  const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < db.ncls) scatter_clause(db, db.c_off[i], cursor);
}
void k_hist(const DB& db, u32* hist, Stream* st) {
  if (!db.ncls) return;
  KLAUNCH(hist_k, grid_for(db.ncls, 256), 256, st, db, hist);
}
void k_scatter(const DB& db, u32* cursor, Stream* st) {
  if (!db.ncls) return;
  KLAUNCH(scatter_k, grid_for(db.ncls, 256), 256, st, db, cursor);
}

// ---------------------------------------------------------------- OT by literal ranges
// The same table without a global atomic per occurrence: the (literal,
// handle) pairs of the live clauses are partitioned by literal range
// (1 << bits literals each, counted per range and tile of clauses, the
// counts scanned range-major), then one workgroup per range counts its
// literals in LDS, scans them into the list offsets (the range's first
// offset is where its partition starts, because the ranges are in literal
// order) and places the handles with LDS cursors.  Order inside a list is
// free, as with the per-occurrence atomics.
constexpr int OTP_UNROLL = 4;
constexpr int OTP_TB = 1024;  // threads per tile workgroup
constexpr u32 OTP_NONE = 0xFFFFFFFFu;  // no clause / no literal in this slot
// calls f(literal, handle) for every literal of the live clauses of the
// tile [c0, c1); each thread holds OTP_UNROLL clauses at a time
template <class F>
__device__ inline void otp_tile(const DB& db, u32 c0, u32 c1, F f) {
  for (u32 i = c0 + threadIdx.x; i < c1; i += OTP_UNROLL * OTP_TB) {
    u32 c[OTP_UNROLL], h[OTP_UNROLL];
#pragma unroll
    for (int q = 0; q < OTP_UNROLL; q++) {
      const u32 ii = i + q * OTP_TB;
      c[q] = ii < c1 ? db.c_off[ii] : OTP_NONE;
    }
#pragma unroll
    for (int q = 0; q < OTP_UNROLL; q++) h[q] = c[q] != OTP_NONE ? db.lits[c[q]] : 0u;
#pragma unroll
    for (int q = 0; q < OTP_UNROLL; q++) {
      if (c[q] == OTP_NONE || st_deleted(hdr_flags(h[q]))) continue;
      const u32* l = db.lits + c[q] + CL_HDR;
      const u32 n = hdr_size(h[q]);
      for (u32 k = 0; k < n; k++) f(l[k], c[q]);
    }
  }
}
// the tile of workgroup b: b itself, or with xcd the tiles taken in runs,
// one run per XCD (workgroups go to the 8 XCDs round robin, so XCD x runs
// the workgroups b = x mod 8; it gets tiles start(x) .. start(x + 1) - 1)
__device__ inline u32 otp_tile_of(u32 b, u32 nblk, int xcd) {
  if (!xcd) return b;
  const u32 q = nblk >> 3, r = nblk & 7u, x = b & 7u;
  return x * q + min(x, r) + (b >> 3);
}
__global__ void __launch_bounds__(OTP_TB) otp_count_k(DB db, u32 ct, int bits, u32 nr, u32 nblk, int xcd,
                                                      u32* counts) {
  extern __shared__ u32 lc[];
  for (u32 j = threadIdx.x; j < nr; j += OTP_TB) lc[j] = 0;
  __syncthreads();
  const u32 t = otp_tile_of(blockIdx.x, nblk, xcd);
  const u32 c0 = t * ct, c1 = min(db.ncls, c0 + ct);
  otp_tile(db, c0, c1, [&](u32 l, u32) { atomicAdd(&lc[l >> bits], 1u); });
  __syncthreads();
  for (u32 j = threadIdx.x; j < nr; j += OTP_TB) counts[(size_t)j * nblk + t] = lc[j];
}
// TABLE: pairs (literal, handle) in one 8-byte word each (one store per
// occurrence); otherwise the literals alone
template <bool TABLE>
__global__ void __launch_bounds__(OTP_TB) otp_part_k(DB db, u32 ct, int bits, u32 nr, u32 nblk, int xcd,
                                                     const u32* offs, void* out) {
  extern __shared__ u32 lc[];
  const u32 t = otp_tile_of(blockIdx.x, nblk, xcd);
  for (u32 j = threadIdx.x; j < nr; j += OTP_TB) lc[j] = offs[(size_t)j * nblk + t];
  __syncthreads();
  const u32 c0 = t * ct, c1 = min(db.ncls, c0 + ct);
  otp_tile(db, c0, c1, [&](u32 l, u32 c) {
    const u32 p = atomicAdd(&lc[l >> bits], 1u);
    if (TABLE)
      ((uint2*)out)[p] = make_uint2(l, c);
    else
      ((u32*)out)[p] = l;
  });
}
// one workgroup per literal range r: literals [r << bits, (r + 1) << bits)
// whose pairs sit at [offs[r * nblk], offs[(r + 1) * nblk]) (the scan's
// total past the last range).  Without TABLE only the counts (hist) are
// written.
constexpr int OTP_LD = 8;
// threads per range workgroup: 1024 (two workgroups per CU) against 256
// (five, bound by the 32 KiB of counters) took the heavy-tailed round's
// range kernel from 2.04 to 1.50 ms and Tseitin's from 0.35 to 0.28 ms:
// more loads in flight per range and fewer table slices open at once per
// XCD; reserving 64 KiB of LDS on top changed nothing
constexpr int OTP_RTB = 1024;
template <bool TABLE>
__global__ void __launch_bounds__(OTP_RTB) otp_range_k(DB db, int bits, u32 n2, u32 nblk, const u32* offs,
                                                  const u32* total, const void* in, u32* hist) {
  const u32* plit = (const u32*)in;
  const uint2* pair = (const uint2*)in;
  extern __shared__ u32 lc[];
  __shared__ u32 lw[OTP_RTB / 64];
  const u32 r = blockIdx.x, nl = 1u << bits, base = r << bits;
  const u32 b0 = offs[(size_t)r * nblk];
  const u32 b1 = r + 1 < gridDim.x ? offs[(size_t)(r + 1) * nblk] : *total;
  for (u32 j = threadIdx.x; j < nl; j += OTP_RTB) lc[j] = 0;
  __syncthreads();
  for (u32 p = b0 + threadIdx.x; p < b1; p += OTP_LD * OTP_RTB) {
    u32 l[OTP_LD];
#pragma unroll
    for (int q = 0; q < OTP_LD; q++)
      l[q] = p + q * OTP_RTB < b1 ? (TABLE ? pair[p + q * OTP_RTB].x : plit[p + q * OTP_RTB]) : OTP_NONE;
#pragma unroll
    for (int q = 0; q < OTP_LD; q++)
      if (l[q] != OTP_NONE) atomicAdd(&lc[l[q] - base], 1u);
  }
  __syncthreads();
  if (!TABLE) {
    for (u32 j = threadIdx.x; j < nl; j += OTP_RTB)
      if (base + j < n2) hist[base + j] = lc[j];
    return;
  }
  // exclusive scan: each thread owns nl / OTP_RTB consecutive counters
  const u32 per = nl / OTP_RTB, j0 = threadIdx.x * per;
  u32 s = 0;
  for (u32 k = 0; k < per; k++) s += lc[j0 + k];
  u32 tot;
  u32 pre = block_excl_scan_w<OTP_RTB / 64>(s, lw, &tot) + b0;
  for (u32 k = 0; k < per; k++) {
    const u32 v = lc[j0 + k];
    lc[j0 + k] = pre;
    pre += v;
  }
  __syncthreads();
  for (u32 j = threadIdx.x; j < nl; j += OTP_RTB)
    if (base + j < n2) db.ot_off[base + j] = lc[j];
  if (r + 1 == gridDim.x && threadIdx.x == 0) db.ot_off[n2] = *total;
  // every offset is read out of LDS before any cursor moves
  __syncthreads();
  for (u32 p = b0 + threadIdx.x; p < b1; p += OTP_LD * OTP_RTB) {
    u32 l[OTP_LD], c[OTP_LD];
#pragma unroll
    for (int q = 0; q < OTP_LD; q++) {
      const uint2 e = p + q * OTP_RTB < b1 ? pair[p + q * OTP_RTB] : make_uint2(OTP_NONE, 0u);
      l[q] = e.x;
      c[q] = e.y;
    }
#pragma unroll
    for (int q = 0; q < OTP_LD; q++)
      if (l[q] != OTP_NONE) db.ot_refs[atomicAdd(&lc[l[q] - base], 1u)] = c[q];
  }
  __syncthreads();
  // the cursors end at the next list's offset: size = cursor - offset
  for (u32 j = threadIdx.x; j < nl; j += OTP_RTB)
    if (base + j < n2) {
      const u32 sz = lc[j] - db.ot_off[base + j];
      db.ot_size[base + j] = sz;
      hist[base + j] = sz;
    }
}
bool otp_plan(size_t n2, size_t ncls, OtpPlan* p, int tile_log2, int xcd) {
  // LDS per workgroup: nr words in the tile kernels, 1 << bits words in
  // the range kernel (both within 64 KiB); ranges of 8192 literals once
  // ranges of 4096 would number 2048 or more
  int bits = 12;
  if ((n2 >> bits) >= 2048) bits = 13;
  const size_t nr = (n2 + ((size_t)1 << bits) - 1) >> bits;
  if (nr > OTP_MAX_RANGES || ncls == 0) return false;
  // tiles of 32 clauses per range (longer segments of pairs per range for
  // the partition's writes, fewer counts), at most 16,384 clauses: on the
  // heavy-tailed round (2,442 ranges) tiles of 16,384 clauses instead of
  // 131,072 gave 2,442 workgroups instead of 306 and took the count and
  // partition kernels from 0.56 + 2.69 to 0.38 + 2.24 ms
  size_t ct = 4096;
  while (ct < 32 * nr && ct < 16384) ct <<= 1;
  if (tile_log2 >= 10) ct = (size_t)1 << tile_log2;
  p->xcd = xcd;
  p->bits = bits;
  p->nr = (u32)nr;
  p->ct = (u32)ct;
  p->nblk = (u32)((ncls + ct - 1) / ct);
  p->ncounts = (size_t)p->nr * p->nblk;
  return true;
}
// counts: p.ncounts + 1 words; pairs: 8-byte aligned, two words per live
// occurrence (table) or one (hist only); tmp: scan_tmp_bytes(p.ncounts)
void k_ot_ranges(const DB& db, const OtpPlan& p, u32 n2, u32* counts, void* pairs, bool table, u32* hist,
                 void* tmp, Stream* st) {
  const size_t lds_r = (size_t)p.nr * 4, lds_b = ((size_t)4 << p.bits);
  KLAUNCH_LDS(otp_count_k, p.nblk, OTP_TB, lds_r, st, db, p.ct, p.bits, p.nr, p.nblk, p.xcd, counts);
  scan_u32(counts, counts, p.ncounts, counts + p.ncounts, tmp, st);
  if (table) {
    KLAUNCH_LDS(otp_part_k<true>, p.nblk, OTP_TB, lds_r, st, db, p.ct, p.bits, p.nr, p.nblk, p.xcd, counts, pairs);
    KLAUNCH_LDS(otp_range_k<true>, p.nr, OTP_RTB, lds_b, st, db, p.bits, n2, p.nblk, counts, counts + p.ncounts, pairs,
                hist);
  } else {
    KLAUNCH_LDS(otp_part_k<false>, p.nblk, OTP_TB, lds_r, st, db, p.ct, p.bits, p.nr, p.nblk, p.xcd, counts,
                pairs);
    KLAUNCH_LDS(otp_range_k<false>, p.nr, OTP_RTB, lds_b, st, db, p.bits, n2, p.nblk, counts, counts + p.ncounts,
                pairs, hist);
  }
}

constexpr u32 SORT_SMALL = 16;
constexpr u32 SORT_BLOCK = 2048;
// wave-aggregated append: one atomic per wavefront per counter
__device__ inline void wave_append(bool pred, u32 val, u32* arr, u32* cnt) {
  const u64 mask = __ballot(pred);
  if (!mask) return;
  const int lane = threadIdx.x & 63;
  const int leader = __ffsll((unsigned long long)mask) - 1;
  u32 base = 0;
  if (lane == leader) base = atomicAdd(cnt, (u32)__popcll(mask));
  base = __shfl(base, leader, 64);
  if (pred) arr[base + __popcll(mask & ((1ull << lane) - 1))] = val;
}
// Lists of LO < len <= W entries (every length up to W when LO is 0)
// sorted by a group of W lanes: bitonic network over wave64 shuffles, keys
// in registers.  Also leaves in lsum[g] the sum of the clause sizes of such
// a list g (a bound for later stages: the lists of elected variables only
// shrink).  Returns the list's length.
template <u32 W, u32 LO>
__device__ inline u32 sort_group(const DB& db, u32 lit, bool valid, u32 j, u32 g, u32* lsum) {
  const u32 len0 = valid ? db.ot_size[lit] : 0;
  const bool mine = valid && len0 <= W && (LO == 0 || len0 > LO);
  const u32 small = mine ? len0 : 0;
  const u32 len = small >= 2 ? small : 0;
  u32* r = db.ot_refs + (small ? db.ot_off[lit] : 0);
  u64 a = ~0ull, b = ~0ull;
  u32 id = 0xFFFFFFFFu, sz = 0;
  if (j < small) {
    const SortKey k = make_key(db, r[j]);
    sz = k.size;
    if (len) {
      a = ((u64)k.size << 32) | k.lit0;
      b = ((u64)k.back << 32) | k.sig;
      id = k.id;
    }
  }
#pragma unroll
  for (u32 d = W / 2; d > 0; d >>= 1) sz += __shfl_xor(sz, d, 64);
  if (j == 0 && mine) lsum[g] = sz;
#pragma unroll
  for (u32 k = 2; k <= W; k <<= 1) {
#pragma unroll
    for (u32 d = k >> 1; d > 0; d >>= 1) {
      const u64 pa = __shfl_xor(a, d, 64), pb = __shfl_xor(b, d, 64);
      const u32 pid = __shfl_xor(id, d, 64);
      const bool mine_less = a != pa ? a < pa : (b != pb ? b < pb : id < pid);
      const bool up = (j & k) == 0, lower = (j & d) == 0;
      const bool keep_min = lower == up;
      if (keep_min != mine_less) {
        a = pa;
        b = pb;
        id = pid;
      }
    }
  }
  if (j < len) r[j] = id;
  return len0;
}
// Every list once by a 4-lane group (wavefront count and dependent loads
// per list, not the sorting, bound these passes): the lists of at most
// SORT_TINY entries, most of an election's, are sorted there; the others
// are flagged (fmid: up to SORT_SMALL entries, flong: beyond), compacted by
// scans and sorted by 16-lane groups, per workgroup or by the long-list
// merge.  (Tried: one 16-lane pass over every list, 1.09 ms on the
// heavy-tailed round's 6.1M lists; a worklist filled with one atomic per
// wavefront on a shared counter; a second pass over every list that skips
// the short ones; a bounded grid striding over the lists.)
constexpr u32 SORT_TINY = 4;
__global__ void sort_tiny_k(DB db, const u32* lits, u32 n, u32* lsum, u32* fmid, u32* flong) {
  const u32 t = blockIdx.x * blockDim.x + threadIdx.x;
  const u32 g = t / SORT_TINY, j = t % SORT_TINY;
  const bool valid = g < n;
  const u32 len = sort_group<SORT_TINY, 0>(db, valid ? lits[g] : 0u, valid, j, g, lsum);
  if (valid && j == 0) {
    fmid[g] = len > SORT_TINY && len <= SORT_SMALL ? 1u : 0u;
    flong[g] = len > SORT_SMALL ? 1u : 0u;
  }
}
// the lists mid[0..*nmid), 16 lanes each; a bounded grid strides over them
// (the loop bound is the same for every lane of a wavefront)
__global__ void sort_mid_k(DB db, const u32* lits, const u32* mid, const u32* nmid, u32* lsum) {
  const u64 m = (u64)*nmid * SORT_SMALL;
  for (u64 base = (u64)blockIdx.x * blockDim.x; base < m; base += (u64)gridDim.x * blockDim.x) {
    const u64 t = base + threadIdx.x;
    const bool valid = t < m;
    const u32 g = valid ? mid[t / SORT_SMALL] : 0u;
    sort_group<SORT_SMALL, SORT_TINY>(db, valid ? lits[g] : 0u, valid, (u32)(t % SORT_SMALL), g, lsum);
  }
}
// bitonic sort of one list per workgroup in LDS (16 < len <= SORT_BLOCK);
// big: positions in lits
__device__ void sort_block_list(DB db, const u32* lits, u32 g, u32* lsum, u64* ka, u64* kb, u32* kid, u32* lw) {
  const u32 lit = lits[g];
  const u32 len = db.ot_size[lit];
  if (len <= SORT_SMALL || len > SORT_BLOCK) return;
  u32* r = db.ot_refs + db.ot_off[lit];
  u32 N = 1;
  while (N < len) N <<= 1;
  u32 sz = 0;
  for (u32 i = threadIdx.x; i < N; i += blockDim.x) {
    if (i < len) {
      const SortKey k = make_key(db, r[i]);
      ka[i] = ((u64)k.size << 32) | k.lit0;
      kb[i] = ((u64)k.back << 32) | k.sig;
      kid[i] = k.id;
      sz += k.size;
    } else {
      ka[i] = ~0ull;
      kb[i] = ~0ull;
      kid[i] = 0xFFFFFFFFu;
    }
  }
  u32 tot;
  block_excl_scan(sz, lw, &tot);
  if (threadIdx.x == 0) lsum[g] = tot;
  for (u32 k = 2; k <= N; k <<= 1) {
    for (u32 j = k >> 1; j > 0; j >>= 1) {
      for (u32 i = threadIdx.x; i < N; i += blockDim.x) {
        const u32 ixj = i ^ j;
        if (ixj > i) {
          const bool up = (i & k) == 0;
          const u64 a0 = ka[i], a1 = ka[ixj], b0 = kb[i], b1 = kb[ixj];
          const u32 i0 = kid[i], i1 = kid[ixj];
          const bool gt = a0 != a1 ? a0 > a1 : (b0 != b1 ? b0 > b1 : i0 > i1);
          if (gt == up) {
            ka[i] = a1; ka[ixj] = a0;
            kb[i] = b1; kb[ixj] = b0;
            kid[i] = i1; kid[ixj] = i0;
          }
        }
      }
      __syncthreads();
    }
  }
  for (u32 i = threadIdx.x; i < len; i += blockDim.x) r[i] = kid[i];
}
// big: positions in lits, *nbig of them (read on the device: a bounded grid
// strides over them, so the launch needs no host read of the count)
__global__ void __launch_bounds__(256) sort_block_k(DB db, const u32* lits, const u32* big, const u32* nbig,
                                                    u32* lsum) {
  __shared__ u64 ka[SORT_BLOCK];
  __shared__ u64 kb[SORT_BLOCK];
  __shared__ u32 kid[SORT_BLOCK];
  __shared__ u32 lw[NW];
  const u32 nb = *nbig;
  for (u32 b = blockIdx.x; b < nb; b += gridDim.x) {
    sort_block_list(db, lits, big[b], lsum, ka, kb, kid, lw);
    __syncthreads();
  }
}
// bottom-up merge sort of one long list by one thread (rare lists above
// SORT_BLOCK entries); 'scratch' has one slot per occurrence entry
__device__ void sort_large_list(DB db, const u32* lits, u32 g, u32* scratch, u32* lsum) {
  const u32 lit = lits[g];
  const u32 len = db.ot_size[lit];
  if (len <= SORT_BLOCK) return;
  u32* a = db.ot_refs + db.ot_off[lit];
  u64 sz = 0;
  for (u32 k = 0; k < len; k++) sz += csize(db, a[k]);
  lsum[g] = sz > 0xFFFFFFFFull ? 0xFFFFFFFFu : (u32)sz;
  u32* b = scratch + db.ot_off[lit];
  u32 *src = a, *dst = b;
  for (u32 w = 1; w < len; w <<= 1) {
    for (u32 lo = 0; lo < len; lo += 2 * w) {
      const u32 mid = lo + w < len ? lo + w : len;
      const u32 hi = lo + 2 * w < len ? lo + 2 * w : len;
      u32 p = lo, q = mid, o = lo;
      while (p < mid && q < hi) {
        if (key_less(make_key(db, src[q]), make_key(db, src[p]))) dst[o++] = src[q++];
        else dst[o++] = src[p++];
      }
      while (p < mid) dst[o++] = src[p++];
      while (q < hi) dst[o++] = src[q++];
    }
    u32* t = src; src = dst; dst = t;
  }
  if (src != a)
    for (u32 k = 0; k < len; k++) a[k] = src[k];
}
// grid-stride over the *nbig collected lists (count read on the device)
__global__ void sort_large_k(DB db, const u32* lits, const u32* big, const u32* nbig, u32* scratch, u32* lsum) {
  const u32 nb = *nbig;
  for (u32 i = blockIdx.x * blockDim.x + threadIdx.x; i < nb; i += gridDim.x * blockDim.x)
    sort_large_list(db, lits, big[i], scratch, lsum);
}
void k_sort_lists(const DB& db, const u32* lits, u32 n, u32* classes, u32* cnt, u32* scratch, u32* lsum,
                  void* tmp, Stream* st) {
  if (!n) return;
  u32* big = classes;           // lists longer than SORT_SMALL
  u32* mid = classes + n;       // lists of SORT_TINY < len <= SORT_SMALL
  u32* fmid = classes + 2 * (size_t)n;
  u32* flong = classes + 3 * (size_t)n;
  u32* pos = classes + 4 * (size_t)n;
  KLAUNCH(sort_tiny_k, grid_for((size_t)n * SORT_TINY, 256), 256, st, db, lits, n, lsum, fmid, flong);
  k_pack_flags(fmid, n, pos, mid, cnt + 1, tmp, st);
  k_pack_flags(flong, n, pos, big, cnt, tmp, st);
  // the collected lists' launches read their counts on the device
  KLAUNCH(sort_mid_k, std::min(grid_for((size_t)n * SORT_SMALL, 256), 4096u), 256, st, db, lits, mid, cnt + 1, lsum);
  KLAUNCH(sort_block_k, std::min(grid_for(n, 1), 1024u), 256, st, db, lits, big, cnt, lsum);
  KLAUNCH(sort_large_k, std::min(grid_for(n, 64), 64u), 64, st, db, lits, big, cnt, scratch, lsum);
}

// ---------------------------------------------------------------- BCP
// Kernels that walk the occurrence lists of a few literals (units): one
// row of workgroups per literal (blockIdx.y) strides over its list, so that
// a handful of hot literals (lists of 10^4-10^5 entries on heavy-tailed
// formulas) does not leave the device to one workgroup each.  The launcher
// aims at about 4096 workgroups in all and at most 65535 literals per
// launch (the grid's y extent).
template <class L>
void launch_per_literal(u32 n, L launch) {
  for (u32 c0 = 0; c0 < n; c0 += 65535) {
    const u32 m = std::min<u32>(n - c0, 65535);
    launch(dim3(std::max<u32>(1, std::min<u32>(64, 4096 / m)), m), c0);
  }
}
__device__ inline u32 list_lane() { return blockIdx.x * blockDim.x + threadIdx.x; }
__device__ inline u32 list_stride() { return gridDim.x * blockDim.x; }
__global__ void bcp_visit_k(DB db, const u32* front, u32* next, u32* nnext, u32* conflict) {
  const u32 t = front[blockIdx.y];
  const OLV ol = olist(db, FLIP(t));
  for (u32 i = list_lane(); i < ol.n; i += list_stride()) bcp_visit(db, ol.r[i], next, nnext, conflict);
}
void k_bcp_visit(const DB& db, const u32* front, u32 nfront, u32* next, u32* nnext, u32* conflict,
                 Stream* st) {
  launch_per_literal(nfront, [&](dim3 grid, u32 c0) {
    KLAUNCH(bcp_visit_k, grid, 256, st, db, front + c0, next, nnext, conflict);
  });
}
__global__ void bcp_mark_k(DB db, const u32* lits, u8* mark) {
  const u32 t = lits[blockIdx.y];
  for (u32 q = 0; q < 2; q++) {
    const OLV ol = olist(db, t ^ q);
    for (u32 i = list_lane(); i < ol.n; i += list_stride()) mark[ol.r[i]] = 1;
  }
}
void k_bcp_mark(const DB& db, const u32* lits, u32 n, u8* mark, Stream* st) {
  launch_per_literal(n, [&](dim3 grid, u32 c0) { KLAUNCH(bcp_mark_k, grid, 256, st, db, lits + c0, mark); });
}
// mark: per handle
__global__ void bcp_clean_k(DB db, u8* mark) {
  const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= db.ncls) return;
  const u32 c = db.c_off[i];
  if (!mark[c]) return;
  mark[c] = 0;
  bcp_clean(db, c);
}
void k_bcp_clean(const DB& db, u8* mark, Stream* st) {
  if (!db.ncls) return;
  KLAUNCH(bcp_clean_k, grid_for(db.ncls, 256), 256, st, db, mark);
}
__global__ void clear_lists_k(DB db, const u32* lits, u32 n) {
  const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const u32 t = lits[i];
  db.ot_size[t] = 0;
  db.ot_size[FLIP(t)] = 0;
}
void k_clear_lists(const DB& db, const u32* lits, u32 n, Stream* st) {
  if (!n) return;
  KLAUNCH(clear_lists_k, grid_for(n, 256), 256, st, db, lits, n);
}
__global__ void set_vstate_k(DB db, const u32* lits, u32 n, u8 state) {
  const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) db.vstate[ABS(lits[i])] = state;
}
void k_set_vstate(const DB& db, const u32* lits, u32 n, u8 state, Stream* st) {
  if (!n) return;
  KLAUNCH(set_vstate_k, grid_for(n, 256), 256, st, db, lits, n, state);
}

// ---------------------------------------------------------------- LCVE
__device__ unsigned long long lcve_classify_var(const DB& db, const LcveArgs& a, u32 v);
__device__ inline unsigned long long wave_min64(unsigned long long x) {
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) {
    const unsigned long long y = __shfl_xor(x, d, 64);
    x = y < x ? y : x;
  }
  return x;
}
// the hard-break key is a minimum over the grid: per lane over a strided
// range of variables, then over the workgroup, then one atomic per
// workgroup into a bounded grid (one atomic per wavefront over 10M
// variables serialised 1 ms on the one word)
__global__ void lcve_classify_k(DB db, LcveArgs a) {
  __shared__ unsigned long long wmin[ST / 64];
  unsigned long long hk = ~0ull;
  for (u32 v = blockIdx.x * blockDim.x + threadIdx.x + 1; v <= db.nvars; v += gridDim.x * blockDim.x) {
    const unsigned long long k = lcve_classify_var(db, a, v);
    hk = k < hk ? k : hk;
  }
  hk = wave_min64(hk);
  if ((threadIdx.x & 63) == 0) wmin[threadIdx.x >> 6] = hk;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (u32 w = 1; w < ST / 64; w++) hk = wmin[w] < hk ? wmin[w] : hk;
    if (hk != ~0ull) at_min64(a.hard, hk);
  }
}
__device__ unsigned long long lcve_classify_var(const DB& db, const LcveArgs& a, u32 v) {
  const u32 ps = a.hist[V2L(v)], ns = a.hist[V2L(v) | 1];
  const u32 score = ps * ns;  // 32-bit product as in the reference
  a.score[v] = score;
  const bool cand = !db.assumed[v] && !db.vstate[v] && (ps || ns);
  // break rules (lcve.cpp:61-66; the GPU build's electionsmax, lcve.cu:94)
  const bool brk = (db.ot_size[V2L(v)] >= a.pt && db.ot_size[V2L(v) | 1] >= a.nt) || ps > a.max_occurs ||
                   ns > a.max_occurs;
  // a breaker that no neighbour can freeze (the freeze rule spares variables
  // at or above both temperatures) ends the election wherever it stands
  const bool hard = brk && ps >= a.pt && ns >= a.nt;
  u8 f = cand ? (brk ? 3 : 1) : 0;
  if (f == 1 && a.clause_max != 0xFFFFFFFFu && lcve_oversize(db, v, a.clause_max)) f = 4;
  a.cflag[v] = f;
  return cand && hard ? ((unsigned long long)score << 32) | v : ~0ull;
}
void k_lcve_classify(const DB& db, const LcveArgs& a, Stream* st) {
  KLAUNCH(lcve_classify_k, std::min(grid_for(db.nvars, ST), 2048u), ST, st, db, a);
}
__global__ void lcve_flag_k(DB db, LcveArgs a, u32* flag) {
  const u32 v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v > db.nvars) return;
  u32 f = 0;
  if (v && a.cflag[v]) f = (((u64)a.score[v] << 32) | v) < *a.hard;
  flag[v] = f;
}
__global__ void lcve_pack_k(DB db, LcveArgs a, const u32* flag, const u32* pos, u32* cand, u32* key) {
  const u32 v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v > db.nvars || !v) return;
  if (!a.cflag[v] || !((((u64)a.score[v] << 32) | v) < *a.hard)) return;
  const u32 i = pos[v];
  cand[i] = v;
  key[i] = a.score[v];
}
void k_lcve_select(const DB& db, const LcveArgs& a, u32* flag_tmp, u32* cand, u32* key, u32* ncand,
                   void* tmp, Stream* st) {
  const size_t n = (size_t)db.nvars + 1;
  KLAUNCH(lcve_flag_k, grid_for(n, 256), 256, st, db, a, flag_tmp);
  scan_u32(flag_tmp, flag_tmp, n, ncand, tmp, st);
  KLAUNCH(lcve_pack_k, grid_for(n, 256), 256, st, db, a, flag_tmp, flag_tmp, cand, key);
}
// packed position and state (undecided) of each candidate, for mis_step
__global__ void lcve_pos_k(const u32* cand, u32 n, u32* pos) {
  const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) pos[cand[i]] = i << 2;
}
void k_lcve_positions(const u32* cand, u32 n, u32* pos, Stream* st) {
  if (!n) return;
  KLAUNCH(lcve_pos_k, grid_for(n, 256), 256, st, cand, n, pos);
}
__global__ void lcve_reject_k(const u32* cand, u32 n, const u8* cflag, u32* pos, u32* state) {
  const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || cflag[cand[i]] != 4) return;
  mis_decide(state, pos, i, cand[i], 2);
}
void k_lcve_reject(const u32* cand, u32 n, const u8* cflag, u32* pos, u32* state, Stream* st) {
  if (!n) return;
  KLAUNCH(lcve_reject_k, grid_for(n, 256), 256, st, cand, n, cflag, pos, state);
}
// One election step by a group of G lanes per candidate (G a power of two,
// groups aligned in the wavefront): the lanes share out the entries of the
// candidate's two lists (positive, then negative: one combined index) and
// combine their findings, an elected earlier neighbour anywhere freezes
// the candidate, else the earliest entry holding an undecided earlier
// neighbour is its blocker's entry.  The outcome is the one-lane scan's
// (mis_step); G lanes keep G chains of loads in flight per candidate.
// Values that steer the group's control flow are read by its first lane and
// broadcast, so every lane of a group takes the same path.  Returns 1 if
// the candidate stays undecided (in every lane of the group).
template <u32 G>
__device__ inline u32 mis_step_group(const DB& db, u32 i, bool valid, const u32* cand, u32* pw, u32* state, u32* blk,
                                     u32* resume) {
  const int lane = threadIdx.x & 63, lead = lane & ~(int)(G - 1), j = lane - lead;
  u32 s0 = (valid && j == 0) ? at_load(&state[i]) : 0u;
  s0 = __shfl(s0, lead, 64);
  if (!valid || s0) return 0;
  const u32 v = cand[i];
  const u32 b = blk[i];
  if (b != NOPOS) {
    u32 sb = j == 0 ? at_load(&state[b]) : 0u;
    sb = __shfl(sb, lead, 64);
    if (sb == 0) return 1;
    if (sb == 1) {
      if (j == 0) mis_decide(state, pw, i, v, 2);
      return 0;
    }
  }
  const u32 r = resume[i];
  const OLV o0 = olist(db, V2L(v)), o1 = olist(db, V2L(v) | 1);
  const u32 n0 = o0.n, ntot = o0.n + o1.n;
  const u32 start = (r >> 31) ? n0 + (r & 0x7FFFFFFFu) : r;
  u32 first = NOPOS, bpos = NOPOS;  // this lane's earliest entry with an undecided earlier neighbour
  bool frozen = false;
  for (u32 e = start + j; e < ntot && !frozen; e += G) {
    const u32 c = e < n0 ? o0.r[e] : o1.r[e - n0];
    const u32 h = db.lits[c];
    if (st_deleted(hdr_flags(h))) continue;
    const u32* l = db.lits + c + CL_HDR;
    const u32 n = hdr_size(h);
    for (u32 q = 0; q < n && !frozen; q += 4) {
      u32 u[4], w[4];
#pragma unroll
      for (u32 k = 0; k < 4; k++) u[k] = q + k < n ? ABS(l[q + k]) : v;
#pragma unroll
      for (u32 k = 0; k < 4; k++) w[k] = u[k] != v ? at_load(&pw[u[k]]) : NOPOS;
#pragma unroll
      for (u32 k = 0; k < 4; k++) {
        const u32 pu = w[k] >> 2;  // NOPOS: past every candidate
        if (pu >= i) continue;
        const u32 su = w[k] & 3u;
        if (su == 1) frozen = true;
        if (su == 0 && first == NOPOS) first = e, bpos = pu;
      }
    }
  }
  u32 fz = frozen ? 1u : 0u;
  u64 m = ((u64)first << 32) | bpos;
#pragma unroll
  for (u32 d = 1; d < G; d <<= 1) {
    fz |= __shfl_xor(fz, d, 64);
    const u64 o = __shfl_xor(m, d, 64);
    m = o < m ? o : m;
  }
  if (fz) {
    if (j == 0) mis_decide(state, pw, i, v, 2);
    return 0;
  }
  const u32 e = (u32)(m >> 32);
  if (e != NOPOS) {
    if (j == 0) {
      blk[i] = (u32)m;
      resume[i] = e < n0 ? e : ((1u << 31) | (e - n0));
    }
    return 1;
  }
  if (j == 0) mis_decide(state, pw, i, v, 1);
  return 0;
}

// one election step over the undecided candidates idx[0..n) (all when idx
// is null), G lanes per candidate; the ones still undecided are appended
// to 'next'.  The launch reserves LDS per one-wavefront workgroup only to
// bound occupancy (option "mislds"; 20 KiB leaves 2 wavefronts per SIMD:
// Tseitin round 0 / 20 / 40 KiB -> mis_step_k 3.34 / 2.76 / 3.18 ms)
template <u32 G>
__global__ void mis_step_k(DB db, const u32* cand, const u32* idx, u32 n, u32* pos, u32* state, u32* blk,
                           u32* resume, u32* next, u32* nnext) {
  const u64 t = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  const u32 q = (u32)(t / G);
  const bool valid = q < n;
  const u32 i = valid ? (idx ? idx[q] : q) : 0u;
  const bool undecided = mis_step_group<G>(db, i, valid, cand, pos, state, blk, resume) != 0;
  wave_append(valid && (threadIdx.x & (G - 1)) == 0 && undecided, i, next, nnext);
}
// A worklist pass whose length is read on the device (n_in): a bounded grid
// strides over the list, so several passes are queued without a host read;
// passes over an empty list return at once.  iters counts the non-empty passes.
// zero: the counter the next pass appends to (the input counter of the
// pass before this one, so nothing reads it any more), cleared here instead
// of by a launch of its own before every pass
template <u32 G>
__global__ void mis_step_dev_k(DB db, const u32* cand, const u32* idx, const u32* n_in, u32* pos, u32* state,
                               u32* blk, u32* resume, u32* next, u32* nnext, u32* iters, u32* zero) {
  const u32 n = *n_in;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    *zero = 0;
    if (n) atomicAdd(iters, 1u);
  }
  const u64 m = (u64)n * G;
  for (u64 base = (u64)blockIdx.x * blockDim.x; base < m; base += (u64)gridDim.x * blockDim.x) {  // wave-uniform
    const u64 t = base + threadIdx.x;
    const u32 q = (u32)(t / G);
    const bool valid = q < n;
    const u32 i = valid ? idx[q] : 0u;
    const bool undecided = mis_step_group<G>(db, i, valid, cand, pos, state, blk, resume) != 0;
    wave_append(valid && (threadIdx.x & (G - 1)) == 0 && undecided, i, next, nnext);
  }
}
#define MIS_LANES(G, call) \
  switch (G) {             \
    case 2: call(2); break; \
    case 4: call(4); break; \
    case 8: call(8); break; \
    default: call(1); break; \
  }
void k_mis_step(const DB& db, const u32* cand, const u32* idx, u32 n, u32* pos, u32* state, u32* blk,
                u32* resume, u32* next, u32* nnext, u32 lds, u32 lanes, Stream* st) {
  if (!n) return;
#define MIS_FIRST(G) \
  KLAUNCH_LDS(mis_step_k<G>, grid_for((size_t)n * G, 64), 64, lds, st, db, cand, idx, n, pos, state, blk, resume, next, nnext)
  MIS_LANES(lanes, MIS_FIRST)
#undef MIS_FIRST
}
void k_mis_step_dev(const DB& db, const u32* cand, const u32* idx, const u32* n_in, u32 n_bound, u32* pos,
                    u32* state, u32* blk, u32* resume, u32* next, u32* nnext, u32* iters, u32* zero, u32 lds,
                    u32 lanes, Stream* st) {
  if (!n_bound) return;
#define MIS_DEV(G)                                                                                           \
  KLAUNCH_LDS(mis_step_dev_k<G>, std::min(grid_for((size_t)n_bound * G, 64), 2048u), 64, lds, st, db, cand, idx, \
              n_in, pos, state, blk, resume, next, nnext, iters, zero)
  MIS_LANES(lanes, MIS_DEV)
#undef MIS_DEV
}
#undef MIS_LANES
__global__ void mis_break_k(const u32* cand, u32 n, const u8* cflag, const u32* state, u32* brk) {
  const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && state[i] == 1 && cflag[cand[i]] == 3) atomicMin(brk, i);
}
void k_mis_break(const u32* cand, u32 n, const u8* cflag, const u32* state, u32* brk, Stream* st) {
  if (!n) return;
  KLAUNCH(mis_break_k, grid_for(n, 256), 256, st, cand, n, cflag, state, brk);
}
__global__ void mis_flag_k(const u32* state, u32 n, const u32* brk, u32* flag) {
  const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) flag[i] = (state[i] == 1 && i < *brk) ? 1u : 0u;
}
__global__ void mis_pack_k(const u32* cand, const u32* state, u32 n, const u32* brk, const u32* pos,
                           u32* out) {
  const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && state[i] == 1 && i < *brk) out[pos[i]] = cand[i];
}
void k_mis_collect(const u32* cand, u32 n, const u32* state, u32 brk_unused, u32* flag_tmp, u32* out,
                   u32* nout, void* tmp, Stream* st) {
  (void)brk_unused;
  // brk lives in nout[1] on the device; nout[0] receives the count
  if (!n) {
    fill_u32(nout, 0, 1, st);
    return;
  }
  KLAUNCH(mis_flag_k, grid_for(n, 256), 256, st, state, n, nout + 1, flag_tmp);
  scan_u32(flag_tmp, flag_tmp, n, nout, tmp, st);
  KLAUNCH(mis_pack_k, grid_for(n, 256), 256, st, cand, state, n, nout + 1, flag_tmp, out);
}

// ---------------------------------------------------------------- per elected variable
constexpr int VB = 64;  // one wavefront per workgroup for divergent per-variable work
// Occupancy of the per-variable kernels that walk the pool (the VE gate
// checks): each of their one-wavefront workgroups reserves PV_LDS bytes of
// LDS, so at most 160 KiB / PV_LDS of them share a CU: 20 KiB gives 8
// wavefronts per CU, 2 per SIMD.  Fewer waves keep more of each wave's
// clauses in L1/L2 while it re-reads them (Tseitin round on MI355X, when
// SUB and all VE decisions walked the pool: 0 / 12 / 20 / 40 KiB -> 18.1 /
// 17.4 / 17.25 / 17.35 ms).  The staged kernels use the same size as their
// working LDS (VR_LDS).
#ifndef PV_LDS
#define PV_LDS 20480
#endif

__global__ void list_sizes_k(DB db, const u32* pv, u32 n, u32 extra, u32* out) {
  const u32 k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) out[k] = db.ot_size[V2L(pv[k])] + db.ot_size[V2L(pv[k]) | 1] + extra;
}
void k_list_sizes(const DB& db, const u32* pv, u32 n, u32 extra, u32* out, Stream* st) {
  if (!n) return;
  KLAUNCH(list_sizes_k, grid_for(n, 256), 256, st, db, pv, n, extra, out);
}
// 'order' (nullable) permutes the elected indices so that the heaviest
// variables start first and share wavefronts with variables of similar work

__global__ void sub_snapshot_k(DB db, const u32* pv, u32 n, SubSnap s) {
  const u32 k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) sub_snapshot_var(db, pv[k], k, s);
}
void k_sub_snapshot(const DB& db, const u32* pv, u32 n, const SubSnap& s, Stream* st) {
  if (!n) return;
  KLAUNCH(sub_snapshot_k, grid_for(n, VB), VB, st, db, pv, n, s);
}
__global__ void sub_restore_k(DB db, const u32* pv, const u32* idx, u32 n, SubSnap s) {
  const u32 t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) sub_restore_var(db, pv[idx[t]], idx[t], s);
}
void k_sub_restore(const DB& db, const u32* pv, const u32* idx, u32 n, const SubSnap& s, Stream* st) {
  if (!n) return;
  KLAUNCH(sub_restore_k, grid_for(n, VB), VB, st, db, pv, idx, n, s);
}
// one workgroup per unit change (elected index, literal, pass): flag the
// owners of the clauses in the literal's list that come after the deriving
// variable and last ran no later than that pass
// one row of workgroups per change (its unit's list: launch_per_literal)
__global__ void sub_affected_k(DB db, SubSnap s, const u32* ch, const u32* lastrun, u32* aff) {
  const u64 c = 3 * (u64)blockIdx.y;
  const u32 k_unit = ch[c], u = ch[c + 1], pass = ch[c + 2];
  const OLV ol = olist(db, u);
  for (u32 i = list_lane(); i < ol.n; i += list_stride()) {
    const u32 o = sub_affected_owner(s, ol.r[i], k_unit, lastrun, pass);
    if (o != NOPOS) aff[o] = 1;
  }
}
void k_sub_affected(const DB& db, const SubSnap& s, const u32* changes, u32 nch, const u32* lastrun, u32* aff,
                    Stream* st) {
  launch_per_literal(nch, [&](dim3 grid, u32 c0) {
    KLAUNCH(sub_affected_k, grid, 256, st, db, s, changes + 3 * (size_t)c0, lastrun, aff);
  });
}
__global__ void set_at_k(u32* dst, const u32* idx, u32 n, u32 value) {
  const u32 t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) dst[idx[t]] = value;
}
void k_set_at(u32* dst, const u32* idx, u32 n, u32 value, Stream* st) {
  if (!n) return;
  KLAUNCH(set_at_k, grid_for(n, 256), 256, st, dst, idx, n, value);
}
// triples (variable, elected index, literal): ust[variable] / ulit[variable]
__global__ void set_vis_k(u32* ust, u32* ulit, const u32* e, u32 n) {
  const u32 t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  ust[e[3 * t]] = e[3 * t + 1];
  ulit[e[3 * t]] = e[3 * t + 2];
}
void k_set_vis(u32* ust, u32* ulit, const u32* entries, u32 n, Stream* st) {
  if (!n) return;
  KLAUNCH(set_vis_k, grid_for(n, 256), 256, st, ust, ulit, entries, n);
}
__global__ void pack_flags_k(const u32* flags, const u32* pos, u32 n, u32* out) {
  const u32 t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n && flags[t]) out[pos[t]] = t;
}
void k_pack_flags(const u32* flags, u32 n, u32* pos, u32* out, u32* count, void* tmp, Stream* st) {
  if (!n) {
    fill_u32(count, 0, 1, st);
    return;
  }
  scan_u32(flags, pos, n, count, tmp, st);
  KLAUNCH(pack_flags_k, grid_for(n, 256), 256, st, flags, pos, n, out);
}
__global__ void oblivion_k(DB db, const u32* lits) {
  const u32 t = lits[blockIdx.y];
  const OLV ol = olist(db, t);
  for (u32 i = list_lane(); i < ol.n; i += list_stride()) set_deleted(db, ol.r[i]);
}
void k_oblivion(const DB& db, const u32* lits, u32 n, Stream* st) {
  launch_per_literal(n, [&](dim3 grid, u32 c0) { KLAUNCH(oblivion_k, grid, 256, st, db, lits + c0); });
}
// Per-variable searches over a variable's own two occurrence lists (SUB,
// the VE decisions that are not settled by the gate checks) re-read those
// lists all along.  Each variable copies them (lview_words / sview_words
// words) into LDS and works there.  A wavefront's LDS (cap words) is
// shared by the variables of one size class: class q < LV_CLASSES - 1
// gives 64 >> q lanes cap / (64 >> q) words each; the last class (lists
// larger than cap) works on the pool, 64 lanes to a wavefront.  A
// wavefront thus runs variables of like size, the largest first, and no
// lane that reads the pool holds up lanes that work in LDS.
//
// work: LV_WORK_HDR words (the first positions of the sort keys 0..8 in
// the sorted order), then n sort keys, then n elected indices.  The key
// of a variable of class c is LV_CLASSES - 1 - c (the largest lists
// first); VE gives the variables its gate checks settle key LV_CLASSES.
// A stable radix sort orders the pairs, lv_bounds_k finds where each key
// starts.  No counter is shared by the variables.
constexpr u32 LV_CLASSES = 8;
constexpr u32 LV_WORK_HDR = 32;
constexpr int LV_KEY_BITS = 4;
static_assert(LV_CLASSES + 2 <= LV_WORK_HDR && LV_CLASSES < (1u << LV_KEY_BITS), "class keys");
__device__ inline u32 lv_class(u32 need, u32 cap) {
  u32 q = 0;
  while (q < LV_CLASSES - 1 && need * (64u >> q) > cap) q++;
  return q;
}
__device__ inline u32 lv_class_lanes(u32 q) { return q < LV_CLASSES - 1 ? 64u >> q : 64u; }
__global__ void lv_bounds_k(const u32* keys, u32 n, u32* starts) {
  const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && (i == 0 || keys[i] != keys[i - 1])) starts[keys[i]] = i;
}
// sorts the (key, elected index) pairs and finds the key starts
void lv_sort(u32* work, u32 n, void* tmp, Stream* st) {
  radix_sort_pairs(work + LV_WORK_HDR, work + LV_WORK_HDR + n, n, LV_KEY_BITS, tmp, st);
  KLAUNCH(lv_bounds_k, grid_for(n, 256), 256, st, work + LV_WORK_HDR, n, work);
}
// one wavefront per workgroup, grid-stride over the chunks of all classes;
// f(k, w, slice, gl, G) runs elected index k with the LDS words
// [w, w + slice) (w null: on the pool) on a group of G lanes (gl = the
// lane's place in it): a chunk of class c < LV_CLASSES - 1 holds 64 >> c
// variables, each on 1 << c lanes.  A class is chosen from a bound on the
// words the variable needs; the staging checks the slice and falls back to
// the pool (one lane per variable).
template <class F>
__device__ inline void lv_chunks(const u32* work, u32 n, u32 cap, u32* lds, F f) {
  const u32 lane = threadIdx.x;
  const u32* placed = work + LV_WORK_HDR + n;
  // key starts (absent keys start where the next key starts)
  u32 start[LV_CLASSES + 1];
  u32 next = n;
#pragma unroll
  for (int key = LV_CLASSES; key >= 0; key--) {
    const u32 s0 = work[key];
    start[key] = s0 < n ? s0 : next;
    next = start[key];
  }
  u32 chunks = 0;
#pragma unroll
  for (u32 key = 0; key < LV_CLASSES; key++) {
    const u32 L = lv_class_lanes(LV_CLASSES - 1 - key);
    chunks += (start[key + 1] - start[key] + L - 1) / L;
  }
  for (u32 J = blockIdx.x; J < chunks; J += gridDim.x) {
    // the key of chunk J and its range, found by a loop over constant
    // indices so that start[] stays in registers (a loop that stops at the
    // key indexes it at run time, which puts the array in scratch memory)
    u32 key = LV_CLASSES - 1, j = J, s0 = start[LV_CLASSES - 1], s1 = start[LV_CLASSES];
    bool found = false;
#pragma unroll
    for (u32 q = 0; q < LV_CLASSES - 1; q++) {
      const u32 L = lv_class_lanes(LV_CLASSES - 1 - q), m = (start[q + 1] - start[q] + L - 1) / L;
      if (!found) {
        if (j < m) {
          key = q, s0 = start[q], s1 = start[q + 1];
          found = true;
        } else {
          j -= m;
        }
      }
    }
    const u32 c = LV_CLASSES - 1 - key, L = lv_class_lanes(c), G = 64 / L;
    const u32 slot = lane / G, r = s0 + j * L + slot;
    if (r < s1)
      f(placed[r], c < LV_CLASSES - 1 ? lds + slot * (cap / L) : nullptr, cap / L, lane % G, G);
    __syncthreads();  // the next chunk reuses the LDS
  }
}
// LDS of each one-wavefront workgroup of the staged launches, and their
// workgroups: as many as fit the CUs
#ifndef VR_LDS
#define VR_LDS PV_LDS
#endif
constexpr u32 LV_BLOCKS = 256 * (160 * 1024 / VR_LDS);

// VE decisions in two launches: the gate checks (pure, equivalence,
// AND/OR) for every elected variable, then, staged, the variables they
// leave open (ITE and XOR searches, resolvent count)
// lsum[2k], lsum[2k + 1]: bounds on the clause size sums of the lists of
// elected variable k (from the list sort).  The searches' work grows with
// the square of the list lengths, their memory linearly: a variable's
// class is raised by VE_CLASS_BUMP (more lanes, a larger slice) over the
// class its lists need.
#ifndef VE_CLASS_BUMP
#define VE_CLASS_BUMP 1
#endif
__global__ void ve_class_k(DB db, const u32* pv, u32 n, const u32* order, u32* work, const u32* lsum, u32 cap) {
  const u32 t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const u32 k = order ? order[t] : t;
  const u32 p = V2L(pv[k]);
  const u64 b = lview_bound(db.ot_size[p] + db.ot_size[p | 1], (u64)lsum[2 * k] + lsum[2 * k + 1]);
  u32 c = lv_class((u32)min(b, (u64)cap + 1), cap);
  if (c < LV_CLASSES - 1) c = min(c + VE_CLASS_BUMP, LV_CLASSES - 2);
  work[LV_WORK_HDR + t] = LV_CLASSES - 1 - c;
  work[LV_WORK_HDR + n + t] = k;
}
// every elected variable on an LDS copy of its lists: the gate checks
// (pure, equivalence, AND/OR) by the group's first lane, then, for the
// variables they leave open, the ITE and XOR searches and the resolvent
// count shared by the group (the checks read each clause of the lists
// once, into LDS, instead of once per check from the pool)
__global__ void ve_decide_k(DB db, Options o, const u32* pv, VeDecision d, const u32* work, u32 n, u32 cap) {
  extern __shared__ u32 lds[];
  lv_chunks(work, n, cap, lds, [&](u32 k, u32* w, u32 slice, u32 gl, u32 G) {
    const u32 v = pv[k];
    LView lv;
    // every lane of the group stages the same words; the writes are the
    // first lane's
    if (w && lview_stage(db, v, w, slice, lv)) {
      group_sync();
      u32 open = 0;
      if (gl == 0) open = ve_decide_gates(lv, o, v, k, d) ? 0u : 1u;
      open = __shfl(open, (int)((threadIdx.x & 63) - gl), 64);
      group_sync();  // the first lane's marks in the copy are seen by the group
      if (open) ve_decide_rest(lv, o, v, k, d, gl, G);
      if (gl == 0) lview_commit(db, lv);
    } else if (gl == 0) {
      if (!ve_decide_gates(db, o, v, k, d)) ve_decide_rest(db, o, v, k, d);
    }
  });
}
void k_ve_decide(const DB& db, const Options& o, const u32* pv, u32 n, const VeDecision& d, const u32* order,
                 u32* work, const u32* lsum, void* tmp, Stream* st) {
  if (!n) return;
  const u32 cap = VR_LDS / 4;
  fill_u32(work, n, LV_WORK_HDR, st);
  KLAUNCH(ve_class_k, grid_for(n, 256), 256, st, db, pv, n, order, work, lsum, cap);
  lv_sort(work, n, tmp, st);
  KLAUNCH_LDS(ve_decide_k, std::min<u32>(grid_for(n, VB), LV_BLOCKS), VB, VR_LDS, st, db, o, pv, d, work, n, cap);
}
// SUB, staged as the VE searches (the class of a variable from the words of
// its SView, raised by SUB_CLASS_BUMP)
#ifndef SUB_CLASS_BUMP
#define SUB_CLASS_BUMP 0
#endif
__global__ void sub_class_k(DB db, const u32* pv, const u32* idx, u32 n, u32* work, const u32* lsum, u32 cap) {
  const u32 t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const u32 k = idx ? idx[t] : t, p = V2L(pv[k]);
  const u64 b = sview_bound(db.ot_size[p] + db.ot_size[p | 1], (u64)lsum[2 * k] + lsum[2 * k + 1]);
  u32 c = lv_class((u32)min(b, (u64)cap + 1), cap);
  if (c < LV_CLASSES - 1) c = min(c + SUB_CLASS_BUMP, LV_CLASSES - 2);
  work[LV_WORK_HDR + t] = LV_CLASSES - 1 - c;
  work[LV_WORK_HDR + n + t] = k;
}
// take_snap: record each variable's state for the SUB re-runs just before
// its run (a run changes only the variable's own lists and clauses, so this
// equals a snapshot before the launch); the staged copy provides it
__global__ void sub_k(DB db, Options o, const u32* pv, const u32* work, u32 n, u32 cap, SubOut out, SubSnap snap,
                      u32 take_snap) {
  extern __shared__ u32 lds[];
  lv_chunks(work, n, cap, lds, [&](u32 k, u32* w, u32 slice, u32 gl, u32 G) {
    const u32 v = pv[k];
    SView sv;
    // every lane of the group stages the same words, so the outcome is the
    // group's; the scans are shared, the writes are the first lane's
    if (w && sview_stage(db, v, w, slice, sv)) {
      const u32 np0 = sv.w[0], nn0 = sv.w[1];
      if (take_snap && gl == 0) sub_snapshot_var(sv, v, k, snap);
      group_sync();
      sub_var(sv, o, v, k, out, gl, G);
      if (gl == 0) sview_commit(db, sv, np0, nn0);
    } else if (gl == 0) {
      if (take_snap) sub_snapshot_var(db, v, k, snap);
      sub_var(db, o, v, k, out);
    }
  });
}
void k_sub(const DB& db, const Options& o, const u32* pv, const u32* idx, u32 n, const SubOut& out, u32* work,
           const u32* lsum, void* tmp, const SubSnap* snap, Stream* st) {
  if (!n) return;
  const u32 cap = VR_LDS / 4;
  fill_u32(work, n, LV_WORK_HDR, st);
  KLAUNCH(sub_class_k, grid_for(n, 256), 256, st, db, pv, idx, n, work, lsum, cap);
  lv_sort(work, n, tmp, st);
  KLAUNCH_LDS(sub_k, std::min<u32>(grid_for(n, VB), LV_BLOCKS), VB, VR_LDS, st, db, o, pv, work, n, cap, out,
              snap ? *snap : SubSnap{}, snap ? 1u : 0u);
}
// VE execution, staged as the searches: the resolvent path reads the
// variable's lists from LDS; EQU substitutions (which rewrite their clauses
// in place) take the last class and work on the pool; variables VE keeps
// are not placed
__global__ void ve_apply_class_k(DB db, const u32* pv, u32 n, const u32* type, u32* work, const u32* lsum,
                                 u32 cap) {
  const u32 k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  u32 key = LV_CLASSES;
  if (type[k] == VT_EQU) {
    key = 0;
  } else if (type[k]) {
    const u32 p = V2L(pv[k]);
    const u64 b = lview_bound(db.ot_size[p] + db.ot_size[p | 1], (u64)lsum[2 * k] + lsum[2 * k + 1]);
    key = LV_CLASSES - 1 - lv_class((u32)min(b, (u64)cap + 1), cap);
  }
  work[LV_WORK_HDR + k] = key;
  work[LV_WORK_HDR + n + k] = k;
}
__global__ void ve_apply_k(DB db, const u32* pv, VeDecision d, VeApply a, const u32* work, u32 n, u32 cap) {
  extern __shared__ u32 lds[];
  lv_chunks(work, n, cap, lds, [&](u32 k, u32* w, u32 slice, u32 gl, u32 G) {
    if (gl) return;  // one lane per variable
    const u32 v = pv[k];
    LView lv;
    if (w && d.type[k] != VT_EQU && lview_stage(db, v, w, slice, lv)) ve_apply(lv, db, v, k, d, a);
    else ve_apply(db, db, v, k, d, a);
  });
}
void k_ve_apply(const DB& db, const u32* pv, u32 n, const VeDecision& d, const VeApply& a, u32* work,
                const u32* lsum, void* tmp, Stream* st) {
  if (!n) return;
  const u32 cap = VR_LDS / 4;
  fill_u32(work, n, LV_WORK_HDR, st);
  KLAUNCH(ve_apply_class_k, grid_for(n, 256), 256, st, db, pv, n, d.type, work, lsum, cap);
  lv_sort(work, n, tmp, st);
  KLAUNCH_LDS(ve_apply_k, std::min<u32>(grid_for(n, VB), LV_BLOCKS), VB, VR_LDS, st, db, pv, d, a, work, n, cap);
}
__global__ void bce_k(DB db, Options o, const u32* pv, const u32* type, u32 n, u32* words,
                      const u32* base, u32* stack, bool count, const u32* order) {
  const u32 t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const u32 k = order ? order[t] : t;
  if (type && type[k]) {
    if (count) words[k] = 0;
    return;
  }
  if (count) words[k] = bce_var(db, o, pv[k], true, nullptr);
  else bce_var(db, o, pv[k], false, stack + base[k]);
}
void k_bce(const DB& db, const Options& o, const u32* pv, const u32* type, u32 n, u32* words,
           const u32* base, u32* stack, bool count, const u32* order, Stream* st) {
  if (!n) return;
  KLAUNCH(bce_k, grid_for(n, VB), VB, st, db, o, pv, type, n, words, base, stack, count, order);
}
// the launch order's key: list occurrences on an 8-bit log-linear scale
// (exact below 64, then 16 steps per power of two up to 2^18), descending,
// so the sort takes two radix passes instead of four; ties keep the elected
// order
__device__ inline u32 work_code(u32 w) {
  if (w < 64) return w;
  const u32 e = 31 - __clz(w);
  return min(255u, 64 + (e - 6) * 16 + ((w >> (e - 4)) & 15u));
}
__global__ void work_keys_k(DB db, const u32* pv, u32 n, u32* key, u32* idx) {
  const u32 k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const u32 w = db.ot_size[V2L(pv[k])] + db.ot_size[V2L(pv[k]) | 1];
  key[k] = 255u - work_code(w);
  idx[k] = k;
}
void k_work_order(const DB& db, const u32* pv, u32 n, u32* key, u32* order, void* tmp, Stream* st) {
  if (!n) return;
  KLAUNCH(work_keys_k, grid_for(n, 256), 256, st, db, pv, n, key, order);
  radix_sort_pairs(key, order, n, 8, tmp, st);
}
__global__ void ere_k(DB db, Options o, const u32* pv, u32 from, u32 n, EreOut out) {
  const u32 k = from + blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) ere_var(db, o, pv[k], k, out);
}
void k_ere(const DB& db, const Options& o, const u32* pv, u32 from, u32 n, const EreOut& out, Stream* st) {
  if (from >= n) return;
  KLAUNCH(ere_k, grid_for(n - from, VB), VB, st, db, o, pv, from, n, out);
}
__global__ void ere_dirty_k(DB db, const u32* pv, u32 from, u32 n, const u32* del_by, u32* first) {
  const u32 k = from + blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n && ere_dirty_var(db, pv[k], k, del_by)) atomicMin(first, k);
}
void k_ere_dirty(const DB& db, const u32* pv, u32 from, u32 n, const u32* del_by, u32* first,
                 Stream* st) {
  if (from >= n) return;
  KLAUNCH(ere_dirty_k, grid_for(n - from, VB), VB, st, db, pv, from, n, del_by, first);
}
__global__ void del_by_k(const u32* rec, u32 nrec, u32* del_by) {
  const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nrec) atomicMin(&del_by[rec[3 * (u64)i + 2]], rec[3 * (u64)i]);
}
void k_del_by(const u32* rec, u32 nrec, u32* del_by, Stream* st) {
  if (!nrec) return;
  KLAUNCH(del_by_k, grid_for(nrec, 256), 256, st, rec, nrec, del_by);
}
__global__ void apply_del_k(DB db, const u32* ids, u32 n) {
  const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) set_deleted(db, ids[i]);
}
void k_apply_deletions(const DB& db, const u32* ids, u32 n, Stream* st) {
  if (!n) return;
  KLAUNCH(apply_del_k, grid_for(n, 256), 256, st, db, ids, n);
}
__global__ void ht_build_k(DB db, HashTable ht, int maxsize) {
  const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < db.ncls) ht_insert_clause(db, ht, db.c_off[i], maxsize);
}
void k_ht_build(const DB& db, const HashTable& ht, int maxsize, Stream* st) {
  fill_u32((u32*)ht.slot, 0xFFFFFFFFu, 2 * ((size_t)ht.mask + 1), st);
  if (!db.ncls) return;
  KLAUNCH(ht_build_k, grid_for(db.ncls, 256), 256, st, db, ht, maxsize);
}
__global__ void ere_pair_counts_k(DB db, Options o, const u32* pv, u32 n, u32* np, unsigned long long* total) {
  const u32 k = blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long c = 0;
  if (k < n) {
    c = ere_pair_count(db, o, pv[k]);
    np[k] = (u32)c;
  }
  c = wave_sum64(c);
  if ((threadIdx.x & 63) == 0 && c) atomicAdd(total, c);
}
void k_ere_pair_counts(const DB& db, const Options& o, const u32* pv, u32 n, u32* np, u64* total, Stream* st) {
  fill_u32((u32*)total, 0, 2, st);
  if (!n) return;
  KLAUNCH(ere_pair_counts_k, grid_for(n, 256), 256, st, db, o, pv, n, np, (unsigned long long*)total);
}
// one thread per resolvent pair; the owning variable is found by a binary
// search over the per-variable pair offsets
__global__ void ere_pairs_k(DB db, ErePairs a, u32 npairs) {
  const u32 t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= npairs) return;
  u32 lo = 0, hi = a.nvars;  // first k with pbase[k] > t
  while (lo < hi) {
    const u32 mid = (lo + hi) >> 1;
    if (a.pbase[mid] <= t) lo = mid + 1;
    else hi = mid;
  }
  const u32 k = lo - 1;
  ere_pair(db, a, k, t - a.pbase[k]);
}
void k_ere_pairs(const DB& db, const ErePairs& a, u32 npairs, Stream* st) {
  if (!npairs) return;
  KLAUNCH(ere_pairs_k, grid_for(npairs, 256), 256, st, db, a, npairs);
}
__global__ void ere_all_matches_k(DB db, ErePairs a, u32 n, const u32* off, u32* out) {
  const u32 o = blockIdx.x * blockDim.x + threadIdx.x;
  if (o < n) ere_all_matches(db, a, o, off, out);
}
void k_ere_all_matches(const DB& db, const ErePairs& a, u32 n, const u32* off, u32* out, Stream* st) {
  if (!n) return;
  KLAUNCH(ere_all_matches_k, grid_for(n, 64), 64, st, db, a, n, off, out);
}

// ---------------------------------------------------------------- small helpers
__global__ void assign_units_k(DB db, const u32* lits, u32 n) {
  const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const u32 l = lits[i];
  db.assign[ABS(l)] = l + 1;
  db.vstate[ABS(l)] = VS_FROZEN;
}
void k_assign_units(const DB& db, const u32* lits, u32 n, Stream* st) {
  if (!n) return;
  KLAUNCH(assign_units_k, grid_for(n, 256), 256, st, db, lits, n);
}
__global__ void pv_lits_k(const u32* pv, u32 n, u32* lits) {
  const u32 k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  lits[2 * k] = V2L(pv[k]);
  lits[2 * k + 1] = V2L(pv[k]) | 1;
}
void k_pv_lits(const u32* pv, u32 n, u32* lits, Stream* st) {
  if (!n) return;
  KLAUNCH(pv_lits_k, grid_for(n, 256), 256, st, pv, n, lits);
}
__global__ void max_list_k(DB db, const u32* lits, u32 n, u32* out) {
  u32 m = 0;
  for (u32 i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    m = max(m, db.ot_size[lits[i]]);
  m = wave_max(m);
  if ((threadIdx.x & 63) == 0) atomicMax(out, m);
}
void k_max_list(const DB& db, const u32* lits, u32 n, u32* out, Stream* st) {
  fill_u32(out, 0, 1, st);
  if (!n) return;
  KLAUNCH(max_list_k, std::min(grid_for(n, 256), RED_BLOCKS), 256, st, db, lits, n, out);
}
__global__ void reduce_lists_k(DB db) {
  const u32 l = blockIdx.x * blockDim.x + threadIdx.x + 2;
  if (l >= 2 * db.nvars + 2) return;
  u32* r = db.ot_refs + db.ot_off[l];
  const u32 n = db.ot_size[l];
  u32 j = 0;
  for (u32 i = 0; i < n; i++)
    if (!is_deleted(db, r[i])) r[j++] = r[i];
  db.ot_size[l] = j;
}
void k_reduce_lists(const DB& db, Stream* st) {
  KLAUNCH(reduce_lists_k, grid_for(2 * (size_t)db.nvars, 256), 256, st, db);
}
__global__ void prep_k(DB db, u32 first) {
  const u32 i = first + blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= db.ncls) return;
  u32* h = chead(db, db.c_off[i]);
  u32* l = h + CL_HDR;
  const u32 n = hdr_size(h[0]);
  for (u32 q = 1; q < n; q++) {
    const u32 t = l[q];
    u32 j = q;
    while (j > 0 && l[j - 1] > t) { l[j] = l[j - 1]; j--; }
    l[j] = t;
  }
  h[1] = calc_sig(l, n);
}
void k_prep(const DB& db, u32 first, Stream* st) {
  if (first >= db.ncls) return;
  KLAUNCH(prep_k, grid_for(db.ncls - first, 256), 256, st, db, first);
}
__global__ void clear_added_k(DB db) {
  const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= db.ncls) return;
  const u32 c = db.c_off[i];
  const u8 f = cflags(db, c);
  if (f & F_ADDED) set_cflags(db, c, (u8)(f & ~F_ADDED));
}
void k_clear_added(const DB& db, Stream* st) {
  if (!db.ncls) return;
  KLAUNCH(clear_added_k, grid_for(db.ncls, 256), 256, st, db);
}
__global__ void zero_sizes_k(DB db, const u32* lits, u32 n) {
  const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) db.ot_size[lits[i]] = 0;
}
void k_zero_sizes(const DB& db, const u32* lits, u32 n, Stream* st) {
  if (!n) return;
  KLAUNCH(zero_sizes_k, grid_for(n, 256), 256, st, db, lits, n);
}

// ---------------------------------------------------------------- DRAT proof
__global__ void proof_sizes_k(DB db, const u32* lines, u32 n, const u32* stamp, u32* bytes) {
  const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) bytes[i] = proof_line_size(db, lines, i, stamp);
}
__global__ void proof_write_k(DB db, const u32* lines, u32 n, const u32* bytes, const u32* off, u8* out) {
  const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && bytes[i]) proof_line_write(db, lines, i, out + off[i]);  // 0 bytes: dropped line
}
void k_proof_sizes(const DB& db, const u32* lines, u32 n, const u32* stamp, u32* bytes, Stream* st) {
  if (!n) return;
  KLAUNCH(proof_sizes_k, grid_for(n, 256), 256, st, db, lines, n, stamp, bytes);
}
void k_proof_write(const DB& db, const u32* lines, u32 n, const u32* bytes, const u32* off, u8* out,
                   Stream* st) {
  if (!n) return;
  KLAUNCH(proof_write_k, grid_for(n, 256), 256, st, db, lines, n, bytes, off, out);
}
__global__ void set_stamps_k(u32* stamp, const u32* pairs, u32 n, int clear) {
  const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) stamp[pairs[2 * i]] = clear ? 0u : pairs[2 * i + 1];
}
void k_set_stamps(u32* stamp, const u32* pairs, u32 n, bool clear, Stream* st) {
  if (!n) return;
  KLAUNCH(set_stamps_k, grid_for(n, 256), 256, st, stamp, pairs, n, clear ? 1 : 0);
}

// ---------------------------------------------------------------- count / compact
__global__ void count_k(DB db, unsigned long long* out) {
  unsigned long long live = 0, lits = 0;
  for (u32 i = blockIdx.x * blockDim.x + threadIdx.x; i < db.ncls; i += gridDim.x * blockDim.x) {
    const u32 h = chead(db, db.c_off[i])[0];
    if (!st_deleted(hdr_flags(h))) {
      live++;
      lits += hdr_size(h);
    }
  }
  live = wave_sum64(live);
  lits = wave_sum64(lits);
  if ((threadIdx.x & 63) == 0) {
    atomicAdd(&out[0], live);
    atomicAdd(&out[1], lits);
  }
}
// occurrences (O_E) and clause bytes (D_E: 12-byte header + 4 bytes per
// literal) of the live clauses in the elected variables' lists; elected
// variables share no clause, so every clause is counted once
__global__ void elected_work_k(DB db, const u32* pv, u32 n, unsigned long long* out) {
  unsigned long long occ = 0, bytes = 0;
  for (u32 k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
    for (u32 s = 0; s < 2; s++) {
      const OLV ol = olist(db, V2L(pv[k]) | s);
      for (u32 i = 0; i < ol.n; i++) {
        const u32 h = chead(db, ol.r[i])[0];
        if (st_deleted(hdr_flags(h))) continue;
        occ++;
        bytes += 12 + 4ull * hdr_size(h);
      }
    }
  }
  occ = wave_sum64(occ);
  bytes = wave_sum64(bytes);
  if ((threadIdx.x & 63) == 0) {
    atomicAdd(&out[0], occ);
    atomicAdd(&out[1], bytes);
  }
}
// occurrences in the election candidates' lists (O_c) with their clauses
// marked (per handle), then the bytes of the distinct marked clauses (D_c);
// the marks are cleared again.  Profiling only.
__global__ void cand_occ_k(DB db, const u32* cand, u32 n, u8* mark, unsigned long long* out) {
  unsigned long long occ = 0;
  for (u32 k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
    for (u32 s = 0; s < 2; s++) {
      const OLV ol = olist(db, V2L(cand[k]) | s);
      for (u32 i = 0; i < ol.n; i++) {
        const u32 c = ol.r[i];
        if (st_deleted(hdr_flags(chead(db, c)[0]))) continue;
        occ++;
        mark[c] = 1;
      }
    }
  }
  occ = wave_sum64(occ);
  if ((threadIdx.x & 63) == 0) atomicAdd(&out[0], occ);
}
__global__ void marked_bytes_k(DB db, u8* mark, unsigned long long* out) {
  unsigned long long bytes = 0;
  for (u32 i = blockIdx.x * blockDim.x + threadIdx.x; i < db.ncls; i += gridDim.x * blockDim.x) {
    const u32 c = db.c_off[i];
    if (!mark[c]) continue;
    mark[c] = 0;
    bytes += 12 + 4ull * csize(db, c);
  }
  bytes = wave_sum64(bytes);
  if ((threadIdx.x & 63) == 0) atomicAdd(&out[1], bytes);
}
void k_cand_work(const DB& db, const u32* cand, u32 n, u8* mark, u64* out2, Stream* st) {
  fill_u32((u32*)out2, 0, 4, st);
  if (!n) return;
  KLAUNCH(cand_occ_k, std::min(grid_for(n, ST), RED_BLOCKS), ST, st, db, cand, n, mark, (unsigned long long*)out2);
  if (db.ncls)
    KLAUNCH(marked_bytes_k, std::min(grid_for(db.ncls, ST), RED_BLOCKS), ST, st, db, mark,
            (unsigned long long*)out2);
}
void k_elected_work(const DB& db, const u32* pv, u32 n, u64* out2, Stream* st) {
  fill_u32((u32*)out2, 0, 4, st);
  if (!n) return;
  KLAUNCH(elected_work_k, std::min(grid_for(n, ST), RED_BLOCKS), ST, st, db, pv, n, (unsigned long long*)out2);
}
void k_count(const DB& db, u64* out2, Stream* st) {
  fill_u32((u32*)out2, 0, 4, st);
  if (!db.ncls) return;
  KLAUNCH(count_k, std::min(grid_for(db.ncls, ST), RED_BLOCKS), ST, st, db, (unsigned long long*)out2);
}
// sel: 0 every live clause, 1 live originals only, 2 live learnt clauses only
__device__ inline bool compact_sel(u32 h, int sel) {
  const u8 f = hdr_flags(h);
  if (st_deleted(f)) return false;
  return sel == 0 || (sel == 1) == !st_learnt(f);
}
__global__ void live_k(DB db, u32* tc, u32* tl, int sel) {
  const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= db.ncls) return;
  const u32 h = chead(db, db.c_off[i])[0];
  const bool live = compact_sel(h, sel);
  tc[i] = live ? 1u : 0u;
  tl[i] = live ? CL_HDR + hdr_size(h) : 0u;  // pool words
}
// base (device, nullable): clause index and pool word where the selection starts
__global__ void compact_k(DB src, DB dst, const u32* tc, const u32* tl, int sel, const u32* base) {
  const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= src.ncls) return;
  const u32* a = chead(src, src.c_off[i]);
  const u32 h = a[0];
  if (!compact_sel(h, sel)) return;
  const u32 nid = tc[i] + (base ? base[0] : 0u), noff = tl[i] + (base ? base[1] : 0u), n = hdr_size(h);
  dst.c_off[nid] = noff;
  u32* b = dst.lits + noff;
  b[0] = make_hdr(n, (u8)(hdr_flags(h) & ~F_MOLTEN));
  for (u32 i = 1; i < CL_HDR + n; i++) b[i] = a[i];
}
void k_compact(const DB& src, DB& dst, u32* tmp_cls, u32* tmp_lit, void* tmp, u32* totals, Stream* st) {
  if (!src.ncls) {
    fill_u32(totals, 0, 2, st);
    return;
  }
  KLAUNCH(live_k, grid_for(src.ncls, 256), 256, st, src, tmp_cls, tmp_lit, 0);
  scan_u32(tmp_cls, tmp_cls, src.ncls, totals, tmp, st);
  scan_u32(tmp_lit, tmp_lit, src.ncls, totals + 1, tmp, st);
  KLAUNCH(compact_k, grid_for(src.ncls, 256), 256, st, src, dst, tmp_cls, tmp_lit, 0, (const u32*)nullptr);
}
// the live clauses with the originals first and the learnt clauses after
// them, each part in clause order (totals[0..1]: clauses and words in all;
// totals[2..3] scratch)
__global__ void add_totals_k(u32* t) {
  t[0] += t[2];
  t[1] += t[3];
}
void k_compact_originals_first(const DB& src, DB& dst, u32* tmp_cls, u32* tmp_lit, void* tmp, u32* totals,
                               Stream* st) {
  if (!src.ncls) {
    fill_u32(totals, 0, 4, st);
    return;
  }
  KLAUNCH(live_k, grid_for(src.ncls, 256), 256, st, src, tmp_cls, tmp_lit, 1);
  scan_u32(tmp_cls, tmp_cls, src.ncls, totals, tmp, st);
  scan_u32(tmp_lit, tmp_lit, src.ncls, totals + 1, tmp, st);
  KLAUNCH(compact_k, grid_for(src.ncls, 256), 256, st, src, dst, tmp_cls, tmp_lit, 1, (const u32*)nullptr);
  KLAUNCH(live_k, grid_for(src.ncls, 256), 256, st, src, tmp_cls, tmp_lit, 2);
  scan_u32(tmp_cls, tmp_cls, src.ncls, totals + 2, tmp, st);
  scan_u32(tmp_lit, tmp_lit, src.ncls, totals + 3, tmp, st);
  KLAUNCH(compact_k, grid_for(src.ncls, 256), 256, st, src, dst, tmp_cls, tmp_lit, 2, (const u32*)totals);
  KLAUNCH(add_totals_k, 1, 1, st, totals);
}

// ---------------------------------------------------------------- variable map
// First root-assigned variable (VMAP::initiate, src/cpu/vmap.h:63-77): the
// lowest set lane of a wave holds the wave's smallest candidate, so each
// wave issues at most one atomic.
__global__ void map_first_k(const u8* vstate, const u32* assign, u32 nvars, u32* first) {
  const u32 v = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const bool cand = v <= nvars && vstate[v] && assign[v];
  const unsigned long long b = __ballot(cand);
  // a wave whose smallest candidate cannot win skips the atomic
  if (cand && (u32)(__ffsll(b) - 1) == (threadIdx.x & 63u) && v < __atomic_load_n(first, __ATOMIC_RELAXED))
    atomicMin(first, v);
}
__global__ void map_keep_k(const u8* vstate, u32 nvars, const u32* first, u32* keep) {
  const u32 v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v > nvars) return;
  const u32 f = *first == 0xFFFFFFFFu ? 0u : *first;
  keep[v] = map_keep(vstate, v, f);
}
// pos: exclusive scan of the keep flags -> mapped
__global__ void map_index_k(const u8* vstate, u32 nvars, const u32* first, const u32* pos, u32* mapped) {
  const u32 v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v > nvars) return;
  const u32 f = *first == 0xFFFFFFFFu ? 0u : *first;
  mapped[v] = map_keep(vstate, v, f) ? pos[v] + 1 : 0u;
}
// info = {first (0: none), new variable count, true literal of first, inactive variables}
__global__ void map_info_k(const u32* assign, const u32* mapped, u32 nvars, const u32* total, u32* info) {
  const u32 f = info[0] == 0xFFFFFFFFu ? 0u : info[0];
  const u32 n = *total;
  info[0] = f;
  info[1] = n;
  info[2] = f ? (V2L(mapped[f]) | SIGN(assign[f] - 1)) : 0u;
  info[3] = nvars - (n - (f ? 1u : 0u));
}
void k_map_build(const u8* vstate, const u32* assign, u32 nvars, u32* pos_tmp, u32* mapped, u32* info,
                 void* tmp, Stream* st) {
  fill_u32(info, 0xFFFFFFFFu, 1, st);
  if (nvars) KLAUNCH(map_first_k, grid_for(nvars, 256), 256, st, vstate, assign, nvars, info);
  KLAUNCH(map_keep_k, grid_for((size_t)nvars + 1, 256), 256, st, vstate, nvars, info, pos_tmp);
  scan_u32(pos_tmp, pos_tmp, (size_t)nvars + 1, info + 4, tmp, st);
  KLAUNCH(map_index_k, grid_for((size_t)nvars + 1, 256), 256, st, vstate, nvars, info, pos_tmp, mapped);
  KLAUNCH(map_info_k, 1, 1, st, assign, mapped, nvars, info + 4, info);
}
__global__ void map_clauses_k(DB db, VarMap m) {
  const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < db.ncls) map_clause(db, db.c_off[i], m);
}
void k_map_clauses(const DB& db, const VarMap& m, Stream* st) {
  if (!db.ncls) return;
  KLAUNCH(map_clauses_k, grid_for(db.ncls, 256), 256, st, db, m);
}
// per-variable arrays in the new numbering (VMAP::mapShrinkVars and mapSP,
// vmap.h:80-100, 132-143): only 'first' stays assigned
__global__ void map_vars_k(VarMap m, u32 nvars_old, const u32* vorg, const u8* vstate, const u8* assumed,
                           u32* vorg2, u8* vstate2, u8* assumed2, u32* assign2) {
  const u32 v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v > nvars_old) return;
  if (v == 0) {
    vorg2[0] = vorg[0];
    vstate2[0] = 0;
    assumed2[0] = 0;
    assign2[0] = 0;
    return;
  }
  const u32 nv = m.mapped[v];
  if (!nv) return;
  vorg2[nv] = vorg[v];
  vstate2[nv] = vstate[v];
  assumed2[nv] = assumed[v];
  assign2[nv] = v == m.first ? m.first_true + 1 : 0u;
}
void k_map_vars(const VarMap& m, u32 nvars_old, const u32* vorg, const u8* vstate, const u8* assumed, u32* vorg2,
                u8* vstate2, u8* assumed2, u32* assign2, Stream* st) {
  KLAUNCH(map_vars_k, grid_for((size_t)nvars_old + 1, 256), 256, st, m, nvars_old, vorg, vstate, assumed, vorg2,
          vstate2, assumed2, assign2);
}
// VMAP::mapOrgs (vmap.h:112-131): 0 stays 0
__global__ void map_lits_k(VarMap m, u32* lits, u64 n) {
  const u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && lits[i]) lits[i] = map_lit(m, lits[i]);
}
void k_map_lits(const VarMap& m, u32* lits, u64 n, Stream* st) {
  if (!n) return;
  KLAUNCH(map_lits_k, grid_for(n, 256), 256, st, m, lits, n);
}

// ---------------------------------------------------------------- fixed device scratch
// Staging memory of the entry points that allocate nothing
// (pfs_map_lits_noalloc): one array per device, held until the process
// ends.  Host code never uses the array's name as an address: its device
// address is looked up once per device (hipGetSymbolAddress, with that
// device current) and kept.  Every call on a device shares the array, so
// calls that overlap on one device race.
static __device__ __align__(256) u32 scratch[SCRATCH_WORDS];
static std::mutex g_scratch_m;
static std::map<int, u32*> g_scratch;
u32* scratch_lookup(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_m);
  auto it = g_scratch.find(device);
  if (it != g_scratch.end()) return it->second;
  int cur = -1;
  PFS_HIP(hipGetDevice(&cur));
  if (cur != device) throw std::runtime_error("scratch lookup: device " + std::to_string(device) + " is not current");
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    (void)hipGetLastError();
    throw std::runtime_error(std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
  }
  g_scratch[device] = (u32*)p;
  return (u32*)p;
}

// ---------------------------------------------------------------- output packing
__global__ void pack_sizes_k(DB db, u32* sizes) {
  const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < db.ncls) sizes[i] = csize(db, db.c_off[i]);
}
__global__ void pack_k(DB db, const u32* start, u64* offsets, u32* lits, u8* flags, u32* lbd) {
  const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < db.ncls) pack_clause(db, i, start, offsets, lits, flags, lbd);
}
void k_pack_cnf(const DB& db, u32* start_tmp, void* tmp, u32* total, u64* offsets, u32* lits, u8* flags, u32* lbd,
                Stream* st) {
  if (!db.ncls) {
    fill_u32((u32*)offsets, 0, 2, st);
    return;
  }
  KLAUNCH(pack_sizes_k, grid_for(db.ncls, 256), 256, st, db, start_tmp);
  scan_u32(start_tmp, start_tmp, db.ncls, total, tmp, st);
  KLAUNCH(pack_k, grid_for(db.ncls, 256), 256, st, db, start_tmp, offsets, lits, flags, lbd);
}

// ---------------------------------------------------------------- DIMACS ingestion
// The clause part of a DIMACS text (src/cpu/dimacs.cpp:41-110, toClause
// :113-170) is lexed on the device in 4096-byte tiles: workgroup b owns
// bytes [4096 b, 4096 b + 4096) of the text, each of its 256 threads 16
// consecutive bytes read with one 16-byte load (the text buffer is padded
// with blanks to whole tiles).  A token starts at a byte that is not blank
// (bytes 9-13 and 32, dimacs.h:33) and follows a blank or the start of the
// clause part.  Bytes of comment lines ('c' tokens, to the end of the line)
// are masked; the clause part ends at the first NUL byte or '%' token.
constexpr u32 LEX_T = 256, LEX_B = 16, LEX_TILE = LEX_T * LEX_B;
template <u32 MODE>
__global__ void __launch_bounds__(LEX_T) lex_k(LexArgs a) {
  __shared__ u32 lw[NW];
  const u64 base = (u64)blockIdx.x * LEX_TILE + (u64)threadIdx.x * LEX_B;
  const uint4 w = *(const uint4*)(a.text + base);
  u8 b[LEX_B];
  *(uint4*)b = w;
  u32 prev = base ? a.text[base - 1] : 32u;
  u32 mine = 0;  // starts of the kind this mode counts or emits
  u32 kinds = 0;  // bit j: byte j starts such a token
#pragma unroll
  for (u32 j = 0; j < LEX_B; j++) {
    const u64 i = base + j;
    const u32 c = b[j];
    const bool start = lex_start(i, c, prev, a.begin, a.end);
    prev = c;
    if (MODE == LEX_SPECIAL) {
      if (i >= a.begin && i < a.end && c == 0) atomicMin(&a.mark[0], (unsigned long long)i);
      if (start && c == '%') atomicMin(&a.mark[1], (unsigned long long)i);
      if (start && c == 'c') kinds |= 1u << j;
    } else if (MODE == LEX_COMMENTS) {
      if (start && c == 'c') kinds |= 1u << j;
    } else if (MODE == LEX_PERCENT) {
      if (start && c == '%' && !a.mask[i]) atomicMin(&a.mark[1], (unsigned long long)i);
    } else {
      if (start && c != 0 && !(a.mask && a.mask[i])) kinds |= 1u << j;
    }
  }
  mine = __popc(kinds);
  if (MODE == LEX_PERCENT) return;
  u32 tot;
  const u32 pre = block_excl_scan(mine, lw, &tot);
  if (MODE == LEX_SPECIAL || MODE == LEX_COUNT) {
    if (threadIdx.x == 0) a.cnt[blockIdx.x] = tot;
    return;
  }
  u32 k = a.cnt[blockIdx.x] + pre;
  while (kinds) {
    const u32 j = __ffs(kinds) - 1;
    kinds &= kinds - 1;
    const u64 i = base + j;
    if (MODE == LEX_COMMENTS) {
      a.pos[k++] = i;
      continue;
    }
    bool bad, big;
    a.lits[k++] = lex_literal(a.text, i, a.end, a.maxvar, bad, big);
    if (bad) {
      atomicOr(a.err, (u32)LEXE_TOKEN);
      atomicMin(&a.mark[2], (unsigned long long)i);
    }
    if (big) atomicOr(a.err, (u32)LEXE_VAR);  // "too many variables"
  }
}
// one thread per comment start: its bytes up to and including the newline
__global__ void lex_mask_k(const u8* text, u64 end, const u64* pos, u32 n, u8* mask) {
  const u32 k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  for (u64 i = pos[k]; i < end; i++) {
    mask[i] = 1;
    if (text[i] == '\n') break;
  }
}
u64 lex_tiles(u64 n) { return (n + LEX_TILE - 1) / LEX_TILE; }
void k_lex(u32 mode, const LexArgs& a, u64 ntiles, Stream* st) {
  if (!ntiles) return;
  switch (mode) {
    case LEX_SPECIAL: KLAUNCH(lex_k<LEX_SPECIAL>, (unsigned)ntiles, LEX_T, st, a); break;
    case LEX_COMMENTS: KLAUNCH(lex_k<LEX_COMMENTS>, (unsigned)ntiles, LEX_T, st, a); break;
    case LEX_PERCENT: KLAUNCH(lex_k<LEX_PERCENT>, (unsigned)ntiles, LEX_T, st, a); break;
    case LEX_COUNT: KLAUNCH(lex_k<LEX_COUNT>, (unsigned)ntiles, LEX_T, st, a); break;
    default: KLAUNCH(lex_k<LEX_EMIT>, (unsigned)ntiles, LEX_T, st, a); break;
  }
}
void k_lex_mask(const u8* text, u64 end, const u64* pos, u32 n, u8* mask, Stream* st) {
  if (!n) return;
  KLAUNCH(lex_mask_k, grid_for(n, 256), 256, st, text, end, pos, n, mask);
}
// clause ends: ends[c] = index of the c-th 0 among the literals
__global__ void zero_flags_k(const u32* lits, u64 n, u32* z) {
  const u64 k = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) z[k] = lits[k] == 0;
}
__global__ void zero_emit_k(const u32* lits, u64 n, const u32* pos, u32* ends) {
  const u64 k = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n && lits[k] == 0) ends[pos[k]] = (u32)k;
}
void k_clause_ends(const u32* lits, u64 n, u32* tmp_pos, void* tmp, u32* total, u32* ends, Stream* st) {
  if (!n) {
    fill_u32(total, 0, 1, st);
    return;
  }
  KLAUNCH(zero_flags_k, grid_for(n, 256), 256, st, lits, n, tmp_pos);
  scan_u32(tmp_pos, tmp_pos, n, total, tmp, st);
  KLAUNCH(zero_emit_k, grid_for(n, 256), 256, st, lits, n, tmp_pos, ends);
}
// clause i holds the literals between the (i-1)-th and the i-th 0; in the
// pool it starts at that literal offset plus two words per earlier clause
__global__ void parse_build_k(DB db, const u32* lits, const u32* ends, u32* err) {
  const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < db.ncls && !parse_build_clause(db, lits, ends, i)) atomicOr(err, (u32)LEXE_SIZE);
}
void k_parse_build(const DB& db, const u32* lits, const u32* ends, u32* err, Stream* st) {
  if (!db.ncls) return;
  KLAUNCH(parse_build_k, grid_for(db.ncls, 256), 256, st, db, lits, ends, err);
}
// ingestion proof: bytes per line (units first, then the text clauses),
// then the lines at their scanned offsets
__global__ void parse_proof_k(DB db, const u32* toks, const u32* ends, const u32* units, u32 nu, u32 nc,
                              const u32* off, u8* out, u32* bytes) {
  const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nu + nc) return;
  u8* o = off ? out + off[i] : nullptr;
  const u32 b = i < nu ? unit_proof_line(db, units[i], o) : parse_proof_clause(db, toks, ends, i - nu, o);
  if (!off) bytes[i] = b;
}
void k_parse_proof(const DB& db, const u32* toks, const u32* ends, const u32* units, u32 nu, u32 nc,
                   const u32* off, u8* out, u32* bytes, Stream* st) {
  if (!(nu + nc)) return;
  KLAUNCH(parse_proof_k, grid_for((size_t)nu + nc, 256), 256, st, db, toks, ends, units, nu, nc, off, out, bytes);
}
// rewrite proof: line i < nu is unit i, then the additions (phase 0) of
// clause indices 0..ncls, then the deletions (phase 1)
__global__ void rewrite_proof_k(DB db, const u32* old, const u32* units, u32 nu, const u32* off, u8* out,
                                u32* bytes) {
  const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nu + 2 * db.ncls) return;
  u8* o = off ? out + off[i] : nullptr;
  u32 b;
  if (i < nu) {
    b = unit_proof_line(db, units[i], o);
  } else {
    const u32 j = i - nu, phase = j >= db.ncls;
    b = rewrite_proof_clause(db, old, db.c_off[phase ? j - db.ncls : j], (int)phase, o);
  }
  if (!off) bytes[i] = b;
}
void k_rewrite_proof(const DB& db, const u32* old, const u32* units, u32 nu, const u32* off, u8* out, u32* bytes,
                     Stream* st) {
  const size_t n = (size_t)nu + 2 * (size_t)db.ncls;
  if (!n) return;
  KLAUNCH(rewrite_proof_k, grid_for(n, 256), 256, st, db, old, units, nu, off, out, bytes);
}
// after the literal sort: duplicates dropped, tautologies removed, empty
// clauses reported, unit clauses turned into unit records (toClause,
// dimacs.cpp:120-166); ulit[i] = the unit literal of clause i or 0;
// kept[0] counts the clauses that stay
__global__ void parse_dedup_k(DB db, u32* ulit, u32* err) {
  const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < db.ncls && parse_dedup_clause(db, db.c_off[i], &ulit[i]) == 2) atomicOr(err, (u32)LEXE_UNSAT);
}
void k_parse_dedup(const DB& db, u32* ulit, u32* err, Stream* st) {
  if (!db.ncls) return;
  KLAUNCH(parse_dedup_k, grid_for(db.ncls, 256), 256, st, db, ulit, err);
}
// unit clauses in text order: the first one of each variable is assigned
// (enqueueUnit); a later one of the other sign is a conflict
__global__ void parse_claim_k(const u32* ulit, u32 n, u32* first) {
  const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && ulit[i]) atomicMin(&first[ABS(ulit[i])], i);
}
__global__ void parse_settle_k(const u32* ulit, u32 n, const u32* first, u32* win, u32* err) {
  const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const u32 u = ulit[i];
  win[i] = 0;
  if (!u) return;
  const u32 f = first[ABS(u)];
  if (ulit[f] != u) atomicOr(err, (u32)LEXE_UNSAT);
  win[i] = f == i;
}
__global__ void parse_trail_k(DB db, const u32* ulit, u32 n, const u32* win, const u32* pos, u32* trail) {
  const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !win[i]) return;
  const u32 u = ulit[i];
  trail[pos[i]] = u;
  db.assign[ABS(u)] = u + 1;
  db.vstate[ABS(u)] = VS_FROZEN;
}
void k_parse_units(const DB& db, const u32* ulit, u32 n, u32* first, u32* win, u32* pos, void* tmp, u32* total,
                   u32* trail, u32* err, Stream* st) {
  fill_u32(total, 0, 1, st);
  if (!n) return;
  fill_u32(first, 0xFFFFFFFFu, (size_t)db.nvars + 1, st);
  KLAUNCH(parse_claim_k, grid_for(n, 256), 256, st, ulit, n, first);
  KLAUNCH(parse_settle_k, grid_for(n, 256), 256, st, ulit, n, first, win, err);
  scan_u32(win, pos, n, total, tmp, st);
  KLAUNCH(parse_trail_k, grid_for(n, 256), 256, st, db, ulit, n, win, pos, trail);
}
__global__ void iota_u32_k(u32* a, u64 n) {
  const u64 k = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) a[k] = (u32)k;
}
void iota_u32(u32* a, size_t n, Stream* st) {
  if (!n) return;
  KLAUNCH(iota_u32_k, grid_for(n, 256), 256, st, a, (u64)n);
}

// ---------------------------------------------------------------- ELS
// Binary implication graph over literals (equivalence.cpp:26-40, 92-139):
// a live clause (a, b) of two active variables gives a -> ... edges
// -a -> b and -b -> a.  Out-edges are a CSR (off, adj); the in-edges of l
// are the negations of the out-edges of -l (the graph is skew-symmetric).
PF_DEV bool els_binary(const DB& db, u32 c, u32& a, u32& b) {
  if (is_deleted(db, c) || csize(db, c) != 2) return false;
  const u32* l = clits(db, c);
  a = l[0], b = l[1];
  return !db.vstate[ABS(a)] && !db.vstate[ABS(b)];
}
__global__ void els_deg_k(DB db, u32* deg) {
  const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
  u32 a, b;
  if (i < db.ncls && els_binary(db, db.c_off[i], a, b)) {
    atomicAdd(&deg[FLIP(a)], 1u);
    atomicAdd(&deg[FLIP(b)], 1u);
  }
}
__global__ void els_fill_k(DB db, const u32* off, u32* cur, u32* adj) {
  const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
  u32 a, b;
  if (i < db.ncls && els_binary(db, db.c_off[i], a, b)) {
    adj[off[FLIP(a)] + atomicAdd(&cur[FLIP(a)], 1u)] = b;
    adj[off[FLIP(b)] + atomicAdd(&cur[FLIP(b)], 1u)] = a;
  }
}
void k_els_graph(const DB& db, u32* deg, u32* off, void* tmp, Stream* st) {
  const size_t n2 = 2 * (size_t)db.nvars + 2;
  fill_u32(deg, 0, n2, st);
  if (db.ncls) KLAUNCH(els_deg_k, grid_for(db.ncls, 256), 256, st, db, deg);
  scan_u32(deg, off, n2, off + n2, tmp, st);
}
void k_els_fill(const DB& db, const u32* off, u32* cur, u32* adj, Stream* st) {
  fill_u32(cur, 0, 2 * (size_t)db.nvars + 2, st);
  if (db.ncls) KLAUNCH(els_fill_k, grid_for(db.ncls, 256), 256, st, db, off, cur, adj);
}
// Trimming: a literal without remaining in-edges or out-edges is an SCC of
// its own.  alive[l] = 1 for active literals still in the graph; a removed
// literal is its own root.  The first frontier holds the literals that
// start with no in- or out-edge.
__global__ void els_trim_init_k(DB db, const u32* off, u32* alive, u32* indeg, u32* outdeg, u32* root, u32* front,
                                u32* nfront) {
  __shared__ u32 lw[NW];
  const u32 n2 = 2 * db.nvars + 2, first = (blockIdx.x * blockDim.x + threadIdx.x) * SIPT;
  u32 bits = 0;
#pragma unroll
  for (u32 t = 0; t < (u32)SIPT; t++) {
    const u32 l = first + t;
    if (l >= n2) break;
    const bool act = l >= 2 && !db.vstate[ABS(l)];
    const u32 out = off[l + 1] - off[l], in = off[FLIP(l) + 1] - off[FLIP(l)];
    indeg[l] = in;
    outdeg[l] = out;
    root[l] = l;
    alive[l] = act && in && out ? 1u : 0u;
    if (act && !(in && out)) bits |= 1u << t;
  }
  u32 k = block_reserve(bits, lw, nfront);
  while (bits) {
    const u32 t = __ffs(bits) - 1;
    bits &= bits - 1;
    front[k++] = first + t;
  }
}
PF_DEV void els_drop(u32 w, u32* alive, u32* root, u32* next, u32* nnext) {
  if (atomicCAS(&alive[w], 1u, 0u) == 1u) {
    root[w] = w;
    next[atomicAdd(nnext, 1u)] = w;
  }
}
__global__ void els_trim_k(const u32* front, u32 n, const u32* off, const u32* adj, u32* alive, u32* indeg,
                           u32* outdeg, u32* root, u32* next, u32* nnext) {
  const u32 k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const u32 v = front[k];
  for (u32 j = off[v]; j < off[v + 1]; j++) {  // out-neighbours lose an in-edge
    const u32 w = adj[j];
    if (atomicSub(&indeg[w], 1u) == 1u) els_drop(w, alive, root, next, nnext);
  }
  const u32 nv = FLIP(v);
  for (u32 j = off[nv]; j < off[nv + 1]; j++) {  // in-neighbours lose an out-edge
    const u32 u = FLIP(adj[j]);
    if (atomicSub(&outdeg[u], 1u) == 1u) els_drop(u, alive, root, next, nnext);
  }
}
void k_els_trim_init(const DB& db, const u32* off, u32* alive, u32* indeg, u32* outdeg, u32* root, u32* front,
                     u32* nfront, Stream* st) {
  fill_u32(nfront, 0, 1, st);
  KLAUNCH(els_trim_init_k, grid_for(2 * (size_t)db.nvars + 2, STILE), ST, st, db, off, alive, indeg, outdeg, root,
          front, nfront);
}
void k_els_trim(const u32* front, u32 n, const u32* off, const u32* adj, u32* alive, u32* indeg, u32* outdeg, u32* root,
                u32* next, u32* nnext, Stream* st) {
  if (!n) return;
  KLAUNCH(els_trim_k, grid_for(n, 256), 256, st, front, n, off, adj, alive, indeg, outdeg, root, next, nnext);
}
// Coloring over the literals that survive trimming (list[0..n)): forward
// minimum labels, then from each root (color == itself) backwards inside
// its color; the marked literals form the root's SCC, whose root is its
// smallest literal.
__global__ void els_alive_k(u32 n2, const u32* alive, u32* list, u32* n) {
  __shared__ u32 lw[NW];
  const u32 first = (blockIdx.x * blockDim.x + threadIdx.x) * SIPT;
  u32 bits = 0;
#pragma unroll
  for (u32 t = 0; t < (u32)SIPT; t++)
    if (first + t < n2 && alive[first + t]) bits |= 1u << t;
  u32 k = block_reserve(bits, lw, n);
  while (bits) {
    const u32 t = __ffs(bits) - 1;
    bits &= bits - 1;
    list[k++] = first + t;
  }
}
__global__ void els_color_init_k(const u32* list, u32 n, u32* color, u32* mark) {
  const u32 k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  color[list[k]] = list[k];
  mark[list[k]] = 0;
}
__global__ void els_fw_k(const u32* list, u32 n, const u32* off, const u32* adj, const u32* alive, u32* color,
                         u32* changed) {
  const u32 k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const u32 v = list[k], c = color[v];
  for (u32 j = off[v]; j < off[v + 1]; j++) {
    const u32 w = adj[j];
    if (alive[w] && c < color[w] && atomicMin(&color[w], c) > c) *changed = 1;
  }
}
__global__ void els_roots_k(const u32* list, u32 n, const u32* color, u32* mark) {
  const u32 k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n && color[list[k]] == list[k]) mark[list[k]] = 1;
}
__global__ void els_bw_k(const u32* list, u32 n, const u32* off, const u32* adj, const u32* alive, const u32* color,
                         u32* mark, u32* changed) {
  const u32 k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const u32 v = list[k];
  if (mark[v]) return;
  const u32 c = color[v];
  for (u32 j = off[v]; j < off[v + 1]; j++) {
    const u32 w = adj[j];
    if (alive[w] && mark[w] && color[w] == c) {
      mark[v] = 1;
      *changed = 1;
      return;
    }
  }
}
__global__ void els_take_k(const u32* list, u32 n, const u32* mark, const u32* color, u32* alive, u32* root) {
  const u32 k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const u32 v = list[k];
  if (mark[v]) {
    root[v] = color[v];
    alive[v] = 0;
  }
}
// degrees inside the literals still alive (all counted before any is
// dropped), then the next trimming frontier
__global__ void els_redeg_k(const u32* list, u32 n, const u32* off, const u32* adj, const u32* alive, u32* indeg,
                            u32* outdeg) {
  const u32 k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const u32 v = list[k];
  if (!alive[v]) return;
  u32 out = 0, in = 0;
  for (u32 j = off[v]; j < off[v + 1]; j++) out += alive[adj[j]];
  const u32 nv = FLIP(v);
  for (u32 j = off[nv]; j < off[nv + 1]; j++) in += alive[FLIP(adj[j])];
  indeg[v] = in;
  outdeg[v] = out;
}
__global__ void els_redrop_k(const u32* list, u32 n, const u32* indeg, const u32* outdeg, u32* alive, u32* root,
                             u32* front, u32* nfront) {
  const u32 k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const u32 v = list[k];
  if (alive[v] && !(indeg[v] && outdeg[v])) els_drop(v, alive, root, front, nfront);
}
void k_els_alive(u32 n2, const u32* alive, u32* list, u32* n, Stream* st) {
  fill_u32(n, 0, 1, st);
  KLAUNCH(els_alive_k, grid_for(n2, STILE), ST, st, n2, alive, list, n);
}
void k_els_color_init(const u32* list, u32 n, u32* color, u32* mark, Stream* st) {
  if (n) KLAUNCH(els_color_init_k, grid_for(n, 256), 256, st, list, n, color, mark);
}
void k_els_fw(const u32* list, u32 n, const u32* off, const u32* adj, const u32* alive, u32* color, u32* changed,
              Stream* st) {
  fill_u32(changed, 0, 1, st);
  if (n) KLAUNCH(els_fw_k, grid_for(n, 256), 256, st, list, n, off, adj, alive, color, changed);
}
void k_els_roots(const u32* list, u32 n, const u32* color, u32* mark, Stream* st) {
  if (n) KLAUNCH(els_roots_k, grid_for(n, 256), 256, st, list, n, color, mark);
}
void k_els_bw(const u32* list, u32 n, const u32* off, const u32* adj, const u32* alive, const u32* color, u32* mark,
              u32* changed, Stream* st) {
  fill_u32(changed, 0, 1, st);
  if (n) KLAUNCH(els_bw_k, grid_for(n, 256), 256, st, list, n, off, adj, alive, color, mark, changed);
}
void k_els_take(const u32* list, u32 n, const u32* mark, const u32* color, u32* alive, u32* root, Stream* st) {
  if (n) KLAUNCH(els_take_k, grid_for(n, 256), 256, st, list, n, mark, color, alive, root);
}
void k_els_redeg(const u32* list, u32 n, const u32* off, const u32* adj, u32* alive, u32* indeg, u32* outdeg,
                 u32* root, u32* front, u32* nfront, Stream* st) {
  fill_u32(nfront, 0, 1, st);
  if (!n) return;
  KLAUNCH(els_redeg_k, grid_for(n, 256), 256, st, list, n, off, adj, alive, indeg, outdeg);
  KLAUNCH(els_redrop_k, grid_for(n, 256), 256, st, list, n, indeg, outdeg, alive, root, front, nfront);
}
// representatives (equivalence.cpp:66-118): rep[l] = root of l's SCC, l
// for assumed variables, 0 for inactive literals; info[0] |= 1 when a
// literal and its negation share an SCC, info[1] counts substituted literals
// grid-stride over a bounded number of workgroups: the count takes one
// atomic per workgroup (a single counter serialises its atomics)
__device__ inline void block_count_add(u32 mine, u32* lw, u32* out) {
  u32 tot;
  block_excl_scan(mine, lw, &tot);
  if (threadIdx.x == 0 && tot) atomicAdd(out, tot);
}
__global__ void els_rep_k(DB db, const u32* root, u32* rep, u32* info) {
  __shared__ u32 lw[NW];
  const u32 n2 = 2 * db.nvars + 2;
  u32 subs = 0;
  for (u32 l = blockIdx.x * blockDim.x + threadIdx.x; l < n2; l += gridDim.x * blockDim.x) {
    const u32 v = ABS(l);
    if (l < 2 || db.vstate[v]) {
      rep[l] = 0;
      continue;
    }
    const u32 r = root[l];
    if (root[FLIP(l)] == r) atomicOr(&info[0], 1u);
    rep[l] = db.assumed[v] ? l : r;
    subs += rep[l] != l;
  }
  block_count_add(subs, lw, &info[1]);
}
__global__ void count_nonzero_k(const u32* a, u32 n, u32* out) {
  __shared__ u32 lw[NW];
  u32 mine = 0;
  for (u32 i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) mine += a[i] != 0;
  block_count_add(mine, lw, out);
}
void k_count_nonzero(const u32* a, u32 n, u32* out, Stream* st) {
  fill_u32(out, 0, 1, st);
  if (n) KLAUNCH(count_nonzero_k, std::min(grid_for(n, ST), RED_BLOCKS), ST, st, a, n, out);
}
void k_els_rep(const DB& db, const u32* root, u32* rep, u32* info, Stream* st) {
  KLAUNCH(els_rep_k, std::min(grid_for(2 * (size_t)db.nvars + 2, ST), RED_BLOCKS), ST, st, db, root, rep, info);
}
// the clauses a substitution touches, with their position in the order
// learnts, then originals (lrank: exclusive scan of the learnt flags)
__global__ void els_learnt_k(DB db, u32* flag) {
  const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < db.ncls) flag[i] = st_learnt(cflags(db, db.c_off[i])) && !is_deleted(db, db.c_off[i]) ? 1u : 0u;
}
__global__ void els_touch_k(DB db, const u32* rep, const u32* lrank, u32 nlearnt, u32* list, u32* pos, u32* n) {
  __shared__ u32 lw[NW];
  const u32 first = (blockIdx.x * blockDim.x + threadIdx.x) * SIPT;
  u32 bits = 0;
#pragma unroll
  for (u32 t = 0; t < (u32)SIPT; t++) {
    const u32 i = first + t;
    if (i >= db.ncls) break;
    const u32 c = db.c_off[i];
    if (is_deleted(db, c)) continue;
    const u32* l = clits(db, c);
    const u32 sz = csize(db, c);
    bool hit = false;
    for (u32 q = 0; q < sz && !hit; q++) hit = rep[l[q]] != l[q];
    if (hit) bits |= 1u << t;
  }
  u32 k = block_reserve(bits, lw, n);
  while (bits) {
    const u32 t = __ffs(bits) - 1;
    bits &= bits - 1;
    const u32 i = first + t, c = db.c_off[i];
    list[k] = c;
    pos[k++] = st_learnt(cflags(db, c)) ? lrank[i] : nlearnt + (i - lrank[i]);
  }
}
void k_els_touch(const DB& db, const u32* rep, u32* lrank, void* tmp, u32* nlearnt, u32* list, u32* pos, u32* n,
                 Stream* st) {
  fill_u32(n, 0, 1, st);
  fill_u32(nlearnt, 0, 1, st);
  if (!db.ncls) return;
  KLAUNCH(els_learnt_k, grid_for(db.ncls, 256), 256, st, db, lrank);
  scan_u32(lrank, lrank, db.ncls, nlearnt, tmp, st);
}
void k_els_touch_list(const DB& db, const u32* rep, const u32* lrank, u32 nlearnt, u32* list, u32* pos, u32* n,
                      Stream* st) {
  if (db.ncls) KLAUNCH(els_touch_k, grid_for(db.ncls, STILE), ST, st, db, rep, lrank, nlearnt, list, pos, n);
}
// one pass of the unit-time fixpoint: the unit each touched clause yields
// under the current times (ptime: 2 p + sign, the smallest position wins)
__global__ void els_check_k(DB db, const u32* list, const u32* pos, u32 n, ElsView e, u32* ptime) {
  const u32 k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  u32 u = 0, m = 0;
  if (els_clause(db, list[k], e, pos[k], false, &u, &m) == 3) atomicMin(&ptime[ABS(u)], 2 * pos[k] + SIGN(u));
}
__global__ void els_commit_k(u32 n1, const u32* ptime, u32* utime, u32* changed) {
  const u32 v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v < n1 && ptime[v] != utime[v]) {
    utime[v] = ptime[v];
    *changed = 1;
  }
}
void k_els_check(const DB& db, const u32* list, const u32* pos, u32 n, const ElsView& e, u32* ptime, Stream* st) {
  fill_u32(ptime, NOPOS, (size_t)db.nvars + 1, st);
  if (n) KLAUNCH(els_check_k, grid_for(n, 256), 256, st, db, list, pos, n, e, ptime);
}
void k_els_commit(u32 n1, const u32* ptime, u32* utime, u32* changed, Stream* st) {
  fill_u32(changed, 0, 1, st);
  KLAUNCH(els_commit_k, grid_for(n1, 256), 256, st, n1, ptime, utime, changed);
}
// the substitution itself; units as (position, literal) records; info[0]
// |= 1 on an empty clause, info[1] counts clauses rewritten to two literals
__global__ void els_apply_k(DB db, const u32* list, const u32* pos, u32 n, ElsView e, u32* ukey, u32* uval, u32* nu,
                            u32* info) {
  const u32 k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  u32 u = 0, m = 0;
  const u32 kind = els_clause(db, list[k], e, pos[k], true, &u, &m);
  if (kind == 2) atomicOr(&info[0], 1u);
  if (kind == 3) {
    const u32 r = atomicAdd(nu, 1u);
    ukey[r] = pos[k];
    uval[r] = u;
  }
  if (kind == 4 && m == 2) atomicAdd(&info[1], 1u);
}
void k_els_apply(const DB& db, const u32* list, const u32* pos, u32 n, const ElsView& e, u32* ukey, u32* uval, u32* nu,
                 u32* info, Stream* st) {
  fill_u32(nu, 0, 1, st);
  if (n) KLAUNCH(els_apply_k, grid_for(n, 256), 256, st, db, list, pos, n, e, ukey, uval, nu, info);
}
// units in position order onto the trail (enqueueUnit)
__global__ void els_enqueue_k(DB db, const u32* uval, u32 n, u32* trail) {
  const u32 k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const u32 u = uval[k];
  trail[k] = u;
  db.assign[ABS(u)] = u + 1;
  db.vstate[ABS(u)] = VS_FROZEN;
}
void k_els_enqueue(const DB& db, const u32* uval, u32 n, u32* trail, Stream* st) {
  if (n) KLAUNCH(els_enqueue_k, grid_for(n, 256), 256, st, db, uval, n, trail);
}
// substituted variables (equivalence.cpp:173-190): an active variable whose
// positive literal has another representative gets a binary record on the
// stack (MODEL::saveBinary, model.cpp:280-288) and, when the
// representative's variable is active, the state SUBSTITUTED
__global__ void els_subst_flags_k(DB db, const u32* rep, u32* flag) {
  const u32 v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v > db.nvars) return;
  flag[v] = v && !db.vstate[v] && rep[V2L(v)] != V2L(v) ? 1u : 0u;
}
__global__ void els_subst_k(DB db, const u32* rep, const u32* flag, const u32* pos, u32* stack, u32* nmarked) {
  __shared__ u32 lw[NW];
  u32 marked = 0;
  for (u32 v = blockIdx.x * blockDim.x + threadIdx.x; v <= db.nvars; v += gridDim.x * blockDim.x) {
    if (!flag[v]) continue;
    const u32 p = V2L(v), other = rep[p];
    u32* w = stack + 5 * (size_t)pos[v];
    w[0] = V2L(db.vorg[v]);
    w[1] = V2L(db.vorg[ABS(other)]) | SIGN(FLIP(other));
    w[2] = 2;
    w[3] = V2L(db.vorg[v]) | 1u;
    w[4] = 1;
    // the representative's variable is the smallest of its SCC, never marked here
    if (!db.vstate[ABS(other)]) {
      db.vstate[v] = VS_SUBSTITUTED;
      marked++;
    }
  }
  block_count_add(marked, lw, nmarked);
}
void k_els_subst_flags(const DB& db, const u32* rep, u32* flag, u32* pos, void* tmp, u32* total, Stream* st) {
  KLAUNCH(els_subst_flags_k, grid_for((size_t)db.nvars + 1, 256), 256, st, db, rep, flag);
  scan_u32(flag, pos, (size_t)db.nvars + 1, total, tmp, st);
}
void k_els_subst(const DB& db, const u32* rep, const u32* flag, const u32* pos, u32* stack, u32* nmarked,
                 Stream* st) {
  fill_u32(nmarked, 0, 1, st);
  KLAUNCH(els_subst_k, std::min(grid_for((size_t)db.nvars + 1, ST), RED_BLOCKS), ST, st, db, rep, flag, pos, stack,
          nmarked);
}

}  // namespace pfs
