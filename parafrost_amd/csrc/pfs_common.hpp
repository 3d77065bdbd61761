// Shared types and per-item helpers of the HIP simplifier.
//
// Everything in this header is plain C++ usable from device code (HIP
// kernels) and host code; PF_HD marks functions compiled for both.
//
// Clause database layout in HBM: see struct DB below.  Clause ids follow
// the simplifier's clause order (src/cpu/simplify.cpp scnf), so the id
// doubles as the deterministic tie-break.  Occurrence table: for literal l,
// ot_refs[ot_off[l] .. ot_off[l] + ot_size[l]) are the clause ids of the
// list.  ot_size follows the size the sequential simplifier would observe
// (entries of deleted clauses stay in a list until the list is compacted or
// cleared).
#pragma once

#include <cstdint>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define PF_HD __host__ __device__ inline
#define PF_DEV __device__ inline
#else
#define PF_HD inline
#define PF_DEV inline
#endif

namespace pfs {

typedef uint32_t u32;
typedef uint64_t u64;
typedef uint8_t u8;
typedef int64_t i64;

constexpr u8 ST_ORIGINAL = 0, ST_LEARNT = 1, ST_DELETED = 2, ST_MASK = 3;
constexpr u8 F_MOLTEN = 4, F_ADDED = 8;
constexpr u8 USAGE_SHIFT = 4, USAGE_MASK = 0x30;
constexpr u8 VS_MELTED = 1, VS_FROZEN = 2, VS_SUBSTITUTED = 4;
constexpr u32 NOPOS = 0xFFFFFFFFu;
constexpr int HSE_MAX_CL_SIZE = 1000;  // src/cpu/subsume.h:25

PF_HD u32 ABS(u32 l) { return l >> 1; }
PF_HD u32 FLIP(u32 l) { return l ^ 1u; }
PF_HD u32 NEGL(u32 l) { return l | 1u; }
PF_HD u32 SIGN(u32 l) { return l & 1u; }
PF_HD u32 V2L(u32 v) { return v << 1; }
PF_HD u32 MAPHASH(u32 l) { return 1u << (l & 31u); }

PF_HD bool st_original(u8 f) { return (f & ST_MASK) == ST_ORIGINAL; }
PF_HD bool st_learnt(u8 f) { return (f & ST_LEARNT) != 0; }
PF_HD bool st_deleted(u8 f) { return (f & ST_DELETED) != 0; }

// Simplifier options (names and defaults of the reference CPU simplifier,
// src/cpu/options.cpp:24-46, normalised as in options.cpp:280-311).
struct Options {
  int32_t phases = 5, mu_pos = 32, mu_neg = 32, lits_min = 500, ve_clause_limit = 100,
          xor_max_arity = 10, sub_limit = 3000, bce_limit = 3000, ere_limit = 3000,
          ere_max_resolvent = 1000, lcve_min = 1, shrink_rate = 2, lbd_tier1 = 2;
  int32_t ve_en = 1, ve_plus_en = 1, sub_en = 1, bce_en = 0, ere_en = 1, all_en = 0,
          ve_lbound_en = 0;
  // launch shaping only (results do not depend on it): process the elected
  // variables heaviest-first
  int32_t work_order = 1;
  int32_t proof_en = 0;  // write binary DRAT lines (src/cpu/proof.cpp)
  // limits of the GPU build's election (src/gpu/options.cpp:34-36, used by
  // lcve_k, src/gpu/lcve.cu:94 and depFreeze_d :43-52): the election stops
  // at the first candidate with more than lcve_max_occurs occurrences of a
  // literal, and a candidate with a live clause longer than lcve_clause_max
  // is passed over.  The CPU simplifier has neither (INT32_MAX: off).
  int32_t lcve_max_occurs = INT32_MAX, lcve_clause_max = INT32_MAX;
  // stable sort of the result by the list key (aggresivesort,
  // src/cpu/simplify.cpp:263, src/gpu/cnf.cu:232-233)
  int32_t aggr_sort = 0;
  // synchronise the stream after every stage (syncalways, src/gpu/options.cu:33)
  int32_t sync_always = 0;
  // occurrence-table build by literal ranges: log2 of the clauses per tile
  // (0: planned, 32 clauses per range and tile up to 16,384 clauses) and
  // the tile-to-XCD mapping (consecutive tiles on one XCD, so that the
  // partition's writes to one range meet in one L2)
  int32_t otp_tile = 0, otp_xcd = 1;
  // the table by literal ranges (1) or by one atomic per occurrence (0,
  // hist_k / scatter_k: also the build past OTP_MAX_RANGES ranges)
  int32_t otp_ranges = 1;
  // LDS bytes reserved per one-wavefront workgroup of the election's steps
  // (occupancy; on the heavy-tailed round 0 / 10 / 20 / 32 KiB gave the
  // first pass 3.6 / 3.8 / 6.0 / 10.2 ms and the worklist passes 1.02 /
  // 0.89 / 0.80 / 0.79 ms; the Tseitin round is fastest at 20 KiB);
  // -1: tuned, else 10240
  int32_t mis_lds = -1;
  // lanes per candidate in the election steps (1, 2, 4, 8; 0: tuned, else 1)
  int32_t mis_lanes = 0;
  // launch-shape tuner: on the first call of a formula class (clause count,
  // variable count and mean clause size, each bucketed) with at least
  // TUNE_MIN_CLAUSES clauses, the occurrence-table tile and the election's
  // LDS reservation that are not set are chosen by timing the candidates on
  // that call's first round; later calls of the class reuse the choice.
  // Results do not depend on either.
  int32_t autotune = 1;
};
constexpr uint64_t TUNE_MIN_CLAUSES = 1ull << 20;

// Device view of the clause database and per-variable state.  Root
// assignments are kept per variable (assign[v] = true literal + 1) so that
// parallel unit propagation can claim a variable with one compare-and-swap.
//
// A clause lives in the word pool 'lits' as a three-word header followed by
// its literals; one clause visit reads one contiguous run:
//   word 0: size (bits 0-23) | flags (bits 24-31: status, molten, added, usage)
//   word 1: signature (OR of 1 << (lit & 31))
//   word 2: LBD (learnt clauses)
//   word 3..: literals, sorted ascending
// A clause is named by its handle, the pool offset of its header (the
// reference's S_REF, src/gpu/datatypes.hpp:35): occurrence lists hold
// handles, so a visit costs no extra indirection.  Handles grow with the
// clause order (clauses are laid out in order; resolvents are appended in
// elected order; compaction is stable), so comparing handles compares
// clause positions.  c_off lists the handles in clause order for the
// clause-parallel kernels.  The flags byte and the size bytes are written
// with separate byte/short stores, so a flag update never rewrites the size.
struct DB {
  u32* lits;      // word pool: headers and literals
  u32* c_off;     // clause index -> handle
  u32* ot_off;
  u32* ot_size;
  u32* ot_refs;   // clause handles
  u32* assign;    // per variable: 0 unassigned, else the true literal + 1
  u8* vstate;     // per variable
  u8* assumed;    // per variable
  u32* vorg;      // per variable: original numbering
  u32 nvars;
  u32 ncls;
};

constexpr u32 CL_HDR = 3;                 // header words per clause
constexpr u32 CL_MAX_SIZE = 0xFFFFFFu;     // largest clause the header holds
PF_HD u32 make_hdr(u32 size, u8 flags) { return size | ((u32)flags << 24); }
PF_HD u32 hdr_size(u32 h) { return h & CL_MAX_SIZE; }
PF_HD u8 hdr_flags(u32 h) { return (u8)(h >> 24); }
PF_HD u32* chead(const DB& db, u32 c) { return db.lits + c; }
PF_HD u32 csize(const DB& db, u32 c) { return hdr_size(db.lits[c]); }
PF_HD u8 cflags(const DB& db, u32 c) { return hdr_flags(db.lits[c]); }
PF_HD u32 csig(const DB& db, u32 c) { return db.lits[c + 1]; }
PF_HD u32 clbd(const DB& db, u32 c) { return db.lits[c + 2]; }
PF_HD u32* clits(const DB& db, u32 c) { return db.lits + c + CL_HDR; }
PF_HD void set_cflags(const DB& db, u32 c, u8 f) { ((u8*)(db.lits + c))[3] = f; }
PF_HD void set_csize(const DB& db, u32 c, u32 n) {
  u32* h = db.lits + c;
  ((uint16_t*)h)[0] = (uint16_t)n;
  ((u8*)h)[2] = (u8)(n >> 16);
}
PF_HD void set_csig(const DB& db, u32 c, u32 sig) { db.lits[c + 1] = sig; }
PF_HD void set_clbd(const DB& db, u32 c, u32 lbd) { db.lits[c + 2] = lbd; }

}  // namespace pfs
