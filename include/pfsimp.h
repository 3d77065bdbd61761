/* pfsimp.h -- C ABI of the MI355X-native GPU simplifier (SIGmA inprocessing
 * engine).  A drop-in for the reference's simplifier boundary: the member
 * Solver::simplify() of the GPU build (src/gpu/simplify.cu:57-75,
 * src/gpu/solver.hpp:772) and Solver::sigmify() of the CPU build
 * (src/cpu/simplify.cpp:115-129).  The host solver hands over its root-level
 * clause database and variable state, runs the simplify rounds, and takes
 * back the simplified clauses, the new root units, the eliminated-variable
 * marks and the eliminated-variable stack for model reconstruction.
 *
 * Conventions: literals are 2*v + sign (v >= 1); plain pointers and sizes
 * only; the library never keeps a caller pointer after a call returns and
 * never exits the process; one thread uses a context at a time; any number
 * of contexts may share a device.
 */
#ifndef PFSIMP_H
#define PFSIMP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pfs_ctx pfs_ctx;

/* Status codes.  PFS_ALLOC_FAIL plays the role of the reference's
 * AWAKEN_FAIL / CNFALLOC_FAIL / OTALLOC_FAIL simplifier states
 * (src/gpu/constants.cuh:28-31): the caller keeps its own clauses and skips
 * GPU simplification. */
typedef enum {
  PFS_OK = 0,
  PFS_ALLOC_FAIL = 1,
  PFS_BAD_INPUT = 2,
  PFS_DEVICE_ERROR = 3,
  PFS_NO_DEVICE = 4,
  PFS_BAD_STATE = 5,
  PFS_UNKNOWN_OPTION = 6
} pfs_status;

/* Outcome of pfs_simplify (the reference's cnfstate after sigmify,
 * src/cpu/simplify.cpp:217-226 and learnEmpty on a refutation). */
typedef enum { PFS_RESULT_UNSAT = 0, PFS_RESULT_SAT = 1, PFS_RESULT_UNSOLVED = 2 } pfs_result;

typedef struct {
  int32_t device;      /* HIP device ordinal */
  uint64_t mem_budget; /* device bytes this context may hold; 0 = no limit */
  int32_t keep_input;  /* keep a device copy of the input for pfs_rewind */
} pfs_config;

typedef struct {
  uint32_t elected, eliminated, units, resolvents;
  uint64_t clauses_after, literals_after;
  uint32_t mis_iterations, sub_passes, ere_passes;
  /* stages that ran in the round: bit s for pfs_stage_name(s) */
  uint32_t stages;
  /* work measures, filled when "profilegpu" is on: live clauses and
   * literals at round start, live occurrences in the elected variables'
   * lists and their clause bytes (12-byte header + 4 bytes per literal) */
  uint64_t clauses_before, literals_before, elected_occurs, elected_bytes;
  /* live occurrences in the election candidates' lists and the bytes of
   * their distinct clauses ("profilegpu"); resolvent words (headers and
   * literals) and eliminated-variable stack words written */
  uint64_t cand_occurs, cand_bytes, resolvent_words, stack_words;
} pfs_round_stats;

/* Creates a context on a device.  Replaces the allocation part of the
 * reference's awaken() (src/gpu/simplify.cu:77-134). */
pfs_status pfs_create(const pfs_config* cfg, pfs_ctx** out);
void pfs_destroy(pfs_ctx* ctx);

/* Device arena.  Every context's device memory is sub-allocated on the host
 * from a pool reserved once per device (the reference's cuArena pool behind
 * cuMM::initDeviceArena, src/gpu/memory.cu:99-166): a stable region (the
 * per-variable arrays, placed from the top) and a dynamic region (clause
 * blocks, occurrence tables, temporaries, placed from the bottom),
 * defragmented by moving the blocks of idle contexts over the holes, dynamic
 * ones down and stable ones up, when an allocation finds no hole
 * (compact_gpu_dynamic and retry, memory.cu:191-219, 360-419).  So the
 * rounds, serving requests and batch runs call neither hipMalloc nor
 * hipFree once the pool is reserved.
 *
 * pfs_arena_configure: size of the default arena's first slab and the most
 * its slabs may reach (0: grows while the device has memory); before the
 * device's first context only (PFS_BAD_STATE afterwards).  Without it the
 * first slab is a quarter of the free device memory (at most 64 GiB), or
 * PFS_ARENA_MB megabytes from the environment.
 * pfs_arena_create: a private arena of exactly 'bytes' on one device that
 * never grows: contexts made with pfs_create_in_arena share it, and an
 * allocation that cannot fit even after compaction fails with
 * PFS_ALLOC_FAIL.  Destroy it after every context in it.
 * pfs_arena_get_stats: of a private arena, or (arena null) of the default
 * arena of 'device'. */
typedef struct pfs_arena pfs_arena;
typedef struct {
  uint64_t capacity, used, peak, largest_hole, holes, slabs;
  uint64_t compactions, moved_blocks, moved_bytes, slab_allocs, pending;
} pfs_arena_stats;
pfs_status pfs_arena_configure(int32_t device, uint64_t first_slab, uint64_t limit);
pfs_status pfs_arena_create(int32_t device, uint64_t bytes, pfs_arena** out);
void pfs_arena_destroy(pfs_arena* arena);
pfs_status pfs_create_in_arena(const pfs_config* cfg, pfs_arena* arena, pfs_ctx** out);
pfs_status pfs_arena_get_stats(pfs_arena* arena, int32_t device, pfs_arena_stats* out);

/* Sets a simplifier option by its reference name (src/cpu/options.cpp:24-46,
 * src/gpu/options.cpp:24-43): booleans "ve", "ve+", "sub", "bce", "ere",
 * "all", "velbound", "profilegpu"; integers "phases", "mupos", "muneg",
 * "vephasemin", "veclausemax", "xormaxarity", "submax", "bcemax", "eremax",
 * "ereclausemax", "lcvemin", "gcfreq", "lbdtier1".  GPU-build aliases
 * ("eliminatedlitsmin", "resolventmax", "submaxoccurs", "bcemaxoccurs",
 * "eremaxoccurs", "electionsmin", "collectfreq", "velitsbound") are accepted.
 * "proof" (boolean, default off) records the binary DRAT lines of the run
 * (pfs_download_proof).
 * "decomposemin" (1-10, default 2) is pfs_decompose's default round count.
 * "mapperc" (real in [0, 1], default 0.2) is the inactive share above which
 * pfs_map renumbers.
 * "workorder" (boolean, default on; no reference counterpart) launches the
 * per-variable kernels heaviest lists first; results do not depend on it.
 * GPU-build options (src/gpu/options.cpp:24-43, options.cu:36-60):
 * "electionsmax" and "lcveclausemax" cap the election (default: no cap, as
 * the CPU simplifier); "aggresivesort" sorts the result's clauses by the
 * list key; "syncalways" synchronises after every stage; "lcvefast" and
 * "vefunction" are accepted off and refused on (PFS_BAD_INPUT with the
 * reason: the rank-based election and the function-table gates are not
 * provided); "nstreams", "*minthreads", "*minblocks", "subclausemax",
 * "ereextend" and the solver's "sigma*" / "solve" options are range-checked
 * and have no effect here.
 * Launch shapes (no reference counterpart; results do not depend on them):
 * "otptile" (0-20: log2 clauses per occurrence-table tile, 0 tuned or
 * planned), "otpxcd" (boolean, default on: consecutive tiles on one XCD),
 * "otpranges" (boolean, default on: the occurrence table by literal ranges;
 * off: one atomic per occurrence, the build used past 12,288 ranges),
 * "mislanes" (0, 1, 2, 4, 8: lanes per election candidate, 0 tuned),
 * "mislds" (-1-65536: LDS bytes reserved per election workgroup, -1 tuned),
 * "autotune" (boolean, default on: see pfs_launch_shape). */
pfs_status pfs_set_option(pfs_ctx* ctx, const char* name, const char* value);

/* Hands over the root-level clause database, as the host solver has it after
 * rootify()/shrinkTop() (src/cpu/simplify.cpp:163-164; the GPU build's
 * extractCNF/reflectCNF, src/gpu/cnf.cu:166-184).
 *   offsets[ncls+1], lits[offsets[ncls]] : clauses (any literal order)
 *   flags[ncls]  (nullable): bit0 learnt, bits 4-5 usage
 *   lbd[ncls]    (nullable): learnt clause LBD
 *   vstate[maxvar+1] (nullable): 1 melted, 2 frozen (root-assigned), 4 substituted
 *   assumed[maxvar+1] (nullable): assumption variables are never eliminated
 *   vorg[maxvar+1]   (nullable): current -> original variable map for the stack
 *   trail[ntrail]: the root units already assigned and propagated */
pfs_status pfs_upload(pfs_ctx* ctx, uint32_t maxvar, uint64_t ncls, const uint64_t* offsets,
                      const uint32_t* lits, const uint8_t* flags, const uint32_t* lbd,
                      const uint8_t* vstate, const uint8_t* assumed, const uint32_t* vorg,
                      const uint32_t* trail, uint32_t ntrail);

/* Equivalent-literal substitution on the resident formula, on the device:
 * the CDCL side's ELS (Solver::ELS/decompose/substitute,
 * src/cpu/equivalence.cpp:41-323).  Up to max_rounds (0: option
 * "decomposemin", default 2, src/cpu/options.cpp:87) decompositions while
 * one succeeds (new units or binary clauses): the SCCs of the binary
 * implication graph over active variables, each represented by its literal
 * of smallest variable (assumed variables keep themselves; a literal and
 * its negation in one SCC refute), clauses rewritten with the
 * representatives (learnts first; duplicates dropped, tautologies and
 * root-satisfied clauses removed, units assigned), the units propagated,
 * substituted variables marked (state 4) with binary records on the stack
 * (MODEL::saveBinary).  Afterwards the clauses are shrunk at the top level
 * (root-satisfied clauses removed, false literals dropped).  Requires a
 * propagated formula (after pfs_upload, pfs_parse_dimacs or pfs_simplify).
 * New units accumulate in pfs_download_units, stack records in
 * pfs_download_stack; info.status 0 means refuted (the next pfs_simplify
 * reports UNSAT).  No proof lines are recorded for it.  Learnt clauses keep
 * their LBD and usage when rewritten (the reference may re-create them). */
typedef struct pfs_els_info {
  uint32_t rounds, status;  /* status 0: refuted, 2: otherwise */
  uint64_t substituted, units, stack_words;
} pfs_els_info;
pfs_status pfs_decompose(pfs_ctx* ctx, uint32_t max_rounds, pfs_els_info* info);

/* DIMACS ingestion on the device: the text (nbytes, need not be
 * NUL-terminated) becomes the resident formula, replacing pfs_upload.
 * Replaces the reference's parser (Solver::parse/toClause,
 * src/cpu/dimacs.cpp:41-170) together with the root propagation and the
 * top-level shrink that precede the first simplify (Solver(), solver.cpp:49;
 * rootify/shrinkTop, simplify.cpp:163-164, shrink.cpp:152-216).  The
 * header ("p cnf V C", after blanks and 'c' comment lines) is read on the
 * host; the clause part is lexed, split, deduplicated and propagated on the
 * device.  Duplicate literals are dropped, tautologies removed, unit
 * clauses assigned in text order, root units propagated, satisfied clauses
 * removed and false literals dropped; the remaining clauses keep their text
 * order.  Comment lines may appear anywhere; the clause part ends at the
 * first NUL byte or '%' token (SATLIB files).  Errors (PFS_BAD_INPUT, the
 * reason in pfs_last_error): no or malformed header, a malformed token
 * (numbers are separated by blanks), a literal above V, an unterminated last
 * clause, more remaining clauses of two or more literals than C.  An empty
 * clause or opposite units make info.status 0 and the next pfs_simplify
 * report UNSAT.  The root trail is pfs_download_trail; the clauses,
 * vstate and sizes are readable before pfs_simplify through
 * pfs_result_sizes, pfs_download_cnf and pfs_download_vstate.
 * With option "proof" the call records DRAT lines from the text's clauses
 * to the resident formula (pfs_download_proof; the parser's lines,
 * dimacs.cpp:245-305, and the top-level shrink's, shrink.cpp:122): every
 * root unit, then per text clause the deletion of a clause that is gone or
 * the addition of what is left of a clause that lost literals with the
 * deletion of its text, and the empty clause for a refuted text. */
typedef struct pfs_parse_info {
  uint32_t maxvar;          /* V of the header */
  uint32_t status;          /* 0: refuted while parsing/propagating, 2: otherwise */
  uint64_t header_clauses;  /* C of the header */
  uint64_t read_clauses;    /* clauses in the text */
  uint64_t clauses, literals;  /* resident formula after propagation */
  uint64_t trail;           /* root units */
} pfs_parse_info;
pfs_status pfs_parse_dimacs(pfs_ctx* ctx, const char* text, uint64_t nbytes, pfs_parse_info* info);
/* The root trail of the resident formula (*n literals; lits nullable to
 * query the size): the input trail, the units of a parse, or after pfs_map
 * its one stand-in literal. */
pfs_status pfs_download_trail(pfs_ctx* ctx, uint32_t* lits, uint64_t* n);

/* Incremental use (the reference's incremental front end iadd/itoClause,
 * src/cpu/addinc.cpp:44-160, and assumptions, src/cpu/assume.cpp): the
 * formula stays resident in device memory between calls.
 * pfs_append adds original clauses over variables 1..maxvar (maxvar may grow)
 * after the current clauses: literals false at the root are dropped,
 * root-satisfied clauses and tautologies skipped, unit clauses assigned
 * (propagated by the next pfs_simplify), an empty clause makes the next
 * pfs_simplify report UNSAT.  The results and stack of later calls refer to
 * the whole resident formula; stack words and units accumulate.
 * pfs_assume replaces the set of assumed variables (never eliminated). */
pfs_status pfs_append(pfs_ctx* ctx, uint32_t maxvar, uint64_t ncls, const uint64_t* offsets,
                      const uint32_t* lits);
pfs_status pfs_assume(pfs_ctx* ctx, const uint32_t* vars, uint32_t n);

/* Restores the last uploaded input from its device copy (keep_input). */
pfs_status pfs_rewind(pfs_ctx* ctx);

/* Runs the simplify rounds (src/cpu/simplify.cpp:157-236; the GPU build's
 * simplifying(), src/gpu/simplify.cu:136-241).  *result receives a pfs_result. */
pfs_status pfs_simplify(pfs_ctx* ctx, int32_t* result);

/* Result sizes: clauses, literals, new root units, stack words, rounds.
 * Before the first pfs_simplify they describe the resident input. */
pfs_status pfs_result_sizes(pfs_ctx* ctx, uint64_t* ncls, uint64_t* nlits, uint64_t* nunits,
                            uint64_t* nstack, uint32_t* nrounds);
/* Simplified clauses in simplifier order (the reference's scnf order):
 * offsets[ncls+1], lits, flags (bit0 learnt, bit1 added resolvent, bits 4-5
 * usage), lbd.  Any pointer may be null.  Replaces cacheCNF/writeBackCNF
 * (src/gpu/cnf.cu:186-236). */
pfs_status pfs_download_cnf(pfs_ctx* ctx, uint64_t* offsets, uint32_t* lits, uint8_t* flags,
                            uint32_t* lbd);
/* New root units in assignment order (cacheUnits, src/gpu/transfer.cu:62-81). */
pfs_status pfs_download_units(pfs_ctx* ctx, uint32_t* lits);
/* Variable states (maxvar+1 bytes): 1 eliminated, 2 root-assigned
 * (cacheEliminated/markEliminated, src/gpu/transfer.cu:42-75). */
pfs_status pfs_download_vstate(pfs_ctx* ctx, uint8_t* vstate);
/* Eliminated-variable stack words to append to the model's stack, in
 * original numbering (model.resolved; cacheResolved, src/gpu/transfer.cu:83-96). */
pfs_status pfs_download_stack(pfs_ctx* ctx, uint32_t* words);
/* Binary DRAT proof of the last pfs_simplify when option "proof" is on:
 * the lines the reference simplifier writes through its PROOF object
 * (src/cpu/proof.cpp:94-229, called from bve.cpp:214-215,
 * elimination.cpp:229-230, equivalence.h:65-83, subsume.h:114,
 * redundancy.h:137, solver.h:251-254 and 475-479; the GPU build's cuPROOF,
 * src/gpu/proof.cu:161-254): 'a'/'d', literals 2*vorg(v)+sign as 7-bit
 * groups, 0.  Lines appended by the host solver's own proof go around it. */
/* Compact variable renumbering after pfs_simplify, on the device: the
 * reference's `if (canMap()) map(true)` (src/cpu/simplify.cpp:227;
 * Solver::map, src/cpu/vmap.cpp:61-135; VMAP::initiate and mapLit,
 * src/cpu/vmap.h:40-78).  Applied when force != 0 or when more than
 * "mapperc" (option, default 0.2, src/cpu/options.cpp:146) of the variables
 * are inactive (eliminated, root-assigned or substituted; canMap,
 * src/cpu/solver.h:164-169); otherwise nothing changes and applied = 0.
 * Variables in state 0 keep their order and become 1..new_maxvar; the first
 * root-assigned variable is kept too and stands for every root-assigned
 * literal (its new literal true at the root is trail_lit, the whole trail
 * after the map; 0 when nothing is assigned).  Afterwards the resident
 * clauses, vstate, vorg and the assumed set use the new numbering (so do
 * pfs_download_cnf, pfs_download_vstate, pfs_download_units and later
 * pfs_append/pfs_simplify calls); the stack and proof stay in original
 * numbering through vorg.  Download units and vstate first when the old
 * numbering is wanted.  Requires no pending root units (call it right after
 * pfs_simplify); pfs_rewind is unavailable afterwards.  Variables added by
 * a later pfs_append receive fresh original ids. */
typedef struct pfs_map_info {
  uint32_t applied, old_maxvar, new_maxvar, first_frozen, trail_lit, inactive;
} pfs_map_info;
pfs_status pfs_map(pfs_ctx* ctx, int32_t force, pfs_map_info* info);
/* Old variable -> new variable of the last pfs_map (old_maxvar+1 words; 0:
 * dropped; the identity when the map was not applied). */
pfs_status pfs_download_map(pfs_ctx* ctx, uint32_t* mapped);
/* VMAP::mapOrgs (src/cpu/vmap.h:112-131) on the device: literals of the old
 * numbering (model.lits, assumptions) mapped in place; 0 stays 0, a
 * root-assigned literal becomes trail_lit or its negation by value, a
 * dropped unassigned literal becomes 0. */
pfs_status pfs_map_lits(pfs_ctx* ctx, uint32_t* lits, uint64_t n);
/* pfs_map_lits for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid allocation); no
 * counterpart in the reference.  The literals are staged through a fixed
 * 4 MiB device array of the library (one per device, held until the
 * process ends) instead of a buffer of the context: the call allocates
 * nothing, host or device.  At most 1,048,576 literals per call
 * (PFS_BAD_INPUT beyond).  Every call on one device shares the array:
 * calls that overlap on a device race. */
pfs_status pfs_map_lits_noalloc(pfs_ctx* ctx, uint32_t* lits, uint64_t n);
/* The binary DRAT lines of the last pfs_simplify or pfs_parse_dimacs call
 * with option "proof" (proof.cpp: the reference's writer): a caller that
 * keeps one proof file appends each call's lines. */
pfs_status pfs_proof_size(pfs_ctx* ctx, uint64_t* nbytes);
pfs_status pfs_download_proof(pfs_ctx* ctx, uint8_t* bytes);
pfs_status pfs_round_stats_get(pfs_ctx* ctx, pfs_round_stats* out, uint32_t max, uint32_t* n);
/* Per-stage device milliseconds (when "profilegpu" is on) and stage calls. */
pfs_status pfs_stage_times(pfs_ctx* ctx, double* ms, uint64_t* calls, uint32_t n);
const char* pfs_stage_name(uint32_t stage);
uint32_t pfs_stage_count(void);
/* Per-kernel device time of the last pfs_simplify when "profilegpu" is on:
 * events recorded on the context's stream around every launch, summed per
 * kernel (the MI355X counterpart of the reference's -profilegpu timers,
 * src/gpu/timer.cuh:27-58). */
uint32_t pfs_kernel_count(pfs_ctx* ctx);
pfs_status pfs_kernel_stats(pfs_ctx* ctx, uint32_t i, const char** name, double* total_ms,
                            uint64_t* launches);
/* The launch shapes of the last pfs_simplify: the occurrence table's tile
 * (log2 of the clauses per tile; 0 where the table was not built by literal
 * ranges), the LDS bytes reserved per election-step workgroup and the lanes
 * per election candidate, and whether that call chose them by timing the
 * candidates (option "autotune": the first call of a formula class with at
 * least 2^20 clauses measures every tile, lane count and reservation not
 * set by "otptile" / "mislanes" / "mislds"; later calls of the class reuse
 * the fastest).  Replaces the reference's launch tuner
 * (src/gpu/grid.cuh:137-229); results do not depend on the shapes. */
pfs_status pfs_launch_shape(pfs_ctx* ctx, int32_t* otp_tile_log2, int32_t* mis_lds, int32_t* mis_lanes,
                            int32_t* tuned);
pfs_status pfs_memory(pfs_ctx* ctx, uint64_t* used, uint64_t* peak);

/* Batched simplification (no reference counterpart: the reference
 * simplifies one formula per process, src/gpu/simplify.cu:43-55).  Runs
 * pfs_simplify on n contexts that already hold their formulas, with
 * 'workers' host threads; each context has its own stream, so the rounds of
 * different formulas overlap on the device and their host-side
 * synchronisation overlaps in time.  A context is used by one worker at a
 * time, so every result equals the context's own pfs_simplify.
 * results[i] receives context i's pfs_result, statuses[i] (nullable) its
 * status, ms[i] (nullable) its wall time.  Returns the first non-OK status. */
pfs_status pfs_batch_simplify(pfs_ctx* const* ctxs, uint32_t n, uint32_t workers, int32_t* results,
                              pfs_status* statuses, double* ms);
/* Request server for incremental sessions (continuous batching; no
 * reference counterpart).  A session is a context whose formula stays
 * resident; a request appends clauses (pfs_append, maxvar may grow),
 * replaces the assumed variables (pfs_assume) and simplifies (pfs_simplify).
 * Requests carry a release time in ms after pfs_server_start.  'workers'
 * host threads take the earliest released request whose session is idle
 * and, when mem_budget > 0, whose session's device memory fits next to the
 * sessions running; requests of one session run one at a time in submission
 * order, so each result equals the session's result alone.  The request's
 * data are copied at submission.  pfs_server_wait blocks until a request
 * is done and returns its status, result and release/start/end times (ms).
 * pfs_server_destroy waits for every submitted request. */
typedef struct pfs_server pfs_server;
pfs_status pfs_server_create(uint32_t workers, uint64_t mem_budget, pfs_server** out);
pfs_status pfs_server_submit(pfs_server* srv, pfs_ctx* session, double release_ms, uint32_t maxvar,
                             uint64_t ncls, const uint64_t* offsets, const uint32_t* lits,
                             const uint32_t* assumed, uint32_t nassumed, uint32_t* ticket);
pfs_status pfs_server_start(pfs_server* srv);
pfs_status pfs_server_wait(pfs_server* srv, uint32_t ticket, int32_t* result, double* release_ms,
                           double* start_ms, double* end_ms);
void pfs_server_destroy(pfs_server* srv);

const char* pfs_last_error(const pfs_ctx* ctx);
const char* pfs_version(void);
int32_t pfs_device_count(void);

#ifdef __cplusplus
}
#endif

#endif /* PFSIMP_H */
