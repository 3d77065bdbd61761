// Host-side launch interface of the simplifier kernels and the small
// device runtime they need.  The round driver (pfs_driver.cpp) only talks
// to the device through these functions.
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "pfs_items.hpp"

namespace pfs {

struct Stream;  // a HIP stream plus reusable pinned staging memory

// ---- runtime
Stream* stream_create(int device);
void stream_destroy(Stream*);
void stream_sync(Stream*);
// waits without throwing and drops staged reads (error paths, destructors);
// false when the stream reported an error
bool stream_quiesce(Stream*);
void stream_bind(Stream*);  // make the stream's device current for the calling thread
void* dev_alloc(size_t bytes);  // throws std::bad_alloc on failure
void dev_free(void*);
void h2d(void* dst, const void* src, size_t bytes, Stream*);
void d2h(void* dst, const void* src, size_t bytes, Stream*);
void d2d(void* dst, const void* src, size_t bytes, Stream*);
void fill_u32(u32* dst, u32 value, size_t n, Stream*);  // fill kernel, no memset nodes
void fill_u8(u8* dst, u8 value, size_t n, Stream*);
// stage timing: events recorded on the stream
// the position of a stream at a point in its queue (a timing-free event)
struct Marker;
Marker* marker_record(Stream*);
bool marker_done(Marker*);  // the stream has passed it (no wait)
void marker_wait(Marker*);
void marker_free(Marker*);
bool stream_wait(Stream*);  // waits for the stream's queued work; false on a device error
void dev_mem_info(size_t* free_bytes, size_t* total_bytes);  // of the current device
struct Event;
Event* event_record(Stream*);
float event_ms(Event* a, Event* b);  // synchronises on b
void event_free(Event*);
const char* device_name(int device);
// per-kernel profiling: events around every launch while enabled
struct KernelTime {
  std::string name;
  double ms;
  uint64_t launches;
};
void prof_enable(Stream*, bool on);
void prof_collect(Stream*, std::vector<KernelTime>& out);  // adds to out, synchronises
int device_count();

// ---- primitives
size_t scan_tmp_bytes(size_t n);
// exclusive prefix sum of n u32 values; out may alias in; *total (device)
// receives the sum when non-null
void scan_u32(const u32* in, u32* out, size_t n, u32* total, void* tmp, Stream*);
size_t radix_tmp_bytes(size_t n);
// stable sort of (key, val) pairs by the low 'bits' bits of key
void radix_sort_pairs(u32* keys, u32* vals, size_t n, int bits, void* tmp, Stream*);
// stable sort of db.c_off[0..ncls) by (size, first, last literal, signature);
// keys: ncls words, tmp: radix_tmp_bytes(ncls)
void k_aggr_sort(const DB& db, u32* keys, void* tmp, Stream*);

// ---- occurrence table
void k_hist(const DB& db, u32* hist, Stream*);
void k_scatter(const DB& db, u32* cursor, Stream*);
// the same table built by literal ranges (no global atomic per occurrence)
constexpr u32 OTP_MAX_RANGES = 12288;
struct OtpPlan {
  int bits;       // literals per range: 1 << bits
  u32 nr;         // ranges
  u32 ct;         // clauses per tile
  u32 nblk;       // tiles
  size_t ncounts; // nr * nblk
  int xcd;        // tiles spread by XCD: each of the 8 XCDs takes a run of consecutive tiles
};
// false when the literal space needs more than OTP_MAX_RANGES ranges;
// tile_log2: clauses per tile (0: planned)
bool otp_plan(size_t n2, size_t ncls, OtpPlan* p, int tile_log2 = 0, int xcd = 0);
// counts: p.ncounts + 1 words; pairs: 8-byte aligned, two words per live
// literal occurrence with 'table', else one.  Without 'table' only hist is
// written; with it also ot_off (n2 + 1 words), ot_size and ot_refs, and
// hist = ot_size.  tmp: scan_tmp_bytes(p.ncounts)
void k_ot_ranges(const DB& db, const OtpPlan& p, u32 n2, u32* counts, void* pairs, bool table, u32* hist,
                 void* tmp, Stream*);
// sort the lists of the given literals (lits[0..n)) by the total order
// classes: 5n words of scratch for the size classes; cnt: 3 device words;
// tmp: scan_tmp_bytes(n)
// lsum[i]: sum of the clause sizes of list lits[i] (n words)
void k_sort_lists(const DB& db, const u32* lits, u32 n, u32* classes, u32* cnt, u32* scratch, u32* lsum, void* tmp,
                  Stream*);

// ---- BCP
void k_bcp_visit(const DB& db, const u32* front, u32 nfront, u32* next, u32* nnext, u32* conflict,
                 Stream*);
void k_bcp_mark(const DB& db, const u32* lits, u32 n, u8* mark, Stream*);
void k_bcp_clean(const DB& db, u8* mark, Stream*);
void k_clear_lists(const DB& db, const u32* lits, u32 n, Stream*);
void k_set_vstate(const DB& db, const u32* lits, u32 n, u8 state, Stream*);

// ---- LCVE
struct LcveArgs {
  const u32* hist;  // per literal
  u32* score;       // per variable
  u64* hard;        // min (score, var) over hard breakers
  u8* cflag;        // per variable: 1 candidate, 3 soft breaker candidate, 4 passed over, 0 none
  u32 pt, nt;
  u32 max_occurs;   // electionsmax: break past this many occurrences of a literal
  u32 clause_max;   // lcveclausemax: pass over candidates with a longer live clause
};
void k_lcve_classify(const DB& db, const LcveArgs& a, Stream*);
// compact candidates with (score, var) < hard into cand/key (var order)
void k_lcve_select(const DB& db, const LcveArgs& a, u32* flag_tmp, u32* cand, u32* key, u32* ncand,
                   void* tmp, Stream*);
void k_lcve_positions(const u32* cand, u32 n, u32* pos, Stream*);
// candidates passed over for an oversize clause (cflag 4) start decided and
// not elected: they neither block nor freeze a neighbour
void k_lcve_reject(const u32* cand, u32 n, const u8* cflag, u32* pos, u32* state, Stream*);
// one election step over the candidates idx[0..n) (all when null); blk: per
// candidate, an undecided earlier neighbour (NOPOS before the first step);
// resume: per candidate, the list entry its scan goes on from (0 before the
// first step); the still undecided candidates are appended to next (count
// in *nnext)
// the same over the list idx[0..*n_in) (n_in on the device, at most n_bound);
// *iters counts the passes over a non-empty list; *zero is cleared (the
// counter the next pass appends to)
// lds: bytes of LDS reserved per one-wavefront workgroup (occupancy);
// lanes: lanes per candidate (1, 2, 4 or 8)
void k_mis_step_dev(const DB& db, const u32* cand, const u32* idx, const u32* n_in, u32 n_bound, u32* pos,
                    u32* state, u32* blk, u32* resume, u32* next, u32* nnext, u32* iters, u32* zero, u32 lds,
                    u32 lanes, Stream*);
void k_mis_step(const DB& db, const u32* cand, const u32* idx, u32 n, u32* pos, u32* state, u32* blk,
                u32* resume, u32* next, u32* nnext, u32 lds, u32 lanes, Stream*);
void k_mis_break(const u32* cand, u32 n, const u8* cflag, const u32* state, u32* brk, Stream*);
void k_mis_collect(const u32* cand, u32 n, const u32* state, u32 brk, u32* flag_tmp, u32* out,
                   u32* nout, void* tmp, Stream*);

// ---- per elected variable stages
void k_list_sizes(const DB& db, const u32* pv, u32 n, u32 extra, u32* out, Stream*);  // |poss|+|negs|+extra
// SUB over the elected indices idx[0..n) (all of 0..n when idx is null)
// work: 2 n + 32 words of scratch (size classes and placement); lsum: the
// list size sums of k_sort_lists over the elected literals (2k, 2k + 1);
// tmp: radix_tmp_bytes(n); snap (nullable): take the re-run snapshot of
// every variable run (as k_sub_snapshot would before the launch)
void k_sub(const DB& db, const Options& o, const u32* pv, const u32* idx, u32 n, const SubOut& out, u32* work,
           const u32* lsum, void* tmp, const SubSnap* snap, Stream*);
void k_sub_snapshot(const DB& db, const u32* pv, u32 n, const SubSnap& s, Stream*);
void k_sub_restore(const DB& db, const u32* pv, const u32* idx, u32 n, const SubSnap& s, Stream*);
// aff[o] = 1 for every owner o of a clause holding a changed unit (changes:
// triples elected index, literal, pass) that comes after the deriving
// variable and last ran (lastrun[o]) no later than the change's pass
void k_sub_affected(const DB& db, const SubSnap& s, const u32* changes, u32 nch, const u32* lastrun, u32* aff,
                    Stream*);
void k_set_at(u32* dst, const u32* idx, u32 n, u32 value, Stream*);  // dst[idx[i]] = value
void k_set_vis(u32* ust, u32* ulit, const u32* entries, u32 n, Stream*);  // triples variable, index, literal
// indices of the nonzero flags, in order; pos: n words of scratch
void k_pack_flags(const u32* flags, u32 n, u32* pos, u32* out, u32* count, void* tmp, Stream*);
void k_oblivion(const DB& db, const u32* lits, u32 n, Stream*);  // delete all clauses of the lists
void k_count_nonzero(const u32* a, u32 n, u32* out, Stream*);    // *out = number of nonzero a[i]
// work, lsum, tmp: as for k_sub (the variables the gate checks leave open
// are staged by size class)
void k_ve_decide(const DB& db, const Options& o, const u32* pv, u32 n, const VeDecision& d, const u32* order,
                 u32* work, const u32* lsum, void* tmp, Stream*);
// work, lsum, tmp: as for k_sub
void k_ve_apply(const DB& db, const u32* pv, u32 n, const VeDecision& d, const VeApply& a, u32* work,
                const u32* lsum, void* tmp,
                Stream*);
void k_bce(const DB& db, const Options& o, const u32* pv, const u32* type, u32 n, u32* words,
           const u32* base, u32* stack, bool count, const u32* order, Stream*);
// heaviest-first permutation of the elected indices (key: scratch of n words)
void k_work_order(const DB& db, const u32* pv, u32 n, u32* key, u32* order, void* tmp, Stream*);
void k_ere(const DB& db, const Options& o, const u32* pv, u32 from, u32 n, const EreOut& out,
           Stream*);
void k_ere_dirty(const DB& db, const u32* pv, u32 from, u32 n, const u32* del_by, u32* first,
                 Stream*);
void k_del_by(const u32* rec, u32 nrec, u32* del_by, Stream*);
void k_apply_deletions(const DB& db, const u32* ids, u32 n, Stream*);
// pair-parallel ERE
void k_ht_build(const DB& db, const HashTable& ht, int maxsize, Stream*);
void k_ere_pair_counts(const DB& db, const Options& o, const u32* pv, u32 n, u32* np, u64* total, Stream*);
void k_ere_pairs(const DB& db, const ErePairs& a, u32 npairs, Stream*);
// the full match lists of the n overflow records a.ovf[0..n) (offsets off[n])
void k_ere_all_matches(const DB& db, const ErePairs& a, u32 n, const u32* off, u32* out, Stream*);

// ---- small helpers
void k_assign_units(const DB& db, const u32* lits, u32 n, Stream*);  // assign + FROZEN
void k_pv_lits(const u32* pv, u32 n, u32* lits, Stream*);            // 2 literals per variable
void k_max_list(const DB& db, const u32* lits, u32 n, u32* out, Stream*);
void k_reduce_lists(const DB& db, Stream*);  // drop deleted entries from every list
void k_prep(const DB& db, u32 first, Stream*);  // clauses [first, ncls): sort literals, signature
void k_zero_sizes(const DB& db, const u32* lits, u32 n, Stream*);
void k_clear_added(const DB& db, Stream*);  // resolvent marks of an earlier simplify call

// ---- DRAT proof (binary format of src/cpu/proof.cpp:94-175)
// lines: (kind, payload) pairs, kinds PR_*; stamp: per variable, the line
// number + 1 of the unit line that assigned it in this batch (0 = assigned
// before).  A PR_ADD line whose clause holds a literal made true before it
// is dropped (PROOF::addResolvent).  bytes[i] = encoded size of line i.
void k_proof_sizes(const DB& db, const u32* lines, u32 n, const u32* stamp, u32* bytes, Stream*);
void k_proof_write(const DB& db, const u32* lines, u32 n, const u32* bytes, const u32* off, u8* out, Stream*);
// stamp[pairs[2i]] = pairs[2i+1], or 0 when clear
void k_set_stamps(u32* stamp, const u32* pairs, u32 n, bool clear, Stream*);

// ---- counting and garbage collection
void k_count(const DB& db, u64* out2, Stream*);  // live clauses, live literals
// live occurrences and clause bytes in the elected variables' lists
void k_elected_work(const DB& db, const u32* pv, u32 n, u64* out2, Stream*);
// O_c and D_c of the election: occurrences in the candidates' lists and the
// bytes of their distinct clauses (mark: per handle, zero before and after)
void k_cand_work(const DB& db, const u32* cand, u32 n, u8* mark, u64* out2, Stream*);
// stable compaction of the live clauses into dst (ids and pool renumbered)
void k_compact(const DB& src, DB& dst, u32* tmp_cls, u32* tmp_lit, void* tmp, u32* totals,
               Stream*);
// the same with the live originals first and the live learnt clauses after
// them (each part in clause order): the order in which the reference
// extracts a formula it wrote back (orgs, then learnts; newBeginning,
// src/cpu/simplify.cpp:251-269, and extract, :147-149).  totals: 4 words.
void k_compact_originals_first(const DB& src, DB& dst, u32* tmp_cls, u32* tmp_lit, void* tmp, u32* totals,
                               Stream*);

// ---- variable map (Solver::map, src/cpu/vmap.cpp:61-135; VMAP, src/cpu/vmap.h)
// mapped[0..nvars] (old variable -> new, 0 dropped) and
// info[0..3] = {first root-assigned variable (0: none), new variable count,
// true literal of first in the new numbering, inactive variables};
// info[4] is scratch; pos_tmp holds nvars+1 words
void k_map_build(const u8* vstate, const u32* assign, u32 nvars, u32* pos_tmp, u32* mapped, u32* info, void* tmp,
                 Stream*);
void k_map_clauses(const DB& db, const VarMap& m, Stream*);
void k_map_vars(const VarMap& m, u32 nvars_old, const u32* vorg, const u8* vstate, const u8* assumed, u32* vorg2,
                u8* vstate2, u8* assumed2, u32* assign2, Stream*);
void k_map_lits(const VarMap& m, u32* lits, u64 n, Stream*);
// the device address of the fixed scratch array (SCRATCH_WORDS words) of
// 'device', which must be current; looked up once per device
constexpr size_t SCRATCH_WORDS = (size_t)1 << 20;  // 4 MiB
u32* scratch_lookup(int device);
// ---- output packing: offsets[ncls+1], literals, flags, lbd of the live
// clauses in clause order (start_tmp: ncls words)
void k_pack_cnf(const DB& db, u32* start_tmp, void* tmp, u32* total, u64* offsets, u32* lits, u8* flags, u32* lbd,
                Stream*);

// ---- DIMACS ingestion (src/cpu/dimacs.cpp:41-170)
// lexer over the clause part of a text in 4096-byte tiles (text padded with
// blanks to whole tiles); modes: 0 special bytes (comment starts counted per
// tile, first NUL and '%' token in mark[0], mark[1]), 1 comment start
// positions, 2 first '%' token outside comments, 3 literal tokens counted
// per tile, 4 literals written (cnt: exclusive scan of mode 3's counts);
// malformed tokens set err bit 1 (first at mark[2]), variables above maxvar
// bit 2
enum LexMode : u32 { LEX_SPECIAL = 0, LEX_COMMENTS = 1, LEX_PERCENT = 2, LEX_COUNT = 3, LEX_EMIT = 4 };
enum LexErr : u32 { LEXE_TOKEN = 1, LEXE_VAR = 2, LEXE_SIZE = 4, LEXE_UNSAT = 8 };
struct LexArgs {
  const u8* text;
  u64 begin, end;                // clause part [begin, end)
  const u8* mask;                // comment bytes (null: none)
  u32* cnt;                      // per tile: counts, or their exclusive scan (emit modes)
  u64* pos;                      // comment starts (mode 1)
  u32* lits;                     // literals, 0 ends a clause (mode 4)
  unsigned long long* mark;      // first NUL, first '%' token, first malformed token
  u32* err;
  u32 maxvar;
};
u64 lex_tiles(u64 n);
void k_lex(u32 mode, const LexArgs& a, u64 ntiles, Stream*);
void k_lex_mask(const u8* text, u64 end, const u64* pos, u32 n, u8* mask, Stream*);
// ends[c] = index of the c-th 0 in lits[0..n); *total = number of 0s
void k_clause_ends(const u32* lits, u64 n, u32* tmp_pos, void* tmp, u32* total, u32* ends, Stream*);
// clause i of the pool from the literals between 0s (err bit 4: too long)
void k_parse_build(const DB& db, const u32* lits, const u32* ends, u32* err, Stream*);
// sorted clauses: duplicates dropped, tautologies and units removed
// (ulit[i]: the unit), err bit 8: empty clause
void k_parse_dedup(const DB& db, u32* ulit, u32* err, Stream*);
// DRAT lines of an in-place rewrite of the clauses (old: the pool before
// it, same handles): the nu units, then the new forms of the changed
// clauses, then the deletions of their old forms and of the clauses gone;
// off null: bytes per line (nu + 2 ncls) into bytes; else written at out + off[i]
void k_rewrite_proof(const DB& db, const u32* old, const u32* units, u32 nu, const u32* off, u8* out, u32* bytes,
                     Stream*);
// DRAT lines of the ingestion from its end state: the nu root units, then
// per text clause its deletion / shrunken addition (parse_proof_clause);
// off null: the byte count of each line into bytes[nu + nc]; else the
// lines written at out + off[i]
void k_parse_proof(const DB& db, const u32* toks, const u32* ends, const u32* units, u32 nu, u32 nc,
                   const u32* off, u8* out, u32* bytes, Stream*);
// unit clauses in order: first of each variable assigned and written to
// the trail (*total of them), opposite units set err bit 8; first needs
// nvars+1 words, win and pos n words
void k_parse_units(const DB& db, const u32* ulit, u32 n, u32* first, u32* win, u32* pos, void* tmp, u32* total,
                   u32* trail, u32* err, Stream*);
void iota_u32(u32* a, size_t n, Stream*);

// ---- equivalent-literal substitution (src/cpu/equivalence.cpp:41-306)
// implication graph CSR over literals: deg and off (2V+3 words, off[2V+2] =
// edges), then adj (k_els_fill)
void k_els_graph(const DB& db, u32* deg, u32* off, void* tmp, Stream*);
void k_els_fill(const DB& db, const u32* off, u32* cur, u32* adj, Stream*);
// trimming: alive, degrees and roots of every literal, first frontier
void k_els_trim_init(const DB& db, const u32* off, u32* alive, u32* indeg, u32* outdeg, u32* root, u32* front,
                     u32* nfront, Stream*);
void k_els_trim(const u32* front, u32 n, const u32* off, const u32* adj, u32* alive, u32* indeg, u32* outdeg,
                u32* root, u32* next, u32* nnext, Stream*);
// coloring over the alive literals
void k_els_alive(u32 n2, const u32* alive, u32* list, u32* n, Stream*);
void k_els_color_init(const u32* list, u32 n, u32* color, u32* mark, Stream*);
void k_els_fw(const u32* list, u32 n, const u32* off, const u32* adj, const u32* alive, u32* color, u32* changed,
              Stream*);
void k_els_roots(const u32* list, u32 n, const u32* color, u32* mark, Stream*);
void k_els_bw(const u32* list, u32 n, const u32* off, const u32* adj, const u32* alive, const u32* color, u32* mark,
              u32* changed, Stream*);
void k_els_take(const u32* list, u32 n, const u32* mark, const u32* color, u32* alive, u32* root, Stream*);
void k_els_redeg(const u32* list, u32 n, const u32* off, const u32* adj, u32* alive, u32* indeg, u32* outdeg,
                 u32* root, u32* front, u32* nfront, Stream*);
// representatives; info[0]: conflict, info[1]: substituted literals
void k_els_rep(const DB& db, const u32* root, u32* rep, u32* info, Stream*);
// touched clauses and their positions (learnts first): lrank = exclusive
// scan of the learnt flags (k_els_touch), then the list (k_els_touch_list)
void k_els_touch(const DB& db, const u32* rep, u32* lrank, void* tmp, u32* nlearnt, u32* list, u32* pos, u32* n,
                 Stream*);
void k_els_touch_list(const DB& db, const u32* rep, const u32* lrank, u32 nlearnt, u32* list, u32* pos, u32* n,
                      Stream*);
void k_els_check(const DB& db, const u32* list, const u32* pos, u32 n, const ElsView& e, u32* ptime, Stream*);
void k_els_commit(u32 n1, const u32* ptime, u32* utime, u32* changed, Stream*);
void k_els_apply(const DB& db, const u32* list, const u32* pos, u32 n, const ElsView& e, u32* ukey, u32* uval,
                 u32* nu, u32* info, Stream*);
void k_els_enqueue(const DB& db, const u32* uval, u32 n, u32* trail, Stream*);
void k_els_subst_flags(const DB& db, const u32* rep, u32* flag, u32* pos, void* tmp, u32* total, Stream*);
void k_els_subst(const DB& db, const u32* rep, const u32* flag, const u32* pos, u32* stack, u32* nmarked, Stream*);

}  // namespace pfs
