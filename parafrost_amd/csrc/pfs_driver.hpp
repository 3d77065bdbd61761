// Round driver of the HIP simplifier: one Simplifier per context owns the
// device state of one formula (clause database, occurrence table,
// per-variable state, eliminated-variable stack) and runs the simplify
// rounds of the reference (src/cpu/simplify.cpp:157-236) on it.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "pfs_arena.hpp"
#include "pfs_launch.hpp"

namespace pfs {

// Device memory owned by a context: blocks of the device arena (one pool
// per device, pfs_arena.hpp) accounted against the context's budget
// (admission control for contexts that share a device).  Per-variable
// arrays live in the arena's stable region; the clause blocks, the
// occurrence table and the temporaries in its dynamic region, which
// compaction slides down between calls.
class DeviceMemory {
 public:
  explicit DeviceMemory(size_t budget) : budget_(budget) {}
  struct Buf {
    void* p = nullptr;
    size_t cap = 0;
  };
  void attach(DeviceArena* arena, Stream* st) {
    arena_ = arena;
    client.arena = arena;
    client.st = st;
  }
  // make b hold at least 'bytes'; keeps the first 'keep' bytes of content
  void ensure(Buf& b, size_t bytes, size_t keep, Stream* st, Region r = R_DYNAMIC);
  void release(Buf& b, bool idle = false);  // idle: the stream has no queued work
  // buffer records hand blocks over only through these, so that the arena
  // keeps the right pointer to rewrite when it moves a block
  void take(Buf& dst, Buf& src);  // dst's block released, dst = src, src empty
  void swap(Buf& a, Buf& b);
  size_t used() const { return client.used; }
  size_t peak() const { return peak_; }
  size_t budget() const { return budget_; }
  void set_budget(size_t b) { budget_ = b; }
  DeviceArena* arena() const { return arena_; }
  ArenaClient client;

 private:
  size_t budget_, peak_ = 0;
  DeviceArena* arena_ = nullptr;
};

struct AllocFail {
  std::string what;
};
struct DeviceFail {
  std::string what;
};

struct RoundStats {
  uint32_t elected = 0, eliminated = 0, units = 0, resolvents = 0;
  uint64_t clauses_after = 0, literals_after = 0;
  uint32_t mis_iterations = 0, sub_passes = 0, ere_passes = 0;
  // work measures (filled when profiling): live clauses/literals at round
  // start, occurrences and clause bytes in the elected variables' lists
  uint64_t clauses_before = 0, literals_before = 0, elected_occurs = 0, elected_bytes = 0;
  uint32_t stages = 0;  // bit s: stage s (enum Stage) ran in this round
  // more work measures (profiling): occurrences and distinct clause bytes in
  // the election candidates' lists; resolvent and stack words written
  uint64_t cand_occurs = 0, cand_bytes = 0, resolvent_words = 0, stack_words = 0;
};

// per-stage device time (ms) and launches, when profiling is enabled
enum Stage { S_GC, S_OT, S_BCP, S_LCVE, S_SORT, S_SUB, S_VE, S_BCE, S_ERE, S_COUNT, S_NSTAGES };
extern const char* const kStageNames[S_NSTAGES];

// a formula class's tuned launch shapes (0 / -1 and false: not measured)
struct Tuned {
  int32_t otp_tile = 0, mis_lds = -1, mis_lanes = 0;
  bool ot = false, mis = false;
};

class Simplifier {
 public:
  // arena: the pool the context's device memory comes from (null: the
  // device's default arena)
  Simplifier(int device, size_t budget, bool keep_input, DeviceArena* arena = nullptr);
  ~Simplifier();
  void free_all();  // every device block back to the arena (the destructor does it too)

  Options opt;
  bool profile = false;

  void upload(uint32_t maxvar, uint64_t ncls, const uint64_t* offsets, const uint32_t* lits,
              const uint8_t* flags, const uint32_t* lbd, const uint8_t* vstate, const uint8_t* assumed,
              const uint32_t* vorg, const uint32_t* trail, uint32_t ntrail);
  // DIMACS text -> resident formula on the device (parse, src/cpu/dimacs.cpp:41-170,
  // then root propagation and the top-level shrink, src/cpu/shrink.cpp:152-216)
  struct ParseInfo {
    uint32_t maxvar = 0, status = 2;
    uint64_t header_clauses = 0, read_clauses = 0, clauses = 0, literals = 0, trail = 0;
  };
  ParseInfo parse_dimacs(const char* text, uint64_t n);
  uint64_t trail_size() const { return trail_n_; }
  // equivalent-literal substitution on the resident formula (Solver::ELS /
  // decompose / substitute, src/cpu/equivalence.cpp:41-323): up to
  // max_rounds decompositions while one succeeds, then the top-level shrink
  struct ElsInfo {
    uint32_t rounds = 0, status = 2;
    uint64_t substituted = 0, units = 0, stack_words = 0;
  };
  ElsInfo decompose(uint32_t max_rounds);
  void download_trail(uint32_t* lits);  // the root trail of the resident formula
  void rewind();                    // restore the last uploaded input (device copy)
  void bind() { stream_bind(st_); }  // select this context's device on the calling thread
  // after an exception: wait for the stream and drop staged reads whose
  // destinations belonged to the unwound call
  void on_error() { stream_quiesce(st_); }
  // the arena client: held locked while a call works on the device; at the
  // start of a call (no device address held yet) the context's own blocks
  // may be moved by a compaction the arena asked for
  ArenaClient& client() { return mem_.client; }
  void call_begin() {
    DeviceArena* a = mem_.arena();
    if (a && a->wants_compaction()) a->compact(&mem_.client, true);
  }
  double lits_mul = 1.0;  // resolvent room at upload, as a multiple of the input (literalsmul)
  // incremental use: add clauses / replace the assumed variables of the
  // resident formula (the state after the last simplify)
  void append(uint32_t maxvar, uint64_t ncls, const uint64_t* offsets, const uint32_t* lits);
  void assume(const uint32_t* vars, uint32_t n);
  int simplify(uint32_t max_rounds); // 0 UNSAT, 1 SAT, 2 UNSOLVED

  // results (valid after simplify)
  uint64_t out_ncls() const { return ncls_; }
  // literals of the result (after the final compaction every clause is live)
  uint64_t out_nlits() const { return lits_end_ - CL_HDR * ncls_; }
  void download_cnf(uint64_t* offsets, uint32_t* lits, uint8_t* flags, uint32_t* lbd);
  const std::vector<uint32_t>& new_units() const { return units_out_; }
  void download_vstate(uint8_t* vstate);
  uint64_t stack_words() const { return stack_n_; }
  void download_stack(uint32_t* words);
  const std::vector<RoundStats>& rounds() const { return rounds_; }
  // binary DRAT proof of the last simplify call (when opt.proof_en)
  uint64_t proof_bytes() const { return proof_n_; }
  void download_proof(uint8_t* bytes);
  double stage_ms[S_NSTAGES] = {0};
  uint64_t stage_calls[S_NSTAGES] = {0};
  std::vector<KernelTime> kernel_times;  // per kernel, when profiling is enabled
  // compact variable renumbering of the resident formula (Solver::map,
  // src/cpu/vmap.cpp:61-135), applied when forced or when more than
  // map_perc of the variables are inactive (canMap, src/cpu/solver.h:164-169)
  struct MapInfo {
    uint32_t applied, old_nvars, new_nvars, first, trail_lit, inactive;
  };
  MapInfo map(bool force, double map_perc);
  void download_map(uint32_t* mapped);  // old variable count + 1 words
  void map_lits(uint32_t* lits, uint64_t n);  // VMAP::mapOrgs on caller literals
  // the same staged through the device's fixed scratch array instead of a
  // context buffer: no allocation of any kind; n <= SCRATCH_WORDS
  void map_lits_noalloc(uint32_t* lits, uint64_t n);
  // the occurrence-table tile (log2 clauses; 0 where the table was not
  // built by ranges) and election LDS bytes of the last simplify, and
  // whether it measured them
  int32_t used_tile() const { return used_tile_; }
  int32_t used_lds() const { return used_lds_; }
  int32_t used_lanes() const { return used_lanes_; }
  bool tuned_now() const { return tuned_now_; }
  size_t mem_used() const { return mem_.used(); }
  size_t mem_peak() const { return mem_.peak(); }
  uint32_t nvars() const { return V_; }

 private:
  typedef DeviceMemory::Buf Buf;
  template <class T>
  T* P(Buf& b) { return (T*)b.p; }

  int simplify_rounds(uint32_t max_rounds);
  void alloc_state();
  DB make_db(int which);
  void gc();
  void originals_first();
  void parse_proof(const u32* toks, const u32* ends, u32 nc);  // the ingestion's DRAT lines
  void rewrite_proof(const u32* old, uint64_t from, bool empty, u32 pre);  // an in-place rewrite's DRAT lines
  void aggr_sort();
  void count(uint64_t& cls, uint64_t& lits);
  void create_ot();
  // live-literal counts into hist_; with 'table' also the occurrence table
  void ot_counts(bool table);
  bool otp_room(const OtpPlan& p);  // pair buffer of the range build; false: use the atomics
  void set_capacities(uint64_t ncls, uint64_t words);
  bool prop();
  bool prop_visit(uint64_t front);  // propagation from trail[front..): assignments only
  void prop_clean(uint64_t from);   // satisfied clauses deleted, false literals dropped
  bool lcve();
  void sort_ot();
  bool sub();
  bool ve();
  void bce();
  void ere();
  bool ere_pairs();   // pair-parallel ERE; false when it must fall back
  void ere_per_var(); // per-variable ERE with speculative re-runs
  void ensure_clause_room(uint64_t add_cls, uint64_t add_lits);
  void ensure_stack(uint64_t words);
  void grow_vars(uint32_t maxvar);
  void commit_units(const std::vector<uint32_t>& lits);
  // proof lines: queued per stage on the host as (kind, payload), encoded
  // on the device by proof_flush() before clause ids can change
  struct ProofEv {
    u32 k, seq, kind, payload;
  };
  void proof_line(u32 kind, u32 payload);
  void proof_events(std::vector<ProofEv>& ev);  // sorts by (k, seq), queues up to the first empty clause
  void proof_flush();
  ProofOut proof_out(u32 cap);                  // device record buffer for a per-variable kernel
  void proof_collect(std::vector<ProofEv>& ev, u32 below_k);  // records of the last kernel with k < below_k
  uint32_t read_u32(const void* dptr);
  const u32* work_order() { return opt.work_order ? P<u32>(order_) : nullptr; }
  void stage_begin(Stage s);
  void stage_end();

  int device_ = 0;
  Stream* st_ = nullptr;
  u32* fixed_scratch_ = nullptr;  // the device's fixed scratch array (its device address)
  DeviceMemory mem_;
  bool keep_input_;

  // sizes
  uint32_t V_ = 0;
  uint64_t ncls_ = 0, lits_end_ = 0, ccap_ = 0, lcap_ = 0;  // lits_end_, lcap_: pool words
  int cur_ = 0;  // which clause block is live
  // clause blocks (dynamic region)
  Buf c_off_[2], lits_[2];  // lits_: word pool of headers + literals; c_off_: handles in clause order
  // occurrence table
  Buf ot_off_, ot_size_, ot_refs_, hist_;
  Buf otp_;  // occurrence-table build: range counts, then (literal, handle) pairs
  // per variable
  Buf assign_, vstate_, assumed_, vorg_, trail_;
  uint64_t trail_n_ = 0, propagated_ = 0;
  // LCVE
  Buf score_, cflag_, flag_, cand_, key_, pos_, state_, pv_;
  uint32_t E_ = 0;
  // per elected variable
  Buf vsz_, vbase_, type_, def_, vncls_, vnlits_, vnstack_, clsb_, litb_, stkb_;
  Buf snap_ref_, snap_flg_, snap_size_, snap_sig_, snap_lbd_, snap_np_, snap_nn_, scratch_;
  Buf elits_, unit_by_, recs_, ids_, mark_, ht_, erecs_, order_, subu_;
  Buf vrest_;  // SUB, VE: size classes and placement of the variables
  Buf lsum_;   // per elected literal (2k, 2k + 1): clause size sum of its list at the list sort
  Buf subaff_, sublast_, ust_, ulit_;  // SUB re-runs: affected flags, last pass, visible stage units
  // proof
  Buf proof_, precs_, pstamp_, plines_, pbytes_;
  uint64_t proof_n_ = 0;
  std::vector<u32> pl_;                          // queued lines: kind, payload
  std::vector<std::pair<u32, u32>> pst_;         // (variable, line number + 1) of queued units
  // stack
  Buf stack_;
  uint64_t stack_n_ = 0;
  // temporaries
  Buf tmp_, tmp2_, tmp3_, scal_;
  // pristine input copy
  Buf in_c_off_, in_lits_, in_assign_, in_vstate_;
  uint64_t in_ncls_ = 0, in_nlits_ = 0, in_trail_ = 0;
  uint64_t in_live_cls_ = 0, in_live_lits_ = 0;  // live clauses and literals of the kept input
  // launch-shape tuning (Options::autotune): the class's choices (ot, mis:
  // known), whether this call's first round measures them, and what the
  // last call used
  Tuned tune_;
  uint64_t tune_key_ = 0;
  bool tune_ot_ = false, tune_mis_ = false, tuned_now_ = false;
  int32_t used_tile_ = 0, used_lds_ = 0, used_lanes_ = 1;
  void tune_begin(uint64_t nC, uint64_t nL);
  void tune_store();
  int32_t tile_choice() const;
  int32_t lds_choice() const;
  int32_t lanes_choice() const;
  bool counts_known_ = false;  // the pool is the kept input as counted (upload or rewind since)
  uint32_t in_unassigned_ = 0;
  Buf in_trail_buf_;
  // variable map of the last map() call: old -> new variables, the old
  // root assignments (for map_lits) and the packed-output staging buffer
  Buf map_, map_assign_, out_;
  void save_input();  // device copy of the resident input for rewind()
  // with profiling on, per-kernel event times of one call replace kernel_times
  struct ProfScope {
    Simplifier* s;
    explicit ProfScope(Simplifier* x);
    ~ProfScope();
  };
  bool map_called_ = false, map_applied_ = false, remapped_ = false;
  uint32_t map_old_nvars_ = 0, map_first_ = 0, map_first_true_ = 0;
  uint32_t orgmax_ = 0;  // largest original variable id handed out

  // host mirrors
  std::vector<uint32_t> h_assign_;  // per variable
  std::vector<uint32_t> h_touched_;  // variables assigned in h_assign_ since it was last set in full
  void set_host_assign(u32 lit) {
    h_assign_[ABS(lit)] = lit + 1;
    h_touched_.push_back(ABS(lit));
  }
  std::vector<uint32_t> units_out_;  // new root units in trail order
  std::vector<uint32_t> in_units_;
  uint32_t unassigned_ = 0;
  int phase_ = 0;
  bool refuted_ = false;           // an empty clause was appended
  bool resident_ = false;          // the clauses come from an earlier simplify call
  bool has_learnts_ = false;       // learnt clauses were uploaded (resolvents of them are learnt too)
  std::vector<u32> pending_units_;  // units of the clauses being appended
  bool prop_changed_ = true;  // BCP touched clauses since the occurrence table was built
  uint32_t mu_inc_ = 0;
  std::vector<RoundStats> rounds_;
  // profiling
  Stage prof_stage_ = S_NSTAGES;
  Event* prof_ev_ = nullptr;
};

}  // namespace pfs
