// Round driver of the HIP simplifier.  Restates the control flow of the
// reference CPU simplifier (src/cpu/simplify.cpp:157-236) over the device
// stages; see pfs_items.hpp for the per-item semantics.
#include "pfs_driver.hpp"

#include <algorithm>
#include <cstring>
#include <iterator>
#include <map>
#include <mutex>
#include <new>
#include <stdexcept>
#include <unordered_map>

namespace pfs {

const char* const kStageNames[S_NSTAGES] = {"gc", "ot", "bcp", "lcve", "sort",
                                            "sub", "ve", "bce", "ere", "count"};
static constexpr u32 INF = 0xFFFFFFFFu;

// ---------------------------------------------------------------- memory
void DeviceMemory::ensure(Buf& b, size_t bytes, size_t keep, Stream* st, Region r) {
  if (bytes <= b.cap) return;
  if (client.used - b.cap + bytes > budget_)
    throw AllocFail{"device memory budget exceeded (" + std::to_string(client.used - b.cap + bytes) + " > " +
                    std::to_string(budget_) + " bytes)"};
  if (b.p && arena_->grow(&client, b.p, bytes)) {  // the space after the block was free
    b.cap = bytes;
    peak_ = std::max(peak_, client.used);
    return;
  }
  void* old = b.p;
  const size_t old_cap = b.cap;
  try {
    arena_->alloc(&client, bytes, r, &b.p);  // sets b.p
  } catch (std::bad_alloc&) {
    b.p = old;
    throw AllocFail{"device arena allocation of " + std::to_string(bytes) + " bytes failed"};
  }
  if (old) {
    if (keep) d2d(b.p, old, std::min(keep, old_cap), st);
    arena_->release(&client, old);  // returns to the pool once the copy has run
  }
  b.cap = bytes;
  peak_ = std::max(peak_, client.used);
}
void DeviceMemory::release(Buf& b, bool idle) {
  if (b.p) arena_->release(&client, b.p, idle);
  b.p = nullptr;
  b.cap = 0;
}
void DeviceMemory::take(Buf& dst, Buf& src) {
  release(dst);
  dst = src;
  src = Buf{};
  arena_->retarget(dst.p, &dst.p);
}
void DeviceMemory::swap(Buf& a, Buf& b) {
  std::swap(a, b);
  arena_->retarget(a.p, &a.p);
  arena_->retarget(b.p, &b.p);
}

// ---------------------------------------------------------------- context
Simplifier::Simplifier(int device, size_t budget, bool keep_input, DeviceArena* arena)
    : device_(device), mem_(budget), keep_input_(keep_input) {
  st_ = stream_create(device);
  mem_.attach(arena ? arena : &DeviceArena::get(device), st_);
  fixed_scratch_ = scratch_lookup(device);  // once per device; stream_create made the device current
}

Simplifier::~Simplifier() {
  free_all();
  stream_destroy(st_);
}

void Simplifier::free_all() {
  Buf* all[] = {&c_off_[0], &lits_[0], &c_off_[1], &lits_[1],
                &ot_off_, &ot_size_, &ot_refs_, &hist_, &otp_, &assign_, &vstate_, &assumed_, &vorg_, &trail_,
                &score_, &cflag_, &flag_, &cand_, &key_, &pos_, &state_, &pv_, &vsz_, &vbase_, &type_,
                &def_, &vncls_, &vnlits_, &vnstack_, &clsb_, &litb_, &stkb_, &snap_ref_, &snap_flg_,
                &snap_size_, &snap_sig_, &snap_lbd_, &snap_np_, &snap_nn_, &scratch_, &elits_, &unit_by_,
                &recs_, &ids_, &mark_, &ht_, &erecs_, &order_, &stack_, &tmp_, &tmp2_, &tmp3_, &scal_, &in_c_off_,
                &in_lits_, &in_assign_, &in_vstate_, &in_trail_buf_, &proof_, &precs_,
                &pstamp_, &plines_, &pbytes_, &subu_, &subaff_, &sublast_, &ust_, &ulit_,
                &map_, &map_assign_, &out_, &vrest_, &lsum_};
  // a destructor must not throw, also after a device error.  Once the
  // stream has finished (or failed) its work, the blocks are free at once,
  // and the releases still pending on it are returned too: no stream
  // marker (a HIP event) outlives the stream it was recorded on, which is
  // destroyed next
  stream_quiesce(st_);
  for (Buf* b : all) mem_.release(*b, true);
  if (mem_.arena()) mem_.arena()->drain(st_);
}

DB Simplifier::make_db(int w) {
  DB db;
  db.lits = P<u32>(lits_[w]);
  db.c_off = P<u32>(c_off_[w]);
  db.ot_off = P<u32>(ot_off_);
  db.ot_size = P<u32>(ot_size_);
  db.ot_refs = P<u32>(ot_refs_);
  db.assign = P<u32>(assign_);
  db.vstate = P<u8>(vstate_);
  db.assumed = P<u8>(assumed_);
  db.vorg = P<u32>(vorg_);
  db.nvars = V_;
  db.ncls = (u32)ncls_;
  return db;
}

uint32_t Simplifier::read_u32(const void* dptr) {
  u32 v;
  d2h(&v, dptr, 4, st_);
  stream_sync(st_);
  return v;
}

void Simplifier::stage_begin(Stage s) {
  stage_calls[s]++;
  if (!rounds_.empty()) rounds_.back().stages |= 1u << s;
  if (!profile) return;
  prof_stage_ = s;
  prof_ev_ = event_record(st_);
}
void Simplifier::stage_end() {
  if (opt.sync_always) stream_sync(st_);  // syncalways: surface device errors at the stage that caused them
  if (!profile || !prof_ev_) return;
  Event* e = event_record(st_);
  stage_ms[prof_stage_] += event_ms(prof_ev_, e);
  event_free(prof_ev_);
  event_free(e);
  prof_ev_ = nullptr;
}

void Simplifier::alloc_state() {
  const size_t V1 = (size_t)V_ + 1, L2 = 2 * (size_t)V_ + 3;
  for (int w = 0; w < 2; w++) {
    mem_.ensure(c_off_[w], ccap_ * 4, 0, st_);
    mem_.ensure(lits_[w], lcap_ * 4, 0, st_);
  }
  mem_.ensure(ot_off_, (L2 + 1) * 4, 0, st_);
  mem_.ensure(ot_size_, L2 * 4, 0, st_);
  mem_.ensure(ot_refs_, lcap_ * 4, 0, st_);
  mem_.ensure(hist_, L2 * 4, 0, st_);
  // per-variable arrays: the arena's stable region
  const Region S = R_STABLE;
  mem_.ensure(assign_, V1 * 4, 0, st_, S);
  mem_.ensure(vstate_, V1, 0, st_, S);
  mem_.ensure(assumed_, V1, 0, st_, S);
  mem_.ensure(vorg_, V1 * 4, 0, st_, S);
  mem_.ensure(trail_, (V1 + 1) * 4, 0, st_, S);
  for (Buf* b : {&score_, &flag_, &cand_, &key_, &pos_, &state_, &pv_, &vsz_, &vbase_, &type_, &def_,
                 &vncls_, &vnlits_, &vnstack_, &clsb_, &litb_, &stkb_, &snap_np_, &snap_nn_, &order_, &sublast_,
                 &ust_, &ulit_})
    mem_.ensure(*b, (V1 + 1) * 4, 0, st_, S);
  mem_.ensure(cflag_, V1, 0, st_, S);
  mem_.ensure(elits_, 2 * (V1 + 1) * 4, 0, st_, S);
  mem_.ensure(unit_by_, L2 * 4, 0, st_);
  mem_.ensure(recs_, 3 * (V1 + 1) * 4, 0, st_, S);
  mem_.ensure(subaff_, 3 * (V1 + 1) * 4, 0, st_, S);
  mem_.ensure(ids_, (V1 + 1) * 4, 0, st_, S);
  mem_.ensure(vrest_, (2 * (V1 + 1) + 32) * 4, 0, st_, S);
  mem_.ensure(lsum_, 2 * (V1 + 1) * 4, 0, st_, S);
  mem_.ensure(mark_, lcap_, 0, st_);  // per clause handle
  const size_t big = std::max<size_t>({V1 + 2, L2 + 1, ccap_});
  mem_.ensure(tmp_, std::max(scan_tmp_bytes(big), radix_tmp_bytes(V1 + 1)), 0, st_);
  mem_.ensure(scal_, 256, 0, st_);
  mem_.ensure(stack_, (size_t)1 << 20, 0, st_);
}

// ---------------------------------------------------------------- upload
void Simplifier::upload(uint32_t maxvar, uint64_t ncls, const uint64_t* offsets, const uint32_t* lits,
                        const uint8_t* flags, const uint32_t* lbd, const uint8_t* vstate,
                        const uint8_t* assumed, const uint32_t* vorg, const uint32_t* trail,
                        uint32_t ntrail) {
  V_ = maxvar;
  // extraction order: originals, then learnts (simplify.cpp:148-149)
  std::vector<uint64_t> order;
  order.reserve(ncls);
  for (uint64_t i = 0; i < ncls; i++)
    if (!(flags && (flags[i] & 1))) order.push_back(i);
  for (uint64_t i = 0; i < ncls; i++)
    if (flags && (flags[i] & 1)) order.push_back(i);
  has_learnts_ = order.size() && flags && (flags[order.back()] & 1);
  // pool words: a header per clause, then its literals
  const uint64_t L = offsets[ncls] - offsets[0] + CL_HDR * ncls;
  if (ncls >= INF / 2 || L >= INF / 4) throw AllocFail{"formula too large for 32-bit clause ids"};
  std::vector<u32> h_off(ncls), h_lits(L);
  uint64_t w = 0;
  for (uint64_t k = 0; k < ncls; k++) {
    const uint64_t i = order[k];
    const uint64_t n = offsets[i + 1] - offsets[i];
    if (n > CL_MAX_SIZE) throw std::invalid_argument("clause longer than 2^24-1 literals");
    h_off[k] = (u32)w;
    const u8 f = flags ? flags[i] : 0;
    const bool learnt = f & 1;
    h_lits[w] = make_hdr((u32)n, learnt ? (u8)(ST_LEARNT | (f & USAGE_MASK)) : (u8)ST_ORIGINAL);
    h_lits[w + 1] = 0;  // signature: set by the literal sort on the device
    h_lits[w + 2] = learnt && lbd ? (lbd[i] & 0x03FFFFFFu) : 0u;
    for (uint64_t j = 0; j < n; j++) {
      const u32 l = lits[offsets[i] + j];
      if (ABS(l) == 0 || ABS(l) > maxvar) throw std::invalid_argument("literal out of range");
      h_lits[w + CL_HDR + j] = l;
    }
    w += CL_HDR + n;
  }
  set_capacities(ncls, L);
  alloc_state();
  cur_ = 0;
  ncls_ = ncls;
  lits_end_ = L;
  h2d(c_off_[0].p, h_off.data(), ncls * 4, st_);
  h2d(lits_[0].p, h_lits.data(), L * 4, st_);
  // per-variable state
  h_assign_.assign((size_t)V_ + 1, 0);
  h_touched_.clear();
  in_units_.assign(trail, trail + ntrail);
  for (u32 i = 0; i < ntrail; i++) h_assign_[ABS(trail[i])] = trail[i] + 1;
  std::vector<u8> vs((size_t)V_ + 1, 0), as((size_t)V_ + 1, 0);
  std::vector<u32> vo((size_t)V_ + 1);
  u32 melted = 0;
  orgmax_ = V_;
  for (u32 v = 0; v <= V_; v++) {
    vs[v] = v && vstate ? vstate[v] : 0;
    as[v] = v && assumed ? assumed[v] : 0;
    vo[v] = vorg ? vorg[v] : v;
    orgmax_ = std::max(orgmax_, vo[v]);
    melted += (vs[v] & (VS_MELTED | VS_SUBSTITUTED)) ? 1 : 0;
  }
  map_called_ = map_applied_ = remapped_ = false;
  for (u32 i = 0; i < ntrail; i++) vs[ABS(trail[i])] = VS_FROZEN;
  h2d(assign_.p, h_assign_.data(), ((size_t)V_ + 1) * 4, st_);
  h2d(vstate_.p, vs.data(), (size_t)V_ + 1, st_);
  h2d(assumed_.p, as.data(), (size_t)V_ + 1, st_);
  h2d(vorg_.p, vo.data(), ((size_t)V_ + 1) * 4, st_);
  h2d(trail_.p, trail, (size_t)ntrail * 4, st_);
  trail_n_ = propagated_ = ntrail;
  unassigned_ = V_ - ntrail - melted;
  stack_n_ = 0;
  units_out_.clear();
  rounds_.clear();
  proof_n_ = 0;
  refuted_ = false;
  resident_ = false;
  fill_u8(P<u8>(mark_), 0, lcap_, st_);
  // extract: sort literals, signatures
  k_prep(make_db(0), 0, st_);
  save_input();
  stream_sync(st_);
}

void Simplifier::save_input() {
  if (!keep_input_) return;
  const uint64_t ncls = ncls_, L = lits_end_, ntrail = trail_n_;
  in_ncls_ = ncls;
  in_nlits_ = L;
  in_trail_ = ntrail;
  in_unassigned_ = unassigned_;
  mem_.ensure(in_c_off_, std::max<uint64_t>(ncls, 1) * 4, 0, st_);
  mem_.ensure(in_lits_, std::max<uint64_t>(L, 1) * 4, 0, st_);
  mem_.ensure(in_assign_, ((size_t)V_ + 1) * 4, 0, st_);
  mem_.ensure(in_vstate_, (size_t)V_ + 1, 0, st_);
  mem_.ensure(in_trail_buf_, ((size_t)ntrail + 1) * 4, 0, st_);
  d2d(in_c_off_.p, c_off_[cur_].p, ncls * 4, st_);
  d2d(in_lits_.p, lits_[cur_].p, L * 4, st_);
  d2d(in_assign_.p, assign_.p, ((size_t)V_ + 1) * 4, st_);
  d2d(in_vstate_.p, vstate_.p, (size_t)V_ + 1, st_);
  d2d(in_trail_buf_.p, trail_.p, (size_t)ntrail * 4, st_);
  // the live counts of the kept input: a simplify right after this or
  // after a rewind starts from them instead of counting the pool again
  u64* d = (u64*)P<u8>(scal_);
  k_count(make_db(cur_), d, st_);
  u64 h[2];
  d2h(h, d, 16, st_);
  stream_sync(st_);
  in_live_cls_ = h[0];
  in_live_lits_ = h[1];
  counts_known_ = true;
}

Simplifier::ProfScope::ProfScope(Simplifier* x) : s(x) {
  if (s->profile) s->kernel_times.clear(), prof_enable(s->st_, true);
}
Simplifier::ProfScope::~ProfScope() {
  if (!s->profile) return;
  try {
    prof_collect(s->st_, s->kernel_times);
    prof_enable(s->st_, false);
  } catch (...) {
  }
}

// ---------------------------------------------------------------- DIMACS ingestion
namespace {
bool dimacs_blank(unsigned char c) { return (c >= 9 && c <= 13) || c == 32; }  // dimacs.h:33
bool dimacs_digit(unsigned char c) { return c >= '0' && c <= '9'; }
// releases temporaries when the scope ends, also on an exception
struct BufScope {
  DeviceMemory& m;
  std::vector<DeviceMemory::Buf*> bufs;
  ~BufScope() {
    for (DeviceMemory::Buf* b : bufs) m.release(*b);
  }
};
}  // namespace

// The header on the host (dimacs.cpp:81-108): blanks, comment lines, then
// "p cnf V C"; the clause part after it on the device.  Parsing drops
// duplicate literals, tautologies and unit clauses (the units are assigned
// in text order) and reports empty or conflicting clauses (toClause,
// dimacs.cpp:113-170); then the root units are propagated and the clauses
// shrunk at the top level like Solver() + rootify/shrinkTop before the
// first simplify (solver.cpp:49, simplify.cpp:163-164,
// shrink.cpp:152-216).  The result equals the reference's simplifier input
// up to the order of literals in a clause (the simplifier sorts them) and
// the order of the trail.
Simplifier::ParseInfo Simplifier::parse_dimacs(const char* text, uint64_t n) {
  ParseInfo info;
  const unsigned char* t = (const unsigned char*)text;
  uint64_t i = 0, V = 0, C = 0;
  bool header = false;
  while (i < n) {
    while (i < n && dimacs_blank(t[i])) i++;
    if (i >= n || t[i] == 0 || t[i] == '%') break;
    if (t[i] == 'c') {  // eatLine
      while (i < n && t[i] != 0 && t[i++] != '\n') {
      }
      continue;
    }
    if (t[i] != 'p') throw std::invalid_argument("clauses before the 'p cnf' header");
    if (n - i < 5 || memcmp(t + i, "p cnf", 5)) throw std::invalid_argument("header has wrong format");
    i += 5;
    auto number = [&](const char* what) -> uint64_t {  // toInteger, dimacs.h:101-111
      while (i < n && dimacs_blank(t[i])) i++;
      bool neg = false;
      if (i < n && (t[i] == '-' || t[i] == '+')) neg = t[i++] == '-';
      if (i >= n || !dimacs_digit(t[i])) throw std::invalid_argument("expected a digit in the header");
      uint64_t v = 0;
      while (i < n && dimacs_digit(t[i])) v = std::min<uint64_t>(v * 10 + (t[i++] - '0'), 1ull << 40);
      if (neg) throw std::invalid_argument(std::string("number of ") + what + " in header is negative");
      return v;
    };
    V = number("variables");
    if (V == 0) throw std::invalid_argument("zero number of variables in header");
    if (V >= 0x7FFFFFFEull) throw std::invalid_argument("number of variables not supported");
    C = number("clauses");
    if (C == 0) throw std::invalid_argument("zero number of clauses in header");
    header = true;
    break;
  }
  if (!header) throw std::invalid_argument("no 'p cnf' header");
  if (n >= (1ull << 40)) throw std::invalid_argument("text too large");
  info.maxvar = (u32)V;
  info.header_clauses = C;
  ProfScope prof(this);

  // ---- lexing on the device
  const uint64_t begin = i;
  const uint64_t ntiles = lex_tiles(n);
  const size_t padded = (size_t)ntiles * 4096 + 16;
  Buf text_d, cnt_d, pmark_d, pos_d, mask_d, lits_d, ends_d, zpos_d, ulit_d, first_d, win_d, upos_d;
  BufScope scope{mem_, {&text_d, &cnt_d, &pmark_d, &pos_d, &mask_d, &lits_d, &ends_d, &zpos_d, &ulit_d, &first_d,
                        &win_d, &upos_d}};
  mem_.ensure(text_d, padded, 0, st_);
  mem_.ensure(cnt_d, ((size_t)ntiles + 1) * 4, 0, st_);
  mem_.ensure(pmark_d, 64, 0, st_);
  fill_u8(P<u8>(text_d) + n, ' ', padded - n, st_);
  h2d(text_d.p, text, n, st_);
  unsigned long long* mark = (unsigned long long*)pmark_d.p;  // 3 words
  u32* err = (u32*)(mark + 3);
  u32* word = err + 1;  // scratch words: totals
  fill_u32((u32*)mark, 0xFFFFFFFFu, 6, st_);
  fill_u32(err, 0, 5, st_);
  mem_.ensure(tmp_, std::max(tmp_.cap, scan_tmp_bytes((size_t)ntiles + 1)), 0, st_);
  LexArgs a{P<u8>(text_d), begin, n, nullptr, P<u32>(cnt_d), nullptr, nullptr, mark, err, (u32)V};
  k_lex(LEX_SPECIAL, a, ntiles, st_);
  scan_u32(P<u32>(cnt_d), P<u32>(cnt_d), ntiles, word, tmp_.p, st_);
  struct {
    unsigned long long mark[3];
    u32 err, word[4];
  } h{};
  d2h(&h, pmark_d.p, sizeof(h), st_);
  stream_sync(st_);
  uint64_t end = std::min<uint64_t>({n, h.mark[0], h.mark[1]});
  const u32 ncom = h.word[0];
  if (ncom) {
    // comment lines: their bytes masked, then the first '%' token outside them
    mem_.ensure(pos_d, (size_t)ncom * 8, 0, st_);
    mem_.ensure(mask_d, padded, 0, st_);
    fill_u8(P<u8>(mask_d), 0, padded, st_);
    a.pos = P<u64>(pos_d);
    k_lex(LEX_COMMENTS, a, ntiles, st_);
    k_lex_mask(P<u8>(text_d), n, P<u64>(pos_d), ncom, P<u8>(mask_d), st_);
    a.mask = P<u8>(mask_d);
    fill_u32((u32*)(mark + 1), 0xFFFFFFFFu, 2, st_);
    k_lex(LEX_PERCENT, a, ntiles, st_);
    d2h(&h, pmark_d.p, sizeof(h), st_);
    stream_sync(st_);
    end = std::min<uint64_t>({n, h.mark[0], h.mark[1]});
  }
  a.end = end;
  k_lex(LEX_COUNT, a, ntiles, st_);
  scan_u32(P<u32>(cnt_d), P<u32>(cnt_d), ntiles, word, tmp_.p, st_);
  u32 T = 0;
  d2h(&T, word, 4, st_);
  stream_sync(st_);
  mem_.ensure(lits_d, ((size_t)T + 1) * 4, 0, st_);
  a.lits = P<u32>(lits_d);
  k_lex(LEX_EMIT, a, ntiles, st_);
  u32 last = 0;
  if (T) d2h(&last, P<u32>(lits_d) + T - 1, 4, st_);
  d2h(&h, pmark_d.p, sizeof(h), st_);
  stream_sync(st_);
  if (h.err & LEXE_TOKEN)
    throw std::invalid_argument("malformed token at byte " + std::to_string(h.mark[2]));
  if (h.err & LEXE_VAR) throw std::invalid_argument("too many variables (a literal above the header's count)");
  if (T && last != 0) throw std::invalid_argument("the last clause is not terminated by 0");

  // ---- clauses into the pool
  mem_.ensure(ends_d, ((size_t)T + 1) * 4, 0, st_);
  mem_.ensure(zpos_d, ((size_t)T + 1) * 4, 0, st_);
  mem_.ensure(tmp_, std::max(tmp_.cap, scan_tmp_bytes((size_t)T + 1)), 0, st_);
  k_clause_ends(P<u32>(lits_d), T, P<u32>(zpos_d), tmp_.p, word, P<u32>(ends_d), st_);
  u32 NC = 0;
  d2h(&NC, word, 4, st_);
  stream_sync(st_);
  mem_.release(zpos_d);
  const uint64_t L = (uint64_t)T - NC + (uint64_t)CL_HDR * NC;
  if (NC >= INF / 2 || L >= INF / 4) throw AllocFail{"formula too large for 32-bit clause ids"};
  info.read_clauses = NC;
  V_ = (u32)V;
  set_capacities(NC, L);
  alloc_state();
  cur_ = 0;
  ncls_ = NC;
  lits_end_ = L;
  DB db = make_db(0);
  k_parse_build(db, P<u32>(lits_d), P<u32>(ends_d), err, st_);
  const size_t V1 = (size_t)V_ + 1;
  fill_u32(P<u32>(assign_), 0, V1, st_);
  fill_u8(P<u8>(vstate_), 0, V1, st_);
  fill_u8(P<u8>(assumed_), 0, V1, st_);
  iota_u32(P<u32>(vorg_), V1, st_);
  fill_u8(P<u8>(mark_), 0, lcap_, st_);
  k_prep(db, 0, st_);
  const bool proof = opt.proof_en != 0;  // the text clauses stay for the proof lines
  if (!proof) {
    mem_.release(lits_d);
    mem_.release(ends_d);
  }
  mem_.ensure(ulit_d, ((size_t)NC + 1) * 4, 0, st_);
  mem_.ensure(first_d, V1 * 4, 0, st_);
  mem_.ensure(win_d, ((size_t)NC + 1) * 4, 0, st_);
  mem_.ensure(upos_d, ((size_t)NC + 1) * 4, 0, st_);
  mem_.ensure(tmp_, std::max(tmp_.cap, scan_tmp_bytes((size_t)NC + 1)), 0, st_);
  fill_u32(word, 0, 2, st_);
  k_parse_dedup(db, P<u32>(ulit_d), err, st_);
  k_parse_units(db, P<u32>(ulit_d), NC, P<u32>(first_d), P<u32>(win_d), P<u32>(upos_d), tmp_.p, word,
                P<u32>(trail_), err, st_);
  k_count(db, (u64*)(P<u32>(scal_) + 40), st_);  // clauses that stay (duplicates, tautologies, units gone)
  uint64_t kept[2];
  d2h(kept, P<u32>(scal_) + 40, sizeof(kept), st_);
  d2h(&h, pmark_d.p, sizeof(h), st_);
  stream_sync(st_);
  if (h.err & LEXE_SIZE) throw std::invalid_argument("clause longer than 2^24-1 literals");
  if (kept[0] > C) throw std::invalid_argument("too many clauses");
  const u32 ntrail = h.word[0];

  // ---- the state of a fresh upload, then root propagation and the shrink
  h_assign_.assign(V1, 0);
  h_touched_.clear();
  std::vector<u32> tr(ntrail);
  if (ntrail) d2h(tr.data(), trail_.p, (size_t)ntrail * 4, st_);
  stream_sync(st_);
  for (u32 u : tr) h_assign_[ABS(u)] = u + 1;
  trail_n_ = ntrail;
  propagated_ = 0;
  unassigned_ = (u32)V - ntrail;
  stack_n_ = 0;
  units_out_.clear();
  rounds_.clear();
  pending_units_.clear();
  refuted_ = (h.err & LEXE_UNSAT) != 0;
  resident_ = false;
  has_learnts_ = false;
  map_called_ = map_applied_ = remapped_ = false;
  orgmax_ = (u32)V;
  proof_n_ = 0;
  pl_.clear();
  pst_.clear();
  if (!refuted_ && ntrail) {
    opt.proof_en = 0;  // the propagation's lines are written from the end state below
    create_ot();
    const bool ok = prop();
    opt.proof_en = proof;
    if (!ok) refuted_ = true;
  }
  if (proof) parse_proof(P<u32>(lits_d), P<u32>(ends_d), refuted_ ? 0 : (u32)NC);
  if (!refuted_) gc();
  units_out_.clear();
  rounds_.clear();
  in_units_.resize(trail_n_);
  if (trail_n_) d2h(in_units_.data(), trail_.p, trail_n_ * 4, st_);
  stream_sync(st_);
  if (refuted_) ncls_ = 0, lits_end_ = 0;
  save_input();
  stream_sync(st_);
  info.status = refuted_ ? 0 : 2;
  info.clauses = ncls_;
  info.literals = out_nlits();
  info.trail = trail_n_;
  return info;
}

void Simplifier::download_trail(uint32_t* lits) {
  if (trail_n_) d2h(lits, trail_.p, trail_n_ * 4, st_);
  stream_sync(st_);
}

void Simplifier::rewind() {
  if (!keep_input_) throw std::logic_error("context was created without keep_input");
  if (remapped_) throw std::logic_error("the variables were renumbered since the upload");
  cur_ = 0;
  ncls_ = in_ncls_;
  lits_end_ = in_nlits_;
  ensure_clause_room(0, 0);
  d2d(c_off_[0].p, in_c_off_.p, in_ncls_ * 4, st_);
  d2d(lits_[0].p, in_lits_.p, in_nlits_ * 4, st_);
  d2d(assign_.p, in_assign_.p, ((size_t)V_ + 1) * 4, st_);
  d2d(vstate_.p, in_vstate_.p, (size_t)V_ + 1, st_);
  d2d(trail_.p, in_trail_buf_.p, in_trail_ * 4, st_);
  trail_n_ = propagated_ = in_trail_;
  unassigned_ = in_unassigned_;
  stack_n_ = 0;
  units_out_.clear();
  refuted_ = false;
  resident_ = false;
  // the host assignments back to the input's: only the entries set since
  // (clearing all 10M entries of the heavy-tailed formula showed as a
  // 0.36 ms gap before each call's first kernel)
  if (h_touched_.size() < h_assign_.size() / 8) {
    for (u32 v : h_touched_) h_assign_[v] = 0;
  } else {
    std::fill(h_assign_.begin(), h_assign_.end(), 0u);
  }
  h_touched_.clear();
  for (u32 u : in_units_) h_assign_[ABS(u)] = u + 1;
  counts_known_ = true;
}

// ---------------------------------------------------------------- incremental use
// Adds original clauses over variables 1..maxvar (maxvar >= the current
// count) to the resident formula, as the reference's incremental front end
// does (iadd / itoClause, src/cpu/addinc.cpp:44-160): literals false at the
// root are dropped, root-satisfied clauses skipped, units assigned (they are
// propagated by the next simplify), an empty clause refutes.
void Simplifier::append(uint32_t maxvar, uint64_t ncls, const uint64_t* offsets, const uint32_t* lits) {
  counts_known_ = false;
  if (maxvar < V_) throw std::invalid_argument("append cannot reduce the number of variables");
  if (maxvar > V_) grow_vars(maxvar);
  std::vector<u32> words, off;
  std::vector<u32> cl;
  uint64_t w = lits_end_;
  for (uint64_t i = 0; i < ncls; i++) {
    cl.clear();
    bool sat = false;
    for (uint64_t j = offsets[i]; j < offsets[i + 1]; j++) {
      const u32 l = lits[j];
      if (ABS(l) == 0 || ABS(l) > maxvar) throw std::invalid_argument("literal out of range");
      const u32 a = h_assign_[ABS(l)];
      if (a) {
        if (a - 1 == l) sat = true;
        continue;
      }
      bool dup = false;
      for (u32 x : cl) {
        if (x == l) dup = true;
        if (x == FLIP(l)) sat = true;  // tautology
      }
      if (!dup) cl.push_back(l);
    }
    if (sat) continue;
    if (cl.empty()) {
      refuted_ = true;
      continue;
    }
    if (cl.size() == 1) {
      const u32 u = cl[0];
      if (h_assign_[ABS(u)]) continue;  // assigned by an earlier clause of this batch
      set_host_assign(u);
      pending_units_.push_back(u);
      continue;
    }
    if (cl.size() > CL_MAX_SIZE) throw std::invalid_argument("clause longer than 2^24-1 literals");
    off.push_back((u32)w);
    words.push_back(make_hdr((u32)cl.size(), ST_ORIGINAL));
    words.push_back(0);  // signature: set by the literal sort
    words.push_back(0);  // lbd
    words.insert(words.end(), cl.begin(), cl.end());
    w += CL_HDR + cl.size();
  }
  const uint64_t nadd = off.size();
  if (nadd) {
    ensure_clause_room(nadd, words.size());
    DB db0 = make_db(cur_);
    h2d(db0.c_off + ncls_, off.data(), nadd * 4, st_);
    h2d(db0.lits + lits_end_, words.data(), words.size() * 4, st_);
    const u32 first = (u32)ncls_;
    ncls_ += nadd;
    lits_end_ += words.size();
    k_prep(make_db(cur_), first, st_);
  }
  if (!pending_units_.empty()) {
    // new root units: assigned and frozen like enqueueUnit, appended to the
    // trail unpropagated
    mem_.ensure(ids_, pending_units_.size() * 4, 0, st_);
    h2d(ids_.p, pending_units_.data(), pending_units_.size() * 4, st_);
    k_assign_units(make_db(cur_), P<u32>(ids_), (u32)pending_units_.size(), st_);
    h2d(P<u32>(trail_) + trail_n_, pending_units_.data(), pending_units_.size() * 4, st_);
    trail_n_ += pending_units_.size();
    for (u32 u : pending_units_) units_out_.push_back(u);
    unassigned_ -= (u32)pending_units_.size();
    pending_units_.clear();
  }
  stream_sync(st_);
}

// Replaces the set of assumed variables (never eliminated; the reference's
// incremental assumptions, src/cpu/assume.cpp, and lcve.cpp's 'iassumed').
void Simplifier::assume(const uint32_t* vars, uint32_t n) {
  counts_known_ = false;
  std::vector<u8> as((size_t)V_ + 1, 0);
  for (uint32_t i = 0; i < n; i++) {
    if (vars[i] == 0 || vars[i] > V_) throw std::invalid_argument("assumed variable out of range");
    as[vars[i]] = 1;
  }
  h2d(assumed_.p, as.data(), as.size(), st_);
  stream_sync(st_);
}

// grows the per-variable state to maxvar variables, keeping its content
void Simplifier::grow_vars(uint32_t maxvar) {
  const size_t o1 = (size_t)V_ + 1, n1 = (size_t)maxvar + 1;
  mem_.ensure(assign_, n1 * 4, o1 * 4, st_, R_STABLE);
  mem_.ensure(vstate_, n1, o1, st_, R_STABLE);
  mem_.ensure(assumed_, n1, o1, st_, R_STABLE);
  mem_.ensure(vorg_, n1 * 4, o1 * 4, st_, R_STABLE);
  mem_.ensure(trail_, (n1 + 1) * 4, trail_n_ * 4, st_, R_STABLE);
  fill_u32(P<u32>(assign_) + o1, 0, n1 - o1, st_);
  fill_u8(P<u8>(vstate_) + o1, 0, n1 - o1, st_);
  fill_u8(P<u8>(assumed_) + o1, 0, n1 - o1, st_);
  // new variables keep their number as original id (iadd, src/cpu/addinc.cpp:48-63);
  // after a renumbering they get fresh original ids past every id handed out
  std::vector<u32> vo(n1 - o1);
  for (size_t v = o1; v < n1; v++) {
    vo[v - o1] = remapped_ ? ++orgmax_ : (u32)v;
    if (!remapped_) orgmax_ = std::max(orgmax_, (u32)v);
  }
  h2d(P<u32>(vorg_) + o1, vo.data(), vo.size() * 4, st_);
  if (pstamp_.p) {
    mem_.ensure(pstamp_, n1 * 4, o1 * 4, st_);
    fill_u32(P<u32>(pstamp_) + o1, 0, n1 - o1, st_);
  }
  unassigned_ += maxvar - V_;
  V_ = maxvar;
  h_assign_.resize(n1, 0);
  alloc_state();  // the per-round buffers (their content is rebuilt every round)
  stream_sync(st_);
}

// ---------------------------------------------------------------- helpers
// Clause-block capacities at upload: the input plus room for resolvents of
// lits_mul times its size (the reference's literalsmul, src/gpu/options.cpp:43,
// applied at src/gpu/simplify.cu:84-92); later growth is by half.
void Simplifier::set_capacities(uint64_t ncls, uint64_t words) {
  const double m = std::min(std::max(lits_mul, 0.0), 4.0);
  ccap_ = std::max<uint64_t>(ncls + (uint64_t)(m * (double)ncls) + 1024, 4096);
  lcap_ = std::max<uint64_t>(words + (uint64_t)(m * (double)words) + 4096, 16384);
  if (ccap_ >= INF / 2 || lcap_ >= INF / 4) {
    ccap_ = std::min<uint64_t>(ccap_, INF / 2 - 1);
    lcap_ = std::min<uint64_t>(lcap_, INF / 4 - 1);
  }
}

void Simplifier::ensure_clause_room(uint64_t add_cls, uint64_t add_lits) {
  const uint64_t need_c = ncls_ + add_cls, need_l = lits_end_ + add_lits;
  if (need_c <= ccap_ && need_l <= lcap_) return;
  const uint64_t nc = std::max(ccap_, need_c + need_c / 2), nl = std::max(lcap_, need_l + need_l / 2);
  if (nc >= INF / 2 || nl >= INF / 4) throw AllocFail{"clause database exceeds 32-bit ids"};
  for (int w = 0; w < 2; w++) {
    const bool keep = (w == cur_);
    mem_.ensure(c_off_[w], nc * 4, keep ? ncls_ * 4 : 0, st_);
    mem_.ensure(lits_[w], nl * 4, keep ? lits_end_ * 4 : 0, st_);
  }
  // the occurrence table is rebuilt every round; its refs array only needs
  // the new capacity from the next round on, but it is resized here too
  mem_.ensure(ot_refs_, nl * 4, lits_end_ * 4, st_);
  const size_t old_mark = mark_.cap;
  mem_.ensure(mark_, nl, old_mark, st_);
  if (nl > old_mark) fill_u8(P<u8>(mark_) + old_mark, 0, nl - old_mark, st_);
  mem_.ensure(tmp_, std::max(tmp_.cap, scan_tmp_bytes(nc)), 0, st_);
  ccap_ = nc;
  lcap_ = nl;
}

void Simplifier::ensure_stack(uint64_t words) {
  if (words * 4 <= stack_.cap) return;
  mem_.ensure(stack_, std::max(words * 4 + words * 2, (uint64_t)stack_.cap * 2), stack_n_ * 4, st_);
}

void Simplifier::commit_units(const std::vector<uint32_t>& lits) {
  if (lits.empty()) return;
  mem_.ensure(ids_, lits.size() * 4, 0, st_);
  h2d(ids_.p, lits.data(), lits.size() * 4, st_);
  k_assign_units(make_db(cur_), P<u32>(ids_), (u32)lits.size(), st_);
  h2d(P<u32>(trail_) + trail_n_, lits.data(), lits.size() * 4, st_);
  trail_n_ += lits.size();
  stream_sync(st_);
}

void Simplifier::count(uint64_t& cls, uint64_t& lits) {
  stage_begin(S_COUNT);
  u64* d = (u64*)P<u8>(scal_);
  k_count(make_db(cur_), d, st_);
  u64 h[2];
  d2h(h, d, 16, st_);
  stream_sync(st_);
  cls = h[0];
  lits = h[1];
  stage_end();
}

// stable compaction of the live clauses (simplify.cpp:238-249); clause ids
// keep their relative order, so the schedule of collections never changes
// the simplifier's result
void Simplifier::gc() {
  stage_begin(S_GC);
  mem_.ensure(tmp2_, std::max<size_t>(ncls_, 1) * 4, 0, st_);
  mem_.ensure(tmp3_, std::max<size_t>(ncls_, 1) * 4, 0, st_);
  DB src = make_db(cur_), dst = make_db(1 - cur_);
  u32* tot = P<u32>(scal_) + 8;
  k_compact(src, dst, P<u32>(tmp2_), P<u32>(tmp3_), tmp_.p, tot, st_);
  u32 h[2];
  d2h(h, tot, 8, st_);
  stream_sync(st_);
  ncls_ = h[0];
  lits_end_ = h[1];
  cur_ = 1 - cur_;
  stage_end();
}

// the live clauses compacted with the originals first, then the learnt
// clauses, each in clause order (newBeginning + extract,
// src/cpu/simplify.cpp:251-269, :147-149)
void Simplifier::originals_first() {
  mem_.ensure(tmp2_, std::max<size_t>(ncls_, 1) * 4, 0, st_);
  mem_.ensure(tmp3_, std::max<size_t>(ncls_, 1) * 4, 0, st_);
  DB src = make_db(cur_), dst = make_db(1 - cur_);
  u32* tot = P<u32>(scal_) + 8;
  k_compact_originals_first(src, dst, P<u32>(tmp2_), P<u32>(tmp3_), tmp_.p, tot, st_);
  u32 h[2];
  d2h(h, tot, 8, st_);
  stream_sync(st_);
  ncls_ = h[0];
  lits_end_ = h[1];
  cur_ = 1 - cur_;
  prop_changed_ = true;
}

// DRAT lines of an ingestion (pfs_parse_dimacs with "proof"): every root
// unit of the trail in trail order (each follows by propagation from the
// text and the units before it), then, per text clause in text order, the
// deletion of a clause that is gone (satisfied, tautology, a unit now on
// the trail) or the addition of what is left of a clause that lost literals
// with the deletion of its text; a refuted text ends with the empty clause.
// The lines describe the parser's, the propagation's and the top-level
// shrink's changes together (dimacs.cpp:245-305, shrink.cpp:122).
void Simplifier::parse_proof(const u32* toks, const u32* ends, u32 nc) {
  const u32 nu = (u32)trail_n_, n = nu + nc;
  DB db = make_db(cur_);
  mem_.ensure(pbytes_, ((size_t)n + 1) * 8, 0, st_);
  u32* bytes = P<u32>(pbytes_);
  u32* off = bytes + n + 1;
  mem_.ensure(tmp_, std::max(tmp_.cap, scan_tmp_bytes((size_t)n + 1)), 0, st_);
  k_parse_proof(db, toks, ends, P<u32>(trail_), nu, nc, nullptr, nullptr, bytes, st_);
  u32* total = P<u32>(scal_) + 50;
  scan_u32(bytes, off, n, total, tmp_.p, st_);
  const u32 add = n ? read_u32(total) : 0;
  const size_t all = (size_t)add + (refuted_ ? 2 : 0);
  mem_.ensure(proof_, std::max<size_t>(all, 4096), 0, st_);
  k_parse_proof(db, toks, ends, P<u32>(trail_), nu, nc, off, P<u8>(proof_), bytes, st_);
  if (refuted_) {
    const u8 empty[2] = {'a', 0};
    h2d(P<u8>(proof_) + add, empty, 2, st_);
  }
  stream_sync(st_);
  proof_n_ = all;
}

// ---------------------------------------------------------------- OT (simplify.cpp:35-56)
// ---------------------------------------------------------------- launch-shape tuning
// One table per process: formula class -> tuned shapes.  The class is the
// device and the binary logarithms of the clause and variable counts, with
// the mean clause size in eighths.
static std::mutex g_tune_m;
static std::map<uint64_t, Tuned> g_tuned;
static uint64_t ilog2(uint64_t x) { return x ? 63 - __builtin_clzll(x) : 0; }

void Simplifier::tune_begin(uint64_t nC, uint64_t nL) {
  tune_key_ = ((uint64_t)device_ << 48) | (ilog2(nC) << 32) | (ilog2(V_) << 16) |
              std::min<uint64_t>(nC ? 8 * nL / nC : 0, 0xFFFF);
  {
    std::lock_guard<std::mutex> lk(g_tune_m);
    auto it = g_tuned.find(tune_key_);
    tune_ = it != g_tuned.end() ? it->second : Tuned{};
  }
  const bool big = opt.autotune && nC >= TUNE_MIN_CLAUSES;
  tune_ot_ = big && opt.otp_tile == 0 && !tune_.ot;
  tune_mis_ = big && (opt.mis_lds < 0 || opt.mis_lanes == 0) && !tune_.mis;
  tuned_now_ = tune_ot_ || tune_mis_;
  used_tile_ = 0;
  used_lds_ = lds_choice();
  used_lanes_ = lanes_choice();
}
void Simplifier::tune_store() {
  std::lock_guard<std::mutex> lk(g_tune_m);
  Tuned& t = g_tuned[tune_key_];
  if (tune_.ot) t.ot = true, t.otp_tile = tune_.otp_tile;
  if (tune_.mis) t.mis = true, t.mis_lds = tune_.mis_lds, t.mis_lanes = tune_.mis_lanes;
}
int32_t Simplifier::tile_choice() const { return opt.otp_tile > 0 ? opt.otp_tile : tune_.ot ? tune_.otp_tile : 0; }
int32_t Simplifier::lds_choice() const {
  return opt.mis_lds >= 0 ? opt.mis_lds : tune_.mis ? tune_.mis_lds : 10240;
}
int32_t Simplifier::lanes_choice() const { return opt.mis_lanes > 0 ? opt.mis_lanes : tune_.mis ? tune_.mis_lanes : 1; }

void Simplifier::create_ot() {
  stage_begin(S_OT);
  prop_changed_ = false;
  ot_counts(true);
  stage_end();
}

// Room for the range build: the range-by-tile counts and one 8-byte
// (literal, handle) pair per literal slot of the pool (an upper bound of the
// occurrences).  Grown by half when short, so that a formula growing request
// by request (serving) does not reallocate it every round.  When the budget
// cannot hold it, the build falls back to per-occurrence atomics, which need
// no extra memory.
bool Simplifier::otp_room(const OtpPlan& p) {
  const size_t slots = lits_end_ >= CL_HDR * ncls_ ? lits_end_ - CL_HDR * ncls_ : lits_end_;
  const size_t counts_words = ((p.ncounts + 1 + 63) & ~(size_t)63);
  const size_t need = counts_words + 2 * slots + 64;
  if (need * 4 <= otp_.cap) return true;
  const size_t counts_cap = (size_t)p.nr * ((std::max<size_t>(ccap_, ncls_) + p.ct - 1) / p.ct);
  const size_t want = ((counts_cap + 1 + 63) & ~(size_t)63) + 2 * (slots + slots / 2) + 64;
  try {
    mem_.ensure(otp_, want * 4, 0, st_);
    return true;
  } catch (AllocFail&) {
  }
  try {
    mem_.ensure(otp_, need * 4, 0, st_);
    return true;
  } catch (AllocFail&) {
  }
  return false;
}

void Simplifier::ot_counts(bool table) {
  const size_t n2 = 2 * (size_t)V_ + 2;
  DB db = make_db(cur_);
  u32* hist = P<u32>(hist_);
  OtpPlan p;
  auto build = [&](int tile) {
    if (!opt.otp_ranges || !otp_plan(n2, db.ncls, &p, tile, opt.otp_xcd) || !otp_room(p)) return false;
    mem_.ensure(tmp_, std::max(tmp_.cap, scan_tmp_bytes(p.ncounts)), 0, st_);
    u32* counts = P<u32>(otp_);
    u32* pairs = counts + ((p.ncounts + 1 + 63) & ~(size_t)63);  // 256-byte aligned
    k_ot_ranges(db, p, (u32)n2, counts, pairs, table, hist, tmp_.p, st_);
    used_tile_ = (int32_t)ilog2(p.ct);
    return true;
  };
  if (table && tune_ot_ && opt.otp_ranges) {
    // every candidate tile builds the whole table, twice over in turn (each
    // build is complete, the last one stands); the fastest is kept for the
    // class
    tune_ot_ = false;
    const int tiles[4] = {11, 12, 13, 14};
    float best[4] = {0, 0, 0, 0};
    bool ok[4] = {false, false, false, false}, any = false;
    for (int rep = 0; rep < 2; rep++)
      for (int j = 0; j < 4; j++) {
        Event* a = event_record(st_);
        const bool built = build(tiles[j]);
        Event* b = event_record(st_);
        const float ms = event_ms(a, b);
        event_free(a);
        event_free(b);
        if (built) best[j] = ok[j] ? std::min(best[j], ms) : ms, ok[j] = true;
      }
    for (int j = 0; j < 4; j++)
      if (ok[j] && (!any || best[j] < best[tune_.otp_tile - tiles[0]]))
        tune_.otp_tile = tiles[j], any = true;
    if (any) {
      tune_.ot = true;
      tune_store();
      return;
    }
  } else if (build(tile_choice())) {
    return;
  }
  used_tile_ = 0;
  fill_u32(hist, 0, n2, st_);
  k_hist(db, hist, st_);
  if (!table) return;
  scan_u32(hist, db.ot_off, n2, db.ot_off + n2, tmp_.p, st_);
  d2d(db.ot_size, hist, n2 * 4, st_);
  fill_u32(hist, 0, n2, st_);
  k_scatter(db, hist, st_);
}

// ---------------------------------------------------------------- BCP (elimination.cpp:64-99)
bool Simplifier::prop() {
  if (trail_n_ == propagated_) return true;
  stage_begin(S_BCP);
  prop_changed_ = true;
  if (!prop_visit(propagated_)) {
    stage_end();
    return false;
  }
  prop_clean(propagated_);
  propagated_ = trail_n_;
  stage_end();
  return true;
}

bool Simplifier::prop_visit(uint64_t front) {
  DB db = make_db(cur_);
  u32* trail = P<u32>(trail_);
  u32* sc = P<u32>(scal_);
  const uint64_t start_new = trail_n_;
  uint64_t nfront = trail_n_ - front;
  while (nfront) {
    fill_u32(sc, 0, 2, st_);
    k_bcp_visit(db, trail + front, (u32)nfront, trail + trail_n_, sc, sc + 1, st_);
    u32 h[2];
    d2h(h, sc, 8, st_);
    stream_sync(st_);
    if (h[1]) {
      proof_line(PR_EMPTY, 0);  // learnEmpty (elimination.cpp:80-86)
      proof_flush();
      return false;
    }
    front = trail_n_;
    nfront = h[0];
    trail_n_ += h[0];
  }
  const uint64_t nnew = trail_n_ - start_new;
  if (nnew) {
    std::vector<u32> fresh(nnew);
    d2h(fresh.data(), trail + start_new, nnew * 4, st_);
    k_set_vstate(db, trail + start_new, (u32)nnew, VS_FROZEN, st_);
    stream_sync(st_);
    for (u32 u : fresh) {
      set_host_assign(u);
      units_out_.push_back(u);
      proof_line(PR_UNIT, u);  // enqueueUnit -> learnUnit (solver.h:430-441, 475-479)
    }
    proof_flush();
    unassigned_ -= (u32)nnew;
    if (!rounds_.empty()) rounds_.back().units += (u32)nnew;
  }
  return true;
}

void Simplifier::prop_clean(uint64_t from) {
  DB db = make_db(cur_);
  u32* trail = P<u32>(trail_);
  const u32 nprop = (u32)(trail_n_ - from);
  k_bcp_mark(db, trail + from, nprop, P<u8>(mark_), st_);
  k_clear_lists(db, trail + from, nprop, st_);
  k_bcp_clean(db, P<u8>(mark_), st_);
  if (!opt.sub_en) k_reduce_lists(db, st_);
  stream_sync(st_);
}

// ---------------------------------------------------------------- LCVE (lcve.cpp)
bool Simplifier::lcve() {
  stage_begin(S_LCVE);
  const size_t n2 = 2 * (size_t)V_ + 2;
  DB db = make_db(cur_);
  u32* hist = P<u32>(hist_);
  // histSimp (lcve.cpp:32): the occurrence counts of the live clauses.  The
  // occurrence-table scatter leaves exactly these counts in hist_ unless the
  // round's BCP changed clauses since.
  if (prop_changed_) ot_counts(false);
  u32* sc = P<u32>(scal_);
  u64* hard = (u64*)(sc + 16);
  fill_u32((u32*)hard, INF, 2, st_);
  LcveArgs a;
  a.hist = hist;
  a.score = P<u32>(score_);
  a.hard = hard;
  a.cflag = P<u8>(cflag_);
  a.pt = (u32)opt.mu_pos << mu_inc_;
  a.nt = (u32)opt.mu_neg << mu_inc_;
  a.max_occurs = opt.lcve_max_occurs == INT32_MAX ? INF : (u32)opt.lcve_max_occurs;
  a.clause_max = opt.lcve_clause_max == INT32_MAX ? INF : (u32)opt.lcve_clause_max;
  k_lcve_classify(db, a, st_);
  k_lcve_select(db, a, P<u32>(flag_), P<u32>(cand_), P<u32>(key_), sc, tmp_.p, st_);
  u32 h[4];
  d2h(h, sc, 4, st_);
  d2h(h + 2, hard, 8, st_);
  stream_sync(st_);
  const u32 ncand = h[0];
  const u64 hk = ((u64)h[3] << 32) | h[2];
  int bits = 32;
  if (hk != ~0ull) {
    const u32 hs = (u32)(hk >> 32);
    bits = 0;
    while (bits < 32 && (hs >> bits)) bits++;
  }
  radix_sort_pairs(P<u32>(key_), P<u32>(cand_), ncand, bits, tmp_.p, st_);
  // positions share a word with the candidate's state (position << 2)
  if (ncand >= (1u << 30)) throw DeviceFail{"LCVE: too many candidates for the packed positions"};
  u32* resume = P<u32>(vsz_);  // free until SUB: where each scan goes on
  // worklists of undecided candidates (the selection buffers are free now);
  // after the first pass over all candidates, passes are queued eight at a
  // time with the list lengths kept on the device in a ring of three
  // counters (sc[12..14]: pass p reads ring[p % 3], appends to
  // ring[(p + 1) % 3] and clears ring[(p + 2) % 3] for the next pass; sc[6]
  // counts the passes that had work)
  u32* wl[2] = {P<u32>(flag_), P<u32>(score_)};
  u32* ring[3] = {sc + 12, sc + 13, sc + 14};
  // the election from scratch with 'lds' bytes reserved per step
  // workgroup and 'lanes' per candidate; sc[6] counts the worklist passes
  // that had work
  auto elect = [&](u32 lds, u32 lanes) {
    fill_u32(P<u32>(pos_), NOPOS, (size_t)V_ + 1, st_);
    k_lcve_positions(P<u32>(cand_), ncand, P<u32>(pos_), st_);
    fill_u32(P<u32>(state_), 0, ncand, st_);
    if (a.clause_max != INF) k_lcve_reject(P<u32>(cand_), ncand, P<u8>(cflag_), P<u32>(pos_), P<u32>(state_), st_);
    fill_u32(P<u32>(key_), NOPOS, ncand, st_);  // the sort keys are consumed: blockers
    fill_u32(resume, 0, ncand, st_);
    fill_u32(sc + 4, 0, 3, st_);
    fill_u32(sc + 12, 0, 3, st_);
    if (!ncand) return;
    k_mis_step(db, P<u32>(cand_), nullptr, ncand, P<u32>(pos_), P<u32>(state_), P<u32>(key_), resume, wl[0],
               ring[0], lds, lanes, st_);
    // the first batch of worklist passes is sized for every candidate (its
    // grid is at the cap either way on large rounds); later batches for the
    // list length read after the previous batch
    u32 p = 0, bound = ncand;
    while (bound) {
      for (int b = 0; b < 8; b++, p++) {
        k_mis_step_dev(db, P<u32>(cand_), wl[p & 1], ring[p % 3], bound, P<u32>(pos_), P<u32>(state_),
                       P<u32>(key_), resume, wl[(p + 1) & 1], ring[(p + 1) % 3], sc + 6, ring[(p + 2) % 3], lds,
                       lanes, st_);
      }
      if (p > ncand + 8) throw DeviceFail{"LCVE election did not converge"};
      bound = read_u32(ring[p % 3]);
    }
  };
  if (tune_mis_ && ncand) {
    // the election is the same set whatever the shape (the
    // lexicographically first independent set is unique): each candidate
    // shape runs it in full, twice over in turn, the last run stands.  The
    // lane count is chosen first (at the default or set reservation), then
    // the reservation at that lane count.
    tune_mis_ = false;
    auto fastest = [&](const std::vector<std::pair<u32, u32>>& shapes) {
      std::vector<float> best(shapes.size(), 0.f);
      for (int rep = 0; rep < 2; rep++)
        for (size_t j = 0; j < shapes.size(); j++) {
          Event* e0 = event_record(st_);
          elect(shapes[j].first, shapes[j].second);
          Event* e1 = event_record(st_);
          const float ms = event_ms(e0, e1);
          event_free(e0);
          event_free(e1);
          best[j] = rep ? std::min(best[j], ms) : ms;
        }
      size_t pick = 0;
      for (size_t j = 1; j < shapes.size(); j++)
        if (best[j] < best[pick]) pick = j;
      return shapes[pick];
    };
    const u32 lds0 = (u32)(opt.mis_lds >= 0 ? opt.mis_lds : 10240);
    std::vector<std::pair<u32, u32>> shapes;
    if (opt.mis_lanes > 0) shapes.push_back({lds0, (u32)opt.mis_lanes});
    else
      for (u32 g : {1u, 2u, 4u, 8u}) shapes.push_back({lds0, g});
    const u32 lanes = fastest(shapes).second;
    shapes.clear();
    if (opt.mis_lds >= 0) shapes.push_back({lds0, lanes});
    else
      for (u32 l : {0u, 10240u, 20480u}) shapes.push_back({l, lanes});
    const std::pair<u32, u32> pick = fastest(shapes);
    tune_.mis_lds = (int32_t)pick.first;
    tune_.mis_lanes = (int32_t)pick.second;
    tune_.mis = true;
    tune_store();
    used_lanes_ = tune_.mis_lanes;
    used_lds_ = tune_.mis_lds;
  } else {
    used_lds_ = lds_choice();
    used_lanes_ = lanes_choice();
    elect((u32)used_lds_, (u32)used_lanes_);
  }
  fill_u32(sc, 0, 1, st_);
  fill_u32(sc + 1, ncand, 1, st_);
  k_mis_break(P<u32>(cand_), ncand, P<u8>(cflag_), P<u32>(state_), sc + 1, st_);
  k_mis_collect(P<u32>(cand_), ncand, P<u32>(state_), 0, P<u32>(flag_), P<u32>(pv_), sc, tmp_.p, st_);
  u32 w7[7];  // sc[0]: elected, sc[6]: worklist passes with work
  d2h(w7, sc, sizeof(w7), st_);
  stream_sync(st_);
  E_ = w7[0];
  rounds_.back().elected = E_;
  rounds_.back().mis_iterations = ncand ? 1 + w7[6] : 0;
  stage_end();
  if (profile) {  // work measures, outside the stage's time
    uint64_t w[2];
    k_cand_work(db, P<u32>(cand_), ncand, P<u8>(mark_), (u64*)(P<u32>(scal_) + 40), st_);
    d2h(w, P<u32>(scal_) + 40, sizeof(w), st_);
    stream_sync(st_);
    rounds_.back().cand_occurs = w[0];
    rounds_.back().cand_bytes = w[1];
  }
  return E_ >= (u32)opt.lcve_min;
}

// ---------------------------------------------------------------- sortOT (simplify.cpp:81-93)
void Simplifier::sort_ot() {
  if (!E_) return;
  stage_begin(S_SORT);
  DB db = make_db(cur_);
  k_pv_lits(P<u32>(pv_), E_, P<u32>(elits_), st_);
  mem_.ensure(scratch_, std::max<size_t>(lits_end_, 6 * (size_t)E_ + 3) * 4, 0, st_);
  mem_.ensure(tmp2_, (10 * (size_t)E_ + 3) * 4, 0, st_);
  mem_.ensure(tmp_, std::max(tmp_.cap, scan_tmp_bytes(2 * (size_t)E_)), 0, st_);
  k_sort_lists(db, P<u32>(elits_), 2 * E_, P<u32>(tmp2_), P<u32>(scal_) + 24, P<u32>(scratch_), P<u32>(lsum_),
               tmp_.p, st_);
  if (opt.work_order) k_work_order(db, P<u32>(pv_), E_, P<u32>(tmp2_), P<u32>(order_), tmp_.p, st_);
  stage_end();
  if (profile) {  // work measures, outside the stage's time
    uint64_t w[2];
    k_elected_work(db, P<u32>(pv_), E_, (u64*)(P<u32>(scal_) + 40), st_);
    d2h(w, P<u32>(scal_) + 40, sizeof(w), st_);
    stream_sync(st_);
    rounds_.back().elected_occurs = w[0];
    rounds_.back().elected_bytes = w[1];
  }
}

// ---------------------------------------------------------------- SUB (elimination.cpp:110-129)
// Every elected variable runs SUB speculatively on the state before the
// stage.  Elected variables own disjoint clauses, so a variable's run can
// only be wrong when an earlier variable derived a unit u while one of its
// clauses holds u (toblivion(ot[u]) deletes that clause first) or when it
// met such a unit as a value.  Those variables (found through the unit's
// list and the clause owners) are restored and run again, seeing the units
// of earlier variables; their new units can affect later variables in turn,
// until no run is stale.  The smallest affected variable sees final units
// only, so every pass finalises at least one variable.  The units are then
// committed in elected order exactly as the sequential loop enqueues them.
bool Simplifier::sub() {
  if (!(opt.sub_en || opt.ve_plus_en) || !E_) return true;
  stage_begin(S_SUB);
  DB db = make_db(cur_);
  u32* sc = P<u32>(scal_);
  k_list_sizes(db, P<u32>(pv_), E_, 0, P<u32>(vsz_), st_);
  scan_u32(P<u32>(vsz_), P<u32>(vbase_), E_, sc + 5, tmp_.p, st_);
  const u32 O = read_u32(sc + 5);
  const size_t ob = std::max<size_t>(O, 1);
  mem_.ensure(snap_ref_, ob * 4, 0, st_);
  mem_.ensure(snap_flg_, ob, 0, st_);
  mem_.ensure(snap_size_, ob * 4, 0, st_);
  mem_.ensure(snap_sig_, ob * 4, 0, st_);
  mem_.ensure(snap_lbd_, ob * 4, 0, st_);
  mem_.ensure(subu_, ob * 8, 0, st_);
  mem_.ensure(unit_by_, std::max<size_t>(lits_end_, 2 * (size_t)V_ + 3) * 4, 0, st_);
  fill_u32(P<u32>(unit_by_), INF, lits_end_, st_);  // clause owners, per handle
  SubSnap snap{P<u32>(vbase_),    P<u32>(snap_np_),  P<u32>(snap_nn_),  P<u32>(snap_ref_), P<u8>(snap_flg_),
               P<u32>(snap_size_), P<u32>(snap_sig_), P<u32>(snap_lbd_), P<u32>(unit_by_)};
  // the snapshot for the re-runs is taken by the first pass (every
  // variable, just before its run)

  // per variable (elected index): its latest units (sequence, literal;
  // literal 0 = refutation) and proof events
  std::map<u32, std::vector<std::pair<u32, u32>>> units;
  std::map<u32, std::vector<ProofEv>> pmap;
  std::vector<u32> changes;  // (elected index, literal, pass)
  const u32 cap = V_ + 1;
  auto run = [&](const u32* idx, u32 n, const std::vector<u32>& ran, u32 pass, bool vis) {
    // vis == false: the first pass (every variable, nothing visible yet)
    fill_u32(sc, 0, 1, st_);
    SubOut out{P<u32>(recs_), sc, cap, P<u32>(subu_), P<u32>(vbase_), proof_out(2 * O + E_ + 64),
               SubVis{vis ? P<u32>(ust_) : nullptr, vis ? P<u32>(ulit_) : nullptr}};
    k_sub(db, opt, P<u32>(pv_), idx, n, out, P<u32>(vrest_), P<u32>(lsum_), tmp_.p, vis ? nullptr : &snap, st_);
    const u32 nrec = read_u32(sc);
    if (nrec > cap) throw DeviceFail{"SUB unit record buffer overflow"};
    std::vector<u32> rec(3 * (size_t)nrec);
    if (nrec) {
      d2h(rec.data(), recs_.p, rec.size() * 4, st_);
      stream_sync(st_);
    }
    std::map<u32, std::vector<std::pair<u32, u32>>> fresh;
    for (u32 i = 0; i < nrec; i++) fresh[rec[3 * i]].push_back({rec[3 * i + 1], rec[3 * i + 2]});
    if (!vis) {
      units = std::move(fresh);
      for (auto& kv : units)
        for (auto& e : kv.second)
          if (e.second) changes.insert(changes.end(), {kv.first, e.second, pass});
      std::vector<ProofEv> pev;
      proof_collect(pev, INF);
      for (const ProofEv& e : pev) pmap[e.k].push_back(e);
      return;
    }
    // unit changes of the variables that ran ('ran' is ascending); only a
    // variable with units before or after can change
    std::vector<u32> keys;
    for (auto& kv : units) keys.push_back(kv.first);
    for (auto& kv : fresh) keys.push_back(kv.first);
    std::sort(keys.begin(), keys.end());
    keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
    for (u32 k : keys) {
      if (!std::binary_search(ran.begin(), ran.end(), k)) continue;
      std::vector<u32> before, after;
      auto it = units.find(k);
      if (it != units.end())
        for (auto& e : it->second)
          if (e.second) before.push_back(e.second);
      auto jt = fresh.find(k);
      if (jt != fresh.end())
        for (auto& e : jt->second)
          if (e.second) after.push_back(e.second);
      std::sort(before.begin(), before.end());
      std::sort(after.begin(), after.end());
      std::vector<u32> diff;
      std::set_symmetric_difference(before.begin(), before.end(), after.begin(), after.end(),
                                    std::back_inserter(diff));
      for (u32 l : diff) changes.insert(changes.end(), {k, l, pass});
      if (jt != fresh.end()) units[k] = jt->second;
      else if (it != units.end()) units.erase(it);
    }
    for (auto it = pmap.begin(); it != pmap.end();) {
      if (std::binary_search(ran.begin(), ran.end(), it->first)) it = pmap.erase(it);
      else ++it;
    }
    std::vector<ProofEv> pev;
    proof_collect(pev, INF);
    for (const ProofEv& e : pev) pmap[e.k].push_back(e);
  };

  // The first refutation in elected order, as the commit below meets it: a
  // variable's own refutation (literal 0) or a unit whose variable is
  // already assigned the other way (at the root or by an earlier variable).
  // Variables after it cannot change the outcome (the sequential stage stops
  // there), so re-runs are limited to the variables up to it; affected
  // variables past it are deferred and run only if the refutation moves past
  // them (an earlier variable's re-run can withdraw the unit it met).
  auto first_refutation = [&]() -> u32 {
    std::unordered_map<u32, u32> over;  // variable -> true literal + 1, over h_assign_
    for (auto& kv : units) {
      std::vector<std::pair<u32, u32>> ev = kv.second;
      std::sort(ev.begin(), ev.end());
      for (auto& e : ev) {
        const u32 u = e.second;
        if (!u) return kv.first;
        auto it = over.find(ABS(u));
        const u32 a = it != over.end() ? it->second : h_assign_[ABS(u)];
        if (a) {
          if (a - 1 != u) return kv.first;
          continue;
        }
        over[ABS(u)] = u + 1;
      }
    }
    return INF;
  };

  // pass 1: every variable, nothing of this stage visible
  u32 pass = 1;
  fill_u32(P<u32>(sublast_), 1, E_, st_);
  run(work_order(), E_, {}, pass, false);
  std::vector<u32> vis_set;   // variables with a visible stage unit
  std::vector<u32> deferred;  // affected variables past the refutation, not re-run (ascending)
  bool vis_ready = false;
  while (!changes.empty() || !deferred.empty()) {
    if (pass > E_ + 2) throw DeviceFail{"SUB re-runs did not converge"};
    u32* aff = P<u32>(subaff_);
    u32* pos = aff + E_;
    u32* idx = aff + 2 * (size_t)E_;
    std::vector<u32> ran;
    if (!changes.empty()) {
      const u32 nch = (u32)(changes.size() / 3);
      mem_.ensure(ids_, changes.size() * 4, 0, st_);
      h2d(ids_.p, changes.data(), changes.size() * 4, st_);
      changes.clear();
      fill_u32(aff, 0, E_, st_);
      k_sub_affected(db, snap, P<u32>(ids_), nch, P<u32>(sublast_), aff, st_);
      k_pack_flags(aff, E_, pos, idx, sc + 6, tmp_.p, st_);
      const u32 m0 = read_u32(sc + 6);
      ran.resize(m0);
      if (m0) d2h(ran.data(), idx, (size_t)m0 * 4, st_);
      stream_sync(st_);
    }
    // the re-runs: affected or deferred variables up to the refutation
    // (the packed list is ascending: merged into the deferred ones, then
    // split at the refutation)
    const u32 kref = first_refutation();
    std::vector<u32> all(deferred.size() + ran.size());
    all.erase(std::set_union(deferred.begin(), deferred.end(), ran.begin(), ran.end(), all.begin()), all.end());
    const auto cut = std::upper_bound(all.begin(), all.end(), kref);
    ran.assign(all.begin(), cut);
    deferred.assign(cut, all.end());
    const u32 m = (u32)ran.size();
    if (!m) break;
    h2d(idx, ran.data(), (size_t)m * 4, st_);
    pass++;
    // visible stage units: per variable, the earliest deriving index
    if (!vis_ready) {
      fill_u32(P<u32>(ust_), INF, (size_t)V_ + 1, st_);
      vis_ready = true;
    } else if (!vis_set.empty()) {
      std::vector<u32> clr;
      for (u32 x : vis_set) clr.insert(clr.end(), {x, INF, 0});
      mem_.ensure(tmp3_, clr.size() * 4, 0, st_);
      h2d(tmp3_.p, clr.data(), clr.size() * 4, st_);
      k_set_vis(P<u32>(ust_), P<u32>(ulit_), P<u32>(tmp3_), (u32)(clr.size() / 3), st_);
    }
    std::map<u32, std::pair<u32, u32>> vis;  // variable -> (index, literal)
    for (auto& kv : units)
      for (auto& e : kv.second)
        if (e.second && !vis.count(ABS(e.second))) vis[ABS(e.second)] = {kv.first, e.second};
    std::vector<u32> ent;
    vis_set.clear();
    for (auto& kv : vis) {
      ent.insert(ent.end(), {kv.first, kv.second.first, kv.second.second});
      vis_set.push_back(kv.first);
    }
    if (!ent.empty()) {
      mem_.ensure(tmp3_, ent.size() * 4, 0, st_);
      h2d(tmp3_.p, ent.data(), ent.size() * 4, st_);
      k_set_vis(P<u32>(ust_), P<u32>(ulit_), P<u32>(tmp3_), (u32)(ent.size() / 3), st_);
    }
    k_sub_restore(db, P<u32>(pv_), idx, m, snap, st_);
    k_set_at(P<u32>(sublast_), idx, m, pass, st_);
    run(idx, m, ran, pass, true);
  }

  // commit the units in elected order (strengthen, elimination.cpp:214-226)
  std::vector<ProofEv> pev;
  for (auto& kv : pmap) pev.insert(pev.end(), kv.second.begin(), kv.second.end());
  std::vector<u32> committed;
  bool ok = true;
  for (auto& kv : units) {
    std::vector<std::pair<u32, u32>> ev = kv.second;
    std::sort(ev.begin(), ev.end());
    for (auto& e : ev) {
      const u32 u = e.second;
      if (!u) {  // refuted by this variable's own strengthening
        ok = false;
        break;
      }
      const u32 a = h_assign_[ABS(u)];
      if (a) {
        if (a - 1 != u) {
          ok = false;
          pev.push_back({kv.first, e.first, PR_EMPTY, 0});
          break;
        }
        continue;
      }
      pev.push_back({kv.first, e.first, PR_UNIT, u});
      set_host_assign(u);
      committed.push_back(u);
      units_out_.push_back(u);
      unassigned_--;
    }
    if (!ok) break;
  }
  if (ok && !committed.empty()) {
    commit_units(committed);
    // toblivion(ot[unit]) deletes the unit's clauses and empties its list
    k_oblivion(db, P<u32>(ids_), (u32)committed.size(), st_);
    k_zero_sizes(db, P<u32>(ids_), (u32)committed.size(), st_);
  }
  proof_events(pev);
  proof_flush();
  stream_sync(st_);
  rounds_.back().units += (u32)committed.size();
  rounds_.back().sub_passes = pass;
  stage_end();
  return ok;
}

// ---------------------------------------------------------------- BVE (bve.cpp:127-243)
bool Simplifier::ve() {
  if (!E_) return true;
  if (!opt.ve_en) {
    fill_u32(P<u32>(type_), 0, E_, st_);
    return true;
  }
  stage_begin(S_VE);
  DB db = make_db(cur_);
  u32* sc = P<u32>(scal_);
  k_list_sizes(db, P<u32>(pv_), E_, 2, P<u32>(vsz_), st_);
  scan_u32(P<u32>(vsz_), P<u32>(vbase_), E_, sc, tmp_.p, st_);
  const u32 S = read_u32(sc);
  mem_.ensure(scratch_, std::max<size_t>(S, 1) * 4, 0, st_);
  VeDecision d{P<u32>(type_), P<u32>(def_), P<u32>(vncls_), P<u32>(vnlits_), P<u32>(vnstack_),
               P<u32>(scratch_), P<u32>(vbase_)};
  k_ve_decide(db, opt, P<u32>(pv_), E_, d, work_order(), P<u32>(vrest_), P<u32>(lsum_), tmp_.p, st_);
  scan_u32(P<u32>(vncls_), P<u32>(clsb_), E_, sc, tmp_.p, st_);
  scan_u32(P<u32>(vnlits_), P<u32>(litb_), E_, sc + 1, tmp_.p, st_);
  scan_u32(P<u32>(vnstack_), P<u32>(stkb_), E_, sc + 2, tmp_.p, st_);
  u32 tot[3];
  d2h(tot, sc, 12, st_);
  stream_sync(st_);
  ensure_clause_room(tot[0], tot[1]);
  ensure_stack(stack_n_ + tot[2]);
  db = make_db(cur_);
  fill_u32(sc + 3, 0, 1, st_);
  VeApply ap;
  ap.cls_base = P<u32>(clsb_);
  ap.lit_base = P<u32>(litb_);
  ap.stack_base = P<u32>(stkb_);
  ap.cls0 = (u32)ncls_;
  ap.lit0 = (u32)lits_end_;
  ap.stack0 = (u32)stack_n_;
  ap.cls_cap = (u32)ccap_;
  ap.lit_cap = (u32)lcap_;
  ap.stack_cap = (u32)(stack_.cap / 4);
  ap.stack = P<u32>(stack_);
  ap.units = UnitOut{P<u32>(recs_), sc + 3, V_ + 1};
  ap.error = sc + 6;
  ap.proof = proof_out(tot[0] + S + V_ + 64);
  fill_u32(sc + 6, 0, 1, st_);
  k_ve_apply(db, P<u32>(pv_), E_, d, ap, P<u32>(vrest_), P<u32>(lsum_), tmp_.p, st_);
  // eliminated variables (markEliminated, solver.h:535-541), counted now so
  // that one read brings it with the unit records' count and the error flag
  k_count_nonzero(P<u32>(type_), E_, sc + 7, st_);
  u32 w5[5];  // sc[3]: unit records, sc[6]: write error, sc[7]: eliminated
  d2h(w5, sc + 3, sizeof(w5), st_);
  stream_sync(st_);
  if (w5[3]) throw DeviceFail{"BVE resolvent writes disagree with the counted space"};
  ncls_ += tot[0];
  lits_end_ += tot[1];
  stack_n_ += tot[2];
  const u32 nrec = w5[0], elim = w5[4];
  if (nrec > V_ + 1) throw DeviceFail{"BVE unit record buffer overflow"};
  bool ok = true;
  std::vector<u32> fresh;
  std::vector<ProofEv> pev;
  proof_collect(pev, INF);
  if (nrec) {
    std::vector<u32> rec(3 * (size_t)nrec);
    d2h(rec.data(), recs_.p, rec.size() * 4, st_);
    stream_sync(st_);
    std::vector<u32> idx(nrec);
    for (u32 i = 0; i < nrec; i++) idx[i] = i;
    std::sort(idx.begin(), idx.end(), [&](u32 a, u32 b) {
      return rec[3 * a] != rec[3 * b] ? rec[3 * a] < rec[3 * b] : rec[3 * a + 1] < rec[3 * b + 1];
    });
    // newResolvent / substitute_single unit handling (bve.cpp:300-311,
    // equivalence.h:58-66): assign if free, refute if false
    for (u32 i : idx) {
      const u32 u = rec[3 * i + 2];
      const u32 a = h_assign_[ABS(u)];
      if (a) {
        if (a - 1 != u) {
          ok = false;
          pev.push_back({rec[3 * i], rec[3 * i + 1], PR_EMPTY, 0});
          break;
        }
        continue;
      }
      pev.push_back({rec[3 * i], rec[3 * i + 1], PR_UNIT, u});
      set_host_assign(u);
      fresh.push_back(u);
      units_out_.push_back(u);
      unassigned_--;
    }
  }
  if (ok) commit_units(fresh);
  proof_events(pev);
  proof_flush();
  unassigned_ -= elim;
  rounds_.back().eliminated = elim;
  rounds_.back().resolvents = tot[0];
  rounds_.back().resolvent_words += tot[1];
  rounds_.back().stack_words += tot[2];
  rounds_.back().units += (u32)fresh.size();
  stage_end();
  return ok;
}

// ---------------------------------------------------------------- BCE (elimination.cpp:131-167)
void Simplifier::bce() {
  if (!opt.bce_en || !E_) return;
  stage_begin(S_BCE);
  DB db = make_db(cur_);
  u32* sc = P<u32>(scal_);
  k_bce(db, opt, P<u32>(pv_), P<u32>(type_), E_, P<u32>(vnstack_), nullptr, nullptr, true, work_order(), st_);
  scan_u32(P<u32>(vnstack_), P<u32>(stkb_), E_, sc, tmp_.p, st_);
  const u32 T = read_u32(sc);
  if (T) {
    ensure_stack(stack_n_ + T);
    k_bce(db, opt, P<u32>(pv_), P<u32>(type_), E_, nullptr, P<u32>(stkb_), P<u32>(stack_) + stack_n_, false,
          work_order(), st_);
    stack_n_ += T;
    rounds_.back().stack_words += T;
  }
  stream_sync(st_);
  stage_end();
}

// ---------------------------------------------------------------- ERE (elimination.cpp:169-197)
void Simplifier::ere() {
  if (!opt.ere_en || !E_) return;
  if (!ere_pairs()) ere_per_var();
}

// Pair-parallel ERE: every resolvent of every elected variable is looked up
// in a hash table of the live clauses; the pairs that found equal clauses are
// then replayed in elected-variable and pair order, which decides exactly
// which clauses the sequential ERE (elimination.cpp:169-197) deletes: a pair
// whose antecedent was deleted by an earlier pair is skipped, and a pair
// deletes its lowest-id match that is still live.  The elected variables are
// taken in chunks of fewer than 2^31 pairs (one thread per pair); a record
// keeps a pair's ERE_K lowest matches, and the pairs with more get their
// full match lists from a second pass.
bool Simplifier::ere_pairs() {
  stage_begin(S_ERE);
  DB db = make_db(cur_);
  u32* sc = P<u32>(scal_);
  u64* tot64 = (u64*)(sc + 32);
  k_ere_pair_counts(db, opt, P<u32>(pv_), E_, P<u32>(vncls_), tot64, st_);
  u64 total = 0;
  d2h(&total, tot64, 8, st_);
  stream_sync(st_);
  // chunks of variables [kb[c], kb[c + 1]) with fewer than 2^31 pairs each
  std::vector<u32> kb{0};
  if (total >= (1ull << 31)) {
    std::vector<u32> np(E_);
    d2h(np.data(), vncls_.p, (size_t)E_ * 4, st_);
    stream_sync(st_);
    u64 acc = 0;
    for (u32 k = 0; k < E_; k++) {
      if (acc + np[k] >= (1ull << 31)) {
        kb.push_back(k);
        acc = 0;
      }
      acc += np[k];
    }
  }
  kb.push_back(E_);
  bool ok = true;
  std::vector<u32> rec;                           // every chunk's records, k made global
  std::unordered_map<u32, std::vector<u32>> full;  // record -> all its matches, ascending
  if (total) {
    size_t cap = 1024;
    while (cap < 2 * (size_t)ncls_) cap <<= 1;
    mem_.ensure(ht_, cap * 8, 0, st_);
    HashTable ht{(u64*)ht_.p, (u32)(cap - 1)};
    k_ht_build(db, ht, opt.ere_max_resolvent, st_);
    u32 rcap = std::max<u32>(1u << 16, V_ + 1), ocap = 1u << 12;
    for (size_t ci = 0; ok && ci + 1 < kb.size(); ci++) {
      const u32 k0 = kb[ci], nk = kb[ci + 1] - k0;
      if (!nk) continue;
      scan_u32(P<u32>(vncls_) + k0, P<u32>(clsb_), nk, sc + 34, tmp_.p, st_);
      const u32 npairs = kb.size() == 2 ? (u32)total : read_u32(sc + 34);  // one chunk: the total
      if (!npairs) continue;
      // records and overflow indices, grown to the counts of a pass that
      // overflowed them and the pass run again
      u32 h[2] = {0, 0};
      ErePairs a{};
      for (int attempt = 0;; attempt++) {
        try {
          mem_.ensure(erecs_, ((size_t)rcap * ERE_REC_WORDS + ocap) * 4, 0, st_);
        } catch (AllocFail&) {
          ok = false;
          break;
        }
        fill_u32(sc, 0, 2, st_);
        a = ErePairs{P<u32>(pv_) + k0, P<u32>(clsb_), nk, P<u32>(erecs_), sc, rcap, sc + 1,
                     P<u32>(erecs_) + (size_t)rcap * ERE_REC_WORDS, ocap, ht, opt.ere_max_resolvent};
        k_ere_pairs(db, a, npairs, st_);
        d2h(h, sc, 8, st_);
        stream_sync(st_);
        if (h[0] <= rcap && h[1] <= ocap) break;
        if (attempt == 2 || h[0] >= (1u << 27)) {
          ok = false;
          break;
        }
        rcap = std::max(rcap, h[0]);
        ocap = std::max(ocap, h[1]);
      }
      if (!ok || !h[0]) continue;
      const u32 nrec = h[0], novf = h[1], first = (u32)(rec.size() / ERE_REC_WORDS);
      rec.resize(rec.size() + (size_t)nrec * ERE_REC_WORDS);
      u32* cr = rec.data() + (size_t)first * ERE_REC_WORDS;
      d2h(cr, erecs_.p, (size_t)nrec * ERE_REC_WORDS * 4, st_);
      stream_sync(st_);
      if (novf) {
        std::vector<u32> ovf(novf), off(novf + 1, 0);
        d2h(ovf.data(), a.ovf, (size_t)novf * 4, st_);
        stream_sync(st_);
        for (u32 o = 0; o < novf; o++) off[o + 1] = off[o] + cr[(size_t)ovf[o] * ERE_REC_WORDS + 4];
        mem_.ensure(tmp3_, ((size_t)novf + 1 + off[novf]) * 4, 0, st_);
        u32* d_off = P<u32>(tmp3_);
        u32* d_out = d_off + novf + 1;
        h2d(d_off, off.data(), ((size_t)novf + 1) * 4, st_);
        k_ere_all_matches(db, a, novf, d_off, d_out, st_);
        std::vector<u32> m(off[novf]);
        d2h(m.data(), d_out, m.size() * 4, st_);
        stream_sync(st_);
        for (u32 o = 0; o < novf; o++) {
          std::vector<u32> l(m.begin() + off[o], m.begin() + off[o + 1]);
          std::sort(l.begin(), l.end());
          full[first + ovf[o]] = std::move(l);
        }
      }
      for (u32 i = 0; i < nrec; i++) cr[(size_t)i * ERE_REC_WORDS] += k0;
    }
  }
  const u32 nrec = (u32)(rec.size() / ERE_REC_WORDS);
  if (ok && nrec) {
    auto R = [&](u32 i, int w) { return rec[(size_t)i * ERE_REC_WORDS + w]; };
    std::vector<u32> idx(nrec);
    for (u32 i = 0; i < nrec; i++) idx[i] = i;
    std::sort(idx.begin(), idx.end(), [&](u32 x, u32 y) {
      return R(x, 0) != R(y, 0) ? R(x, 0) < R(y, 0) : R(x, 1) < R(y, 1);
    });
    std::unordered_map<u32, bool> dead;
    std::vector<u32> del;
    for (u32 i : idx) {
      if (dead.count(R(i, 2)) || dead.count(R(i, 3))) continue;
      const u32 nm = R(i, 4);
      u32 target = INF;
      if (nm > (u32)ERE_K) {
        for (u32 c : full[i])
          if (!dead.count(c)) {
            target = c;
            break;
          }
      } else {
        for (u32 q = 0; q < nm; q++)
          if (!dead.count(R(i, 5 + q))) {
            target = R(i, 5 + q);
            break;
          }
      }
      if (target == INF) continue;
      dead[target] = true;
      del.push_back(target);
    }
    for (u32 c : del) proof_line(PR_DEL, c);  // removeClause (redundancy.h:137)
    proof_flush();
    if (!del.empty()) {
      mem_.ensure(ids_, del.size() * 4, 0, st_);
      h2d(ids_.p, del.data(), del.size() * 4, st_);
      k_apply_deletions(db, P<u32>(ids_), (u32)del.size(), st_);
      stream_sync(st_);
    }
  }
  if (ok) rounds_.back().ere_passes = 1;
  stage_end();
  return ok;
}

void Simplifier::ere_per_var() {
  if (!opt.ere_en || !E_) return;
  stage_begin(S_ERE);
  DB db = make_db(cur_);
  u32* sc = P<u32>(scal_);
  const u32 cap = V_ + 1;
  mem_.ensure(unit_by_, std::max<size_t>(lits_end_, 2 * (size_t)V_ + 3) * 4, 0, st_);
  u32 from = 0, passes = 0;
  std::vector<ProofEv> pev;
  while (true) {
    if (++passes > E_ + 2) throw DeviceFail{"ERE re-runs did not converge"};
    fill_u32(sc, 0, 2, st_);
    EreOut out{P<u32>(recs_), sc, cap, sc + 1};
    k_ere(db, opt, P<u32>(pv_), from, E_, out, st_);
    u32 h[2];
    d2h(h, sc, 8, st_);
    stream_sync(st_);
    if (h[1]) throw DeviceFail{"a variable deleted more clauses in ERE than the per-variable buffer holds"};
    if (h[0] > cap) throw DeviceFail{"ERE record buffer overflow"};
    const u32 nrec = h[0];
    if (!nrec) break;
    std::vector<u32> rec(3 * (size_t)nrec);
    d2h(rec.data(), recs_.p, rec.size() * 4, st_);
    stream_sync(st_);
    // a variable whose target was deleted by an earlier variable must re-run
    std::unordered_map<u32, u32> first_by;
    for (u32 i = 0; i < nrec; i++) {
      const u32 k = rec[3 * i], c = rec[3 * i + 2];
      auto it = first_by.find(c);
      if (it == first_by.end() || k < it->second) first_by[c] = k;
    }
    u32 first_dirty = INF;
    for (u32 i = 0; i < nrec; i++) {
      const u32 k = rec[3 * i], c = rec[3 * i + 2];
      if (first_by[c] < k) first_dirty = std::min(first_dirty, k);
    }
    fill_u32(P<u32>(unit_by_), INF, lits_end_, st_);  // per handle
    k_del_by(P<u32>(recs_), nrec, P<u32>(unit_by_), st_);
    fill_u32(sc + 2, INF, 1, st_);
    k_ere_dirty(db, P<u32>(pv_), from, E_, P<u32>(unit_by_), sc + 2, st_);
    first_dirty = std::min(first_dirty, read_u32(sc + 2));
    std::vector<u32> del;
    for (u32 i = 0; i < nrec; i++)
      if (rec[3 * i] < first_dirty) {
        del.push_back(rec[3 * i + 2]);
        pev.push_back({rec[3 * i], rec[3 * i + 1], PR_DEL, rec[3 * i + 2]});
      }
    if (!del.empty()) {
      mem_.ensure(ids_, del.size() * 4, 0, st_);
      h2d(ids_.p, del.data(), del.size() * 4, st_);
      k_apply_deletions(db, P<u32>(ids_), (u32)del.size(), st_);
    }
    if (first_dirty == INF) break;
    from = first_dirty;
  }
  proof_events(pev);
  proof_flush();
  stream_sync(st_);
  rounds_.back().ere_passes = passes;
  stage_end();
}

// ---------------------------------------------------------------- DRAT proof
// The sequential simplifier writes its proof as it goes (src/cpu/proof.cpp):
// units (learnUnit), strengthened clauses, resolvents and substituted
// clauses (addResolvent: dropped when already satisfied), SUB and ERE
// deletions (removeClause), the empty clause (learnEmpty).  Here every stage
// queues its lines in that order and encodes them on the device before the
// next stage can renumber or change clauses.
void Simplifier::proof_line(u32 kind, u32 payload) {
  if (!opt.proof_en) return;
  if (kind == PR_UNIT) pst_.push_back({ABS(payload), (u32)(pl_.size() / 2 + 1)});
  pl_.push_back(kind);
  pl_.push_back(payload);
}
void Simplifier::proof_events(std::vector<ProofEv>& ev) {
  if (!opt.proof_en) return;
  std::stable_sort(ev.begin(), ev.end(), [](const ProofEv& a, const ProofEv& b) {
    return a.k != b.k ? a.k < b.k : a.seq < b.seq;
  });
  for (const ProofEv& e : ev) {
    proof_line(e.kind, e.payload);
    if (e.kind == PR_EMPTY) break;  // the simplifier stops at a refutation
  }
  ev.clear();
}
ProofOut Simplifier::proof_out(u32 cap) {
  if (!opt.proof_en) return ProofOut{nullptr, nullptr, 0, nullptr};
  mem_.ensure(precs_, (size_t)cap * 16, 0, st_);
  u32* sc = P<u32>(scal_) + 48;
  fill_u32(sc, 0, 2, st_);
  return ProofOut{P<u32>(precs_), sc, cap, sc + 1};
}
void Simplifier::proof_collect(std::vector<ProofEv>& ev, u32 below_k) {
  if (!opt.proof_en) return;
  u32 h[2];
  d2h(h, P<u32>(scal_) + 48, 8, st_);
  stream_sync(st_);
  if (h[1]) throw DeviceFail{"proof record buffer overflow"};
  if (!h[0]) return;
  std::vector<u32> r((size_t)h[0] * 4);
  d2h(r.data(), precs_.p, r.size() * 4, st_);
  stream_sync(st_);
  for (u32 i = 0; i < h[0]; i++)
    if (r[4 * i] < below_k) ev.push_back({r[4 * i], r[4 * i + 1], r[4 * i + 2], r[4 * i + 3]});
}
void Simplifier::proof_flush() {
  if (!opt.proof_en || pl_.empty()) {
    pl_.clear();
    pst_.clear();
    return;
  }
  const u32 n = (u32)(pl_.size() / 2), np = (u32)pst_.size();
  for (const auto& q : pst_) {
    pl_.push_back(q.first);
    pl_.push_back(q.second);
  }
  mem_.ensure(plines_, pl_.size() * 4, 0, st_);
  h2d(plines_.p, pl_.data(), pl_.size() * 4, st_);
  const u32* lines = P<u32>(plines_);
  const u32* pairs = lines + 2 * (size_t)n;
  if (pstamp_.cap < ((size_t)V_ + 1) * 4) {
    mem_.ensure(pstamp_, ((size_t)V_ + 1) * 4, 0, st_);
    fill_u32(P<u32>(pstamp_), 0, (size_t)V_ + 1, st_);
  }
  mem_.ensure(pbytes_, 2 * (size_t)n * 4, 0, st_);
  mem_.ensure(tmp_, std::max(tmp_.cap, scan_tmp_bytes(n)), 0, st_);
  u32* bytes = P<u32>(pbytes_);
  u32* off = bytes + n;
  u32* total = P<u32>(scal_) + 50;
  DB db = make_db(cur_);
  k_set_stamps(P<u32>(pstamp_), pairs, np, false, st_);
  k_proof_sizes(db, lines, n, P<u32>(pstamp_), bytes, st_);
  scan_u32(bytes, off, n, total, tmp_.p, st_);
  const u32 add = read_u32(total);
  if (proof_.cap < proof_n_ + add) mem_.ensure(proof_, std::max<size_t>((proof_n_ + add) * 3 / 2, 4096), proof_n_, st_);
  k_proof_write(db, lines, n, bytes, off, P<u8>(proof_) + proof_n_, st_);
  k_set_stamps(P<u32>(pstamp_), pairs, np, true, st_);
  stream_sync(st_);
  proof_n_ += add;
  pl_.clear();
  pst_.clear();
}
void Simplifier::download_proof(uint8_t* bytes) {
  if (!proof_n_) return;
  d2h(bytes, proof_.p, proof_n_, st_);
  stream_sync(st_);
}

// ---------------------------------------------------------------- simplify (simplify.cpp:157-236)
int Simplifier::simplify(uint32_t max_rounds) {
  rounds_.clear();
  for (int s = 0; s < S_NSTAGES; s++) stage_ms[s] = 0, stage_calls[s] = 0;
  proof_n_ = 0;
  pl_.clear();
  pst_.clear();
  // a resident formula simplified before: its resolvents are no longer new,
  // and, as the reference's next call extracts the clauses it wrote back
  // (orgs, then learnts), its originals come first
  if (resident_) {
    k_clear_added(make_db(cur_), st_);
    if (has_learnts_) originals_first();
  }
  resident_ = true;
  kernel_times.clear();
  prof_enable(st_, profile);
  const int r = simplify_rounds(max_rounds);
  if (profile) prof_collect(st_, kernel_times);
  prof_enable(st_, false);
  return r;
}

int Simplifier::simplify_rounds(uint32_t max_rounds) {
  if (refuted_) {
    proof_line(PR_EMPTY, 0);
    proof_flush();
    return 0;
  }
  if (!ncls_) return 2;
  phase_ = 0;
  mu_inc_ = 0;
  uint64_t nC, nL;
  if (counts_known_) {  // the kept input, untouched since it was counted
    nC = in_live_cls_;
    nL = in_live_lits_;
  } else {
    count(nC, nL);
  }
  counts_known_ = false;
  tune_begin(nC, nL);
  int64_t litsbefore = (int64_t)nL, diff = INT64_MAX;
  u32 r = 0;
  while (nC && nL) {
    if (max_rounds && r >= max_rounds) break;
    rounds_.emplace_back();
    rounds_.back().clauses_before = nC;
    rounds_.back().literals_before = nL;
    // compaction keeps clause order, so when it runs does not change the
    // result; collect once more than half of the slots are dead
    if (ncls_ > 2 * nC + 1024) gc();
    create_ot();
    if (!prop()) return 0;
    if (!lcve()) break;
    sort_ot();
    if ((phase_ == opt.phases) || (diff <= opt.lits_min && phase_ > 2)) {
      ere();
      break;
    }
    if (!sub()) return 0;
    if (!ve()) return 0;
    bce();
    count(nC, nL);
    rounds_.back().clauses_after = nC;
    rounds_.back().literals_after = nL;
    diff = litsbefore - (int64_t)nL;
    litsbefore = (int64_t)nL;
    phase_++;
    mu_inc_++;
    mu_inc_ += phase_ == opt.phases;
    r++;
  }
  if (!prop()) return 0;
  gc();
  nC = ncls_;  // the compaction keeps exactly the live clauses
  if (!unassigned_ || !nC) return 1;
  if (opt.aggr_sort) aggr_sort();
  return 2;
}

// aggresivesort: the result's clauses in the stable order of the list key
// (src/cpu/simplify.cpp:263).  The sorted handles are compacted into the
// other block, so handles grow with the clause order again.
void Simplifier::aggr_sort() {
  if (ncls_ < 2) return;
  stage_begin(S_GC);
  mem_.ensure(tmp2_, ncls_ * 4, 0, st_);
  mem_.ensure(tmp_, std::max(tmp_.cap, radix_tmp_bytes(ncls_)), 0, st_);
  k_aggr_sort(make_db(cur_), P<u32>(tmp2_), tmp_.p, st_);
  stage_end();
  gc();
}

// ---------------------------------------------------------------- ELS
// Solver::ELS (src/cpu/equivalence.cpp:316-323) on the resident formula:
// decompose (:41-199) repeated while it succeeds.  One round: the binary
// implication graph over active literals, its SCCs (trimming, then
// coloring), representatives (smallest variable of the SCC; assumed
// variables keep themselves; a literal with its negation refutes), the
// substitution pass (substitute, :201-306, learnts before originals, with
// the units it finds visible to the clauses after them: the times of
// those units are the fixpoint of re-running the pass's unit detection),
// propagation of the new units (assignments only, as the reference's BCP),
// then SUBSTITUTED marks and binary stack records (:173-190).  After the
// last round the clauses are shrunk at the top level (rootify/shrinkTop)
// and compacted.
Simplifier::ElsInfo Simplifier::decompose(uint32_t max_rounds) {
  counts_known_ = false;
  if (refuted_) throw std::logic_error("the formula is refuted");
  if (!pending_units_.empty() || trail_n_ != propagated_)
    throw std::logic_error("root units are not propagated: simplify or parse first");
  ElsInfo info;
  ProfScope prof(this);
  // proof lines: written at the end from the clauses before the call (a
  // copy of the pool: the rounds rewrite clauses in place and keep their
  // handles until the final compaction) and after it
  proof_n_ = 0;
  pl_.clear();
  pst_.clear();
  const bool proof = opt.proof_en != 0;
  u32 els_conflict = 0;  // with a refutation: the unit written before the empty clause
  Buf before;
  BufScope before_scope{mem_, {&before}};
  if (proof) {
    mem_.ensure(before, std::max<uint64_t>(lits_end_, 1) * 4, 0, st_);
    d2d(before.p, lits_[cur_].p, lits_end_ * 4, st_);
  }
  struct NoProof {
    Options& o;
    int32_t saved;
    explicit NoProof(Options& x) : o(x), saved(x.proof_en) { o.proof_en = 0; }
    ~NoProof() { o.proof_en = saved; }
  } noproof(opt);
  const uint64_t start = trail_n_, stack0 = stack_n_;
  const size_t n2 = 2 * (size_t)V_ + 2, n1 = (size_t)V_ + 1;
  Buf deg, off, cur, adj, alive, indeg, outdeg, root, color, mark, fa, fb, list, rep, utime, ptime, lrank, tl, tpos,
      ukey, uval, sflag, spos, words;
  BufScope scope{mem_, {&deg, &off, &cur, &adj, &alive, &indeg, &outdeg, &root, &color, &mark, &fa, &fb, &list, &rep,
                        &utime, &ptime, &lrank, &tl, &tpos, &ukey, &uval, &sflag, &spos, &words}};
  for (Buf* b : {&deg, &cur, &alive, &indeg, &outdeg, &root, &color, &mark, &fa, &fb, &list, &rep})
    mem_.ensure(*b, n2 * 4, 0, st_);
  mem_.ensure(off, (n2 + 1) * 4, 0, st_);
  for (Buf* b : {&utime, &ptime, &sflag, &spos}) mem_.ensure(*b, n1 * 4, 0, st_);
  mem_.ensure(words, 64, 0, st_);
  u32* w = P<u32>(words);  // [0] count A, [1] count B, [2] changed, [3] alive, [4..5] rep info, [6] learnts,
                           // [7] touched, [8] units, [9..10] apply info, [11] substituted vars, [12] marked
  auto rd = [&](u32 k) {
    u32 x;
    d2h(&x, w + k, 4, st_);
    stream_sync(st_);
    return x;
  };
  auto trim = [&](u32 nf) {  // frontier in fa, count in w[0]
    for (u32 it = 0; nf && it <= n2; it++) {
      fill_u32(w + 1, 0, 1, st_);
      k_els_trim(P<u32>(fa), nf, P<u32>(off), P<u32>(adj), P<u32>(alive), P<u32>(indeg), P<u32>(outdeg),
                 P<u32>(root), P<u32>(fb), w + 1, st_);
      nf = rd(1);
      mem_.swap(fa, fb);
    }
  };
  bool unsat = false;
  for (uint32_t r = 0; r < max_rounds && !unsat; r++) {
    info.rounds++;
    DB db = make_db(cur_);
    mem_.ensure(tmp_, std::max({tmp_.cap, scan_tmp_bytes(n2 + 1), scan_tmp_bytes((size_t)ncls_ + 1),
                                radix_tmp_bytes((size_t)ncls_ + 1)}), 0, st_);
    // ---- graph and SCCs
    k_els_graph(db, P<u32>(deg), P<u32>(off), tmp_.p, st_);
    u32 E = 0;
    d2h(&E, P<u32>(off) + n2, 4, st_);
    stream_sync(st_);
    mem_.ensure(adj, ((size_t)E + 1) * 4, 0, st_);
    k_els_fill(db, P<u32>(off), P<u32>(cur), P<u32>(adj), st_);
    k_els_trim_init(db, P<u32>(off), P<u32>(alive), P<u32>(indeg), P<u32>(outdeg), P<u32>(root), P<u32>(fa), w, st_);
    trim(rd(0));
    for (u32 it = 0; it <= n2; it++) {  // each pass removes at least one SCC
      k_els_alive((u32)n2, P<u32>(alive), P<u32>(list), w + 3, st_);
      const u32 na = rd(3);
      if (!na) break;
      k_els_color_init(P<u32>(list), na, P<u32>(color), P<u32>(mark), st_);
      for (u32 k = 0; k <= na; k++) {
        k_els_fw(P<u32>(list), na, P<u32>(off), P<u32>(adj), P<u32>(alive), P<u32>(color), w + 2, st_);
        if (!rd(2)) break;
      }
      k_els_roots(P<u32>(list), na, P<u32>(color), P<u32>(mark), st_);
      for (u32 k = 0; k <= na; k++) {
        k_els_bw(P<u32>(list), na, P<u32>(off), P<u32>(adj), P<u32>(alive), P<u32>(color), P<u32>(mark), w + 2, st_);
        if (!rd(2)) break;
      }
      k_els_take(P<u32>(list), na, P<u32>(mark), P<u32>(color), P<u32>(alive), P<u32>(root), st_);
      k_els_redeg(P<u32>(list), na, P<u32>(off), P<u32>(adj), P<u32>(alive), P<u32>(indeg), P<u32>(outdeg),
                  P<u32>(root), P<u32>(fa), w, st_);
      trim(rd(0));
    }
    fill_u32(w + 4, 0, 2, st_);
    k_els_rep(db, P<u32>(root), P<u32>(rep), w + 4, st_);
    u32 ri[2];
    d2h(ri, w + 4, 8, st_);
    stream_sync(st_);
    if (ri[0]) {
      if (proof) {  // the literal whose negation shares its SCC (the first one)
        std::vector<u32> rt(n2);
        std::vector<u8> vs(n1);
        d2h(rt.data(), root.p, n2 * 4, st_);
        d2h(vs.data(), vstate_.p, n1, st_);
        stream_sync(st_);
        for (u32 l = 2; l < n2 && !els_conflict; l++)
          if (!vs[ABS(l)] && rt[l] == rt[FLIP(l)]) els_conflict = FLIP(l);
      }
      unsat = true;
      break;
    }
    const u32 nsub = ri[1];
    u32 nunits = 0, nbin = 0;
    const uint64_t visited = trail_n_;
    if (nsub) {
      // ---- substitution pass
      mem_.ensure(lrank, ((size_t)ncls_ + 1) * 4, 0, st_);
      mem_.ensure(tl, ((size_t)ncls_ + 1) * 4, 0, st_);
      mem_.ensure(tpos, ((size_t)ncls_ + 1) * 4, 0, st_);
      k_els_touch(db, P<u32>(rep), P<u32>(lrank), tmp_.p, w + 6, P<u32>(tl), P<u32>(tpos), w + 7, st_);
      const u32 nlearnt = rd(6);
      k_els_touch_list(db, P<u32>(rep), P<u32>(lrank), nlearnt, P<u32>(tl), P<u32>(tpos), w + 7, st_);
      const u32 nt = rd(7);
      fill_u32(P<u32>(utime), NOPOS, n1, st_);
      const ElsView ev{P<u32>(rep), P<u32>(utime)};
      for (u32 it = 0; it <= nt + 1; it++) {
        k_els_check(db, P<u32>(tl), P<u32>(tpos), nt, ev, P<u32>(ptime), st_);
        k_els_commit((u32)n1, P<u32>(ptime), P<u32>(utime), w + 2, st_);
        if (!rd(2)) break;
      }
      mem_.ensure(ukey, ((size_t)nt + 1) * 4, 0, st_);
      mem_.ensure(uval, ((size_t)nt + 1) * 4, 0, st_);
      fill_u32(w + 9, 0, 2, st_);
      k_els_apply(db, P<u32>(tl), P<u32>(tpos), nt, ev, P<u32>(ukey), P<u32>(uval), w + 8, w + 9, st_);
      u32 ai[3];
      d2h(ai, w + 8, 12, st_);
      stream_sync(st_);
      nunits = ai[0];
      nbin = ai[2];
      if (ai[1]) {
        unsat = true;
        break;
      }
      if (nunits) {
        radix_sort_pairs(P<u32>(ukey), P<u32>(uval), nunits, 32, tmp_.p, st_);
        k_els_enqueue(db, P<u32>(uval), nunits, P<u32>(trail_) + trail_n_, st_);
        std::vector<u32> us(nunits);
        d2h(us.data(), P<u32>(uval), (size_t)nunits * 4, st_);
        stream_sync(st_);
        for (u32 u : us) {
          set_host_assign(u);
          units_out_.push_back(u);
        }
        trail_n_ += nunits;
        unassigned_ -= nunits;
      }
      // ---- propagation of the new units (assignments only)
      if (trail_n_ > visited) {
        create_ot();
        if (!prop_visit(visited)) {
          unsat = true;
          break;
        }
      }
    }
    // ---- substituted variables and their stack records
    DB db2 = make_db(cur_);
    k_els_subst_flags(db2, P<u32>(rep), P<u32>(sflag), P<u32>(spos), tmp_.p, w + 11, st_);
    const u32 ns = rd(11);
    if (ns) {
      ensure_stack(stack_n_ + 5 * (uint64_t)ns);
      k_els_subst(db2, P<u32>(rep), P<u32>(sflag), P<u32>(spos), P<u32>(stack_) + stack_n_, w + 12, st_);
      const u32 marked = rd(12);
      stack_n_ += 5 * (uint64_t)ns;
      unassigned_ -= marked;
      info.substituted += marked;
    }
    const bool success = nsub && (nunits || nbin);
    if (!success) break;
  }
  info.units = trail_n_ - start;
  info.stack_words = stack_n_ - stack0;
  if (unsat) {
    if (proof) rewrite_proof(nullptr, start, true, els_conflict);
    refuted_ = true;
    info.status = 0;
    return info;
  }
  // rootify + shrinkTop over the units of the rounds, then compaction
  if (trail_n_ > start) {
    create_ot();
    prop_clean(start);
  }
  if (proof) rewrite_proof(P<u32>(before), start, false, 0);
  propagated_ = trail_n_;
  prop_changed_ = true;
  gc();
  return info;
}

// DRAT lines of a call that rewrote the clauses in place (decompose): the
// root units from trail[from..), then the new form of every clause that
// changed, then the deletions of the old forms and of the clauses gone
// (equivalence.cpp:250-270 writes per clause the addition and then the
// deletion; all additions first keep every old clause, the SCC binaries
// included, available to the later additions' unit propagation); a
// refutation ends with the empty clause, after the unit 'pre' when it is not
// 0 (a literal whose negation shares its SCC: the unit of its negation
// follows by propagation along the cycle, and then the empty clause).  old:
// the pool before the call (null: units and the empty clause only).
void Simplifier::rewrite_proof(const u32* old, uint64_t from, bool empty, u32 pre) {
  const u32 nu = (u32)(trail_n_ - from);
  DB db = make_db(cur_);
  if (!old) db.ncls = 0;
  const size_t n = (size_t)nu + 2 * (size_t)db.ncls;
  mem_.ensure(pbytes_, (n + 1) * 8, 0, st_);
  u32* bytes = P<u32>(pbytes_);
  u32* off = bytes + n + 1;
  mem_.ensure(tmp_, std::max(tmp_.cap, scan_tmp_bytes(n + 1)), 0, st_);
  u32* units = P<u32>(trail_) + from;
  k_rewrite_proof(db, old, units, nu, nullptr, nullptr, bytes, st_);
  u32* total = P<u32>(scal_) + 50;
  scan_u32(bytes, off, n, total, tmp_.p, st_);
  const u32 add = n ? read_u32(total) : 0;
  u8 tail[16];
  size_t nt = 0;
  if (empty && pre) {
    tail[nt++] = 'a';
    u32 vo = 0;  // the proof numbers literals by original variable
    d2h(&vo, P<u32>(vorg_) + ABS(pre), 4, st_);
    stream_sync(st_);
    u32 r = V2L(vo) | SIGN(pre);
    while (r > 127) {
      tail[nt++] = (u8)((r & 127) | 128);
      r >>= 7;
    }
    tail[nt++] = (u8)r;
    tail[nt++] = 0;
  }
  if (empty) {
    tail[nt++] = 'a';
    tail[nt++] = 0;
  }
  const size_t all = (size_t)add + nt;
  mem_.ensure(proof_, std::max<size_t>(all, 4096), 0, st_);
  k_rewrite_proof(db, old, units, nu, off, P<u8>(proof_), bytes, st_);
  if (nt) h2d(P<u8>(proof_) + add, tail, nt, st_);
  stream_sync(st_);
  proof_n_ = all;
}

// ---------------------------------------------------------------- results
void Simplifier::download_cnf(uint64_t* offsets, uint32_t* lits, uint8_t* flags, uint32_t* lbd) {
  // after the final compaction c_off lists every clause in order; the
  // clauses are packed on the device and only the requested arrays cross
  // the bus
  const size_t n = ncls_, nl = out_nlits();
  const size_t o_off = 0, o_lits = o_off + 8 * (n + 1), o_lbd = o_lits + 4 * nl, o_flags = o_lbd + 4 * n;
  mem_.ensure(out_, o_flags + n + 8, 0, st_);
  mem_.ensure(tmp2_, std::max<size_t>(n, 1) * 4, 0, st_);
  uint8_t* base = P<uint8_t>(out_);
  k_pack_cnf(make_db(cur_), P<u32>(tmp2_), tmp_.p, P<u32>(scal_) + 56, (u64*)(base + o_off), (u32*)(base + o_lits),
             base + o_flags, (u32*)(base + o_lbd), st_);
  if (offsets) d2h(offsets, base + o_off, 8 * (n + 1), st_);
  if (lits) d2h(lits, base + o_lits, 4 * nl, st_);
  if (flags) d2h(flags, base + o_flags, n, st_);
  if (lbd) d2h(lbd, base + o_lbd, 4 * n, st_);
  stream_sync(st_);
}

// ---------------------------------------------------------------- variable map
Simplifier::MapInfo Simplifier::map(bool force, double map_perc) {
  counts_known_ = false;
  if (refuted_) throw std::logic_error("the formula is refuted");
  if (!pending_units_.empty() || trail_n_ != propagated_)
    throw std::logic_error("root units are not propagated: simplify before mapping");
  const u32 V = V_;
  const size_t n1 = (size_t)V + 1;
  ProfScope prof(this);
  mem_.ensure(map_, n1 * 4, 0, st_);
  mem_.ensure(tmp2_, n1 * 4, 0, st_);
  mem_.ensure(tmp_, std::max(tmp_.cap, scan_tmp_bytes(n1)), 0, st_);
  u32* info = P<u32>(scal_) + 56;
  k_map_build(P<u8>(vstate_), P<u32>(assign_), V, P<u32>(tmp2_), P<u32>(map_), info, tmp_.p, st_);
  u32 h[4];
  d2h(h, info, sizeof(h), st_);
  stream_sync(st_);
  MapInfo r{0, V, V, h[0], 0, h[3]};
  map_called_ = true;
  map_applied_ = false;
  map_old_nvars_ = V;
  // canMap: inactive > map_perc * maxVar (DL() is 0 at the boundary)
  if (!force && !((double)h[3] > map_perc * (double)V)) return r;
  const u32 nV = h[1];
  map_first_ = h[0];
  map_first_true_ = h[2];
  VarMap m{P<u32>(map_), P<u32>(assign_), map_first_, map_first_true_};
  k_map_clauses(make_db(cur_), m, st_);
  // per-variable state in the new numbering; the old assignments stay for map_lits
  const size_t m1 = (size_t)nV + 1;
  Buf vorg2, vstate2, assumed2, assign2;
  mem_.ensure(vorg2, m1 * 4, 0, st_, R_STABLE);
  mem_.ensure(vstate2, m1, 0, st_, R_STABLE);
  mem_.ensure(assumed2, m1, 0, st_, R_STABLE);
  mem_.ensure(assign2, m1 * 4, 0, st_, R_STABLE);
  k_map_vars(m, V, P<u32>(vorg_), P<u8>(vstate_), P<u8>(assumed_), P<u32>(vorg2), P<u8>(vstate2), P<u8>(assumed2),
             P<u32>(assign2), st_);
  // trail: only 'first' survives (VMAP::mapShrinkLits, vmap.h:158-169)
  trail_n_ = propagated_ = 0;
  if (map_first_) {
    h2d(P<u32>(trail_), &map_first_true_, 4, st_);
    trail_n_ = propagated_ = 1;
  }
  stream_sync(st_);
  mem_.take(map_assign_, assign_);
  mem_.take(assign_, assign2);
  mem_.swap(vorg_, vorg2);
  mem_.swap(vstate_, vstate2);
  mem_.swap(assumed_, assumed2);
  mem_.release(vorg2);
  mem_.release(vstate2);
  mem_.release(assumed2);
  std::vector<u32> units;
  for (u32 u : units_out_)
    if (ABS(u) == map_first_) units.push_back(map_first_true_);
  units_out_.swap(units);
  h_assign_.assign(m1, 0);
  h_touched_.clear();
  if (map_first_) h_assign_[ABS(map_first_true_)] = map_first_true_ + 1;
  V_ = nV;
  remapped_ = true;
  map_applied_ = true;
  alloc_state();
  stream_sync(st_);
  r.applied = 1;
  r.new_nvars = nV;
  r.trail_lit = map_first_ ? map_first_true_ : 0;
  return r;
}

void Simplifier::download_map(uint32_t* mapped) {
  if (!map_called_) throw std::logic_error("no variable map was computed");
  if (!map_applied_) {
    for (u32 v = 0; v <= map_old_nvars_; v++) mapped[v] = v;
    return;
  }
  d2h(mapped, map_.p, ((size_t)map_old_nvars_ + 1) * 4, st_);
  stream_sync(st_);
}

void Simplifier::map_lits(uint32_t* lits, uint64_t n) {
  if (!map_called_) throw std::logic_error("no variable map was computed");
  for (uint64_t i = 0; i < n; i++)
    if (lits[i] && (ABS(lits[i]) == 0 || ABS(lits[i]) > map_old_nvars_))
      throw std::invalid_argument("literal out of range of the mapped numbering");
  if (!map_applied_ || !n) return;
  mem_.ensure(tmp3_, n * 4, 0, st_);
  h2d(tmp3_.p, lits, n * 4, st_);
  VarMap m{P<u32>(map_), P<u32>(map_assign_), map_first_, map_first_true_};
  k_map_lits(m, P<u32>(tmp3_), n, st_);
  d2h(lits, tmp3_.p, n * 4, st_);
  stream_sync(st_);
}

void Simplifier::map_lits_noalloc(uint32_t* lits, uint64_t n) {
  if (!map_called_) throw std::logic_error("no variable map was computed");
  if (n > SCRATCH_WORDS)
    throw std::invalid_argument("pfs_map_lits_noalloc maps at most " + std::to_string(SCRATCH_WORDS) +
                                " literals per call (" + std::to_string(n) + " given)");
  for (uint64_t i = 0; i < n; i++)
    if (lits[i] && (ABS(lits[i]) == 0 || ABS(lits[i]) > map_old_nvars_))
      throw std::invalid_argument("literal out of range of the mapped numbering");
  if (!map_applied_ || !n) return;
  // the scratch is written in full before the kernel reads it: nothing to clear
  h2d(fixed_scratch_, lits, n * 4, st_);
  VarMap m{P<u32>(map_), P<u32>(map_assign_), map_first_, map_first_true_};
  k_map_lits(m, fixed_scratch_, n, st_);
  d2h(lits, fixed_scratch_, n * 4, st_);
  stream_sync(st_);
}

void Simplifier::download_vstate(uint8_t* vs) {
  d2h(vs, vstate_.p, (size_t)V_ + 1, st_);
  stream_sync(st_);
}
void Simplifier::download_stack(uint32_t* words) {
  d2h(words, stack_.p, stack_n_ * 4, st_);
  stream_sync(st_);
}

}  // namespace pfs
