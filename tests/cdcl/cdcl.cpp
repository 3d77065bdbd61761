// A small CDCL SAT solver, model extension over the eliminated-variable
// stack, and a model check.
//
// TEST INFRASTRUCTURE ONLY: built into tests/cdcl/libcdcl.so and called by
// the tests (tests/cdcl/__init__.py) to check that a simplified formula is
// equisatisfiable with its input and that a model of the output, extended
// over the stack, satisfies the original formula.  Nothing in the product
// library links or calls it.
//
// Literals are 2v + s (v >= 1, s = 1 for the negative literal), the encoding
// of the simplifier's clause database and of its stack.
//
// Solver: two watched literals with a blocking literal, first-UIP learning
// with local minimisation, VSIDS on a binary heap, phase saving, Luby
// restarts (unit 100 conflicts) and LBD-ranked halving of the learnt
// clauses.  Assumptions are decided first, one per level.
//
// Model extension follows the reference's MODEL::extend
// (src/gpu/model.cpp:101-162): the stack holds records [witness, other
// literals..., size]; they are replayed from the end and a record with no
// true literal sets its witness true.

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

typedef uint32_t u32;
typedef uint64_t u64;
const u32 NOREF = 0xFFFFFFFFu;
const uint8_t UNDEF = 2;

class Solver {
 public:
  explicit Solver(u32 nvars)
      : n_(nvars), assign_(nvars + 1, UNDEF), level_(nvars + 1, 0), reason_(nvars + 1, NOREF),
        phase_(nvars + 1, 0), seen_(nvars + 1, 0), act_(nvars + 1, 0.0), hpos_(nvars + 1, -1),
        watches_(2 * (size_t)nvars + 2) {
    for (u32 v = 1; v <= n_; v++) heap_insert(v);
  }

  // false when the formula is refuted at the root
  bool add_clause(const u32* lits, u64 n) {
    if (!ok_) return false;
    tmp_.assign(lits, lits + n);
    std::sort(tmp_.begin(), tmp_.end());
    size_t j = 0;
    for (size_t i = 0; i < tmp_.size(); i++) {
      const u32 l = tmp_[i];
      if (j && tmp_[j - 1] == l) continue;                 // duplicate literal
      if (j && tmp_[j - 1] == (l ^ 1)) return true;        // tautology
      const uint8_t v = val(l);
      if (v == 1) return true;                             // satisfied at the root
      if (v == 0) continue;                                // false at the root
      tmp_[j++] = l;
    }
    tmp_.resize(j);
    if (tmp_.empty()) return ok_ = false;
    if (tmp_.size() == 1) {
      enqueue(tmp_[0], NOREF);
      if (propagate() != NOREF) ok_ = false;
      return ok_;
    }
    attach(new_clause(tmp_, false, 0));
    return true;
  }

  // 10 SAT, 20 UNSAT (under the assumptions), 0 conflict limit reached
  int solve(const u32* assumptions, u32 na, u64 conflict_limit) {
    if (!ok_) return 20;
    if (propagate() != NOREF) return (ok_ = false), 20;
    u64 conflicts = 0, restart_at = 100 * luby(0), restarts = 0, since_restart = 0;
    u64 reduce_at = std::max<u64>(2000, nclauses_ / 4);
    std::vector<u32> learnt;
    for (;;) {
      const u32 confl = propagate();
      if (confl != NOREF) {
        conflicts++, since_restart++;
        if (decision_level() == 0) return (ok_ = false), 20;
        u32 bt, lbd;
        analyze(confl, learnt, bt, lbd);
        backtrack(bt);
        if (learnt.size() == 1) {
          enqueue(learnt[0], NOREF);
        } else {
          const u32 cr = new_clause(learnt, true, lbd);
          attach(cr);
          learnts_.push_back(cr);
          enqueue(learnt[0], cr);
        }
        var_inc_ /= 0.95;
        if (conflict_limit && conflicts >= conflict_limit) {
          backtrack(0);
          return 0;
        }
        continue;
      }
      if (since_restart >= restart_at) {
        backtrack(0);
        since_restart = 0;
        restart_at = 100 * luby(++restarts);
      }
      if (learnts_.size() >= reduce_at) {
        reduce();
        reduce_at += reduce_at / 10 + 300;
      }
      u32 next = 0;
      while (decision_level() < na) {  // assumptions, one per level
        const u32 a = assumptions[decision_level()];
        const uint8_t v = val(a);
        if (v == 1) {
          trail_lim_.push_back((u32)trail_.size());  // already true: an empty level
        } else if (v == 0) {
          backtrack(0);
          return 20;
        } else {
          next = a;
          break;
        }
      }
      if (!next) {
        u32 v = 0;
        while (!heap_.empty()) {
          const u32 x = heap_pop();
          if (assign_[x] == UNDEF) {
            v = x;
            break;
          }
        }
        if (!v) return 10;  // every variable assigned: a model
        next = 2 * v + phase_[v];
      }
      trail_lim_.push_back((u32)trail_.size());
      enqueue(next, NOREF);
    }
  }

  void model(uint8_t* out) const {
    out[0] = 0;
    for (u32 v = 1; v <= n_; v++) out[v] = assign_[v] == 1 ? 1 : 0;
  }

 private:
  u32 n_;
  bool ok_ = true;
  std::vector<uint8_t> assign_;  // per variable: 0 false, 1 true, 2 unassigned
  std::vector<u32> level_, reason_;
  std::vector<uint8_t> phase_, seen_;
  std::vector<double> act_;
  double var_inc_ = 1.0;
  std::vector<int> hpos_;
  std::vector<u32> heap_;
  struct Watch {
    u32 cref, blocker;
  };
  std::vector<std::vector<Watch>> watches_;  // [l]: clauses watching l ^ 1
  // clause arena: [size, lbd << 2 | deleted << 1 | learnt, literals...]
  std::vector<u32> arena_;
  std::vector<u32> learnts_;
  u64 nclauses_ = 0;
  std::vector<u32> trail_, trail_lim_;
  size_t qhead_ = 0;
  std::vector<u32> tmp_, stack_, toclear_;

  static u64 luby(u64 i) {  // 1 1 2 1 1 2 4 1 1 2 ...
    u64 size = 1, seq = 0;
    while (size < i + 1) size = 2 * size + 1, seq++;
    while (size - 1 != i) {
      size = (size - 1) >> 1;
      seq--;
      i = i % size;
    }
    return (u64)1 << seq;
  }
  uint8_t val(u32 l) const {
    const uint8_t a = assign_[l >> 1];
    return a == UNDEF ? UNDEF : (uint8_t)(a ^ (l & 1));
  }
  u32 decision_level() const { return (u32)trail_lim_.size(); }
  u32 csize(u32 cr) const { return arena_[cr]; }
  u32* clits(u32 cr) { return &arena_[cr + 2]; }
  bool learnt(u32 cr) const { return arena_[cr + 1] & 1; }
  u32 lbd(u32 cr) const { return arena_[cr + 1] >> 2; }
  bool deleted(u32 cr) const { return arena_[cr + 1] & 2; }

  u32 new_clause(const std::vector<u32>& lits, bool is_learnt, u32 l) {
    if (arena_.size() + lits.size() + 2 >= NOREF) throw std::bad_alloc();
    const u32 cr = (u32)arena_.size();
    arena_.push_back((u32)lits.size());
    arena_.push_back((l << 2) | (is_learnt ? 1u : 0u));
    arena_.insert(arena_.end(), lits.begin(), lits.end());
    if (!is_learnt) nclauses_++;
    return cr;
  }
  void attach(u32 cr) {
    const u32* c = clits(cr);
    watches_[c[0] ^ 1].push_back({cr, c[1]});
    watches_[c[1] ^ 1].push_back({cr, c[0]});
  }
  void enqueue(u32 l, u32 from) {
    const u32 v = l >> 1;
    assign_[v] = (uint8_t)(!(l & 1));
    level_[v] = decision_level();
    reason_[v] = from;
    trail_.push_back(l);
  }
  u32 propagate() {
    u32 confl = NOREF;
    while (qhead_ < trail_.size()) {
      const u32 p = trail_[qhead_++];  // true; clauses watching p ^ 1 are visited
      const u32 fl = p ^ 1;
      std::vector<Watch>& ws = watches_[p];
      size_t i = 0, j = 0;
      const size_t n = ws.size();
      while (i < n) {
        const Watch w = ws[i++];
        if (val(w.blocker) == 1) {
          ws[j++] = w;
          continue;
        }
        u32* c = clits(w.cref);
        const u32 sz = csize(w.cref);
        if (c[0] == fl) std::swap(c[0], c[1]);
        const u32 first = c[0];
        const Watch nw{w.cref, first};
        if (first != w.blocker && val(first) == 1) {
          ws[j++] = nw;
          continue;
        }
        bool moved = false;
        for (u32 k = 2; k < sz; k++) {
          if (val(c[k]) != 0) {
            c[1] = c[k];
            c[k] = fl;
            watches_[c[1] ^ 1].push_back(nw);
            moved = true;
            break;
          }
        }
        if (moved) continue;
        ws[j++] = nw;
        if (val(first) == 0) {
          confl = w.cref;
          qhead_ = trail_.size();
          while (i < n) ws[j++] = ws[i++];
        } else {
          enqueue(first, w.cref);
        }
      }
      ws.resize(j);
      if (confl != NOREF) break;
    }
    return confl;
  }
  void bump(u32 v) {
    if ((act_[v] += var_inc_) > 1e100) {
      for (u32 x = 1; x <= n_; x++) act_[x] *= 1e-100;
      var_inc_ *= 1e-100;
    }
    if (hpos_[v] >= 0) heap_up((size_t)hpos_[v]);
  }
  // first UIP; learnt[0] is the asserting literal, learnt[1] the one of the
  // backtrack level
  void analyze(u32 confl, std::vector<u32>& out, u32& bt, u32& lbd_out) {
    out.clear();
    out.push_back(0);
    int path = 0;
    u32 p = 0;
    bool first = true;
    size_t idx = trail_.size();
    do {
      u32* c = clits(confl);
      const u32 sz = csize(confl);
      for (u32 k = first ? 0 : 1; k < sz; k++) {
        const u32 q = c[k], v = q >> 1;
        if (seen_[v] || level_[v] == 0) continue;
        seen_[v] = 1;
        bump(v);
        if (level_[v] >= decision_level()) path++;
        else out.push_back(q);
      }
      while (!seen_[trail_[--idx] >> 1]) {
      }
      p = trail_[idx];
      confl = reason_[p >> 1];
      seen_[p >> 1] = 0;
      path--;
      first = false;
    } while (path > 0);
    out[0] = p ^ 1;
    // local minimisation: drop a literal whose reason's other literals are
    // all in the clause (or fixed at the root)
    toclear_.assign(out.begin(), out.end());
    size_t j = 1;
    for (size_t i = 1; i < out.size(); i++) {
      const u32 v = out[i] >> 1;
      const u32 r = reason_[v];
      bool keep = true;
      if (r != NOREF) {
        keep = false;
        const u32* c = clits(r);
        for (u32 k = 1; k < csize(r); k++) {
          const u32 u = c[k] >> 1;
          if (!seen_[u] && level_[u] > 0) {
            keep = true;
            break;
          }
        }
      }
      if (keep) out[j++] = out[i];
    }
    out.resize(j);
    for (u32 l : toclear_) seen_[l >> 1] = 0;
    bt = 0;
    if (out.size() > 1) {
      size_t m = 1;
      for (size_t i = 2; i < out.size(); i++)
        if (level_[out[i] >> 1] > level_[out[m] >> 1]) m = i;
      std::swap(out[1], out[m]);
      bt = level_[out[1] >> 1];
    }
    // LBD: distinct decision levels
    stack_.clear();
    for (u32 l : out) stack_.push_back(level_[l >> 1]);
    std::sort(stack_.begin(), stack_.end());
    lbd_out = (u32)(std::unique(stack_.begin(), stack_.end()) - stack_.begin());
  }
  void backtrack(u32 lvl) {
    if (decision_level() <= lvl) return;
    for (size_t i = trail_.size(); i-- > trail_lim_[lvl];) {
      const u32 v = trail_[i] >> 1;
      phase_[v] = (uint8_t)(trail_[i] & 1);
      assign_[v] = UNDEF;
      reason_[v] = NOREF;
      if (hpos_[v] < 0) heap_insert(v);
    }
    trail_.resize(trail_lim_[lvl]);
    trail_lim_.resize(lvl);
    qhead_ = trail_.size();
  }
  bool locked(u32 cr) {
    const u32 l = clits(cr)[0];
    return val(l) == 1 && reason_[l >> 1] == cr;
  }
  // keep the better half of the learnt clauses (low LBD first; LBD <= 2 kept)
  void reduce() {
    std::sort(learnts_.begin(), learnts_.end(), [&](u32 a, u32 b) { return lbd(a) < lbd(b); });
    const size_t half = learnts_.size() / 2;
    std::vector<u32> kept;
    for (size_t i = 0; i < learnts_.size(); i++) {
      const u32 cr = learnts_[i];
      if (i < half || lbd(cr) <= 2 || locked(cr)) kept.push_back(cr);
      else arena_[cr + 1] |= 2;
    }
    learnts_.swap(kept);
    for (auto& ws : watches_) {
      size_t j = 0;
      for (size_t i = 0; i < ws.size(); i++)
        if (!deleted(ws[i].cref)) ws[j++] = ws[i];
      ws.resize(j);
    }
  }
  // max-heap of variables by activity
  void heap_insert(u32 v) {
    hpos_[v] = (int)heap_.size();
    heap_.push_back(v);
    heap_up(heap_.size() - 1);
  }
  void heap_up(size_t i) {
    const u32 v = heap_[i];
    while (i) {
      const size_t p = (i - 1) / 2;
      if (act_[heap_[p]] >= act_[v]) break;
      heap_[i] = heap_[p];
      hpos_[heap_[i]] = (int)i;
      i = p;
    }
    heap_[i] = v;
    hpos_[v] = (int)i;
  }
  u32 heap_pop() {
    const u32 top = heap_[0];
    hpos_[top] = -1;
    const u32 last = heap_.back();
    heap_.pop_back();
    if (!heap_.empty()) {
      size_t i = 0;
      const size_t n = heap_.size();
      for (;;) {
        size_t c = 2 * i + 1;
        if (c >= n) break;
        if (c + 1 < n && act_[heap_[c + 1]] > act_[heap_[c]]) c++;
        if (act_[heap_[c]] <= act_[last]) break;
        heap_[i] = heap_[c];
        hpos_[heap_[i]] = (int)i;
        i = c;
      }
      heap_[i] = last;
      hpos_[last] = (int)i;
    }
    return top;
  }
};

}  // namespace

extern "C" {

// Solves the clauses (offsets: ncls + 1 entries) plus the given units under
// the assumptions.  Returns 10 (SAT; model[0..maxvar] filled when model is
// not null), 20 (UNSAT), 0 (conflict limit reached, 0 = none) or -1 (out of
// memory).
int cdcl_solve(u32 maxvar, u64 ncls, const u64* offsets, const u32* lits, const u32* units, u64 nunits,
               const u32* assumptions, u32 nassumptions, u64 conflict_limit, uint8_t* model) {
  try {
    Solver s(maxvar);
    bool ok = true;
    for (u64 i = 0; ok && i < nunits; i++) ok = s.add_clause(units + i, 1);
    for (u64 c = 0; ok && c < ncls; c++) ok = s.add_clause(lits + offsets[c], offsets[c + 1] - offsets[c]);
    const int r = ok ? s.solve(assumptions, nassumptions, conflict_limit) : 20;
    if (r == 10 && model) s.model(model);
    return r;
  } catch (const std::bad_alloc&) {
    return -1;
  }
}

// MODEL::extend over the stack (src/gpu/model.cpp:101-162): values[v] in
// {0, 1} for v in 1..maxvar, updated in place.  Returns the number of
// witnesses flipped.
u64 cdcl_extend(const u32* stack, u64 nwords, uint8_t* values) {
  u64 flipped = 0;
  u64 i = nwords;
  while (i > 1) {
    const u32 k = stack[i - 1];
    if (!k || k > i - 1) return ~(u64)0;  // malformed record
    const u64 first = i - 1 - k;
    bool sat = false;
    for (u64 j = first + 1; j < i - 1 && !sat; j++) sat = values[stack[j] >> 1] == !(stack[j] & 1);
    if (!sat) {
      const u32 w = stack[first];
      values[w >> 1] = (uint8_t)!(w & 1);
      flipped++;
    }
    i = first;
  }
  return flipped;
}

// The number of clauses (and units) the values falsify; the index of the
// first falsified clause in *first (ncls + unit index for a unit, ~0 when
// none).
u64 cdcl_verify(u64 ncls, const u64* offsets, const u32* lits, const u32* units, u64 nunits,
                const uint8_t* values, u64* first) {
  u64 bad = 0;
  *first = ~(u64)0;
  for (u64 c = 0; c < ncls; c++) {
    bool sat = false;
    for (u64 k = offsets[c]; k < offsets[c + 1] && !sat; k++) sat = values[lits[k] >> 1] == !(lits[k] & 1);
    if (!sat) {
      if (!bad) *first = c;
      bad++;
    }
  }
  for (u64 i = 0; i < nunits; i++) {
    if (values[units[i] >> 1] != !(units[i] & 1)) {
      if (!bad) *first = ncls + i;
      bad++;
    }
  }
  return bad;
}

}  // extern "C"
