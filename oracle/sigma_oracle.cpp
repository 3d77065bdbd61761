// ORACLE -- test infrastructure only.
//
// A from-scratch CPU restatement of ParaFROST's CPU simplifier ("SIGmA on
// host", src/cpu/simplify.cpp:115-268 and the headers it includes).  It is
// the checker for the HIP simplifier in parafrost_amd/: only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
// product library never links or calls it.
//
// Two tie modes for the occurrence-list sort (src/cpu/simplify.cpp:81-93):
//   REF  -- insertion sort for <= 20 entries, std::sort above
//           (src/cpu/sort.h:97-109) with CNF_CMP_KEY (simplify.h:29-42),
//           which reproduces the compiled reference bit for bit;
//   DET  -- the deterministic total order CNF_CMP_KEY then clause index,
//           which the GPU path implements.
// Everything else is identical between the modes.
//
// Pinning: checked bit-for-bit against oracle/_ref/pfref (the reference
// CPU simplifier compiled from its own sources) on the committed golden
// fixtures and on seeded random formulas of every generator family;
// the gate-free pieces (merge, tautology, subsumption, blocked clause,
// resolvent counting) are additionally checked against brute force in
// tests/test_oracle_bruteforce.py.

#include <algorithm>
#include <cassert>
#include <chrono>
#include <climits>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

namespace {

typedef uint32_t u32;
typedef int64_t i64;

// clause status (src/cpu/constants.h:69-71)
constexpr uint8_t ORIGINAL = 0, LEARNT = 1, DELETED = 2;
// variable states (src/cpu/constants.h:59-61)
constexpr uint8_t MELTED_M = 1, FROZEN_M = 2, SUBSTITUTED_M = 4;
constexpr uint8_t USAGET3 = 1;

inline u32 ABS(u32 l) { return l >> 1; }
inline u32 FLIP(u32 l) { return l ^ 1u; }
inline u32 NEG(u32 l) { return l | 1u; }
inline u32 SIGN(u32 l) { return l & 1u; }
inline u32 V2L(u32 v) { return v << 1; }
inline u32 MAPHASH(u32 l) { return 1u << (l & 31u); }

struct Unsat {};

struct Opts {
  int32_t phases, mu_pos, mu_neg, lits_min, ve_clause_limit, xor_max_arity;
  int32_t sub_limit, bce_limit, ere_limit, ere_max_resolvent, lcve_min, shrink_rate, lbd_tier1;
  int32_t ve_en, ve_plus_en, sub_en, bce_en, ere_en, all_en, ve_lbound_en;
  int32_t det;
  int32_t proof_en;
  // GPU-build election limits (electionsmax / lcveclausemax, src/gpu/lcve.cu:43-52, 94)
  // and the aggressive output sort (src/cpu/simplify.cpp:263); INT32_MAX / 0: off
  int32_t lcve_max_occurs, lcve_clause_max, aggr_sort;
};

// Clause record restating SCLAUSE (src/cpu/sclause.h:29-160); literals live
// in one pool at [off, off + size).
struct SC {
  uint8_t st = ORIGINAL, molten = 0, added = 0, usage = 0;
  u32 lbd = 0, sig = 0;
  int size = 0;
  u32 off = 0;
};

typedef std::vector<u32> OL;  // occurrence list of clause ids

struct Oracle {
  Opts opts{};
  u32 maxVar = 0;
  std::vector<SC> cls;
  std::vector<u32> pool;
  std::vector<u32> scnf;
  std::vector<OL> ot;
  std::vector<int8_t> value;  // per literal: -1 unassigned, 0 false, 1 true
  std::vector<uint8_t> vstate, frozen, assumed;
  std::vector<u32> vorg;
  std::vector<u32> trail;
  size_t propagated = 0;
  std::vector<u32> resolved;
  std::vector<u32> PVs, eligible, scores, frozenStack;
  std::vector<u32> ops, ons;  // occurs[v].ps / .ns
  int phase = 0;
  u32 mu_inc = 0;
  u32 unassigned = 0;
  u32 nClauses = 0, nLiterals = 0, n_cls_after = 0, n_lits_after = 0;
  std::vector<u32> elected_per_round;
  // per-round diagnostics
  std::vector<u32> round_units;

  // ------------------------------------------------------------ clause access
  u32* L(u32 c) { return pool.data() + cls[c].off; }
  u32 lit(u32 c, int i) { return pool[cls[c].off + i]; }
  bool original(u32 c) const { return cls[c].st == ORIGINAL; }
  bool learnt(u32 c) const { return cls[c].st & LEARNT; }
  bool deleted(u32 c) const { return cls[c].st & DELETED; }
  void calcSig(u32 c) {  // sclause.h:121-125
    u32 s = 0;
    for (int i = 0; i < cls[c].size; i++) s |= MAPHASH(lit(c, i));
    cls[c].sig = s;
  }
  u32 back(u32 c) { return lit(c, cls[c].size - 1); }
  bool has(u32 c, u32 l) {  // sclause.h:126-145 on a sorted clause: membership
    for (int i = 0; i < cls[c].size; i++)
      if (lit(c, i) == l) return true;
    return false;
  }

  // ------------------------------------------------------------ trail / model
  // ------------------------------------------------------------ DRAT proof
  // Binary DRAT as src/cpu/proof.cpp:94-175 writes it: 'a' or 'd', each
  // literal as 2*vorg(v)+sign in 7-bit groups (low first, bit 7 set on all
  // but the last byte), then a 0 byte.
  std::vector<uint8_t> proof;
  void proofLits(const u32* l, int n) {
    for (int i = 0; i < n; i++) {
      u32 r = (vorg[ABS(l[i])] << 1) | SIGN(l[i]);
      while (r > 127) { proof.push_back((uint8_t)((r & 127) | 128)); r >>= 7; }
      proof.push_back((uint8_t)r);
    }
    proof.push_back(0);
  }
  void proofAdd(const u32* l, int n) { proof.push_back('a'); proofLits(l, n); }
  void proofDelete(u32 c) {  // removeClause -> PROOF::deleteClause (sclause.cpp:77-84)
    if (!opts.proof_en) return;
    proof.push_back('d');
    proofLits(L(c), cls[c].size);
  }
  void proofResolvent(const u32* l, int n) {  // PROOF::addResolvent (proof.cpp:214-229)
    if (!opts.proof_en) return;
    for (int i = 0; i < n; i++)
      if (value[l[i]] > 0) return;  // satisfied: no line
    proofAdd(l, n);
  }

  // ------------------------------------------------------------ trail / model
  void learnEmpty() {  // solver.h:251-254
    if (opts.proof_en) { proof.push_back('a'); proof.push_back(0); }
    throw Unsat();
  }
  void enqueueUnit(u32 l) {  // solver.h:430-441 (learnUnit marks FROZEN, solver.h:475-479)
    if (opts.proof_en) proofAdd(&l, 1);
    vstate[ABS(l)] = FROZEN_M;
    value[l] = 1;
    value[FLIP(l)] = 0;
    trail.push_back(l);
    unassigned--;
  }
  void markEliminated(u32 v) {  // solver.h:535-541
    vstate[v] = MELTED_M;
    unassigned--;
  }
  u32 orgLit(u32 l) { return V2L(vorg[ABS(l)]) | SIGN(l); }
  void saveWitness(u32 w) {  // model.cpp:261-269
    resolved.push_back(orgLit(w));
    resolved.push_back(1);
  }
  void saveClause(u32 c, u32 witlit) {  // model.cpp:289-305
    const int size = cls[c].size;
    const size_t last = resolved.size();
    int pos = -1;
    for (int i = 0; i < size; i++) {
      const u32 l = lit(c, i);
      resolved.push_back(orgLit(l));
      if (l == witlit) pos = i;
    }
    if (pos) std::swap(resolved[pos + last], resolved[last]);
    resolved.push_back((u32)size);
  }

  // ------------------------------------------------------------ helpers (simplify.h)
  static bool subSig(u32 A, u32 B) { return !(A & ~B); }  // simplify.h:84
  static bool selfsubSig(u32 A, u32 B) {                  // simplify.h:86-90
    const u32 Bt = B | ((B & 0xAAAAAAAAu) >> 1) | ((B & 0x55555555u) << 1);
    return !(A & ~Bt);
  }
  bool isEqual(u32 c, const std::vector<u32>& o) {  // simplify.h:92-103
    for (size_t i = 0; i < o.size(); i++)
      if (lit(c, (int)i) != o[i]) return false;
    return true;
  }
  bool isTautology(u32 x, u32 c1, u32 c2) {  // simplify.h:119-137
    const int n1 = cls[c1].size, n2 = cls[c2].size;
    int i1 = 0, i2 = 0;
    while (i1 < n1 && i2 < n2) {
      const u32 l1 = lit(c1, i1), l2 = lit(c2, i2);
      const u32 v1 = ABS(l1), v2 = ABS(l2);
      if (v1 == x) i1++;
      else if (v2 == x) i2++;
      else if ((l1 ^ l2) == 1u) return true;
      else if (v1 < v2) i1++;
      else if (v2 < v1) i2++;
      else { i1++; i2++; }
    }
    return false;
  }
  bool mergeOut(u32 x, u32 c1, u32 c2, std::vector<u32>& out) {  // simplify.h:139-173
    out.clear();
    const int n1 = cls[c1].size, n2 = cls[c2].size;
    int i1 = 0, i2 = 0;
    while (i1 < n1 && i2 < n2) {
      const u32 l1 = lit(c1, i1), l2 = lit(c2, i2);
      const u32 v1 = ABS(l1), v2 = ABS(l2);
      if (v1 == x) i1++;
      else if (v2 == x) i2++;
      else if ((l1 ^ l2) == 1u) return false;
      else if (v1 < v2) { i1++; out.push_back(l1); }
      else if (v2 < v1) { i2++; out.push_back(l2); }
      else { i1++; i2++; out.push_back(l1); }
    }
    while (i1 < n1) { const u32 l = lit(c1, i1++); if (ABS(l) != x) out.push_back(l); }
    while (i2 < n2) { const u32 l = lit(c2, i2++); if (ABS(l) != x) out.push_back(l); }
    return true;
  }
  int mergeLen(u32 x, u32 c1, u32 c2) {  // simplify.h:175-202
    const int n1 = cls[c1].size, n2 = cls[c2].size;
    int i1 = 0, i2 = 0, len = n1 + n2 - 2;
    while (i1 < n1 && i2 < n2) {
      const u32 l1 = lit(c1, i1), l2 = lit(c2, i2);
      const u32 v1 = ABS(l1), v2 = ABS(l2);
      if (v1 == x) i1++;
      else if (v2 == x) i2++;
      else if ((l1 ^ l2) == 1u) return 0;
      else if (v1 < v2) i1++;
      else if (v2 < v1) i2++;
      else { i1++; i2++; len--; }
    }
    return len;
  }
  void freezeBinaries(OL& l) {  // simplify.h:204-211
    for (u32 c : l)
      if (original(c) && cls[c].size == 2) cls[c].molten = 0;
  }
  int countOrgs(OL& l) {  // simplify.h:221-227
    int n = 0;
    for (u32 c : l) n += original(c);
    return n;
  }
  int countLitsBefore(OL& l) {  // simplify.h:229-235
    int n = 0;
    for (u32 c : l)
      if (original(c)) n += cls[c].size;
    return n;
  }
  // simplify.h:237-266 (substitute == true) and 268-295 (false)
  bool countRes(u32 x, int clsbefore, OL& me, OL& other, int& nAddedCls, bool substitute) {
    int nAddedLits = 0;
    const int rlimit = opts.ve_clause_limit;
    for (u32 ci : me) {
      if (!original(ci)) continue;
      const bool a = cls[ci].molten;
      for (u32 cj : other) {
        if (!original(cj)) continue;
        const bool b = cls[cj].molten;
        if (substitute && a == b) continue;
        int rsize;
        if ((rsize = mergeLen(x, ci, cj)) > 1) {
          if (++nAddedCls > clsbefore || (rlimit && rsize > rlimit)) return true;
          nAddedLits += rsize;
        }
      }
    }
    if (opts.ve_lbound_en) {
      int before = countLitsBefore(me) + countLitsBefore(other);
      if (nAddedLits > before) return true;
    }
    return false;
  }
  void toblivion(OL& l) {  // simplify.h:297-303
    for (u32 c : l) cls[c].st = DELETED;
    OL().swap(l);
  }
  void toblivion(u32 p, int pOrgs, int nOrgs, OL& poss, OL& negs) {  // simplify.h:305-332
    const u32 n = NEG(p);
    if (pOrgs > nOrgs) {
      for (u32 c : negs) {
        if (original(c)) saveClause(c, n);
        cls[c].st = DELETED;
      }
      saveWitness(p);
      OL().swap(negs);
      toblivion(poss);
    } else {
      for (u32 c : poss) {
        if (original(c)) saveClause(c, p);
        cls[c].st = DELETED;
      }
      saveWitness(n);
      OL().swap(poss);
      toblivion(negs);
    }
  }

  // ------------------------------------------------------------ OT (simplify.cpp)
  void createOT() {  // simplify.cpp:35-56 (reset = true)
    for (u32 v = 1; v <= maxVar; v++) { ot[V2L(v)].clear(); ot[NEG(V2L(v))].clear(); }
    for (u32 c : scnf) {
      if (learnt(c) || original(c))
        for (int k = 0; k < cls[c].size; k++) ot[lit(c, k)].push_back(c);
    }
  }
  void reduceOL(OL& l) {  // simplify.cpp:58-68
    size_t j = 0;
    for (size_t i = 0; i < l.size(); i++)
      if (!deleted(l[i])) l[j++] = l[i];
    l.resize(j);
  }
  void reduceOT() {  // simplify.cpp:70-79
    for (u32 v = 1; v <= maxVar; v++) { reduceOL(ot[V2L(v)]); reduceOL(ot[NEG(V2L(v))]); }
  }
  // CNF_CMP_KEY (simplify.h:29-42)
  bool keyLess(u32 x, u32 y) {
    const int xs = cls[x].size, ys = cls[y].size;
    if (xs < ys) return true;
    if (xs > ys) return false;
    const u32 a = lit(x, 0), b = lit(y, 0);
    if (a < b) return true;
    if (a > b) return false;
    const u32 xb = back(x), yb = back(y);
    if (xb < yb) return true;
    if (xb > yb) return false;
    return cls[x].sig < cls[y].sig;
  }
  void sortList(OL& l) {
    if (l.size() < 2) return;
    if (opts.det) {
      std::sort(l.begin(), l.end(), [&](u32 a, u32 b) {
        if (keyLess(a, b)) return true;
        if (keyLess(b, a)) return false;
        return a < b;
      });
      return;
    }
    auto cmp = [&](u32 a, u32 b) { return keyLess(a, b); };
    if (l.size() <= 20) {  // sort.h:43-57 insertion sort
      for (size_t i = 1; i < l.size(); i++) {
        const u32 t = l[i];
        size_t j;
        for (j = i; j > 0 && cmp(t, l[j - 1]); j--) l[j] = l[j - 1];
        l[j] = t;
      }
    } else {
      std::sort(l.begin(), l.end(), cmp);
    }
  }
  void sortOT() {  // simplify.cpp:81-93
    for (u32 v : PVs) {
      OL& poss = ot[V2L(v)];
      OL& negs = ot[NEG(V2L(v))];
      if (poss.size() > 1) sortList(poss);
      if (negs.size() > 1) sortList(negs);
    }
  }
  void shrinkSimp() {  // simplify.cpp:238-249
    size_t j = 0;
    for (size_t i = 0; i < scnf.size(); i++)
      if (!deleted(scnf[i])) scnf[j++] = scnf[i];
    scnf.resize(j);
  }
  void resizeCNF() {  // solver.h:685-689
    const int times = phase + 1;
    if (opts.shrink_rate > 0 && times > 1 && times != opts.phases && (times % opts.shrink_rate) == 0)
      shrinkSimp();
  }

  // ------------------------------------------------------------ BCP (elimination.cpp:41-99)
  bool propClause(u32 l, u32 c) {
    u32 sig = 0;
    u32* d = L(c);
    int j = 0;
    for (int i = 0; i < cls[c].size; i++) {
      const u32 o = d[i];
      if (o != l) {
        if (value[o] > 0) return true;
        d[j++] = o;
        sig |= MAPHASH(o);
      }
    }
    cls[c].sig = sig;
    cls[c].size--;
    return false;
  }
  bool prop() {
    const size_t before = propagated;
    while (propagated < trail.size()) {
      const u32 assign = trail[propagated++], f = FLIP(assign);
      OL& fl = ot[f];
      for (size_t i = 0; i < fl.size(); i++) {
        const u32 c = fl[i];
        if (deleted(c)) continue;
        if (propClause(f, c)) cls[c].st = DELETED;
        else {
          const int size = cls[c].size;
          if (!size) learnEmpty();  // elimination.cpp:80 (learnEmpty; prop fails)
          if (size == 1) {
            const u32 unit = lit(c, 0);
            if (value[unit] < 0) enqueueUnit(unit);
            else learnEmpty();
          }
        }
      }
      OL().swap(ot[f]);
      toblivion(ot[assign]);
    }
    if (propagated != before && !opts.sub_en) reduceOT();
    return true;
  }

  // ------------------------------------------------------------ LCVE (lcve.cpp)
  void histSimp() {  // histogram.cpp:82-95
    std::fill(ops.begin(), ops.end(), 0);
    std::fill(ons.begin(), ons.end(), 0);
    for (u32 c : scnf) {
      if (deleted(c)) continue;
      for (int k = 0; k < cls[c].size; k++) {
        const u32 l = lit(c, k);
        if (SIGN(l)) ons[ABS(l)]++;
        else ops[ABS(l)]++;
      }
    }
  }
  void varReorder() {  // lcve.cpp:25-44
    eligible.resize(maxVar);
    histSimp();
    for (u32 v = 1; v <= maxVar; v++) {
      eligible[v - 1] = v;
      scores[v] = ops[v] * ons[v];
    }
    // rSort with LCV_CMP / LCV_RANK: a stable sort by score of 1..maxVar
    std::stable_sort(eligible.begin(), eligible.end(),
                     [&](u32 a, u32 b) { return scores[a] < scores[b]; });
  }
  void depFreeze(OL& l, u32 cand, u32 pt, u32 nt) {  // lcve.cpp:83-99
    for (u32 c : l) {
      if (deleted(c)) continue;
      for (int k = 0; k < cls[c].size; k++) {
        const u32 v = ABS(lit(c, k));
        if (!frozen[v] && v != cand && (ops[v] < pt || ons[v] < nt)) {
          frozen[v] = 1;
          frozenStack.push_back(v);
        }
      }
    }
  }
  bool oversize(const OL& l) {
    for (u32 c : l)
      if (!deleted(c) && cls[c].size > opts.lcve_clause_max) return true;
    return false;
  }
  bool LCVE() {  // lcve.cpp:46-81
    varReorder();
    PVs.clear();
    frozenStack.clear();
    for (size_t i = 0; i < eligible.size(); i++) {
      const u32 cand = eligible[i];
      if (assumed[cand]) continue;
      if (vstate[cand]) continue;
      if (frozen[cand]) continue;
      const u32 p = V2L(cand), n = NEG(p);
      const u32 psz = (u32)ot[p].size(), nsz = (u32)ot[n].size();
      if (ops[cand] == 0 && ons[cand] == 0) continue;
      // electionsmax (the GPU build's lcve_k, lcve.cu:94): histogram counts
      const u32 maxocc = (u32)opts.lcve_max_occurs;
      if (ops[cand] > maxocc || ons[cand] > maxocc) break;
      const u32 pt = (u32)opts.mu_pos << mu_inc, nt = (u32)opts.mu_neg << mu_inc;
      if (psz >= pt && nsz >= nt) break;
      // lcveclausemax (depFreeze_d, lcve.cu:43-52): a candidate with a live
      // clause longer than the limit is passed over; its freezes are undone
      if (opts.lcve_clause_max != INT32_MAX && (oversize(ot[p]) || oversize(ot[n]))) continue;
      PVs.push_back(cand);
      depFreeze(ot[p], cand, pt, nt);
      depFreeze(ot[n], cand, pt, nt);
    }
    for (u32 v : frozenStack) frozen[v] = 0;
    elected_per_round.push_back((u32)PVs.size());
    return PVs.size() >= (size_t)opts.lcve_min;
  }

  // ------------------------------------------------------------ SUB (subsume.h)
  void updateOL(OL& l) {  // subsume.h:27-37
    size_t j = 0;
    for (size_t i = 0; i < l.size(); i++) {
      const u32 c = l[i];
      if (cls[c].molten) cls[c].molten = 0;
      else if (!deleted(c)) l[j++] = c;
    }
    l.resize(j);
  }
  bool subLits(u32 sm, u32 lr) {  // subsume.h:39-56
    int i1 = 0, i2 = 0, s = 0;
    const int n1 = cls[sm].size, n2 = cls[lr].size;
    while (i1 < n1 && i2 < n2) {
      const u32 a = lit(sm, i1), b = lit(lr, i2);
      if (a < b) i1++;
      else if (b < a) i2++;
      else { s++; i1++; i2++; }
    }
    return s == n1;
  }
  bool selfsubLits(u32 x, u32 fx, u32 sm, u32 lr) {  // subsume.h:75-103
    int i1 = 0, i2 = 0, s = 0;
    const int n1 = cls[sm].size, n2 = cls[lr].size;
    bool self = false;
    while (i1 < n1 && i2 < n2) {
      const u32 a = lit(sm, i1), b = lit(lr, i2);
      if (a == fx) i1++;
      else if (b == x) { self = true; i2++; }
      else if (a < b) i1++;
      else if (b < a) i2++;
      else { s++; i1++; i2++; }
    }
    if (s + 1 == n1) {
      if (self) return true;
      while (i2 < n2) {
        if (lit(lr, i2) == x) return true;
        i2++;
      }
    }
    return false;
  }
  void bumpShrunken(u32 c) {  // elimination.cpp:28-39
    const int old = (int)cls[c].lbd;
    if (old <= opts.lbd_tier1) return;
    const int nl = std::min(cls[c].size - 1, old);
    if (nl >= old) return;
    cls[c].lbd = (u32)nl;
    cls[c].usage = USAGET3;
  }
  void strengthen(u32 c, u32 me) {  // elimination.cpp:199-234
    u32 sig = 0;
    int n = 0;
    u32* d = L(c);
    for (int k = 0; k < cls[c].size; k++) {
      const u32 l = d[k];
      if (l != me) { d[n++] = l; sig |= MAPHASH(l); }
    }
    cls[c].sig = sig;
    cls[c].size--;
    if (n == 1) {
      const u32 unit = d[0];
      const int8_t val = value[unit];
      if (val < 0) {
        enqueueUnit(unit);
        toblivion(ot[unit]);
      } else if (!val) {
        learnEmpty();
      }
    } else {
      proofResolvent(d, n);
      if (learnt(c)) bumpShrunken(c);
    }
  }
  bool subsume(OL& l, size_t end, u32 cand) {  // subsume.h:105-121
    const int csz = cls[cand].size;
    for (size_t j = 0; j < end; j++) {
      const u32 s = l[j];
      if (deleted(s)) continue;
      if (cls[cand].molten && cls[s].size > csz) continue;
      if (cls[s].size > 1 && subSig(cls[s].sig, cls[cand].sig) && subLits(s, cand)) {
        if (learnt(s) && original(cand)) cls[s].st = ORIGINAL;
        proofDelete(cand);
        cls[cand].st = DELETED;  // removeClause (sclause.cpp:77-84)
        return true;
      }
    }
    return false;
  }
  bool selfsubsume(u32 x, u32 fx, OL& l, u32 cand) {  // subsume.h:123-142
    const int csz = cls[cand].size;
    const u32 csig = cls[cand].sig;
    for (size_t j = 0; j < l.size(); j++) {
      const u32 s = l[j];
      const int ssz = cls[s].size;
      if (ssz > csz) break;
      if (deleted(s)) continue;
      if (ssz > 1 && selfsubSig(cls[s].sig, csig) && selfsubLits(x, fx, s, cand)) {
        strengthen(cand, x);
        cls[cand].molten = 1;
        return true;
      }
    }
    return false;
  }
  void self_sub_x(u32 p, OL& poss, OL& negs) {  // subsume.h:144-179
    const u32 n = NEG(p);
    for (size_t i = 0; i < poss.size(); i++) {
      const u32 c = poss[i];
      if (cls[c].size > 1000) break;
      if (deleted(c)) continue;
      selfsubsume(p, n, negs, c);
      subsume(poss, i, c);
    }
    updateOL(poss);
    for (size_t i = 0; i < negs.size(); i++) {
      const u32 c = negs[i];
      if (cls[c].size > 1000) break;
      if (deleted(c)) continue;
      selfsubsume(n, p, poss, c);
      subsume(negs, i, c);
    }
    updateOL(negs);
  }
  void SUB() {  // elimination.cpp:110-129
    if (!(opts.sub_en || opts.ve_plus_en)) return;
    for (u32 v : PVs) {
      const u32 p = V2L(v), n = NEG(p);
      OL& poss = ot[p];
      OL& negs = ot[n];
      if ((int)poss.size() <= opts.sub_limit && (int)negs.size() <= opts.sub_limit)
        self_sub_x(p, poss, negs);
    }
  }

  // ------------------------------------------------------------ gates
  // equivalence.h:25-43
  u32 substitute_single(u32 dx, u32 c, u32 def) {
    u32* d = L(c);
    int n = 0;
    for (int i = 0; i < cls[c].size; i++) {
      const u32 l = d[i];
      if (l == dx) d[n++] = def;
      else if (l != def) d[n++] = l;
    }
    cls[c].size = n;
    if (n == 1) return d[0];
    std::sort(d, d + n);  // Sort(.., LESS): distinct values, order unique
    calcSig(c);
    return 0;
  }
  // equivalence.h:45-87
  bool substitute_single(u32 p, u32 def) {
    const u32 n = NEG(p), def_f = FLIP(def);
    OL& poss = ot[p];
    OL& negs = ot[n];
    for (u32 c : negs) {
      if (learnt(c) || cls[c].molten || has(c, def)) cls[c].st = DELETED;
      else if (original(c)) {
        const u32 unit = substitute_single(n, c, def_f);
        if (unit) {
          const int8_t val = value[unit];
          if (val < 0) enqueueUnit(unit);
          else if (!val) return true;
        } else {
          proofResolvent(L(c), cls[c].size);
        }
      }
    }
    for (u32 c : poss) {
      if (learnt(c) || cls[c].molten || has(c, def_f)) cls[c].st = DELETED;
      else if (original(c)) {
        const u32 unit = substitute_single(p, c, def);
        if (unit) {
          const int8_t val = value[unit];
          if (val < 0) enqueueUnit(unit);
          else if (!val) return true;
        } else {
          proofResolvent(L(c), cls[c].size);
        }
      }
    }
    return false;
  }
  u32 find_sfanin(u32 gate_out, OL& l) {  // equivalence.h:89-107
    u32 imp = 0;
    int nImps = 0;
    for (u32 c : l) {
      if (original(c) && cls[c].size == 2) {
        imp = FLIP(lit(c, 0) ^ lit(c, 1) ^ gate_out);
        cls[c].molten = 1;
        nImps++;
      }
      if (nImps > 1) return 0;
    }
    return imp;
  }
  u32 find_BN_gate(u32 p, OL& poss, OL& negs) {  // equivalence.h:109-131
    const u32 n = NEG(p);
    u32 first = find_sfanin(p, poss);
    if (first) {
      u32 second = n, def = first;
      if (second < first) first = second, second = def;
      for (u32 c : negs) {
        if (original(c) && cls[c].size == 2 && lit(c, 0) == first && lit(c, 1) == second) {
          cls[c].molten = 1;
          return def;
        }
      }
    }
    freezeBinaries(poss);
    return 0;
  }
  void save_BN_gate(u32 p, int pOrgs, int nOrgs, OL& poss, OL& negs) {  // equivalence.h:133-154
    const u32 n = NEG(p);
    if (pOrgs > nOrgs) {
      for (u32 c : negs)
        if (original(c)) saveClause(c, n);
      saveWitness(p);
    } else {
      for (u32 c : poss)
        if (original(c)) saveClause(c, p);
      saveWitness(n);
    }
  }
  // and.h:25-86
  void find_fanin(u32 gate_out, OL& l, std::vector<u32>& out, u32& sig) {
    out.clear();
    sig = 0;
    for (u32 c : l) {
      if (original(c) && cls[c].size == 2) {
        const u32 imp = FLIP(lit(c, 0) ^ lit(c, 1) ^ gate_out);
        out.push_back(imp);
        sig |= MAPHASH(imp);
        cls[c].molten = 1;
      }
    }
  }
  bool find_AO_gate(u32 dx, int nOrgCls, std::vector<u32>& out, int& nAddedCls) {
    out.clear();
    u32 sig;
    const u32 x = ABS(dx);
    OL& it = ot[dx];
    find_fanin(dx, it, out, sig);
    if (out.size() > 1) {
      const u32 f = FLIP(dx);
      out.push_back(f);
      sig |= MAPHASH(f);
      std::sort(out.begin(), out.end());
      OL& ot2 = ot[f];
      for (u32 c : ot2) {
        if (original(c) && cls[c].size == (int)out.size() && subSig(cls[c].sig, sig) && isEqual(c, out)) {
          cls[c].molten = 1;
          nAddedCls = 0;
          if (countRes(x, nOrgCls, it, ot2, nAddedCls, true)) {
            cls[c].molten = 0;
            break;
          }
          return true;
        }
      }
    }
    freezeBinaries(it);
    return false;
  }
  // ifthenelse.h:25-98
  u32 fast_equality_check(u32 x, u32 y, u32 z) {  // returns clause id + 1, 0 if none
    if (ot[y].size() > ot[z].size()) std::swap(y, z);
    if (ot[x].size() > ot[y].size()) std::swap(x, y);
    OL& l = ot[x];
    u32 a[3] = {x, y, z};
    std::sort(a, a + 3);  // sort3
    for (u32 c : l) {
      if (cls[c].molten) continue;
      if (original(c) && cls[c].size == 3 && lit(c, 0) == a[0] && lit(c, 1) == a[1] && lit(c, 2) == a[2])
        return c + 1;
    }
    return 0;
  }
  bool find_ITE_gate(u32 dx, int nOrgCls, int& nAddedCls) {
    const u32 fx = FLIP(dx);
    OL& it = ot[dx];
    for (size_t i = 0; i < it.size(); i++) {
      const u32 ci = it[i];
      if (!(original(ci) && cls[ci].size == 3)) continue;
      u32 xi = lit(ci, 0), yi = lit(ci, 1), zi = lit(ci, 2);
      if (yi == dx) std::swap(xi, yi);
      if (zi == dx) std::swap(xi, zi);
      for (size_t j = i + 1; j < it.size(); j++) {
        const u32 cj = it[j];
        if (!(original(cj) && cls[cj].size == 3)) continue;
        u32 xj = lit(cj, 0), yj = lit(cj, 1), zj = lit(cj, 2);
        if (yj == dx) std::swap(xj, yj);
        if (zj == dx) std::swap(xj, zj);
        if (ABS(yi) == ABS(zj)) std::swap(yj, zj);
        if (ABS(zi) == ABS(zj)) continue;
        if (yi != FLIP(yj)) continue;
        const u32 d1 = fast_equality_check(fx, yi, FLIP(zi));
        if (!d1) continue;
        const u32 d2 = fast_equality_check(fx, yj, FLIP(zj));
        if (!d2) continue;
        cls[ci].molten = 1; cls[cj].molten = 1;
        cls[d1 - 1].molten = 1; cls[d2 - 1].molten = 1;
        const u32 v = ABS(dx);
        OL& ot2 = ot[fx];
        nAddedCls = 0;
        if (countRes(v, nOrgCls, it, ot2, nAddedCls, true)) {
          cls[ci].molten = 0; cls[cj].molten = 0;
          cls[d1 - 1].molten = 0; cls[d2 - 1].molten = 0;
          return false;
        }
        return true;
      }
    }
    return false;
  }
  // xor.h:25-150
  void freezeArities(OL& me, OL& other) {
    for (u32 c : me)
      if (cls[c].size > 2 && cls[c].molten) cls[c].molten = 0;
    for (u32 c : other)
      if (cls[c].size > 2 && cls[c].molten) cls[c].molten = 0;
  }
  bool checkArity(u32 c, const u32* ls, int size) {
    for (int i = 0; i < cls[c].size; i++) {
      const u32 l = lit(c, i);
      int j;
      for (j = 0; j < size; j++)
        if (l == ls[j]) break;
      if (j == size) return false;
    }
    return true;
  }
  bool makeArity(u32& parity, u32* ls, int size) {
    const u32 old = parity;
    while (__builtin_parity(++parity));
    for (int k = 0; k < size; k++) {
      const u32 bit = 1u << k;
      if ((parity & bit) != (old & bit)) ls[k] = FLIP(ls[k]);
    }
    u32 best = ls[0];
    int minsize = (int)ot[best].size();
    for (int k = 1; k < size; k++) {
      const int s = (int)ot[ls[k]].size();
      if (s < minsize) { minsize = s; best = ls[k]; }
    }
    for (u32 c : ot[best]) {
      if (original(c) && cls[c].size == size && checkArity(c, ls, size)) {
        cls[c].molten = 1;
        return true;
      }
    }
    return false;
  }
  bool find_XOR_gate(u32 dx, int nOrgCls, std::vector<u32>& out, int& nAddedCls) {
    const u32 fx = FLIP(dx), v = ABS(dx);
    OL& it = ot[dx];
    OL& ot2 = ot[fx];
    for (size_t i = 0; i < it.size(); i++) {
      const u32 ci = it[i];
      if (!original(ci)) continue;
      const int size = cls[ci].size;
      const int arity = size - 1;
      if (size < 3 || arity > opts.xor_max_arity) continue;
      out.assign(L(ci), L(ci) + size);
      u32 parity = 0;
      int itargets = 1 << arity;
      while (--itargets && makeArity(parity, out.data(), size));
      if (itargets) freezeArities(it, ot2);
      else {
        cls[ci].molten = 1;
        nAddedCls = 0;
        if (countRes(v, nOrgCls, it, ot2, nAddedCls, true)) {
          freezeArities(it, ot2);
          break;
        }
        return true;
      }
    }
    return false;
  }

  // ------------------------------------------------------------ BVE (bve.cpp)
  void newResolvent(const std::vector<u32>& r) {  // bve.cpp:289-320
    const int size = (int)r.size();
    if (size == 1) {
      const u32 unit = r[0];
      const int8_t val = value[unit];
      if (val < 0) enqueueUnit(unit);
      else if (!val) learnEmpty();
      return;
    }
    proofResolvent(r.data(), size);
    SC s;
    s.st = ORIGINAL;
    s.size = size;
    s.off = (u32)pool.size();
    pool.insert(pool.end(), r.begin(), r.end());
    cls.push_back(s);
    const u32 c = (u32)cls.size() - 1;
    calcSig(c);
    cls[c].added = 1;
    scnf.push_back(c);
  }
  void xresolve(u32 x, std::vector<u32>& out, bool substitute) {  // bve.cpp:245-287
    u32 dx = V2L(x), fx = NEG(dx);
    if (ot[dx].size() > ot[fx].size()) std::swap(dx, fx);
    // copy the lists: newResolvent may grow 'cls' but never the lists
    OL& me = ot[dx];
    OL& other = ot[fx];
    for (size_t i = 0; i < me.size(); i++) {
      const u32 ci = me[i];
      if (!original(ci)) continue;
      const bool a = cls[ci].molten;
      for (size_t j = 0; j < other.size(); j++) {
        const u32 cj = other[j];
        if (!original(cj)) continue;
        if (substitute && a == (bool)cls[cj].molten) continue;
        if (mergeOut(x, ci, cj, out)) newResolvent(out);
      }
    }
  }
  void bve() {  // bve.cpp:127-243
    std::vector<u32> out;
    for (size_t i = 0; i < PVs.size(); i++) {
      u32 v = PVs[i];
      const u32 p = V2L(v), n = NEG(p);
      OL& poss = ot[p];
      OL& negs = ot[n];
      const int pOrgs = countOrgs(poss), nOrgs = countOrgs(negs);
      if (!pOrgs || !nOrgs) {
        toblivion(p, pOrgs, nOrgs, poss, negs);
        v = 0;
      } else if (u32 def = find_BN_gate(p, poss, negs)) {
        save_BN_gate(p, pOrgs, nOrgs, poss, negs);
        if (substitute_single(p, def)) learnEmpty();
        v = 0;
      } else {
        out.clear();
        const int nOrgCls = pOrgs + nOrgs;
        int nAddedCls = 0;
        int type = 0;  // 1 resolution, 2 substitution
        if (nOrgCls > 2) {
          if (find_AO_gate(n, nOrgCls, out, nAddedCls)) type = 2;
          else if (!nAddedCls && find_AO_gate(p, nOrgCls, out, nAddedCls)) type = 2;
        }
        if (!type && nOrgCls > 3) {
          if (find_ITE_gate(p, nOrgCls, nAddedCls)) type = 2;
          else if (!nAddedCls && find_ITE_gate(n, nOrgCls, nAddedCls)) type = 2;
          else if (find_XOR_gate(p, nOrgCls, out, nAddedCls)) type = 2;
          else if (!nAddedCls && find_XOR_gate(n, nOrgCls, out, nAddedCls)) type = 2;
        }
        if (!type && !nAddedCls && !countRes(v, nOrgCls, poss, negs, nAddedCls, false)) type = 1;
        if (type == 2) {
          if (nAddedCls) xresolve(v, out, true);
          toblivion(p, pOrgs, nOrgs, ot[p], ot[n]);
          v = 0;
        } else if (type == 1) {
          if (nAddedCls) xresolve(v, out, false);
          toblivion(p, pOrgs, nOrgs, ot[p], ot[n]);
          v = 0;
        }
      }
      if (!v) {
        markEliminated(PVs[i]);
        PVs[i] = 0;
      }
    }
  }
  void VE() {
    if (opts.ve_en) bve();
  }

  // ------------------------------------------------------------ BCE (elimination.cpp:131-167)
  void BCE() {
    if (!opts.bce_en) return;
    for (u32 v : PVs) {
      if (!v) continue;
      const u32 p = V2L(v), n = NEG(p);
      OL& poss = ot[p];
      OL& negs = ot[n];
      if ((int)poss.size() <= opts.bce_limit && (int)negs.size() <= opts.bce_limit) {
        for (u32 ng : negs) {
          if (!original(ng)) continue;
          bool all = true;
          for (u32 ps : poss) {
            if (original(ps) && !isTautology(v, ng, ps)) { all = false; break; }
          }
          if (all) {
            saveClause(ng, n);
            cls[ng].st = DELETED;
          }
        }
      }
    }
  }

  // ------------------------------------------------------------ ERE (redundancy.h, elimination.cpp:169-197)
  bool merge_ere(u32 x, u32 c1, u32 c2, u32 t) {  // redundancy.h:218-265 (file lines 27-61)
    const int n1 = cls[c1].size, n2 = cls[c2].size, n = cls[t].size;
    int i1 = 0, i2 = 0, it = 0;
    while (it < n && i1 < n1 && i2 < n2) {
      const u32 l = lit(t, it);
      if (ABS(l) == x) return false;
      const u32 l1 = lit(c1, i1), l2 = lit(c2, i2);
      const u32 v1 = ABS(l1), v2 = ABS(l2);
      if (v1 == x) i1++;
      else if (v2 == x) i2++;
      else if (v1 < v2) { if (l1 != l) return false; i1++, it++; }
      else if (v2 < v1) { if (l2 != l) return false; i2++, it++; }
      else { if (l1 != l) return false; i1++, i2++, it++; }
    }
    while (it < n && i1 < n1) {
      const u32 l1 = lit(c1, i1);
      if (ABS(l1) != x && l1 != lit(t, it++)) return false;
      i1++;
    }
    while (it < n && i2 < n2) {
      const u32 l2 = lit(c2, i2);
      if (ABS(l2) != x && l2 != lit(t, it++)) return false;
      i2++;
    }
    return true;
  }
  void forward_equ(u32 x, u32 c1, u32 c2, int maxsize) {  // redundancy.h:67-121
    const int n1 = cls[c1].size, n2 = cls[c2].size;
    int len = n1 + n2 - 2;
    int i1 = 0, i2 = 0;
    int minsize = INT_MAX;
    u32 msig = 0, best = 0;
    while (i1 < n1 && i2 < n2) {
      const u32 l1 = lit(c1, i1), l2 = lit(c2, i2);
      const u32 v1 = ABS(l1), v2 = ABS(l2);
      if (v1 == x) i1++;
      else if (v2 == x) i2++;
      else if ((l1 ^ l2) == 1u) return;
      else if (v1 < v2) {
        i1++;
        const int s = (int)ot[l1].size();
        if (s < minsize) minsize = s, best = l1;
        msig |= MAPHASH(l1);
      } else if (v2 < v1) {
        i2++;
        const int s = (int)ot[l2].size();
        if (s < minsize) minsize = s, best = l2;
        msig |= MAPHASH(l2);
      } else {
        i1++, i2++, len--;
        const int s = (int)ot[l1].size();
        if (s < minsize) minsize = s, best = l1;
        msig |= MAPHASH(l1);
      }
    }
    if (len > 1 && (!maxsize || len <= maxsize)) {
      const uint8_t type = (learnt(c1) || learnt(c2)) ? LEARNT : ORIGINAL;
      OL& ml = ot[best];
      for (int i = 0; i < minsize; i++) {
        const u32 c = ml[i];
        if (len == cls[c].size && (learnt(c) || cls[c].st == type) && subSig(msig, cls[c].sig) &&
            merge_ere(x, c1, c2, c)) {
          proofDelete(c);
          cls[c].st = DELETED;
          break;
        }
      }
    }
  }
  void ERE() {
    if (!opts.ere_en) return;
    const int maxsize = opts.ere_max_resolvent;
    for (size_t k = 0; k < PVs.size(); k++) {
      const u32 x = PVs[k];
      u32 dx = V2L(x), fx = NEG(dx);
      if (ot[dx].size() > ot[fx].size()) std::swap(dx, fx);
      OL& me = ot[dx];
      OL& other = ot[fx];
      if ((int)me.size() <= opts.ere_limit && (int)other.size() <= opts.ere_limit) {
        for (size_t i = 0; i < me.size(); i++) {
          const u32 ci = me[i];
          if (deleted(ci)) continue;
          for (size_t j = 0; j < other.size(); j++) {
            const u32 cj = other[j];
            if (deleted(cj)) continue;
            forward_equ(x, ci, cj, maxsize);
          }
        }
      }
    }
  }

  // ------------------------------------------------------------ driver
  void countAll() {  // solver.h:725-735
    n_cls_after = n_lits_after = 0;
    for (u32 c : scnf)
      if (original(c) || learnt(c)) { n_cls_after++; n_lits_after += cls[c].size; }
  }
  bool stop(i64 lr) { return (phase == opts.phases) || (lr <= opts.lits_min && phase > 2); }
  void filterPVs() {  // solver.h:698-708
    size_t n = 0;
    for (u32 x : PVs)
      if (x && !vstate[x]) PVs[n++] = x;
    PVs.resize(n);
  }

  // returns 0 UNSAT, 1 SAT, 2 UNSOLVED (simplify.cpp:157-226)
  int run() {
    try {
      if (scnf.empty()) return 2;
      phase = 0;
      mu_inc = 0;
      nClauses = (u32)scnf.size();
      nLiterals = 0;
      for (u32 c : scnf) nLiterals += cls[c].size;
      i64 litsbefore = nLiterals, diff = INT64_MAX;
      while (nClauses && nLiterals) {
        resizeCNF();
        createOT();
        if (!prop()) return 0;
        if (!LCVE()) break;
        sortOT();
        if (stop(diff)) { ERE(); break; }
        SUB();
        VE();
        BCE();
        countAll();
        filterPVs();
        nClauses = n_cls_after, nLiterals = n_lits_after;
        diff = litsbefore - (i64)nLiterals, litsbefore = nLiterals;
        phase++, mu_inc++;
        mu_inc += phase == opts.phases;
      }
      if (!prop()) return 0;
      countAll();
      nClauses = n_cls_after, nLiterals = n_lits_after;
      shrinkSimp();
      if (!unassigned || !nClauses) return 1;
      // aggresivesort (newBeginning, simplify.cpp:263): stable by CNF_CMP_KEY
      if (opts.aggr_sort)
        std::stable_sort(scnf.begin(), scnf.end(), [&](u32 x, u32 y) { return keyLess(x, y); });
      return 2;
    } catch (Unsat&) {
      return 0;
    }
  }
};

struct Result {
  int status = 2;
  u32 maxvar = 0;
  std::vector<u32> sizes, lbds, lits, units, resolved;
  std::vector<uint8_t> flags, vstate;
  std::vector<u32> elected;
  std::vector<uint8_t> proof;
  double seconds = 0;
};

}  // namespace

extern "C" {

typedef struct {
  int32_t phases, mu_pos, mu_neg, lits_min, ve_clause_limit, xor_max_arity;
  int32_t sub_limit, bce_limit, ere_limit, ere_max_resolvent, lcve_min, shrink_rate, lbd_tier1;
  int32_t ve_en, ve_plus_en, sub_en, bce_en, ere_en, all_en, ve_lbound_en;
  int32_t det;
  int32_t proof_en;
  int32_t lcve_max_occurs, lcve_clause_max, aggr_sort;
} oracle_opts;

// Runs the restated simplifier.  Input: maxvar, clauses as offsets/lits
// (2v+s, any literal order), per-clause flags (bit0 learnt, bits 4-5 usage)
// and lbd (nullable), per-var vstate (nullable), assumed flags (nullable),
// vorg (nullable = identity), the number of root-assigned variables.
// Returns an opaque handle; read it with oracle_get_* and free it.
// Incremental form: 'trail' holds the root-assigned literals, of which the
// first 'npropagated' are propagated already (their clauses are gone) and the
// rest still have to be (the next prop() does it, as after the reference's
// incremental clause addition, src/cpu/addinc.cpp).
void* oracle_run_trail(const oracle_opts* o, uint32_t maxvar, uint64_t ncls, const uint64_t* offsets,
                       const uint32_t* lits, const uint8_t* flags, const uint32_t* lbd,
                       const uint8_t* vstate, const uint8_t* assumed, const uint32_t* vorg,
                       const uint32_t* trail, uint32_t ntrail, uint32_t npropagated);

void* oracle_run(const oracle_opts* o, uint32_t maxvar, uint64_t ncls, const uint64_t* offsets,
                 const uint32_t* lits, const uint8_t* flags, const uint32_t* lbd,
                 const uint8_t* vstate, const uint8_t* assumed, const uint32_t* vorg,
                 uint32_t nassigned) {
  return oracle_run_trail(o, maxvar, ncls, offsets, lits, flags, lbd, vstate, assumed, vorg, nullptr, nassigned,
                          nassigned);
}

void* oracle_run_trail(const oracle_opts* o, uint32_t maxvar, uint64_t ncls, const uint64_t* offsets,
                       const uint32_t* lits, const uint8_t* flags, const uint32_t* lbd,
                       const uint8_t* vstate, const uint8_t* assumed, const uint32_t* vorg,
                       const uint32_t* trail, uint32_t ntrail, uint32_t npropagated) {
  const uint32_t nassigned = ntrail;
  Oracle* s = new Oracle();
  Result* r = new Result();
  s->opts = *reinterpret_cast<const Opts*>(o);
  if (s->opts.all_en) s->opts.ve_en = s->opts.ve_plus_en = s->opts.bce_en = s->opts.ere_en = 1;
  s->maxVar = maxvar;
  const size_t nd = 2 * (size_t)maxvar + 2;
  s->ot.resize(nd);
  s->value.assign(nd, -1);
  s->vstate.assign(maxvar + 1, 0);
  s->frozen.assign(maxvar + 1, 0);
  s->assumed.assign(maxvar + 1, 0);
  s->vorg.resize(maxvar + 1);
  s->scores.assign(maxvar + 1, 0);
  s->ops.assign(maxvar + 1, 0);
  s->ons.assign(maxvar + 1, 0);
  u32 melted = 0;
  for (u32 v = 0; v <= maxvar; v++) {
    s->vorg[v] = vorg ? vorg[v] : v;
    if (vstate) { s->vstate[v] = v ? vstate[v] : 0; melted += v && (vstate[v] & (MELTED_M | SUBSTITUTED_M)); }
    if (assumed) s->assumed[v] = assumed[v];
  }
  s->unassigned = maxvar - nassigned - melted;
  if (trail) {
    for (uint32_t i = 0; i < ntrail; i++) {
      s->value[trail[i]] = 1;
      s->value[FLIP(trail[i])] = 0;
      s->trail.push_back(trail[i]);
    }
    s->propagated = npropagated;
  }
  s->cls.resize(ncls);
  s->pool.resize(offsets[ncls]);
  s->scnf.resize(ncls);
  // extract (simplify.cpp:95-113, 148-149): originals then learnts, each
  // in input order; literals sorted; signature computed.  Clause ids follow
  // the extraction order.
  std::vector<uint64_t> order(ncls);
  for (uint64_t i = 0; i < ncls; i++) order[i] = i;
  std::stable_partition(order.begin(), order.end(),
                        [&](uint64_t i) { return !(flags && (flags[i] & 1)); });
  u32 wr = 0;
  for (uint64_t k = 0; k < ncls; k++) {
    const uint64_t i = order[k];
    SC& c = s->cls[k];
    c.off = wr;
    c.size = (int)(offsets[i + 1] - offsets[i]);
    const uint8_t f = flags ? flags[i] : 0;
    c.st = (f & 1) ? LEARNT : ORIGINAL;
    if (c.st == LEARNT) {
      c.lbd = lbd ? (lbd[i] & 0x03FFFFFFu) : 0;
      c.usage = (f >> 4) & 3;
    }
    std::copy(lits + offsets[i], lits + offsets[i + 1], s->pool.begin() + wr);
    std::sort(s->pool.begin() + wr, s->pool.begin() + wr + c.size);
    wr += (u32)c.size;
    s->calcSig((u32)k);
    s->scnf[k] = (u32)k;
  }
  const auto t0 = std::chrono::steady_clock::now();
  r->status = s->run();
  r->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  r->maxvar = maxvar;
  r->elected = s->elected_per_round;
  r->proof = s->proof;  // kept for UNSAT too (ends with the empty clause)
  if (r->status) {
    for (u32 c : s->scnf) {
      const SC& x = s->cls[c];
      r->sizes.push_back((u32)x.size);
      r->flags.push_back((uint8_t)((x.st == LEARNT ? 1 : 0) | (x.added ? 2 : 0) | ((x.usage & 3) << 4)));
      r->lbds.push_back(x.st == LEARNT ? x.lbd : 0);
      for (int k = 0; k < x.size; k++) r->lits.push_back(s->pool[x.off + k]);
    }
    r->units = s->trail;
    r->resolved = s->resolved;
    r->vstate = s->vstate;
  }
  delete s;
  return r;
}

int oracle_status(void* h) { return ((Result*)h)->status; }
double oracle_seconds(void* h) { return ((Result*)h)->seconds; }
// sizes of the result arrays: [nclauses, nlits, nunits, nresolved, nvstate, nrounds]
void oracle_sizes(void* h, uint64_t* out6) {
  Result* r = (Result*)h;
  out6[0] = r->sizes.size();
  out6[1] = r->lits.size();
  out6[2] = r->units.size();
  out6[3] = r->resolved.size();
  out6[4] = r->vstate.size();
  out6[5] = r->elected.size();
}
void oracle_get(void* h, uint32_t* sizes, uint8_t* flags, uint32_t* lbds, uint32_t* lits,
                uint32_t* units, uint32_t* resolved, uint8_t* vstate, uint32_t* elected) {
  Result* r = (Result*)h;
  if (sizes) std::copy(r->sizes.begin(), r->sizes.end(), sizes);
  if (flags) std::copy(r->flags.begin(), r->flags.end(), flags);
  if (lbds) std::copy(r->lbds.begin(), r->lbds.end(), lbds);
  if (lits) std::copy(r->lits.begin(), r->lits.end(), lits);
  if (units) std::copy(r->units.begin(), r->units.end(), units);
  if (resolved) std::copy(r->resolved.begin(), r->resolved.end(), resolved);
  if (vstate) std::copy(r->vstate.begin(), r->vstate.end(), vstate);
  if (elected) std::copy(r->elected.begin(), r->elected.end(), elected);
}
uint64_t oracle_proof_size(void* h) { return ((Result*)h)->proof.size(); }
void oracle_get_proof(void* h, uint8_t* out) {
  Result* r = (Result*)h;
  std::copy(r->proof.begin(), r->proof.end(), out);
}
void oracle_free(void* h) { delete (Result*)h; }

}  // extern "C"
