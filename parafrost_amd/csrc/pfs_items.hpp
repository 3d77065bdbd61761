// Per-item work of the simplifier stages: one call handles one clause, one
// occurrence list or one elected variable.  The HIP kernels in
// pfs_kernels.hip map work-items onto these functions.
//
// The per-variable functions restate the sequential stages of the
// reference CPU simplifier for one elected variable.  Elected variables
// share no live clause (LCVE), so each function only writes clauses in its
// own occurrence lists; the few effects that cross variables (units found
// by strengthening, ERE deletions, root units) are returned to the driver,
// which applies them in elected order.
#pragma once

#include "pfs_common.hpp"

namespace pfs {

// ---------------------------------------------------------------- atomics
PF_HD u32 at_add(u32* p, u32 v) {
#if defined(__HIP_DEVICE_COMPILE__)
  return atomicAdd(p, v);
#else
  return __atomic_fetch_add(p, v, __ATOMIC_RELAXED);
#endif
}
PF_HD u32 at_min(u32* p, u32 v) {
#if defined(__HIP_DEVICE_COMPILE__)
  return atomicMin(p, v);
#else
  u32 o = __atomic_load_n(p, __ATOMIC_RELAXED);
  while (v < o && !__atomic_compare_exchange_n(p, &o, v, true, __ATOMIC_RELAXED, __ATOMIC_RELAXED)) {
  }
  return o;
#endif
}
PF_HD u32 at_cas(u32* p, u32 cmp, u32 v) {
#if defined(__HIP_DEVICE_COMPILE__)
  return atomicCAS(p, cmp, v);
#else
  u32 e = cmp;
  __atomic_compare_exchange_n(p, &e, v, false, __ATOMIC_RELAXED, __ATOMIC_RELAXED);
  return e;
#endif
}
PF_HD u32 at_load(const u32* p) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __atomic_load_n(p, __ATOMIC_RELAXED);
#else
  return __atomic_load_n(p, __ATOMIC_RELAXED);
#endif
}
PF_HD u64 at_min64(u64* p, u64 v) {
#if defined(__HIP_DEVICE_COMPILE__)
  return atomicMin((unsigned long long*)p, (unsigned long long)v);
#else
  u64 o = __atomic_load_n(p, __ATOMIC_RELAXED);
  while (v < o && !__atomic_compare_exchange_n(p, &o, v, true, __ATOMIC_RELAXED, __ATOMIC_RELAXED)) {
  }
  return o;
#endif
}

// ---------------------------------------------------------------- clause access
struct CL {
  u32* l;
  u32 n;
};
PF_HD CL clause(const DB& db, u32 c) { return CL{clits(db, c), csize(db, c)}; }
PF_HD u8 flags(const DB& db, u32 c) { return cflags(db, c); }
PF_HD bool is_original(const DB& db, u32 c) { return st_original(cflags(db, c)); }
PF_HD bool is_learnt(const DB& db, u32 c) { return st_learnt(cflags(db, c)); }
PF_HD bool is_deleted(const DB& db, u32 c) { return st_deleted(cflags(db, c)); }
PF_HD bool is_molten(const DB& db, u32 c) { return (cflags(db, c) & F_MOLTEN) != 0; }
PF_HD void set_deleted(const DB& db, u32 c) {
  const u8 f = cflags(db, c);
  if (!st_deleted(f)) set_cflags(db, c, (u8)((f & ~ST_MASK) | ST_DELETED));
}
PF_HD void set_status(const DB& db, u32 c, u8 st) { set_cflags(db, c, (u8)((cflags(db, c) & ~ST_MASK) | st)); }
PF_HD void melt(const DB& db, u32 c) { set_cflags(db, c, (u8)(cflags(db, c) | F_MOLTEN)); }
PF_HD void freeze(const DB& db, u32 c) { set_cflags(db, c, (u8)(cflags(db, c) & ~F_MOLTEN)); }
PF_HD u32 calc_sig(const u32* l, u32 n) {
  u32 s = 0;
  for (u32 i = 0; i < n; i++) s |= MAPHASH(l[i]);
  return s;
}
// value of a literal under the root assignment: -1 unassigned, 0 false, 1 true
PF_HD int lit_value(const DB& db, u32 l) {
  const u32 a = db.assign[ABS(l)];
  if (!a) return -1;
  return (a - 1) == l ? 1 : 0;
}

// an occurrence list as the sequential simplifier sees it
struct OLV {
  u32* r;
  u32 n;
};
PF_HD OLV olist(const DB& db, u32 lit) { return OLV{db.ot_refs + db.ot_off[lit], db.ot_size[lit]}; }
PF_HD u32 cid(const DB&, u32 c) { return c; }  // clause identity (the pool handle)

// A variable's two occurrence lists copied to a private buffer (LDS on the
// device) for the searches that read them many times.  Layout: the sizes of
// the positive and the negative list, then each list's record offsets, then
// the records: pool handle, header word (size and flags), literals.  A
// clause is named by its record offset; cid() gives its pool handle, which
// orders clauses as in the pool.  Only the variable's own literals have
// lists here, and only the molten mark changes (lview_commit returns it).
struct LView {
  u32* w;
  u32 lit;  // the variable's positive literal
};
PF_HD OLV olist(const LView& v, u32 lit) {
  return lit == v.lit ? OLV{v.w + 2, v.w[0]} : OLV{v.w + 2 + v.w[0], v.w[1]};
}
PF_HD u32 cid(const LView& v, u32 c) { return v.w[c]; }
PF_HD u32 csize(const LView& v, u32 c) { return hdr_size(v.w[c + 1]); }
PF_HD u8 cflags(const LView& v, u32 c) { return hdr_flags(v.w[c + 1]); }
PF_HD void set_cflags(const LView& v, u32 c, u8 f) { v.w[c + 1] = (v.w[c + 1] & CL_MAX_SIZE) | ((u32)f << 24); }
PF_HD u32* clits(const LView& v, u32 c) { return v.w + c + 2; }
// the clause's signature (not staged: the OR of MAPHASH over its literals,
// as the pool's signature word holds it)
PF_HD u32 csig(const LView& v, u32 c) {
  u32 sig = 0;
  const u32* l = v.w + c + 2;
  for (u32 q = 0; q < hdr_size(v.w[c + 1]); q++) sig |= MAPHASH(l[q]);
  return sig;
}
PF_HD CL clause(const LView& v, u32 c) { return CL{clits(v, c), csize(v, c)}; }
PF_HD bool is_original(const LView& v, u32 c) { return st_original(cflags(v, c)); }
PF_HD bool is_molten(const LView& v, u32 c) { return (cflags(v, c) & F_MOLTEN) != 0; }
PF_HD void melt(const LView& v, u32 c) { set_cflags(v, c, (u8)(cflags(v, c) | F_MOLTEN)); }
PF_HD void freeze(const LView& v, u32 c) { set_cflags(v, c, (u8)(cflags(v, c) & ~F_MOLTEN)); }
// words the view of variable v takes; lview_bound / sview_bound bound them
// from the list lengths and a bound on the sum of the clause sizes of the
// two lists
PF_HD u32 lview_words(const DB& db, u32 v) {
  const OLV a = olist(db, V2L(v)), b = olist(db, NEGL(V2L(v)));
  u32 n = 2 + a.n + b.n;
  for (u32 i = 0; i < a.n; i++) n += 2 + csize(db, a.r[i]);
  for (u32 i = 0; i < b.n; i++) n += 2 + csize(db, b.r[i]);
  return n;
}
PF_HD u64 lview_bound(u32 nlists, u64 sizes) { return 2 + 3 * (u64)nlists + sizes; }
// copies the lists of v into w[0, cap); false (w partly written) when they
// do not fit
PF_HD bool lview_stage(const DB& db, u32 v, u32* w, u32 cap, LView& lv) {
  const u32 p = V2L(v);
  const OLV a = olist(db, p), b = olist(db, NEGL(p));
  u32 at = 2 + a.n + b.n;
  if ((u64)at > cap) return false;
  w[0] = a.n;
  w[1] = b.n;
  for (u32 s = 0; s < 2; s++) {
    const OLV ol = s ? b : a;
    u32* offs = w + 2 + (s ? a.n : 0);
    for (u32 i = 0; i < ol.n; i++) {
      const u32 c = ol.r[i];
      const u32 h = db.lits[c];
      const u32 n = hdr_size(h);
      if ((u64)at + 2 + n > cap) return false;
      const u32* l = clits(db, c);
      offs[i] = at;
      w[at] = c;
      w[at + 1] = h;
      for (u32 q = 0; q < n; q++) w[at + 2 + q] = l[q];
      at += 2 + n;
    }
  }
  lv = LView{w, p};
  return true;
}
// The same for SUB, which also changes clauses and compacts the lists:
// records hold the pool handle, the three header words and the literals;
// w[0] and w[1] are the current sizes of the positive and the negative
// list, w[2] and w[3] their first words (offset arrays), w[4] the first
// record.  sview_commit writes back what changed.
struct SView {
  u32* w;
  u32 lit;
  const u32* assign;
};
PF_HD OLV olist(const SView& v, u32 lit) {
  return lit == v.lit ? OLV{v.w + v.w[2], v.w[0]} : OLV{v.w + v.w[3], v.w[1]};
}
PF_HD u32 cid(const SView& v, u32 c) { return v.w[c]; }
PF_HD u32 csize(const SView& v, u32 c) { return hdr_size(v.w[c + 1]); }
PF_HD u8 cflags(const SView& v, u32 c) { return hdr_flags(v.w[c + 1]); }
PF_HD u32 csig(const SView& v, u32 c) { return v.w[c + 2]; }
PF_HD u32 clbd(const SView& v, u32 c) { return v.w[c + 3]; }
PF_HD u32* clits(const SView& v, u32 c) { return v.w + c + 4; }
PF_HD void set_cflags(const SView& v, u32 c, u8 f) { v.w[c + 1] = (v.w[c + 1] & CL_MAX_SIZE) | ((u32)f << 24); }
PF_HD void set_csize(const SView& v, u32 c, u32 n) { v.w[c + 1] = (v.w[c + 1] & ~CL_MAX_SIZE) | n; }
PF_HD void set_csig(const SView& v, u32 c, u32 sig) { v.w[c + 2] = sig; }
PF_HD void set_clbd(const SView& v, u32 c, u32 lbd) { v.w[c + 3] = lbd; }
PF_HD CL clause(const SView& v, u32 c) { return CL{clits(v, c), csize(v, c)}; }
PF_HD bool is_original(const SView& v, u32 c) { return st_original(cflags(v, c)); }
PF_HD bool is_learnt(const SView& v, u32 c) { return st_learnt(cflags(v, c)); }
PF_HD bool is_deleted(const SView& v, u32 c) { return st_deleted(cflags(v, c)); }
PF_HD bool is_molten(const SView& v, u32 c) { return (cflags(v, c) & F_MOLTEN) != 0; }
PF_HD void set_deleted(const SView& v, u32 c) {
  const u8 f = cflags(v, c);
  if (!st_deleted(f)) set_cflags(v, c, (u8)((f & ~ST_MASK) | ST_DELETED));
}
PF_HD void set_status(const SView& v, u32 c, u8 st) { set_cflags(v, c, (u8)((cflags(v, c) & ~ST_MASK) | st)); }
PF_HD void melt(const SView& v, u32 c) { set_cflags(v, c, (u8)(cflags(v, c) | F_MOLTEN)); }
PF_HD void freeze(const SView& v, u32 c) { set_cflags(v, c, (u8)(cflags(v, c) & ~F_MOLTEN)); }
PF_HD int lit_value(const SView& v, u32 l) {
  const u32 a = v.assign[ABS(l)];
  if (!a) return -1;
  return (a - 1) == l ? 1 : 0;
}
PF_HD u32 sview_words(const DB& db, u32 v) {
  const OLV a = olist(db, V2L(v)), b = olist(db, NEGL(V2L(v)));
  u32 n = 4 + a.n + b.n;
  for (u32 i = 0; i < a.n; i++) n += CL_HDR + 1 + csize(db, a.r[i]);
  for (u32 i = 0; i < b.n; i++) n += CL_HDR + 1 + csize(db, b.r[i]);
  return n;
}
PF_HD u64 sview_bound(u32 nlists, u64 sizes) { return 4 + (CL_HDR + 2) * (u64)nlists + sizes; }
PF_HD bool sview_stage(const DB& db, u32 v, u32* w, u32 cap, SView& sv) {
  const u32 p = V2L(v);
  const OLV a = olist(db, p), b = olist(db, NEGL(p));
  u32 at = 4 + a.n + b.n;
  if ((u64)at > cap) return false;
  w[0] = a.n;
  w[1] = b.n;
  w[2] = 4;
  w[3] = 4 + a.n;
  for (u32 s = 0; s < 2; s++) {
    const OLV ol = s ? b : a;
    u32* offs = w + w[2 + s];
    for (u32 i = 0; i < ol.n; i++) {
      const u32 c = ol.r[i];
      const u32* h = chead(db, c);
      const u32 n = hdr_size(h[0]);
      if ((u64)at + CL_HDR + 1 + n > cap) return false;
      offs[i] = at;
      w[at] = c;
      for (u32 q = 0; q < CL_HDR + n; q++) w[at + 1 + q] = h[q];
      at += CL_HDR + 1 + n;
    }
  }
  sv = SView{w, p, db.assign};
  return true;
}
// writes back the changed clauses (walked in staging order: a record spans
// the clause's size in the pool, which changes only here) and the lists
// that lost entries
PF_HD void sview_commit(const DB& db, const SView& v, u32 np0, u32 nn0) {
  u32 at = 4 + np0 + nn0;
  for (u32 i = 0; i < np0 + nn0; i++) {
    const u32 c = v.w[at];
    u32* h = chead(db, c);
    const u32 n0 = hdr_size(h[0]);
    const u32 hl = v.w[at + 1];
    if (hdr_size(hl) != n0) {
      const u32 n = hdr_size(hl);
      for (u32 q = 0; q < n; q++) h[CL_HDR + q] = v.w[at + 4 + q];
    }
    if (hl != h[0]) h[0] = hl;
    if (v.w[at + 2] != h[1]) h[1] = v.w[at + 2];
    if (v.w[at + 3] != h[2]) h[2] = v.w[at + 3];
    at += CL_HDR + 1 + n0;
  }
  for (u32 s = 0; s < 2; s++) {
    const u32 lit = v.lit | s, n = v.w[s];
    if (n == (s ? nn0 : np0)) continue;
    u32* r = db.ot_refs + db.ot_off[lit];
    const u32* offs = v.w + v.w[2 + s];
    for (u32 j = 0; j < n; j++) r[j] = v.w[offs[j]];
    db.ot_size[lit] = n;
  }
}
// writes the molten marks the searches changed back to the pool
PF_HD void lview_commit(const DB& db, const LView& v) {
  const u32 n = v.w[0] + v.w[1];
  for (u32 i = 0; i < n; i++) {
    const u32 r = v.w[2 + i];
    const u32 c = v.w[r];
    const u8 lf = hdr_flags(v.w[r + 1]), gf = cflags(db, c);
    if ((lf ^ gf) & F_MOLTEN) set_cflags(db, c, (u8)((gf & ~F_MOLTEN) | (lf & F_MOLTEN)));
  }
}

// ---------------------------------------------------------------- OT construction
// histogram of live literals (simplify.cpp:45-54 counts, histogram.cpp:82-95)
PF_HD void hist_clause(const DB& db, u32 c, u32* hist) {
  if (is_deleted(db, c)) return;
  const CL cl = clause(db, c);
  for (u32 i = 0; i < cl.n; i++) at_add(&hist[cl.l[i]], 1u);
}
// scatter of clause ids into the lists (simplify.cpp:45-54); order within a
// list is fixed afterwards where it matters (elected lists are sorted)
PF_HD void scatter_clause(const DB& db, u32 c, u32* cursor) {
  if (is_deleted(db, c)) return;
  const CL cl = clause(db, c);
  for (u32 i = 0; i < cl.n; i++) {
    const u32 l = cl.l[i];
    const u32 pos = at_add(&cursor[l], 1u);
    db.ot_refs[db.ot_off[l] + pos] = c;
  }
}

// CNF_CMP_KEY (src/cpu/simplify.h:29-42) completed by the clause id
struct SortKey {
  u32 size, lit0, back, sig, id;
};
PF_HD SortKey make_key(const DB& db, u32 c) {
  const u32 n = csize(db, c);
  const u32* l = clits(db, c);
  SortKey k;
  k.size = n;
  k.lit0 = n ? l[0] : 0;
  k.back = n ? l[n - 1] : 0;
  k.sig = csig(db, c);
  k.id = c;
  return k;
}
PF_HD bool key_less(const SortKey& a, const SortKey& b) {
  if (a.size != b.size) return a.size < b.size;
  if (a.lit0 != b.lit0) return a.lit0 < b.lit0;
  if (a.back != b.back) return a.back < b.back;
  if (a.sig != b.sig) return a.sig < b.sig;
  return a.id < b.id;
}
// insertion sort of one short list (used for lists up to a few dozen)
PF_HD void sort_list_insertion(const DB& db, u32* r, u32 n) {
  for (u32 i = 1; i < n; i++) {
    const u32 t = r[i];
    const SortKey kt = make_key(db, t);
    u32 j = i;
    while (j > 0) {
      const SortKey kj = make_key(db, r[j - 1]);
      if (!key_less(kt, kj)) break;
      r[j] = r[j - 1];
      j--;
    }
    r[j] = t;
  }
}

// ---------------------------------------------------------------- BCP
// Visit clause c after one of its literals became false.  Claims a new unit
// with a compare-and-swap on its variable and appends it to 'next'.
PF_HD void bcp_visit(const DB& db, u32 c, u32* next, u32* nnext, u32* conflict) {
  if (is_deleted(db, c)) return;
  const CL cl = clause(db, c);
  u32 nun = 0, last = 0;
  for (u32 i = 0; i < cl.n; i++) {
    const u32 l = cl.l[i];
    const u32 a = at_load(&db.assign[ABS(l)]);
    if (!a) {
      nun++;
      last = l;
    } else if (a - 1 == l) {
      return;  // satisfied
    }
  }
  if (nun == 0) {
    *conflict = 1;
    return;
  }
  if (nun == 1) {
    const u32 old = at_cas(&db.assign[ABS(last)], 0u, last + 1);
    if (old == 0) next[at_add(nnext, 1u)] = last;
    else if (old != last + 1) *conflict = 1;
  }
}
// After the fixpoint: delete satisfied clauses, strip false literals
// (elimination.cpp:41-62 applied for every assigned literal).
PF_HD void bcp_clean(const DB& db, u32 c) {
  if (is_deleted(db, c)) return;
  u32* l = clits(db, c);
  const u32 n = csize(db, c);
  u32 j = 0, sig = 0;
  for (u32 i = 0; i < n; i++) {
    const int v = lit_value(db, l[i]);
    if (v > 0) {
      set_deleted(db, c);
      return;
    }
    if (v < 0) {
      l[j++] = l[i];
      sig |= MAPHASH(l[i]);
    }
  }
  if (j != n) {
    set_csize(db, c, j);
    set_csig(db, c, sig);
  }
}

// ---------------------------------------------------------------- LCVE
// Greedy election (lcve.cpp:46-99) as the lexicographically-first
// independent set over the candidate order: a candidate is elected iff no
// earlier candidate sharing a live clause is elected.  state: 0 undecided,
// 1 elected, 2 frozen.  Returns 1 if the candidate stays undecided.
// The elected set does not depend on the schedule (the lexicographically
// first independent set is unique), only the number of passes does.
// pw[u] packs the candidate position of variable u and its state
// (position << 2 | state; NOPOS for variables that are not candidates), so
// a neighbour costs one load; state[] holds the same states by position.
// A scan reads the whole neighbourhood (an elected earlier neighbour
// freezes the candidate wherever it sits) and remembers the first earlier
// neighbour that is still undecided in blk[i] and the list entry where it
// met it in resume[i] (sign << 31 | entry).  States only move from
// undecided to elected or frozen, so everything before that entry stays
// irrelevant (later or frozen neighbours); while the blocker is undecided
// the candidate waits without a scan, an elected blocker freezes it, and a
// frozen one makes the next scan start from the entry where it met it.
// lcveclausemax (the GPU build's depFreeze_d, src/gpu/lcve.cu:43-52): a
// candidate with a live clause longer than the limit is passed over (it
// elects nothing and freezes nothing; the reference's partial freezes of a
// passed-over candidate's positive list are not reproduced)
PF_HD bool lcve_oversize(const DB& db, u32 v, u32 limit) {
  for (u32 s = 0; s < 2; s++) {
    const OLV ol = olist(db, V2L(v) | s);
    for (u32 k = 0; k < ol.n; k++) {
      const u32 c = ol.r[k];
      if (!is_deleted(db, c) && csize(db, c) > limit) return true;
    }
  }
  return false;
}
PF_HD void mis_decide(u32* state, u32* pw, u32 i, u32 v, u32 s) {
  state[i] = s;
  pw[v] = (i << 2) | s;
}
// ---------------------------------------------------------------- shared helpers
PF_HD bool sub_sig(u32 A, u32 B) { return !(A & ~B); }  // simplify.h:84
PF_HD bool selfsub_sig(u32 A, u32 B) {                   // simplify.h:86-90
  const u32 Bt = B | ((B & 0xAAAAAAAAu) >> 1) | ((B & 0x55555555u) << 1);
  return !(A & ~Bt);
}
// simplify.h:175-202: resolvent length on x, 0 when tautological
PF_HD u32 merge_len(u32 x, CL a, CL b) {
  u32 i1 = 0, i2 = 0;
  u32 len = a.n + b.n - 2;
  while (i1 < a.n && i2 < b.n) {
    const u32 l1 = a.l[i1], l2 = b.l[i2];
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
// simplify.h:139-173 for a resolvent known not to be tautological
PF_HD u32 merge_out(u32 x, CL a, CL b, u32* out) {
  u32 i1 = 0, i2 = 0, n = 0;
  while (i1 < a.n && i2 < b.n) {
    const u32 l1 = a.l[i1], l2 = b.l[i2];
    const u32 v1 = ABS(l1), v2 = ABS(l2);
    if (v1 == x) i1++;
    else if (v2 == x) i2++;
    else if (v1 < v2) { i1++; out[n++] = l1; }
    else if (v2 < v1) { i2++; out[n++] = l2; }
    else { i1++; i2++; out[n++] = l1; }
  }
  while (i1 < a.n) { const u32 l = a.l[i1++]; if (ABS(l) != x) out[n++] = l; }
  while (i2 < b.n) { const u32 l = b.l[i2++]; if (ABS(l) != x) out[n++] = l; }
  return n;
}
// simplify.h:119-137
PF_HD bool is_tautology(u32 x, CL a, CL b) {
  u32 i1 = 0, i2 = 0;
  while (i1 < a.n && i2 < b.n) {
    const u32 l1 = a.l[i1], l2 = b.l[i2];
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

// ---------------------------------------------------------------- SUB (subsume.h)
// DRAT proof events of the per-variable stages: records (elected index,
// event sequence, kind, payload).  The sequence numbers of one variable are
// shared with its unit records, so sorting everything by (index, sequence)
// gives the order in which the sequential simplifier writes its proof.
constexpr u32 PR_ADD = 1;     // 'a' + clause payload, unless satisfied when written (addResolvent)
constexpr u32 PR_DEL = 2;     // 'd' + clause payload (removeClause)
constexpr u32 PR_UNIT = 3;    // 'a' + literal payload (enqueueUnit)
constexpr u32 PR_EMPTY = 4;   // 'a' + empty clause (learnEmpty)
struct ProofOut {
  u32* rec;       // 4 words per record; null when proof output is off
  u32* n;
  u32 cap;
  u32* overflow;
};
PF_HD void push_proof(const ProofOut& po, u32 k, u32 seq, u32 kind, u32 payload) {
  if (!po.rec) return;
  const u32 r = at_add(po.n, 1u);
  if (r >= po.cap) {
    at_add(po.overflow, 1u);
    return;
  }
  po.rec[4 * (u64)r] = k;
  po.rec[4 * (u64)r + 1] = seq;
  po.rec[4 * (u64)r + 2] = kind;
  po.rec[4 * (u64)r + 3] = payload;
}

// encoding of one queued proof line (kind, payload) as the binary DRAT
// writer does it (src/cpu/proof.cpp:94-135): literal 2*vorg(v)+sign in
// 7-bit groups, low group first, continuation bit 7
PF_HD u32 proof_lit_value(const DB& db, u32 l) { return (db.vorg[ABS(l)] << 1) | SIGN(l); }
PF_HD u32 varint_bytes(u32 r) {
  u32 b = 1;
  while (r > 127) {
    r >>= 7;
    b++;
  }
  return b;
}
// PROOF::addResolvent (proof.cpp:214-229) drops a clause that holds a
// literal already true: assigned before this batch (stamp 0) or by a unit
// line earlier in it
PF_HD bool proof_line_dropped(const DB& db, const u32* lines, u32 i, const u32* stamp) {
  if (lines[2 * i] != PR_ADD) return false;
  const CL c = clause(db, lines[2 * i + 1]);
  for (u32 j = 0; j < c.n; j++) {
    const u32 l = c.l[j];
    if (lit_value(db, l) > 0) {
      const u32 s = stamp[ABS(l)];
      if (s == 0 || s < i + 1) return true;
    }
  }
  return false;
}
PF_HD u32 proof_line_size(const DB& db, const u32* lines, u32 i, const u32* stamp) {
  const u32 kind = lines[2 * i], pay = lines[2 * i + 1];
  if (kind == PR_EMPTY) return 2;
  if (kind == PR_UNIT) return 2 + varint_bytes(proof_lit_value(db, pay));
  if (proof_line_dropped(db, lines, i, stamp)) return 0;
  const CL c = clause(db, pay);
  u32 b = 2;
  for (u32 j = 0; j < c.n; j++) b += varint_bytes(proof_lit_value(db, c.l[j]));
  return b;
}
PF_HD u8* put_varint(u8* o, u32 r) {
  while (r > 127) {
    *o++ = (u8)((r & 127) | 128);
    r >>= 7;
  }
  *o++ = (u8)r;
  return o;
}
PF_HD void proof_line_write(const DB& db, const u32* lines, u32 i, u8* o) {
  const u32 kind = lines[2 * i], pay = lines[2 * i + 1];
  *o++ = kind == PR_DEL ? 'd' : 'a';
  if (kind == PR_UNIT) {
    o = put_varint(o, proof_lit_value(db, pay));
  } else if (kind != PR_EMPTY) {
    const CL c = clause(db, pay);
    for (u32 j = 0; j < c.n; j++) o = put_varint(o, proof_lit_value(db, c.l[j]));
  }
  *o = 0;
}

// Units derived by other elected variables of the same SUB stage, as a
// variable must see them when it runs after them: ust[x] = elected index of
// the variable whose unit assigned x (NOPOS: none), ulit[x] = that literal.
struct SubVis {
  const u32* ust;  // null: no stage units are visible (the first pass)
  const u32* ulit;
};

struct SubOut {
  // (elected index, sequence, literal) records of the units a variable
  // derives; literal 0 records a refutation at that sequence
  u32* rec;
  u32* n;         // number of records
  u32 cap;
  // per elected variable k, 2 words (literal, sequence) per unit it derives
  // at ubuf + 2 * ubase[k]; ubase is the scan of |poss| + |negs| and every
  // unit consumes one clause of those lists, so a slice never overflows
  u32* ubuf;
  const u32* ubase;
  ProofOut proof;
  SubVis vis;
};

// state of one variable's SUB run; units live in the variable's slice of
// SubOut::ubuf rather than in a private array (no scratch memory)
struct SubLocal {
  u32* ub;   // this variable's unit slice
  u32 n;     // units derived so far
  u32 ev;    // event counter of this variable
  bool unsat;
  u32 unsat_seq;
  u32 k;     // elected index
  ProofOut proof;
  SubVis vis;
};
// value of a literal seen by one variable: committed root assignment plus
// the units this variable derived so far
template <class D>
PF_HD int sub_value(const D& db, const SubLocal& sl, u32 l) {
  const int v = lit_value(db, l);
  if (v >= 0) return v;
  if (sl.vis.ust && sl.vis.ust[ABS(l)] < sl.k) return sl.vis.ulit[ABS(l)] == l ? 1 : 0;
  for (u32 i = 0; i < sl.n; i++) {
    if (sl.ub[2 * i] == l) return 1;
    if (sl.ub[2 * i] == FLIP(l)) return 0;
  }
  return -1;
}
PF_HD bool has_lit(CL c, u32 l) {
  for (u32 i = 0; i < c.n; i++)
    if (c.l[i] == l) return true;
  return false;
}
// deletes the clauses of this variable's lists that contain 'unit'
// (the part of toblivion(ot[unit]) this variable observes)
template <class D>
PF_HD void sub_local_oblivion(const D& db, u32 p, u32 unit) {
  for (u32 s = 0; s < 2; s++) {
    const OLV ol = olist(db, p | s);
    for (u32 k = 0; k < ol.n; k++) {
      const u32 c = ol.r[k];
      if (!is_deleted(db, c) && has_lit(clause(db, c), unit)) set_deleted(db, c);
    }
  }
}
template <class D>
PF_HD void bump_shrunken(const D& db, u32 c, int lbd_tier1) {  // elimination.cpp:28-39
  const int old = (int)clbd(db, c);
  if (old <= lbd_tier1) return;
  int nl = (int)csize(db, c) - 1;
  if (nl > old) nl = old;
  if (nl >= old) return;
  set_clbd(db, c, (u32)nl);
  set_cflags(db, c, (u8)((cflags(db, c) & ~USAGE_MASK) | (1u << USAGE_SHIFT)));
}
// elimination.cpp:199-234
template <class D>
PF_HD void strengthen(const D& db, u32 c, u32 me, u32 p, SubLocal& sl, int lbd_tier1) {
  u32* l = clits(db, c);
  const u32 n0 = csize(db, c);
  u32 n = 0, sig = 0;
  for (u32 k = 0; k < n0; k++) {
    const u32 x = l[k];
    if (x != me) { l[n++] = x; sig |= MAPHASH(x); }
  }
  set_csig(db, c, sig);
  set_csize(db, c, n0 - 1);
  if (n == 1) {
    const u32 unit = l[0];
    const int val = sub_value(db, sl, unit);
    if (val < 0) {
      sl.ub[2 * sl.n] = unit;
      sl.ub[2 * sl.n + 1] = sl.ev++;
      sl.n++;
      sub_local_oblivion(db, p, unit);
    } else if (!val) {
      if (!sl.unsat) {
        sl.unsat_seq = sl.ev++;
        push_proof(sl.proof, sl.k, sl.unsat_seq, PR_EMPTY, 0);
      }
      sl.unsat = true;
    }
  } else {
    push_proof(sl.proof, sl.k, sl.ev++, PR_ADD, cid(db, c));
    if (is_learnt(db, c)) bump_shrunken(db, c, lbd_tier1);
  }
}
// subsume.h:39-56
PF_HD bool sub_lits(CL sm, CL lr) {
  u32 i1 = 0, i2 = 0, s = 0;
  while (i1 < sm.n && i2 < lr.n) {
    const u32 a = sm.l[i1], b = lr.l[i2];
    if (a < b) i1++;
    else if (b < a) i2++;
    else { s++; i1++; i2++; }
  }
  return s == sm.n;
}
// subsume.h:75-103
PF_HD bool selfsub_lits(u32 x, u32 fx, CL sm, CL lr) {
  u32 i1 = 0, i2 = 0, s = 0;
  bool self = false;
  while (i1 < sm.n && i2 < lr.n) {
    const u32 a = sm.l[i1], b = lr.l[i2];
    if (a == fx) i1++;
    else if (b == x) { self = true; i2++; }
    else if (a < b) i1++;
    else if (b < a) i2++;
    else { s++; i1++; i2++; }
  }
  if (s + 1 == sm.n) {
    if (self) return true;
    while (i2 < lr.n) {
      if (lr.l[i2] == x) return true;
      i2++;
    }
  }
  return false;
}
// subsume.h:105-121: is 'cand' subsumed by an earlier entry of the list
// Groups of G consecutive lanes of a wavefront may share one variable
// (gl = 0..G-1 is the lane's place in its group; G = 1 is one lane per
// variable).  group_first returns the lowest j in [0, n) with pred(j), n if
// none: the group tests j = gl, gl + G, ... and takes the lowest hit of the
// first stride that has one, which is the sequential scan's answer; pred
// must not change memory.  The group's first lane does every write (the
// SUB actions), group_sync orders those writes before the group's next
// reads of the variable's LDS copy.
template <class P>
PF_HD u32 group_first(u32 n, u32 gl, u32 G, P pred) {
#ifdef __HIP_DEVICE_COMPILE__
  if (G > 1) {
    const u32 g0 = (threadIdx.x & 63) - gl;
    const u64 gmask = (G >= 64 ? ~0ull : ((1ull << G) - 1)) << g0;
    for (u32 j0 = 0; j0 < n; j0 += G) {
      const u32 j = j0 + gl;
      const bool hit = j < n && pred(j);
      const u64 m = __ballot(hit) & gmask;
      if (m) return j0 + (u32)(__ffsll((long long)m) - 1) - g0;
    }
    return n;
  }
#endif
  for (u32 j = 0; j < n; j++)
    if (pred(j)) return j;
  return n;
}
PF_HD void group_sync() {
#ifdef __HIP_DEVICE_COMPILE__
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
#endif
}
// the lowest j in [from, n) with pred(j), n if none (as group_first)
template <class P>
PF_HD u32 group_first_from(u32 from, u32 n, u32 gl, u32 G, P pred) {
  if (from >= n) return n;
  return from + group_first(n - from, gl, G, [&](u32 j) { return pred(from + j); });
}
// reductions over the G lanes of a group (G a power of two, groups aligned)
PF_HD u32 group_min(u32 x, u32 G) {
#ifdef __HIP_DEVICE_COMPILE__
  for (u32 d = 1; d < G; d <<= 1) x = min(x, (u32)__shfl_xor(x, d, 64));
#endif
  return x;
}
PF_HD u32 group_sum(u32 x, u32 G) {
#ifdef __HIP_DEVICE_COMPILE__
  for (u32 d = 1; d < G; d <<= 1) x += (u32)__shfl_xor(x, d, 64);
#endif
  return x;
}
// subsume.h:105-121: is 'cand' subsumed by an earlier entry of the list
template <class D>
PF_HD bool subsume_cand(const D& db, OLV ol, u32 end, u32 cand, SubLocal& sl, u32 gl = 0, u32 G = 1) {
  const u32 csz = csize(db, cand), csg = csig(db, cand);
  const bool cm = is_molten(db, cand);
  const CL cc = clause(db, cand);
  const u32 j = group_first(end, gl, G, [&](u32 j) {
    const u32 s = ol.r[j];
    if (is_deleted(db, s)) return false;
    const u32 ssz = csize(db, s);
    if (cm && ssz > csz) return false;
    return ssz > 1 && sub_sig(csig(db, s), csg) && sub_lits(clause(db, s), cc);
  });
  if (j == end) return false;
  if (gl == 0) {
    const u32 s = ol.r[j];
    if (is_learnt(db, s) && is_original(db, cand)) set_status(db, s, ST_ORIGINAL);
    push_proof(sl.proof, sl.k, sl.ev++, PR_DEL, cid(db, cand));
    set_deleted(db, cand);
  }
  group_sync();
  return true;
}
// subsume.h:123-142: strengthen 'cand' by removing x if a clause of 'ol'
// self-subsumes it (the scan stops at the first clause longer than cand)
template <class D>
PF_HD bool selfsubsume_cand(const D& db, u32 x, u32 fx, OLV ol, u32 cand, u32 p, SubLocal& sl,
                            int lbd_tier1, u32 gl = 0, u32 G = 1) {
  const u32 csz = csize(db, cand);
  const u32 cand_sig = csig(db, cand);
  const CL cc = clause(db, cand);
  const u32 j = group_first(ol.n, gl, G, [&](u32 j) {
    const u32 s = ol.r[j];
    const u32 ssz = csize(db, s);
    if (ssz > csz) return true;  // the stop
    if (is_deleted(db, s)) return false;
    return ssz > 1 && selfsub_sig(csig(db, s), cand_sig) && selfsub_lits(x, fx, clause(db, s), cc);
  });
  if (j == ol.n || csize(db, ol.r[j]) > csz) return false;
  if (gl == 0) {
    strengthen(db, cand, x, p, sl, lbd_tier1);
    melt(db, cand);
  }
  group_sync();
  return true;
}
// subsume.h:27-37
PF_HD void update_ol(const DB& db, u32 lit) {
  u32* r = db.ot_refs + db.ot_off[lit];
  const u32 n = db.ot_size[lit];
  u32 j = 0;
  for (u32 i = 0; i < n; i++) {
    const u32 c = r[i];
    if (is_molten(db, c)) freeze(db, c);
    else if (!is_deleted(db, c)) r[j++] = c;
  }
  db.ot_size[lit] = j;
}
PF_HD void update_ol(const SView& v, u32 lit) {
  const u32 s = lit == v.lit ? 0u : 1u;
  u32* r = v.w + v.w[2 + s];
  const u32 n = v.w[s];
  u32 j = 0;
  for (u32 i = 0; i < n; i++) {
    const u32 c = r[i];
    if (is_molten(v, c)) freeze(v, c);
    else if (!is_deleted(v, c)) r[j++] = c;
  }
  v.w[s] = j;
}
// subsume.h:144-179 + the list-size gate of elimination.cpp:121-124, for
// elected variable v (index k in elected order)
PF_HD bool sub_visible_unit(const SubVis& vis, u32 k, CL c) {
  for (u32 i = 0; i < c.n; i++) {
    const u32 x = ABS(c.l[i]);
    if (vis.ust[x] < k && vis.ulit[x] == c.l[i]) return true;
  }
  return false;
}
// D: the pool (DB) or the variable's lists staged in an SView
// gl, G: the lane's place in the group of lanes running this variable
template <class D>
PF_HD void sub_var(const D& db, const Options& o, u32 v, u32 k, const SubOut& out, u32 gl = 0, u32 G = 1) {
  const u32 p = V2L(v), n = NEGL(p);
  u32 np = olist(db, p).n, nn = olist(db, n).n;
  if (out.vis.ust) {
    // toblivion(ot[unit]) of the units earlier variables derived: their
    // clauses are deleted and the unit's own list emptied before this
    // variable runs (elimination.cpp:214-220, simplify.h:297-303)
    if (out.vis.ust[v] < k) {
      if (out.vis.ulit[v] == p) np = 0;
      else nn = 0;
    }
    if (gl == 0)
      for (u32 q = 0; q < 2; q++) {
        const OLV ol = olist(db, p | q);
        for (u32 i = 0; i < ol.n; i++) {
          const u32 c = ol.r[i];
          if (!is_deleted(db, c) && sub_visible_unit(out.vis, k, clause(db, c))) set_deleted(db, c);
        }
      }
    group_sync();
  }
  if ((int)np > o.sub_limit || (int)nn > o.sub_limit) return;
  SubLocal sl;
  sl.ub = out.ubuf + 2 * (u64)out.ubase[k];
  sl.n = 0;
  sl.ev = 0;
  sl.unsat = false;
  sl.unsat_seq = 0;
  sl.k = k;
  sl.proof = out.proof;
  sl.vis = out.vis;
  {
    const OLV poss = olist(db, p);
    for (u32 i = 0; i < poss.n; i++) {
      const u32 c = poss.r[i];
      if ((int)csize(db, c) > HSE_MAX_CL_SIZE) break;
      if (is_deleted(db, c)) continue;
      selfsubsume_cand(db, p, n, olist(db, n), c, p, sl, o.lbd_tier1, gl, G);
      subsume_cand(db, poss, i, c, sl, gl, G);
    }
    if (gl == 0) update_ol(db, p);
    group_sync();
  }
  {
    const OLV negs = olist(db, n);
    for (u32 i = 0; i < negs.n; i++) {
      const u32 c = negs.r[i];
      if ((int)csize(db, c) > HSE_MAX_CL_SIZE) break;
      if (is_deleted(db, c)) continue;
      selfsubsume_cand(db, n, p, olist(db, p), c, p, sl, o.lbd_tier1, gl, G);
      subsume_cand(db, negs, i, c, sl, gl, G);
    }
    if (gl == 0) update_ol(db, n);
    group_sync();
  }
  if (gl) return;  // the group's first lane holds the units
  for (u32 i = 0; i < sl.n; i++) {
    const u32 r = at_add(out.n, 1u);
    if (r < out.cap) {
      out.rec[3 * (u64)r] = k;
      out.rec[3 * (u64)r + 1] = sl.ub[2 * i + 1];
      out.rec[3 * (u64)r + 2] = sl.ub[2 * i];
    }
  }
  if (sl.unsat) {
    const u32 r = at_add(out.n, 1u);
    if (r < out.cap) {
      out.rec[3 * (u64)r] = k;
      out.rec[3 * (u64)r + 1] = sl.unsat_seq;
      out.rec[3 * (u64)r + 2] = 0;
    }
  }
}

// snapshot / restore of what sub_var may change (list contents and the
// clauses in them), so that a variable can be re-run after earlier units
struct SubSnap {
  u32* base;   // per elected variable: start in the entry arrays
  u32* np;     // per elected variable: list sizes before SUB
  u32* nn;
  u32* ref;    // per entry
  u8* flg;
  u32* size;
  u32* sig;
  u32* lbd;
  u32* owner;  // per clause: elected index of the variable whose lists hold it live
};
// D: the pool, or the variable's staged SView before its SUB run changes it
template <class D>
PF_HD void sub_snapshot_var(const D& db, u32 v, u32 k, const SubSnap& s) {
  u32 e = s.base[k];
  for (u32 q = 0; q < 2; q++) {
    const OLV ol = olist(db, V2L(v) | q);
    if (q == 0) s.np[k] = ol.n;
    else s.nn[k] = ol.n;
    for (u32 j = 0; j < ol.n; j++, e++) {
      const u32 c = ol.r[j];
      if (!is_deleted(db, c)) s.owner[cid(db, c)] = k;  // live clauses have one elected owner
      s.ref[e] = cid(db, c);
      s.flg[e] = cflags(db, c);
      s.size[e] = csize(db, c);
      s.sig[e] = csig(db, c);
      s.lbd[e] = clbd(db, c);
    }
  }
}
PF_HD void sub_restore_var(const DB& db, u32 v, u32 k, const SubSnap& s) {
  u32 e = s.base[k];
  for (u32 q = 0; q < 2; q++) {
    const u32 lit = V2L(v) | q;
    const u32 n = q == 0 ? s.np[k] : s.nn[k];
    u32* r = db.ot_refs + db.ot_off[lit];
    for (u32 j = 0; j < n; j++, e++) {
      const u32 c = s.ref[e];
      r[j] = c;
      set_cflags(db, c, s.flg[e]);
      set_csig(db, c, s.sig[e]);
      set_clbd(db, c, s.lbd[e]);
      const u32 cur = csize(db, c);
      if (cur < s.size[e]) {
        // the only literal SUB removes from a clause of this list is 'lit'
        u32* l = clits(db, c);
        u32 pos = cur;
        while (pos > 0 && l[pos - 1] > lit) {
          l[pos] = l[pos - 1];
          pos--;
        }
        l[pos] = lit;
        set_csize(db, c, s.size[e]);
      }
    }
    db.ot_size[lit] = n;
  }
}
// A variable must run again when one of its clauses contains a unit that an
// earlier elected variable derived (toblivion(ot[unit]) deletes that clause
// before the variable runs) and its last run did not see that unit, or saw
// one that is gone.  The clauses containing the unit are the unit's list;
// their owners after the deriving variable are the ones affected.
PF_HD u32 sub_affected_owner(const SubSnap& s, u32 c, u32 k_unit, const u32* lastrun, u32 pass) {
  const u32 o = s.owner[c];
  return (o != NOPOS && o > k_unit && lastrun[o] <= pass) ? o : NOPOS;
}

// ---------------------------------------------------------------- BVE (bve.cpp, gates)
enum : u32 { VT_NONE = 0, VT_PURE = 1, VT_EQU = 2, VT_RES = 3, VT_SUBST = 4 };
struct VeDecision {
  u32* type;      // per elected variable
  u32* def;       // EQU definition literal
  u32* ncls;      // resolvents of size > 1
  u32* nlits;     // their pool words (headers + literals)
  u32* nstack;    // eliminated-stack words
  u32* scratch;   // per variable scratch (fan-in literals)
  u32* scr_base;  // per variable start in scratch
};

template <class D>
PF_HD u32 count_orgs(const D& db, OLV ol) {
  u32 n = 0;
  for (u32 i = 0; i < ol.n; i++) n += is_original(db, ol.r[i]);
  return n;
}
template <class D>
PF_HD u32 count_lits_before(const D& db, OLV ol) {
  u32 n = 0;
  for (u32 i = 0; i < ol.n; i++)
    if (is_original(db, ol.r[i])) n += csize(db, ol.r[i]);
  return n;
}
// simplify.h:237-295: true when the bound is exceeded; nAddedCls / lits
// On a group (G > 1) the pairs are shared out and the counts summed: the
// bound is exceeded at some pair iff the total exceeds it, and a caller
// reads nAddedCls after a true return only as zero / not zero (every call
// starts from zero; a true return has counted at least one resolvent).
template <class D>
PF_HD bool count_res(const D& db, const Options& o, u32 x, int clsbefore, OLV me, OLV other,
                     int& nAddedCls, u32& nAddedLits, bool substitute, u32 gl = 0, u32 G = 1) {
  nAddedLits = 0;
  const int rlimit = o.ve_clause_limit;
  if (G > 1) {
    u32 cls = 0, lits = 0, over = 0;
    const u32 np = me.n * other.n;
    for (u32 q = gl; q < np; q += G) {
      const u32 ci = me.r[q / other.n], cj = other.r[q % other.n];
      if (!is_original(db, ci) || !is_original(db, cj)) continue;
      if (substitute && is_molten(db, ci) == is_molten(db, cj)) continue;
      const u32 rsize = merge_len(x, clause(db, ci), clause(db, cj));
      if (rsize > 1) {
        cls++;
        lits += rsize;
        if (rlimit && (int)rsize > rlimit) over = 1;
      }
    }
    nAddedCls += (int)group_sum(cls, G);
    nAddedLits = group_sum(lits, G);
    if (nAddedCls > clsbefore || group_sum(over, G)) return true;
  } else {
  for (u32 i = 0; i < me.n; i++) {
    const u32 ci = me.r[i];
    if (!is_original(db, ci)) continue;
    const bool a = is_molten(db, ci);
    const CL A = clause(db, ci);
    for (u32 j = 0; j < other.n; j++) {
      const u32 cj = other.r[j];
      if (!is_original(db, cj)) continue;
      if (substitute && a == is_molten(db, cj)) continue;
      const u32 rsize = merge_len(x, A, clause(db, cj));
      if (rsize > 1) {
        if (++nAddedCls > clsbefore || (rlimit && (int)rsize > rlimit)) return true;
        nAddedLits += rsize;
      }
    }
  }
  }
  if (o.ve_lbound_en) {
    const u32 before = count_lits_before(db, me) + count_lits_before(db, other);
    if (nAddedLits > before) return true;
  }
  return false;
}
template <class D>
PF_HD void freeze_binaries(const D& db, OLV ol) {  // simplify.h:204-211
  for (u32 i = 0; i < ol.n; i++) {
    const u32 c = ol.r[i];
    if (is_original(db, c) && csize(db, c) == 2) freeze(db, c);
  }
}
// equivalence.h:89-131
template <class D>
PF_HD u32 find_sfanin(const D& db, u32 gate_out, OLV ol) {
  u32 imp = 0;
  int nImps = 0;
  for (u32 i = 0; i < ol.n; i++) {
    const u32 c = ol.r[i];
    if (is_original(db, c) && csize(db, c) == 2) {
      const u32* l = clits(db, c);
      imp = FLIP(l[0] ^ l[1] ^ gate_out);
      melt(db, c);
      nImps++;
    }
    if (nImps > 1) return 0;
  }
  return imp;
}
template <class D>
PF_HD u32 find_equ_gate(const D& db, u32 p) {
  const u32 n = NEGL(p);
  const OLV poss = olist(db, p), negs = olist(db, n);
  u32 first = find_sfanin(db, p, poss);
  if (first) {
    u32 second = n, def = first;
    if (second < first) first = second, second = def;
    for (u32 i = 0; i < negs.n; i++) {
      const u32 c = negs.r[i];
      if (is_original(db, c) && csize(db, c) == 2) {
        const u32* l = clits(db, c);
        if (l[0] == first && l[1] == second) {
          melt(db, c);
          return def;
        }
      }
    }
  }
  freeze_binaries(db, poss);
  return 0;
}
// and.h:25-86; 'out' is per-variable scratch
template <class D>
PF_HD bool find_ao_gate(const D& db, const Options& o, u32 dx, int nOrgCls, u32* out, int& nAddedCls,
                        u32& nAddedLits) {
  const u32 x = ABS(dx);
  const OLV it = olist(db, dx);
  u32 nout = 0, sig = 0;
  for (u32 i = 0; i < it.n; i++) {
    const u32 c = it.r[i];
    if (is_original(db, c) && csize(db, c) == 2) {
      const u32* l = clits(db, c);
      const u32 imp = FLIP(l[0] ^ l[1] ^ dx);
      out[nout++] = imp;
      sig |= MAPHASH(imp);
      melt(db, c);
    }
  }
  if (nout > 1) {
    const u32 f = FLIP(dx);
    out[nout++] = f;
    sig |= MAPHASH(f);
    for (u32 i = 1; i < nout; i++) {  // ascending order (distinct literals)
      const u32 t = out[i];
      u32 j = i;
      while (j > 0 && out[j - 1] > t) { out[j] = out[j - 1]; j--; }
      out[j] = t;
    }
    const OLV ot2 = olist(db, f);
    for (u32 i = 0; i < ot2.n; i++) {
      const u32 c = ot2.r[i];
      if (!is_original(db, c) || csize(db, c) != nout || !sub_sig(csig(db, c), sig)) continue;
      const u32* l = clits(db, c);
      bool eq = true;
      for (u32 q = 0; q < nout && eq; q++) eq = (l[q] == out[q]);
      if (!eq) continue;
      melt(db, c);
      nAddedCls = 0;
      if (count_res(db, o, x, nOrgCls, it, ot2, nAddedCls, nAddedLits, true)) {
        freeze(db, c);
        break;
      }
      return true;
    }
  }
  freeze_binaries(db, it);
  return false;
}
// ifthenelse.h:25-45: first (lowest id) live non-molten original ternary
// equal to {x, y, z}; NOPOS if none.  The reference scans the shortest of
// the three lists; every such clause is in all three and the lowest id is
// taken, so scanning the list of x (the eliminated variable's literal,
// whose clauses the gate search reads anyway) gives the same clause.
template <class D>
PF_HD u32 fast_equality_check(const D& db, u32 x, u32 y, u32 z, u32 gl = 0, u32 G = 1) {
  const OLV ol = olist(db, x);
  u32 a = x, b = y, c = z, t;
  if (b > c) { t = b; b = c; c = t; }
  if (a > c) { t = a; a = c; c = t; }
  if (a > b) { t = a; a = b; b = t; }
  u32 best = NOPOS, best_id = NOPOS;
  for (u32 i = gl; i < ol.n; i += G) {
    const u32 cl = ol.r[i];
    if (is_molten(db, cl)) continue;
    if (!is_original(db, cl) || csize(db, cl) != 3) continue;
    const u32* l = clits(db, cl);
    if (l[0] == a && l[1] == b && l[2] == c && cid(db, cl) < best_id) {
      best = cl;
      best_id = cid(db, cl);
    }
  }
  if (G > 1) {  // the lowest id over the group, then its clause
    const u32 m = group_min(best_id, G);
    best = group_min(best_id == m ? best : NOPOS, G);
  }
  return best;
}
// the second clause of an ITE candidate pair: cj ternary original, and after
// ordering (x first, then y matching yi's variable) yi = -yj, zi and zj of
// different variables
template <class D>
PF_HD bool ite_pair(const D& db, u32 dx, u32 yi, u32 zi, u32 cj, u32& yj_out, u32& zj_out) {
  if (!(is_original(db, cj) && csize(db, cj) == 3)) return false;
  const u32* lj = clits(db, cj);
  u32 xj = lj[0], yj = lj[1], zj = lj[2], t;
  if (yj == dx) { t = xj; xj = yj; yj = t; }
  if (zj == dx) { t = xj; xj = zj; zj = t; }
  if (ABS(yi) == ABS(zj)) { t = yj; yj = zj; zj = t; }
  if (ABS(zi) == ABS(zj)) return false;
  if (yi != FLIP(yj)) return false;
  yj_out = yj;
  zj_out = zj;
  return true;
}
// pairs (i, j) in the reference's order; on a group, the candidates j of
// each i are found with group_first_from and the two equality checks are
// shared
template <class D>
PF_HD bool find_ite_gate(const D& db, const Options& o, u32 dx, int nOrgCls, int& nAddedCls,
                         u32& nAddedLits, u32 gl = 0, u32 G = 1) {
  const u32 fx = FLIP(dx);
  const OLV it = olist(db, dx);
  for (u32 i = 0; i < it.n; i++) {
    const u32 ci = it.r[i];
    if (!(is_original(db, ci) && csize(db, ci) == 3)) continue;
    const u32* li = clits(db, ci);
    u32 xi = li[0], yi = li[1], zi = li[2], t;
    if (yi == dx) { t = xi; xi = yi; yi = t; }
    if (zi == dx) { t = xi; xi = zi; zi = t; }
    for (u32 j = i + 1; j < it.n; j++) {
      u32 yj = 0, zj = 0;
      j = group_first_from(j, it.n, gl, G, [&](u32 q) { u32 a, b; return ite_pair(db, dx, yi, zi, it.r[q], a, b); });
      if (j == it.n) break;
      const u32 cj = it.r[j];
      ite_pair(db, dx, yi, zi, cj, yj, zj);
      const u32 d1 = fast_equality_check(db, fx, yi, FLIP(zi), gl, G);
      if (d1 == NOPOS) continue;
      const u32 d2 = fast_equality_check(db, fx, yj, FLIP(zj), gl, G);
      if (d2 == NOPOS) continue;
      if (gl == 0) { melt(db, ci); melt(db, cj); melt(db, d1); melt(db, d2); }
      group_sync();
      nAddedCls = 0;
      if (count_res(db, o, ABS(dx), nOrgCls, it, olist(db, fx), nAddedCls, nAddedLits, true, gl, G)) {
        if (gl == 0) { freeze(db, ci); freeze(db, cj); freeze(db, d1); freeze(db, d2); }
        group_sync();
        return false;
      }
      return true;
    }
  }
  return false;
}
// xor.h:31-43
template <class D>
PF_HD void freeze_arities(const D& db, OLV me, OLV other, u32 gl = 0, u32 G = 1) {
  if (gl == 0) {
    for (u32 i = 0; i < me.n; i++) {
      const u32 c = me.r[i];
      if (csize(db, c) > 2 && is_molten(db, c)) freeze(db, c);
    }
    for (u32 i = 0; i < other.n; i++) {
      const u32 c = other.r[i];
      if (csize(db, c) > 2 && is_molten(db, c)) freeze(db, c);
    }
  }
  group_sync();
}
// xor.h:70-102: flip the literals for the next parity pattern and look for
// the matching clause (lowest id).  The pattern is the base clause with
// literal k flipped where bit k of 'parity' is set; clauses hold distinct
// variables in ascending order, so a clause of the same size holds exactly
// the pattern's literals iff it equals the pattern position by position.
// The reference scans the shortest of the pattern's lists; the matches are
// in every one of them and the lowest id is taken, so the list of the
// pattern's literal of the eliminated variable (at position kx of the base)
// gives the same clause.
PF_HD u32 xor_lit(const u32* base, u32 parity, u32 k) { return base[k] ^ ((parity >> k) & 1u); }
template <class D>
PF_HD bool make_arity(const D& db, u32& parity, const u32* base, u32 size, u32 kx, u32 gl = 0, u32 G = 1) {
  do { parity++; } while (__builtin_popcount(parity) & 1);
  const OLV ol = olist(db, xor_lit(base, parity, kx));
  u32 found = NOPOS, found_id = NOPOS;
  for (u32 i = gl; i < ol.n; i += G) {
    const u32 c = ol.r[i];
    const u32 id = cid(db, c);
    if (id >= found_id) continue;
    if (!is_original(db, c) || csize(db, c) != size) continue;
    const u32* l = clits(db, c);
    bool all = true;
    for (u32 q = 0; q < size && all; q++) all = (l[q] == xor_lit(base, parity, q));
    if (all) {
      found = c;
      found_id = id;
    }
  }
  if (G > 1) {  // the lowest id over the group, then its clause
    const u32 m = group_min(found_id, G);
    found = group_min(found_id == m ? found : NOPOS, G);
  }
  if (found == NOPOS) return false;
  if (gl == 0) melt(db, found);
  group_sync();
  return true;
}
template <class D>
PF_HD bool find_xor_gate(const D& db, const Options& o, u32 dx, int nOrgCls, int& nAddedCls,
                         u32& nAddedLits, u32 gl = 0, u32 G = 1) {
  const u32 fx = FLIP(dx), v = ABS(dx);
  const OLV it = olist(db, dx), ot2 = olist(db, fx);
  for (u32 i = 0; i < it.n; i++) {
    const u32 ci = it.r[i];
    if (!is_original(db, ci)) continue;
    const u32 size = csize(db, ci);
    const int arity = (int)size - 1;
    if (size < 3 || arity > o.xor_max_arity || size > 32) continue;
    const u32* base = clits(db, ci);
    u32 kx = 0;
    while (kx < size && base[kx] != dx) kx++;
    if (kx == size) continue;  // not reached: the clause is in the list of dx
    u32 parity = 0;
    int itargets = 1 << arity;
    while (--itargets && make_arity(db, parity, base, size, kx, gl, G)) {
    }
    if (itargets) freeze_arities(db, it, ot2, gl, G);
    else {
      if (gl == 0) melt(db, ci);
      group_sync();
      nAddedCls = 0;
      if (count_res(db, o, v, nOrgCls, it, ot2, nAddedCls, nAddedLits, true, gl, G)) {
        freeze_arities(db, it, ot2, gl, G);
        break;
      }
      return true;
    }
  }
  return false;
}
// stack words written by toblivion / save_BN_gate (simplify.h:305-332,
// equivalence.h:133-154): the originals of the side with fewer originals
template <class D>
PF_HD u32 stack_words(const D& db, OLV side) {
  u32 w = 2;
  for (u32 i = 0; i < side.n; i++) {
    const u32 c = side.r[i];
    if (is_original(db, c)) w += csize(db, c) + 1;
  }
  return w;
}
// bve.cpp:138-241, decision part, for elected variable v at index k, in
// two parts: the pure, equivalence and AND/OR checks (ve_decide_gates) for
// every elected variable, then the ITE and XOR checks and the resolvent
// count (ve_decide_rest) for the variables the first part leaves open.
// The first part of a failed substitution count (nAddedCls > 0 disables
// the negative-side searches and the plain count) is carried in d.ncls[k].
template <class D>
PF_HD void ve_write(const D& db, u32 k, const VeDecision& d, OLV side, u32 type, u32 def, int nAddedCls,
                    u32 nAddedLits) {
  d.type[k] = type;
  d.def[k] = def;
  const bool adds = (type == VT_RES || type == VT_SUBST) && nAddedCls;
  d.ncls[k] = adds ? (u32)nAddedCls : 0;
  d.nlits[k] = adds ? nAddedLits + CL_HDR * (u32)nAddedCls : 0;  // pool words
  d.nstack[k] = type ? stack_words(db, side) : 0;
}
// returns false when the variable is left to ve_decide_rest
template <class D>
PF_HD bool ve_decide_gates(const D& db, const Options& o, u32 v, u32 k, const VeDecision& d) {
  const u32 p = V2L(v), n = NEGL(p);
  const OLV poss = olist(db, p), negs = olist(db, n);
  const u32 pOrgs = count_orgs(db, poss), nOrgs = count_orgs(db, negs);
  const OLV side = pOrgs > nOrgs ? negs : poss;
  u32 type = VT_NONE, def = 0;
  int nAddedCls = 0;
  u32 nAddedLits = 0;
  if (!pOrgs || !nOrgs) {
    type = VT_PURE;
  } else if ((def = find_equ_gate(db, p)) != 0) {
    type = VT_EQU;
  } else {
    const int nOrgCls = (int)(pOrgs + nOrgs);
    u32* out = d.scratch + d.scr_base[k];
    if (nOrgCls > 2) {
      if (find_ao_gate(db, o, n, nOrgCls, out, nAddedCls, nAddedLits)) type = VT_SUBST;
      else if (!nAddedCls && find_ao_gate(db, o, p, nOrgCls, out, nAddedCls, nAddedLits)) type = VT_SUBST;
    }
    if (!type && nOrgCls > 3) {
      d.ncls[k] = (u32)nAddedCls;
      return false;
    }
    if (!type && !nAddedCls && !count_res(db, o, v, nOrgCls, poss, negs, nAddedCls, nAddedLits, false))
      type = VT_RES;
  }
  ve_write(db, k, d, side, type, def, nAddedCls, nAddedLits);
  return true;
}
// D: the pool (DB) or the variable's lists staged in an LView
template <class D>
PF_HD void ve_decide_rest(const D& db, const Options& o, u32 v, u32 k, const VeDecision& d, u32 gl = 0,
                          u32 G = 1) {
  const u32 p = V2L(v), n = NEGL(p);
  const OLV poss = olist(db, p), negs = olist(db, n);
  const u32 pOrgs = count_orgs(db, poss), nOrgs = count_orgs(db, negs);
  const OLV side = pOrgs > nOrgs ? negs : poss;
  const int nOrgCls = (int)(pOrgs + nOrgs);
  u32 type = VT_NONE;
  int nAddedCls = (int)d.ncls[k];
  u32 nAddedLits = 0;
  if (find_ite_gate(db, o, p, nOrgCls, nAddedCls, nAddedLits, gl, G)) type = VT_SUBST;
  else if (!nAddedCls && find_ite_gate(db, o, n, nOrgCls, nAddedCls, nAddedLits, gl, G)) type = VT_SUBST;
  else if (find_xor_gate(db, o, p, nOrgCls, nAddedCls, nAddedLits, gl, G)) type = VT_SUBST;
  else if (!nAddedCls && find_xor_gate(db, o, n, nOrgCls, nAddedCls, nAddedLits, gl, G)) type = VT_SUBST;
  if (!type && !nAddedCls && !count_res(db, o, v, nOrgCls, poss, negs, nAddedCls, nAddedLits, false, gl, G))
    type = VT_RES;
  if (gl == 0) ve_write(db, k, d, side, type, 0, nAddedCls, nAddedLits);
}
PF_HD void ve_decide(const DB& db, const Options& o, u32 v, u32 k, const VeDecision& d) {
  if (!ve_decide_gates(db, o, v, k, d)) ve_decide_rest(db, o, v, k, d);
}

// unit records appended by VE / BCP: (elected index, sequence, literal)
struct UnitOut {
  u32* rec;   // 3 words per record
  u32* n;
  u32 cap;
};
PF_HD void push_unit(const UnitOut& u, u32 k, u32 seq, u32 lit) {
  const u32 i = at_add(u.n, 1u);
  if (i < u.cap) {
    u.rec[3 * (u64)i] = k;
    u.rec[3 * (u64)i + 1] = seq;
    u.rec[3 * (u64)i + 2] = lit;
  }
}

// model.cpp:261-305
PF_HD u32 org_lit(const DB& db, u32 l) { return V2L(db.vorg[ABS(l)]) | SIGN(l); }
// D: where the clause is read (the pool or a staged LView); the original
// variable numbers come from the pool's vorg
template <class D>
PF_HD u32 save_clause(const D& view, const DB& db, u32 c, u32 witlit, u32* st) {
  const CL cl = clause(view, c);
  u32 pos = 0;
  for (u32 i = 0; i < cl.n; i++) {
    st[i] = org_lit(db, cl.l[i]);
    if (cl.l[i] == witlit) pos = i;
  }
  if (pos) { const u32 t = st[pos]; st[pos] = st[0]; st[0] = t; }
  st[cl.n] = cl.n;
  return cl.n + 1;
}
PF_HD u32 save_witness(const DB& db, u32 w, u32* st) {
  st[0] = org_lit(db, w);
  st[1] = 1;
  return 2;
}
// writes the eliminated-stack records of toblivion / save_BN_gate
template <class D>
PF_HD void save_side(const D& view, const DB& db, u32 p, u32 pOrgs, u32 nOrgs, u32* st) {
  const u32 n = NEGL(p);
  const bool negside = pOrgs > nOrgs;
  const OLV side = olist(view, negside ? n : p);
  const u32 wit = negside ? n : p;
  u32 w = 0;
  for (u32 i = 0; i < side.n; i++) {
    const u32 c = side.r[i];
    if (is_original(view, c)) w += save_clause(view, db, c, wit, st + w);
  }
  save_witness(db, negside ? p : n, st + w);
}
PF_HD void delete_lists(const DB& db, u32 p) {
  for (u32 q = 0; q < 2; q++) {
    const OLV ol = olist(db, p | q);
    for (u32 i = 0; i < ol.n; i++) set_deleted(db, ol.r[i]);
    db.ot_size[p | q] = 0;
  }
}
// equivalence.h:25-43: substitute def for dx in clause c (kept sorted);
// returns the unit literal when one literal remains, else 0
PF_HD u32 substitute_clause(const DB& db, u32 dx, u32 c, u32 def) {
  u32* l = clits(db, c);
  const u32 n0 = csize(db, c);
  u32 n = 0;
  for (u32 i = 0; i < n0; i++) {
    const u32 x = l[i];
    if (x == dx) l[n++] = def;
    else if (x != def) l[n++] = x;
  }
  set_csize(db, c, n);
  if (n == 1) return l[0];
  for (u32 i = 1; i < n; i++) {
    const u32 t = l[i];
    u32 j = i;
    while (j > 0 && l[j - 1] > t) { l[j] = l[j - 1]; j--; }
    l[j] = t;
  }
  set_csig(db, c, calc_sig(l, n));
  return 0;
}
struct VeApply {
  const u32* cls_base;   // per elected variable: first new clause id - cls0
  const u32* lit_base;   // per elected variable: first literal slot - lit0
  const u32* stack_base; // per elected variable: first stack word - stack0
  u32 cls0, lit0, stack0;
  u32 cls_cap, lit_cap, stack_cap;  // capacities: writes past them are refused
  u32* stack;
  UnitOut units;
  u32* error;  // set when a write would leave its buffer
  ProofOut proof;
};
// bve.cpp:138-241, execution part
// L: where the variable's lists and clauses are read: the pool (DB), or a
// staged LView for the resolvent path (which changes none of them until it
// deletes the lists; the EQU path rewrites its clauses and reads the pool)
template <class L>
PF_HD void ve_apply(const L& lv, const DB& db, u32 v, u32 k, const VeDecision& d, const VeApply& a) {
  const u32 type = d.type[k];
  if (!type) return;
  const u32 p = V2L(v), n = NEGL(p);
  const OLV poss = olist(lv, p), negs = olist(lv, n);
  const u32 pOrgs = count_orgs(lv, poss), nOrgs = count_orgs(lv, negs);
  const u32 sbase = a.stack0 + a.stack_base[k];
  if (sbase + d.nstack[k] > a.stack_cap) {
    at_add(a.error, 1u);
    return;
  }
  u32* st = a.stack + sbase;
  u32 seq = 0;
  if (type == VT_EQU) {
    save_side(db, db, p, pOrgs, nOrgs, st);
    // equivalence.h:45-87
    const u32 def = d.def[k], def_f = FLIP(def);
    for (u32 q = 0; q < 2; q++) {
      const OLV ol = q == 0 ? negs : poss;
      const u32 dx = q == 0 ? n : p;
      const u32 sub = q == 0 ? def_f : def;  // replacement literal
      for (u32 i = 0; i < ol.n; i++) {
        const u32 c = ol.r[i];
        if (is_learnt(db, c) || is_molten(db, c) || has_lit(clause(db, c), FLIP(sub))) set_deleted(db, c);
        else if (is_original(db, c)) {
          const u32 unit = substitute_clause(db, dx, c, sub);
          if (unit) push_unit(a.units, k, seq++, unit);
          else push_proof(a.proof, k, seq++, PR_ADD, c);
        }
      }
    }
  } else {
    // resolvents are produced only when some resolvent is longer than one
    // literal ("if (nAddedCls) xresolve(...)", bve.cpp:226-235): a variable
    // whose resolvents are all units or tautologies is eliminated without
    // emitting those units
    if ((type == VT_RES || type == VT_SUBST) && d.ncls[k]) {
      const bool substitute = type == VT_SUBST;
      u32 dx = p, fx = n;
      if (olist(lv, dx).n > olist(lv, fx).n) { dx = n; fx = p; }
      const OLV me = olist(lv, dx), other = olist(lv, fx);
      u32 cid = a.cls0 + a.cls_base[k], lpos = a.lit0 + a.lit_base[k];
      for (u32 i = 0; i < me.n; i++) {
        const u32 ci = me.r[i];
        if (!is_original(lv, ci)) continue;
        const bool ma = is_molten(lv, ci);
        const CL A = clause(lv, ci);
        for (u32 j = 0; j < other.n; j++) {
          const u32 cj = other.r[j];
          if (!is_original(lv, cj)) continue;
          if (substitute && ma == is_molten(lv, cj)) continue;
          const CL B = clause(lv, cj);
          const u32 len = merge_len(v, A, B);
          if (!len) {
            // tautology, or a unit resolvent of two literals (len counts
            // duplicates once; a length of 0 only arises from a tautology)
            continue;
          }
          if (len == 1) {
            u32 one[2];
            merge_out(v, A, B, one);
            push_unit(a.units, k, seq++, one[0]);
            continue;
          }
          if (cid >= a.cls_cap || lpos + CL_HDR + len > a.lit_cap) {
            at_add(a.error, 1u);
            return;
          }
          u32* h = db.lits + lpos;
          merge_out(v, A, B, h + CL_HDR);
          h[0] = make_hdr(len, F_ADDED);
          h[1] = calc_sig(h + CL_HDR, len);
          h[2] = 0;
          db.c_off[cid] = lpos;
          push_proof(a.proof, k, seq++, PR_ADD, lpos);
          cid++;
          lpos += CL_HDR + len;
        }
      }
    }
    save_side(lv, db, p, pOrgs, nOrgs, st);
    delete_lists(db, p);
  }
  db.vstate[v] = VS_MELTED;
}

// ---------------------------------------------------------------- BCE (elimination.cpp:131-167)
// count == true: returns the stack words; otherwise writes them at 'st'
PF_HD u32 bce_var(const DB& db, const Options& o, u32 v, bool count, u32* st) {
  const u32 p = V2L(v), n = NEGL(p);
  if ((int)db.ot_size[p] > o.bce_limit || (int)db.ot_size[n] > o.bce_limit) return 0;
  const OLV poss = olist(db, p), negs = olist(db, n);
  u32 w = 0;
  for (u32 i = 0; i < negs.n; i++) {
    const u32 ng = negs.r[i];
    if (!is_original(db, ng)) continue;
    const CL A = clause(db, ng);
    bool all = true;
    for (u32 j = 0; j < poss.n && all; j++) {
      const u32 ps = poss.r[j];
      if (is_original(db, ps) && !is_tautology(v, A, clause(db, ps))) all = false;
    }
    if (all) {
      if (count) w += A.n + 1;
      else {
        w += save_clause(db, db, ng, n, st + w);
        set_deleted(db, ng);
      }
    }
  }
  return w;
}

// ---------------------------------------------------------------- ERE (redundancy.h)
constexpr int ERE_MAX_DEL = 32;
struct EreOut {
  u32* rec;       // (elected index, sequence, clause id)
  u32* n;
  u32 cap;
  u32* overflow;  // set when one variable exceeds ERE_MAX_DEL deletions
};
// redundancy.h:27-61
PF_HD bool merge_ere(u32 x, CL c1, CL c2, CL t) {
  u32 i1 = 0, i2 = 0, it = 0;
  while (it < t.n && i1 < c1.n && i2 < c2.n) {
    const u32 l = t.l[it];
    if (ABS(l) == x) return false;
    const u32 l1 = c1.l[i1], l2 = c2.l[i2];
    const u32 v1 = ABS(l1), v2 = ABS(l2);
    if (v1 == x) i1++;
    else if (v2 == x) i2++;
    else if (v1 < v2) { if (l1 != l) return false; i1++; it++; }
    else if (v2 < v1) { if (l2 != l) return false; i2++; it++; }
    else { if (l1 != l) return false; i1++; i2++; it++; }
  }
  while (it < t.n && i1 < c1.n) {
    const u32 l1 = c1.l[i1];
    if (ABS(l1) != x && l1 != t.l[it++]) return false;
    i1++;
  }
  while (it < t.n && i2 < c2.n) {
    const u32 l2 = c2.l[i2];
    if (ABS(l2) != x && l2 != t.l[it++]) return false;
    i2++;
  }
  return true;
}
// redundancy.h:67-121 for one resolvent; 'dl' holds this variable's own
// deletions so far
PF_HD void forward_equ(const DB& db, u32 x, u32 c1, u32 c2, int maxsize, u32* dl, u32& nd, bool& ovf) {
  const CL A = clause(db, c1), B = clause(db, c2);
  u32 len = A.n + B.n - 2;
  u32 i1 = 0, i2 = 0;
  u32 minsize = 0xFFFFFFFFu, msig = 0, best = 0;
  while (i1 < A.n && i2 < B.n) {
    const u32 l1 = A.l[i1], l2 = B.l[i2];
    const u32 v1 = ABS(l1), v2 = ABS(l2);
    u32 lit;
    if (v1 == x) { i1++; continue; }
    else if (v2 == x) { i2++; continue; }
    else if ((l1 ^ l2) == 1u) return;
    else if (v1 < v2) { i1++; lit = l1; }
    else if (v2 < v1) { i2++; lit = l2; }
    else { i1++; i2++; len--; lit = l1; }
    const u32 s = db.ot_size[lit];
    if (s < minsize) { minsize = s; best = lit; }
    msig |= MAPHASH(lit);
  }
  if (!(len > 1 && (!maxsize || (int)len <= maxsize))) return;
  const u8 type = (is_learnt(db, c1) || is_learnt(db, c2)) ? ST_LEARNT : ST_ORIGINAL;
  const OLV ml = olist(db, best);
  u32 found = NOPOS;
  for (u32 i = 0; i < ml.n; i++) {
    const u32 c = ml.r[i];
    if (c >= found) continue;
    const u8 f = cflags(db, c);
    if (csize(db, c) != len) continue;
    if (!(st_learnt(f) || (f & ST_MASK) == type)) continue;
    bool mine = false;
    for (u32 q = 0; q < nd && q < (u32)ERE_MAX_DEL; q++) mine |= (dl[q] == c);
    if (mine) continue;
    if (!sub_sig(msig, csig(db, c))) continue;
    if (!merge_ere(x, A, B, clause(db, c))) continue;
    found = c;
  }
  if (found == NOPOS) return;
  if (nd < (u32)ERE_MAX_DEL) dl[nd] = found;
  else ovf = true;
  nd++;
}
// elimination.cpp:169-197 for elected variable v at index k: records the
// clauses it deletes without applying them
PF_HD void ere_var(const DB& db, const Options& o, u32 v, u32 k, const EreOut& out) {
  u32 dx = V2L(v), fx = NEGL(dx);
  if (db.ot_size[dx] > db.ot_size[fx]) { const u32 t = dx; dx = fx; fx = t; }
  const OLV me = olist(db, dx), other = olist(db, fx);
  if ((int)me.n > o.ere_limit || (int)other.n > o.ere_limit) return;
  u32 dl[ERE_MAX_DEL];
  u32 nd = 0;
  bool ovf = false;
  for (u32 i = 0; i < me.n; i++) {
    const u32 ci = me.r[i];
    if (is_deleted(db, ci)) continue;
    for (u32 j = 0; j < other.n; j++) {
      const u32 cj = other.r[j];
      if (is_deleted(db, cj)) continue;
      forward_equ(db, v, ci, cj, o.ere_max_resolvent, dl, nd, ovf);
    }
  }
  for (u32 q = 0; q < nd && q < (u32)ERE_MAX_DEL; q++) {
    const u32 i = at_add(out.n, 1u);
    if (i < out.cap) {
      out.rec[3 * (u64)i] = k;
      out.rec[3 * (u64)i + 1] = q;
      out.rec[3 * (u64)i + 2] = dl[q];
    }
  }
  if (ovf) at_add(out.overflow, 1u);
}
// a variable's ERE result is exact unless an earlier variable deleted one
// of the clauses it read: its own list entries or one of its targets
PF_HD bool ere_dirty_var(const DB& db, u32 v, u32 k, const u32* del_by) {
  for (u32 q = 0; q < 2; q++) {
    const OLV ol = olist(db, V2L(v) | q);
    for (u32 i = 0; i < ol.n; i++)
      if (del_by[ol.r[i]] < k) return true;
  }
  return false;
}

// ---------------------------------------------------------------- ERE, pair-parallel
// ERE deletes, for each resolvent, the first live clause equal to it in the
// shortest occurrence list of its literals (redundancy.h:67-121).  A clause
// equal to the resolvent contains every resolvent literal, so that list
// holds exactly the live clauses equal to the resolvent, and "first" is the
// lowest clause id (lists of non-elected literals keep clause order).  The
// pair-parallel form therefore looks every resolvent up in a hash table of
// the live clauses, keyed by the literal sequence, and returns all matches;
// the order-dependent part (which match is still live) is resolved
// afterwards over the few pairs that matched, in elected-then-pair order.
PF_HD u32 hash_step(u32 h, u32 lit) { return (h ^ lit) * 0x01000193u; }
PF_HD u32 hash_final(u32 h) {
  h ^= h >> 16;
  h *= 0x85ebca6bu;
  h ^= h >> 13;
  h *= 0xc2b2ae35u;
  h ^= h >> 16;
  return h;
}
constexpr u32 HASH_SEED = 0x811c9dc5u;
constexpr u64 HT_EMPTY = ~0ull;
struct HashTable {
  u64* slot;  // (hash << 32) | clause id, HT_EMPTY when free
  u32 mask;   // capacity - 1 (capacity a power of two, load <= 1/2)
};
PF_HD u64 at_cas64(u64* p, u64 cmp, u64 v) {
#if defined(__HIP_DEVICE_COMPILE__)
  return atomicCAS((unsigned long long*)p, (unsigned long long)cmp, (unsigned long long)v);
#else
  u64 e = cmp;
  __atomic_compare_exchange_n(p, &e, v, false, __ATOMIC_RELAXED, __ATOMIC_RELAXED);
  return e;
#endif
}
PF_HD void ht_insert_clause(const DB& db, const HashTable& ht, u32 c, int maxsize) {
  if (is_deleted(db, c)) return;
  const CL cl = clause(db, c);
  if (cl.n < 2 || (maxsize && (int)cl.n > maxsize)) return;
  u32 h = HASH_SEED;
  for (u32 i = 0; i < cl.n; i++) h = hash_step(h, cl.l[i]);
  h = hash_final(h);
  const u64 e = ((u64)h << 32) | c;
  u32 s = h & ht.mask;
  while (at_cas64(&ht.slot[s], HT_EMPTY, e) != HT_EMPTY) s = (s + 1) & ht.mask;
}
constexpr int ERE_K = 4;        // matches kept per resolvent
constexpr int ERE_REC_WORDS = 5 + ERE_K;  // k, pair, ci, cj, nmatch, ids
struct ErePairs {
  const u32* pv;
  const u32* pbase;  // per elected variable: first pair index (exclusive scan)
  u32 nvars;
  u32* rec;
  u32* nrec;
  u32 cap;
  u32* overflow;  // records with more matches than ERE_K
  u32* ovf;       // their record indices (up to ovf_cap)
  u32 ovf_cap;
  HashTable ht;
  int maxsize;
};
// the lists of elected variable v in ERE order (elimination.cpp:177-182)
PF_HD void ere_lists(const DB& db, u32 v, OLV& me, OLV& other) {
  u32 dx = V2L(v), fx = NEGL(dx);
  if (db.ot_size[dx] > db.ot_size[fx]) { const u32 t = dx; dx = fx; fx = t; }
  me = olist(db, dx);
  other = olist(db, fx);
}
PF_HD u32 ere_pair_count(const DB& db, const Options& o, u32 v) {
  OLV me, other;
  ere_lists(db, v, me, other);
  if ((int)me.n > o.ere_limit || (int)other.n > o.ere_limit) return 0;
  return me.n * other.n;
}
// The resolvent of pair 'local' of elected variable index k and every live
// clause equal to it (found through the hash table): f(c) per match.
// Returns false when the pair has no resolvent to look up (an antecedent
// deleted, tautology, size limits); c1, c2 are the antecedents.
template <class F>
PF_HD bool ere_matches(const DB& db, const ErePairs& a, u32 k, u32 local, u32& c1, u32& c2, F f) {
  const u32 x = a.pv[k];
  OLV me, other;
  ere_lists(db, x, me, other);
  const u32 i = local / other.n, j = local % other.n;
  c1 = me.r[i];
  c2 = other.r[j];
  if (is_deleted(db, c1) || is_deleted(db, c2)) return false;
  const CL A = clause(db, c1), B = clause(db, c2);
  u32 i1 = 0, i2 = 0, len = 0, h = HASH_SEED;
  while (i1 < A.n && i2 < B.n) {
    const u32 l1 = A.l[i1], l2 = B.l[i2];
    const u32 v1 = ABS(l1), v2 = ABS(l2);
    if (v1 == x) i1++;
    else if (v2 == x) i2++;
    else if ((l1 ^ l2) == 1u) return false;
    else if (v1 < v2) { i1++; h = hash_step(h, l1); len++; }
    else if (v2 < v1) { i2++; h = hash_step(h, l2); len++; }
    else { i1++; i2++; h = hash_step(h, l1); len++; }
  }
  while (i1 < A.n) { const u32 l = A.l[i1++]; if (ABS(l) != x) { h = hash_step(h, l); len++; } }
  while (i2 < B.n) { const u32 l = B.l[i2++]; if (ABS(l) != x) { h = hash_step(h, l); len++; } }
  if (!(len > 1 && (!a.maxsize || (int)len <= a.maxsize))) return false;
  h = hash_final(h);
  const u8 type = (is_learnt(db, c1) || is_learnt(db, c2)) ? ST_LEARNT : ST_ORIGINAL;
  for (u32 s = h & a.ht.mask;; s = (s + 1) & a.ht.mask) {
    const u64 e = a.ht.slot[s];
    if (e == HT_EMPTY) break;
    if ((u32)(e >> 32) != h) continue;
    const u32 c = (u32)e;
    const u8 fl = cflags(db, c);
    if (st_deleted(fl) || csize(db, c) != len) continue;
    if (!(st_learnt(fl) || (fl & ST_MASK) == type)) continue;
    if (!merge_ere(x, A, B, clause(db, c))) continue;
    f(c);
  }
  return true;
}
// one resolvent (pair 'local' of elected variable index k): a record with
// its ERE_K lowest matching clause ids; a pair with more matches is listed
// in ovf, whose full match lists ere_all_matches writes afterwards
PF_HD void ere_pair(const DB& db, const ErePairs& a, u32 k, u32 local) {
  u32 ids[ERE_K];
  u32 nm = 0, c1, c2;
  ere_matches(db, a, k, local, c1, c2, [&](u32 c) {
    // keep the ERE_K lowest ids, ascending
    u32 ins = c;
    for (u32 q = 0; q < nm && q < (u32)ERE_K; q++) {
      if (ins < ids[q]) { const u32 t = ids[q]; ids[q] = ins; ins = t; }
    }
    if (nm < (u32)ERE_K) ids[nm] = ins;
    nm++;
  });
  if (!nm) return;
  const u32 r = at_add(a.nrec, 1u);
  if (r >= a.cap) return;
  u32* w = a.rec + (u64)r * ERE_REC_WORDS;
  w[0] = k;
  w[1] = local;
  w[2] = c1;
  w[3] = c2;
  w[4] = nm;
  for (u32 q = 0; q < (u32)ERE_K; q++) w[5 + q] = q < nm ? ids[q] : 0xFFFFFFFFu;
  if (nm > (u32)ERE_K) {
    const u32 o = at_add(a.overflow, 1u);
    if (o < a.ovf_cap) a.ovf[o] = r;
  }
}
// every match of overflow record ovf[o] into out[off[o] ..) (off[o] + nm
// words: the record's count)
PF_HD void ere_all_matches(const DB& db, const ErePairs& a, u32 o, const u32* off, u32* out) {
  const u32* w = a.rec + (u64)a.ovf[o] * ERE_REC_WORDS;
  u32 c1, c2, n = 0;
  const u32 base = off[o], nm = w[4];
  ere_matches(db, a, w[0], w[1], c1, c2, [&](u32 c) {
    if (n < nm) out[base + n] = c;
    n++;
  });
}

// ---------------------------------------------------------------- variable map
// Compact renumbering after a simplify call (VMAP, src/cpu/vmap.h:26-78):
// variables in state 0 keep their order and become 1..n; the first
// root-assigned variable ('first') is kept as well and stands for every
// root-assigned literal.
struct VarMap {
  const u32* mapped;  // old variable -> new variable, 0 when dropped
  const u32* assign;  // root assignments in the old numbering (true literal + 1)
  u32 first;          // first root-assigned old variable, 0 when there is none
  u32 first_true;     // the new literal of 'first' that is true at the root
};
// variable v survives the map (VMAP::initiate, vmap.h:63-77)
PF_HD u32 map_keep(const u8* vstate, u32 v, u32 first) {
  return (v && (vstate[v] == 0 || v == first)) ? 1u : 0u;
}
// VMAP::mapLit (vmap.h:45-60): a surviving variable keeps its sign; a
// root-assigned literal becomes the literal of 'first' with the same value;
// a dropped unassigned literal becomes 0
PF_HD u32 map_lit(const VarMap& m, u32 lit) {
  const u32 v = ABS(lit), nv = m.mapped[v];
  if (nv && v != m.first) return V2L(nv) | SIGN(lit);
  const u32 a = m.assign[v];
  if (!a || !m.first) return 0;
  return (lit == a - 1) ? m.first_true : FLIP(m.first_true);
}
// one clause in place: literals and signature (VMAP::mapClause, vmap.h:183-193).
// Clause literals are unassigned and the map keeps the variable order, so
// sorted literals stay sorted.
PF_HD void map_clause(const DB& db, u32 c, const VarMap& m) {
  if (is_deleted(db, c)) return;
  u32* l = clits(db, c);
  const u32 n = csize(db, c);
  u32 sig = 0;
  for (u32 q = 0; q < n; q++) {
    const u32 x = map_lit(m, l[q]);
    l[q] = x;
    sig |= MAPHASH(x);
  }
  set_csig(db, c, sig);
}
// output packing of clause index i (the host write-back, src/cpu/sclause.cpp:22-57):
// literals at start[i], flags bit0 learnt, bit1 added, bits 4-5 usage
PF_HD void pack_clause(const DB& db, u32 i, const u32* start, u64* offsets, u32* lits, u8* flags, u32* lbd) {
  const u32 c = db.c_off[i];
  const u32 n = csize(db, c), s = start[i];
  const u8 f = cflags(db, c);
  const u32* src = clits(db, c);
  offsets[i] = s;
  for (u32 q = 0; q < n; q++) lits[s + q] = src[q];
  flags[i] = (u8)((st_learnt(f) ? 1 : 0) | ((f & F_ADDED) ? 2 : 0) | (f & USAGE_MASK));
  lbd[i] = st_learnt(f) ? clbd(db, c) : 0u;
  if (i + 1 == db.ncls) offsets[db.ncls] = (u64)s + n;
}

// ---------------------------------------------------------------- DIMACS ingestion
PF_HD bool lex_blank(u32 c) { return (c >= 9 && c <= 13) || c == 32; }  // isSpace, dimacs.h:33
PF_HD bool lex_digit(u32 c) { return c >= '0' && c <= '9'; }
// byte i (value c, byte before it 'prev') starts a token of the clause part [begin, end)
PF_HD bool lex_start(u64 i, u32 c, u32 prev, u64 begin, u64 end) {
  return i >= begin && i < end && !lex_blank(c) && (i == begin || lex_blank(prev));
}
// toInteger (dimacs.h:101-111) for the token at byte i: the literal
// 2 v + sign (0 ends a clause); bad: not a signed number followed by a
// blank; big: v above maxvar
PF_HD u32 lex_literal(const u8* text, u64 i, u64 end, u32 maxvar, bool& bad, bool& big) {
  u64 p = i;
  u32 sign = 0, c = text[p];
  if (c == '-' || c == '+') sign = c == '-', p++;
  u64 v = 0;
  u32 nd = 0;
  while (p < end && lex_digit(c = text[p])) {
    v = v * 10 + (c - '0');
    if (v > 0x7FFFFFFFull) v = 0x80000000ull;  // saturates: reported as above maxvar
    p++, nd++;
  }
  bad = nd == 0 || (p < end && !lex_blank(text[p]));
  big = !bad && v > maxvar;
  return bad || !v ? 0u : (u32)(2 * v + sign);
}
// clause i of a parse from the literals between the (i-1)-th and the i-th 0:
// in the pool it starts at its literal offset plus two words per earlier
// clause; returns false when it is too long for the header
PF_HD bool parse_build_clause(const DB& db, const u32* lits, const u32* ends, u32 i) {
  const u32 s = i ? ends[i - 1] + 1 : 0u, n = ends[i] - s, off = s + 2 * i;
  u32* h = db.lits + off;
  const u32 m = n < CL_MAX_SIZE ? n : CL_MAX_SIZE;
  h[0] = make_hdr(m, ST_ORIGINAL);
  h[1] = 0;
  h[2] = 0;
  for (u32 q = 0; q < m; q++) h[CL_HDR + q] = lits[s + q];
  db.c_off[i] = off;
  return n <= CL_MAX_SIZE;
}
// a parsed clause with sorted literals (toClause, dimacs.cpp:113-170):
// duplicates dropped; a tautology is removed, a unit clause becomes the
// unit (*unit) and is removed too.  Returns 0 kept, 1 tautology, 2 empty, 3 unit.
PF_HD u32 parse_dedup_clause(const DB& db, u32 c, u32* unit) {
  u32* l = clits(db, c);
  const u32 n = csize(db, c);
  u32 m = 0, prev = 0;
  bool taut = false;
  for (u32 q = 0; q < n; q++) {
    const u32 x = l[q];
    if (m) {
      if (x == prev) continue;
      if ((prev ^ 1u) == x) taut = true;
    }
    if (m != q) l[m] = x;  // a literal moves only after a dropped duplicate
    m++;
    prev = x;
  }
  *unit = 0;
  if (taut || m < 2) {
    if (m == 1 && !taut) *unit = prev;
    set_cflags(db, c, ST_DELETED);
    return taut ? 1u : (m == 0 ? 2u : 3u);
  }
  // dropping duplicates keeps the signature (the same set of literals)
  if (m != n) set_csize(db, c, m);
  return 0;
}

// DRAT lines of the ingestion (the parser's deletions of satisfied clauses
// and its additions of shrunken ones, dimacs.cpp:245-305; the top-level
// shrink's, shrink.cpp:122), written from the end state: per text clause i
// (tokens toks[s..ends[i]) with s = ends[i - 1] + 1) whose pool clause is
// gone, a deletion of the text clause; per text clause that lost literals
// (duplicates, literals false at the root), an addition of what is left and
// the deletion of the text clause.  Bytes of the lines; written to o when
// not null.
PF_HD u32 parse_proof_clause(const DB& db, const u32* toks, const u32* ends, u32 i, u8* o) {
  const u32 s = i ? ends[i - 1] + 1 : 0u, n = ends[i] - s;
  const u32 c = db.c_off[i];
  const u32 h = db.lits[c];
  const bool gone = st_deleted(hdr_flags(h));
  const u32 m = hdr_size(h);
  if (!gone && m >= n) return 0;
  u32 b = 0;
  if (!gone) {
    const u32* l = clits(db, c);
    b += 2;
    for (u32 q = 0; q < m; q++) b += varint_bytes(proof_lit_value(db, l[q]));
    if (o) {
      *o++ = 'a';
      for (u32 q = 0; q < m; q++) o = put_varint(o, proof_lit_value(db, l[q]));
      *o++ = 0;
    }
  }
  b += 2;
  for (u32 q = 0; q < n; q++) b += varint_bytes(proof_lit_value(db, toks[s + q]));
  if (o) {
    *o++ = 'd';
    for (u32 q = 0; q < n; q++) o = put_varint(o, proof_lit_value(db, toks[s + q]));
    *o = 0;
  }
  return b;
}
// a root unit as a line of its own ('a' u)
PF_HD u32 unit_proof_line(const DB& db, u32 u, u8* o) {
  const u32 v = proof_lit_value(db, u);
  if (o) {
    *o++ = 'a';
    o = put_varint(o, v);
    *o = 0;
  }
  return 2 + varint_bytes(v);
}

// DRAT lines of a call that rewrites clauses in place (equivalent-literal
// substitution and its top-level shrink), from the clause at handle c
// before the call (pool 'old') and now: a clause gone now is deleted; a
// clause whose literals changed gets its new form added (phase 0) and its
// old form deleted (phase 1).  Bytes of the phase's line; written to o
// when not null.
PF_HD u32 rewrite_proof_clause(const DB& db, const u32* old, u32 c, int phase, u8* o) {
  const u32 ho = old[c], hn = db.lits[c];
  if (st_deleted(hdr_flags(ho))) return 0;
  const u32 no = hdr_size(ho), nn = hdr_size(hn);
  const u32* lo = old + c + CL_HDR;
  const u32* ln = db.lits + c + CL_HDR;
  const bool gone = st_deleted(hdr_flags(hn));
  if (!gone) {
    bool same = no == nn;
    for (u32 q = 0; same && q < no; q++) same = lo[q] == ln[q];
    if (same) return 0;
  }
  if (phase == 0 && gone) return 0;
  const u32* l = phase == 0 ? ln : lo;
  const u32 n = phase == 0 ? nn : no;
  u32 b = 2;
  for (u32 q = 0; q < n; q++) b += varint_bytes(proof_lit_value(db, l[q]));
  if (o) {
    *o++ = phase == 0 ? 'a' : 'd';
    for (u32 q = 0; q < n; q++) o = put_varint(o, proof_lit_value(db, l[q]));
    *o = 0;
  }
  return b;
}

// ---------------------------------------------------------------- ELS
// Equivalent-literal substitution (Solver::decompose/substitute,
// src/cpu/equivalence.cpp:41-306).  rep[l]: the representative of literal
// l (its SCC's literal of smallest variable; l itself for assumed
// variables; 0 for inactive literals, as the reference's fresh array).
// The substitution pass runs over the clauses in the order learnts, then
// originals; a unit it finds at position p is visible to the clauses after
// p: utime[x] = 2 p + sign of that unit, NOPOS when there is none.
struct ElsView {
  const u32* rep;
  const u32* utime;
};
// 1 true, 0 false, -1 unassigned at position p
PF_HD int els_value(const DB& db, const ElsView& e, u32 lit, u32 p) {
  const u32 a = db.assign[ABS(lit)];
  if (a) return (a - 1) == lit ? 1 : 0;
  const u32 t = e.utime[ABS(lit)];
  if (t != NOPOS && (t >> 1) < p) return (t & 1u) == SIGN(lit) ? 1 : 0;
  return -1;
}
// one clause at position p (equivalence.cpp:214-256).  Returns 0 untouched
// (every literal is its own representative), 1 satisfied or tautological,
// 2 empty, 3 unit (*unit), 4 rewritten with *m >= 2 literals.  With write,
// the kept literals (mapped, duplicates dropped, first occurrence order)
// are compacted in place, sorted, and size and signature updated; satisfied
// and unit clauses are marked deleted.  Without write the clause is only
// classified (an earlier kept literal is found again from the original
// literals).
PF_HD u32 els_clause(const DB& db, u32 c, const ElsView& e, u32 p, bool write, u32* unit, u32* m_out) {
  u32* l = clits(db, c);
  const u32 n = csize(db, c);
  bool touched = false;
  for (u32 q = 0; q < n && !touched; q++) touched = e.rep[l[q]] != l[q];
  if (!touched) return 0;
  u32 m = 0, first = 0;
  bool sat = false;
  for (u32 q = 0; q < n && !sat; q++) {
    const u32 lit = l[q];
    const int x = els_value(db, e, lit, p);
    if (x > 0) sat = true;
    if (x != -1) continue;
    const u32 other = e.rep[lit];
    const int y = els_value(db, e, other, p);
    if (y > 0) sat = true;
    if (y != -1) continue;
    // the marks of equivalence.cpp:231-236: same variable kept before?
    int seen = -1;
    if (write) {
      for (u32 r = 0; r < m && seen < 0; r++)
        if (ABS(l[r]) == ABS(other)) seen = l[r] == other ? 0 : 1;
    } else {
      for (u32 r = 0; r < q && seen < 0; r++) {
        const u32 lr = l[r];
        if (els_value(db, e, lr, p) != -1) continue;
        const u32 orr = e.rep[lr];
        if (els_value(db, e, orr, p) != -1) continue;
        if (ABS(orr) == ABS(other)) seen = orr == other ? 0 : 1;
      }
    }
    if (seen == 1) sat = true;
    if (seen != -1) continue;
    if (write) l[m] = other;
    if (!m) first = other;
    m++;
  }
  *m_out = m;
  if (sat) {
    if (write) set_deleted(db, c);
    return 1;
  }
  if (!m) return 2;
  if (m == 1) {
    *unit = first;
    if (write) set_deleted(db, c);
    return 3;
  }
  if (write) {
    for (u32 q = 1; q < m; q++) {
      const u32 t = l[q];
      u32 j = q;
      while (j > 0 && l[j - 1] > t) l[j] = l[j - 1], j--;
      l[j] = t;
    }
    set_csize(db, c, m);
    set_csig(db, c, calc_sig(l, m));
  }
  return 4;
}

}  // namespace pfs
