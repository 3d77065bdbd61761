// Seeded synthetic CNF generators for the simplifier workloads, plus the
// root-level preparation a host solver performs before calling the
// simplifier (unit propagation, removal of satisfied clauses and of false
// literals; the reference does this in its parser, root BCP and
// rootify/shrinkTop before extraction, src/cpu/dimacs.cpp:220-277,
// src/cpu/simplify.cpp:163-164).
//
// RNG: SplitMix64 seeds a xoshiro256** state; every generator consumes
// the stream in a fixed order so that the same seed always yields the
// same formula.  Literals use the simplifier encoding 2*v + sign.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

namespace {

struct Rng {
  uint64_t s[4];
  static uint64_t splitmix(uint64_t& x) {
    uint64_t z = (x += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
  }
  explicit Rng(uint64_t seed) {
    uint64_t x = seed;
    for (int i = 0; i < 4; i++) s[i] = splitmix(x);
  }
  static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  uint64_t next() {
    const uint64_t r = rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return r;
  }
  // uniform integer in [0, n) by the multiply-high method
  uint64_t below(uint64_t n) { return (uint64_t)(((__uint128_t)next() * n) >> 64); }
  uint32_t bit() { return (uint32_t)(next() >> 63); }
  double unit() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

struct Cnf {
  uint32_t maxvar = 0;
  std::vector<uint64_t> off{0};
  std::vector<uint32_t> lits;
  void add(const uint32_t* l, int n) {
    lits.insert(lits.end(), l, l + n);
    off.push_back(lits.size());
  }
};

inline uint32_t mk(uint32_t v, uint32_t s) { return (v << 1) | s; }

}  // namespace

extern "C" {

typedef struct {
  uint32_t maxvar;
  uint64_t nclauses;
  uint64_t nlits;
  uint64_t* offsets;  // nclauses + 1
  uint32_t* lits;     // 2*v + sign
} pfg_cnf;

static int export_cnf(Cnf& c, pfg_cnf* out) {
  out->maxvar = c.maxvar;
  out->nclauses = c.off.size() - 1;
  out->nlits = c.lits.size();
  out->offsets = (uint64_t*)malloc(sizeof(uint64_t) * c.off.size());
  out->lits = (uint32_t*)malloc(sizeof(uint32_t) * (c.lits.size() ? c.lits.size() : 1));
  if (!out->offsets || !out->lits) return -1;
  memcpy(out->offsets, c.off.data(), sizeof(uint64_t) * c.off.size());
  if (!c.lits.empty()) memcpy(out->lits, c.lits.data(), sizeof(uint32_t) * c.lits.size());
  return 0;
}

void pfg_free(pfg_cnf* c) {
  if (!c) return;
  free(c->offsets);
  free(c->lits);
  c->offsets = nullptr;
  c->lits = nullptr;
}

// Uniform random k-SAT: each clause has k distinct variables drawn
// uniformly (rejection), each sign Bernoulli(0.5).
int pfg_uniform_ksat(uint64_t seed, uint32_t nvars, uint64_t nclauses, uint32_t k, pfg_cnf* out) {
  if (!out || nvars < k || k == 0 || k > 64) return -2;
  Rng r(seed);
  Cnf c;
  c.maxvar = nvars;
  c.lits.reserve(nclauses * k);
  c.off.reserve(nclauses + 1);
  uint32_t vs[64], ls[64];
  for (uint64_t i = 0; i < nclauses; i++) {
    for (uint32_t j = 0; j < k; j++) {
      uint32_t v;
      bool dup;
      do {
        v = 1 + (uint32_t)r.below(nvars);
        dup = false;
        for (uint32_t q = 0; q < j; q++) dup |= (vs[q] == v);
      } while (dup);
      vs[j] = v;
      ls[j] = mk(v, r.bit());
    }
    c.add(ls, k);
  }
  return export_cnf(c, out);
}

// Tseitin-encoded random circuit.  Inputs are variables 1..ninputs; gate g
// (1-based) defines variable ninputs+g.  Gate type: AND with probability
// pand, XOR with probability pxor, otherwise ITE.  Fan-ins are distinct
// earlier nodes chosen uniformly, each with a random sign.  Encodings:
//   AND o = a & b      : (-o a) (-o b) (o -a -b)
//   XOR o = a ^ b      : (-o a b) (-o -a -b) (o -a b) (o a -b)
//   ITE o = c ? a : b  : (-o -c a) (-o c b) (o -c -a) (o c -b)
// Finally noutputs distinct gate outputs are asserted as unit clauses
// (appended after all gate clauses).  With sat_outputs = 0 each output gets
// a random sign; with sat_outputs = 1 the inputs get random values, the
// circuit is simulated, and each output is asserted with its simulated
// value, so the formula is satisfiable by construction.
int pfg_tseitin(uint64_t seed, uint32_t ninputs, uint32_t ngates, uint32_t noutputs,
                double pand, double pxor, int sat_outputs, pfg_cnf* out) {
  if (!out || ninputs < 3 || noutputs > ngates) return -2;
  Rng r(seed);
  Cnf c;
  c.maxvar = ninputs + ngates;
  c.lits.reserve((uint64_t)ngates * 10 + noutputs);
  c.off.reserve((uint64_t)ngates * 4 + noutputs + 1);
  std::vector<uint8_t> gtype(sat_outputs ? ngates + 1 : 0);
  std::vector<uint32_t> gin(sat_outputs ? 3 * ((size_t)ngates + 1) : 0);
  for (uint32_t g = 1; g <= ngates; g++) {
    const uint32_t o = ninputs + g;
    const uint32_t earlier = o - 1;
    const double u = r.unit();
    const int type = u < pand ? 0 : (u < pand + pxor ? 1 : 2);
    const int nin = type == 2 ? 3 : 2;
    uint32_t fv[3], fl[3];
    for (int j = 0; j < nin; j++) {
      uint32_t v;
      bool dup;
      do {
        v = 1 + (uint32_t)r.below(earlier);
        dup = false;
        for (int q = 0; q < j; q++) dup |= (fv[q] == v);
      } while (dup);
      fv[j] = v;
      fl[j] = mk(v, r.bit());
    }
    if (sat_outputs) {
      gtype[g] = (uint8_t)type;
      for (int j = 0; j < nin; j++) gin[3 * (size_t)g + j] = fl[j];
    }
    const uint32_t po = mk(o, 0), no = mk(o, 1);
    if (type == 0) {
      const uint32_t a = fl[0], b = fl[1];
      uint32_t c1[2] = {no, a}, c2[2] = {no, b}, c3[3] = {po, a ^ 1, b ^ 1};
      c.add(c1, 2); c.add(c2, 2); c.add(c3, 3);
    } else if (type == 1) {
      const uint32_t a = fl[0], b = fl[1];
      uint32_t c1[3] = {no, a, b}, c2[3] = {no, a ^ 1, b ^ 1};
      uint32_t c3[3] = {po, a ^ 1, b}, c4[3] = {po, a, b ^ 1};
      c.add(c1, 3); c.add(c2, 3); c.add(c3, 3); c.add(c4, 3);
    } else {
      const uint32_t s = fl[0], a = fl[1], b = fl[2];
      uint32_t c1[3] = {no, s ^ 1, a}, c2[3] = {no, s, b};
      uint32_t c3[3] = {po, s ^ 1, a ^ 1}, c4[3] = {po, s, b ^ 1};
      c.add(c1, 3); c.add(c2, 3); c.add(c3, 3); c.add(c4, 3);
    }
  }
  // simulated values (1 = true) of every node under random inputs
  std::vector<uint8_t> val(sat_outputs ? (size_t)ninputs + ngates + 1 : 0);
  if (sat_outputs) {
    for (uint32_t v = 1; v <= ninputs; v++) val[v] = (uint8_t)r.bit();
    auto lv = [&](uint32_t l) { return (uint8_t)(val[l >> 1] ^ (l & 1)); };
    for (uint32_t g = 1; g <= ngates; g++) {
      const uint32_t* f = &gin[3 * (size_t)g];
      uint8_t x;
      if (gtype[g] == 0) x = lv(f[0]) & lv(f[1]);
      else if (gtype[g] == 1) x = lv(f[0]) ^ lv(f[1]);
      else x = lv(f[0]) ? lv(f[1]) : lv(f[2]);
      val[ninputs + g] = x;
    }
  }
  // distinct outputs by partial Fisher-Yates over the gate range
  std::vector<uint32_t> pick(ngates);
  for (uint32_t i = 0; i < ngates; i++) pick[i] = ninputs + 1 + i;
  for (uint32_t i = 0; i < noutputs; i++) {
    const uint32_t j = i + (uint32_t)r.below(ngates - i);
    std::swap(pick[i], pick[j]);
    const uint32_t sign = sat_outputs ? (uint32_t)(val[pick[i]] ^ 1u) : r.bit();
    uint32_t u = mk(pick[i], sign);
    c.add(&u, 1);
  }
  return export_cnf(c, out);
}

// Heavy-tailed k-SAT: clause length 2 + G, G ~ Geometric(1/2) on
// {0,1,...}, capped at maxlen.  Variables are drawn with weight
// rank^(-alpha) over a seeded random permutation of 1..nvars, by inverting
// the continuous power-law CDF on [1, nvars+1).  Duplicate variables in a
// clause are dropped (the first sign is kept); clauses left with fewer
// than 2 literals are redrawn.
int pfg_heavytail(uint64_t seed, uint32_t nvars, uint64_t nclauses, double alpha,
                  uint32_t maxlen, pfg_cnf* out) {
  if (!out || nvars < 2 || maxlen < 2 || maxlen > 64 || alpha <= 0.0 || alpha == 1.0) return -2;
  Rng r(seed);
  Cnf c;
  c.maxvar = nvars;
  std::vector<uint32_t> perm(nvars);
  for (uint32_t i = 0; i < nvars; i++) perm[i] = i + 1;
  for (uint32_t i = nvars - 1; i > 0; i--) {
    const uint32_t j = (uint32_t)r.below((uint64_t)i + 1);
    std::swap(perm[i], perm[j]);
  }
  const double e = 1.0 - alpha;
  const double top = std::pow((double)nvars + 1.0, e) - 1.0;
  c.lits.reserve(nclauses * 3);
  c.off.reserve(nclauses + 1);
  uint32_t vs[64], ls[64];
  for (uint64_t i = 0; i < nclauses; i++) {
    int n;
    do {
      const uint64_t w = r.next();
      uint32_t g = w ? (uint32_t)__builtin_ctzll(w) : 64u;
      uint32_t len = 2 + std::min<uint32_t>(g, maxlen - 2);
      n = 0;
      for (uint32_t j = 0; j < len; j++) {
        const double x = std::pow(1.0 + r.unit() * top, 1.0 / e);
        uint64_t rank = (uint64_t)x;
        if (rank < 1) rank = 1;
        if (rank > nvars) rank = nvars;
        const uint32_t v = perm[rank - 1];
        const uint32_t s = r.bit();
        bool dup = false;
        for (int q = 0; q < n; q++) dup |= (vs[q] == v);
        if (dup) continue;
        vs[n] = v;
        ls[n] = mk(v, s);
        n++;
      }
    } while (n < 2);
    c.add(ls, n);
  }
  return export_cnf(c, out);
}

// Root-level preparation: propagate unit clauses to a fixpoint, drop
// satisfied clauses and tautologies, remove false and duplicate literals,
// keep the surviving clauses in their input order.  'values' receives
// maxvar+1 entries (-1 unassigned, 0 false, 1 true for the positive
// literal) and 'trail' the assigned literals in propagation order.
// Returns 0, or 1 if the formula is refuted at the root.
int pfg_prepare(const pfg_cnf* in, pfg_cnf* out, int8_t* values, uint32_t* trail,
                uint32_t* ntrail) {
  const uint32_t V = in->maxvar;
  std::vector<int8_t> val(2 * (size_t)V + 2, -1);
  std::vector<uint32_t> tr;
  std::vector<std::vector<uint32_t>> occ(2 * (size_t)V + 2);
  std::vector<uint8_t> mark(2 * (size_t)V + 2, 0);
  // clean clauses (duplicates / tautologies) and collect units
  Cnf c0;
  c0.maxvar = V;
  std::vector<uint32_t> tmp;
  bool conflict = false;
  auto assign = [&](uint32_t l) {
    if (val[l] == 1) return true;
    if (val[l] == 0) return false;
    val[l] = 1;
    val[l ^ 1] = 0;
    tr.push_back(l);
    return true;
  };
  for (uint64_t i = 0; i < in->nclauses; i++) {
    tmp.clear();
    bool taut = false;
    for (uint64_t k = in->offsets[i]; k < in->offsets[i + 1]; k++) {
      const uint32_t l = in->lits[k];
      if (mark[l]) continue;
      if (mark[l ^ 1]) { taut = true; }
      mark[l] = 1;
      tmp.push_back(l);
    }
    for (uint32_t l : tmp) mark[l] = 0;
    if (taut) continue;
    if (tmp.empty()) { conflict = true; continue; }
    if (tmp.size() == 1) {
      if (!assign(tmp[0])) conflict = true;
      continue;
    }
    c0.add(tmp.data(), (int)tmp.size());
  }
  const uint64_t n0 = c0.off.size() - 1;
  for (uint64_t i = 0; i < n0; i++)
    for (uint64_t k = c0.off[i]; k < c0.off[i + 1]; k++) occ[c0.lits[k]].push_back((uint32_t)i);
  std::vector<uint8_t> sat(n0, 0);
  size_t head = 0;
  while (!conflict && head < tr.size()) {
    const uint32_t t = tr[head++];
    for (uint32_t ci : occ[t]) sat[ci] = 1;
    for (uint32_t ci : occ[t ^ 1]) {
      if (sat[ci]) continue;
      int nun = 0;
      uint32_t last = 0;
      bool s = false;
      for (uint64_t k = c0.off[ci]; k < c0.off[ci + 1]; k++) {
        const uint32_t l = c0.lits[k];
        if (val[l] == 1) { s = true; break; }
        if (val[l] == -1) { nun++; last = l; }
      }
      if (s) { sat[ci] = 1; continue; }
      if (nun == 0) { conflict = true; break; }
      if (nun == 1) { assign(last); sat[ci] = 1; }
    }
  }
  if (conflict) {
    Cnf e;
    e.maxvar = V;
    export_cnf(e, out);
    return 1;
  }
  Cnf c;
  c.maxvar = V;
  for (uint64_t i = 0; i < n0; i++) {
    tmp.clear();
    bool s = false;
    for (uint64_t k = c0.off[i]; k < c0.off[i + 1]; k++) {
      const uint32_t l = c0.lits[k];
      if (val[l] == 1) { s = true; break; }
      if (val[l] == -1) tmp.push_back(l);
    }
    if (s) continue;
    c.add(tmp.data(), (int)tmp.size());
  }
  if (values)
    for (uint32_t v = 0; v <= V; v++) values[v] = v ? val[2 * (size_t)v] : -1;
  if (trail) memcpy(trail, tr.data(), tr.size() * sizeof(uint32_t));
  if (ntrail) *ntrail = (uint32_t)tr.size();
  return export_cnf(c, out);
}

// DIMACS writer (1-based signed integers).
int pfg_write_dimacs(const pfg_cnf* c, const char* path) {
  FILE* f = fopen(path, "wb");
  if (!f) return -1;
  fprintf(f, "p cnf %u %llu\n", c->maxvar, (unsigned long long)c->nclauses);
  std::vector<char> buf;
  buf.reserve(1 << 20);
  char tmp[32];
  for (uint64_t i = 0; i < c->nclauses; i++) {
    for (uint64_t k = c->offsets[i]; k < c->offsets[i + 1]; k++) {
      const uint32_t l = c->lits[k];
      const int n = snprintf(tmp, sizeof tmp, "%s%u ", (l & 1) ? "-" : "", l >> 1);
      buf.insert(buf.end(), tmp, tmp + n);
    }
    buf.push_back('0');
    buf.push_back('\n');
    if (buf.size() > (1u << 20)) {
      fwrite(buf.data(), 1, buf.size(), f);
      buf.clear();
    }
  }
  fwrite(buf.data(), 1, buf.size(), f);
  fclose(f);
  return 0;
}

}  // extern "C"
