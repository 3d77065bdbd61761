// extern "C" entry points of libpfsimp.so (declared in include/pfsimp.h).
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <chrono>
#include <atomic>
#include <exception>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../include/pfsimp.h"
#include "pfs_driver.hpp"

using namespace pfs;

struct pfs_ctx {
  std::unique_ptr<Simplifier> s;
  std::string err;
  bool uploaded = false, simplified = false;
  int result = PFS_RESULT_UNSOLVED;
  // options as given (before the normalisation of options.cpp:280-311)
  int32_t phases = 5, mu_pos = 32, mu_neg = 32, lits_min = 500, ve_clause_limit = 100,
          xor_max_arity = 10, sub_limit = 3000, bce_limit = 3000, ere_limit = 3000,
          ere_max_resolvent = 1000, lcve_min = 1, shrink_rate = 2, lbd_tier1 = 2;
  bool ve = true, ve_plus = true, sub = true, bce = false, ere = true, all = false, velbound = false,
       profile = false, work_order = true, proof = false, aggr_sort = false, sync_always = false;
  // limits of the GPU build's election (electionsmax, lcveclausemax); off by
  // default so that the election is the CPU simplifier's
  int32_t lcve_max_occurs = INT32_MAX, lcve_clause_max = INT32_MAX;
  double lits_mul = 1.0;  // literalsmul: resolvent room at upload, as a multiple of the literals
  double map_perc = 0.2;  // mapperc, src/cpu/options.cpp:146
  int32_t decompose_min = 2;  // decomposemin, src/cpu/options.cpp:87
  // launch shaping of the occurrence-table build (no reference counterpart):
  // log2 of the clauses per tile (0: planned from the literal ranges) and
  // the tile-to-XCD mapping (consecutive tiles on one XCD)
  int32_t otp_tile = 0;
  bool otp_xcd = true;
  bool otp_ranges = true;  // the occurrence table by literal ranges (off: one atomic per occurrence)
  int32_t mis_lds = -1;  // LDS reserved per election-step workgroup (occupancy; -1: tuned, else 10240)
  int32_t mis_lanes = 0;  // lanes per election candidate (0: tuned, else 1)
  bool autotune = true;  // launch-shape tuner (Options::autotune)
};

namespace {

thread_local std::string g_create_error;

// src/cpu/options.cpp:280-311
Options normalise(const pfs_ctx* c) {
  Options o;
  o.all_en = c->all;
  o.ve_en = c->ve || c->ve_plus;
  o.ve_plus_en = c->ve_plus;
  o.ve_lbound_en = c->velbound;
  o.ve_clause_limit = c->ve_clause_limit;
  o.xor_max_arity = c->xor_max_arity;
  o.bce_en = c->bce;
  o.bce_limit = c->bce_limit;
  o.ere_en = c->ere;
  o.ere_limit = c->ere_limit;
  o.ere_max_resolvent = c->ere_max_resolvent;
  o.sub_en = c->sub;
  o.sub_limit = c->sub_limit;
  o.lcve_min = c->lcve_min;
  o.lits_min = c->lits_min;
  o.mu_pos = c->mu_pos;
  o.mu_neg = c->mu_neg;
  o.phases = c->phases;
  o.shrink_rate = c->shrink_rate;
  o.lbd_tier1 = c->lbd_tier1;
  o.work_order = c->work_order;
  o.proof_en = c->proof;
  o.lcve_max_occurs = c->lcve_max_occurs;
  o.lcve_clause_max = c->lcve_clause_max;
  o.aggr_sort = c->aggr_sort;
  o.sync_always = c->sync_always;
  o.otp_tile = c->otp_tile;
  o.otp_xcd = c->otp_xcd;
  o.otp_ranges = c->otp_ranges;
  o.mis_lds = c->mis_lds;
  o.mis_lanes = c->mis_lanes;
  o.autotune = c->autotune;
  if (o.all_en) o.ve_en = 1, o.ve_plus_en = 1, o.bce_en = 1, o.ere_en = 1;
  if (!o.phases && (o.ve_en || o.sub_en || o.bce_en)) o.phases = 1;
  if (o.phases && !(o.ve_en || o.sub_en || o.bce_en)) o.phases = 0;
  if (o.phases > 1 && !o.ve_en) o.phases = 1;
  return o;
}

bool parse_bool(const char* v, bool& out) {
  if (!v || !*v || !strcmp(v, "1") || !strcmp(v, "true") || !strcmp(v, "on")) { out = true; return true; }
  if (!strcmp(v, "0") || !strcmp(v, "false") || !strcmp(v, "off")) { out = false; return true; }
  return false;
}
bool parse_int(const char* v, int32_t lo, int32_t hi, int32_t& out) {
  if (!v || !*v) return false;
  char* end = nullptr;
  const double d = strtod(v, &end);  // accepts 3e3 like the reference parser
  if (*end || d < lo || d > hi) return false;
  out = (int32_t)d;
  return true;
}

template <class F>
pfs_status guard_body(pfs_ctx* c, F&& f) {
  try {
    if (c && c->s) {
      c->s->bind();  // the calling thread may be a worker of pfs_batch_simplify
      c->s->call_begin();
    }
    return f();
  } catch (AllocFail& e) {
    if (c) c->err = e.what;
    return PFS_ALLOC_FAIL;
  } catch (DeviceFail& e) {
    if (c) c->err = e.what;
    return PFS_DEVICE_ERROR;
  } catch (std::invalid_argument& e) {
    if (c) c->err = e.what();
    return PFS_BAD_INPUT;
  } catch (std::logic_error& e) {
    if (c) c->err = e.what();
    return PFS_BAD_STATE;
  } catch (std::bad_alloc&) {
    if (c) c->err = "host allocation failed";
    return PFS_ALLOC_FAIL;
  } catch (std::exception& e) {
    if (c) c->err = e.what();
    return PFS_DEVICE_ERROR;
  }
}
// Every call on a context holds the context's arena client lock: the arena
// moves a context's blocks only while that lock is free (or, at the start of
// a call, for the caller itself).
template <class F>
pfs_status guard(pfs_ctx* c, F&& f) {
  if (!c || !c->s) return guard_body(c, std::forward<F>(f));
  std::lock_guard<std::mutex> lk(c->s->client().busy);
  const pfs_status s = guard_body(c, std::forward<F>(f));
  if (s != PFS_OK) c->s->on_error();
  return s;
}

}  // namespace

extern "C" {

static pfs_status create_in(const pfs_config* cfg, DeviceArena* arena, pfs_ctx** out) {
  if (!out) return PFS_BAD_INPUT;
  *out = nullptr;
  const int dev = cfg ? cfg->device : 0;
  if (pfs::device_count() <= dev || dev < 0) {
    g_create_error = "no HIP device " + std::to_string(dev);
    return PFS_NO_DEVICE;
  }
  if (arena && arena->device() != dev) {
    g_create_error = "the arena belongs to another device";
    return PFS_BAD_INPUT;
  }
  pfs_ctx* c = new pfs_ctx();
  const pfs_status s = guard(c, [&] {
    const size_t budget = cfg && cfg->mem_budget ? (size_t)cfg->mem_budget : ~(size_t)0;
    c->s.reset(new Simplifier(dev, budget, cfg ? cfg->keep_input != 0 : false, arena));
    return PFS_OK;
  });
  if (s != PFS_OK) {
    g_create_error = c->err;
    delete c;
    return s;
  }
  *out = c;
  return PFS_OK;
}

pfs_status pfs_create(const pfs_config* cfg, pfs_ctx** out) { return create_in(cfg, nullptr, out); }

pfs_status pfs_create_in_arena(const pfs_config* cfg, pfs_arena* arena, pfs_ctx** out) {
  if (!arena) return PFS_BAD_INPUT;
  return create_in(cfg, reinterpret_cast<DeviceArena*>(arena), out);
}

void pfs_destroy(pfs_ctx* c) {
  if (!c) return;
  try {
    if (c->s) {
      std::lock_guard<std::mutex> lk(c->s->client().busy);
      c->s->free_all();  // the blocks go back to the arena while no compaction can touch them
    }
    c->s.reset();
  } catch (...) {
  }
  delete c;
}

pfs_status pfs_arena_create(int32_t device, uint64_t bytes, pfs_arena** out) {
  if (!out || !bytes) return PFS_BAD_INPUT;
  *out = nullptr;
  if (pfs::device_count() <= device || device < 0) return PFS_NO_DEVICE;
  *out = reinterpret_cast<pfs_arena*>(DeviceArena::create_private(device, bytes));
  return PFS_OK;
}

void pfs_arena_destroy(pfs_arena* a) { delete reinterpret_cast<DeviceArena*>(a); }

pfs_status pfs_arena_configure(int32_t device, uint64_t first_slab, uint64_t limit) {
  if (device < 0) return PFS_BAD_INPUT;
  return DeviceArena::configure(device, first_slab, limit) ? PFS_OK : PFS_BAD_STATE;
}

pfs_status pfs_arena_get_stats(pfs_arena* a, int32_t device, pfs_arena_stats* out) {
  if (!out) return PFS_BAD_INPUT;
  DeviceArena* ar = a ? reinterpret_cast<DeviceArena*>(a) : nullptr;
  if (!ar) {
    if (pfs::device_count() <= device || device < 0) return PFS_NO_DEVICE;
    ar = &DeviceArena::get(device);
  }
  const ArenaStats s = ar->stats();
  out->capacity = s.capacity;
  out->used = s.used;
  out->peak = s.peak;
  out->largest_hole = s.largest_hole;
  out->holes = s.holes;
  out->slabs = s.slabs;
  out->compactions = s.compactions;
  out->moved_blocks = s.moved_blocks;
  out->moved_bytes = s.moved_bytes;
  out->slab_allocs = s.slab_allocs;
  out->pending = s.pending;
  return PFS_OK;
}

pfs_status pfs_set_option(pfs_ctx* c, const char* name, const char* value) {
  if (!c || !name) return PFS_BAD_INPUT;
  const std::string n = name;
  struct IntOpt { const char* a; const char* b; int32_t pfs_ctx::*f; int32_t lo, hi; };
  static const IntOpt ints[] = {
      {"phases", nullptr, &pfs_ctx::phases, 0, INT32_MAX},
      {"mupos", nullptr, &pfs_ctx::mu_pos, 10, INT32_MAX},
      {"muneg", nullptr, &pfs_ctx::mu_neg, 10, INT32_MAX},
      {"vephasemin", "eliminatedlitsmin", &pfs_ctx::lits_min, 1, INT32_MAX},
      {"veclausemax", "resolventmax", &pfs_ctx::ve_clause_limit, 0, INT32_MAX},
      {"xormaxarity", nullptr, &pfs_ctx::xor_max_arity, 2, 20},
      {"submax", "submaxoccurs", &pfs_ctx::sub_limit, 100, INT32_MAX},
      {"bcemax", "bcemaxoccurs", &pfs_ctx::bce_limit, 100, INT32_MAX},
      {"eremax", "eremaxoccurs", &pfs_ctx::ere_limit, 100, INT32_MAX},
      {"ereclausemax", nullptr, &pfs_ctx::ere_max_resolvent, 0, INT32_MAX},
      {"lcvemin", "electionsmin", &pfs_ctx::lcve_min, 1, INT32_MAX},
      {"gcfreq", "collectfreq", &pfs_ctx::shrink_rate, 0, 5},
      {"lbdtier1", nullptr, &pfs_ctx::lbd_tier1, 1, INT32_MAX},
      {"decomposemin", nullptr, &pfs_ctx::decompose_min, 1, 10},
      {"electionsmax", nullptr, &pfs_ctx::lcve_max_occurs, 1, INT32_MAX},
      {"lcveclausemax", nullptr, &pfs_ctx::lcve_clause_max, 1, INT32_MAX},
      {"otptile", nullptr, &pfs_ctx::otp_tile, 0, 20},
      {"mislds", nullptr, &pfs_ctx::mis_lds, -1, 65536},
      {"mislanes", nullptr, &pfs_ctx::mis_lanes, 0, 8},
  };
  for (const IntOpt& o : ints) {
    if (n == o.a || (o.b && n == o.b)) {
      int32_t v;
      if (!parse_int(value, o.lo, o.hi, v)) {
        c->err = "bad value for option " + n;
        return PFS_BAD_INPUT;
      }
      if (n == "mislanes" && (v & (v - 1))) {
        c->err = "mislanes must be 0 (tuned), 1, 2, 4 or 8";
        return PFS_BAD_INPUT;
      }
      c->*(o.f) = v;
      return PFS_OK;
    }
  }
  struct BoolOpt { const char* a; const char* b; bool pfs_ctx::*f; };
  static const BoolOpt bools[] = {
      {"ve", nullptr, &pfs_ctx::ve},   {"ve+", "veextend", &pfs_ctx::ve_plus},
      {"sub", nullptr, &pfs_ctx::sub}, {"bce", nullptr, &pfs_ctx::bce},
      {"ere", nullptr, &pfs_ctx::ere}, {"all", nullptr, &pfs_ctx::all},
      {"velbound", "velitsbound", &pfs_ctx::velbound},
      {"profilegpu", "profilesimp", &pfs_ctx::profile},
      {"workorder", nullptr, &pfs_ctx::work_order},
      {"proof", nullptr, &pfs_ctx::proof},
      {"aggresivesort", nullptr, &pfs_ctx::aggr_sort},
      {"syncalways", nullptr, &pfs_ctx::sync_always},
      {"otpxcd", nullptr, &pfs_ctx::otp_xcd},
      {"otpranges", nullptr, &pfs_ctx::otp_ranges},
      {"autotune", nullptr, &pfs_ctx::autotune},
  };
  for (const BoolOpt& o : bools) {
    if (n == o.a || (o.b && n == o.b)) {
      bool v;
      if (!parse_bool(value, v)) {
        c->err = "bad value for option " + n;
        return PFS_BAD_INPUT;
      }
      c->*(o.f) = v;
      return PFS_OK;
    }
  }
  // GPU-build heuristics that this build does not provide (src/gpu/options.cpp:32,
  // src/gpu/options.cu:50): the rank-based MIS election (nondeterministic) and
  // the function-table gates.  Off is accepted; on is refused with a reason.
  if (n == "lcvefast" || n == "vefunction") {
    bool v;
    if (!parse_bool(value, v)) {
      c->err = "bad value for option " + n;
      return PFS_BAD_INPUT;
    }
    if (v) {
      c->err = n == "lcvefast"
                   ? "lcvefast=1 (rank-based MIS election) is not provided: the election is the deterministic "
                     "sequential one (the lexicographically first independent set, computed in parallel)"
                   : "vefunction=1 (function-table gates) is not provided: variable elimination detects EQU, "
                     "AND/OR, ITE and XOR gates as the CPU simplifier does";
      return PFS_BAD_INPUT;
    }
    return PFS_OK;
  }
  // GPU-build launch shaping and host-side options, accepted with their
  // reference ranges and without effect here: stream count (one stream per
  // context), the launch tuner's floors (this build sizes its launches from
  // the work), the option the GPU kernels never read (subclausemax,
  // ereextend), and the solver's own scheduling of simplify calls.
  {
    struct Range { const char* name; double lo, hi; bool boolean; };
    static const Range noop[] = {
        {"nstreams", 6, 32, false},        {"veminthreads", 2, 1024, false}, {"subminthreads", 2, 1024, false},
        {"ereminthreads", 2, 1024, false}, {"veminblocks", 0, 1, false},     {"subminblocks", 0, 1, false},
        {"ereminblocks", 0, 1, false},     {"subclausemax", 0, INT32_MAX, false},
        {"ereextend", 0, 3, false},        {"solve", 0, 1, true},            {"sigma", 0, 1, true},
        {"sigmalive", 0, 1, true},         {"sigmasleep", 0, 1, true},       {"sigmainc", 1, INT32_MAX, false},
        {"sigmamin", 1, INT32_MAX, false},
    };
    for (const Range& r : noop) {
      if (n != r.name) continue;
      bool ok;
      if (r.boolean) {
        bool b;
        ok = parse_bool(value, b);
      } else {
        char* end = nullptr;
        const double d = value && *value ? strtod(value, &end) : -1.0;
        ok = end && !*end && d >= r.lo && d <= r.hi;
      }
      if (!ok) {
        c->err = "bad value for option " + n;
        return PFS_BAD_INPUT;
      }
      return PFS_OK;
    }
  }
  if (n == "literalsmul") {  // FP64R(0, 4), src/gpu/options.cpp:43
    char* end = nullptr;
    const double d = value && *value ? strtod(value, &end) : -1.0;
    if (!end || *end || !(d >= 0.0 && d <= 4.0)) {
      c->err = "bad value for option " + n;
      return PFS_BAD_INPUT;
    }
    c->lits_mul = d;
    return PFS_OK;
  }
  if (n == "mapperc") {
    char* end = nullptr;
    const double d = value && *value ? strtod(value, &end) : -1.0;
    if (!end || *end || !(d >= 0.0 && d <= 1.0)) {  // FP64R(0, 1)
      c->err = "bad value for option " + n;
      return PFS_BAD_INPUT;
    }
    c->map_perc = d;
    return PFS_OK;
  }
  c->err = "unknown option " + n;
  return PFS_UNKNOWN_OPTION;
}

pfs_status pfs_upload(pfs_ctx* c, uint32_t maxvar, uint64_t ncls, const uint64_t* offsets,
                      const uint32_t* lits, const uint8_t* flags, const uint32_t* lbd,
                      const uint8_t* vstate, const uint8_t* assumed, const uint32_t* vorg,
                      const uint32_t* trail, uint32_t ntrail) {
  if (!c || !offsets || (!lits && offsets[ncls]) || (!trail && ntrail) || (!maxvar && ncls)) return PFS_BAD_INPUT;
  return guard(c, [&] {
    for (uint64_t i = 0; i < ncls; i++)
      if (offsets[i + 1] < offsets[i] + 1) throw std::invalid_argument("empty clause or bad offsets");
    for (uint32_t i = 0; i < ntrail; i++)
      if ((trail[i] >> 1) == 0 || (trail[i] >> 1) > maxvar) throw std::invalid_argument("trail literal out of range");
    c->uploaded = false;
    c->s->lits_mul = c->lits_mul;
    c->s->upload(maxvar, ncls, offsets, lits, flags, lbd, vstate, assumed, vorg, trail, ntrail);
    c->uploaded = true;
    c->simplified = false;
    c->result = PFS_RESULT_UNSOLVED;
    return PFS_OK;
  });
}

pfs_status pfs_parse_dimacs(pfs_ctx* c, const char* text, uint64_t nbytes, pfs_parse_info* info) {
  if (!c || (!text && nbytes)) return PFS_BAD_INPUT;
  return guard(c, [&] {
    c->uploaded = false;
    c->simplified = false;
    c->s->opt = normalise(c);  // "proof" and the propagation's options
    c->s->profile = c->profile;
    c->s->lits_mul = c->lits_mul;
    const Simplifier::ParseInfo r = c->s->parse_dimacs(text, nbytes);
    c->uploaded = true;
    c->result = r.status == 0 ? PFS_RESULT_UNSAT : PFS_RESULT_UNSOLVED;
    if (info) {
      info->maxvar = r.maxvar;
      info->status = r.status;
      info->header_clauses = r.header_clauses;
      info->read_clauses = r.read_clauses;
      info->clauses = r.clauses;
      info->literals = r.literals;
      info->trail = r.trail;
    }
    return PFS_OK;
  });
}

pfs_status pfs_download_trail(pfs_ctx* c, uint32_t* lits, uint64_t* n) {
  if (!c) return PFS_BAD_INPUT;
  if (!c->uploaded) {
    c->err = "nothing uploaded";
    return PFS_BAD_STATE;
  }
  return guard(c, [&] {
    if (n) *n = c->s->trail_size();
    if (lits) c->s->download_trail(lits);
    return PFS_OK;
  });
}

pfs_status pfs_rewind(pfs_ctx* c) {
  if (!c) return PFS_BAD_INPUT;
  if (!c->uploaded) {
    c->err = "nothing uploaded";
    return PFS_BAD_STATE;
  }
  return guard(c, [&] {
    c->s->rewind();
    c->simplified = false;
    return PFS_OK;
  });
}

pfs_status pfs_append(pfs_ctx* c, uint32_t maxvar, uint64_t ncls, const uint64_t* offsets, const uint32_t* lits) {
  if (!c || (ncls && (!offsets || !lits))) return PFS_BAD_INPUT;
  if (!c->uploaded) {
    c->err = "nothing uploaded";
    return PFS_BAD_STATE;
  }
  return guard(c, [&] {
    c->s->append(maxvar, ncls, offsets, lits);
    c->simplified = false;
    return PFS_OK;
  });
}

pfs_status pfs_assume(pfs_ctx* c, const uint32_t* vars, uint32_t n) {
  if (!c || (n && !vars)) return PFS_BAD_INPUT;
  if (!c->uploaded) {
    c->err = "nothing uploaded";
    return PFS_BAD_STATE;
  }
  return guard(c, [&] {
    c->s->assume(vars, n);
    return PFS_OK;
  });
}

pfs_status pfs_simplify(pfs_ctx* c, int32_t* result) {
  if (!c) return PFS_BAD_INPUT;
  if (!c->uploaded) {
    c->err = "nothing uploaded";
    return PFS_BAD_STATE;
  }
  return guard(c, [&] {
    c->s->opt = normalise(c);
    c->s->profile = c->profile;
    c->result = c->s->simplify(0);
    c->simplified = true;
    if (result) *result = c->result;
    return PFS_OK;
  });
}

pfs_status pfs_result_sizes(pfs_ctx* c, uint64_t* ncls, uint64_t* nlits, uint64_t* nunits, uint64_t* nstack,
                            uint32_t* nrounds) {
  if (!c) return PFS_BAD_INPUT;
  if (!c->uploaded) {
    c->err = "nothing uploaded";
    return PFS_BAD_STATE;
  }
  const bool unsat = c->result == PFS_RESULT_UNSAT;
  if (ncls) *ncls = unsat ? 0 : c->s->out_ncls();
  if (nlits) *nlits = unsat ? 0 : c->s->out_nlits();
  if (nunits) *nunits = c->s->new_units().size();
  if (nstack) *nstack = c->s->stack_words();
  if (nrounds) *nrounds = (uint32_t)c->s->rounds().size();
  return PFS_OK;
}

pfs_status pfs_download_cnf(pfs_ctx* c, uint64_t* offsets, uint32_t* lits, uint8_t* flags, uint32_t* lbd) {
  if (!c) return PFS_BAD_INPUT;
  if (!c->uploaded) return PFS_BAD_STATE;
  if (c->result == PFS_RESULT_UNSAT) {
    if (offsets) offsets[0] = 0;
    return PFS_OK;
  }
  return guard(c, [&] {
    c->s->download_cnf(offsets, lits, flags, lbd);
    return PFS_OK;
  });
}

pfs_status pfs_download_units(pfs_ctx* c, uint32_t* lits) {
  if (!c || !lits) return PFS_BAD_INPUT;
  if (!c->uploaded) return PFS_BAD_STATE;
  const auto& u = c->s->new_units();
  std::memcpy(lits, u.data(), u.size() * 4);
  return PFS_OK;
}

pfs_status pfs_download_vstate(pfs_ctx* c, uint8_t* vstate) {
  if (!c || !vstate) return PFS_BAD_INPUT;
  if (!c->uploaded) return PFS_BAD_STATE;
  return guard(c, [&] {
    c->s->download_vstate(vstate);
    return PFS_OK;
  });
}

pfs_status pfs_map(pfs_ctx* c, int32_t force, pfs_map_info* info) {
  if (!c) return PFS_BAD_INPUT;
  if (!c->simplified) {
    c->err = "simplify has not run";
    return PFS_BAD_STATE;
  }
  if (c->result == PFS_RESULT_UNSAT) {
    c->err = "the formula is refuted";
    return PFS_BAD_STATE;
  }
  return guard(c, [&] {
    c->s->profile = c->profile;
    const Simplifier::MapInfo r = c->s->map(force != 0, c->map_perc);
    if (info) {
      info->applied = r.applied;
      info->old_maxvar = r.old_nvars;
      info->new_maxvar = r.new_nvars;
      info->first_frozen = r.first;
      info->trail_lit = r.trail_lit;
      info->inactive = r.inactive;
    }
    return PFS_OK;
  });
}

pfs_status pfs_decompose(pfs_ctx* c, uint32_t max_rounds, pfs_els_info* info) {
  if (!c) return PFS_BAD_INPUT;
  if (!c->uploaded) {
    c->err = "nothing uploaded";
    return PFS_BAD_STATE;
  }
  if (c->result == PFS_RESULT_UNSAT) {
    c->err = "the formula is refuted";
    return PFS_BAD_STATE;
  }
  return guard(c, [&] {
    c->s->opt = normalise(c);
    c->s->profile = c->profile;
    const Simplifier::ElsInfo r = c->s->decompose(max_rounds ? max_rounds : (uint32_t)c->decompose_min);
    if (r.status == 0) c->result = PFS_RESULT_UNSAT;
    if (info) {
      info->rounds = r.rounds;
      info->status = r.status;
      info->substituted = r.substituted;
      info->units = r.units;
      info->stack_words = r.stack_words;
    }
    return PFS_OK;
  });
}

pfs_status pfs_download_map(pfs_ctx* c, uint32_t* mapped) {
  if (!c || !mapped) return PFS_BAD_INPUT;
  return guard(c, [&] {
    c->s->download_map(mapped);
    return PFS_OK;
  });
}

pfs_status pfs_map_lits(pfs_ctx* c, uint32_t* lits, uint64_t n) {
  if (!c || (!lits && n)) return PFS_BAD_INPUT;
  return guard(c, [&] {
    c->s->map_lits(lits, n);
    return PFS_OK;
  });
}

pfs_status pfs_map_lits_noalloc(pfs_ctx* c, uint32_t* lits, uint64_t n) {
  if (!c || (!lits && n)) return PFS_BAD_INPUT;
  return guard(c, [&] {
    c->s->map_lits_noalloc(lits, n);
    return PFS_OK;
  });
}

pfs_status pfs_download_stack(pfs_ctx* c, uint32_t* words) {
  if (!c || !words) return PFS_BAD_INPUT;
  if (!c->uploaded) return PFS_BAD_STATE;
  return guard(c, [&] {
    c->s->download_stack(words);
    return PFS_OK;
  });
}

pfs_status pfs_proof_size(pfs_ctx* c, uint64_t* nbytes) {
  if (!c || !nbytes) return PFS_BAD_INPUT;
  if (!c->uploaded) return PFS_BAD_STATE;
  *nbytes = c->s->proof_bytes();
  return PFS_OK;
}

pfs_status pfs_download_proof(pfs_ctx* c, uint8_t* bytes) {
  if (!c || (!bytes && c->simplified && c->s->proof_bytes())) return PFS_BAD_INPUT;
  if (!c->uploaded) return PFS_BAD_STATE;
  return guard(c, [&] {
    c->s->download_proof(bytes);
    return PFS_OK;
  });
}

pfs_status pfs_round_stats_get(pfs_ctx* c, pfs_round_stats* out, uint32_t max, uint32_t* n) {
  if (!c) return PFS_BAD_INPUT;
  const auto& r = c->s->rounds();
  const uint32_t k = (uint32_t)std::min<size_t>(r.size(), max);
  for (uint32_t i = 0; i < k; i++) {
    out[i].elected = r[i].elected;
    out[i].eliminated = r[i].eliminated;
    out[i].units = r[i].units;
    out[i].resolvents = r[i].resolvents;
    out[i].clauses_after = r[i].clauses_after;
    out[i].literals_after = r[i].literals_after;
    out[i].mis_iterations = r[i].mis_iterations;
    out[i].sub_passes = r[i].sub_passes;
    out[i].ere_passes = r[i].ere_passes;
    out[i].stages = r[i].stages;
    out[i].clauses_before = r[i].clauses_before;
    out[i].literals_before = r[i].literals_before;
    out[i].elected_occurs = r[i].elected_occurs;
    out[i].elected_bytes = r[i].elected_bytes;
    out[i].cand_occurs = r[i].cand_occurs;
    out[i].cand_bytes = r[i].cand_bytes;
    out[i].resolvent_words = r[i].resolvent_words;
    out[i].stack_words = r[i].stack_words;
  }
  if (n) *n = (uint32_t)r.size();
  return PFS_OK;
}

pfs_status pfs_stage_times(pfs_ctx* c, double* ms, uint64_t* calls, uint32_t n) {
  if (!c) return PFS_BAD_INPUT;
  for (uint32_t i = 0; i < n && i < (uint32_t)S_NSTAGES; i++) {
    if (ms) ms[i] = c->s->stage_ms[i];
    if (calls) calls[i] = c->s->stage_calls[i];
  }
  return PFS_OK;
}

const char* pfs_stage_name(uint32_t stage) { return stage < (uint32_t)S_NSTAGES ? kStageNames[stage] : ""; }
uint32_t pfs_stage_count(void) { return (uint32_t)S_NSTAGES; }

uint32_t pfs_kernel_count(pfs_ctx* c) { return c ? (uint32_t)c->s->kernel_times.size() : 0; }

pfs_status pfs_kernel_stats(pfs_ctx* c, uint32_t i, const char** name, double* total_ms, uint64_t* launches) {
  if (!c || i >= c->s->kernel_times.size()) return PFS_BAD_INPUT;
  const KernelTime& k = c->s->kernel_times[i];
  if (name) *name = k.name.c_str();
  if (total_ms) *total_ms = k.ms;
  if (launches) *launches = k.launches;
  return PFS_OK;
}

pfs_status pfs_launch_shape(pfs_ctx* c, int32_t* otp_tile_log2, int32_t* mis_lds, int32_t* mis_lanes,
                            int32_t* tuned) {
  if (!c) return PFS_BAD_INPUT;
  if (otp_tile_log2) *otp_tile_log2 = c->s->used_tile();
  if (mis_lds) *mis_lds = c->s->used_lds();
  if (mis_lanes) *mis_lanes = c->s->used_lanes();
  if (tuned) *tuned = c->s->tuned_now() ? 1 : 0;
  return PFS_OK;
}

pfs_status pfs_memory(pfs_ctx* c, uint64_t* used, uint64_t* peak) {
  if (!c) return PFS_BAD_INPUT;
  if (used) *used = c->s->mem_used();
  if (peak) *peak = c->s->mem_peak();
  return PFS_OK;
}

const char* pfs_last_error(const pfs_ctx* c) { return c ? c->err.c_str() : g_create_error.c_str(); }
pfs_status pfs_batch_simplify(pfs_ctx* const* ctxs, uint32_t n, uint32_t workers, int32_t* results,
                              pfs_status* statuses, double* ms) {
  if (!ctxs || (n && !results)) return PFS_BAD_INPUT;
  for (uint32_t i = 0; i < n; i++)
    if (!ctxs[i]) return PFS_BAD_INPUT;
  const uint32_t nw = std::max<uint32_t>(1, std::min(workers ? workers : 1, n));
  std::atomic<uint32_t> next{0};
  std::vector<pfs_status> st(n, PFS_OK);
  auto work = [&] {
    for (uint32_t i; (i = next.fetch_add(1)) < n;) {
      const auto t0 = std::chrono::steady_clock::now();
      st[i] = pfs_simplify(ctxs[i], &results[i]);
      if (ms) ms[i] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
  };
  std::vector<std::thread> pool;
  for (uint32_t w = 1; w < nw; w++) pool.emplace_back(work);
  work();
  for (auto& t : pool) t.join();
  pfs_status first = PFS_OK;
  for (uint32_t i = 0; i < n; i++) {
    if (statuses) statuses[i] = st[i];
    if (first == PFS_OK && st[i] != PFS_OK) first = st[i];
  }
  return first;
}

const char* pfs_version(void) { return "pfsimp 0.1 (gfx950)"; }
int32_t pfs_device_count(void) { return pfs::device_count(); }

}  // extern "C"
