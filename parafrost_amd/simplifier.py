"""Python binding of libpfsimp.so (the C ABI in include/pfsimp.h).

The binding is thin: it marshals numpy arrays into the C entry points and
back.  There is no fallback: if the HIP library is missing or no GPU is
visible, constructing a ``Simplifier`` raises.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

_LIB = None
_LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib", "libpfsimp.so")

PFS_OK = 0
STATUS_NAMES = {0: "OK", 1: "ALLOC_FAIL", 2: "BAD_INPUT", 3: "DEVICE_ERROR", 4: "NO_DEVICE", 5: "BAD_STATE",
                6: "UNKNOWN_OPTION"}
RESULT_UNSAT, RESULT_SAT, RESULT_UNSOLVED = 0, 1, 2

# exported symbols declared in include/pfsimp.h
SYMBOLS = ("pfs_create", "pfs_destroy", "pfs_set_option", "pfs_upload", "pfs_rewind", "pfs_simplify",
           "pfs_result_sizes", "pfs_download_cnf", "pfs_download_units", "pfs_download_vstate",
           "pfs_download_stack", "pfs_proof_size", "pfs_download_proof", "pfs_round_stats_get", "pfs_stage_times", "pfs_stage_name", "pfs_stage_count",
           "pfs_kernel_count", "pfs_kernel_stats", "pfs_memory", "pfs_last_error", "pfs_version",
           "pfs_device_count", "pfs_batch_simplify", "pfs_append", "pfs_assume", "pfs_server_create",
           "pfs_server_submit", "pfs_server_start", "pfs_server_wait", "pfs_server_destroy", "pfs_map",
           "pfs_download_map", "pfs_map_lits", "pfs_parse_dimacs", "pfs_download_trail",
           "pfs_decompose", "pfs_arena_configure", "pfs_arena_create", "pfs_arena_destroy",
           "pfs_create_in_arena", "pfs_arena_get_stats", "pfs_map_lits_noalloc", "pfs_launch_shape")


class PfsError(RuntimeError):
    pass


class _Config(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int32), ("mem_budget", ctypes.c_uint64), ("keep_input", ctypes.c_int32)]


class _ArenaStats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint64) for n in ("capacity", "used", "peak", "largest_hole", "holes", "slabs",
                                               "compactions", "moved_blocks", "moved_bytes", "slab_allocs",
                                               "pending")]


class _RoundStats(ctypes.Structure):
    _fields_ = [("elected", ctypes.c_uint32), ("eliminated", ctypes.c_uint32), ("units", ctypes.c_uint32),
                ("resolvents", ctypes.c_uint32), ("clauses_after", ctypes.c_uint64),
                ("literals_after", ctypes.c_uint64), ("mis_iterations", ctypes.c_uint32),
                ("sub_passes", ctypes.c_uint32), ("ere_passes", ctypes.c_uint32), ("stages", ctypes.c_uint32),
                ("clauses_before", ctypes.c_uint64), ("literals_before", ctypes.c_uint64),
                ("elected_occurs", ctypes.c_uint64), ("elected_bytes", ctypes.c_uint64),
                ("cand_occurs", ctypes.c_uint64), ("cand_bytes", ctypes.c_uint64),
                ("resolvent_words", ctypes.c_uint64), ("stack_words", ctypes.c_uint64)]


class _ParseInfo(ctypes.Structure):
    _fields_ = [("maxvar", ctypes.c_uint32), ("status", ctypes.c_uint32), ("header_clauses", ctypes.c_uint64),
                ("read_clauses", ctypes.c_uint64), ("clauses", ctypes.c_uint64), ("literals", ctypes.c_uint64),
                ("trail", ctypes.c_uint64)]


class _ElsInfo(ctypes.Structure):
    _fields_ = [("rounds", ctypes.c_uint32), ("status", ctypes.c_uint32), ("substituted", ctypes.c_uint64),
                ("units", ctypes.c_uint64), ("stack_words", ctypes.c_uint64)]


class _MapInfo(ctypes.Structure):
    _fields_ = [("applied", ctypes.c_uint32), ("old_maxvar", ctypes.c_uint32), ("new_maxvar", ctypes.c_uint32),
                ("first_frozen", ctypes.c_uint32), ("trail_lit", ctypes.c_uint32), ("inactive", ctypes.c_uint32)]


def lib_path() -> str:
    return _LIB_PATH


def load_library():
    """Loads libpfsimp.so (raises if it is not built)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(_LIB_PATH):
        raise PfsError(f"HIP simplifier library not built: {_LIB_PATH} (run __graft_entry__.build())")
    L = ctypes.CDLL(_LIB_PATH)
    vp, u32, u64, i32 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int32
    L.pfs_create.argtypes = [ctypes.POINTER(_Config), ctypes.POINTER(vp)]
    L.pfs_destroy.argtypes = [vp]
    L.pfs_destroy.restype = None
    L.pfs_set_option.argtypes = [vp, ctypes.c_char_p, ctypes.c_char_p]
    L.pfs_upload.argtypes = [vp, u32, u64, vp, vp, vp, vp, vp, vp, vp, vp, u32]
    L.pfs_rewind.argtypes = [vp]
    L.pfs_simplify.argtypes = [vp, ctypes.POINTER(i32)]
    L.pfs_result_sizes.argtypes = [vp, ctypes.POINTER(u64), ctypes.POINTER(u64), ctypes.POINTER(u64),
                                   ctypes.POINTER(u64), ctypes.POINTER(u32)]
    L.pfs_download_cnf.argtypes = [vp, vp, vp, vp, vp]
    L.pfs_download_units.argtypes = [vp, vp]
    L.pfs_download_vstate.argtypes = [vp, vp]
    L.pfs_download_stack.argtypes = [vp, vp]
    L.pfs_proof_size.argtypes = [vp, ctypes.POINTER(u64)]
    L.pfs_download_proof.argtypes = [vp, vp]
    L.pfs_round_stats_get.argtypes = [vp, ctypes.POINTER(_RoundStats), u32, ctypes.POINTER(u32)]
    L.pfs_stage_times.argtypes = [vp, vp, vp, u32]
    L.pfs_stage_name.argtypes = [u32]
    L.pfs_stage_name.restype = ctypes.c_char_p
    L.pfs_stage_count.restype = u32
    L.pfs_memory.argtypes = [vp, ctypes.POINTER(u64), ctypes.POINTER(u64)]
    L.pfs_launch_shape.argtypes = [vp, ctypes.POINTER(i32), ctypes.POINTER(i32), ctypes.POINTER(i32),
                                   ctypes.POINTER(i32)]
    L.pfs_kernel_count.argtypes = [vp]
    L.pfs_kernel_count.restype = u32
    L.pfs_kernel_stats.argtypes = [vp, u32, ctypes.POINTER(ctypes.c_char_p), ctypes.POINTER(ctypes.c_double),
                                   ctypes.POINTER(u64)]
    L.pfs_last_error.argtypes = [vp]
    L.pfs_last_error.restype = ctypes.c_char_p
    L.pfs_version.restype = ctypes.c_char_p
    L.pfs_batch_simplify.argtypes = [vp, u32, u32, vp, vp, vp]
    L.pfs_append.argtypes = [vp, u32, u64, vp, vp]
    L.pfs_assume.argtypes = [vp, vp, u32]
    L.pfs_server_create.argtypes = [u32, u64, ctypes.POINTER(vp)]
    L.pfs_server_submit.argtypes = [vp, vp, ctypes.c_double, u32, u64, vp, vp, vp, u32, ctypes.POINTER(u32)]
    L.pfs_server_start.argtypes = [vp]
    L.pfs_server_wait.argtypes = [vp, u32, ctypes.POINTER(i32), ctypes.POINTER(ctypes.c_double),
                                  ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
    L.pfs_server_destroy.argtypes = [vp]
    L.pfs_map.argtypes = [vp, i32, ctypes.POINTER(_MapInfo)]
    L.pfs_download_map.argtypes = [vp, vp]
    L.pfs_map_lits.argtypes = [vp, vp, u64]
    L.pfs_map_lits_noalloc.argtypes = [vp, vp, u64]
    L.pfs_parse_dimacs.argtypes = [vp, ctypes.c_char_p, u64, ctypes.POINTER(_ParseInfo)]
    L.pfs_download_trail.argtypes = [vp, vp, ctypes.POINTER(u64)]
    L.pfs_decompose.argtypes = [vp, u32, ctypes.POINTER(_ElsInfo)]
    L.pfs_device_count.restype = i32
    L.pfs_arena_configure.argtypes = [i32, u64, u64]
    L.pfs_arena_create.argtypes = [i32, u64, ctypes.POINTER(vp)]
    L.pfs_arena_destroy.argtypes = [vp]
    L.pfs_arena_destroy.restype = None
    L.pfs_create_in_arena.argtypes = [ctypes.POINTER(_Config), vp, ctypes.POINTER(vp)]
    L.pfs_arena_get_stats.argtypes = [vp, i32, ctypes.POINTER(_ArenaStats)]
    _LIB = L
    return L


def _ptr(a):
    return None if a is None else a.ctypes.data


@dataclass
class Result:
    status: int  # 0 UNSAT, 1 SAT, 2 UNSOLVED
    maxvar: int
    sizes: np.ndarray
    flags: np.ndarray
    lbd: np.ndarray
    lits: np.ndarray
    units: np.ndarray
    resolved: np.ndarray
    vstate: np.ndarray
    rounds: list
    proof: bytes = b""  # binary DRAT lines (option "proof")


def arena_stats(arena=None, device: int = 0) -> dict:
    """pfs_arena_get_stats of a private Arena, or of the device's default arena."""
    L = load_library()
    out = _ArenaStats()
    h = arena._h if arena is not None else None
    st = L.pfs_arena_get_stats(h, device, ctypes.byref(out))
    if st != PFS_OK:
        raise PfsError(f"pfs_arena_get_stats: {STATUS_NAMES.get(st, st)}")
    return {n: int(getattr(out, n)) for n, _ in _ArenaStats._fields_}


def arena_configure(device: int, first_slab: int, limit: int = 0) -> bool:
    """pfs_arena_configure: False once the device's default arena exists."""
    return load_library().pfs_arena_configure(device, first_slab, limit) == PFS_OK


class Arena:
    """A private device arena of a fixed capacity (pfs_arena_create):
    contexts created with arena=... share it."""

    def __init__(self, nbytes: int, device: int = 0):
        L = load_library()
        self._L = L
        self._h = ctypes.c_void_p()
        st = L.pfs_arena_create(device, nbytes, ctypes.byref(self._h))
        if st != PFS_OK:
            raise PfsError(f"pfs_arena_create: {STATUS_NAMES.get(st, st)}")
        self.device = device

    def stats(self) -> dict:
        return arena_stats(self, self.device)

    def close(self):
        """Destroys the arena: every context in it must be closed first."""
        if self._h:
            self._L.pfs_arena_destroy(self._h)
            self._h = ctypes.c_void_p()


class Simplifier:
    """One simplifier context (device state of one formula)."""

    def __init__(self, device: int = 0, mem_budget: int = 0, keep_input: bool = False, arena: Arena = None,
                 **options):
        L = load_library()
        self._L = L
        self._h = ctypes.c_void_p()
        cfg = _Config(device, mem_budget, 1 if keep_input else 0)
        if arena is not None:
            st = L.pfs_create_in_arena(ctypes.byref(cfg), arena._h, ctypes.byref(self._h))
        else:
            st = L.pfs_create(ctypes.byref(cfg), ctypes.byref(self._h))
        if st != PFS_OK:
            raise PfsError(f"pfs_create: {STATUS_NAMES.get(st, st)}: {L.pfs_last_error(None).decode()}")
        self.maxvar = 0
        for k, v in options.items():
            self.set_option(k, v)

    def _check(self, st, what):
        if st != PFS_OK:
            msg = self._L.pfs_last_error(self._h).decode()
            raise PfsError(f"{what}: {STATUS_NAMES.get(st, st)}: {msg}")

    def close(self):
        if self._h:
            self._L.pfs_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_option(self, name: str, value) -> None:
        if isinstance(value, bool):
            value = "1" if value else "0"
        self._check(self._L.pfs_set_option(self._h, name.encode(), str(value).encode()), f"option {name}")

    def upload(self, maxvar, offsets, lits, flags=None, lbd=None, vstate=None, assumed=None, vorg=None,
               trail=None) -> None:
        offsets = np.ascontiguousarray(offsets, np.uint64)
        lits = np.ascontiguousarray(lits, np.uint32)
        flags = None if flags is None else np.ascontiguousarray(flags, np.uint8)
        lbd = None if lbd is None else np.ascontiguousarray(lbd, np.uint32)
        vstate = None if vstate is None else np.ascontiguousarray(vstate, np.uint8)
        assumed = None if assumed is None else np.ascontiguousarray(assumed, np.uint8)
        vorg = None if vorg is None else np.ascontiguousarray(vorg, np.uint32)
        trail = np.zeros(0, np.uint32) if trail is None else np.ascontiguousarray(trail, np.uint32)
        self.maxvar = int(maxvar)
        st = self._L.pfs_upload(self._h, maxvar, offsets.shape[0] - 1, _ptr(offsets), _ptr(lits), _ptr(flags),
                                _ptr(lbd), _ptr(vstate), _ptr(assumed), _ptr(vorg), _ptr(trail), trail.shape[0])
        self._check(st, "pfs_upload")

    def upload_state(self, st) -> None:
        """Uploads a parafrost_amd.pfd.State input dump."""
        off = np.zeros(st.sizes.shape[0] + 1, np.uint64)
        np.cumsum(st.sizes, out=off[1:])
        self.upload(st.maxvar, off, st.lits, st.flags, st.lbd, st.vstate, None, None, st.units)

    def append(self, maxvar, offsets, lits) -> None:
        """Adds original clauses to the resident formula (pfs_append)."""
        offsets = np.ascontiguousarray(offsets, np.uint64)
        lits = np.ascontiguousarray(lits, np.uint32)
        self.maxvar = max(self.maxvar, int(maxvar))
        self._check(self._L.pfs_append(self._h, maxvar, offsets.shape[0] - 1, _ptr(offsets), _ptr(lits)),
                    "pfs_append")

    def assume(self, variables) -> None:
        """Replaces the assumed (never eliminated) variables (pfs_assume)."""
        v = np.ascontiguousarray(variables, np.uint32)
        self._check(self._L.pfs_assume(self._h, _ptr(v), v.shape[0]), "pfs_assume")

    def parse_dimacs(self, text: bytes) -> dict:
        """DIMACS text -> resident formula on the device (pfs_parse_dimacs)."""
        if isinstance(text, str):
            text = text.encode()
        info = _ParseInfo()
        self._check(self._L.pfs_parse_dimacs(self._h, text, len(text), ctypes.byref(info)), "pfs_parse_dimacs")
        out = {k: int(getattr(info, k)) for k, _ in _ParseInfo._fields_}
        self.maxvar = out["maxvar"]
        return out

    def decompose(self, rounds: int = 0) -> dict:
        """Equivalent-literal substitution on the resident formula (pfs_decompose)."""
        info = _ElsInfo()
        self._check(self._L.pfs_decompose(self._h, rounds, ctypes.byref(info)), "pfs_decompose")
        return {k: int(getattr(info, k)) for k, _ in _ElsInfo._fields_}

    def proof(self) -> bytes:
        """The binary DRAT lines of the last call (pfs_proof_size,
        pfs_download_proof): a simplify or an ingestion with "proof" on."""
        n = ctypes.c_uint64()
        self._check(self._L.pfs_proof_size(self._h, ctypes.byref(n)), "pfs_proof_size")
        if not n.value:
            return b""
        pf = np.zeros(n.value, np.uint8)
        self._check(self._L.pfs_download_proof(self._h, _ptr(pf)), "pfs_download_proof")
        return pf.tobytes()

    def trail(self) -> np.ndarray:
        """The root trail of the resident formula (pfs_download_trail)."""
        n = ctypes.c_uint64()
        self._check(self._L.pfs_download_trail(self._h, None, ctypes.byref(n)), "pfs_download_trail")
        t = np.zeros(max(n.value, 1), np.uint32)
        self._check(self._L.pfs_download_trail(self._h, _ptr(t), ctypes.byref(n)), "pfs_download_trail")
        return t[: n.value]

    def rewind(self) -> None:
        self._check(self._L.pfs_rewind(self._h), "pfs_rewind")

    def simplify(self) -> int:
        r = ctypes.c_int32(0)
        self._check(self._L.pfs_simplify(self._h, ctypes.byref(r)), "pfs_simplify")
        return int(r.value)

    def result(self, status: int) -> Result:
        L = self._L
        nc, nl, nu, ns = (ctypes.c_uint64() for _ in range(4))
        nr = ctypes.c_uint32()
        self._check(L.pfs_result_sizes(self._h, ctypes.byref(nc), ctypes.byref(nl), ctypes.byref(nu),
                                       ctypes.byref(ns), ctypes.byref(nr)), "pfs_result_sizes")
        off = np.zeros(nc.value + 1, np.uint64)
        lits = np.zeros(max(nl.value, 1), np.uint32)
        flags = np.zeros(max(nc.value, 1), np.uint8)
        lbd = np.zeros(max(nc.value, 1), np.uint32)
        self._check(L.pfs_download_cnf(self._h, _ptr(off), _ptr(lits), _ptr(flags), _ptr(lbd)), "pfs_download_cnf")
        units = np.zeros(max(nu.value, 1), np.uint32)
        self._check(L.pfs_download_units(self._h, _ptr(units)), "pfs_download_units")
        stack = np.zeros(max(ns.value, 1), np.uint32)
        self._check(L.pfs_download_stack(self._h, _ptr(stack)), "pfs_download_stack")
        vst = np.zeros(self.maxvar + 1, np.uint8)
        self._check(L.pfs_download_vstate(self._h, _ptr(vst)), "pfs_download_vstate")
        rs = (_RoundStats * max(nr.value, 1))()
        n = ctypes.c_uint32()
        L.pfs_round_stats_get(self._h, rs, nr.value, ctypes.byref(n))
        rounds = [{k: getattr(rs[i], k) for k, _ in _RoundStats._fields_} for i in range(n.value)]
        sizes = np.diff(off).astype(np.uint32)
        npf = ctypes.c_uint64()
        self._check(L.pfs_proof_size(self._h, ctypes.byref(npf)), "pfs_proof_size")
        proof = b""
        if npf.value:
            pf = np.zeros(npf.value, np.uint8)
            self._check(L.pfs_download_proof(self._h, _ptr(pf)), "pfs_download_proof")
            proof = pf.tobytes()
        return Result(status, self.maxvar, sizes, flags[: nc.value], lbd[: nc.value], lits[: nl.value],
                      units[: nu.value], stack[: ns.value], vst, rounds, proof)

    def map(self, force: bool = False) -> dict:
        """Compact variable renumbering of the resident formula (pfs_map)."""
        info = _MapInfo()
        self._check(self._L.pfs_map(self._h, 1 if force else 0, ctypes.byref(info)), "pfs_map")
        out = {k: int(getattr(info, k)) for k, _ in _MapInfo._fields_}
        if out["applied"]:
            self.maxvar = out["new_maxvar"]
        return out

    def download_map(self, old_maxvar: int) -> np.ndarray:
        """Old variable -> new variable of the last map (0: dropped)."""
        m = np.zeros(old_maxvar + 1, np.uint32)
        self._check(self._L.pfs_download_map(self._h, _ptr(m)), "pfs_download_map")
        return m

    def map_lits(self, lits) -> np.ndarray:
        """VMAP::mapOrgs on a copy of 'lits' (pfs_map_lits)."""
        a = np.array(lits, dtype=np.uint32, copy=True)
        self._check(self._L.pfs_map_lits(self._h, _ptr(a), a.shape[0]), "pfs_map_lits")
        return a

    def map_lits_noalloc(self, lits) -> np.ndarray:
        """pfs_map_lits_noalloc on a copy of 'lits': the same mapping staged
        through the library's fixed device scratch (no allocation)."""
        a = np.array(lits, dtype=np.uint32, copy=True)
        self._check(self._L.pfs_map_lits_noalloc(self._h, _ptr(a), a.shape[0]), "pfs_map_lits_noalloc")
        return a

    def stage_times(self) -> dict:
        L = self._L
        n = L.pfs_stage_count()
        ms = np.zeros(n, np.float64)
        calls = np.zeros(n, np.uint64)
        L.pfs_stage_times(self._h, _ptr(ms), _ptr(calls), n)
        return {L.pfs_stage_name(i).decode(): (float(ms[i]), int(calls[i])) for i in range(n)}

    def kernel_times(self) -> dict:
        """{kernel: (total ms, launches)} of the last simplify (profilegpu on)."""
        L = self._L
        out = {}
        for i in range(L.pfs_kernel_count(self._h)):
            name, ms, n = ctypes.c_char_p(), ctypes.c_double(), ctypes.c_uint64()
            L.pfs_kernel_stats(self._h, i, ctypes.byref(name), ctypes.byref(ms), ctypes.byref(n))
            out[name.value.decode()] = (ms.value, n.value)
        return out

    def launch_shape(self) -> dict:
        """pfs_launch_shape: the occurrence-table tile (log2 clauses), the
        election's LDS bytes and lanes per candidate the last simplify used,
        and whether it tuned them."""
        t, l, g, tuned = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        self._check(self._L.pfs_launch_shape(self._h, ctypes.byref(t), ctypes.byref(l), ctypes.byref(g),
                                             ctypes.byref(tuned)), "pfs_launch_shape")
        return {"otp_tile_log2": t.value, "mis_lds": l.value, "mis_lanes": g.value, "tuned": bool(tuned.value)}

    def memory(self) -> tuple[int, int]:
        used, peak = ctypes.c_uint64(), ctypes.c_uint64()
        self._L.pfs_memory(self._h, ctypes.byref(used), ctypes.byref(peak))
        return used.value, peak.value


def simplify_state(st, device: int = 0, **options) -> Result:
    """Simplifies one PFD input state on the GPU and returns the result."""
    s = Simplifier(device, **options)
    try:
        s.upload_state(st)
        status = s.simplify()
        return s.result(status)
    finally:
        s.close()


def batch_simplify(simplifiers, workers: int = 8):
    """Runs simplify() on every Simplifier of the list (each already holding
    its formula) with 'workers' native threads, one stream per context
    (pfs_batch_simplify).  Returns (results, per-formula wall ms)."""
    n = len(simplifiers)
    if not n:
        return [], []
    L = simplifiers[0]._L
    handles = (ctypes.c_void_p * n)(*[s._h.value for s in simplifiers])
    res = np.zeros(n, np.int32)
    sts = np.zeros(n, np.int32)
    ms = np.zeros(n, np.float64)
    st = L.pfs_batch_simplify(handles, n, workers, _ptr(res), _ptr(sts), _ptr(ms))
    if st != PFS_OK:
        bad = int(np.nonzero(sts != PFS_OK)[0][0])
        simplifiers[bad]._check(int(sts[bad]), f"pfs_batch_simplify (formula {bad})")
    return [int(x) for x in res], ms.tolist()


class Server:
    """Request server over incremental sessions (pfs_server_*): each session
    is a Simplifier whose formula stays resident; a request appends clauses,
    replaces the assumed variables and simplifies, released at a given time
    (ms after start()) and served by 'workers' native threads."""

    def __init__(self, workers: int = 8, mem_budget: int = 0):
        self._L = load_library()
        self._h = ctypes.c_void_p()
        st = self._L.pfs_server_create(workers, mem_budget, ctypes.byref(self._h))
        if st != PFS_OK:
            raise PfsError(f"pfs_server_create: {STATUS_NAMES.get(st, st)}")

    def submit(self, session: "Simplifier", release_ms: float, maxvar: int, offsets, lits, assumed) -> int:
        offsets = np.ascontiguousarray(offsets, np.uint64)
        lits = np.ascontiguousarray(lits, np.uint32)
        assumed = np.ascontiguousarray(assumed, np.uint32)
        t = ctypes.c_uint32()
        session.maxvar = max(session.maxvar, int(maxvar))
        st = self._L.pfs_server_submit(self._h, session._h, float(release_ms), maxvar, offsets.shape[0] - 1,
                                       _ptr(offsets), _ptr(lits), _ptr(assumed), assumed.shape[0], ctypes.byref(t))
        if st != PFS_OK:
            raise PfsError(f"pfs_server_submit: {STATUS_NAMES.get(st, st)}")
        return t.value

    def start(self) -> None:
        st = self._L.pfs_server_start(self._h)
        if st != PFS_OK:
            raise PfsError(f"pfs_server_start: {STATUS_NAMES.get(st, st)}")

    def wait(self, ticket: int):
        """-> (status, result, release_ms, start_ms, end_ms)"""
        r = ctypes.c_int32()
        a, b, c = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        st = self._L.pfs_server_wait(self._h, ticket, ctypes.byref(r), ctypes.byref(a), ctypes.byref(b),
                                     ctypes.byref(c))
        return st, r.value, a.value, b.value, c.value

    def close(self) -> None:
        if self._h:
            self._L.pfs_server_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
