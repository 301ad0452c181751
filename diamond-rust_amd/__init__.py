"""Python binding of the MI355X search core (libdmr.so, C ABI in include/dmr.h).

Mirrors the reference's search surface: SearchConfig (src/search.rs:19-80),
blastp/blastx (search::run + merge_frames, src/search.rs:95-146, 357-396) and
the drop-in calls seed::build, seed::join_into, ungapped::xdrop,
gapped::extend/extend_span, gapped::band_for. Every call runs on the GPU
through libdmr.so; there is no CPU fallback. Importing works without a GPU
(the library only needs the HIP runtime); creating a Context needs one.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import os
import pathlib

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
LIB_PATH = HERE / "libdmr.so"

DMR_OK = 0
DMR_E_ARG = -1
DMR_E_DEVICE = -2
DMR_E_OOM = -3
DMR_E_CAPACITY = -4
DMR_E_UNSUPPORTED = -5
DMR_E_TIMEOUT = -6

# Symbols declared by include/dmr.h.
EXPORTS = [
    "dmr_search_config_default", "dmr_last_error", "dmr_version", "dmr_ctx_create",
    "dmr_ctx_set_tuning", "dmr_ctx_get_tuning", "dmr_ctx_load_tuning", "dmr_sw16_variants",
    "dmr_ctx_destroy", "dmr_ctx_share_db", "dmr_device_count", "dmr_band_for", "dmr_evalue", "dmr_seed_build",
    "dmr_seed_join", "dmr_xdrop", "dmr_xdrop_batch", "dmr_extend_span", "dmr_extend_span_noalloc", "dmr_extend",
    "dmr_extend_span_batch", "dmr_db_load", "dmr_db_load_file", "dmr_db_ids", "dmr_db_headers", "dmr_db_build_index", "dmr_blastp", "dmr_blastx",
    "dmr_six_frames", "dmr_queries_load", "dmr_search_loaded", "dmr_blastp_plain", "dmr_last_class_profile",
    "dmr_result_count", "dmr_result_hsps", "dmr_result_ops", "dmr_result_format",
    "dmr_result_stats", "dmr_last_dp_cells", "dmr_result_destroy", "dmr_score_layout",
    "dmr_fasta_parse", "dmr_fasta_sizes", "dmr_fasta_residues", "dmr_fasta_text", "dmr_fasta_destroy",
    "dmr_server_create", "dmr_server_submit", "dmr_server_wait", "dmr_server_stats", "dmr_server_destroy",
]

# Packed score classes (tuning keys lanes_<c> / threads_<c>): band_for widths
# and merged-band width bounds.
PACKED_CLASSES = ["w25", "w33", "w61", "w81", "w129", "m97", "m129", "m161", "m193", "m225", "m257", "m385",
                  "m513"]
# Score-kernel classes, in the order of dmr_last_class_profile (DMR_SCORE_CLASSES).
SCORE_CLASSES = PACKED_CLASSES + ["wave64", "wave128", "wave256", "wave512", "wave1024", "wave2048", "rowexact"]


def _class_kernel(name: str, lanes: int, threads: int, f16: bool) -> str:
    """Kernel that scores class `name` with the given lanes per DP pair and
    block size: its template instance as rocprofv3 names it, without the dmr
    namespaces."""
    if name.startswith("wave"):
        return f"k_sw_wave<{int(name[4:]) // 64}, false>"
    if name == "rowexact":
        return "k_sw_block / k_sw_generic<false>"
    tr = "F16" if f16 else "I16"
    w = int(name[1:])
    if name.startswith("m"):  # merged bands: per-DP walls
        return f"k_sw16p<{w}, {lanes}, {threads}, {tr}, true>"
    return f"k_sw16p<{w}, {lanes}, {threads}, {tr}, false>" if lanes > 1 else f"k_sw16<{w}, {tr}>"


HSP_FIELDS = ["query", "subject", "frame", "q_begin", "q_end", "s_begin", "s_end", "score",
              "identities", "positives", "mismatches", "gap_openings", "gaps", "length", "n_ops"]

STAT_FIELDS = ["query_seeds", "runs", "hits", "hit_runs", "segments", "groups", "jobs", "maybe",
               "selected", "traceback_cells", "ms_seed", "ms_join", "ms_sort_hits",
               "ms_ungapped", "ms_bands", "ms_sw", "ms_select", "ms_traceback", "ms_total",
               "ms_host", "ms_translate"]


def sw16_variants():
    """Compiled packed score kernels (dmr_sw16_variants): dicts of class name,
    width (bound), lanes per DP pair and threads per block."""
    n = C.c_uint32()
    _check(lib().dmr_sw16_variants(None, 0, C.byref(n)))
    out = np.zeros(4 * n.value, np.int32)
    _check(lib().dmr_sw16_variants(out.ctypes.data, n.value, C.byref(n)))
    return [dict(cls=PACKED_CLASSES[int(r[0])], width=int(r[1]), lanes=int(r[2]), threads=int(r[3]))
            for r in out.reshape(-1, 4)]


class DmrError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"dmr error {code}: {msg}")
        self.code = code


class SeedEntry(C.Structure):
    _fields_ = [("key", C.c_uint64), ("seq", C.c_uint32), ("pos", C.c_uint32)]


class SeedHit(C.Structure):
    _fields_ = [("query", C.c_uint32), ("qpos", C.c_uint32), ("subject", C.c_uint32),
                ("spos", C.c_uint32)]


class Segment(C.Structure):
    _fields_ = [("q_start", C.c_uint64), ("s_start", C.c_uint64), ("len", C.c_uint64),
                ("score", C.c_int32), ("reserved", C.c_int32)]


class HspStats(C.Structure):
    _fields_ = [(n, C.c_uint32) for n in
                ("identities", "positives", "mismatches", "gap_openings", "gaps", "length")]


class Aligned(C.Structure):
    _fields_ = [("found", C.c_int32), ("score", C.c_int32), ("q_begin", C.c_uint64),
                ("q_end", C.c_uint64), ("s_begin", C.c_uint64), ("s_end", C.c_uint64),
                ("stats", HspStats), ("n_ops", C.c_uint32), ("reserved", C.c_uint32)]


class SpanJob(C.Structure):
    _fields_ = [("q_off", C.c_uint64), ("q_len", C.c_uint64), ("s_off", C.c_uint64),
                ("s_len", C.c_uint64), ("d0", C.c_int64), ("d1", C.c_int64)]


class _Cfg(C.Structure):
    _fields_ = [("sensitivity", C.c_char_p), ("evalue", C.c_double),
                ("max_target_seqs", C.c_uint32), ("max_hsps", C.c_uint32),
                ("index_chunks", C.c_int32), ("gap_open", C.c_int32), ("gap_extend", C.c_int32),
                ("min_orf", C.c_int32), ("seed_cut", C.c_double)]


@dataclasses.dataclass
class SearchConfig:
    """search::SearchConfig defaults (src/search.rs:36-48, src/cli.rs:125-140)."""
    sensitivity: str = "default"
    evalue: float = 0.001
    max_target_seqs: int = 25
    max_hsps: int = 1
    index_chunks: int = -1
    gap_open: int = -1
    gap_extend: int = -1
    min_orf: int = -1
    seed_cut: float = -1.0

    def _c(self):
        self._keep = self.sensitivity.encode()
        return _Cfg(self._keep, self.evalue, self.max_target_seqs, self.max_hsps,
                    self.index_chunks, self.gap_open, self.gap_extend, self.min_orf,
                    self.seed_cut)


_lib = None


def lib():
    """Loads libdmr.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is missing: run diamond-rust_amd/build.py")
        L = C.CDLL(str(LIB_PATH))
        P, U64, I64, U32, I32, D = C.c_void_p, C.c_uint64, C.c_int64, C.c_uint32, C.c_int32, C.c_double
        sig = {
            "dmr_last_error": ([], C.c_char_p), "dmr_version": ([], C.c_char_p),
            "dmr_ctx_create": ([C.c_int, P], C.c_int), "dmr_ctx_destroy": ([P], None),
            "dmr_ctx_share_db": ([P, P], C.c_int),
            "dmr_device_count": ([P], C.c_int), "dmr_band_for": ([U64], U32),
            "dmr_ctx_set_tuning": ([P, C.c_char_p, D], C.c_int),
            "dmr_ctx_get_tuning": ([P, C.c_char_p, P], C.c_int),
            "dmr_ctx_load_tuning": ([P, C.c_char_p], C.c_int),
            "dmr_sw16_variants": ([P, U32, P], C.c_int),
            "dmr_evalue": ([I32, I32, I32, U64, U64, U64, P, P], C.c_int),
            "dmr_seed_build": ([P, P, P, U32, C.c_char_p, C.c_char_p, U32, D, P, U64, P], C.c_int),
            "dmr_seed_join": ([P, P, U64, P, U64, U64, P, U64, P], C.c_int),
            "dmr_xdrop": ([P, P, U64, P, U64, U64, U64, I32, P], C.c_int),
            "dmr_xdrop_batch": ([P, P, U64, P, U64, P, P, U64, I32, P], C.c_int),
            "dmr_extend_span": ([P, P, U64, P, U64, I64, I64, I32, I32, P, P, U32], C.c_int),
            "dmr_extend": ([P, P, U64, P, U64, U64, U64, U64, I32, I32, P, P, U32], C.c_int),
            "dmr_extend_span_noalloc": ([P, P, U64, P, U64, I64, I64, I32, I32, P, P, U32], C.c_int),
            "dmr_extend_span_batch": ([P, P, U64, P, U64, P, U64, I32, I32, C.c_int, P, P, U64, P, P],
                                      C.c_int),
            "dmr_db_load": ([P, P, P, U32], C.c_int),
            "dmr_db_load_file": ([P, C.c_char_p, P], C.c_int),
            "dmr_db_ids": ([P, P, U64, P], C.c_int),
            "dmr_db_headers": ([P, P, U64, P], C.c_int),
            "dmr_db_build_index": ([P, C.c_char_p, I32, D, P], C.c_int),
            "dmr_blastp": ([P, P, P, U32, P, P], C.c_int),
            "dmr_blastx": ([P, P, P, U32, P, P], C.c_int),
            "dmr_six_frames": ([P, P, P, U32, I32, P, U64, P, P], C.c_int),
            "dmr_queries_load": ([P, P, P, U32], C.c_int),
            "dmr_search_loaded": ([P, P, P], C.c_int),
            "dmr_blastp_plain": ([P, P, P, U32, P, P], C.c_int),
            "dmr_result_count": ([P], I64), "dmr_result_hsps": ([P, P, P], C.c_int),
            "dmr_result_ops": ([P, I64, P, U32, P], C.c_int),
            "dmr_result_format": ([P, C.c_char_p, C.c_char_p, C.c_char_p, C.c_char_p, C.c_char_p, P, U64,
                                   P], C.c_int),
            "dmr_result_stats": ([P, P], C.c_int), "dmr_last_dp_cells": ([P, P], C.c_int),
            "dmr_last_class_profile": ([P, P, P, P], C.c_int),
            "dmr_score_layout": ([P, P, P], C.c_int),
            "dmr_result_destroy": ([P], None), "dmr_search_config_default": ([P], None),
            "dmr_fasta_parse": ([C.c_char_p, U64, P], C.c_int),
            "dmr_fasta_sizes": ([P, P, P, P, P, P], C.c_int),
            "dmr_fasta_residues": ([P, P, P, P], C.c_int),
            "dmr_fasta_text": ([P, P, P], C.c_int),
            "dmr_fasta_destroy": ([P], None),
            "dmr_server_create": ([P, U32, U64, C.c_char_p, P], C.c_int),
            "dmr_server_submit": ([P, P, U64, P, P], C.c_int),
            "dmr_server_wait": ([P, U64, I64, P, P], C.c_int),
            "dmr_server_stats": ([P, P], C.c_int),
            "dmr_server_destroy": ([P], None),
        }
        for name, (args, res) in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


def _check(rc):
    if rc != DMR_OK:
        raise DmrError(rc, lib().dmr_last_error().decode(errors="replace"))


def _ptr(a):
    return a.ctypes.data if a is not None and len(a) else None


def band_for(length: int) -> int:
    """gapped::band_for (src/gapped.rs:27-32)."""
    return int(lib().dmr_band_for(length))


def evalue(raw: int, qlen: int, slen: int, db_letters: int, gap_open=11, gap_extend=1):
    e, b = C.c_double(), C.c_double()
    _check(lib().dmr_evalue(gap_open, gap_extend, raw, qlen, slen, db_letters, C.byref(e), C.byref(b)))
    return e.value, b.value


def pack(seqs):
    """Concatenated uint8 letters + uint64 offsets (n+1) of a list of arrays."""
    seqs = [np.asarray(s, dtype=np.uint8) for s in seqs]
    off = np.zeros(len(seqs) + 1, dtype=np.uint64)
    if seqs:
        np.cumsum([len(s) for s in seqs], out=off[1:])
    letters = np.concatenate(seqs) if seqs else np.zeros(0, np.uint8)
    return np.ascontiguousarray(letters, dtype=np.uint8), off


@dataclasses.dataclass
class FastaRecords:
    """io::parse_fasta output (src/io.rs:40-102) in the library's layout:
    raw residues (upper-cased, layout characters dropped), the same encoded
    (alphabet::encode), n + 1 offsets into either, ids and headers."""
    raw: bytes
    letters: np.ndarray
    offsets: np.ndarray
    ids: list
    headers: list
    looks_like_dna: bool


def parse_fasta(data: bytes) -> FastaRecords:
    """dmr_fasta_parse of FASTA text held in memory (host only, no device)."""
    L = lib()
    h = C.c_void_p()
    _check(L.dmr_fasta_parse(data, len(data), C.byref(h)))
    try:
        nrec, nres, nid, nhd, dna = C.c_uint64(), C.c_uint64(), C.c_uint64(), C.c_uint64(), C.c_int()
        _check(L.dmr_fasta_sizes(h, C.byref(nrec), C.byref(nres), C.byref(nid), C.byref(nhd), C.byref(dna)))
        raw = C.create_string_buffer(max(1, nres.value))
        letters = np.zeros(nres.value, np.uint8)
        offsets = np.zeros(nrec.value + 1, np.uint64)
        _check(L.dmr_fasta_residues(h, raw, _ptr(letters), offsets.ctypes.data))
        ids = C.create_string_buffer(max(1, nid.value))
        heads = C.create_string_buffer(max(1, nhd.value))
        _check(L.dmr_fasta_text(h, ids, heads))
    finally:
        L.dmr_fasta_destroy(h)
    split = lambda b, n: b.raw[:n].decode().split("\n")[:-1] if n else []
    return FastaRecords(raw.raw[:nres.value], letters, offsets, split(ids, nid.value),
                        split(heads, nhd.value), bool(dna.value))


def ids_text(ids) -> bytes:
    """Id or header text for dmr_result_format (each line ended by a newline,
    so empty entries keep their place) from a list, or a str/bytes as is."""
    if isinstance(ids, bytes):
        return ids
    if isinstance(ids, str):
        return ids.encode()
    return "".join(i + "\n" for i in ids).encode()


class Result:
    def __init__(self, handle):
        self._h = C.c_void_p(handle)

    def __del__(self):
        if getattr(self, "_h", None) and lib is not None:
            try:
                lib().dmr_result_destroy(self._h)
            except Exception:
                pass
            self._h = None

    def __len__(self):
        return int(lib().dmr_result_count(self._h))

    def hsps(self):
        n = len(self)
        ints = np.zeros((n, 16), dtype=np.int64)
        dbl = np.zeros((n, 2), dtype=np.float64)
        if n:
            _check(lib().dmr_result_hsps(self._h, ints.ctypes.data, dbl.ctypes.data))
        out = {k: ints[:, i] for i, k in enumerate(HSP_FIELDS)}
        out["bits"] = dbl[:, 0]
        out["evalue"] = dbl[:, 1]
        return out

    def ops(self, i):
        n = C.c_uint32()
        buf = np.zeros(1 << 16, dtype=np.uint32)
        _check(lib().dmr_result_ops(self._h, i, buf.ctypes.data, len(buf), C.byref(n)))
        return buf[: n.value].copy()

    def format(self, query_ids, subject_ids, fields="", query_titles=None, subject_titles=None) -> str:
        """outfmt-6 text; ids and titles (full headers, for qtitle/stitle) as
        lists of str, or as one newline-separated str/bytes (as dmr_db_ids and
        dmr_db_headers give them)."""
        q = ids_text(query_ids)
        s = ids_text(subject_ids)
        qt = ids_text(query_titles) if query_titles is not None else None
        st = ids_text(subject_titles) if subject_titles is not None else None
        f = fields.encode() if isinstance(fields, str) else " ".join(fields).encode()
        n = C.c_uint64()
        # One call when the text fits a buffer of ~160 bytes per line (the
        # default fields take ~80); else the size, then the fill.
        guess = 160 * len(self) + 256
        buf = C.create_string_buffer(guess)
        rc = lib().dmr_result_format(self._h, q, s, qt, st, f, buf, guess, C.byref(n))
        if rc == DMR_E_CAPACITY:
            buf = C.create_string_buffer(n.value + 1)
            rc = lib().dmr_result_format(self._h, q, s, qt, st, f, buf, n.value + 1, C.byref(n))
        _check(rc)
        return C.string_at(buf, n.value).decode()

    def stats(self):
        v = np.zeros(len(STAT_FIELDS), dtype=np.float64)
        _check(lib().dmr_result_stats(self._h, v.ctypes.data))
        return dict(zip(STAT_FIELDS, v.tolist()))


SERVER_STATS = ["requests", "batches", "mean_batch", "max_batch", "peak_admitted_bytes", "max_queue",
                "mean_queue_wait_ms", "budget_bytes"]


class Server:
    """dmr_server: a serving loop over a context's resident database
    (continuous batching, device-memory admission, per-request cap)."""

    def __init__(self, ctx: "Context", workers=2, hbm_budget=8 << 30, presets="default,sensitive,very-sensitive"):
        h = C.c_void_p()
        _check(lib().dmr_server_create(ctx._h, workers, hbm_budget, presets.encode() if presets else None,
                                       C.byref(h)))
        self._h = h
        self._ctx = ctx  # the context outlives the server
        self._cfgs = {}

    def submit(self, letters, cfg: "SearchConfig | None" = None) -> int:
        cfg = cfg or SearchConfig()
        q = np.ascontiguousarray(letters, dtype=np.uint8)
        c = cfg._c()
        t = C.c_uint64()
        _check(lib().dmr_server_submit(self._h, _ptr(q), len(q), C.byref(c), C.byref(t)))
        return t.value

    def wait(self, ticket: int, timeout_ms: int = -1):
        """(Result, latency_ms) of a request; raises DmrError (DMR_E_TIMEOUT
        when it is not done within timeout_ms)."""
        h = C.c_void_p()
        lat = C.c_double()
        _check(lib().dmr_server_wait(self._h, ticket, timeout_ms, C.byref(h), C.byref(lat)))
        return Result(h.value), lat.value

    def stats(self):
        v = np.zeros(8, dtype=np.float64)
        _check(lib().dmr_server_stats(self._h, v.ctypes.data))
        return dict(zip(SERVER_STATS, v.tolist()))

    def close(self):
        if getattr(self, "_h", None):
            lib().dmr_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Context:
    """One GPU, its stream, the resident database and seed index."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        _check(lib().dmr_ctx_create(device, C.byref(h)))
        self._h = h

    def close(self):
        if self._h:
            lib().dmr_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_tuning(self, key: str, value: float):
        """dmr_ctx_set_tuning: slice_hits, profile, and the score-pass kernels
        (lanes_<class>, threads_<class>, transpose_<class>, transpose,
        merged_mask, f16, score_plan, graph, fuse)."""
        _check(lib().dmr_ctx_set_tuning(self._h, key.encode(), float(value)))

    def load_tuning(self, path):
        """dmr_ctx_load_tuning: the tuning file (None = none)."""
        _check(lib().dmr_ctx_load_tuning(self._h, None if path is None else str(path).encode()))

    def layout(self) -> dict:
        """Score-pass settings of the last search (every layout key)."""
        keys = [f"{k}_{c}" for c in PACKED_CLASSES for k in ("lanes", "threads", "transpose")]
        keys += ["merged_mask", "f16", "score_plan", "graph", "fuse"]
        return {k: int(self.get_tuning(k)) for k in keys}

    def get_tuning(self, key: str) -> float:
        v = C.c_double()
        _check(lib().dmr_ctx_get_tuning(self._h, key.encode(), C.byref(v)))
        return v.value

    # -- database / search ------------------------------------------------
    def share_db(self, other: "Context"):
        """dmr_ctx_share_db: use other's resident database and seed indexes
        (no copy); the two contexts may then search concurrently."""
        _check(lib().dmr_ctx_share_db(self._h, other._h))
        self._db = getattr(other, "_db", None)

    def db_load(self, letters, offsets):
        letters = np.ascontiguousarray(letters, dtype=np.uint8)
        offsets = np.ascontiguousarray(offsets, dtype=np.uint64)
        self._db = (letters, offsets)
        _check(lib().dmr_db_load(self._h, _ptr(letters), offsets.ctypes.data, len(offsets) - 1))

    def db_load_file(self, path) -> int:
        """Database::open of a DRDMND01 file (src/db.rs:159-224); returns its sequence count."""
        n = C.c_uint32()
        _check(lib().dmr_db_load_file(self._h, os.fsencode(path), C.byref(n)))
        self._db = None
        return n.value

    def db_ids(self) -> str:
        """Ids of the file-loaded database, one per line (each followed by a newline)."""
        need = C.c_uint64()
        rc = lib().dmr_db_ids(self._h, None, 0, C.byref(need))
        if rc not in (DMR_OK, DMR_E_CAPACITY):
            _check(rc)
        buf = C.create_string_buffer(max(1, need.value))
        _check(lib().dmr_db_ids(self._h, buf, need.value, C.byref(need)))
        return buf.raw[:need.value].decode()

    def db_headers(self) -> str:
        """Headers of the file-loaded database, one per line (the stitle text)."""
        need = C.c_uint64()
        rc = lib().dmr_db_headers(self._h, None, 0, C.byref(need))
        if rc not in (DMR_OK, DMR_E_CAPACITY):
            _check(rc)
        buf = C.create_string_buffer(max(1, need.value))
        _check(lib().dmr_db_headers(self._h, buf, need.value, C.byref(need)))
        return buf.raw[:need.value].decode()

    def build_index(self, sensitivity="default", index_chunks=-1, seed_cut=-1.0) -> float:
        ms = C.c_double()
        _check(lib().dmr_db_build_index(self._h, sensitivity.encode(), index_chunks, seed_cut,
                                        C.byref(ms)))
        return ms.value

    def blastp(self, letters, offsets, cfg: SearchConfig | None = None) -> Result:
        cfg = cfg or SearchConfig()
        letters = np.ascontiguousarray(letters, dtype=np.uint8)
        offsets = np.ascontiguousarray(offsets, dtype=np.uint64)
        c = cfg._c()
        h = C.c_void_p()
        _check(lib().dmr_blastp(self._h, _ptr(letters), offsets.ctypes.data, len(offsets) - 1,
                                C.byref(c), C.byref(h)))
        return Result(h.value)

    def queries_load(self, letters, offsets):
        """Uploads a blastp query batch that stays resident for search_loaded."""
        letters = np.ascontiguousarray(letters, dtype=np.uint8)
        offsets = np.ascontiguousarray(offsets, dtype=np.uint64)
        _check(lib().dmr_queries_load(self._h, _ptr(letters), offsets.ctypes.data, len(offsets) - 1))

    def search_loaded(self, cfg: SearchConfig | None = None) -> Result:
        cfg = cfg or SearchConfig()
        c = cfg._c()
        h = C.c_void_p()
        _check(lib().dmr_search_loaded(self._h, C.byref(c), C.byref(h)))
        return Result(h.value)

    def blastp_plain(self, letters, offsets, cfg: SearchConfig | None = None) -> Result:
        """The per-call drop-in form (one kernel launch per reference call)."""
        cfg = cfg or SearchConfig()
        letters = np.ascontiguousarray(letters, dtype=np.uint8)
        offsets = np.ascontiguousarray(offsets, dtype=np.uint64)
        c = cfg._c()
        h = C.c_void_p()
        _check(lib().dmr_blastp_plain(self._h, _ptr(letters), offsets.ctypes.data, len(offsets) - 1,
                                      C.byref(c), C.byref(h)))
        return Result(h.value)

    def blastx(self, reads, cfg: SearchConfig | None = None) -> Result:
        cfg = cfg or SearchConfig()
        data = "".join(reads).encode()
        off = np.zeros(len(reads) + 1, dtype=np.uint64)
        np.cumsum([len(r) for r in reads], out=off[1:])
        c = cfg._c()
        h = C.c_void_p()
        buf = C.create_string_buffer(data, len(data) + 1)
        _check(lib().dmr_blastx(self._h, buf, off.ctypes.data, len(reads), C.byref(c), C.byref(h)))
        return Result(h.value)

    def six_frames(self, reads, min_orf=-1):
        """dmr_six_frames: the six masked frames of every read (a list of
        6 * len(reads) uint8 arrays, frame k of read r at 6r + k)."""
        data = "".join(reads).encode()
        off = np.zeros(len(reads) + 1, dtype=np.uint64)
        np.cumsum([len(r) for r in reads], out=off[1:])
        fo = np.zeros(6 * len(reads) + 1, dtype=np.uint64)
        n = C.c_uint64()
        buf = C.create_string_buffer(data, len(data) + 1)
        rc = lib().dmr_six_frames(self._h, buf, off.ctypes.data, len(reads), min_orf, None, 0,
                                  fo.ctypes.data, C.byref(n))
        if rc not in (DMR_OK, DMR_E_CAPACITY):
            _check(rc)
        letters = np.zeros(max(1, n.value), np.uint8)
        _check(lib().dmr_six_frames(self._h, buf, off.ctypes.data, len(reads), min_orf, letters.ctypes.data,
                                    n.value, fo.ctypes.data, C.byref(n)))
        return [letters[fo[k]:fo[k + 1]].copy() for k in range(6 * len(reads))]

    def class_profile(self):
        """Per score-kernel class of the last search: packed 16-bit by band width
        (w25..w129, merged bands m97..m513), wave-per-DP int32 by width bound
        (wave64..wave2048), and the row-exact int32 kernel: jobs, DP cells,
        kernel ms."""
        names = SCORE_CLASSES
        j = np.zeros(len(names), np.uint64)
        c = np.zeros(len(names), np.uint64)
        m = np.zeros(len(names), np.float64)
        _check(lib().dmr_last_class_profile(self._h, j.ctypes.data, c.ctypes.data, m.ctypes.data))
        lay = self.layout()
        return {n: dict(jobs=int(j[i]), cells=int(c[i]), ms=float(m[i]),
                        kernel=_class_kernel(n, lay.get(f"lanes_{n}", 0), lay.get(f"threads_{n}", 0),
                                             bool(lay["f16"])))
                for i, n in enumerate(names)}

    def score_layout(self):
        """(split mask, f16) of the packed score classes: bit c set = class c
        split over several lanes per DP pair (two for w81, four otherwise);
        f16 True = pairs in f16, False = saturating int16."""
        mask, f16 = C.c_uint32(), C.c_uint32()
        _check(lib().dmr_score_layout(self._h, C.byref(mask), C.byref(f16)))
        return mask.value, bool(f16.value)

    def last_dp_cells(self) -> int:
        v = C.c_uint64()
        _check(lib().dmr_last_dp_cells(self._h, C.byref(v)))
        return v.value

    # -- drop-in calls ------------------------------------------------------
    def seed_build(self, seqs, shape_code, reduction_def, base=0, cut=0.0):
        letters, off = pack(seqs)
        n = C.c_uint64()
        cap = max(1, int(len(letters)))
        out = (SeedEntry * cap)()
        _check(lib().dmr_seed_build(self._h, _ptr(letters), off.ctypes.data, len(seqs),
                                    shape_code.encode(), reduction_def.encode(), base, cut,
                                    out, cap, C.byref(n)))
        a = np.ctypeslib.as_array(out)[: n.value]
        return a["key"].copy(), a["seq"].copy(), a["pos"].copy()

    def seed_join(self, q, r, max_expand=4096):
        """q, r: (keys, seqs, poss) tuples sorted by key. Returns an (n, 4)
        array of (query, qpos, subject, spos)."""
        def entries(t):
            k, s, p = (np.asarray(x) for x in t)
            arr = (SeedEntry * max(1, len(k)))()
            a = np.ctypeslib.as_array(arr)
            a["key"][: len(k)] = k
            a["seq"][: len(k)] = s
            a["pos"][: len(k)] = p
            return arr, len(k)
        qa, nq = entries(q)
        ra, nr = entries(r)
        n = C.c_uint64()
        rc = lib().dmr_seed_join(self._h, qa, nq, ra, nr, max_expand, None, 0, C.byref(n))
        if rc not in (DMR_OK, DMR_E_CAPACITY):
            _check(rc)
        out = (SeedHit * max(1, n.value))()
        _check(lib().dmr_seed_join(self._h, qa, nq, ra, nr, max_expand, out, n.value, C.byref(n)))
        a = np.ctypeslib.as_array(out)[: n.value]
        return np.stack([a["query"], a["qpos"], a["subject"], a["spos"]], axis=1).astype(np.int64)

    def xdrop(self, query, subject, qa, sa, xdrop):
        q = np.ascontiguousarray(query, dtype=np.uint8)
        s = np.ascontiguousarray(subject, dtype=np.uint8)
        o = Segment()
        _check(lib().dmr_xdrop(self._h, _ptr(q), len(q), _ptr(s), len(s), qa, sa, xdrop, C.byref(o)))
        return (o.q_start, o.s_start, o.len, o.score)

    def xdrop_batch(self, query, subject, qa, sa, xdrop):
        q = np.ascontiguousarray(query, dtype=np.uint8)
        s = np.ascontiguousarray(subject, dtype=np.uint8)
        qa = np.ascontiguousarray(qa, dtype=np.uint32)
        sa = np.ascontiguousarray(sa, dtype=np.uint32)
        out = (Segment * max(1, len(qa)))()
        _check(lib().dmr_xdrop_batch(self._h, _ptr(q), len(q), _ptr(s), len(s), _ptr(qa), _ptr(sa),
                                     len(qa), xdrop, out))
        a = np.ctypeslib.as_array(out)[: len(qa)]
        return np.stack([a["q_start"], a["s_start"], a["len"], a["score"]], axis=1).astype(np.int64)

    def extend_span(self, query, subject, d0, d1, gap_open=11, gap_extend=1):
        """gapped::extend_span: None or a dict with score, coordinates, stats, ops."""
        q = np.ascontiguousarray(query, dtype=np.uint8)
        s = np.ascontiguousarray(subject, dtype=np.uint8)
        out = Aligned()
        ops = np.zeros(len(q) + len(s) + 1, dtype=np.uint32)
        _check(lib().dmr_extend_span(self._h, _ptr(q), len(q), _ptr(s), len(s), d0, d1, gap_open,
                                     gap_extend, C.byref(out), ops.ctypes.data, len(ops)))
        return _aligned_dict(out, ops[: out.n_ops])

    def extend_span_noalloc(self, query, subject, d0, d1, gap_open=11, gap_extend=1):
        """dmr_extend_span_noalloc: extend_span without device allocation."""
        q = np.ascontiguousarray(query, dtype=np.uint8)
        s = np.ascontiguousarray(subject, dtype=np.uint8)
        out = Aligned()
        ops = np.zeros(len(q) + len(s) + 1, dtype=np.uint32)
        _check(lib().dmr_extend_span_noalloc(self._h, _ptr(q), len(q), _ptr(s), len(s), d0, d1, gap_open,
                                             gap_extend, C.byref(out), ops.ctypes.data, len(ops)))
        return _aligned_dict(out, ops[: out.n_ops])

    def extend(self, query, subject, anchor, band, gap_open=11, gap_extend=1):
        q = np.ascontiguousarray(query, dtype=np.uint8)
        s = np.ascontiguousarray(subject, dtype=np.uint8)
        out = Aligned()
        ops = np.zeros(len(q) + len(s) + 1, dtype=np.uint32)
        _check(lib().dmr_extend(self._h, _ptr(q), len(q), _ptr(s), len(s), anchor[0], anchor[1],
                                band, gap_open, gap_extend, C.byref(out), ops.ctypes.data, len(ops)))
        return _aligned_dict(out, ops[: out.n_ops])

    def extend_span_batch(self, qbuf, sbuf, jobs, traceback=True, gap_open=11, gap_extend=1):
        """jobs: iterable of (q_off, q_len, s_off, s_len, d0, d1)."""
        qb = np.ascontiguousarray(qbuf, dtype=np.uint8)
        sb = np.ascontiguousarray(sbuf, dtype=np.uint8)
        jobs = list(jobs)
        J = (SpanJob * max(1, len(jobs)))(*[SpanJob(*j) for j in jobs])
        out = (Aligned * max(1, len(jobs)))()
        offs = np.zeros(len(jobs) + 1, dtype=np.uint64)
        need = C.c_uint64()
        cap = sum(j[1] + j[3] + 1 for j in jobs) if traceback else 0
        pool = np.zeros(max(1, cap), dtype=np.uint32)
        _check(lib().dmr_extend_span_batch(self._h, _ptr(qb), len(qb), _ptr(sb), len(sb), J,
                                           len(jobs), gap_open, gap_extend, 1 if traceback else 0,
                                           out, pool.ctypes.data, len(pool), offs.ctypes.data,
                                           C.byref(need)))
        res = []
        for i in range(len(jobs)):
            res.append(_aligned_dict(out[i], pool[offs[i]:offs[i + 1]]) if traceback
                       else (out[i].score if out[i].found else None))
        return res


def _aligned_dict(a: Aligned, ops):
    if not a.found:
        return None
    st = a.stats
    return dict(score=a.score, q_begin=a.q_begin, q_end=a.q_end, s_begin=a.s_begin, s_end=a.s_end,
                identities=st.identities, positives=st.positives, mismatches=st.mismatches,
                gap_openings=st.gap_openings, gaps=st.gaps, length=st.length,
                ops=[(int(o) & 3, int(o) >> 2) for o in ops])
