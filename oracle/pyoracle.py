"""ctypes binding of the oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: loaded by tests/, the smoke check and the CPU
baseline of bench.py, to compare results with; never by the product.
"""
from __future__ import annotations

import ctypes as C
import pathlib
import subprocess

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
LIB = HERE / "_build" / "liboracle.so"

_lib = None


def build():
    subprocess.run(["make", "-C", str(HERE), "-j4"], check=True, capture_output=True)


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        L = C.CDLL(str(LIB))
        P, U64, I64, U32, D = C.c_void_p, C.c_uint64, C.c_int64, C.c_uint32, C.c_double
        sig = {
            "orc_encode": ([C.c_char_p, P, C.c_int], C.c_int),
            "orc_shape_key": ([C.c_char_p, P, C.c_int, C.c_int, C.c_char_p, P], C.c_int),
            "orc_seed_is_complex": ([P, C.c_char_p, C.c_char_p, D], C.c_int),
            "orc_shapes_for": ([C.c_char_p, C.c_int, P, C.c_int], C.c_int),
            "orc_reduction_map": ([C.c_char_p, P], C.c_int),
            "orc_seed_cut": ([C.c_char_p, D], D),
            "orc_index_chunks": ([C.c_char_p, C.c_long], C.c_long),
            "orc_seed_build": ([P, P, U32, C.c_char_p, C.c_char_p, U32, D, P, P, P, I64], I64),
            "orc_seed_join": ([P, P, P, I64, P, P, P, I64, U32, I64, P, I64], I64),
            "orc_chunks": ([I64, I64, P], C.c_int),
            "orc_matrix_score": ([C.c_int, C.c_int], C.c_int),
            "orc_matrix_ok": ([C.c_int, C.c_int], C.c_int),
            "orc_xdrop": ([P, I64, P, I64, I64, I64, C.c_int, P], None),
            "orc_extend_span": ([P, I64, P, I64, I64, I64, C.c_int, C.c_int, P, P, C.c_int, P], C.c_int),
            "orc_band_for": ([I64], C.c_int),
            "orc_evalue": ([C.c_int, I64, I64, U64], D),
            "orc_bit_score": ([C.c_int], D),
            "orc_raw_score": ([D], C.c_int),
            "orc_log_erfc": ([D], D),
            "orc_fmt_double": ([D, P, C.c_int], C.c_int),
            "orc_fmt_evalue": ([D, P, C.c_int], C.c_int),
            "orc_six_frames": ([C.c_char_p, P, I64, P, C.c_long], C.c_int),
            "orc_q_frame": ([C.c_int], C.c_int),
            "orc_nt_range": ([C.c_int, I64, I64, I64, P], None),
            "orc_alignment_text": ([C.c_int, P, I64, P, I64, P, I64, P, I64], I64),
            "orc_search": ([P, P, U32, P, P, U32, C.c_char_p, C.c_int, D, C.c_int, C.c_int,
                            C.c_long, D, C.c_int, C.c_long, U64], P),
            "orc_result_count": ([P], I64),
            "orc_result_hsps": ([P, P, P], None),
            "orc_result_counters": ([P, P], None),
            "orc_result_outfmt6": ([P, C.c_char_p, C.c_char_p, C.c_char_p, C.c_char_p, C.c_char_p,
                                    P, I64], I64),
            "orc_result_free": ([P], None),
        }
        for name, (args, res) in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


def encode(s: str) -> np.ndarray:
    out = np.zeros(len(s) + 1, dtype=np.uint8)
    n = lib().orc_encode(s.encode(), out.ctypes.data, len(out))
    return out[:n].copy()


def _p(a):
    return a.ctypes.data if a is not None and len(a) else None


def shape_key(code, seq, pos, reduction):
    seq = np.ascontiguousarray(seq, dtype=np.uint8)
    k = C.c_uint64()
    ok = lib().orc_shape_key(code.encode(), _p(seq), len(seq), pos, reduction.encode(), C.byref(k))
    return k.value if ok else None


def seed_is_complex(seq, code, reduction, cut):
    seq = np.ascontiguousarray(seq, dtype=np.uint8)
    return bool(lib().orc_seed_is_complex(_p(seq), code.encode(), reduction.encode(), cut))


def shapes_for(sens, limit=0):
    buf = C.create_string_buffer(8192)
    lib().orc_shapes_for(sens.encode(), limit, buf, len(buf))
    return [x for x in buf.value.decode().split("\n") if x]


def seed_build(seqs, code, reduction, base=0, cut=0.0):
    seqs = [np.asarray(s, dtype=np.uint8) for s in seqs]
    off = np.zeros(len(seqs) + 1, dtype=np.uint64)
    np.cumsum([len(s) for s in seqs], out=off[1:])
    letters = np.concatenate(seqs) if seqs else np.zeros(0, np.uint8)
    cap = max(1, len(letters))
    k = np.zeros(cap, np.uint64)
    s = np.zeros(cap, np.uint32)
    p = np.zeros(cap, np.uint32)
    n = lib().orc_seed_build(_p(letters), off.ctypes.data, len(seqs), code.encode(),
                             reduction.encode(), base, cut, k.ctypes.data, s.ctypes.data,
                             p.ctypes.data, cap)
    return k[:n].copy(), s[:n].copy(), p[:n].copy()


def seed_join(q, r, ntargets, max_expand=4096):
    qk, qs, qp = (np.ascontiguousarray(x) for x in q)
    rk, rs, rp = (np.ascontiguousarray(x) for x in r)
    qk, rk = qk.astype(np.uint64), rk.astype(np.uint64)
    qs, qp, rs, rp = (x.astype(np.uint32) for x in (qs, qp, rs, rp))
    n = lib().orc_seed_join(_p(qk), _p(qs), _p(qp), len(qk), _p(rk), _p(rs), _p(rp), len(rk),
                            ntargets, max_expand, None, 0)
    out = np.zeros(4 * max(n, 1), np.uint32)
    lib().orc_seed_join(_p(qk), _p(qs), _p(qp), len(qk), _p(rk), _p(rs), _p(rp), len(rk),
                        ntargets, max_expand, out.ctypes.data, n)
    return out[: 4 * n].reshape(n, 4).astype(np.int64)


def chunks(n, parts):
    out = np.zeros(2 * max(parts, 1) + 2, np.int64)
    k = lib().orc_chunks(n, parts, out.ctypes.data)
    return [(int(out[2 * i]), int(out[2 * i + 1])) for i in range(k)]


def xdrop(q, s, qa, sa, xd):
    q = np.ascontiguousarray(q, dtype=np.uint8)
    s = np.ascontiguousarray(s, dtype=np.uint8)
    o = np.zeros(4, np.int64)
    lib().orc_xdrop(_p(q), len(q), _p(s), len(s), qa, sa, xd, o.ctypes.data)
    return tuple(int(x) for x in o)


def extend_span(q, s, d0, d1, go=11, ge=1):
    """None or a dict like the product's extend_span (plus 'cells')."""
    q = np.ascontiguousarray(q, dtype=np.uint8)
    s = np.ascontiguousarray(s, dtype=np.uint8)
    o = np.zeros(12, np.int64)
    ops = np.zeros(len(q) + len(s) + 2, np.uint32)
    nops = C.c_int()
    ok = lib().orc_extend_span(_p(q), len(q), _p(s), len(s), d0, d1, go, ge, o.ctypes.data,
                               ops.ctypes.data, len(ops), C.byref(nops))
    if ok <= 0:
        return None
    keys = ["score", "q_begin", "q_end", "s_begin", "s_end", "identities", "positives",
            "mismatches", "gap_openings", "gaps", "length", "cells"]
    d = {k: int(v) for k, v in zip(keys, o)}
    d["ops"] = [(int(x) & 3, int(x) >> 2) for x in ops[: nops.value]]
    return d


def dp_cells(q, s, d0, d1):
    q = np.ascontiguousarray(q, dtype=np.uint8)
    s = np.ascontiguousarray(s, dtype=np.uint8)
    o = np.zeros(12, np.int64)
    ops = np.zeros(len(q) + len(s) + 2, np.uint32)
    nops = C.c_int()
    lib().orc_extend_span(_p(q), len(q), _p(s), len(s), d0, d1, 11, 1, o.ctypes.data,
                          ops.ctypes.data, len(ops), C.byref(nops))
    return int(o[11])


def band_for(n):
    return lib().orc_band_for(n)


def evalue(raw, qlen, slen, db):
    return lib().orc_evalue(raw, qlen, slen, db)


def bit_score(raw):
    return lib().orc_bit_score(raw)


def raw_score(bits):
    return lib().orc_raw_score(bits)


def log_erfc(x):
    return lib().orc_log_erfc(x)


def fmt_double(x):
    b = C.create_string_buffer(64)
    lib().orc_fmt_double(x, b, 64)
    return b.value.decode()


def fmt_evalue(x):
    b = C.create_string_buffer(64)
    lib().orc_fmt_evalue(x, b, 64)
    return b.value.decode()


def six_frames(nt, min_orf=-1):
    cap = 2 * len(nt) + 8  # six frames of about len / 3 letters each
    out = np.zeros(cap, np.uint8)
    lens = np.zeros(6, np.int64)
    n = lib().orc_six_frames(nt.encode(), out.ctypes.data, cap, lens.ctypes.data, min_orf)
    if n > cap:
        raise RuntimeError("six_frames output larger than its buffer")
    frames, o = [], 0
    for L in lens:
        frames.append(out[o:o + L].copy())
        o += L
    return frames


COUNTER_FIELDS = ["query_seeds", "ref_seeds", "hits", "xdrops", "segments", "bands", "dps",
                  "dp_cells", "clipped_dps", "band_tie_events", "anchor_tie_events", "t_seed",
                  "t_join", "t_align"]


class SearchResult:
    def __init__(self, h):
        self._h = h

    def __del__(self):
        if self._h:
            lib().orc_result_free(self._h)
            self._h = None

    def __len__(self):
        return int(lib().orc_result_count(self._h))

    def hsps(self):
        n = len(self)
        ints = np.zeros((n, 16), np.int64)
        d = np.zeros((n, 2), np.float64)
        if n:
            lib().orc_result_hsps(self._h, ints.ctypes.data, d.ctypes.data)
        names = ["query", "subject", "frame", "q_begin", "q_end", "s_begin", "s_end", "score",
                 "identities", "positives", "mismatches", "gap_openings", "gaps", "length", "n_ops"]
        out = {k: ints[:, i] for i, k in enumerate(names)}
        out["bits"] = d[:, 0]
        out["evalue"] = d[:, 1]
        return out

    def counters(self):
        v = np.zeros(14, np.float64)
        lib().orc_result_counters(self._h, v.ctypes.data)
        return dict(zip(COUNTER_FIELDS, v.tolist()))

    def format(self, qids, sids, fields="", qtitles=None, stitles=None):
        """outfmt-6 text; qtitles/stitles (lists of headers) feed qtitle/stitle."""
        lines = lambda v: "".join(x + "\n" for x in v).encode()
        q, s = lines(qids), lines(sids)
        qt = lines(qtitles) if qtitles is not None else None
        st = lines(stitles) if stitles is not None else None
        f = fields.encode() if isinstance(fields, str) else " ".join(fields).encode()
        n = lib().orc_result_outfmt6(self._h, q, s, qt, st, f, None, 0)
        buf = C.create_string_buffer(n + 1)
        lib().orc_result_outfmt6(self._h, q, s, qt, st, f, buf, n + 1)
        return buf.raw[:n].decode()


def search(db_letters, db_offsets, q_letters, q_offsets, sensitivity="default", threads=1,
           evalue=1e-3, max_target_seqs=25, max_hsps=1, index_chunks=-1, seed_cut=-1.0,
           translated=False, min_orf=-1, header_letters=0):
    dbl = np.ascontiguousarray(db_letters, dtype=np.uint8)
    dbo = np.ascontiguousarray(db_offsets, dtype=np.uint64)
    ql = np.ascontiguousarray(q_letters, dtype=np.uint8)
    qo = np.ascontiguousarray(q_offsets, dtype=np.uint64)
    h = lib().orc_search(_p(dbl), dbo.ctypes.data, len(dbo) - 1, _p(ql), qo.ctypes.data,
                         len(qo) - 1, sensitivity.encode(), threads, evalue, max_target_seqs,
                         max_hsps, index_chunks, seed_cut, 1 if translated else 0, min_orf,
                         header_letters)
    return SearchResult(h)


def q_frame(frame):
    """Hsp::q_frame (src/hsp.rs:84-90)."""
    return int(lib().orc_q_frame(frame))


def nt_range(frame, q_begin, q_end, dna_len):
    """Hsp::nt_range (src/hsp.rs:97-110)."""
    out = np.zeros(2, np.int64)
    lib().orc_nt_range(frame, q_begin, q_end, dna_len, out.ctypes.data)
    return int(out[0]), int(out[1])


def alignment_text(which, ops, q, q_begin, s, s_begin):
    """CIGAR / BTOP / qseq / sseq (src/hsp.rs:152-241) of ops (op | count << 2,
    op 0 M, 1 I, 2 D) over encoded q and s from (q_begin, s_begin)."""
    code = {"cigar": 0, "btop": 1, "qseq": 2, "sseq": 3}[which]
    o = np.ascontiguousarray(ops, dtype=np.uint32)
    qa = np.ascontiguousarray(q, dtype=np.uint8)
    sa = np.ascontiguousarray(s, dtype=np.uint8)
    L = lib()
    n = L.orc_alignment_text(code, _p(o), len(o), _p(qa), q_begin, _p(sa), s_begin, None, 0)
    buf = C.create_string_buffer(n + 1)
    L.orc_alignment_text(code, _p(o), len(o), _p(qa), q_begin, _p(sa), s_begin, buf, n + 1)
    return buf.raw[:n].decode()


# ---- FASTA ingress (pure Python) -------------------------------------------

# char::is_whitespace, which str::trim uses: the Unicode White_Space property
# (PropList.txt). Python's str.strip differs (it also strips U+001C..U+001F).
WHITE_SPACE = frozenset([*range(0x09, 0x0E), 0x20, 0x85, 0xA0, 0x1680, *range(0x2000, 0x200B), 0x2028,
                         0x2029, 0x202F, 0x205F, 0x3000])
_UPPER = bytes.maketrans(b"abcdefghijklmnopqrstuvwxyz", b"ABCDEFGHIJKLMNOPQRSTUVWXYZ")
_LAYOUT = b" \t\n\x0c\r-."  # u8::is_ascii_whitespace, '-' and '.'


def rust_trim(s: str) -> str:
    a, b = 0, len(s)
    while a < b and ord(s[a]) in WHITE_SPACE:
        a += 1
    while b > a and ord(s[b - 1]) in WHITE_SPACE:
        b -= 1
    return s[a:b]


def _finish(header: str, seq: bytearray):
    """io::finish (src/io.rs:76-94): id = first ASCII-whitespace token, else seq<len>."""
    n = 0
    while n < len(header) and header[n] not in " \t\n\x0c\r":
        n += 1
    rid = header[:n] if n else f"seq{max(len(seq), 1)}"
    return header, rid, bytes(seq)


def parse_fasta(data: bytes):
    """io::parse_fasta (src/io.rs:40-74) with split_lines (src/io.rs:96-102):
    a list of (header, id, raw residues); raises ValueError with the
    reference's message. from_utf8_lossy is Python's 'replace' decoding (both
    substitute maximal invalid subparts)."""
    out = []
    header, seq, seen = None, bytearray(), False
    for line in data.split(b"\n"):
        if line.endswith(b"\r"):
            line = line[:-1]
        if line[:1] == b">":
            if header is not None:
                out.append(_finish(header, seq))
                seq = bytearray()
            seen = True
            header = rust_trim(line[1:].decode("utf-8", "replace"))
        elif header is not None:
            seq += line.translate(_UPPER, _LAYOUT)
        elif line and line[:1] not in (b";", b"#"):
            raise ValueError("malformed FASTA: sequence data found before the first '>' header")
    if header is not None:
        out.append(_finish(header, seq))
    if not seen and data:
        raise ValueError("malformed FASTA: no '>' header found")
    return out


def looks_like_dna(records) -> bool:
    """io::looks_like_dna (src/io.rs:104-121) over parse_fasta's records."""
    canonical = total = 0
    for _, _, raw in records:
        part = raw[:20000]
        total += len(part)
        canonical += sum(part.count(c) for c in b"ACGTUN")
        if total >= 20000:
            break
    return total > 0 and canonical * 100 >= total * 85
