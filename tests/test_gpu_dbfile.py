"""DRDMND01 database files (src/db.rs): dmr_db_load_file against dmr_db_load
on the same letters, the ids it reads, and Database::open's error cases.

write_db below lays a file out as db::write_db does (src/db.rs:86-149): a
72-byte header, n + 1 letter offsets, n + 1 header offsets, the encoded
letters, then the headers.
"""
import struct

import numpy as np
import pytest

from test_gpu_search import EXTENDED

pytestmark = pytest.mark.gpu


def write_db(path, headers, seqs, magic=b"DRDMND01", version=1, total_letters=None, lead=b""):
    """total_letters: the header's letter count (default: the letters
    written); lead: bytes before the first sequence in the letters section
    (the letter offsets then start at len(lead))."""
    n = len(seqs)
    so = np.zeros(n + 1, np.uint64)
    np.cumsum([len(s) for s in seqs], out=so[1:])
    so += len(lead)
    hb = [h.encode() if isinstance(h, str) else h for h in headers]
    do = np.zeros(n + 1, np.uint64)
    np.cumsum([len(h) for h in hb], out=do[1:])
    seq_off_start = 72
    def_off_start = seq_off_start + 8 * (n + 1)
    letters_start = def_off_start + 8 * (n + 1)
    defs_start = letters_start + int(so[-1])
    tl = int(so[-1]) - len(lead) if total_letters is None else total_letters
    head = magic + struct.pack("<II7Q", version, 1, n, tl, int(do[-1]), seq_off_start,
                               def_off_start, letters_start, defs_start)
    body = (so.tobytes() + do.tobytes() + lead +
            b"".join(np.asarray(s, np.uint8).tobytes() for s in seqs))
    data = head + body + b"".join(hb)
    with open(path, "wb") as f:
        f.write(data)
    return data


def split(w):
    off = w["db_offsets"].astype(np.int64)
    return [w["db_letters"][off[i]:off[i + 1]] for i in range(len(off) - 1)]


@pytest.fixture
def fresh(dmr):
    c = dmr.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("sens", ["default", "sensitive"])
def test_file_search_matches_letters(dmr, ctx, fresh, e2e, tmp_path, sens):
    """A search over the file-loaded database prints what the same letters
    loaded with dmr_db_load print, with the ids the file holds."""
    write_db(tmp_path / "db.dmnd", e2e["db_ids"], split(e2e))
    n = fresh.db_load_file(tmp_path / "db")  # db::resolve appends .dmnd
    assert n == len(e2e["db_ids"])
    ids = fresh.db_ids()
    assert ids == "".join(i + "\n" for i in e2e["db_ids"])
    cfg = dmr.SearchConfig(sensitivity=sens)
    got = fresh.blastp(e2e["q_letters"], e2e["q_offsets"], cfg).format(e2e["q_ids"], ids[:-1].split("\n"),
                                                                       EXTENDED)
    ctx.db_load(e2e["db_letters"], e2e["db_offsets"])
    exp = ctx.blastp(e2e["q_letters"], e2e["q_offsets"], cfg).format(e2e["q_ids"], e2e["db_ids"], EXTENDED)
    assert got == exp
    assert got.count("\n") == {"default": 4, "sensitive": 7}[sens]


def test_ids_follow_seq_id(fresh, tmp_path):
    """alphabet::seq_id (src/alphabet.rs:87-93): leading '>' and ' ' trimmed,
    cut at the first ASCII whitespace (space, tab, form feed, CR; not vertical
    tab); a header that is not UTF-8 reads as empty (Database::header)."""
    heads = [b"sp|P1|A desc", b">> x\ty", b"", b"a\x0bb c", b"\xff\xfe bad", b"q\x0cz", b"\xc3\xa9t\xc3\xa9 z",
             b"\xed\xa0\x80 s", b"  >lead"]
    exp = ["sp|P1|A", "x", "", "a\x0bb", "", "q", "été", "", "lead"]
    seqs = [np.array([i % 20, 3, 4], np.uint8) for i in range(len(heads))]
    write_db(tmp_path / "ids.dmnd", heads, seqs)
    assert fresh.db_load_file(tmp_path / "ids.dmnd") == len(heads)
    assert fresh.db_ids() == "".join(e + "\n" for e in exp)


def corrupt_cases(good):
    n = struct.unpack_from("<Q", good, 16)[0]
    cases = {
        "short": (good[:71], "file too short"),
        "magic": (b"DRDMND02" + good[8:], "bad magic"),
        "version": (good[:8] + struct.pack("<I", 2) + good[12:], "database version 2 is not supported (expected 1)"),
        "truncated": (good[:-1], "truncated or corrupt database"),
        "seq_table": (good[:40] + struct.pack("<Q", len(good)) + good[48:], "corrupt offset table"),
        "def_table": (good[:48] + struct.pack("<Q", len(good) - 8) + good[56:], "corrupt offset table"),
    }
    # Letter offsets that run backwards, and a letter code outside the alphabet.
    b = bytearray(good)
    struct.pack_into("<Q", b, 72 + 8, 10**6)
    cases["offsets"] = (bytes(b), "corrupt offset table")
    b = bytearray(good)
    letters_start = struct.unpack_from("<Q", good, 56)[0]
    b[letters_start + 1] = 30
    cases["letter"] = (bytes(b), "corrupt letters section")
    assert n > 1
    return cases


def test_open_errors(dmr, fresh, e2e, tmp_path):
    """Database::open's checks (src/db.rs:172-209) and db::resolve's message;
    a failed load leaves the previous database resident."""
    good = write_db(tmp_path / "good.dmnd", e2e["db_ids"], split(e2e))
    fresh.db_load_file(tmp_path / "good.dmnd")
    ids = fresh.db_ids()
    cfg = dmr.SearchConfig(sensitivity="sensitive")
    before = fresh.blastp(e2e["q_letters"], e2e["q_offsets"], cfg).format(e2e["q_ids"], e2e["db_ids"])
    for name, (data, msg) in corrupt_cases(good).items():
        p = tmp_path / f"{name}.dmnd"
        p.write_bytes(data)
        with pytest.raises(dmr.DmrError) as ei:
            fresh.db_load_file(p)
        assert msg in str(ei.value), name
        assert ei.value.code == dmr.DMR_E_ARG
    for missing in [tmp_path / "nothere", tmp_path / "good.fa"]:
        with pytest.raises(dmr.DmrError) as ei:
            fresh.db_load_file(missing)
        assert f"Database file not found: {missing}" in str(ei.value)
    assert fresh.db_ids() == ids
    after = fresh.blastp(e2e["q_letters"], e2e["q_offsets"], cfg).format(e2e["q_ids"], e2e["db_ids"])
    assert after == before


def test_letters_outside_alphabet_rejected(dmr, fresh, e2e):
    """Encoded letters are alphabet indexes below 25 (src/alphabet.rs:15-16);
    the entry points refuse other codes before any device work."""
    bad = e2e["db_letters"].copy()
    bad[5] = 25
    with pytest.raises(dmr.DmrError, match="outside the 25-letter alphabet"):
        fresh.db_load(bad, e2e["db_offsets"])
    with pytest.raises(dmr.DmrError, match="no database file|not loaded from a file"):
        fresh.db_ids()
    fresh.db_load(e2e["db_letters"], e2e["db_offsets"])
    q = e2e["q_letters"].copy()
    q[-1] = 200
    with pytest.raises(dmr.DmrError, match="outside the 25-letter alphabet"):
        fresh.blastp(q, e2e["q_offsets"], dmr.SearchConfig())


def test_drop_in_calls_refuse_codes_outside_alphabet(dmr, fresh):
    """The per-call entry points check their letters the same way."""
    q = np.array([0, 1, 2, 3, 4, 5, 6, 7], np.uint8)
    bad = q.copy()
    bad[3] = 31
    with pytest.raises(dmr.DmrError, match="outside the 25-letter alphabet"):
        fresh.xdrop(bad, q, 2, 2, 20)
    with pytest.raises(dmr.DmrError, match="outside the 25-letter alphabet"):
        fresh.extend_span(q, bad, -2, 2)
    with pytest.raises(dmr.DmrError, match="outside the 25-letter alphabet"):
        fresh.seed_build([bad], "111101011", "A KR EDNQ C G H ILVM FYW P ST")
    assert fresh.extend_span(q, q, -2, 2)["score"] > 0


@pytest.mark.parametrize("sens", ["default", "sensitive"])
def test_fasta_ingress_search(dmr, fresh, orc, e2e, sens):
    """FASTA text -> dmr_fasta_parse -> database and queries -> dmr_blastp ->
    outfmt 6 with the parsed ids equals the oracle's search of the same files."""
    import gzip
    import pathlib
    g = pathlib.Path(__file__).resolve().parent / "golden" / "e2e"
    db = dmr.parse_fasta(gzip.open(g / "db_prot.fa.gz", "rb").read())
    qs = dmr.parse_fasta(gzip.open(g / "query.fa.gz", "rb").read())
    assert np.array_equal(db.letters, e2e["db_letters"]) and np.array_equal(qs.letters, e2e["q_letters"])
    fresh.db_load(db.letters, db.offsets)
    got = fresh.blastp(qs.letters, qs.offsets, dmr.SearchConfig(sensitivity=sens)).format(qs.ids, db.ids, EXTENDED)
    ref = orc.search(e2e["db_letters"], e2e["db_offsets"], e2e["q_letters"], e2e["q_offsets"],
                     sensitivity=sens, threads=8)
    assert got == ref.format(e2e["q_ids"], e2e["db_ids"], EXTENDED)
    assert got.count("\n") == {"default": 4, "sensitive": 7}[sens]


def test_headers_and_titles(dmr, fresh, orc, e2e, tmp_path):
    """dmr_db_headers gives Database::header of each sequence (the bytes as
    text, empty when not UTF-8); qtitle / stitle print the full headers
    (src/output.rs:201-202) as the oracle does from the reference's files."""
    write_db(tmp_path / "db.dmnd", e2e["db_headers"], split(e2e))
    fresh.db_load_file(tmp_path / "db.dmnd")
    heads = fresh.db_headers()
    assert heads == "".join(h + "\n" for h in e2e["db_headers"])
    fields = "qseqid qtitle sseqid stitle score evalue"
    got = fresh.blastp(e2e["q_letters"], e2e["q_offsets"], dmr.SearchConfig(sensitivity="sensitive")).format(
        e2e["q_ids"], fresh.db_ids(), fields, e2e["q_headers"], heads)
    ref = orc.search(e2e["db_letters"], e2e["db_offsets"], e2e["q_letters"], e2e["q_offsets"],
                     sensitivity="sensitive", threads=8)
    exp = ref.format(e2e["q_ids"], e2e["db_ids"], fields, e2e["q_headers"], e2e["db_headers"])
    assert got == exp and got.count("\n") == 7
    r = fresh.blastp(e2e["q_letters"], e2e["q_offsets"], dmr.SearchConfig(sensitivity="sensitive"))
    with pytest.raises(dmr.DmrError, match="qtitle needs the query headers"):
        r.format(e2e["q_ids"], fresh.db_ids(), "qseqid qtitle")
    with pytest.raises(dmr.DmrError, match="stitle needs the database headers"):
        r.format(e2e["q_ids"], fresh.db_ids(), "stitle", e2e["q_headers"])
    # Headers that are not UTF-8 read as empty; others keep every byte.
    hs = [b"a b\tc", b"\xff bad", b"", "été x".encode()]
    write_db(tmp_path / "h.dmnd", hs, [np.array([1, 2, 3], np.uint8)] * 4)
    fresh.db_load_file(tmp_path / "h.dmnd")
    assert fresh.db_headers() == "a b\tc\n\n\nété x\n"
    write_db(tmp_path / "nl.dmnd", [b"x\ny", b"z"], [np.array([1, 2, 3], np.uint8)] * 2)
    fresh.db_load_file(tmp_path / "nl.dmnd")
    with pytest.raises(dmr.DmrError, match="holds a newline"):
        fresh.db_headers()


def test_header_letter_count_sets_evalues(dmr, fresh, orc, e2e, tmp_path):
    """E-values use Database::total_letters, the header's count
    (src/db.rs:199,236-238; src/search.rs:158), even where it differs from
    the letters the offsets hold; letters before the first offset belong to
    no sequence (Database::seq)."""
    n_letters = int(e2e["db_offsets"][-1])
    cfg = dmr.SearchConfig(sensitivity="sensitive", evalue=10.0)
    for total, lead in [(n_letters * 10, b""), (n_letters // 3, b""), (n_letters, b"\x05\x06\x07")]:
        write_db(tmp_path / "t.dmnd", e2e["db_ids"], split(e2e), total_letters=total, lead=lead)
        fresh.db_load_file(tmp_path / "t.dmnd")
        got = fresh.blastp(e2e["q_letters"], e2e["q_offsets"], cfg).format(e2e["q_ids"], e2e["db_ids"], EXTENDED)
        ref = orc.search(e2e["db_letters"], e2e["db_offsets"], e2e["q_letters"], e2e["q_offsets"],
                         sensitivity="sensitive", threads=8, evalue=10.0, header_letters=total)
        exp = ref.format(e2e["q_ids"], e2e["db_ids"], EXTENDED)
        assert got == exp and len(ref) > 0, (total, lead)
    assert len(ref) > 0
