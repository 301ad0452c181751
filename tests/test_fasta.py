"""FASTA ingress (dmr_fasta_parse, host code) against the oracle's
restatement of io::parse_fasta, and both against the reference's own tests
(src/io.rs:124-153) and the Unicode Standard's U+FFFD substitution example.
"""
import gzip
import pathlib

import numpy as np
import pytest
from hypothesis import given, settings, strategies as st

from conftest import read_fasta_gz

GOLD = pathlib.Path(__file__).resolve().parent / "golden"


def lib_parse(dmr, data):
    """(records, looks_like_dna, letters) from the library; records as the oracle's tuples."""
    f = dmr.parse_fasta(data)
    off = f.offsets.astype(np.int64)
    recs = [(h, i, f.raw[off[k]:off[k + 1]]) for k, (h, i) in enumerate(zip(f.headers, f.ids))]
    assert len(recs) == len(off) - 1
    return recs, f.looks_like_dna, f.letters


def both(dmr, orc, data):
    try:
        exp = orc.parse_fasta(data)
    except ValueError as e:
        with pytest.raises(dmr.DmrError) as ei:
            dmr.parse_fasta(data)
        assert str(e) in str(ei.value)
        assert ei.value.code == dmr.DMR_E_ARG
        return None
    got, dna, letters = lib_parse(dmr, data)
    assert got == exp
    assert dna == orc.looks_like_dna(exp)
    assert bytes(letters) == bytes(encode_table(orc)[np.frombuffer(b"".join(r for _, _, r in exp), np.uint8)])
    return got


_TAB = None


def encode_table(orc):
    """alphabet::encode per byte from the oracle's encoder; bytes outside
    ASCII map to X (23), the table's default (src/alphabet.rs:24-44)."""
    global _TAB
    if _TAB is None:
        t = np.full(256, 23, np.uint8)
        for b in range(1, 128):
            e = orc.encode(chr(b))
            if len(e) == 1:
                t[b] = e[0]
        _TAB = t
    return _TAB


def test_reference_io_tests(dmr, orc):
    """src/io.rs:128-152: multi-line CRLF records, data before the header, DNA guess."""
    for parse in (orc.parse_fasta, lambda d: lib_parse(dmr, d)[0]):
        recs = parse(b">seq1 first description\r\nACGT\r\nACGT\r\n>seq2\nAAAA\n")
        assert len(recs) == 2
        assert recs[0] == ("seq1 first description", "seq1", b"ACGTACGT")
        assert recs[1][1:] == ("seq2", b"AAAA")
    with pytest.raises(ValueError):
        orc.parse_fasta(b"ACGT\n>a\nAC\n")
    with pytest.raises(dmr.DmrError, match="sequence data found before the first '>' header"):
        dmr.parse_fasta(b"ACGT\n>a\nAC\n")
    assert lib_parse(dmr, b">a\nACGTACGTTT\n")[1] and orc.looks_like_dna(orc.parse_fasta(b">a\nACGTACGTTT\n"))
    prot = b">a\nMKWVTFISLLFLFSSAYSR\n"
    assert not lib_parse(dmr, prot)[1] and not orc.looks_like_dna(orc.parse_fasta(prot))


def test_lossy_headers_follow_unicode_table_3_8(dmr, orc):
    """The Unicode Standard, section 3.9, Table 3-8 (U+FFFD for maximal
    subparts): 61 F1 80 80 E1 80 C2 62 80 63 80 BF 64 reads as
    a FFFD FFFD FFFD b FFFD c FFFD FFFD d (String::from_utf8_lossy does the same)."""
    raw = bytes.fromhex("61F18080E180C262806380BF64")
    exp = "a���b�c��d"
    for parse in (orc.parse_fasta, lambda d: lib_parse(dmr, d)[0]):
        (h, i, r), = parse(b">" + raw + b"\nAC\n")
        assert h == exp and i == exp and r == b"AC"
    # Surrogate encodings, overlong forms, a truncated 3-byte form at the end.
    for raw, exp in [(b"\xed\xa0\x80", "�" * 3), (b"\xc0\xaf", "�" * 2), (b"\xe0\x80\xaf", "�" * 3),
                     (b"x\xe1\x80", "x�"), (b"\xf4\x90\x80\x80", "�" * 4), (b"\xf0\x9f\x98\x80", "\U0001F600")]:
        for parse in (orc.parse_fasta, lambda d: lib_parse(dmr, d)[0]):
            assert parse(b">" + raw + b"\n")[0][0] == exp, raw


def test_header_trim_and_ids(dmr, orc):
    """str::trim removes White_Space code points (U+0085, U+00A0, U+3000 ...)
    but not U+001C or U+200B; the id is cut at ASCII whitespace only; an
    empty id becomes seq<len> with len at least 1 (src/io.rs:76-94)."""
    cases = [
        (">\u0085  id rest　 \nAC\n", "id rest", "id"),
        (">\u001cid​\nAC\n", "\u001cid​", "\u001cid​"),
        (">a b c\nAC\n", "a b c", "a b"),
        (">\nACD\n", "", "seq3"),
        (">  \n\n", "", "seq1"),
        ("> x\x0by z\n", "x\x0by z", "x\x0by"),
    ]
    for text, head, rid in cases:
        for parse in (orc.parse_fasta, lambda d: lib_parse(dmr, d)[0]):
            (h, i, _), = parse(text.encode())
            assert (h, i) == (head, rid), text


def test_comments_empty_and_errors(dmr, orc):
    assert both(dmr, orc, b"") == []
    assert both(dmr, orc, b"\n\n;c\n#c\n>a\nM\n") == [("a", "a", b"M")]
    both(dmr, orc, b";only a comment\n")  # no '>' header found
    both(dmr, orc, b"\n \n>a\n")          # a blank line is data before the header
    assert both(dmr, orc, b">a\n;x#.-\tm \r\n>b") == [("a", "a", b";X#M"), ("b", "b", b"")]


_PIECES = [b">", b"\n", b"\r\n", b"\r", b" ", b"\t", b"-", b".", b";", b"#", b"ACGT", b"mkw", b"x*", b"\x0b",
           b"\x0c", b"\xc3\xa9", b"\xe1\x80", b"\xff", b"\x85", b"\xc2\xa0", b"\xe3\x80\x80", b"\x00", b"id"]


@settings(max_examples=400, deadline=None)
@given(st.lists(st.sampled_from(_PIECES), max_size=60))
def test_random_texts_match_oracle(dmr, orc, pieces):
    both(dmr, orc, b"".join(pieces))


def test_large_text_in_parts(dmr, orc):
    """Several MB: the library cuts the text at '>' lines and parses the parts
    on its worker threads; the records equal the oracle's whole-text parse."""
    rng = np.random.default_rng(7)
    sym = np.frombuffer(b"ARNDCQEGHILKMFPSTWYVBZXacgt-.", np.uint8)
    out = [b";comment\n\n"]
    size = 0
    k = 0
    while size < 9 << 20:
        n = int(rng.integers(0, 3000))
        s = sym[rng.integers(0, len(sym), n)].tobytes()
        w = int(rng.integers(40, 121))
        eol = b"\r\n" if k % 5 == 0 else b"\n"
        body = eol.join(s[i:i + w] for i in range(0, len(s), w))
        head = b">" if k % 97 == 0 else b">r%d desc %d" % (k, n)
        rec = head + eol + body + eol + (b">\n" if k % 1000 == 999 else b"")
        out.append(rec)
        size += len(rec)
        k += 1
    data = b"".join(out)
    got = both(dmr, orc, data)
    assert len(got) >= k


def test_reference_fixture_files(dmr, orc):
    """The reference's end-to-end FASTA files parse to the records the tests'
    reader gives, with first-token ids."""
    for name in ("db_prot.fa.gz", "query.fa.gz", "subject.fa.gz"):
        data = gzip.open(GOLD / "e2e" / name, "rb").read()
        got = both(dmr, orc, data)
        exp = read_fasta_gz(GOLD / "e2e" / name)
        assert [(h, r) for h, _, r in got] == exp
        assert [i for _, i, _ in got] == [h.split()[0] for h, _ in exp]
