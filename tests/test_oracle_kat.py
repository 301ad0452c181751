"""The oracle against the reference's own known-answer tests.

Each test cites the reference test it transcribes (file:line); the expected
values are the reference's literals.
"""
import json
import math
import pathlib

import numpy as np
import pytest

GOLD = pathlib.Path(__file__).resolve().parent / "golden"
MURPHY10 = "A KR EDNQ C G H ILVM FYW P ST"
NONE20 = "A R N D C Q E G H I L K M F P S T W Y V"
LN2 = math.log(2.0)


def test_alphabet_roundtrip(orc):  # src/alphabet.rs:99-112
    syms = "ARNDCQEGHILKMFPSTWYVBJZX*"
    assert list(orc.encode(syms)) == list(range(25))
    assert list(orc.encode("a")) == [0]
    assert list(orc.encode("U")) == [4]
    assert list(orc.encode("?")) == [23]
    assert list(orc.encode("AC-DE\nFG")) == [0, 4, 3, 6, 13, 7]


def test_reductions(orc):  # src/reduction.rs:86-116
    m = np.zeros(25, np.uint8)
    assert orc.lib().orc_reduction_map(b"default", m.ctypes.data) == 10
    enc = {c: int(orc.encode(c)[0]) for c in "AKRLM"}
    assert m[enc["A"]] == 0 and m[enc["K"]] == 1 and m[enc["R"]] == 1
    assert m[enc["L"]] == m[enc["M"]]
    assert orc.lib().orc_reduction_map(b"sensitive", m.ctypes.data) == 10
    assert orc.lib().orc_reduction_map(b"very-sensitive", m.ctypes.data) == 9
    assert orc.lib().orc_reduction_map(b"ultra-sensitive", m.ctypes.data) == 10


def test_shape_keys(orc):  # src/shape.rs:231-265
    s = orc.encode("MKWV")
    assert orc.shape_key("111", s, 0, NONE20) == 12 * 400 + 11 * 20 + 17
    assert orc.shape_key("111", s, 2, NONE20) is None
    assert orc.shape_key("101", orc.encode("ACD"), 0, NONE20) == 3
    assert orc.shape_key("11", orc.encode("LI"), 0, MURPHY10) == orc.shape_key("11", orc.encode("LL"), 0, MURPHY10)
    # Letters >= 20 break seeds (src/shape.rs:68).
    assert orc.shape_key("11", orc.encode("AX"), 0, NONE20) is None


def test_shape_presets(orc):  # src/shape.rs:267-299
    d = orc.shapes_for("default")
    assert d == ["111101110111", "111011010010111"]
    assert all(c.count("1") == 10 for c in d)
    assert len(orc.shapes_for("sensitive", 4)) == 4
    assert orc.shapes_for("more-sensitive") == orc.shapes_for("sensitive")
    assert len(orc.shapes_for("sensitive")) == 16
    v = orc.shapes_for("very-sensitive")
    assert len(v) == 16 and v[0] == "11110111" and v[0].count("1") == 7


def test_seed_build_and_join(orc):  # src/seed.rs:175-196
    a = orc.encode("MKWVTFGCPD")
    cut = 0.8 * LN2 * 4
    q = orc.seed_build([a], "1111", MURPHY10, 0, cut)
    r = orc.seed_build([a], "1111", MURPHY10, 7, cut)
    assert len(q[0]) == 7
    hits = orc.seed_join(q, r, 8, 4096)
    h = hits[(hits[:, 0] == 0) & (hits[:, 2] == 7)]
    assert len(h) == 7
    h = h[np.argsort(h[:, 1], kind="stable")]
    assert (h[0, 1], h[0, 3]) == (0, 0)


def test_complexity_filter(orc):  # src/seed.rs:198-209
    cut = 0.8 * LN2 * 8
    s8 = "11111111"
    assert not orc.seed_is_complex(orc.encode("AAAAAAAAAA"), s8, MURPHY10, cut)
    assert orc.seed_is_complex(orc.encode("MKWVTFGCPD"), s8, MURPHY10, cut)
    assert not orc.seed_is_complex(orc.encode("AAAAAACC"), s8, MURPHY10, cut)
    assert orc.seed_is_complex(orc.encode("AAAAAACC"), s8, MURPHY10, 0.8)


def test_chunks(orc):  # src/seed.rs:211-216
    assert orc.chunks(10, 4) == [(0, 3), (3, 6), (6, 8), (8, 10)]
    assert sum(b - a for a, b in orc.chunks(10, 4)) == 10
    assert orc.chunks(0, 4) == []


def test_xdrop(orc):  # src/ungapped.rs:99-129
    q = orc.encode("MKWVTFISLLFLFSSAYSR")
    qs, ss, ln, sc = orc.xdrop(q, q, 0, 0, 40)
    assert qs == 0 and ln == len(q)
    assert sc == sum(orc.lib().orc_matrix_score(int(c), int(c)) for c in q)
    seg = orc.xdrop(orc.encode("WWWWWWWWWWCCCCCCCCCC"), orc.encode("WWWWWWWWWWDDDDDDDDDD"), 0, 0, 5)
    assert seg[2] == 10 and seg[3] == 110
    seg = orc.xdrop(orc.encode("AAAAAA"), orc.encode("AAAAAA"), 0, 0, 30)
    assert (seg[0], seg[1], seg[2]) == (0, 0, 6)


def test_gapped(orc):  # src/gapped.rs:299-338
    q = orc.encode("MKWVTFISLLFLFSSAYSR")
    score = sum(orc.lib().orc_matrix_score(int(c), int(c)) for c in q)
    a = orc.extend_span(q, q, -30, 30)
    assert a["score"] == score
    assert (a["q_begin"], a["s_begin"], a["q_end"], a["s_end"]) == (0, 0, len(q), len(q))
    assert a["identities"] == len(q) and a["mismatches"] == 0 and a["gaps"] == 0
    s = orc.encode("MKWVTFISLLLFSSAYSR")
    a = orc.extend_span(q, s, -30, 30)
    assert a["gaps"] > 0 and a["identities"] == len(s)
    assert a["score"] < sum(orc.lib().orc_matrix_score(int(c), int(c)) for c in s)
    assert orc.extend_span(orc.encode("WWWWWWWWWW"), orc.encode("AAAAAAAAAA"), -15, 15) is None
    assert [orc.band_for(n) for n in (10, 50, 120, 400, 900)] == [12, 16, 30, 64, 64]


def test_matrix(orc):  # src/matrix.rs:278-321
    assert orc.lib().orc_matrix_score(0, 0) == 4
    assert orc.lib().orc_matrix_score(17, 17) == 11
    assert orc.lib().orc_matrix_score(0, 17) == -3
    b = orc.bit_score(100)
    assert abs(b - 43.1) < 0.1
    assert orc.raw_score(b) == 100
    assert abs(orc.evalue(198, 497, 601, 2263996) / 7.60e-17 - 1) < 0.01
    assert orc.lib().orc_matrix_ok(11, 1) == 1
    assert orc.lib().orc_matrix_ok(42, 7) == 0


def test_matrix_table_matches_reference(orc):
    g = json.loads((GOLD / "blosum62.json").read_text())
    for a in range(25):
        for b in range(25):
            assert orc.lib().orc_matrix_score(a, b) == g["scores"][a][b]


def test_matrix_is_symmetric():
    """The packed score kernels may run a DP with the subject on the rows
    (the best score of a band is the same either way when the score matrix
    is symmetric, which the reference's BLOSUM62 is)."""
    g = json.loads((GOLD / "blosum62.json").read_text())["scores"]
    assert all(g[a][b] == g[b][a] for a in range(25) for b in range(25))


def test_log_erfc(orc):  # src/alp.rs:387-396
    assert orc.log_erfc(0.0) == 0.0
    assert abs(orc.log_erfc(1.0) - (-1.8496055099332485)) < 1e-12
    assert abs(orc.log_erfc(4.0) - (-17.987778312103007)) < 1e-11
    assert abs(orc.log_erfc(5.0) - (-27.200889545537436)) < 1e-11
    assert abs(orc.log_erfc(-5.0) - 0.6931471805591766) < 1e-12
    assert abs(orc.log_erfc(-20.0) - LN2) < 1e-15


def test_alp_evalues(orc):  # src/alp.rs:398-428
    for raw, slen, exp in [(818, 454, 3.21e-105), (262, 456, 3.89e-25), (198, 601, 7.60e-17)]:
        assert abs(orc.evalue(raw, 497, slen, 2263996) / exp - 1) < 0.01
    e = orc.evalue(60, 300, 300, 1_000_000)
    blast = 300 * 1e6 * math.exp(-0.267 * 60 + math.log(0.041))
    assert e > 0 and math.isfinite(e) and abs(math.log(e / blast)) < 3


def test_formatting(orc):  # src/output.rs:282-296
    assert orc.fmt_evalue(0.0) == "0.0"
    assert orc.fmt_evalue(0.001) == "1.00e-3"
    assert orc.fmt_evalue(4.7e-45) == "4.70e-45"
    assert orc.fmt_evalue(3.21e-105) == "3.21e-105"
    assert orc.fmt_double(319.7) == "319"
    assert orc.fmt_double(105.49) == "105"
    assert orc.fmt_double(80.87) == "80.9"
    assert orc.fmt_double(22.581) == "22.6"


def test_presets(orc):  # src/search.rs:410-454
    for name, cut in [("faster", 0.9), ("fast", 0.9), ("default", 0.8), ("mid-sensitive", 1.0),
                      ("sensitive", 1.0), ("more-sensitive", 1.0), ("very-sensitive", 1.0),
                      ("ultra-sensitive", 1.0)]:
        assert orc.lib().orc_seed_cut(name.encode(), -1.0) == cut
    assert orc.lib().orc_seed_cut(b"very-sensitive", 0.5) == 0.5
    for name, n in [("faster", 4), ("fast", 4), ("default", 4), ("mid-sensitive", 4),
                    ("sensitive", 4), ("more-sensitive", 4), ("very-sensitive", 1),
                    ("ultra-sensitive", 1)]:
        assert orc.lib().orc_index_chunks(name.encode(), -1) == n
    assert orc.lib().orc_index_chunks(b"default", 7) == 7


def test_translation(orc):  # src/translate.rs:150-183
    aa = lambda s: list(orc.encode(s))
    f = orc.six_frames("ATGGCCTAA", min_orf=1)
    assert list(f[0]) == aa("MA*")
    assert list(orc.six_frames("TTTTTCCTT", min_orf=1)[0]) == aa("FFL")
    assert list(orc.six_frames("atgaaa", min_orf=1)[0]) == aa("MK")
    assert list(orc.six_frames("ATGNNTAAA", min_orf=1)[0]) == aa("MXK")
    f = orc.six_frames("ATGAAATTT", min_orf=1)
    assert list(f[0]) == aa("MKF") and list(f[3]) == aa("KFH")
    # mask_short_orfs(MK*MKF*WWWW, 3) -> XX*MKF*WWWW: frame 0 of the DNA below
    # translates to MK*MKF*WWWW.
    nt = "ATGAAATAAATGAAATTTTAATGGTGGTGGTGG"
    assert list(orc.six_frames(nt, min_orf=3)[0]) == aa("XX*MKF*WWWW")
