"""Pins of the oracle's outfmt-6 renderings (src/hsp.rs:84-241,
src/output.rs:150-223) that the reference's own tests leave open.

- The reference's hsp tests, transcribed (frame numbering, nt_range).
- nt_range against translation: the nucleotides it names translate back to
  the frame's letters (translation is pinned by src/translate.rs's tests).
- CIGAR, BTOP and the aligned strings of hand-made alignments, written out
  by hand from the BLAST conventions the reference documents (BTOP: a number
  for identical columns, query letter then subject letter otherwise).
- On a real search with every field: each line's fields agree with one
  another and with the sequences (CIGAR lengths equal the coordinate spans,
  BTOP expands to qseq/sseq, qseq/sseq without gaps are the sequence slices,
  counts and percentages recomputed from the columns, the raw score rescored
  from BLOSUM62 with gap cost 11 + L, bits from lambda and K).
- qtitle / stitle are the FASTA headers of the reference's e2e files.
"""
import json
import math
import pathlib
import re

import numpy as np
import pytest

GOLD = pathlib.Path(__file__).resolve().parent / "golden"
SYM = "ARNDCQEGHILKMFPSTWYVBJZX*"
ALL = ("qseqid sseqid pident length mismatch gapopen qstart qend sstart send evalue bitscore score "
       "nident positive gaps ppos qlen slen qcovhsp scovhsp cigar btop qseq sseq qframe qstrand "
       "full_sseq qseq_translated")


def test_frames_follow_diamond_numbering(orc):  # src/hsp.rs:248-266
    for frame, expect in [(0, 1), (1, 2), (2, 3), (3, -1), (4, -2), (5, -3)]:
        assert orc.q_frame(frame) == expect


def test_reverse_frame_maps_back_to_forward_strand(orc):  # src/hsp.rs:268-287
    assert orc.nt_range(3, 2, 5, 30) == (30 - 15, 30 - 6)


def test_nt_range_translates_back(orc):
    """The nucleotides nt_range names are the codons of the frame's letters:
    forward frames read them directly, reverse frames on the reverse
    complement."""
    rng = np.random.default_rng(5)
    comp = str.maketrans("ACGT", "TGCA")
    for _ in range(40):
        n = int(rng.integers(9, 80))
        dna = "".join(rng.choice(list("ACGT"), size=n))
        frames = orc.six_frames(dna, min_orf=1)
        for f in range(6):
            L = len(frames[f])
            if L == 0:
                continue
            qb = int(rng.integers(0, L))
            qe = int(rng.integers(qb, L + 1))
            a, b = orc.nt_range(f, qb, qe, n)
            assert b - a == 3 * (qe - qb) and 0 <= a <= b <= n
            piece = dna[a:b] if f < 3 else dna[a:b].translate(comp)[::-1]
            got = orc.six_frames(piece, min_orf=1)[0] if piece else np.zeros(0, np.uint8)
            assert list(got) == list(frames[f][qb:qe]), (f, qb, qe)


def ops(*runs):
    code = {"M": 0, "I": 1, "D": 2}
    return np.array([code[o] | (n << 2) for o, n in runs], np.uint32)


def test_alignment_text_by_hand(orc):
    """A 3M 1I 3M 1D 2M alignment with one substitution, written out by hand."""
    q = orc.encode("ACDEFGHIK")
    s = orc.encode("ACDFAHWIK")
    o = ops(("M", 3), ("I", 1), ("M", 3), ("D", 1), ("M", 2))
    assert orc.alignment_text("cigar", o, q, 0, s, 0) == "3M1I3M1D2M"
    assert orc.alignment_text("qseq", o, q, 0, s, 0) == "ACDEFGH-IK"
    assert orc.alignment_text("sseq", o, q, 0, s, 0) == "ACD-FAHWIK"
    # 3 identical, E against a gap, F=F, G/A, H=H, a gap against W, then IK.
    assert orc.alignment_text("btop", o, q, 0, s, 0) == "3E-1GA1-W2"
    # Offsets into longer sequences; runs of several gaps.
    q2 = orc.encode("WWACDEEFG")
    s2 = orc.encode("YACDFGGG")
    o2 = ops(("M", 3), ("I", 2), ("M", 2), ("D", 2))
    assert orc.alignment_text("cigar", o2, q2, 2, s2, 1) == "3M2I2M2D"
    assert orc.alignment_text("qseq", o2, q2, 2, s2, 1) == "ACDEEFG--"
    assert orc.alignment_text("sseq", o2, q2, 2, s2, 1) == "ACD--FGGG"
    assert orc.alignment_text("btop", o2, q2, 2, s2, 1) == "3E-E-2-G-G"
    # X (unknown) against X counts as identical in BTOP, as in the codes.
    q3, s3 = orc.encode("AXB"), orc.encode("AXZ")
    assert orc.alignment_text("btop", ops(("M", 3)), q3, 0, s3, 0) == "2BZ"


def blosum62():
    """BLOSUM62 in NCBI letter order (the reference's matrix data)."""
    return json.loads((GOLD / "blosum62.json").read_text())["scores"]


def rows(text):
    names = ALL.split()
    return [dict(zip(names, line.split("\t"))) for line in text.splitlines()]


def close_to(field, value):
    """A fmt_double field (floor at >= 100, one decimal below) holds value."""
    x = float(field)
    if value >= 100:
        return x == math.floor(value)
    return abs(x - value) <= 0.05 + 1e-9


def expand_btop(btop, qslice):
    """BTOP back to (qseq, sseq): a number copies that many query letters to
    both strings; a pair gives the query then the subject character."""
    qa, sa, qi, k = [], [], 0, 0
    while k < len(btop):
        if btop[k].isdigit():
            j = k
            while j < len(btop) and btop[j].isdigit():
                j += 1
            n = int(btop[k:j])
            qa.append(qslice[qi:qi + n])
            sa.append(qslice[qi:qi + n])
            qi += n
            k = j
        else:
            a, b = btop[k], btop[k + 1]
            qa.append(a)
            sa.append(b)
            qi += a != "-"
            k += 2
    return "".join(qa), "".join(sa)


@pytest.mark.parametrize("sens", ["sensitive", "default"])
def test_fields_agree_on_a_search(orc, synth, sens):
    mat = blosum62()
    sc = {(a, b): mat[i][j] for i, a in enumerate(SYM) for j, b in enumerate(SYM)}
    w = synth.make("tiny")
    ref = orc.search(w.db_letters, w.db_offsets, w.q_letters, w.q_offsets, sensitivity=sens, threads=4)
    out = rows(ref.format(w.q_ids, w.db_ids, ALL))
    assert len(out) > 10
    qmap = {q: i for i, q in enumerate(w.q_ids)}
    smap = {s: i for i, s in enumerate(w.db_ids)}
    dec = lambda a: "".join(SYM[c] for c in a)
    gaps_total = 0
    for r in out:
        qi, si = qmap[r["qseqid"]], smap[r["sseqid"]]
        qseq_full = dec(w.q_letters[w.q_offsets[qi]:w.q_offsets[qi + 1]])
        sseq_full = dec(w.db_letters[w.db_offsets[si]:w.db_offsets[si + 1]])
        qb, qe, sb, se = int(r["qstart"]) - 1, int(r["qend"]), int(r["sstart"]) - 1, int(r["send"])
        assert int(r["qlen"]) == len(qseq_full) and int(r["slen"]) == len(sseq_full)
        assert r["full_sseq"] == sseq_full and r["qseq_translated"] == qseq_full[qb:qe]
        assert r["qframe"] == "1" and r["qstrand"] == "+"
        # CIGAR: M + I runs span the query range, M + D runs the subject range.
        runs = [(int(n), op) for n, op in re.findall(r"(\d+)([MID])", r["cigar"])]
        assert "".join(f"{n}{op}" for n, op in runs) == r["cigar"]
        assert sum(n for n, op in runs if op in "MI") == qe - qb
        assert sum(n for n, op in runs if op in "MD") == se - sb
        qa, sa = r["qseq"], r["sseq"]
        assert len(qa) == len(sa) == int(r["length"]) == sum(n for n, _ in runs)
        assert qa.replace("-", "") == qseq_full[qb:qe] and sa.replace("-", "") == sseq_full[sb:se]
        assert expand_btop(r["btop"], qseq_full[qb:qe]) == (qa, sa)
        cols = [(a, b) for a, b in zip(qa, sa) if a != "-" and b != "-"]
        ident = sum(a == b for a, b in cols)
        pos = sum(sc[(a, b)] > 0 for a, b in cols)
        ngaps = sum(c == "-" for c in qa + sa)
        gaps_total += ngaps
        opens = len(re.findall(r"-+", qa)) + len(re.findall(r"-+", sa))
        assert int(r["nident"]) == ident and int(r["mismatch"]) == len(cols) - ident
        assert int(r["positive"]) == pos and int(r["gaps"]) == ngaps and int(r["gapopen"]) == opens
        L = len(qa)
        assert close_to(r["pident"], ident * 100.0 / L) and close_to(r["ppos"], pos * 100.0 / L)
        assert close_to(r["qcovhsp"], (qe - qb) * 100.0 / len(qseq_full))
        assert close_to(r["scovhsp"], (se - sb) * 100.0 / len(sseq_full))
        # Raw score: BLOSUM62 over aligned columns, 11 + L per gap run.
        raw = sum(sc[(a, b)] for a, b in cols)
        for g in re.findall(r"-+", qa) + re.findall(r"-+", sa):
            raw -= 11 + len(g)
        assert int(r["score"]) == raw
        assert close_to(r["bitscore"], (0.267 * raw - math.log(0.041)) / math.log(2.0))
    assert gaps_total > 0  # gapped alignments were covered


def test_titles_are_the_fasta_headers(orc, e2e):
    """qtitle / stitle print the whole header (src/output.rs:201-202), ids
    only its first token."""
    ref = orc.search(e2e["db_letters"], e2e["db_offsets"], e2e["q_letters"], e2e["q_offsets"],
                     sensitivity="sensitive", threads=4)
    text = ref.format(e2e["q_ids"], e2e["db_ids"], "qseqid qtitle sseqid stitle",
                      qtitles=e2e["q_headers"], stitles=e2e["db_headers"])
    lines = [l.split("\t") for l in text.splitlines()]
    assert len(lines) == 7
    by_id = dict(zip(e2e["db_ids"], e2e["db_headers"]))
    for qid, qt, sid, st in lines:
        assert qt == e2e["q_headers"][e2e["q_ids"].index(qid)]
        assert st == by_id[sid] and st.startswith(sid + " ") and len(st) > len(sid) + 5
