"""GPU parity of the drop-in calls (through the C ABI) against the oracle.

Bit-exact for every field: seeds (key, seq, pos), hits, segments and the
full gapped::Aligned (score, coordinates, statistics, run-length ops).
"""
import numpy as np
import pytest

from conftest import mutate, rand_seq

pytestmark = pytest.mark.gpu

MURPHY10 = "A KR EDNQ C G H ILVM FYW P ST"
DIAMOND9 = "A G P CH ST NDE RQK FWY ILMV"


def _eq_aligned(a, b):
    if a is None or b is None:
        return a is None and b is None
    keys = ["score", "q_begin", "q_end", "s_begin", "s_end", "identities", "positives",
            "mismatches", "gap_openings", "gaps", "length", "ops"]
    return all(a[k] == b[k] for k in keys)


def span_cases(seed, n_cases):
    rng = np.random.default_rng(seed)
    for _ in range(n_cases):
        n = int(rng.integers(1, 400))
        q = rand_seq(rng, n, canonical_only=rng.random() < 0.8)
        kind = rng.integers(0, 4)
        if kind == 0:
            s = q.copy()
        elif kind == 1:
            s = mutate(rng, q, float(rng.uniform(0.1, 0.7)), float(rng.uniform(0.0, 0.08)))
        else:
            s = rand_seq(rng, int(rng.integers(1, 500)))
        if len(s) == 0:
            s = rand_seq(rng, 1)
        m = len(s)
        # Bands inside, across and entirely outside the matrix (clamped rows).
        c = int(rng.integers(-m - 20, n + 20))
        w = int(rng.choice([0, 1, 5, 12, 30, 64, 200, 300, 700, 2000]))
        yield q, s, c - w, c + w + int(rng.integers(0, 3))


def test_extend_span_single(ctx, orc):
    for q, s, d0, d1 in span_cases(1, 60):
        assert _eq_aligned(ctx.extend_span(q, s, d0, d1), orc.extend_span(q, s, d0, d1)), (len(q), len(s), d0, d1)


def test_extend_span_reference_cases(ctx, orc):  # src/gapped.rs:299-329
    q = orc.encode("MKWVTFISLLFLFSSAYSR")
    a = ctx.extend(q, q, (0, 0), 30)
    assert a["score"] == sum(orc.lib().orc_matrix_score(int(c), int(c)) for c in q)
    assert (a["q_begin"], a["q_end"], a["identities"]) == (0, len(q), len(q))
    s = orc.encode("MKWVTFISLLLFSSAYSR")
    a = ctx.extend(q, s, (0, 0), 30)
    assert a["gaps"] > 0 and a["identities"] == len(s)
    assert ctx.extend(orc.encode("WWWWWWWWWW"), orc.encode("AAAAAAAAAA"), (0, 0), 15) is None


@pytest.mark.parametrize("traceback", [True, False])
def test_extend_span_batch(ctx, orc, traceback):
    cases = list(span_cases(7, 400))
    qbuf = np.concatenate([c[0] for c in cases])
    sbuf = np.concatenate([c[1] for c in cases])
    jobs, qo, so = [], 0, 0
    for q, s, d0, d1 in cases:
        jobs.append((qo, len(q), so, len(s), d0, d1))
        qo += len(q)
        so += len(s)
    got = ctx.extend_span_batch(qbuf, sbuf, jobs, traceback=traceback)
    for (q, s, d0, d1), g in zip(cases, got):
        ref = orc.extend_span(q, s, d0, d1)
        if traceback:
            assert _eq_aligned(g, ref)
        else:
            assert g == (ref["score"] if ref else None)


def test_xdrop_batch(ctx, orc):
    rng = np.random.default_rng(11)
    for _ in range(20):
        q = rand_seq(rng, int(rng.integers(1, 300)), canonical_only=False)
        s = mutate(rng, q, 0.3, 0.0) if rng.random() < 0.6 else rand_seq(rng, int(rng.integers(1, 300)))
        if len(s) == 0:
            continue
        qa = rng.integers(0, len(q) + 1, size=50).astype(np.uint32)
        sa = rng.integers(0, len(s) + 1, size=50).astype(np.uint32)
        xd = int(rng.integers(1, 40))
        got = ctx.xdrop_batch(q, s, qa, sa, xd)
        for k in range(50):
            assert tuple(int(x) for x in got[k]) == orc.xdrop(q, s, int(qa[k]), int(sa[k]), xd)
    q = orc.encode("WWWWWWWWWWCCCCCCCCCC")  # src/ungapped.rs:99-129
    assert ctx.xdrop(q, orc.encode("WWWWWWWWWWDDDDDDDDDD"), 0, 0, 5)[2:] == (10, 110)
    # Long extensions (many 16-step batches, then single steps up to both
    # sequence ends), anchors at and next to the ends, and an X-drop of 0 or
    # less, where the reference's loop test fails before the first step.
    q = rand_seq(rng, 1500)
    s = np.concatenate([rand_seq(rng, 37), mutate(rng, q, 0.1, 0.0), rand_seq(rng, 21)])
    qa = np.array([0, 1, 15, 16, 17, 700, 1499, 1500, 3, 1000], np.uint32)
    sa = qa + 37
    for xd in (-3, 0, 1, 20, 60, 400):
        got = ctx.xdrop_batch(q, s, qa, sa, xd)
        for k in range(len(qa)):
            assert tuple(int(x) for x in got[k]) == orc.xdrop(q, s, int(qa[k]), int(sa[k]), xd)


@pytest.mark.parametrize("code,red,cut", [
    ("111101110111", MURPHY10, 0.8 * np.log(2) * 10),
    ("1011110111", MURPHY10, 1.0 * np.log(2) * 8),
    ("11110111", DIAMOND9, 1.0 * np.log(2) * 7),
    ("1101", MURPHY10, 0.0),
    ("1111111111111111", MURPHY10, 0.8 * np.log(2) * 16),
])
def test_seed_build(ctx, orc, code, red, cut):
    rng = np.random.default_rng(5)
    seqs = [rand_seq(rng, int(rng.integers(0, 300)), canonical_only=rng.random() < 0.7) for _ in range(40)]
    seqs.append(np.zeros(40, np.uint8))  # homopolymer: low complexity
    got = ctx.seed_build(seqs, code, red, base=17, cut=cut)
    exp = orc.seed_build(seqs, code, red, base=17, cut=cut)
    for g, e in zip(got, exp):
        assert np.array_equal(g.astype(np.int64), e.astype(np.int64))


def test_seed_build_reference_case(ctx, orc):  # src/seed.rs:175-196
    a = orc.encode("MKWVTFGCPD")
    k, s, p = ctx.seed_build([a], "1111", MURPHY10, 0, 0.8 * np.log(2) * 4)
    assert len(k) == 7


@pytest.mark.parametrize("cap", [4096, 12, 1])
def test_seed_join(ctx, orc, cap):
    rng = np.random.default_rng(9)
    qs = [rand_seq(rng, int(rng.integers(10, 200))) for _ in range(12)]
    rs = [mutate(rng, qs[i % 12], 0.25, 0.02) for i in range(30)] + [rand_seq(rng, 150) for _ in range(20)]
    rs += [np.zeros(60, np.uint8)]  # repeats push some keys over the cap
    q = orc.seed_build(qs, "110101", MURPHY10, 0, 0.0)
    r = orc.seed_build(rs, "110101", MURPHY10, 100, 0.0)
    got = ctx.seed_join(q, r, cap)
    exp = orc.seed_join(q, r, 12, cap)
    assert np.array_equal(got, exp)


def test_seed_join_cap_boundary(ctx):
    z = np.zeros
    q = (z(64, np.uint64), np.arange(64) % 3, np.arange(64))
    assert len(ctx.seed_join(q, (z(64, np.uint64), np.arange(64), np.arange(64)), 4096)) == 4096
    assert len(ctx.seed_join(q, (z(65, np.uint64), np.arange(65), np.arange(65)), 4096)) == 0


def _score_batch(ctx, cases):
    qbuf = np.concatenate([c[0] for c in cases])
    sbuf = np.concatenate([c[1] for c in cases])
    jobs, qo, so = [], 0, 0
    for q, s, d0, d1 in cases:
        jobs.append((qo, len(q), so, len(s), d0, d1))
        qo += len(q)
        so += len(s)
    return ctx.extend_span_batch(qbuf, sbuf, jobs, traceback=False)


@pytest.mark.parametrize("seed", [0, 1])
def test_int16_score_pass_band_widths(ctx, orc, seed):
    """Score-only batches run the packed int16 kernel for the standard band
    widths 25/33/61/81/129 (2 * band_for + 1); every geometry must match."""
    rng = np.random.default_rng(100 + seed)
    cases = []
    for _ in range(600):
        band = int(rng.choice([12, 16, 30, 40, 64]))
        n = int(rng.integers(1, 1200 if rng.random() < 0.2 else 350))
        q = rand_seq(rng, n, canonical_only=rng.random() < 0.9)
        r = rng.random()
        s = q.copy() if r < 0.15 else (mutate(rng, q, float(rng.uniform(0.1, 0.7)), 0.05) if r < 0.6
                                       else rand_seq(rng, int(rng.integers(1, 800))))
        if len(s) == 0:
            s = rand_seq(rng, 2)
        c = int(rng.integers(-len(s) - band, n + band))
        cases.append((q, s, c - band, c + band))
    got = _score_batch(ctx, cases)
    for (q, s, d0, d1), g in zip(cases, got):
        ref = orc.extend_span(q, s, d0, d1)
        assert g == (ref["score"] if ref else None), (len(q), len(s), d0, d1)


def test_int16_tail_rows_below_band(ctx, orc):
    """Rows past m + d1 keep one cell at column m (the reference clamps them
    there); with an in-band best below 11 that column decides the score."""
    enc = orc.encode
    cases = []
    for k in (20, 50, 99):
        q = enc("D" * k + "W" + "D" * (120 - k))
        cases.append((q, enc("AAW"), -12, 12))       # W = 25
        cases.append((q, enc("AAAAW"), -20, 20))     # W = 41: exact int32 path
        cases.append((q, enc("AW"), 5 - 30, 5 + 30))  # W = 61
    got = _score_batch(ctx, cases)
    for (q, s, d0, d1), g in zip(cases, got):
        ref = orc.extend_span(q, s, d0, d1)
        assert g == (ref["score"] if ref else None) and g == 11


def test_int16_saturation_reruns_exact(ctx, orc):
    """Self-alignments scoring above 32767 saturate int16 and are re-scored in int32."""
    rng = np.random.default_rng(3)
    hi = np.array([17, 4, 8, 18, 14, 13], np.uint8)  # W C H Y P F: self-scores 11, 9, 8, 7, 7, 6
    q = hi[rng.integers(0, len(hi), size=4500)]
    cases = [(q, q.copy(), -64, 64), (q[:300], q[:300].copy(), -64, 64)]
    got = _score_batch(ctx, cases)
    ref = [orc.extend_span(*c)["score"] for c in cases]
    assert ref[0] > 32767
    assert got == ref


def test_six_frames_match_oracle(ctx, orc):
    """dmr_six_frames (device translation + --min-orf masking) against the
    oracle's translate::six_frames + mask_short_orfs (src/translate.rs:54-137)
    on reads with both cases, U, IUPAC and other bytes, lengths 0..3000, and
    default, zero, one, small and huge min_orf values."""
    rng = np.random.default_rng(11)
    alphabet = list("ACGTACGTACGTACGTacgtuUNRYKMSWBDHVnx-*")
    reads = []
    for L in list(range(0, 12)) + [29, 30, 31, 99, 100, 101] + list(rng.integers(100, 3000, size=40)):
        p = rng.random()
        src = alphabet if p < 0.3 else list("ACGT")
        reads.append("".join(rng.choice(src, size=int(L))))
    # Reads of stops and short ORFs only, and one long ORF.
    reads.append("TAA" * 60)
    reads.append(("ATG" * 7 + "TAG") * 20)
    reads.append("ATG" * 300)
    for min_orf in [-1, 0, 1, 5, 40, 1000]:
        got = ctx.six_frames(reads, min_orf)
        for r, nt in enumerate(reads):
            exp = orc.six_frames(nt, min_orf=min_orf)
            for k in range(6):
                assert list(got[6 * r + k]) == list(exp[k]), (min_orf, r, k)


def test_wide_bands_block_kernel(ctx, orc):
    """Bands wider than one wave covers (2049 .. 16384 offsets: the
    block-per-DP kernel, K = 4, 8, 16) and past it (the row-exact kernel):
    scores of long related and unrelated pairs, bands inside, across and
    outside the matrix, against the oracle."""
    rng = np.random.default_rng(23)
    cases = []
    for w in [1100, 1500, 2100, 3000, 4500, 7000, 8300]:
        for kind in range(3):
            n = int(rng.integers(600, 2600))
            q = rand_seq(rng, n)
            s = q.copy() if kind == 0 else (mutate(rng, q, 0.3, 0.03) if kind == 1 else rand_seq(rng, int(rng.integers(400, 3000))))
            c = int(rng.integers(-len(s), n))
            cases.append((q, s, c - w, c + w + int(rng.integers(0, 3))))
    # Outside the matrix on both sides, and one very long unrelated pair.
    q, s = rand_seq(rng, 900), rand_seq(rng, 700)
    cases += [(q, s, 1000, 1000 + 3000), (q, s, -5000, -1500), (q, s, -2500, 2500)]
    qbuf = np.concatenate([c[0] for c in cases])
    sbuf = np.concatenate([c[1] for c in cases])
    jobs, qo, so = [], 0, 0
    for q, s, d0, d1 in cases:
        jobs.append((qo, len(q), so, len(s), d0, d1))
        qo += len(q)
        so += len(s)
    got = ctx.extend_span_batch(qbuf, sbuf, jobs, traceback=False)
    for (q, s, d0, d1), g in zip(cases, got):
        ref = orc.extend_span(q, s, d0, d1)
        assert g == (ref["score"] if ref else None), (len(q), len(s), d0, d1)
