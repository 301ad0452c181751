"""Pins of the oracle beyond the reference's unit tests.

- extend_span with an unbounded band reduces to textbook local Gotoh SW
  (open + L * extend gap cost): score and first-maximum cell against an
  independent O(nm) implementation.
- Alignment invariants: rescoring the ops gives the score, column counts add up.
- Band monotonicity: a wider band never scores lower.
- Complexity filter equals the integer rule multinomial(w; counts) >= 2^(c*w).
- Join equals a nested-loop equality join with the max_expand cap, including
  the cap boundary (nq*nr = 4096 kept, 4097 dropped).
- X-drop equals a prefix-sum re-derivation of the stopping rule.
- End-to-end on the reference's bench/e2e data: the hit counts, identities,
  bit scores and coordinates the reference README reports.
"""
import math

import numpy as np
import pytest

from conftest import mutate, rand_seq

MURPHY10 = "A KR EDNQ C G H ILVM FYW P ST"


def b62(orc):
    return np.array([[orc.lib().orc_matrix_score(a, b) for b in range(25)] for a in range(25)])


def gotoh_local(q, s, M, go=11, ge=1):
    """Textbook local affine alignment; returns (best, i, j) of the first
    strict maximum in row-major order."""
    n, m = len(q), len(s)
    NEG = -10 ** 9
    H = [[0] * (m + 1) for _ in range(n + 1)]
    E = [[NEG] * (m + 1) for _ in range(n + 1)]
    F = [[NEG] * (m + 1) for _ in range(n + 1)]
    best, bi, bj = 0, 0, 0
    for i in range(1, n + 1):
        for j in range(1, m + 1):
            E[i][j] = max(H[i - 1][j] - go - ge, E[i - 1][j] - ge)
            F[i][j] = max(H[i][j - 1] - go - ge, F[i][j - 1] - ge)
            H[i][j] = max(0, H[i - 1][j - 1] + M[q[i - 1]][s[j - 1]], E[i][j], F[i][j])
            if H[i][j] > best:
                best, bi, bj = H[i][j], i, j
    return best, bi, bj


def rescore(a, q, s, M, go=11, ge=1):
    i, j, sc = a["q_begin"], a["s_begin"], 0
    cols = ident = mism = gaps = 0
    for op, n in a["ops"]:
        if op == 0:
            for _ in range(n):
                sc += M[q[i]][s[j]]
                ident += int(q[i] == s[j])
                mism += int(q[i] != s[j])
                i += 1
                j += 1
        else:
            sc -= go + ge * n
            gaps += n
            if op == 1:
                i += n
            else:
                j += n
        cols += n
    return sc, i, j, cols, ident, mism, gaps


@pytest.mark.parametrize("seed", range(40))
def test_unbanded_equals_textbook_sw(orc, seed):
    rng = np.random.default_rng(seed)
    M = b62(orc)
    n, m = int(rng.integers(1, 28)), int(rng.integers(1, 28))
    q = rand_seq(rng, n, canonical_only=seed % 3 != 0)
    s = mutate(rng, q, 0.4, 0.1) if seed % 2 else rand_seq(rng, m)
    if len(s) == 0:
        s = rand_seq(rng, 3)
    a = orc.extend_span(q, s, -len(s) - 1, len(q) + 1)
    best, bi, bj = gotoh_local(list(q), list(s), M)
    if best <= 0:
        assert a is None
        return
    assert a["score"] == best
    assert (a["q_end"], a["s_end"]) == (bi, bj)
    sc, ie, je, cols, ident, mism, gaps = rescore(a, q, s, M)
    assert sc == best and (ie, je) == (a["q_end"], a["s_end"])
    assert cols == a["length"] and ident == a["identities"] and mism == a["mismatches"]
    assert gaps == a["gaps"] and ident + mism + gaps == a["length"]


@pytest.mark.parametrize("seed", range(30))
def test_banded_invariants(orc, seed):
    rng = np.random.default_rng(100 + seed)
    M = b62(orc)
    q = rand_seq(rng, int(rng.integers(5, 120)))
    s = mutate(rng, q, 0.35, 0.05)
    if len(s) < 2:
        return
    d = int(rng.integers(-len(s) + 1, len(q)))
    prev = 0
    for band in (0, 2, 5, 12, 30, 200):
        a = orc.extend_span(q, s, d - band, d + band)
        sc = a["score"] if a else 0
        assert sc >= prev
        prev = sc
        if a:
            r = rescore(a, q, s, M)
            assert r[0] == a["score"] and r[3] == a["length"]
            assert (r[1], r[2]) == (a["q_end"], a["s_end"])


def test_complexity_equals_integer_rule(orc):
    rng = np.random.default_rng(7)
    for w, code in [(10, "111101110111"), (8, "1011110111"), (4, "1111")]:
        cut = 0.8 * math.log(2) * w
        red = orc.lib()
        for _ in range(400):
            # bias towards low complexity so both outcomes occur
            alph = rng.integers(1, 6)
            seq = rng.choice(rng.integers(0, 20, size=alph), size=len(code)).astype(np.uint8)
            m = np.zeros(25, np.uint8)
            orc.lib().orc_reduction_map(b"default", m.ctypes.data)
            counts = np.bincount([m[seq[i]] for i, c in enumerate(code) if c == "1"], minlength=10)
            multinom = math.factorial(w) // math.prod(math.factorial(int(c)) for c in counts)
            expect = multinom >= 2 ** (0.8 * w)
            assert orc.seed_is_complex(seq, code, MURPHY10, cut) == expect


def nested_join(q, r, cap):
    out = []
    qk, qs, qp = q
    rk, rs, rp = r
    keys = sorted(set(qk.tolist()) & set(rk.tolist()))
    for k in keys:
        qi = np.nonzero(qk == k)[0]
        ri = np.nonzero(rk == k)[0]
        if len(qi) * len(ri) > cap:
            continue
        for a in qi:
            for b in ri:
                out.append((int(qs[a]), int(qp[a]), int(rs[b]), int(rp[b])))
    out.sort(key=lambda h: (h[0], h[2], h[1] - h[3], h[3]))
    return np.array(out, dtype=np.int64).reshape(-1, 4)


@pytest.mark.parametrize("seed", range(6))
def test_join_equals_nested_loop(orc, seed):
    rng = np.random.default_rng(seed)
    qs = [rand_seq(rng, int(rng.integers(5, 60))) for _ in range(5)]
    rs = [mutate(rng, qs[i % 5], 0.2, 0.02) for i in range(8)] + [rand_seq(rng, 50) for _ in range(4)]
    q = orc.seed_build(qs, "1101", MURPHY10, 0, 0.0)
    r = orc.seed_build(rs, "1101", MURPHY10, 0, 0.0)
    for cap in (4096, 6, 1):
        got = orc.seed_join(q, r, 5, cap)
        assert np.array_equal(got, nested_join(q, r, cap))


def test_join_cap_boundary(orc):
    # 64 x 64 = 4096 hits are kept, 64 x 65 = 4160 > 4096 are dropped.
    z = np.zeros
    q = (z(64, np.uint64), np.arange(64, dtype=np.uint32) % 3, np.arange(64, dtype=np.uint32))
    r64 = (z(64, np.uint64), np.arange(64, dtype=np.uint32), np.arange(64, dtype=np.uint32))
    r65 = (z(65, np.uint64), np.arange(65, dtype=np.uint32), np.arange(65, dtype=np.uint32))
    assert len(orc.seed_join(q, r64, 3, 4096)) == 4096
    assert len(orc.seed_join(q, r65, 3, 4096)) == 0


def xdrop_prefix(q, s, qa, sa, xd, M):
    def side(steps):
        best, run, arg, k = 0, 0, 0, 0
        for k, v in enumerate(steps, start=1):
            if best - run >= xd:
                break
            run += v
            if run > best:
                best, arg = run, k
        return best, arg
    left = [M[q[qa - 1 - t]][s[sa - 1 - t]] for t in range(min(qa, sa))]
    lb, delta = side(left)
    right = [M[q[qa + t]][s[sa + t]] for t in range(min(len(q) - qa, len(s) - sa))]
    # the right side continues from the left-hand best
    best, run, ln = lb, lb, 0
    for k, v in enumerate(right, start=1):
        if best - run >= xd:
            break
        run += v
        if run > best:
            best, ln = run, k
    return qa - delta, sa - delta, ln + delta, best


@pytest.mark.parametrize("seed", range(20))
def test_xdrop_prefix_rederivation(orc, seed):
    rng = np.random.default_rng(seed)
    M = b62(orc)
    q = rand_seq(rng, int(rng.integers(2, 80)), canonical_only=False)
    s = mutate(rng, q, 0.3, 0.0)
    if len(s) < 2:
        return
    for _ in range(10):
        qa, sa = int(rng.integers(0, len(q) + 1)), int(rng.integers(0, len(s) + 1))
        xd = int(rng.integers(1, 40))
        assert orc.xdrop(q, s, qa, sa, xd) == xdrop_prefix(list(q), list(s), qa, sa, xd, M)


def _rows(text):
    return [l.split("\t") for l in text.strip().split("\n") if l]


def test_e2e_reference_readme_facts(orc, e2e):
    """README.md:136-179 of the reference: hit counts per preset on bench/e2e,
    WP_048121132.1 at 26.2% / 83.2 bits under --default, and
    WP_048121137.1 as a 612-column, 91.7-bit alignment of query 4-478 under
    --sensitive."""
    args = (e2e["db_letters"], e2e["db_offsets"], e2e["q_letters"], e2e["q_offsets"])
    fmt = lambda r: _rows(r.format(e2e["q_ids"], e2e["db_ids"]))
    d = fmt(orc.search(*args, sensitivity="default", threads=8))
    assert len(d) == 4
    r = [x for x in d if x[1] == "WP_048121132.1"][0]
    assert r[2] == "26.2" and r[11] == "83.2"
    assert len(fmt(orc.search(*args, sensitivity="fast", threads=8))) == 1
    s = fmt(orc.search(*args, sensitivity="sensitive", threads=8))
    assert len(s) == 7
    r = [x for x in s if x[1] == "WP_048121137.1"][0]
    assert r[3] == "612" and r[11] == "91.7" and (r[6], r[7]) == ("4", "478")
    # The three ALP E-values the reference checks come from these HSPs.
    ev = {x[1]: x[10] for x in d}
    assert ev["WP_014867732.1"] == "3.21e-105" and ev["WP_048156290.1"] == "3.89e-25"
    assert ev["WP_048121137.1"] == "7.60e-17"
