"""dmr_extend_span_noalloc (no device allocation, a static per-GPU array)
against dmr_extend_span, the entry point it stands beside: the same found
flag, score, coordinates, statistics and ops, on two calls in a row (the
second smaller, so it runs over what the first left in the array), then a
few more shapes, and the size limit.
"""
import numpy as np
import pytest

from conftest import mutate, rand_seq

pytestmark = pytest.mark.gpu


@pytest.fixture
def fresh(dmr):
    c = dmr.Context(0)
    yield c
    c.close()


def same(a, b):
    if a is None or b is None:
        return a is None and b is None
    return all(np.array_equal(np.asarray(a[k]), np.asarray(b[k])) for k in a) and a.keys() == b.keys()


def test_noalloc_matches_extend_span(fresh):
    rng = np.random.default_rng(2024)
    big_q = rand_seq(rng, 900)
    big_s = mutate(rng, big_q, sub=0.3, indel=0.04)
    small_q = rand_seq(rng, 140)
    small_s = mutate(rng, small_q, sub=0.4, indel=0.03)
    # Two calls in a row, the second on a smaller pair.
    for q, s, d0, d1 in [(big_q, big_s, -300, 300), (small_q, small_s, -20, 20)]:
        exp = fresh.extend_span(q, s, d0, d1)
        got = fresh.extend_span_noalloc(q, s, d0, d1)
        assert exp is not None and same(got, exp)
    # Bands across, inside and outside the matrix, unrelated pairs, gap costs.
    cases = [(small_q, small_s, 5, 400), (small_q, small_s, -400, -50), (big_q[:60], rand_seq(rng, 75), -8, 8),
             (rand_seq(rng, 1), rand_seq(rng, 1), 0, 0), (small_q, small_q, -1000, 1000)]
    for q, s, d0, d1 in cases:
        assert same(fresh.extend_span_noalloc(q, s, d0, d1), fresh.extend_span(q, s, d0, d1))
    q, s = big_q[:300], big_s[:320]
    assert same(fresh.extend_span_noalloc(q, s, -50, 50, 10, 1), fresh.extend_span(q, s, -50, 50, 10, 1))


def test_noalloc_size_limit(dmr, fresh):
    """A DP past the 32 MiB array is refused with DMR_E_UNSUPPORTED before any
    device work; the next call works."""
    rng = np.random.default_rng(7)
    q = rand_seq(rng, 20000)
    with pytest.raises(dmr.DmrError) as ei:
        fresh.extend_span_noalloc(q, q, -2000, 2000)
    assert ei.value.code == dmr.DMR_E_UNSUPPORTED and "bytes of device scratch" in str(ei.value)
    s = q[:500]
    assert same(fresh.extend_span_noalloc(q[:400], s, -64, 64), fresh.extend_span(q[:400], s, -64, 64))
