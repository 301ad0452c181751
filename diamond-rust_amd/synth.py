"""Deterministic synthetic protein workloads for tests and the benchmark.

Residues are drawn i.i.d. from the BLOSUM62 background frequencies
(reference src/matrix.rs:16-19, NCBI letter order) and encoded as uint8
codes 0..19. Sequence lengths are log-normal and clipped. Homologs are
planted by substituting residues from the BLOSUM62 conditional distribution
P(b|a) = p_b * exp(lambda_u * s_ab) (lambda_u solves
sum_ab p_a p_b exp(lambda_u s_ab) = 1, the method of src/matrix.rs:245-272),
mixed with identity-keeping or background draws to reach a target identity,
plus insertions/deletions at a per-site rate with geometric lengths.

Everything is a pure function of the workload name and seed, so the CPU
reference and the GPU path always see identical bytes.
"""
from __future__ import annotations

import numpy as np

BACKGROUND = np.array(
    [0.074, 0.052, 0.045, 0.054, 0.025, 0.034, 0.054, 0.074, 0.026, 0.068,
     0.099, 0.058, 0.025, 0.047, 0.039, 0.057, 0.051, 0.013, 0.032, 0.073])
BACKGROUND = BACKGROUND / BACKGROUND.sum()

AA = "ARNDCQEGHILKMFPSTWYVBJZX*"

# BLOSUM62 restricted to the 20 canonical letters (NCBI order).
_B62 = np.array([
    [4, -1, -2, -2, 0, -1, -1, 0, -2, -1, -1, -1, -1, -2, -1, 1, 0, -3, -2, 0],
    [-1, 5, 0, -2, -3, 1, 0, -2, 0, -3, -2, 2, -1, -3, -2, -1, -1, -3, -2, -3],
    [-2, 0, 6, 1, -3, 0, 0, 0, 1, -3, -3, 0, -2, -3, -2, 1, 0, -4, -2, -3],
    [-2, -2, 1, 6, -3, 0, 2, -1, -1, -3, -4, -1, -3, -3, -1, 0, -1, -4, -3, -3],
    [0, -3, -3, -3, 9, -3, -4, -3, -3, -1, -1, -3, -1, -2, -3, -1, -1, -2, -2, -1],
    [-1, 1, 0, 0, -3, 5, 2, -2, 0, -3, -2, 1, 0, -3, -1, 0, -1, -2, -1, -2],
    [-1, 0, 0, 2, -4, 2, 5, -2, 0, -3, -3, 1, -2, -3, -1, 0, -1, -3, -2, -2],
    [0, -2, 0, -1, -3, -2, -2, 6, -2, -4, -4, -2, -3, -3, -2, 0, -2, -2, -3, -3],
    [-2, 0, 1, -1, -3, 0, 0, -2, 8, -3, -3, -1, -2, -1, -2, -1, -2, -2, 2, -3],
    [-1, -3, -3, -3, -1, -3, -3, -4, -3, 4, 2, -3, 1, 0, -3, -2, -1, -3, -1, 3],
    [-1, -2, -3, -4, -1, -2, -3, -4, -3, 2, 4, -2, 2, 0, -3, -2, -1, -2, -1, 1],
    [-1, 2, 0, -1, -3, 1, 1, -2, -1, -3, -2, 5, -1, -3, -1, 0, -1, -3, -2, -2],
    [-1, -1, -2, -3, -1, 0, -2, -3, -2, 1, 2, -1, 5, 0, -2, -1, -1, -1, -1, 1],
    [-2, -3, -3, -3, -2, -3, -3, -3, -1, 0, 0, -3, 0, 6, -4, -2, -2, 1, 3, -1],
    [-1, -2, -2, -1, -3, -1, -1, -2, -2, -3, -3, -1, -2, -4, 7, -1, -1, -4, -3, -2],
    [1, -1, 1, 0, -1, 0, 0, 0, -1, -2, -2, 0, -1, -2, -1, 4, 1, -3, -2, -2],
    [0, -1, 0, -1, -1, -1, -1, -2, -2, -1, -1, -1, -1, -2, -1, 1, 5, -2, -2, 0],
    [-3, -3, -4, -4, -2, -2, -3, -2, -2, -3, -2, -3, -1, 1, -4, -3, -2, 11, 2, -3],
    [-2, -2, -2, -3, -2, -1, -2, -3, 2, -1, -1, -2, -1, 3, -3, -2, -2, 2, 7, -1],
    [0, -3, -3, -3, -1, -2, -2, -3, -3, 3, 1, -2, 1, -1, -2, -2, 0, -3, -1, 4],
], dtype=np.float64)


def solve_lambda_u() -> float:
    """Ungapped lambda of BLOSUM62 under the background (bisection)."""
    pp = np.outer(BACKGROUND, BACKGROUND)

    def f(lam):
        return float((pp * np.exp(lam * _B62)).sum() - 1.0)

    lo, hi = 1e-6, 1.0
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if f(mid) > 0:
            hi = mid
        else:
            lo = mid
    return 0.5 * (lo + hi)


LAMBDA_U = solve_lambda_u()
_COND = BACKGROUND[None, :] * np.exp(LAMBDA_U * _B62)
_COND = _COND / _COND.sum(axis=1, keepdims=True)
_COND_CUM = np.cumsum(_COND, axis=1)
COND_IDENTITY = float((BACKGROUND * np.diag(_COND)).sum())
BG_IDENTITY = float((BACKGROUND ** 2).sum())
_BG_CUM = np.cumsum(BACKGROUND)


def _residues(rng, n):
    u = rng.random(n)
    return np.minimum(np.searchsorted(_BG_CUM, u, side="right"), 19).astype(np.uint8)


def _lengths(rng, n, median, sigma, lo, hi):
    x = np.exp(np.log(median) + sigma * rng.standard_normal(n))
    return np.clip(np.rint(x), lo, hi).astype(np.int64)


def mutate(rng, seq: np.ndarray, identity: float, indel_rate: float) -> np.ndarray:
    """A homolog of `seq` at roughly `identity` with geometric indels (p=0.5)."""
    n = len(seq)
    if identity >= COND_IDENTITY:
        keep = (identity - COND_IDENTITY) / (1.0 - COND_IDENTITY)
        bg = 0.0
    else:
        keep = 0.0
        bg = (COND_IDENTITY - identity) / (COND_IDENTITY - BG_IDENTITY)
    u = rng.random(n)
    v = rng.random(n)
    cond = np.minimum((_COND_CUM[seq] < v[:, None]).sum(axis=1), 19).astype(np.uint8)
    bgr = _residues(rng, n)
    out = np.where(u < keep, seq, np.where(rng.random(n) < bg, bgr, cond)).astype(np.uint8)
    # Indels: at each site, with probability indel_rate, insert or delete a
    # geometric(0.5) number of residues.
    ev = np.nonzero(rng.random(n) < indel_rate)[0]
    if len(ev) == 0:
        return out
    pieces = []
    last = 0
    for p in ev:
        if p < last:
            continue
        L = int(rng.geometric(0.5))
        pieces.append(out[last:p])
        if rng.random() < 0.5:
            pieces.append(_residues(rng, L))
            last = p
        else:
            last = min(n, p + L)
    pieces.append(out[last:])
    return np.concatenate(pieces).astype(np.uint8)


def _pack(seqs):
    lens = np.array([len(s) for s in seqs], dtype=np.uint64)
    off = np.zeros(len(seqs) + 1, dtype=np.uint64)
    np.cumsum(lens, out=off[1:])
    letters = np.concatenate(seqs).astype(np.uint8) if seqs else np.zeros(0, np.uint8)
    return letters, off


class Workload:
    """Packed DB + queries: letters (uint8 codes), offsets (uint64, n+1), ids."""

    def __init__(self, name, db_seqs, db_ids, q_seqs, q_ids, meta):
        self.name = name
        self.db_letters, self.db_offsets = _pack(db_seqs)
        self.q_letters, self.q_offsets = _pack(q_seqs)
        self.db_ids = db_ids
        self.q_ids = q_ids
        self.meta = meta

    @property
    def n_db(self):
        return len(self.db_offsets) - 1

    @property
    def n_q(self):
        return len(self.q_offsets) - 1

    def subset_queries(self, idx):
        qs = [self.q_letters[self.q_offsets[i]:self.q_offsets[i + 1]] for i in idx]
        w = Workload.__new__(Workload)
        w.name = self.name + f"[{len(idx)}q]"
        w.db_letters, w.db_offsets = self.db_letters, self.db_offsets
        w.q_letters, w.q_offsets = _pack(qs)
        w.db_ids = self.db_ids
        w.q_ids = [self.q_ids[i] for i in idx]
        w.meta = dict(self.meta, queries=len(idx))
        return w


def protein_db_2p26(seed=20240601, n_queries=1024):
    """The 7,860-protein (~2.26 Maa) DB shared by the --default (1x497 aa)
    and --sensitive (1024 queries) workloads, with the planted homologs of
    both query sets."""
    rng = np.random.default_rng(seed)
    n_db = 7860
    q1 = _residues(rng, 497)
    qlens = _lengths(rng, n_queries, 300, 0.57, 50, 1000)
    qs = [_residues(rng, int(L)) for L in qlens]
    planted = []  # (query set, query index, homolog)
    for _ in range(7):
        planted.append((1, 0, mutate(rng, q1, rng.uniform(0.30, 0.60), 0.02)))
    for qi in range(n_queries):
        for _ in range(int(rng.integers(0, 11))):
            planted.append((2, qi, mutate(rng, qs[qi], rng.uniform(0.25, 0.60), 0.02)))
    n_bg = max(0, n_db - len(planted))
    # Background lengths are scaled so the DB totals ~2,264,000 letters.
    planted_letters = sum(len(p[2]) for p in planted)
    bg_median = max(30.0, (2_264_000 - planted_letters) / max(n_bg, 1) / np.exp(0.57 ** 2 / 2))
    bg_lens = _lengths(rng, n_bg, bg_median, 0.57, 30, 5000)
    db = [_residues(rng, int(L)) for L in bg_lens] + [p[2] for p in planted]
    tags = ["bg"] * n_bg + [f"h{p[0]}q{p[1]}" for p in planted]
    perm = rng.permutation(len(db))
    db = [db[i] for i in perm]
    ids = [f"s{k:06d}_{tags[i]}" for k, i in enumerate(perm)]
    return db, ids, q1, qs


def protein_db_60m(seed=20240601, n_queries=4096, n_db=200_000):
    """The 200,000-protein (~60 M letters) DB of the --very-sensitive and
    blastx workloads: background lengths log-normal (median 250, sigma 0.57,
    clipped 30..5000), 4096 queries (median 300, clipped 50..1000), each with
    0-10 remote homologs planted at 20-35% identity with indels at 0.04 per
    site."""
    rng = np.random.default_rng(seed + 60)
    qlens = _lengths(rng, n_queries, 300, 0.57, 50, 1000)
    q_all = _residues(rng, int(qlens.sum()))
    q_off = np.concatenate([[0], np.cumsum(qlens)])
    qs = [q_all[q_off[i]:q_off[i + 1]] for i in range(n_queries)]
    planted = []
    for qi in range(n_queries):
        for _ in range(int(rng.integers(0, 11))):
            planted.append((qi, mutate(rng, qs[qi], rng.uniform(0.20, 0.35), 0.04)))
    n_bg = n_db - len(planted)
    bg_lens = _lengths(rng, n_bg, 250, 0.57, 30, 5000)
    bg_all = _residues(rng, int(bg_lens.sum()))
    bg_off = np.concatenate([[0], np.cumsum(bg_lens)])
    db = [bg_all[bg_off[i]:bg_off[i + 1]] for i in range(n_bg)] + [p[1] for p in planted]
    tags = ["bg"] * n_bg + [f"hq{p[0]}" for p in planted]
    perm = rng.permutation(len(db))
    db = [db[i] for i in perm]
    ids = [f"v{k:06d}_{tags[i]}" for k, i in enumerate(perm)]
    return db, ids, qs


# 16-bit lookup of the background distribution (for the largest DB, where
# drawing doubles per letter costs seconds).
_LUT16 = np.minimum(np.searchsorted(_BG_CUM, (np.arange(65536) + 0.5) / 65536.0, side="right"),
                    19).astype(np.uint8)


def protein_db_serving(seed=20240601, n_db=1_000_000, n_pool=4096):
    """The serving workload: 1,000,000 proteins (~300 M letters; lengths
    log-normal, median 255, sigma 0.57, clipped 30..5000, letters from the
    background through a 16-bit table) and a pool of 4096 queries with
    heavy-tailed lengths (log-normal median 350, sigma 0.8, clipped
    30..5000), each with 0-5 homologs planted at 25-60% identity (indels
    0.02). Returns (db letters, db offsets, db ids, query list)."""
    rng = np.random.default_rng(seed + 1000)
    qlens = _lengths(rng, n_pool, 350, 0.8, 30, 5000)
    qs = [_residues(rng, int(L)) for L in qlens]
    planted = []
    for qi in range(n_pool):
        for _ in range(int(rng.integers(0, 6))):
            planted.append(mutate(rng, qs[qi], rng.uniform(0.25, 0.60), 0.02))
    n_bg = n_db - len(planted)
    lens = np.concatenate([_lengths(rng, n_bg, 255, 0.57, 30, 5000),
                           np.array([len(p) for p in planted], np.int64)])
    perm = rng.permutation(n_db)  # DB slot -> source (background first, then planted)
    slot_lens = lens[perm]
    off = np.zeros(n_db + 1, np.uint64)
    np.cumsum(slot_lens, out=off[1:])
    letters = _LUT16[rng.integers(0, 65536, size=int(off[-1]), dtype=np.uint16)]
    for k in np.nonzero(perm >= n_bg)[0]:
        letters[int(off[k]):int(off[k + 1])] = planted[int(perm[k]) - n_bg]
    ids = [f"u{k:07d}" for k in range(n_db)]
    return letters, off, ids, qs


class ServingWorkload:
    """A resident DB and a query pool for the serving loop."""

    def __init__(self, name, db_letters, db_offsets, db_ids, pool, meta):
        self.name = name
        self.db_letters, self.db_offsets, self.db_ids = db_letters, db_offsets, db_ids
        self.pool = pool
        self.meta = meta

    @property
    def n_db(self):
        return len(self.db_offsets) - 1

    def stream(self, n, qps, seed=7):
        """n requests of a Poisson arrival stream at qps: (arrival s, pool
        index, preset), presets 50% default, 40% sensitive, 10%
        very-sensitive."""
        rng = np.random.default_rng(seed)
        t = np.cumsum(rng.exponential(1.0 / qps, size=n))
        q = rng.integers(0, len(self.pool), size=n)
        u = rng.random(n)
        presets = np.where(u < 0.5, "default", np.where(u < 0.9, "sensitive", "very-sensitive"))
        return list(zip(t.tolist(), q.tolist(), presets.tolist()))


def make_serving(name="serve-1m", seed=20240601) -> ServingWorkload:
    """'serve-1m': 1,000,000 proteins (~300 M letters), 4096 pooled queries;
    'serve-tiny': the same recipe at 20,000 proteins (tests)."""
    n_db = {"serve-1m": 1_000_000, "serve-tiny": 20_000}[name]
    letters, off, ids, pool = protein_db_serving(seed, n_db=n_db, n_pool=4096 if n_db > 100_000 else 256)
    return ServingWorkload(name, letters, off, ids, pool, dict(seed=seed, proteins=n_db))


# Standard genetic code, amino acid -> synonymous codons (reverse translation
# of planted proteins into reads).
_CODONS = {
    "A": ["GCT", "GCC", "GCA", "GCG"], "R": ["CGT", "CGC", "CGA", "CGG", "AGA", "AGG"],
    "N": ["AAT", "AAC"], "D": ["GAT", "GAC"], "C": ["TGT", "TGC"], "Q": ["CAA", "CAG"],
    "E": ["GAA", "GAG"], "G": ["GGT", "GGC", "GGA", "GGG"], "H": ["CAT", "CAC"],
    "I": ["ATT", "ATC", "ATA"], "L": ["TTA", "TTG", "CTT", "CTC", "CTA", "CTG"],
    "K": ["AAA", "AAG"], "M": ["ATG"], "F": ["TTT", "TTC"], "P": ["CCT", "CCC", "CCA", "CCG"],
    "S": ["TCT", "TCC", "TCA", "TCG", "AGT", "AGC"], "T": ["ACT", "ACC", "ACA", "ACG"],
    "W": ["TGG"], "Y": ["TAT", "TAC"], "V": ["GTT", "GTC", "GTA", "GTG"],
}
_COMP = str.maketrans("ACGT", "TGCA")

# Approximate E. coli K-12 codon usage (per thousand codons), the weights the
# reads' reverse translation draws synonymous codons with.
_ECOLI_USAGE = {
    "GCG": 33.7, "GCC": 25.8, "GCA": 20.1, "GCT": 15.3, "CGT": 20.9, "CGC": 22.0, "CGA": 3.6,
    "CGG": 5.4, "AGA": 2.1, "AGG": 1.2, "AAC": 21.7, "AAT": 17.7, "GAT": 32.1, "GAC": 19.1,
    "TGC": 6.5, "TGT": 5.2, "CAG": 28.8, "CAA": 15.3, "GAA": 39.4, "GAG": 17.8, "GGC": 29.6,
    "GGT": 24.7, "GGG": 11.1, "GGA": 8.0, "CAT": 12.9, "CAC": 9.7, "ATT": 30.3, "ATC": 25.1,
    "ATA": 4.4, "CTG": 52.6, "TTA": 13.9, "TTG": 13.7, "CTT": 11.0, "CTC": 11.0, "CTA": 3.9,
    "AAA": 33.6, "AAG": 10.3, "ATG": 27.9, "TTT": 22.3, "TTC": 16.6, "CCG": 23.2, "CCA": 8.4,
    "CCT": 7.0, "CCC": 5.5, "AGC": 16.1, "TCT": 8.5, "TCC": 8.6, "AGT": 8.8, "TCG": 8.9,
    "TCA": 7.2, "ACC": 23.4, "ACG": 14.4, "ACT": 8.9, "ACA": 7.1, "TGG": 15.2, "TAT": 16.2,
    "TAC": 12.2, "GTG": 26.4, "GTT": 18.3, "GTC": 15.3, "GTA": 10.9,
}
# Per letter code 0..19: its codons (padded to 6) and cumulative usage.
_CODON_BYTES = np.zeros((20, 6, 3), np.uint8)
_CODON_CUM = np.ones((20, 6))
for _a in range(20):
    _cs = _CODONS[AA[_a]]
    _w = np.array([_ECOLI_USAGE[c] for c in _cs])
    _CODON_CUM[_a, :len(_cs)] = np.cumsum(_w) / _w.sum()
    for _k, _c in enumerate(_cs):
        _CODON_BYTES[_a, _k] = np.frombuffer(_c.encode(), np.uint8)


def reverse_translate(rng, prot):
    """DNA of a protein (codes 0..19) with E. coli-like synonymous codons."""
    u = rng.random(len(prot))
    k = np.minimum((_CODON_CUM[prot] < u[:, None]).sum(axis=1), 5)
    return _CODON_BYTES[prot, k].reshape(-1).tobytes().decode()


def sequencing_errors(rng, seq, sub=0.01, indel=0.001):
    """1% base substitutions, 0.1% single-base insertions or deletions."""
    b = np.frombuffer(seq.encode(), np.uint8).copy()
    m = rng.random(len(b)) < sub
    b[m] = np.frombuffer(b"ACGT", np.uint8)[rng.integers(0, 4, size=int(m.sum()))]
    ev = np.nonzero(rng.random(len(b)) < indel)[0]
    if len(ev) == 0:
        return b.tobytes().decode()
    out, last = [], 0
    for p in ev:
        out.append(b[last:p].tobytes())
        if rng.random() < 0.5:
            out.append(b"ACGT"[int(rng.integers(0, 4))].to_bytes(1, "little"))
            last = p
        else:
            last = p + 1
    out.append(b[last:].tobytes())
    return b"".join(out).decode()


def _planted_reads(rng, db_seqs, n_reads, lo_nt, hi_nt):
    """Reads of lo_nt..hi_nt bases (uniform): a segment of a random DB
    protein reverse-translated with E. coli codon usage, with sequencing
    errors, reverse-complemented half of the time."""
    reads = []
    for _ in range(n_reads):
        n_nt = int(rng.integers(lo_nt, hi_nt + 1))
        prot = db_seqs[int(rng.integers(0, len(db_seqs)))]
        n_aa = min(len(prot), max(1, n_nt // 3))
        a = int(rng.integers(0, len(prot) - n_aa + 1))
        seq = sequencing_errors(rng, reverse_translate(rng, np.asarray(prot[a:a + n_aa], np.int64)))
        if rng.random() < 0.5:
            seq = seq.translate(_COMP)[::-1]
        reads.append(seq)
    return reads


class ReadWorkload:
    """Nucleotide reads (ASCII) against a protein DB, for blastx."""

    def __init__(self, name, db: Workload, reads, read_ids, meta):
        self.name = name
        self.db_letters, self.db_offsets, self.db_ids = db.db_letters, db.db_offsets, db.db_ids
        self.n_db = db.n_db
        self.reads = reads
        self.read_ids = read_ids
        self.meta = meta


def _reads(rng, db_seqs, n_reads, lo_nt, hi_nt, planted_frac):
    """Reads of lo_nt..hi_nt bases: reverse-translated segments of DB
    proteins (random synonymous codons, 1% base substitutions, 0.1% single-base
    indels, either strand) or random bases."""
    reads = []
    for _ in range(n_reads):
        n_nt = int(rng.integers(lo_nt, hi_nt + 1))
        if rng.random() < planted_frac:
            prot = db_seqs[int(rng.integers(0, len(db_seqs)))]
            n_aa = min(len(prot), max(1, n_nt // 3))
            a = int(rng.integers(0, len(prot) - n_aa + 1))
            seq = "".join(rng.choice(_CODONS[AA[c]]) for c in prot[a:a + n_aa])
            out = []
            for ch in seq:
                r = rng.random()
                if r < 0.0005:
                    continue  # deletion
                if r < 0.001:
                    out.append("ACGT"[int(rng.integers(0, 4))])  # insertion
                out.append("ACGT"[int(rng.integers(0, 4))] if rng.random() < 0.01 else ch)
            seq = "".join(out)
            if rng.random() < 0.5:
                seq = seq.translate(_COMP)[::-1]
        else:
            seq = "".join(rng.choice(list("ACGT"), size=n_nt))
        reads.append(seq)
    return reads


def make_reads(name: str, seed: int = 20240601) -> ReadWorkload:
    """Named blastx workloads:
    'blastx-tiny'  : 40 reads of 60..900 bases vs the 'tiny' 400-protein DB
    'blastx-2048'  : 2048 reads of 150..3000 bases (uniform) vs the 60 M letter
                     DB: reverse-translated DB proteins (E. coli codon usage),
                     1% substitutions, 0.1% indels, either strand
    """
    rng = np.random.default_rng(seed + 7)
    if name == "blastx-tiny":
        db = make("tiny", seed)
        db_seqs = [db.db_letters[db.db_offsets[i]:db.db_offsets[i + 1]] for i in range(db.n_db)]
        reads = _reads(rng, db_seqs, 40, 60, 900, 0.7)
    elif name == "blastx-2048":
        db = make("blastp-very-sensitive-4096", seed)
        db_seqs = [db.db_letters[db.db_offsets[i]:db.db_offsets[i + 1]] for i in range(db.n_db)]
        reads = _planted_reads(rng, db_seqs, 2048, 150, 3000)
    else:
        raise ValueError(f"unknown read workload {name}")
    return ReadWorkload(name, db, reads, [f"r{i:05d}" for i in range(len(reads))],
                        dict(seed=seed, reads=len(reads)))


def make(name: str, seed: int = 20240601) -> Workload:
    """Named workloads:
    'blastp-default-1x497'   : 1 query of 497 aa vs the 2.26 Maa DB
    'blastp-sensitive-1024'  : 1024 queries (median 300, 50..1000) vs the same DB
    'blastp-very-sensitive-4096': 4096 queries vs 200,000 proteins (~60 M letters),
                               remote homologs (20-35% identity, indels 0.04)
    'tiny'                   : 24 queries vs 400 proteins (fast parity cases)
    """
    if name in ("blastp-default-1x497", "blastp-sensitive-1024"):
        db, ids, q1, qs = protein_db_2p26(seed)
        if name == "blastp-default-1x497":
            return Workload(name, db, ids, [q1], ["query1"],
                            dict(sensitivity="default", seed=seed, queries=1))
        return Workload(name, db, ids, qs, [f"q{i:04d}" for i in range(len(qs))],
                        dict(sensitivity="sensitive", seed=seed, queries=len(qs)))
    if name == "blastp-very-sensitive-4096":
        db, ids, qs = protein_db_60m(seed)
        return Workload(name, db, ids, qs, [f"v{i:04d}" for i in range(len(qs))],
                        dict(sensitivity="very-sensitive", seed=seed, queries=len(qs)))
    if name.startswith("tiny"):
        rng = np.random.default_rng(seed)
        n_q = 24
        qlens = _lengths(rng, n_q, 200, 0.6, 20, 900)
        qs = [_residues(rng, int(L)) for L in qlens]
        planted = []
        for qi in range(n_q):
            for _ in range(int(rng.integers(0, 4))):
                planted.append(mutate(rng, qs[qi], rng.uniform(0.25, 0.8), 0.03))
        bg = [_residues(rng, int(L)) for L in _lengths(rng, 400 - len(planted), 200, 0.6, 10, 2000)]
        db = bg + planted
        perm = rng.permutation(len(db))
        db = [db[i] for i in perm]
        return Workload(name, db, [f"t{k:04d}" for k in range(len(db))], qs,
                        [f"tq{i:02d}" for i in range(n_q)], dict(seed=seed, queries=n_q))
    raise ValueError(f"unknown workload {name}")
