"""Writes the oracle's outfmt-6 text of whole synthetic workloads as fixtures.

Usage: python tests/golden/make_search_golden.py NAME [NAME ...]
  blastp-sensitive-1024   1024 queries vs the 2.26 Maa DB, --sensitive (one
                          batch: the max_expand cap counts the whole batch)
  blastp-very-sensitive-4096-q64
                          queries 0, 64, 128, ... (64) of the 4096-query
                          workload vs its 60 M letter DB, --very-sensitive,
                          as one batch of those 64
  blastx-2048-r64         reads 0, 32, 64, ... (64) of the 2048-read blastx
                          workload vs the same 60 M letter DB, --default, as
                          one batch; also the frame and strand fields

Each fixture is tests/golden/search/<NAME>.tsv.gz: the 12 default fields,
the raw score and the CIGAR, one line per HSP, in output order. The
workloads come from the seeded generator diamond-rust_amd/synth.py, so the
GPU tests and the bench regenerate the same bytes and compare with these
files instead of running the oracle at full size on the GPU machine.
"""
import gzip
import os
import pathlib
import sys
import time

ROOT = pathlib.Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT / "tests"))
sys.path.insert(0, str(ROOT / "oracle"))

from conftest import load_synth  # noqa: E402
import pyoracle  # noqa: E402

FIELDS = ("qseqid sseqid pident length mismatch gapopen qstart qend sstart send evalue bitscore "
          "score cigar")


SUBSETS = {"blastp-very-sensitive-4096-q64": ("blastp-very-sensitive-4096", range(0, 4096, 64))}
READ_SUBSETS = {"blastx-2048-r64": ("blastx-2048", range(0, 2048, 32), "default")}
READ_FIELDS = FIELDS + " qframe qstrand qlen"


def read_workload(name):
    """(read workload, read indexes, sensitivity) of a blastx fixture name."""
    synth = load_synth()
    base, idx, sens = READ_SUBSETS[name]
    return synth.make_reads(base), list(idx), sens


def workload(name):
    """(workload, sensitivity) of a fixture name; a subset fixture searches
    the listed queries of its workload as one batch."""
    synth = load_synth()
    if name == "blastp-sensitive-1024":
        return synth.make(name), "sensitive"
    if name in SUBSETS:
        base, idx = SUBSETS[name]
        w = synth.make(base)
        return w.subset_queries(list(idx)), w.meta["sensitivity"]
    raise ValueError(name)


def main(names):
    out_dir = pathlib.Path(__file__).resolve().parent / "search"
    out_dir.mkdir(exist_ok=True)
    for name in names:
        if name in READ_SUBSETS:
            import numpy as np
            w, idx, sens = read_workload(name)
            reads = [w.reads[i] for i in idx]
            nt = np.frombuffer("".join(reads).encode(), np.uint8)
            off = np.zeros(len(reads) + 1, np.uint64)
            np.cumsum([len(r) for r in reads], out=off[1:])
            t0 = time.time()
            r = pyoracle.search(w.db_letters, w.db_offsets, nt, off, sensitivity=sens,
                                threads=os.cpu_count() or 1, translated=True)
            text = r.format([w.read_ids[i] for i in idx], w.db_ids, READ_FIELDS)
            with gzip.GzipFile(out_dir / f"{name}.tsv.gz", "wb", mtime=0) as f:
                f.write(text.encode())
            print(f"{name}: {len(r)} HSPs in {time.time() - t0:.1f} s; counters {r.counters()}")
            continue
        w, sens = workload(name)
        t0 = time.time()
        r = pyoracle.search(w.db_letters, w.db_offsets, w.q_letters, w.q_offsets, sensitivity=sens,
                            threads=os.cpu_count() or 1)
        text = r.format(w.q_ids, w.db_ids, FIELDS)
        c = r.counters()
        with gzip.GzipFile(out_dir / f"{name}.tsv.gz", "wb", mtime=0) as f:
            f.write(text.encode())
        print(f"{name}: {len(r)} HSPs in {time.time() - t0:.1f} s; counters {c}")


if __name__ == "__main__":
    main(sys.argv[1:] or ["blastp-sensitive-1024"])
