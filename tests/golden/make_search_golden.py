"""Writes the oracle's outfmt-6 text of whole synthetic workloads as fixtures.

Usage: python tests/golden/make_search_golden.py NAME [NAME ...]
  blastp-sensitive-1024   1024 queries vs the 2.26 Maa DB, --sensitive (one
                          batch: the max_expand cap counts the whole batch)

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


def workload(name):
    synth = load_synth()
    if name == "blastp-sensitive-1024":
        return synth.make(name), "sensitive"
    raise ValueError(name)


def main(names):
    out_dir = pathlib.Path(__file__).resolve().parent / "search"
    out_dir.mkdir(exist_ok=True)
    for name in names:
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
