"""Writes the golden fixtures under tests/golden/ from the reference repository.

Reads (as text/data only):
  src/matrix_data.rs   -> blosum62.json (25x25 scores + 11/1 gap row)
  bench/e2e/*.fa       -> e2e/*.fa.gz (the reference's bundled benchmark data)

Usage: python tests/golden/make_golden.py [/root/reference]
"""
from __future__ import annotations

import gzip
import json
import pathlib
import re
import sys

REF = pathlib.Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
OUT = pathlib.Path(__file__).resolve().parent


def blosum62():
    text = (REF / "src" / "matrix_data.rs").read_text()
    start = text.index('name: "BLOSUM62"')
    block = text[start:text.index("StdMatrix {", start)]
    scores = block[block.index("scores: ["):block.index("params: &[")]
    rows = re.findall(r"\[([-\d,\s]+)\]", scores)
    mat = [[int(x) for x in r.replace("\n", " ").split(",") if x.strip()] for r in rows]
    mat = [r for r in mat if len(r) == 25]
    assert len(mat) == 25, len(mat)
    params = re.findall(r"\((\d+), (\d+), ([\d.]+), ([\d.]+), ([\d.]+), ([\d.]+), ([\d.]+)\)",
                        block[block.index("params: &["):])
    ung = re.search(r"ungapped: \(([\d.]+), ([\d.]+)\)", block)
    return {
        "scores": mat,
        "params": [[int(p[0]), int(p[1])] + [float(x) for x in p[2:]] for p in params],
        "ungapped": [float(ung.group(1)), float(ung.group(2))],
        "source": "src/matrix_data.rs (BLOSUM62 block)",
    }


def main():
    (OUT / "blosum62.json").write_text(json.dumps(blosum62(), indent=1))
    e2e = OUT / "e2e"
    e2e.mkdir(exist_ok=True)
    for name in ("db_prot.fa", "query.fa", "subject.fa"):
        data = (REF / "bench" / "e2e" / name).read_bytes()
        with gzip.GzipFile(e2e / (name + ".gz"), "wb", mtime=0) as f:
            f.write(data)
    print("wrote", sorted(p.name for p in OUT.rglob("*") if p.is_file()))


if __name__ == "__main__":
    main()
