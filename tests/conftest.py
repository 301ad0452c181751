import gzip
import importlib.util
import os
import pathlib
import sys

import numpy as np
import pytest

ROOT = pathlib.Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT / "oracle"))
sys.path.insert(0, str(ROOT / "tests"))
sys.path.insert(0, str(ROOT / "tests" / "golden"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")


def load_package():
    """The product package (directory name is not a Python identifier)."""
    name = "diamond_rust_amd"
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(
        name, ROOT / "diamond-rust_amd" / "__init__.py",
        submodule_search_locations=[str(ROOT / "diamond-rust_amd")])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def load_synth():
    name = "diamond_rust_amd.synth"
    if name in sys.modules:
        return sys.modules[name]
    load_package()
    spec = importlib.util.spec_from_file_location(name, ROOT / "diamond-rust_amd" / "synth.py")
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def read_fasta_gz(path):
    """FASTA records as (header, raw residues) like io::parse_fasta
    (src/io.rs:40-92): upper-cased, whitespace, '-' and '.' stripped."""
    recs, head, seq = [], None, []
    for line in gzip.open(path, "rb").read().split(b"\n"):
        if line.endswith(b"\r"):
            line = line[:-1]
        if line.startswith(b">"):
            if head is not None:
                recs.append((head, b"".join(seq)))
            head = line[1:].decode().strip()
            seq = []
        elif head is not None:
            seq.append(bytes(c for c in line.upper() if c not in b" \t\x0c\r-."))
    if head is not None:
        recs.append((head, b"".join(seq)))
    return recs


@pytest.fixture(scope="session")
def dmr():
    return load_package()


@pytest.fixture(scope="session")
def synth():
    return load_synth()


@pytest.fixture(scope="session")
def orc():
    import pyoracle
    pyoracle.lib()
    return pyoracle


@pytest.fixture(scope="session")
def ctx(dmr):
    c = dmr.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="session")
def e2e(orc):
    g = ROOT / "tests" / "golden" / "e2e"
    db = read_fasta_gz(g / "db_prot.fa.gz")
    q = read_fasta_gz(g / "query.fa.gz")
    enc = lambda s: orc.encode(s.decode())
    dbs = [enc(s) for _, s in db]
    qs = [enc(s) for _, s in q]
    off = np.zeros(len(dbs) + 1, np.uint64)
    np.cumsum([len(x) for x in dbs], out=off[1:])
    qo = np.zeros(len(qs) + 1, np.uint64)
    np.cumsum([len(x) for x in qs], out=qo[1:])
    # Full headers (qtitle / stitle) as io::parse_fasta keeps them.
    heads = lambda name: [h for h, _, _ in orc.parse_fasta(gzip.open(g / name, "rb").read())]
    return dict(db_letters=np.concatenate(dbs), db_offsets=off, q_letters=np.concatenate(qs),
                q_offsets=qo, db_ids=[h.split()[0] for h, _ in db],
                q_ids=[h.split()[0] for h, _ in q], db_headers=heads("db_prot.fa.gz"),
                q_headers=heads("query.fa.gz"))


def rand_seq(rng, n, canonical_only=True):
    hi = 20 if canonical_only else 25
    return rng.integers(0, hi, size=n, dtype=np.uint8)


def mutate(rng, s, sub=0.3, indel=0.03):
    out = []
    for c in s:
        r = rng.random()
        if r < indel / 2:
            continue
        if r < indel:
            out.append(int(rng.integers(0, 20)))
        out.append(int(rng.integers(0, 20)) if rng.random() < sub else int(c))
    return np.array(out, dtype=np.uint8)
