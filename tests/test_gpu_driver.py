"""The C++ host driver (bruno_amd/host/compact_dbg, over the C ABI) writes
the same .branch.fasta / .chain.fasta / .chain.components / .branch.edges /
.chain.edges / .graph.all bytes as the oracle for the reference fixtures."""
import os
import subprocess

import numpy as np
import pytest

from helpers import REF_DATA, ROOT, pack, read_fastq, thresholds_L
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

DRIVER = os.path.join(ROOT, "bruno_amd", "host", "compact_dbg")


@pytest.mark.parametrize("name,args,th,variant", [
    ("test.unitiqs.fastq", [], None, 0),
    ("test.star.unitiqs.fastq", ["-R"], None, 1),
    ("debruijn.palindrome_chr14.fastq", ["-R", "-T", "-L", "2"], thresholds_L(2), 1),
    ("test.medium.fastq", ["-R", "-T", "-L", "1"], thresholds_L(1), 1),
])
def test_driver_files_match_oracle(tmp_path, name, args, th, variant):
    path = os.path.join(REF_DATA, name)
    prefix = str(tmp_path / "out")
    r = subprocess.run([DRIVER] + args + ["-O", prefix, path], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    seqs, _ = read_fastq(path)
    b, off = pack(seqs)
    o = Oracle(31, None if th is None else np.array(th, np.uint64), 1)
    o.add_reads(b, off)
    o.run(variant)
    assert open(prefix + ".chain.fasta", "rb").read() == o.chain_fasta()
    assert open(prefix + ".branch.fasta", "rb").read() == o.branch_fasta()
    assert open(prefix + ".chain.components", "rb").read() == o.chain_components()
    assert open(prefix + ".branch.edges", "rb").read() == o.branch_edges()
    assert open(prefix + ".chain.edges", "rb").read() == o.chain_edges()
    assert open(prefix + ".graph.all", "rb").read() == graph_all(o.nodes(), 31)


def graph_all(nodes, k):
    """.graph.all from the oracle's filtered graph (P2): nodes sorted by
    k-mer, print_graph_node columns (k-mer, in A C G T, out A C G T, self)."""
    order = np.lexsort((nodes["kmer_lo"], nodes["kmer_hi"]))
    out = []
    for r in nodes[order]:
        v = int(r["kmer_lo"]) | (int(r["kmer_hi"]) << 64)
        ks = "".join("ACGT"[(v >> (2 * (k - 1 - i))) & 3] for i in range(k))
        c = [int(x) for x in r["counts"]]
        out.append("\t".join([ks] + [str(x) for x in c[4:8] + c[0:4] + [c[8]]]) + "\n")
    return "".join(out).encode()


def test_driver_several_files_grow_the_table(tmp_path):
    """A small first file sizes the table (from its sketch); the second,
    with many more distinct k-mers, makes the job move its table into a
    larger one.  The outputs equal the oracle on all reads together."""
    from helpers import random_reads
    _, reads = random_reads(21, 30000, 3000, 120, 0.01)
    files = []
    for i, part in enumerate((reads[:30], reads[30:])):
        fn = str(tmp_path / ("part%d.fastq" % i))
        with open(fn, "w") as f:
            f.write("".join("@r%d\n%s\n+\n%s\n" % (j, s, "I" * len(s)) for j, s in enumerate(part)))
        files.append(fn)
    prefix = str(tmp_path / "out")
    r = subprocess.run([DRIVER, "-R", "-T", "-L", "2", "-O", prefix] + files, capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    b, off = pack(reads)
    o = Oracle(31, np.array(thresholds_L(2), np.uint64), 1)
    o.add_reads(b, off)
    o.run(1)
    assert open(prefix + ".chain.fasta", "rb").read() == o.chain_fasta()
    assert open(prefix + ".branch.fasta", "rb").read() == o.branch_fasta()
    assert open(prefix + ".graph.all", "rb").read() == graph_all(o.nodes(), 31)
