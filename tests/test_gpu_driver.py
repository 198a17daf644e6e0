"""The C++ host driver (bruno_amd/host/compact_dbg, over the C ABI) writes
the same .branch.fasta / .chain.fasta / .chain.components bytes as the
oracle for the reference fixtures and a synthetic batch."""
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
