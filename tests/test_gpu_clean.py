"""Graph cleaning on the GPU (bruno_clean) against the oracle.

After the first compaction, bruno_clean runs bruno's deadend / bubble
cleaning loop; the job then holds the cleaned graph, and every getter must
equal the oracle's .clean.* outputs: unitig FASTA, branch FASTA, chain nodes
(P4), the graph's node records, and the branch / chain edge outputs, with
the same number of rounds.
"""
import os

import numpy as np
import pytest

import bruno_amd
from gpu_common import node_table, oracle_node_table
from helpers import REF_DATA, pack, random_reads, rc, read_fastq, thresholds_L
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

FIXTURES = ["debruijn.bubble.fastq", "debruijn.bubble2.fastq", "debruijn.palindrome_chr14.fastq",
            "debruijn.shifted_cycle_chr14.fastq", "test.debruijn.small.fastq", "test.star.unitiqs.fastq",
            "debruijn.palindrome.fastq", "natural.fastq"]


def clean_both(k, t, reads, variant=1, capacity=0):
    th = thresholds_L(t) if t else None
    k1_lo = th[2] if th else 65536
    b, off = pack(reads)
    job = bruno_amd.GraphJob(k, th, capacity=capacity)
    job.parse_count(b, off)
    job.run(variant)
    pre = (job.unitigs(), job.branch_fasta())
    st = job.clean(variant, k1_lo)
    g = dict(stats=st, pre=pre, fasta=job.unitigs(), branch_fasta=job.branch_fasta(), nodes=job.nodes(),
             chain_nodes=job.chain_nodes(), branch_edges=job.branch_edges_text(), chain_edges=job.chain_edges_text(),
             job_stats=job.stats())
    job.close()
    o = Oracle(k, None if th is None else np.array(th, np.uint64))
    o.add_reads(b, off)
    o.run(variant)
    opre = (o.chain_fasta(), o.branch_fasta())
    o.run_clean(variant)
    return g, o, opre


def check(k, g, o, opre):
    assert g["pre"] == opre
    st = g["stats"]
    assert st["rounds"] == o.stat("clean_rounds")
    assert st["deadends"] == o.stat("clean_deadends")
    assert st["bubble_chains"] == o.stat("clean_bubbles")
    assert g["fasta"] == o.clean_output("chain_fasta")
    assert g["branch_fasta"] == o.clean_output("branch_fasta")
    assert node_table(g["nodes"]) == oracle_node_table(o.clean_nodes())
    on, gn = o.clean_chain_nodes(), g["chain_nodes"]
    assert len(gn) == len(on)
    if len(on):
        assert np.array_equal(gn["k"][:, 0], on["k_lo"])
        assert np.array_equal(gn["rep"][:, 0], on["rep_lo"])
        assert np.array_equal(gn["dist"], on["dist"])
        if k > 31:
            assert np.array_equal(gn["k"][:, 1], on["k_hi"])
            assert np.array_equal(gn["rep"][:, 1], on["rep_hi"])
    assert g["branch_edges"] == o.clean_output("branch_edges")
    assert g["chain_edges"] == o.clean_output("chain_edges")


@pytest.mark.parametrize("name", FIXTURES)
@pytest.mark.parametrize("k,t", [(21, None), (31, None), (31, 2), (21, 3)])
def test_fixture_clean(name, k, t):
    reads, _ = read_fastq(os.path.join(REF_DATA, name))
    g, o, opre = clean_both(k, t, reads)
    check(k, g, o, opre)


def tip_bubble_reads(seed, glen=400):
    rng = np.random.default_rng(100 + seed)
    g, reads = random_reads(seed, glen, 120, (40, 80), 0.0)
    for _ in range(6):
        st = int(rng.integers(0, len(g) - 60))
        s = list(g[st:st + 60])
        s[30] = "ACGT"[("ACGT".index(s[30]) + 1 + int(rng.integers(0, 3))) % 4]
        reads += ["".join(s)] * int(rng.integers(1, 3))
    for _ in range(8):
        st = int(rng.integers(0, len(g) - 50))
        s = list(g[st:st + 50])
        s[-3] = "ACGT"[("ACGT".index(s[-3]) + 1) % 4]
        reads.append("".join(s) if rng.random() < 0.5 else rc("".join(s)))
    return reads


@pytest.mark.parametrize("seed", range(8))
@pytest.mark.parametrize("variant", [0, 1])
def test_random_tips_and_bubbles(seed, variant):
    k = [11, 15, 21, 33, 45, 63, 12, 32][seed]
    reads = tip_bubble_reads(seed)
    for t in (None, 2, 3):
        g, o, opre = clean_both(k, t, reads, variant)
        check(k, g, o, opre)


def test_region_table_clean():
    """A region-layout table (counted by regions; count_regions 2 allows
    them from 2^22 slots) cleans the same."""
    from bruno_amd.autotune import get_tuning, set_tuning
    old = get_tuning()
    set_tuning(count_regions=2)
    try:
        reads = tip_bubble_reads(3) * 3
        g, o, opre = clean_both(31, 2, reads, capacity=1 << 22)
    finally:
        set_tuning(**old)
    assert g["job_stats"]["count_by_regions"] == 1
    check(31, g, o, opre)


def test_clean_state_errors():
    b, off = pack(["ACGTACGTTGCAACGTAGGCTAGCTAGGACT" * 3])
    job = bruno_amd.GraphJob(15, thresholds_L(1))
    job.parse_count(b, off)
    with pytest.raises(bruno_amd.BrunoError) as e:
        job.clean(1, 1)
    assert e.value.code == 5  # BRUNO_STATE: not compacted yet
    job.run(1)
    st = job.clean(1, 1)
    assert st["rounds"] >= 0
    job.close()


@pytest.mark.parametrize("name,args,t", [
    ("debruijn.bubble.fastq", ["-R"], None),
    ("debruijn.palindrome_chr14.fastq", ["-R", "-T", "-L", "2"], 2),
    ("test.debruijn.small.fastq", ["-T", "-L", "2"], 2),
])
def test_driver_clean_files(tmp_path, name, args, t):
    """compact_dbg --clean writes the first compaction's files and the
    .clean.* files of the cleaned graph, each equal to the oracle's."""
    import subprocess
    from helpers import ROOT
    driver = os.path.join(ROOT, "bruno_amd", "host", "compact_dbg")
    path = os.path.join(REF_DATA, name)
    prefix = str(tmp_path / "out")
    r = subprocess.run([driver] + args + ["--clean", "-O", prefix, path], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    seqs, _ = read_fastq(path)
    b, off = pack(seqs)
    th = thresholds_L(t) if t else None
    variant = 1 if "-R" in args else 0
    o = Oracle(31, None if th is None else np.array(th, np.uint64))
    o.add_reads(b, off)
    o.run(variant)
    assert open(prefix + ".chain.fasta", "rb").read() == o.chain_fasta()
    assert open(prefix + ".branch.fasta", "rb").read() == o.branch_fasta()
    o.run_clean(variant)
    for suffix, name_ in ((".chain.fasta", "chain_fasta"), (".branch.fasta", "branch_fasta"),
                          (".chain.components", "chain_components"), (".branch.edges", "branch_edges"),
                          (".chain.edges", "chain_edges")):
        assert open(prefix + ".clean" + suffix, "rb").read() == o.clean_output(name_), suffix
