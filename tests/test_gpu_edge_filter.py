"""The block-construct driver's k+1-mer edge-frequency filter
(compact_debruijn_graph_block_construct.cpp:415-480) on the GPU
(bruno_job_desc.thresholding 2) against the oracle's edge mode: every parity
object, on random genomes of odd and even k, the rc-palindrome special
cases, spilled k2 entries, the region count, a batch and the host driver."""
import os
import subprocess

import numpy as np
import pytest

import bruno_amd
from bruno_amd.autotune import get_tuning, set_tuning
from gpu_common import compare_all, node_table, run_gpu
from helpers import DEFAULT_T, REF_DATA, ROOT, pack, random_reads, read_fastq
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

CASES = [
    (1, 300, 40, 30, 0.0, 7),
    (3, 2000, 300, 50, 0.01, 11),
    (4, 200, 200, 25, 0.02, 5),
    (6, 4000, 400, 100, 0.01, 31),
    (7, 600, 120, (40, 120), 0.01, 31),
    (8, 300, 60, 40, 0.01, 8),
    (9, 3000, 300, 120, 0.01, 63),
    (10, 150, 150, 30, 0.0, 6),
    (12, 2000, 200, 80, 0.01, 40),
]


def edge_thresholds(lo, hi=DEFAULT_T):
    return [0, DEFAULT_T, lo, hi, 0, DEFAULT_T]


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("lo,hi", [(1, DEFAULT_T), (2, DEFAULT_T), (3, DEFAULT_T), (2, 6)])
def test_edge_filter_matches_oracle(case, lo, hi):
    seed, G, n, L, err, k = case
    _, reads = random_reads(seed, G, n, L, err)
    b, off = pack(reads)
    compare_all(k, edge_thresholds(lo, hi), b, off, edge_filter=True)


@pytest.mark.parametrize("k,reads,lo", [
    # ACGCGT is its own reverse complement: two k2 entries of 3 for 3 windows
    (5, ["TACGCGTCA"] * 3, 3),
    (5, ["TACGCGTCA"] * 3, 4),
    # GAATTC is its own reverse complement: windows seen from both sides
    (6, ["TTGAATTCAA"] * 2 + ["CAGAATTCGT"] * 3 + ["ATGAATTCGC"] + ["TAGAATTCCG"] * 4 + ["AGAATTC"] * 2, 3),
    (6, ["TTGAATTCAA"] * 2 + ["CAGAATTCGT"] * 3 + ["ATGAATTCGC"] + ["TAGAATTCCG"] * 4 + ["AGAATTC"] * 2, 5),
    # one read "A GAATTC T": the k-mer occurs once, its k+1-mer twice
    (6, ["AGAATTCT"], 2),
    (6, ["AGAATTCT"], 3),
])
def test_palindromes(k, reads, lo):
    b, off = pack(reads)
    compare_all(k, edge_thresholds(lo), b, off, edge_filter=True)


def test_window_count_twice_the_kmer_count():
    """The filter pre-pass keeps a k-mer whose only window count is twice
    its own count: a node survives at lower 2 although its k-mer occurs
    once."""
    b, off = pack(["AGAATTCT"])
    g = run_gpu(6, edge_thresholds(2), b, off, edge_filter=True)
    assert any(c[8] == 1 for c in node_table(g["nodes"]).values())


@pytest.mark.parametrize("case", CASES[:4])
def test_lower_zero_is_the_plain_graph(case):
    seed, G, n, L, err, k = case
    _, reads = random_reads(seed, G, n, L, err)
    b, off = pack(reads)
    e = run_gpu(k, edge_thresholds(0), b, off, edge_filter=True, stages=False)
    p = run_gpu(k, None, b, off, stages=False)
    assert e["fasta"] == p["fasta"] and e["branch_fasta"] == p["branch_fasta"]
    assert node_table(e["nodes"]) == node_table(p["nodes"])


def test_spilled_entries():
    """k-mers with all 16 biedges: the candidate kernel for slots with side
    table entries applies the same filter."""
    rng = np.random.default_rng(5)
    reads = []
    for _ in range(30):
        core = "".join(rng.choice(list("ACGT"), 31))
        for i, a in enumerate("ACGT"):
            for j, c in enumerate("ACGT"):
                r = "".join(rng.choice(list("ACGT"), 10)) + a + core + c + "".join(rng.choice(list("ACGT"), 10))
                reads += [r] * (1 + (i + j) % 3)
    b, off = pack(reads)
    for lo in (2, 4):
        g, _ = compare_all(31, edge_thresholds(lo), b, off, edge_filter=True)
        assert g["stats"]["n_overflow"] >= 30


@pytest.fixture
def region_count():
    old = get_tuning()
    set_tuning(count_regions=2)
    yield
    set_tuning(**old)


@pytest.mark.parametrize("k", [15, 31])
def test_region_count(k, region_count):
    """The region build's filter bits use the same pre-pass bound."""
    _, reads = random_reads(21, 50000, 3000, 150, 0.02)
    b, off = pack(reads)
    for lo in (1, 3):
        g, _ = compare_all(k, edge_thresholds(lo), b, off, capacity=1 << 22, edge_filter=True)
        assert g["stats"]["count_by_regions"] == 1


def test_batch_jobs_match_oracle():
    jobs = [random_reads(30 + i, G, n, 100, 0.01)[1] for i, (G, n) in enumerate([(3000, 300), (800, 100), (5000, 400)])]
    jobs.append(["AGAATTCT", "TTGAATTCAA"])
    reads = [r for j in jobs for r in j]
    b, off = pack(reads)
    first = np.zeros(len(jobs) + 1, np.uint64)
    first[1:] = np.cumsum([len(j) for j in jobs])
    th = edge_thresholds(2)
    k = 31
    bt = bruno_amd.GraphBatch(k, th, [max(256, 2 * sum(len(r) for r in j)) for j in jobs], edge_filter=True)
    bt.parse_count(b, off, first)
    bt.run()
    for i, j in enumerate(jobs):
        o = Oracle(k, np.array(th, np.uint64), 1, edge_filter=True)
        jb, jo = pack(j)
        o.add_reads(jb, jo)
        o.run(1)
        v = bt.job(i)
        assert v.unitigs() == o.chain_fasta()
        assert v.branch_fasta() == o.branch_fasta()
    bt.close()


def test_driver_edge_flag(tmp_path):
    """compact_dbg -E -T -L 2 writes the oracle's edge-filtered files."""
    driver = os.path.join(ROOT, "bruno_amd", "host", "compact_dbg")
    path = os.path.join(REF_DATA, "test.medium.fastq")
    prefix = str(tmp_path / "out")
    r = subprocess.run([driver, "-E", "-R", "-T", "-L", "2", "-O", prefix, path], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    seqs, _ = read_fastq(path)
    b, off = pack(seqs)
    o = Oracle(31, np.array([0, DEFAULT_T, 2, DEFAULT_T, 0, DEFAULT_T], np.uint64), 1, edge_filter=True)
    o.add_reads(b, off)
    o.run(1)
    assert open(prefix + ".chain.fasta", "rb").read() == o.chain_fasta()
    assert open(prefix + ".branch.fasta", "rb").read() == o.branch_fasta()
    assert open(prefix + ".chain.components", "rb").read() == o.chain_components()
    assert open(prefix + ".branch.edges", "rb").read() == o.branch_edges()
    assert open(prefix + ".chain.edges", "rb").read() == o.chain_edges()
