"""Pin the oracle's non-benchmark outputs .branch.edges
(print_branch_edge_frequencies, compact_dbg_stats.hpp:219-262) and
.chain.edges (compute_freq_map + print_chain_frequencies,
compact_dbg_stats.hpp:783-1065) against the string-level builder: branch
k-mers with their counts, and per walked unitig its end k-mers, their outer
edge counts and the float32 mean / min / max of its self counts."""
import numpy as np
import pytest

import naive_dbg as nd
from helpers import random_reads, thresholds_L
from test_oracle_naive import CASES, run


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("tl", [None, 2])
def test_edge_outputs_match_naive(case, tl):
    seed, G, n, L, err, k = case
    _, reads = random_reads(seed, G, n, L, err)
    th = None if tl is None else thresholds_L(tl)
    o = run(reads, k, th)
    g = nd.unfiltered(reads, k) if th is None else nd.filtered(reads, k, th)
    assert o.stat("chain_edge_misses") == 0
    assert o.branch_edges().decode() == nd.branch_edges(g)
    assert o.chain_edges().decode() == nd.chain_edges(g, k)


def test_chain_edges_worked_example():
    """k=5, the 12-base genome CAGATTTTCATA read 3 times and its first 9
    bases twice more: one unitig (no branch) of 8 k-mers, L = CAGAT (its rep
    end), R = TCATA, no outer edges; the first 5 k-mers lie in all 5 reads,
    the last 3 in 3: sum 34 over 8 nodes, mean 4 (4.25 truncated), min 3,
    max 5."""
    genome = "CAGATTTTCATA"
    k = 5
    reads = [genome] * 3 + [genome[:9]] * 2
    o = run(reads, k)
    g = nd.unfiltered(reads, k)
    assert nd.classify(g)[0] == set()
    (rep, seq), = nd.unitigs(g)[0]
    assert seq == genome
    assert [g[nd.canon(seq[i:i + k])][8] for i in range(8)] == [5, 5, 5, 5, 5, 3, 3, 3]
    line = o.chain_edges().decode()
    assert line == "CAGAT\tTCATA\t" + "0\t" * 8 + "4\t3\t5\t\n"
    assert line == nd.chain_edges(g, k)
    assert o.branch_edges() == b""


def test_chain_edges_float_mean():
    """k=5, TTATGCAGA x4, its first 7 bases x2 and bases 1..7 once: the
    unitig is printed from its rep end TCTGC (the reverse strand); self
    counts 4 5 7 7 6 on that strand, sum 29 over 5, mean 5 (5.8 truncated,
    as the float -> 16-bit count conversion does)."""
    k = 5
    reads = ["TTATGCAGA"] * 4 + ["TTATGCA"] * 2 + ["TATGCAG"]
    o = run(reads, k)
    g = nd.unfiltered(reads, k)
    (rep, seq), = nd.unitigs(g)[0]
    assert seq == "TCTGCATAA"
    assert [g[nd.canon(seq[i:i + k])][8] for i in range(5)] == [4, 5, 7, 7, 6]
    line = o.chain_edges().decode()
    assert line == "TCTGC\tCATAA\t" + "0\t" * 8 + "5\t4\t7\t\n"
    assert line == nd.chain_edges(g, k)


def test_branch_edges_worked_example():
    """k=5, two reads sharing a k-mer then diverging: the shared k-mer has
    two out edges (a branch); its line lists in A C G T, out A C G T, self."""
    k = 5
    reads = ["GACTTAGCA"] * 2 + ["GACTTAGTA"]
    o = run(reads, k)
    g = nd.unfiltered(reads, k)
    br = sorted(nd.classify(g)[0])
    assert o.branch_edges().decode() == nd.branch_edges(g)
    assert len(br) >= 1
    for b in br:
        line = [ln for ln in o.branch_edges().decode().splitlines() if ln.startswith(b)][0]
        n = g[b]
        assert line.split("\t")[1:] == [str(x) for x in n[4:8] + n[0:4] + [n[8]]]


@pytest.mark.parametrize("name,tl", [("test.unitiqs.fastq", None), ("test.short.unitiqs.fastq", None),
                                     ("debruijn.palindrome_chr14.fastq", 2)])
def test_fixture_edge_outputs(name, tl):
    """The reference's own test reads: every chain end finds its chain's
    frequency summary, and both files equal the string-level builder's."""
    import os
    from helpers import REF_DATA, pack, read_fastq
    from oracle.oracle import Oracle
    seqs, _ = read_fastq(os.path.join(REF_DATA, name))
    th = None if tl is None else thresholds_L(tl)
    o = Oracle(31, None if th is None else np.array(th, np.uint64), 1)
    b, off = pack(seqs)
    o.add_reads(b, off)
    o.run(1)
    g = nd.unfiltered(seqs, 31) if th is None else nd.filtered(seqs, 31, th)
    assert o.stat("chain_edge_misses") == 0
    assert o.branch_edges().decode() == nd.branch_edges(g)
    assert o.chain_edges().decode() == nd.chain_edges(g, 31)


@pytest.mark.parametrize("case", CASES[:6])
def test_kmer_only_thresholds_do_not_break_edge_outputs(case):
    """-L 3 0 0 keeps edges into removed k-mers, so a ranked chain can carry
    a removed k-mer as its far end (the reference asserts there).  The
    oracle still finishes; ends it cannot find count as misses with zero
    edge counts, and the branch lines are unaffected."""
    seed, G, n, L, err, k = case
    _, reads = random_reads(seed, G, n, L, err)
    o = run(reads, k, [3, 65536, 0, 65536, 0, 65536])
    g = nd.filtered(reads, k, [3, 65536, 0, 65536, 0, 65536])
    assert o.branch_edges().decode() == nd.branch_edges(g)
    lines = o.chain_edges().decode().splitlines()
    assert len(lines) == len(o.chain_fasta().decode().split(">")) - 1


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("lo,hi", [(2, 65536), (3, 40)])
def test_edge_filter_mode_outputs_match_recount(case, lo, hi):
    """-E (block-construct driver): its .branch.edges / .chain.edges counts
    come from recounting the reads' records after the k+1-mer filter
    (count_edges / count_kmers, compact_dbg_stats.hpp:144-170, :645-689:
    parse_and_filter_nodes, then saturating 16-bit counters for the selected
    k-mers).  The string-level builder's edge-filtered graph holds exactly
    those counts (kept records, kept edges, saturating), computed from the
    read strings, so the oracle's edge-mode outputs are pinned by it."""
    from oracle.oracle import Oracle
    from helpers import pack
    seed, G, n, L, err, k = case
    _, reads = random_reads(seed, G, n, L, err)
    th = [0, 65536, lo, hi, 0, 65536]
    o = Oracle(k, np.array(th, np.uint64), 1, edge_filter=True)
    b, off = pack(reads)
    o.add_reads(b, off)
    o.run(1)
    g = nd.edge_filtered(reads, k, lo, hi)
    assert o.stat("chain_edge_misses") == 0
    assert o.branch_edges().decode() == nd.branch_edges(g)
    assert o.chain_edges().decode() == nd.chain_edges(g, k)
