"""Pin the oracle's k+1-mer edge-frequency filter (the block-construct
driver: compute_edge_frequency + transform_biedges_by_frequency +
filter_isolated_nodes, compact_debruijn_graph_block_construct.cpp:415-480)
against naive_dbg.edge_filtered, which counts k+1-mer windows in the reads
directly instead of deriving them from the k2 multiset."""
import pytest

import naive_dbg as nd
from helpers import DEFAULT_T, pack, random_reads
from oracle.oracle import Oracle
from test_oracle_naive import CASES, kstr, oracle_graph


def edge_thresholds(lo, hi=DEFAULT_T):
    """-T -L lo -U hi of the block-construct driver: lower (inclusive) and
    upper (exclusive) canonical k+1-mer counts; the other entries unused."""
    return [0, DEFAULT_T, lo, hi, 0, DEFAULT_T]


def run(reads, k, th, variant=1):
    o = Oracle(k, th, 1, edge_filter=True)
    b, off = pack(reads)
    o.add_reads(b, off)
    o.run(variant)
    return o


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("lo,hi", [(1, DEFAULT_T), (2, DEFAULT_T), (3, DEFAULT_T), (2, 6)])
def test_edge_filter_matches_window_counts(case, lo, hi):
    seed, G, n, L, err, k = case
    _, reads = random_reads(seed, G, n, L, err)
    o = run(reads, k, edge_thresholds(lo, hi))
    exp = nd.edge_filtered(reads, k, lo, hi)
    assert oracle_graph(o, k, o.nodes()) == exp
    assert o.branch_fasta().decode() == nd.branch_fasta(exp)
    assert o.chain_fasta().decode() == nd.chain_fasta(exp)


def test_palindromic_k1mer_counts_windows_once():
    """k=5, reads 'TACGCGTCA' x3: the 6-mer ACGCGT is its own reverse
    complement, so each of its 3 windows gives two records of ACGCG with
    out-edge T, one per strand: in-edge T from the forward strand, in-edge
    G (the complement of the C after the window) from the reverse one, so
    two k2 entries of 3.  The window count is 3; the k2 filter halves each
    entry with rounding up, 2 + 2 = 4.  At a threshold of 4 the edge filter
    drops the edge and the k2 filter keeps it."""
    k = 5
    reads = ["TACGCGTCA"] * 3
    x = "ACGCG"
    assert nd.is_pal(x + "T")
    assert nd.kp1_counts(reads, k)["ACGCGT"] == 3
    for lo in (3, 4):
        o = run(reads, k, edge_thresholds(lo))
        g = oracle_graph(o, k, o.nodes())
        assert g == nd.edge_filtered(reads, k, lo, DEFAULT_T)
        assert (g.get(x, [0] * 9)[3] > 0) == (lo <= 3)
    main = Oracle(k, edge_thresholds(4), 1)
    b, off = pack(reads)
    main.add_reads(b, off)
    main.run(1)
    assert oracle_graph(main, k, main.nodes())[x][3] > 0


@pytest.mark.parametrize("case", CASES[:6])
def test_unthresholded_edge_filter_is_the_plain_graph(case):
    """With lower 0 every edge passes: the block driver's output equals the
    unthresholded graph, as the reference's own comparisons expect."""
    seed, G, n, L, err, k = case
    _, reads = random_reads(seed, G, n, L, err)
    o = run(reads, k, edge_thresholds(0))
    plain = Oracle(k, None, 1)
    b, off = pack(reads)
    plain.add_reads(b, off)
    plain.run(1)
    assert o.chain_fasta() == plain.chain_fasta()
    assert o.branch_fasta() == plain.branch_fasta()


@pytest.mark.parametrize("lo", [1, 2, 3, 4, 5])
def test_palindromic_kmer_windows(lo):
    """k=6 around GAATTC (its own reverse complement): the instances of a
    palindromic k-mer are recorded from the reverse strand, so a window
    c.GAATTC appears as an out-edge comp(c) as well as an in-edge c.  Flanks
    of both strands at several counts, each threshold against the window
    count."""
    k = 6
    core = "GAATTC"
    assert nd.is_pal(core)
    reads = (["TT" + core + "AA"] * 2 + ["CA" + core + "GT"] * 3 + ["AT" + core + "GC"]
             + [nd.rc("CG" + core + "TA")] * 4 + ["A" + core] * 2)
    o = run(reads, k, edge_thresholds(lo))
    exp = nd.edge_filtered(reads, k, lo, DEFAULT_T)
    assert oracle_graph(o, k, o.nodes()) == exp
    assert o.chain_fasta().decode() == nd.chain_fasta(exp)


@pytest.mark.parametrize("lo", [1, 2, 3])
def test_window_count_twice_the_kmer_count(lo):
    """One read 'A GAATTC T' (k=6): the palindromic k-mer occurs once, the
    canonical k+1-mer of its two windows (AGAATTC, GAATTCT) twice."""
    reads = ["AGAATTCT"]
    assert nd.kp1_counts(reads, 6)["AGAATTC"] + nd.kp1_counts(reads, 6)["GAATTCT"] == 2
    o = run(reads, 6, edge_thresholds(lo))
    exp = nd.edge_filtered(reads, 6, lo, DEFAULT_T)
    assert oracle_graph(o, 6, o.nodes()) == exp
    assert ("GAATTC" in exp) == (lo <= 2)
