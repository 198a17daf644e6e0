"""Pin the CPU oracle against an independent string-level builder.

naive_dbg counts edges from k+1-mer windows and walks unitigs edge by edge,
so it shares neither the k2-block aggregation nor the list ranking with the
oracle.  Small k makes branches, palindromes and cycles frequent.
"""
import numpy as np
import pytest

import naive_dbg as nd
from helpers import pack, random_reads, thresholds_L
from oracle.oracle import Oracle


def kstr(lo, hi, k):
    v = (int(hi) << 64) | int(lo)
    return "".join("ACGT"[(v >> (2 * (k - 1 - i))) & 3] for i in range(k))


def oracle_graph(o, k, nodes):
    return {kstr(n["kmer_lo"], n["kmer_hi"], k): list(map(int, n["counts"])) for n in nodes}


def naive_components(g, k):
    """.chain.components from the walked unitigs: node d of a unitig S (on
    the strand it is printed on) is K = S[d:d+k], printed as canon(K), the
    rep S[0:k], d and whether K is canonical, lines sorted by K
    (print_chain_node, debruijn_chain_operations.hpp:427-442;
    print_chain_nodes, compact_dbg_io.hpp:463-498)."""
    rows = []
    for rep, seq in nd.unitigs(g)[0]:
        assert seq[:k] == rep
        for d in range(len(seq) - k + 1):
            K = seq[d:d + k]
            rows.append((K, "%s\t%s\t%d\t%d\n" % (nd.canon(K), rep, d, int(K == nd.canon(K)))))
    rows.sort()
    return "".join(r for _, r in rows)


def naive_k2(reads, k):
    """The k2 multiset: distinct canonical (k-mer, biedge byte) pairs and
    their instance counts, from the string-level records."""
    from collections import Counter
    return Counter((x, (nd.DNA16[a] << 4) | nd.DNA16[b]) for x, a, b in nd.records(reads, k))


def oracle_k2(o, k):
    lo, hi, e, c = o.k2()
    got = {}
    for i in range(len(e)):
        key = (kstr(lo[i], hi[i], k), int(e[i]))
        assert key not in got
        got[key] = int(c[i])
    return got


def run(reads, k, th=None, variant=1, nthreads=1):
    o = Oracle(k, th, nthreads)
    b, off = pack(reads)
    o.add_reads(b, off)
    o.run(variant)
    return o


CASES = [
    # (seed, genome_len, n_reads, read_len, err, k)
    (1, 300, 40, 30, 0.0, 7),
    (2, 500, 80, (20, 60), 0.01, 9),
    (3, 2000, 300, 50, 0.01, 11),
    (4, 200, 200, 25, 0.02, 5),
    (5, 1000, 200, 40, 0.005, 15),
    (6, 4000, 400, 100, 0.01, 31),
    (7, 600, 120, (40, 120), 0.01, 31),
    (8, 300, 60, 40, 0.01, 8),     # even k: k-mer palindromes possible
    (9, 3000, 300, 120, 0.01, 63),
    (10, 150, 150, 30, 0.0, 6),
]


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("tl", [None, 1, 2, 3])
def test_oracle_matches_naive(case, tl):
    seed, G, n, L, err, k = case
    _, reads = random_reads(seed, G, n, L, err)
    th = None if tl is None else thresholds_L(tl)
    o = run(reads, k, th)
    # P1 always
    assert oracle_graph(o, k, o.unfiltered()) == nd.unfiltered(reads, k)
    # P2
    exp = nd.unfiltered(reads, k) if th is None else nd.filtered(reads, k, th)
    got = oracle_graph(o, k, o.nodes())
    assert got == exp
    # P3
    assert o.branch_fasta().decode() == nd.branch_fasta(exp)
    # P5 (compare as sets of (header, seq) and exact order when consistent)
    assert o.chain_fasta().decode() == nd.chain_fasta(exp)
    # P4: every chain node's (canonical k-mer, rep, dist, strand), derived
    # from the walked unitig strings
    assert o.chain_components().decode() == naive_components(exp, k)
    # the k2 counter before any filter
    assert oracle_k2(o, k) == dict(naive_k2(reads, k))


def test_planted_component_mistakes_are_caught():
    """A strand flag or distance taken from the wrong end would change the
    derived components: the derivation is not a restatement of the oracle."""
    _, reads = random_reads(3, 2000, 300, 50, 0.01)
    k = 11
    o = run(reads, k, thresholds_L(2))
    exp = nd.filtered(reads, k, thresholds_L(2))
    good = naive_components(exp, k)
    assert o.chain_components().decode() == good
    lines = good.splitlines()
    assert any(l.split("\t")[3] == "0" for l in lines) and any(l.split("\t")[3] == "1" for l in lines)
    flipped = "".join("\t".join(l.split("\t")[:3] + [str(1 - int(l.split("\t")[3]))]) + "\n" for l in lines)
    assert o.chain_components().decode() != flipped


def test_multithreaded_oracle_is_identical():
    _, reads = random_reads(11, 20000, 3000, 100, 0.01)
    for th in (None, thresholds_L(2)):
        a = run(reads, 31, th, nthreads=1)
        b = run(reads, 31, th, nthreads=4)
        assert a.chain_fasta() == b.chain_fasta()
        assert a.branch_fasta() == b.branch_fasta()
        assert np.array_equal(a.nodes(), b.nodes())
        assert np.array_equal(a.unfiltered(), b.unfiltered())


def test_crafted_palindromic_edge_halving():
    """k=5: ACGT|A... a k+1-mer that is its own reverse complement forces the
    (cnt+1)>>1 halving (debruijn_graph_map.hpp:1478, :1499).  Worked by hand:
    the 6-mer ACGCGT is an rc-palindrome; read 'TACGCGTA' (x3) gives k-mer
    ACGCG with out edge T from both strands."""
    k = 5
    reads = ["TACGCGTAA"] * 3
    o = run(reads, k, thresholds_L(1))
    got = oracle_graph(o, k, o.nodes())
    assert got == nd.filtered(reads, k, thresholds_L(1))
    # hand check of the halving: k-mer ACGCG (canonical) sees out 'T' from the
    # forward window and from the reverse strand window CGCGT -> rc ACGCG
    x = "ACGCG"
    assert nd.is_pal(x + "T")
    recs = [r for r in nd.records(reads, k) if r[0] == x]
    # 3 forward (in T, out T) + 3 reverse (in A... ) instances
    assert len(recs) == 6
    # with threshold 4 the halved count 3 fails while the raw count 6 would pass
    o4 = run(reads, k, thresholds_L(4))
    assert oracle_graph(o4, k, o4.nodes()) == nd.filtered(reads, k, thresholds_L(4))
    assert x not in oracle_graph(o4, k, o4.nodes()) or \
        oracle_graph(o4, k, o4.nodes())[x][3] == 0


def test_crafted_cycle_is_removed():
    """A circular genome read with wrap-around gives a pure cycle of chain
    nodes: all of them are removed and no unitig is printed."""
    k = 7
    rng = np.random.default_rng(5)
    g = "".join(rng.choice(list("ACGT"), 60))
    circ = g + g[:k + 10]
    reads = [circ[i:i + 20] for i in range(0, 60, 3)]
    o = run(reads, k)
    us, ncyc = nd.unitigs(nd.unfiltered(reads, k))
    assert o.chain_fasta().decode() == nd.chain_fasta(nd.unfiltered(reads, k))
    if not us:
        assert o.stat("n_cycle_nodes") == ncyc > 0


def test_reads_with_n_are_dropped_and_short_reads_yield_nothing():
    k = 7
    reads = ["ACGTACGTTGCA", "ACGTNACGTTGCAAT", "ACGTACGT"]  # 12 ok, N dropped, 8 < k+2
    o = run(reads, k)
    assert o.stat("n_instances") == 12 - k + 1
