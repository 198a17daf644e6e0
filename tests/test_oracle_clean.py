"""Pin the oracle's graph cleaning (deadends, bubbles, recompaction).

bruno's cleaning loop (clean_deadend_and_bubbles_naive,
test/common/compact_dbg_clean.hpp:560-850) has no golden output in the
reference.  The oracle follows it through chain summaries built from list
ranking results (to_summarized_chains, debruijn_chain_graph.hpp:2044-2180);
naive_dbg.clean restates the same rules on strings with chains found by
walking, so a mistake in the oracle's summary orientation, merging, edge
erasure or loop order shows up as a difference.  Hand-made graphs pin what
the rules must do: a weak tip is cut, a weak bubble path is cut and the
strong path joins its flanks into one unitig, a long tip survives, a strong
bubble survives.
"""
import numpy as np
import pytest

import naive_dbg as nd
from helpers import REF_DATA, pack, random_reads, rc, read_fastq, thresholds_L
from oracle.oracle import Oracle

import os

FIXTURES = ["debruijn.bubble.fastq", "debruijn.bubble2.fastq", "debruijn.palindrome_chr14.fastq",
            "debruijn.shifted_cycle_chr14.fastq", "test.debruijn.small.fastq", "test.star.unitiqs.fastq",
            "debruijn.palindrome.fastq", "natural.fastq"]


def kstr(lo, hi, k):
    v = (int(hi) << 64) | int(lo)
    return "".join("ACGT"[(v >> (2 * (k - 1 - i))) & 3] for i in range(k))


def run_both(reads, k, t):
    th = thresholds_L(t) if t else None
    o = Oracle(k, np.array(th, np.uint64) if th else None)
    b, off = pack(reads)
    o.add_reads(b, off)
    o.run(1)
    pre = (o.chain_fasta(), o.branch_fasta())
    o.run_clean(1)
    g = nd.filtered(reads, k, th) if th else nd.unfiltered(reads, k)
    g2, rounds = nd.clean(g, k, th[2] if th else 65536)
    return o, g, g2, rounds, pre


def check_equal(o, g2, rounds, k):
    assert o.stat("clean_rounds") == rounds
    assert o.clean_output("chain_fasta") == nd.chain_fasta(g2).encode()
    assert o.clean_output("branch_fasta") == nd.branch_fasta(g2).encode()
    assert o.clean_output("branch_edges") == nd.branch_edges(g2).encode()
    got = {kstr(n["kmer_lo"], n["kmer_hi"], k): list(map(int, n["counts"])) for n in o.clean_nodes()}
    assert got == g2


@pytest.mark.parametrize("name", FIXTURES)
@pytest.mark.parametrize("k", [21, 31])
@pytest.mark.parametrize("t", [None, 2, 3])
def test_fixtures_clean_like_string_builder(name, k, t):
    reads, _ = read_fastq(os.path.join(REF_DATA, name))
    o, _, g2, rounds, _ = run_both(reads, k, t)
    check_equal(o, g2, rounds, k)


@pytest.mark.parametrize("seed", range(12))
def test_random_tips_and_bubbles_clean_like_string_builder(seed):
    """Random genomes with low-coverage variant reads (bubbles) and reads
    whose ends carry errors (tips), small k so chains are short."""
    rng = np.random.default_rng(100 + seed)
    k = [11, 13, 15, 21][seed % 4]
    g, reads = random_reads(seed, 400, 120, (40, 80), 0.0)
    for _ in range(6):  # variant copies: a substitution in the middle of a read-sized window
        st = int(rng.integers(0, len(g) - 60))
        s = list(g[st:st + 60])
        s[30] = "ACGT"[("ACGT".index(s[30]) + 1 + int(rng.integers(0, 3))) % 4]
        reads += ["".join(s)] * int(rng.integers(1, 3))
    for _ in range(8):  # tips: errors in the last bases of a read
        st = int(rng.integers(0, len(g) - 50))
        s = list(g[st:st + 50])
        s[-3] = "ACGT"[("ACGT".index(s[-3]) + 1) % 4]
        reads.append("".join(s) if rng.random() < 0.5 else rc("".join(s)))
    for t in (None, 2, 3):
        o, _, g2, rounds, _ = run_both(reads, k, t)
        check_equal(o, g2, rounds, k)


def unitigs_of(fasta):
    lines = fasta.decode().split("\n")
    return sorted(lines[1::2])


def test_weak_bubble_path_is_cut_and_strong_path_joins_flanks():
    """A 300-base genome read 8 times, and a copy with one substitution at
    position 150 read twice.  With -L 1 every edge is kept: the variant
    makes a bubble (two paths of k + 1 nodes between the same branch
    points).  Its edges have count 2 < 1 + 1 is false, so with -L 1 nothing
    is cut; with -L 2 the variant path's end edges (count 2 < 3) are cut and
    the genome is one unitig again."""
    rng = np.random.default_rng(7)
    k = 15
    g = "".join(rng.choice(list("ACGT"), 300))
    v = list(g)
    v[150] = "ACGT"[("ACGT".index(v[150]) + 2) % 4]
    v = "".join(v)
    reads = [g] * 8 + [v] * 2
    o, _, g2, rounds, pre = run_both(reads, k, 2)
    check_equal(o, g2, rounds, k)
    before = unitigs_of(pre[0])
    after = unitigs_of(o.clean_output("chain_fasta"))
    assert len(before) == 4 and rounds >= 1 and o.stat("clean_bubbles") == 2
    canon_g = min(g, rc(g))
    assert canon_g in after  # the strong path and its flanks, one unitig
    assert o.clean_output("branch_fasta") == b""
    # the variant path survives as its own unitig, cut off from the genome
    assert any(v[150 - k + 1:150 + k] in u or rc(v[150 - k + 1:150 + k]) in u for u in after)
    # a strong bubble (both paths read 8 times) is kept
    o2, _, g3, rounds2, pre2 = run_both([g] * 8 + [v] * 8, k, 2)
    check_equal(o2, g3, rounds2, k)
    assert rounds2 == 0 and o2.clean_output("chain_fasta") == pre2[0]


def test_weak_short_tip_is_cut_long_tip_kept():
    """Reads of a genome plus reads that leave it after position 200 with a
    different tail: a tip of 10 bases (k-mers < 2k - 1) read twice is cut
    under -L 2; a tip of 60 bases is not a deadend by length and stays."""
    rng = np.random.default_rng(11)
    k = 15
    g = "".join(rng.choice(list("ACGT"), 400))
    tail = "".join(rng.choice(list("ACGT"), 60))
    short = g[150:200] + tail[:10]
    long_ = g[150:200] + tail
    o, _, g2, rounds, pre = run_both([g] * 6 + [short] * 2, k, 2)
    check_equal(o, g2, rounds, k)
    assert o.stat("clean_deadends") == 1
    after = unitigs_of(o.clean_output("chain_fasta"))
    assert min(g, rc(g)) in after
    o, _, g2, rounds, pre = run_both([g] * 6 + [long_] * 2, k, 2)
    check_equal(o, g2, rounds, k)
    assert rounds == 0 and o.clean_output("chain_fasta") == pre[0]
