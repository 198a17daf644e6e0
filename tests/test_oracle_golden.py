"""Pin the CPU oracle on the reference's own known-answer fixtures.

* test.unitiqs.fastq -> test.unitiqs.fasta: the reference's regression script
  diffs the single-process output against this file
  (test/test/test_compact_debruijn_graph_all.sh:157-158); sequence lines must
  match exactly and in order.
* test.short.unitiqs.fastq: six disjoint reads whose quality lines hold the
  expected unitig in canonical orientation.
"""
import os

from helpers import REF_DATA, pack, read_fastq, thresholds_L
from oracle.oracle import Oracle


def run_oracle(seqs, k=31, th=None, variant=1, nthreads=1):
    o = Oracle(k, th, nthreads)
    b, off = pack(seqs)
    o.add_reads(b, off)
    o.run(variant)
    return o


def test_unitiqs_fasta_known_answer():
    seqs, _ = read_fastq(os.path.join(REF_DATA, "test.unitiqs.fastq"))
    with open(os.path.join(REF_DATA, "test.unitiqs.fasta")) as f:
        exp = [l for l in f.read().split("\n")[1::2] if l]
    for variant in (0, 1):
        o = run_oracle(seqs, variant=variant)
        got = o.chain_fasta().decode().split("\n")[1::2]
        assert got == exp
        assert o.stat("n_branches") == 0


def test_short_unitiqs_quality_lines():
    seqs, quals = read_fastq(os.path.join(REF_DATA, "test.short.unitiqs.fastq"))
    o = run_oracle(seqs)
    fa = o.chain_fasta().decode().split("\n")
    got = fa[1::2]
    assert sorted(got) == sorted(quals)
    # output order is by representative k-mer: sequence lines start with rep
    reps = [s[:31] for s in got]
    assert reps == sorted(reps)
    # header is the canonical form of the representative
    from naive_dbg import canon
    assert [h[1:] for h in fa[0::2]] == [canon(r) for r in reps]


def test_unitiqs_thresholded_l1_same_as_unthresholded():
    """-T -L 1 keeps every edge seen once: the reference script compares
    tl1 outputs against the unthresholded gold (test_compact_debruijn_graph_all.sh:204-222)."""
    for name in ("test.unitiqs.fastq", "test.short.unitiqs.fastq", "test.star.unitiqs.fastq",
                 "test.star2.unitiqs.fastq"):
        seqs, _ = read_fastq(os.path.join(REF_DATA, name))
        a = run_oracle(seqs)
        b = run_oracle(seqs, th=thresholds_L(1))
        assert a.chain_fasta() == b.chain_fasta(), name
        assert a.branch_fasta() == b.branch_fasta(), name


def test_list_rank_variants_agree_on_fixtures():
    for name in sorted(os.listdir(REF_DATA)):
        if not name.endswith(".fastq"):
            continue
        seqs, _ = read_fastq(os.path.join(REF_DATA, name))
        for th in (None, thresholds_L(2)):
            a = run_oracle(seqs, th=th, variant=0)
            b = run_oracle(seqs, th=th, variant=1)
            assert a.chain_fasta() == b.chain_fasta(), name
            assert a.chain_components() == b.chain_components(), name
