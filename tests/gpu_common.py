"""Helpers for the GPU parity tests: run the HIP path and the oracle on the
same reads and compare every parity object."""
import numpy as np

import bruno_amd
from oracle.oracle import Oracle


def kv(words, k):
    """k-mer words (lo, hi) -> python int."""
    return int(words[0]) | (int(words[1]) << 64 if len(words) > 1 else 0)


def run_gpu(k, th, bases, offs, variant=1, capacity=0, stages=True, edge_filter=False, fused=False):
    job = bruno_amd.GraphJob(k, th, capacity=capacity, edge_filter=edge_filter)
    if fused:
        job.set_fused_filter(True)
    job.parse_count(bases, offs)
    res = {}
    if stages and not fused:
        res["counts"] = job.counts()
        res["k2"] = job.k2()
    job.filter()
    res["nodes"] = job.nodes()
    job.build_chains()
    res["rounds"] = job.list_rank(variant)
    job.compact()
    res["fasta"] = job.unitigs()
    res["branch_fasta"] = job.branch_fasta()
    res["chain_nodes"] = job.chain_nodes()
    res["branches"] = job.branches()
    res["stats"] = job.stats()
    job.close()
    return res


def run_oracle(k, th, bases, offs, variant=1, nthreads=1, edge_filter=False):
    o = Oracle(k, None if th is None else np.array(th, np.uint64), nthreads, edge_filter=edge_filter)
    o.add_reads(bases, offs)
    o.run(variant)
    return o


def node_table(arr):
    """{kmer int: tuple(counts)} from NODE_DTYPE (gpu) records."""
    return {(int(r["kmer"][0]) | (int(r["kmer"][1]) << 64)): tuple(int(x) for x in r["counts"]) for r in arr}


def oracle_node_table(arr):
    return {(int(r["kmer_lo"]) | (int(r["kmer_hi"]) << 64)): tuple(int(x) for x in r["counts"]) for r in arr}


def compare_all(k, th, bases, offs, variant=1, nthreads=1, stages=True, capacity=0, edge_filter=False, fused=False):
    stages = stages and not fused  # a fused count leaves node records, not k2 entries
    g = run_gpu(k, th, bases, offs, variant, capacity=capacity, stages=stages, edge_filter=edge_filter, fused=fused)
    o = run_oracle(k, th, bases, offs, variant, nthreads, edge_filter=edge_filter)
    compare_results(k, g, o, variant, stages)
    return g, o


def compare_results(k, g, o, variant=1, stages=True):
    """Every parity object of one job's GPU results `g` against the oracle."""
    st = g["stats"]
    if variant == 1:
        # list_rank_min_update only sends to ends that still point to a chain
        # node: on graphs built from reads no end gets two updates per round
        assert st["n_conflicts"] == 0
    assert st["n_instances"] == o.stat("n_instances")
    if stages:
        # P1: unthresholded counters
        assert node_table(g["counts"]) == oracle_node_table(o.unfiltered())
        # k2 multiset
        km, e, c = g["k2"]
        got = sorted(((int(km[i][0]) | (int(km[i][1]) << 64 if km.shape[1] > 1 else 0)), int(e[i]), int(c[i]))
                     for i in range(len(e)))
        lo, hi, oe, oc = o.k2()
        exp = sorted(((int(lo[i]) | (int(hi[i]) << 64)), int(oe[i]), int(oc[i])) for i in range(len(oe)))
        assert got == exp
        assert st["n_k2"] == len(exp)
    # P2
    assert node_table(g["nodes"]) == oracle_node_table(o.nodes())
    # P3
    assert g["branch_fasta"] == o.branch_fasta()
    assert st["n_branches"] == o.stat("n_branches")
    # P4 (sorted by rep, |dist|)
    on = o.chain_nodes()
    gn = g["chain_nodes"]
    assert len(gn) == len(on)
    if len(on):
        assert np.array_equal(gn["k"][:, 0], on["k_lo"])
        assert np.array_equal(gn["rep"][:, 0], on["rep_lo"])
        assert np.array_equal(gn["dist"], on["dist"])
        if k > 31:
            assert np.array_equal(gn["k"][:, 1], on["k_hi"])
            assert np.array_equal(gn["rep"][:, 1], on["rep_hi"])
    # P5
    assert g["fasta"] == o.chain_fasta()
    assert g["rounds"] == o.stat("iterations")
    assert st["n_cycle_nodes"] == o.stat("n_cycle_nodes")
    assert st["n_isolated"] == o.stat("n_isolated")
