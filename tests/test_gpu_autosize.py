"""Jobs created with capacity 0 size their table from a sketch of the
batches counted so far; a later batch with many more distinct k-mers moves
the table into a larger one (every k-mer and k2 entry with its counts).
The results equal the oracle on all reads, at k = 31 and 63, whichever count
path (atomic or region layout) the old and new tables use."""
import numpy as np
import pytest

import bruno_amd
from bruno_amd.autotune import get_tuning, set_tuning
from gpu_common import compare_results, run_oracle
from helpers import pack, random_reads, thresholds_L

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("k", [31, 63, 21])
@pytest.mark.parametrize("th", [None, thresholds_L(2)])
def test_growing_table_equals_oracle(k, th):
    _, reads = random_reads(40 + k, 40000, 4000, (90, 160), 0.01)
    parts = [reads[:40], reads[40:400], reads[400:]]
    job = bruno_amd.GraphJob(k, th, capacity=0)
    caps = []
    for p in parts:
        b, off = pack(p)
        job.parse_count(b, off)
        caps.append(job.stats()["capacity"])
    assert caps[0] < caps[-1]  # the table grew
    res = {"counts": job.counts(), "k2": job.k2()}
    job.filter()
    res["nodes"] = job.nodes()
    job.build_chains()
    res["rounds"] = job.list_rank(1)
    job.compact()
    res.update(fasta=job.unitigs(), branch_fasta=job.branch_fasta(), chain_nodes=job.chain_nodes(),
               branches=job.branches(), stats=job.stats())
    job.close()
    b, off = pack(reads)
    compare_results(k, res, run_oracle(k, th, b, off, 1), 1)


def test_growth_into_a_region_table():
    """Count regions allowed from 2^22 slots: a first batch sizes a small
    atomic-layout table, a large second batch needs more than 2^22 slots, so
    the table moves into the region layout and later counts use regions."""
    knobs = get_tuning()
    try:
        set_tuning(count_regions=2)
        cfg = bruno_amd.SynthConfig(seed=5, genome_len=3_000_000, n_reads=400_000, read_len=150, sub_rate=0.01)
        b, off = bruno_amd.synth_reads_host(cfg)
        cut = [0, 2000, 300_000, 400_000]
        job = bruno_amd.GraphJob(31, thresholds_L(3), capacity=0)
        caps = []
        for a, z in zip(cut[:-1], cut[1:]):
            job.parse_count(b[int(off[a]):int(off[z])], off[a:z + 1] - off[a])
            caps.append(job.stats()["capacity"])
        st = job.stats()
        assert caps[0] < (1 << 22) <= caps[-1]
        assert st["count_by_regions"] == 1  # the last batch counted by regions
        job.run(1)
        got = (job.unitigs(), job.branch_fasta())
        job.close()
        o = run_oracle(31, thresholds_L(3), b, off, 1, nthreads=8)
        assert got == (o.chain_fasta(), o.branch_fasta())
    finally:
        set_tuning(**knobs)
