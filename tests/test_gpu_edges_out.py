"""The non-benchmark outputs .branch.edges and .chain.edges on the GPU
(bruno_get_branch_nodes, bruno_get_chain_edges, through the C ABI) against
the oracle's bytes: random genomes of odd and even k, thresholds, the region
count, a batch, saturated counts, and the recycling job's refusal."""
import numpy as np
import pytest

import bruno_amd
from bruno_amd.autotune import get_tuning, set_tuning
from gpu_common import run_oracle
from helpers import pack, random_reads, thresholds_L

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
    (13, 30000, 3000, 150, 0.01, 31),
]


def gpu_edges(k, th, b, off, capacity=0, edge_filter=False):
    job = bruno_amd.GraphJob(k, th, capacity=capacity, edge_filter=edge_filter)
    job.parse_count(b, off)
    job.run()
    out = (job.branch_edges_text(), job.chain_edges_text(), job.stats())
    job.close()
    return out


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("tl", [None, 2])
def test_edge_outputs_match_oracle(case, tl):
    seed, G, n, L, err, k = case
    _, reads = random_reads(seed, G, n, L, err)
    b, off = pack(reads)
    th = None if tl is None else thresholds_L(tl)
    be, ce, st = gpu_edges(k, th, b, off)
    o = run_oracle(k, th, b, off)
    assert o.stat("chain_edge_misses") == 0
    assert be == o.branch_edges()
    assert ce == o.chain_edges()
    assert len(ce.splitlines()) == st["n_unitigs"]


def test_saturated_counts():
    """Self counts at 65535 (many copies of a few reads): the sum runs past
    float32's exact integers and the mean is still the oracle's."""
    _, reads = random_reads(5, 400, 6, 120, 0.0)
    reads = reads * 70000
    b, off = pack(reads)
    be, ce, _ = gpu_edges(31, None, b, off)
    o = run_oracle(31, None, b, off)
    assert be == o.branch_edges() and ce == o.chain_edges()
    assert b"65535" in ce


@pytest.fixture
def region_count():
    old = get_tuning()
    set_tuning(count_regions=2)
    yield
    set_tuning(**old)


def test_region_table(region_count):
    """Lookups in a region-layout table."""
    _, reads = random_reads(21, 50000, 3000, 150, 0.02)
    b, off = pack(reads)
    be, ce, st = gpu_edges(31, thresholds_L(2), b, off, capacity=1 << 22)
    assert st["count_by_regions"] == 1
    o = run_oracle(31, thresholds_L(2), b, off)
    assert be == o.branch_edges() and ce == o.chain_edges()


def test_edge_filter_mode():
    _, reads = random_reads(3, 2000, 300, 50, 0.01)
    b, off = pack(reads)
    th = [0, 65536, 2, 65536, 0, 65536]
    be, ce, _ = gpu_edges(11, th, b, off, edge_filter=True)
    o = run_oracle(11, th, b, off, edge_filter=True)
    assert be == o.branch_edges() and ce == o.chain_edges()


def test_batch_jobs():
    jobs = [random_reads(40 + i, G, n, 100, 0.01)[1] for i, (G, n) in enumerate([(3000, 300), (800, 100), (5000, 400)])]
    jobs.insert(1, [])
    reads = [r for j in jobs for r in j]
    b, off = pack(reads)
    first = np.zeros(len(jobs) + 1, np.uint64)
    first[1:] = np.cumsum([len(j) for j in jobs])
    th = thresholds_L(2)
    bt = bruno_amd.GraphBatch(31, th, [max(256, 2 * sum(len(r) for r in j)) for j in jobs])
    bt.parse_count(b, off, first)
    bt.run()
    for i, j in enumerate(jobs):
        jb, jo = pack(j)
        o = run_oracle(31, th, jb, jo)
        v = bt.job(i)
        assert v.branch_edges_text() == o.branch_edges()
        assert v.chain_edges_text() == o.chain_edges()
    bt.close()


def test_recycled_table_is_refused():
    """A job that recycles its table no longer holds the counts after
    compact: the getters report the state error instead of reading it."""
    _, reads = random_reads(3, 2000, 300, 50, 0.01)
    b, off = pack(reads)
    job = bruno_amd.GraphJob(11, thresholds_L(2))
    job.set_recycle(True)
    job.parse_count(b, off)
    job.run()
    with pytest.raises(bruno_amd.BrunoError):
        job.chain_edges()
    with pytest.raises(bruno_amd.BrunoError):
        job.branch_nodes()
    job.close()
