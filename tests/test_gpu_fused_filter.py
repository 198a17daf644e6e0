"""The filter fused into the region count (bruno_job_set_fused_filter): the
region build runs local_insert_by_freqencies and the branch / chain
classification on every candidate slot in LDS and appends the nodes;
bruno_filter closes the node lists' holes and finishes the slots with
side-table entries.  Every parity object must equal the oracle's, for
one- and two-word k-mers, thresholded, unthresholded and edge-filtered
jobs, node lists that run out of room (the rest left to bruno_filter),
spilled slots, cleaning, and the state rules that come with it.
count_regions 2 gives tables of >= 2^22 slots the region layout."""
import numpy as np
import pytest

import bruno_amd
from bruno_amd.autotune import get_tuning, set_tuning
from gpu_common import compare_all, run_gpu
from helpers import DEFAULT_T, pack, random_reads, thresholds_L
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

REGION_CAP = 1 << 22

CASES = [
    (1, 300, 40, 30, 0.0, 7),
    (3, 2000, 300, 50, 0.01, 11),
    # k = 5 on a short genome: most k-mers carry more biedges than a slot holds
    (4, 200, 200, 25, 0.02, 5),
    (6, 4000, 400, 100, 0.01, 31),
    (7, 600, 120, (40, 300), 0.01, 31),
    (13, 30000, 3000, 150, 0.01, 31),
    (21, 6000, 500, (80, 250), 0.01, 63),
    (22, 30000, 2000, 250, 0.01, 47),
    (23, 4000, 400, 120, 0.02, 32),
]


@pytest.fixture(autouse=True)
def small_region_tables():
    old = get_tuning()
    set_tuning(count_regions=2)
    yield
    set_tuning(**old)


def _reads(case):
    seed, G, n, L, err, k = case
    _, reads = random_reads(seed, G, n, L, err)
    return k, reads


def edge_thresholds(lo):
    return [0, DEFAULT_T, lo, DEFAULT_T, 0, DEFAULT_T]


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("tl", [None, 2, 3])
def test_fused_matches_oracle(case, tl):
    k, reads = _reads(case)
    b, off = pack(reads)
    g, o = compare_all(k, None if tl is None else thresholds_L(tl), b, off, capacity=REGION_CAP, fused=True)
    st = g["stats"]
    assert st["count_by_regions"] == 1 and st["filter_fused"] == 1
    assert st["n_nodes"] == o.stat("n_nodes")


@pytest.mark.parametrize("case", CASES[1:7])
@pytest.mark.parametrize("lo", [2, 3])
def test_fused_edge_filter(case, lo):
    k, reads = _reads(case)
    b, off = pack(reads)
    g, _ = compare_all(k, edge_thresholds(lo), b, off, capacity=REGION_CAP, edge_filter=True, fused=True)
    assert g["stats"]["filter_fused"] == 1


@pytest.mark.parametrize("case", [CASES[2], CASES[5], CASES[6]])
@pytest.mark.parametrize("list_cap", [1, 100, 1000])
def test_fused_lists_run_out(case, list_cap):
    """Node lists with room for list_cap nodes: the build places what fits,
    bruno_filter grows the lists and places the rest; same results."""
    old = get_tuning()
    set_tuning(fuse_list_cap=list_cap)
    try:
        k, reads = _reads(case)
        b, off = pack(reads)
        for tl in (None, 2):
            g, o = compare_all(k, None if tl is None else thresholds_L(tl), b, off, capacity=REGION_CAP, fused=True)
            assert g["stats"]["filter_fused"] == 1
            assert g["stats"]["n_nodes"] == o.stat("n_nodes")
    finally:
        set_tuning(**old)


@pytest.mark.parametrize("case", [CASES[3], CASES[6]])
def test_fused_equals_unfused(case):
    """The same job fused and unfused: the same statistics and outputs."""
    k, reads = _reads(case)
    b, off = pack(reads)
    th = thresholds_L(2)
    a = run_gpu(k, th, b, off, capacity=REGION_CAP, stages=False, fused=True)
    c = run_gpu(k, th, b, off, capacity=REGION_CAP, stages=False, fused=False)
    assert a["stats"]["filter_fused"] == 1 and c["stats"]["filter_fused"] == 0
    for key in ("n_distinct", "n_k2", "n_overflow", "n_nodes", "n_branches", "n_chain_nodes", "n_unitigs"):
        assert a["stats"][key] == c["stats"][key], key
    for key in ("fasta", "branch_fasta"):
        assert a[key] == c[key]


def test_fused_state_rules():
    k, reads = _reads(CASES[3])
    b, off = pack(reads)
    job = bruno_amd.GraphJob(k, thresholds_L(2), capacity=REGION_CAP)
    job.set_fused_filter(True)
    job.parse_count(b, off)
    assert job.stats()["filter_fused"] == 1
    for call in (lambda: job.parse_count(b, off), job.k2, job.counts):
        with pytest.raises(bruno_amd.BrunoError) as e:
            call()
        assert e.value.code == 5  # BRUNO_STATE
    job.run(1)
    first = job.unitigs()
    # after a reset the job counts again (fused again), and two batches
    # without fusion merge as before
    job.reset()
    job.parse_count(b, off)
    job.run(1)
    assert job.unitigs() == first
    job.reset()
    job.set_fused_filter(False)
    half = len(reads) // 2
    b1, o1 = pack(reads[:half])
    b2, o2 = pack(reads[half:])
    job.parse_count(b1, o1)
    job.parse_count(b2, o2)
    assert job.stats()["filter_fused"] == 0
    job.run(1)
    assert job.unitigs() == first
    job.close()


def test_fused_then_clean():
    """Cleaning works on the node lists a fused count left."""
    k, reads = _reads(CASES[5])
    b, off = pack(reads)
    th = thresholds_L(2)
    job = bruno_amd.GraphJob(k, th, capacity=REGION_CAP)
    job.set_fused_filter(True)
    job.parse_count(b, off)
    job.run(1)
    st = job.clean(1, th[2])
    fasta = job.unitigs()
    job.close()
    o = Oracle(k, np.array(th, np.uint64))
    o.add_reads(b, off)
    o.run(1)
    o.run_clean(1)
    assert st["rounds"] == o.stat("clean_rounds")
    assert fasta == o.clean_output("chain_fasta")
