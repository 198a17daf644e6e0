"""Counting by table region (bucketed instance records, per-region builds in
LDS) against the oracle, and against the atomic count on the same inputs,
for one-word (k <= 31) and two-word (k <= 63) keys.
count_regions 2 gives tables of >= 2^22 slots of plain k <= 31 jobs the
region layout (the default bar is the same, 2^22 slots)."""
import numpy as np
import pytest

import bruno_amd
from bruno_amd.autotune import get_tuning, set_tuning
from gpu_common import compare_all, compare_results, node_table, run_oracle
from helpers import pack, random_reads, thresholds_L

pytestmark = pytest.mark.gpu

REGION_CAP = 1 << 22

CASES = [
    (1, 300, 40, 30, 0.0, 7),
    (3, 2000, 300, 50, 0.01, 11),
    (6, 4000, 400, 100, 0.01, 31),
    (7, 600, 120, (40, 300), 0.01, 31),
    (11, 5000, 500, (60, 200), 0.02, 25),
    (13, 30000, 3000, 150, 0.01, 31),
    # two-word keys: 64-byte slots, 16-byte records carrying the k-mer's high word
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


@pytest.fixture
def direct_count():
    set_tuning(count_regions=0)
    yield


def _reads(case):
    seed, G, n, L, err, k = case
    _, reads = random_reads(seed, G, n, L, err)
    return k, reads


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("tl", [None, 2])
def test_region_count_matches_oracle(case, tl):
    k, reads = _reads(case)
    b, off = pack(reads)
    g, _ = compare_all(k, None if tl is None else thresholds_L(tl), b, off, capacity=REGION_CAP)
    assert g["stats"]["count_by_regions"] == 1
    assert g["stats"]["capacity"] >= REGION_CAP


@pytest.mark.parametrize("case", CASES[2:4])
def test_atomic_count_on_large_table(case, direct_count):
    """The knob off: one probe window, atomic counting, same results."""
    k, reads = _reads(case)
    b, off = pack(reads)
    g, _ = compare_all(k, thresholds_L(2), b, off, capacity=REGION_CAP)
    assert g["stats"]["count_by_regions"] == 0


def test_counts_accumulate_over_calls():
    """Three parse_count calls: later calls merge into regions the first
    call wrote."""
    k, reads = _reads(CASES[5])
    job = bruno_amd.GraphJob(k, thresholds_L(2), capacity=REGION_CAP)
    parts = np.array_split(np.arange(len(reads)), 3)
    for idx in parts:
        pb, po = pack([reads[i] for i in idx])
        job.parse_count(pb, po)
        assert job.stats()["count_by_regions"] == 1
    b, off = pack(reads)
    o = run_oracle(k, thresholds_L(2), b, off)
    g = {"counts": job.counts(), "k2": job.k2()}
    job.filter()
    g["nodes"] = job.nodes()
    job.build_chains()
    g["rounds"] = job.list_rank(1)
    job.compact()
    g.update(fasta=job.unitigs(), branch_fasta=job.branch_fasta(), chain_nodes=job.chain_nodes(),
             branches=job.branches(), stats=job.stats())
    job.close()
    compare_results(k, g, o, 1, True)


def test_skewed_input_falls_back_to_atomic_count():
    """One k-mer in most instances overfills its record bucket: the call
    is counted with atomics instead, with the same results."""
    rng = np.random.default_rng(9)
    reads = ["A" * 200] * 800 + ["".join(rng.choice(list("ACGT"), 150)) for _ in range(200)]
    b, off = pack(reads)
    g, _ = compare_all(31, None, b, off, capacity=REGION_CAP)
    assert g["stats"]["count_by_regions"] == 0
    g, _ = compare_all(31, thresholds_L(1), b, off, capacity=REGION_CAP)


def test_edge_reads():
    """Empty, short, N-containing reads (the N within the first, second and
    later 64-base chunks) and one 20 kbp read among 3000 short ones (the
    long read's records overflow the workgroup's staging and are appended
    one by one)."""
    long_read = "".join(np.random.default_rng(3).choice(list("ACGT"), 20000))
    _, short = random_reads(17, 40000, 3000, 120, 0.01)
    with_n = [long_read[100:500], long_read[600:1000], long_read[1100:1300]]
    with_n = [with_n[0][:300] + "N" + with_n[0][301:], with_n[1][:150] + "N" + with_n[1][151:],
              with_n[2][:191] + "N" + with_n[2][192:]]
    reads = ["", "ACGT" * 5, "A" * 33, "ACGTN" * 10, "C" * 31 + "A", long_read, long_read[5000:9000],
             long_read[12000:15000][::-1].translate(str.maketrans("ACGT", "TGCA"))] + with_n + short
    b, off = pack(reads)
    for th in (None, thresholds_L(1)):
        g, _ = compare_all(31, th, b, off, capacity=REGION_CAP)
        assert g["stats"]["count_by_regions"] == 1


def test_empty_and_reused_job():
    """A job that counts nothing, then reset and reused with reads, then
    reset with recycling on: each run equals a fresh job's."""
    job = bruno_amd.GraphJob(31, thresholds_L(2), capacity=REGION_CAP)
    job.parse_count(b"", np.zeros(1, np.uint64))
    job.run()
    assert job.unitigs() == b"" and job.branch_fasta() == b""
    for case, recycle in ((CASES[2], False), (CASES[4], True), (CASES[5], True)):
        _, reads = _reads(case)
        b, off = pack(reads)
        job.reset()
        job.set_recycle(recycle)
        job.parse_count(b, off)
        job.run()
        fresh = bruno_amd.GraphJob(31, thresholds_L(2), capacity=REGION_CAP)
        fresh.parse_count(b, off)
        fresh.run()
        assert job.unitigs() == fresh.unitigs()
        assert job.branch_fasta() == fresh.branch_fasta()
        assert job.stats()["count_by_regions"] == 1
        fresh.close()
    job.close()


@pytest.mark.parametrize("k", [15, 31, 45, 63])
def test_region_vs_atomic_tables_agree(k, direct_count):
    """Counts, k2 entries and every output of the region build equal the
    atomic count's on a bigger random genome (side-table spills included)."""
    _, reads = random_reads(21, 200000, 6000, 150, 0.02)
    b, off = pack(reads)
    outs = []
    for regions in (2, 0):
        set_tuning(count_regions=regions)
        job = bruno_amd.GraphJob(k, thresholds_L(2), capacity=REGION_CAP)
        job.parse_count(b, off)
        st = job.stats()
        assert st["count_by_regions"] == (regions > 0)
        c = job.counts()
        kk = job.k2()
        job.run()
        k2 = sorted(zip(map(tuple, kk[0].tolist()), kk[1].tolist(), kk[2].tolist()))
        outs.append((node_table(c), k2, job.unitigs(), job.branch_fasta(), st["n_distinct"], st["n_k2"],
                     st["n_instances"]))
        job.close()
    assert outs[0] == outs[1]


def test_long_reads_count_by_regions():
    """Reads of 60 to 3000 bases in one workgroup: the parser walks to the
    longest read of its workgroup, appending every few positions, with the
    same results."""
    _, reads = random_reads(19, 30000, 60, 3000, 0.01)
    b, off = pack(reads)
    g, _ = compare_all(31, thresholds_L(2), b, off, capacity=REGION_CAP)
    assert g["stats"]["count_by_regions"] == 1


@pytest.mark.parametrize("k", [31, 63])
def test_spilled_entries_in_regions(k):
    """k-mers with more distinct biedges than a slot holds: the region build
    sends the extra entries to the side table, keyed by table slot."""
    rng = np.random.default_rng(5)
    reads = []
    for _ in range(40):
        core = "".join(rng.choice(list("ACGT"), k))
        for a in "ACGT":
            for c in "ACGT":
                r = "".join(rng.choice(list("ACGT"), 10)) + a + core + c + "".join(rng.choice(list("ACGT"), 10))
                reads += [r] * 5 + [r.translate(str.maketrans("ACGT", "TGCA"))[::-1]] * 2
    order = rng.permutation(len(reads))
    reads = [reads[i] for i in order]
    b, off = pack(reads)
    for th in (None, thresholds_L(2)):
        g, _ = compare_all(k, th, b, off, capacity=REGION_CAP)
        assert g["stats"]["count_by_regions"] == 1
        assert g["stats"]["n_overflow"] >= 40


def spill_reads(k, seed=5):
    rng = np.random.default_rng(seed)
    reads = []
    for _ in range(40):
        core = "".join(rng.choice(list("ACGT"), k))
        for a in "ACGT":
            for c in "ACGT":
                r = "".join(rng.choice(list("ACGT"), 10)) + a + core + c + "".join(rng.choice(list("ACGT"), 10))
                reads += [r] * 5 + [r.translate(str.maketrans("ACGT", "TGCA"))[::-1]] * 2
    return [reads[i] for i in rng.permutation(len(reads))]


@pytest.mark.parametrize("case", [CASES[3], CASES[5], CASES[7], CASES[8]])
@pytest.mark.parametrize("region_slots,capacity", [(1280, 1 << 24), (2560, 1 << 25)])
def test_pipelined_region_builds(case, region_slots, capacity):
    """2^14 regions and more, 16 pipeline groups (count_groups; the
    automatic choice pipelines only the largest tables): the
    second pass splits the sub-buckets in groups, and each group's region
    builds run on a second stream while the next group is split (small regions: the 256-thread build; large
    ones: the 512-thread build).  Same results, spilled slots included."""
    old = get_tuning()
    set_tuning(region_slots=region_slots, count_groups=16)
    try:
        k, reads = _reads(case)
        for rs in (reads, reads + spill_reads(k)):
            b, off = pack(rs)
            for th in (None, thresholds_L(2)):
                g, _ = compare_all(k, th, b, off, capacity=capacity)
                assert g["stats"]["count_by_regions"] == 1
    finally:
        set_tuning(**old)


@pytest.mark.parametrize("case", [CASES[5], CASES[7]])
def test_one_job_batch_counts_by_regions(case):
    """A batch of one job is a plain job: its large table takes the region
    layout and counts by regions, with the results of the job run alone."""
    k, reads = _reads(case)
    b, off = pack(reads)
    th = thresholds_L(2)
    bt = bruno_amd.GraphBatch(k, th, [REGION_CAP])
    bt.parse_count(b, off, [0, len(reads)])
    bt.run(1)
    v = bt.job(0)
    st, u, br = v.stats(), v.unitigs(), v.branch_fasta()
    bt.close()
    assert st["count_by_regions"] == 1
    job = bruno_amd.GraphJob(k, th, capacity=REGION_CAP)
    job.parse_count(b, off)
    job.run(1)
    assert (u, br) == (job.unitigs(), job.branch_fasta())
    assert st["n_unitigs"] == job.stats()["n_unitigs"]
    job.close()


@pytest.mark.parametrize("groups", [1, 3, 64])
def test_pipeline_group_counts(groups):
    """Any number of pipeline groups (one, uneven groups, one sub-bucket
    per group) gives the same results."""
    old = get_tuning()
    set_tuning(region_slots=1280, count_groups=groups)
    try:
        k, reads = _reads(CASES[5])
        b, off = pack(reads + spill_reads(k))
        g, _ = compare_all(k, thresholds_L(2), b, off, capacity=1 << 24)
        assert g["stats"]["count_by_regions"] == 1
    finally:
        set_tuning(**old)
