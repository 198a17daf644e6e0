"""HIP path vs the CPU oracle, bit-exact on every parity object.

P0 records, P1 counters, the k2 multiset, P2 filtered graph, P3 branch
FASTA, P4 listranked chain nodes, P5 unitig FASTA, plus iteration, cycle and
isolated counts.  All calls go through the C ABI.
"""
import os

import numpy as np
import pytest

import bruno_amd
from bruno_amd.configs import config1, config2, config3
from gpu_common import compare_all, run_gpu, run_oracle
from helpers import REF_DATA, pack, random_reads, read_fastq, thresholds_L

pytestmark = pytest.mark.gpu

FIXTURES = sorted(f for f in os.listdir(REF_DATA) if f.endswith(".fastq"))


@pytest.mark.parametrize("name", FIXTURES)
@pytest.mark.parametrize("tl", [None, 1, 2])
def test_reference_fixtures(name, tl):
    seqs, _ = read_fastq(os.path.join(REF_DATA, name))
    b, off = pack(seqs)
    th = None if tl is None else thresholds_L(tl)
    compare_all(31, th, b, off)


def test_unitiqs_known_answer_on_gpu():
    seqs, _ = read_fastq(os.path.join(REF_DATA, "test.unitiqs.fastq"))
    b, off = pack(seqs)
    g = run_gpu(31, None, b, off)
    with open(os.path.join(REF_DATA, "test.unitiqs.fasta")) as f:
        exp = [l for l in f.read().split("\n")[1::2] if l]
    assert g["fasta"].decode().split("\n")[1::2] == exp


CASES = [
    (1, 300, 40, 30, 0.0, 7),
    (2, 500, 80, (20, 60), 0.01, 9),
    (3, 2000, 300, 50, 0.01, 11),
    (4, 200, 200, 25, 0.02, 5),
    (5, 1000, 200, 40, 0.005, 15),
    (6, 4000, 400, 100, 0.01, 31),
    (7, 600, 120, (40, 300), 0.01, 31),
    (8, 300, 60, 40, 0.01, 8),
    (9, 3000, 300, 120, 0.01, 63),
    (10, 150, 150, 30, 0.0, 6),
    (11, 5000, 500, (60, 200), 0.02, 33),
    (12, 400, 400, 70, 0.0, 47),
]


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("tl", [None, 2, 3])
@pytest.mark.parametrize("variant", [0, 1])
def test_random_small(case, tl, variant):
    seed, G, n, L, err, k = case
    _, reads = random_reads(seed, G, n, L, err)
    b, off = pack(reads)
    compare_all(k, None if tl is None else thresholds_L(tl), b, off, variant=variant)


@pytest.mark.parametrize("case", [CASES[2], CASES[5], CASES[8], CASES[10]])
def test_output_order_paths(case, monkeypatch):
    """Unitig order placed from the chain offsets (after the contraction
    ranking) and by the full sort (after round emulation): both equal the
    oracle's sorted output."""
    seed, G, n, L, err, k = case
    _, reads = random_reads(seed, G, n, L, err)
    b, off = pack(reads)
    g, _ = compare_all(k, None, b, off, variant=1, stages=False)
    assert g["stats"]["n_unitig_nodes"] > 0
    assert g["stats"]["output_by_chains"] == 1
    monkeypatch.setenv("BRUNO_LR_ROUNDS", "1")
    g2, _ = compare_all(k, None, b, off, variant=1, stages=False)
    assert g2["stats"]["output_by_chains"] == 0
    assert g2["fasta"] == g["fasta"]


@pytest.mark.parametrize("case", CASES[:8])
@pytest.mark.parametrize("variant", [0, 1])
def test_random_small_round_emulation(case, variant, monkeypatch):
    """The round-by-round list ranking (used when a chain graph is not
    well-formed) on the same inputs: same outputs."""
    monkeypatch.setenv("BRUNO_LR_ROUNDS", "1")
    seed, G, n, L, err, k = case
    _, reads = random_reads(seed, G, n, L, err)
    b, off = pack(reads)
    for tl in (None, 2):
        compare_all(k, None if tl is None else thresholds_L(tl), b, off, variant=variant)


def test_long_chains_and_cycles():
    """Single long paths (ceil(log2 L) rounds) and circular genomes (cycles
    that are removed), alone and mixed."""
    rng = np.random.default_rng(77)
    reads = []
    for L in (3, 4, 5, 17, 1000, 4097, 20001):
        reads.append("".join(rng.choice(list("ACGT"), L + 30)))
    g = "".join(rng.choice(list("ACGT"), 3000))
    circ = g + g[:200]
    cyc_reads = [circ[i:i + 150] for i in range(0, 3000, 20)]
    for batch in (reads, cyc_reads, reads + cyc_reads):
        b, off = pack(batch)
        for variant in (0, 1):
            compare_all(31, None, b, off, variant=variant)


def test_edge_cases():
    k = 31
    long_read = "".join(np.random.default_rng(3).choice(list("ACGT"), 20000))
    reads = [
        "",                                    # empty read
        "ACGT" * 5,                            # shorter than k
        "A" * 33,                              # exactly k+2, homopolymer
        "ACGTN" * 10,                          # contains N: dropped
        "acgtacgtacgtacgtacgtacgtacgtacgtacgtac",  # lowercase
        long_read,                              # one long read
        long_read[5000:9000],                   # overlapping piece, other strand below
        long_read[12000:15000][::-1].translate(str.maketrans("ACGT", "TGCA")),
        "C" * 31 + "A",                         # k+1: no records
    ]
    b, off = pack(reads)
    for th in (None, thresholds_L(1), thresholds_L(2)):
        compare_all(k, th, b, off)


def test_empty_job():
    job = bruno_amd.GraphJob(31, thresholds_L(3), capacity=1024)
    job.parse_count(b"", np.zeros(1, np.uint64))
    job.run()
    assert job.unitigs() == b""
    assert job.branch_fasta() == b""
    st = job.stats()
    assert st["n_instances"] == 0 and st["n_unitigs"] == 0
    job.close()


def test_counts_saturate_at_65535():
    """70000 copies of one read: node counters cap at 65535
    (sat_incr / sat_add, debruijn_graph_node.hpp:81-91); k2 counts do not."""
    read = "".join(np.random.default_rng(9).choice(list("ACGT"), 40))
    reads = [read] * 70000
    b, off = pack(reads)
    for th in (None, thresholds_L(3)):
        compare_all(31, th, b, off)


def test_incremental_batches_equal_single_batch():
    """parse_count called per file (the incremental loop,
    compact_dbg_freq_construct_clean.cpp:313-330) equals one big batch."""
    _, reads = random_reads(21, 20000, 4000, 100, 0.01)
    th = thresholds_L(2)
    b, off = pack(reads)
    one = run_gpu(31, th, b, off, stages=False)
    job = bruno_amd.GraphJob(31, th, capacity=200000)
    for part in (reads[:1000], reads[1000:2500], reads[2500:]):
        pb, po = pack(part)
        job.parse_count(pb, po)
    job.run()
    assert job.unitigs() == one["fasta"]
    assert job.branch_fasta() == one["branch_fasta"]
    job.close()


def test_parse_records_p0():
    """P0: the per-instance canonical (k-mer, biedge) records in read order."""
    import ctypes
    from bruno_amd.device import DeviceBuffer
    from oracle.oracle import Oracle
    _, reads = random_reads(31, 3000, 300, (30, 150), 0.01)
    reads.append("ACGTNACGT" * 10)
    L = bruno_amd.lib()
    for k in (11, 31, 63):
        b, off = pack(reads)
        db = DeviceBuffer.from_numpy(np.frombuffer(b, np.uint8))
        do = DeviceBuffer.from_numpy(off)
        cnt = ctypes.c_uint64(0)
        assert L.bruno_parse_records(k, db.ptr, do.ptr, len(reads), None, None, ctypes.byref(cnt), None) == 0
        W = 1 if k <= 31 else 2
        km = DeviceBuffer(max(1, cnt.value * W * 8))
        be = DeviceBuffer(max(1, cnt.value))
        assert L.bruno_parse_records(k, db.ptr, do.ptr, len(reads), km.ptr, be.ptr, ctypes.byref(cnt), None) == 0
        kmh = km.to_numpy(np.uint64, cnt.value * W).reshape(-1, W)
        beh = be.to_numpy(np.uint8, cnt.value)
        got = sorted(((int(x[0]) | (int(x[1]) << 64 if W == 2 else 0)), int(e)) for x, e in zip(kmh, beh))
        o = Oracle(k, None, 1)
        o.add_reads(b, off)
        o.run(1)
        lo, hi, e = o.records()
        exp = sorted(((int(lo[i]) | (int(hi[i]) << 64)), int(e[i])) for i in range(len(e)))
        assert got == exp


def test_device_synth_matches_host():
    from bruno_amd.device import DeviceBuffer
    from bruno_amd.pipeline import synth_reads_device
    cfg = config2(seed=5, scale=0.001)["synth"]
    b, o = bruno_amd.synth_reads_host(cfg)
    db = DeviceBuffer(cfg.n_bases)
    do = DeviceBuffer((cfg.n_reads + 1) * 8)
    synth_reads_device(cfg, db.ptr, do.ptr)
    assert np.array_equal(db.to_numpy(np.uint8, cfg.n_bases), b)
    assert np.array_equal(do.to_numpy(np.uint64, cfg.n_reads + 1), o)


def test_device_batches_equal_host_batches():
    from bruno_amd.device import DeviceBuffer
    cfg = config1(seed=17, scale=0.05)
    b, off = bruno_amd.synth_reads_host(cfg["synth"])
    db = DeviceBuffer.from_numpy(b)
    do = DeviceBuffer.from_numpy(off)
    j1 = bruno_amd.GraphJob(31, cfg["thresholds"])
    j1.parse_count_device(db.ptr, do.ptr, len(off) - 1)
    j1.run()
    j2 = bruno_amd.GraphJob(31, cfg["thresholds"])
    j2.parse_count(b, off)
    j2.run()
    assert j1.unitigs() == j2.unitigs() and j1.branch_fasta() == j2.branch_fasta()
    # reuse of one job's table after reset gives the same result
    j1.reset()
    j1.parse_count_device(db.ptr, do.ptr, len(off) - 1)
    j1.run()
    assert j1.unitigs() == j2.unitigs()


@pytest.mark.parametrize("mk", [lambda: config1(seed=11, scale=0.2), lambda: config2(seed=12, scale=0.002),
                                lambda: config3(seed=13, scale=0.01)])
def test_configs_scaled(mk):
    cfg = mk()
    b, off = bruno_amd.synth_reads_host(cfg["synth"])
    compare_all(cfg["k"], cfg["thresholds"], b, off, nthreads=4)


def test_config1_full():
    cfg = config1(seed=1)
    b, off = bruno_amd.synth_reads_host(cfg["synth"])
    g, o = compare_all(cfg["k"], cfg["thresholds"], b, off, nthreads=8, stages=False)
    assert g["stats"]["n_instances"] == 300000 * 70


def test_spilled_k2_entries_under_contention():
    """Many k-mers with more distinct biedges than fit in their slot, hit
    concurrently: every spilled (k-mer, biedge) pair must be counted.  A read
    set with one centre k-mer flanked by all 16 (in, out) base pairs,
    repeated, plus the reverse strand."""
    rng = np.random.default_rng(5)
    core = "".join(rng.choice(list("ACGT"), 31))
    reads = []
    for a in "ACGT":
        for b in "ACGT":
            r = "".join(rng.choice(list("ACGT"), 10)) + a + core + b + "".join(rng.choice(list("ACGT"), 10))
            reads += [r] * 37
            reads += [r.translate(str.maketrans("ACGT", "TGCA"))[::-1]] * 11
    order = rng.permutation(len(reads))
    reads = [reads[i] for i in order]
    b, off = pack(reads)
    for k in (31, 63):
        for th in (None, thresholds_L(2)):
            compare_all(k, th, b, off)


@pytest.mark.parametrize("case", CASES[:6])
@pytest.mark.parametrize("variant", [0, 1])
def test_kmer_only_threshold(case, variant):
    """-L 3 0 0 (k-mer threshold only) keeps edges into removed k-mers, so
    chain ends can point at nodes that are not in the graph: the list
    ranking falls back to round-by-round emulation."""
    seed, G, n, L, err, k = case
    _, reads = random_reads(seed, G, n, L, err)
    b, off = pack(reads)
    compare_all(k, [3, 65536, 0, 65536, 0, 65536], b, off, variant=variant)


@pytest.mark.parametrize("knobs", [dict(parse_block=128, parse_grid=64, ruler_sample=8, scan_grid=1),
                                   dict(parse_block=512, parse_grid=65536, ruler_sample=4096, scan_grid=32768),
                                   dict(parse_block=256, parse_grid=1000, ruler_sample=16, scan_grid=7)])
def test_tuning_knobs_do_not_change_results(knobs):
    """Every knob the autotuner may set leaves every parity object unchanged
    (tiny grids included: every read, slot and ruler walk is still covered)."""
    from bruno_amd.autotune import get_tuning, set_tuning
    before = get_tuning()
    try:
        set_tuning(**knobs)
        for case in (CASES[2], CASES[8]):
            seed, G, n, L, err, k = case
            _, reads = random_reads(seed, G, n, L, err)
            b, off = pack(reads)
            compare_all(k, thresholds_L(2), b, off, variant=1)
        _, reads = random_reads(3, 30000, 3000, 150, 0.0)
        b, off = pack(reads)
        compare_all(31, None, b, off, variant=1, stages=False)
    finally:
        set_tuning(**before)


def test_recycle_mode_repeats_exactly():
    """Recycle mode clears the table beside list ranking and output; three
    runs of one job in a row equal the oracle, and the filtered graph is
    unavailable once the chains are built."""
    _, reads = random_reads(31, 30000, 6000, 150, 0.01)
    b, off = pack(reads)
    th = thresholds_L(2)
    o = run_oracle(31, th, b, off, 1)
    job = bruno_amd.GraphJob(31, th, capacity=1 << 20)
    job.set_recycle(True)
    for _ in range(3):
        job.reset()
        job.parse_count(b, off)
        job.filter()
        job.build_chains()
        with pytest.raises(bruno_amd.BrunoError):
            job.nodes()
        job.list_rank(1)
        job.compact()
        assert job.unitigs() == o.chain_fasta()
        assert job.branch_fasta() == o.branch_fasta()
    job.close()
