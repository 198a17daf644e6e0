"""Batches: many independent jobs through one set of launches.

Each job of a batch owns a segment of one HBM table.  Every parity object of
every job must equal the oracle on that job's reads alone, and the same job
run by itself on the GPU (including its list-ranking round count), whatever
the other jobs hold.  All calls go through the C ABI."""
import numpy as np
import pytest

import bruno_amd
from gpu_common import compare_results, run_gpu, run_oracle
from helpers import pack, random_reads, thresholds_L

pytestmark = pytest.mark.gpu

K_ONLY_3 = [3, 65536, 0, 65536, 0, 65536]  # -L 3 0 0: k-mer threshold only


def mixed_jobs(seed):
    """Jobs of uneven size and read length, plus the awkward ones: no reads,
    only reads shorter than k + 2, reads with N, a circular genome (cycles),
    one long read (a long chain) and an exact copy of another job."""
    rng = np.random.default_rng(seed)
    jobs = []
    for i, (G, n, L, err) in enumerate([(3000, 300, 100, 0.01), (800, 120, (40, 150), 0.01),
                                        (6000, 600, 250, 0.02), (400, 80, 60, 0.0)]):
        _, reads = random_reads(seed * 10 + i, G, n, L, err)
        jobs.append(reads)
    jobs.append([])
    jobs.append(["ACGT" * 5, "A" * 20, "CCGGTTAA"])
    _, r = random_reads(seed * 10 + 7, 1000, 150, 80, 0.01)
    jobs.append([s if j % 7 else s[:30] + "N" + s[31:] for j, s in enumerate(r)])
    g = "".join(rng.choice(list("ACGT"), 2000))
    circ = g + g[:200]
    jobs.append([circ[i:i + 150] for i in range(0, 2000, 25)])
    jobs.append(["".join(rng.choice(list("ACGT"), 9000))])
    jobs.append(list(jobs[0]))
    return jobs


def pack_jobs(jobs):
    reads = [r for job in jobs for r in job]
    b, off = pack(reads)
    first = np.zeros(len(jobs) + 1, np.uint64)
    first[1:] = np.cumsum([len(j) for j in jobs])
    return b, off, first


def capacity(job, k):
    inst = sum(max(0, len(r) - k + 1) for r in job)
    return max(256, 2 * inst + 64)


def run_batch(k, th, jobs, variant=1, stages=True):
    b, off, first = pack_jobs(jobs)
    bt = bruno_amd.GraphBatch(k, th, [capacity(j, k) for j in jobs])
    bt.parse_count(b, off, first)
    res = [dict() for _ in jobs]
    if stages:
        for i in range(len(jobs)):
            res[i]["counts"] = bt.job(i).counts()
            res[i]["k2"] = bt.job(i).k2()
            res[i]["n_k2"] = len(res[i]["k2"][2])
    bt.filter()
    for i in range(len(jobs)):
        res[i]["nodes"] = bt.job(i).nodes()
    bt.build_chains()
    rounds = bt.list_rank(variant)
    bt.compact()
    for i in range(len(jobs)):
        v = bt.job(i)
        res[i]["fasta"] = v.unitigs()
        res[i]["branch_fasta"] = v.branch_fasta()
        res[i]["chain_nodes"] = v.chain_nodes()
        res[i]["branches"] = v.branches()
        st = v.stats()
        st["n_conflicts"] = 0  # batch-wide only
        st["n_k2"] = res[i].get("n_k2", 0)  # batch-wide only; the k2 list itself is compared
        res[i]["stats"] = st
        res[i]["rounds"] = st["iterations"]
    total = bt.stats()
    bt.close()
    return res, rounds, total


@pytest.mark.parametrize("k", [31, 15, 63])
@pytest.mark.parametrize("th", [None, thresholds_L(2), K_ONLY_3])
@pytest.mark.parametrize("variant", [0, 1])
def test_batch_equals_oracle_per_job(k, th, variant):
    jobs = mixed_jobs(5 + k)
    res, rounds, total = run_batch(k, th, jobs, variant)
    assert rounds == max(r["rounds"] for r in res)
    for i, job in enumerate(jobs):
        b, off = pack(job)
        o = run_oracle(k, th, b, off, variant)
        compare_results(k, res[i], o, variant)
    assert total["n_instances"] == sum(r["stats"]["n_instances"] for r in res)
    assert total["n_unitigs"] == sum(r["stats"]["n_unitigs"] for r in res)
    assert total["n_branches"] == sum(r["stats"]["n_branches"] for r in res)


@pytest.mark.parametrize("variant", [0, 1])
def test_batch_round_emulation_per_job(variant, monkeypatch):
    """Forced round-by-round list ranking: each job stops at its own round
    count, as it would alone."""
    monkeypatch.setenv("BRUNO_LR_ROUNDS", "1")
    k, th = 31, thresholds_L(2)
    jobs = mixed_jobs(41)
    res, _, _ = run_batch(k, th, jobs, variant, stages=False)
    for i, job in enumerate(jobs):
        b, off = pack(job)
        compare_results(k, res[i], run_oracle(k, th, b, off, variant), variant, stages=False)


def test_result_alone_equals_result_in_the_mix():
    """The same job alone (one-job launch) and inside two different batches."""
    k, th = 31, thresholds_L(3)
    target = random_reads(99, 20000, 5000, 150, 0.01)[1]
    others_a = [random_reads(100 + i, 3000 * (i + 1), 500 * (i + 1), 100, 0.01)[1] for i in range(5)]
    others_b = [random_reads(200 + i, 1000, 300, (50, 250), 0.02)[1] for i in range(12)]
    b, off = pack(target)
    alone = run_gpu(k, th, b, off, 1, stages=False)
    for mix, pos in ((others_a, 2), (others_b, 0), (others_b, 12)):
        jobs = list(mix)
        jobs.insert(pos, target)
        res, _, _ = run_batch(k, th, jobs, 1, stages=False)
        got = res[pos]
        assert got["fasta"] == alone["fasta"]
        assert got["branch_fasta"] == alone["branch_fasta"]
        assert np.array_equal(got["chain_nodes"], alone["chain_nodes"])
        assert got["rounds"] == alone["rounds"]
        for f in ("n_instances", "n_unitigs", "n_branches", "n_cycle_nodes", "n_isolated", "n_chain_nodes"):
            assert got["stats"][f] == alone["stats"][f], f


def test_batch_reset_and_reuse():
    k, th = 31, thresholds_L(2)
    jobs = mixed_jobs(3)
    b, off, first = pack_jobs(jobs)
    caps = [capacity(j, k) for j in jobs]
    bt = bruno_amd.GraphBatch(k, th, caps)
    outs = []
    for _ in range(2):
        bt.reset()
        bt.parse_count(b, off, first)
        bt.run(1)
        outs.append([bt.job(i).unitigs() for i in range(len(jobs))])
    assert outs[0] == outs[1]
    bt.close()
    for i, job in enumerate(jobs):
        bj, oj = pack(job)
        assert outs[0][i] == run_oracle(k, th, bj, oj).chain_fasta()


def test_batch_parse_in_several_calls():
    """bruno_batch_parse_count called per chunk of reads (each call names
    every job's share) equals one call."""
    k, th = 31, thresholds_L(2)
    jobs = mixed_jobs(8)
    halves = [(j[:len(j) // 2], j[len(j) // 2:]) for j in jobs]
    bt = bruno_amd.GraphBatch(k, th, [capacity(j, k) for j in jobs])
    for part in (0, 1):
        b, off, first = pack_jobs([h[part] for h in halves])
        bt.parse_count(b, off, first)
    bt.run(1)
    for i, job in enumerate(jobs):
        bj, oj = pack(job)
        assert bt.job(i).unitigs() == run_oracle(k, th, bj, oj).chain_fasta()
        assert bt.job(i).stats()["n_reads"] == len(job)
    bt.close()


def test_batch_segment_full_reports_capacity():
    """A job whose segment is too small fails the batch with CAPACITY (the
    probe stops after one sweep of the segment), and the handle stays
    usable for reset and destroy."""
    k = 31
    jobs = [random_reads(1, 5000, 400, 100, 0.01)[1], random_reads(2, 5000, 400, 100, 0.01)[1]]
    b, off, first = pack_jobs(jobs)
    bt = bruno_amd.GraphBatch(k, None, [capacity(jobs[0], k), 64])
    with pytest.raises(bruno_amd.BrunoError) as e:
        bt.parse_count(b, off, first)
    assert e.value.code == 3
    bt.close()


def test_whole_job_getters_refuse_a_batch():
    bt = bruno_amd.GraphBatch(31, None, [256, 256])
    b, off, first = pack_jobs([["ACGT" * 10], ["TTGCA" * 10]])
    bt.parse_count(b, off, first)
    with pytest.raises(bruno_amd.BrunoError):
        bt.counts()
    with pytest.raises(bruno_amd.BrunoError):
        bruno_amd.GraphJob.parse_count_device(bt, 0, 0, 0)
    bt.close()


def test_synth_batch_device_equals_host():
    from bruno_amd.configs import config4_jobs
    from bruno_amd.device import DeviceBuffer, synchronize
    from bruno_amd.pipeline import synth_batch_device, synth_batch_host
    jobs = config4_jobs(seed=4, n_jobs=6, scale=0.003)
    b, o, first = synth_batch_host(jobs)
    db, do = DeviceBuffer(max(1, b.size)), DeviceBuffer(o.size * 8)
    f2 = synth_batch_device(jobs, db.ptr, do.ptr)
    synchronize()
    assert np.array_equal(f2, first)
    assert np.array_equal(do.to_numpy(np.uint64, o.size), o)
    assert np.array_equal(db.to_numpy(np.uint8, b.size), b)
    db.free()
    do.free()


def test_device_pool_reuses_and_coalesces():
    """Blocks freed in any order coalesce back into one free block of the
    chunk; a reserved chunk serves later allocations without growing."""
    from bruno_amd.device import DeviceBuffer, pool_reserve, pool_stats
    pool_reserve(1 << 30)
    before = pool_stats()
    bufs = [DeviceBuffer(s) for s in (100, 1 << 20, 12345, 1 << 26, 7, 3 << 20)]
    mid = pool_stats()
    assert mid["reserved"] == before["reserved"]
    assert mid["in_use"] >= before["in_use"] + sum(b.nbytes for b in bufs)
    for i in (3, 0, 5, 1, 4, 2):
        bufs[i].free()
    after = pool_stats()
    assert after["in_use"] == before["in_use"]
    assert after["largest_free"] == before["largest_free"]
    # a job's table and phase buffers come from the pool as well
    k, th = 31, thresholds_L(2)
    jobs = mixed_jobs(3)[:3]
    b, off, first = pack_jobs(jobs)
    bt = bruno_amd.GraphBatch(k, th, [capacity(j, k) for j in jobs])
    bt.parse_count(b, off, first)
    bt.run(1)
    assert pool_stats()["in_use"] > after["in_use"]
    bt.close()
    assert pool_stats()["in_use"] == after["in_use"]
