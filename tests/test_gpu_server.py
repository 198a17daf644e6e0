"""The request loop: jobs submitted concurrently from several threads, with
different k and thresholds, are admitted by footprint, batched with whatever
else is waiting, and each job's result is the same as the job run alone
(and as the oracle)."""
import threading
import time

import numpy as np
import pytest

import bruno_amd
from bruno_amd.server import Server
from gpu_common import run_gpu, run_oracle
from helpers import pack, random_reads, thresholds_L

pytestmark = pytest.mark.gpu


def job_set(seed, n):
    rng = np.random.default_rng(seed)
    jobs = []
    for i in range(n):
        k = [31, 31, 31, 63, 21][i % 5]
        th = [thresholds_L(2), thresholds_L(3), None][i % 3]
        G = int(rng.integers(500, 20000))
        L = int(rng.choice([80, 100, 150, 250]))
        n_reads = int(G * rng.uniform(5, 30) / L)
        _, reads = random_reads(seed * 1000 + i, G, n_reads, L, 0.01)
        jobs.append((k, th, reads))
    return jobs


def cap_of(reads, k):
    return max(4096, 2 * sum(max(0, len(r) - k + 1) for r in reads))


def test_results_in_the_mix_equal_alone():
    jobs = job_set(1, 24)
    tickets = [None] * len(jobs)
    with Server(hbm_budget=16 << 30, n_workers=2, max_batch_jobs=16, reserve_pool=False) as srv:
        def submit(idx):
            for i in idx:
                k, th, reads = jobs[i]
                b, off = pack(reads)
                tickets[i] = srv.submit_reads(k, th, b, off, capacity=cap_of(reads, k))
                time.sleep(0.002 * (i % 3))
        threads = [threading.Thread(target=submit, args=(range(t, len(jobs), 3),)) for t in range(3)]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
        infos = [srv.wait(t) for t in tickets]
        outs = [(srv.unitigs(t), srv.branch_fasta(t)) for t in tickets]
        st = srv.stats()
        for t in tickets:
            srv.release(t)
    assert st["jobs_done"] == len(jobs) and st["jobs_failed"] == 0
    assert st["max_batch_jobs"] > 1  # some jobs ran together
    for i, (k, th, reads) in enumerate(jobs):
        assert infos[i]["status"] == 0
        b, off = pack(reads)
        alone = run_gpu(k, th, b, off, 1, stages=False)
        assert outs[i][0] == alone["fasta"], i
        assert outs[i][1] == alone["branch_fasta"], i
        assert infos[i]["iterations"] == alone["rounds"]
        assert infos[i]["n_instances"] == alone["stats"]["n_instances"]
        assert infos[i]["end_s"] >= infos[i]["start_s"] >= infos[i]["submit_s"]
        if i % 4 == 0:
            o = run_oracle(k, th, b, off, 1)
            assert outs[i][0] == o.chain_fasta() and outs[i][1] == o.branch_fasta()


def test_small_table_is_retried_alone_with_more_slots():
    k, th = 31, thresholds_L(2)
    _, reads = random_reads(5, 20000, 3000, 120, 0.01)
    _, other = random_reads(6, 3000, 400, 100, 0.01)
    b, off = pack(reads)
    b2, off2 = pack(other)
    with Server(hbm_budget=4 << 30, n_workers=1, reserve_pool=False) as srv:
        t1 = srv.submit_reads(k, th, b, off, capacity=8192)  # far too few slots
        t2 = srv.submit_reads(k, th, b2, off2, capacity=cap_of(other, k))
        i1, i2 = srv.wait(t1), srv.wait(t2)
        u1, u2 = srv.unitigs(t1), srv.unitigs(t2)
        assert srv.stats()["retries"] >= 1
    assert i1["status"] == 0 and i1["attempts"] > 1
    assert i2["status"] == 0
    assert u1 == run_oracle(k, th, b, off).chain_fasta()
    assert u2 == run_oracle(k, th, b2, off2).chain_fasta()


def test_job_larger_than_the_budget_fails_with_capacity():
    _, reads = random_reads(7, 5000, 500, 100, 0.01)
    b, off = pack(reads)
    with Server(hbm_budget=1 << 20, n_workers=1, reserve_pool=False) as srv:
        t = srv.submit_reads(31, None, b, off, capacity=1 << 20)
        info = srv.wait(t)
    assert info["status"] == 3


def test_synthetic_requests_equal_host_reads():
    from bruno_amd.configs import config4_jobs
    from bruno_amd.planner import capacity
    cfgs = config4_jobs(seed=9, n_jobs=5, scale=0.01)
    th = thresholds_L(3)
    with Server(hbm_budget=16 << 30, n_workers=2, reserve_pool=False) as srv:
        ts = [srv.submit_synth(31, th, c, capacity=capacity(c, 31)) for c in cfgs]
        res = [(srv.wait(t), srv.unitigs(t)) for t in ts]
    for c, (info, u) in zip(cfgs, res):
        b, o = bruno_amd.synth_reads_host(c)
        job = bruno_amd.GraphJob(31, th, capacity=capacity(c, 31))
        job.parse_count(b, o)
        job.run(1)
        assert u == job.unitigs()
        assert info["n_instances"] == job.stats()["n_instances"]
        job.close()


def test_destroy_finishes_queued_jobs():
    jobs = job_set(3, 6)
    srv = Server(hbm_budget=8 << 30, n_workers=2, reserve_pool=False)
    for k, th, reads in jobs:
        b, off = pack(reads)
        srv.submit_reads(k, th, b, off, capacity=cap_of(reads, k))
    st = srv.stats()
    assert st["submitted"] == len(jobs)
    srv.close()  # returns once every job has run
