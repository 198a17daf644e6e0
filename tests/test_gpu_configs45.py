"""The batched (config4) and served (config5) workloads at their real shape.

config4: the whole 256-job list, packed by estimated footprint into several
launch groups, each group one batch.  At a reduced genome scale every job is
compared with the oracle on its own reads and with the same job run alone
(one job per launch); at full size (27 G k-mer instances) every job's outputs
pass the size-independent checks of the full-size tests and equal the job run
alone.

config5: the Poisson request stream (k = 31 / 63 at 80/20) through the
request loop, whose admission batches whatever is waiting.  At a reduced
scale every result equals the job run alone and the oracle, including a job
whose table is too small and is retried; at full size every result passes
the size-independent checks and equals the job run alone."""
import hashlib
import os
import time

import numpy as np
import pytest

import bruno_amd
from bruno_amd.configs import config4_jobs, config5_jobs
from bruno_amd.device import DeviceBuffer, mem_info, pool_trim, synchronize
from bruno_amd.pipeline import synth_batch_device, synth_reads_device, thresholds_L
from bruno_amd.planner import SOLO_SLOTS, capacity, footprint, plan_groups
from bruno_amd.server import Server
from gpu_common import run_oracle

pytestmark = pytest.mark.gpu

TH = thresholds_L(3)
SMALL = 0.005
COMP = bytes.maketrans(b"ACGT", b"TGCA")
THREADS = min(16, os.cpu_count() or 1)


def check_outputs(k, fasta, branch, st):
    """Size-independent properties of one job's outputs: counts add up,
    headers are the canonical first k-mers, records sorted by them, one
    node per sequence position, branch records of fixed size."""
    assert st["n_nodes"] == st["n_branches"] + st["n_chain_nodes"]
    assert st["n_chain_nodes"] == st["n_unitig_nodes"] + st["n_cycle_nodes"] + st["n_isolated"]
    lines = fasta.split(b"\n") if fasta else []
    heads, seqs = lines[0::2], lines[1::2]
    assert len(heads) == len(seqs) == st["n_unitigs"]
    firsts = [q[:k] for q in seqs]
    assert all(len(q) >= k for q in seqs)
    assert all(h[1:] == min(f, f.translate(COMP)[::-1]) and h[:1] == b">" for h, f in zip(heads, firsts))
    assert all(firsts[i] < firsts[i + 1] for i in range(len(firsts) - 1))
    assert sum(len(q) - k + 1 for q in seqs) == st["n_unitig_nodes"]
    assert len(branch) == st["n_branches"] * (2 * k + 3)


def digest(b):
    return hashlib.sha1(b).hexdigest()


def run_group(k, jobs, idx, keep_text=True):
    cfgs = [jobs[i] for i in idx]
    nb, nr = sum(c.n_bases for c in cfgs), sum(c.n_reads for c in cfgs)
    bases, offs = DeviceBuffer(max(1, nb)), DeviceBuffer((nr + 1) * 8)
    first = synth_batch_device(cfgs, bases.ptr, offs.ptr)
    synchronize()
    bt = bruno_amd.GraphBatch(k, TH, [capacity(c, k) for c in cfgs])
    bt.parse_count_device(bases.ptr, offs.ptr, nr, first)
    bt.run(1)
    out = {}
    for q, i in enumerate(idx):
        v = bt.job(q)
        u, b = v.unitigs(), v.branch_fasta()
        out[i] = dict(stats=v.stats(), fasta=u if keep_text else None, branch=b if keep_text else None,
                      digest=(digest(u), digest(b)))
    bt.close()
    bases.free()
    offs.free()
    return out


def run_alone(k, c, cap=None):
    bases, offs = DeviceBuffer(max(1, c.n_bases)), DeviceBuffer((c.n_reads + 1) * 8)
    synth_reads_device(c, bases.ptr, offs.ptr)
    synchronize()
    job = bruno_amd.GraphJob(k, TH, capacity=capacity(c, k) if cap is None else cap)
    job.parse_count_device(bases.ptr, offs.ptr, c.n_reads)
    job.run(1)
    u, b, st = job.unitigs(), job.branch_fasta(), job.stats()
    job.close()
    bases.free()
    offs.free()
    return dict(stats=st, fasta=u, branch=b, digest=(digest(u), digest(b)))


def same_job(a, b):
    assert a["digest"] == b["digest"]
    for f in ("n_reads", "n_instances", "n_unitigs", "n_unitig_nodes", "n_branches", "n_chain_nodes",
              "n_cycle_nodes", "n_isolated", "iterations"):
        assert a["stats"][f] == b["stats"][f], f


def test_config4_all_jobs_in_launch_groups_equal_oracle_and_alone():
    k = 31
    jobs = config4_jobs(n_jobs=256, scale=SMALL)
    total = sum(footprint(c, k) for c in jobs)
    groups = plan_groups(jobs, k, total / 5)
    assert len(groups) >= 4 and sorted(i for g in groups for i in g) == list(range(256))
    res = {}
    for g in groups:
        res.update(run_group(k, jobs, g))
    for i, c in enumerate(jobs):
        b, o = bruno_amd.synth_reads_host(c)
        orc = run_oracle(k, TH, b, o, 1, nthreads=THREADS)
        r = res[i]
        assert r["fasta"] == orc.chain_fasta(), i
        assert r["branch"] == orc.branch_fasta(), i
        assert r["stats"]["n_instances"] == orc.stat("n_instances")
        assert r["stats"]["iterations"] == orc.stat("iterations")
        assert r["stats"]["n_cycle_nodes"] == orc.stat("n_cycle_nodes")
        assert r["stats"]["n_isolated"] == orc.stat("n_isolated")
        same_job(r, run_alone(k, c))


def test_config4_full_size_batched_invariants_and_alone():
    k = 31
    jobs = config4_jobs(n_jobs=256)
    pool_trim()
    free, _ = mem_info()
    groups = plan_groups(jobs, k, 0.8 * free)
    assert len(groups) > 1
    inst = 0
    for g in groups:
        res = run_group(k, jobs, g)
        if len(g) == 1 and capacity(jobs[g[0]], k) >= SOLO_SLOTS:
            assert res[g[0]]["stats"]["count_by_regions"] == 1  # a large job alone counts by regions
        for i in g:
            r = res[i]
            c = jobs[i]
            assert r["stats"]["n_instances"] == c.n_reads * (c.read_len - k + 1)
            check_outputs(k, r["fasta"], r["branch"], r["stats"])
            same_job(r, run_alone(k, c))
            inst += r["stats"]["n_instances"]
    assert inst > 2.5e10


def serve_stream(cfgs, ks, arrivals, caps, budget, workers=2):
    """Submit each job at its arrival time; return (infos, texts, stats)."""
    with Server(hbm_budget=int(budget), n_workers=workers, max_batch_jobs=64, reserve_pool=False) as srv:
        t0 = time.perf_counter()
        tickets = []
        for c, k, a, cap in zip(cfgs, ks, arrivals, caps):
            d = t0 + a - time.perf_counter()
            if d > 0:
                time.sleep(d)
            tickets.append(srv.submit_synth(k, TH, c, capacity=cap))
        infos, texts = [], []
        for t in tickets:
            infos.append(srv.wait(t))
            u, b = srv.unitigs(t), srv.branch_fasta(t)
            texts.append((digest(u), digest(b), u if len(u) < (64 << 20) else None, b))
            srv.release(t)
        st = srv.stats()
    return infos, texts, st


def test_config5_poisson_stream_equals_alone_and_oracle():
    cfgs, ks, arrivals = config5_jobs(n_jobs=256, rate=2000.0, scale=SMALL)
    assert 20 < sum(k == 63 for k in ks) < 80
    caps = [capacity(c, k) for c, k in zip(cfgs, ks)]
    caps[7] = 2048  # far too small: the loop retries this job alone with more slots
    infos, texts, st = serve_stream(cfgs, ks, arrivals, caps, 8 << 30)
    assert st["jobs_done"] == 256 and st["jobs_failed"] == 0
    assert st["max_batch_jobs"] > 1 and st["retries"] >= 1
    assert infos[7]["attempts"] > 1
    for i, (c, k) in enumerate(zip(cfgs, ks)):
        info, (du, db, u, b) = infos[i], texts[i]
        assert info["status"] == 0
        alone = run_alone(k, c)
        assert (du, db) == alone["digest"], i
        assert info["iterations"] == alone["stats"]["iterations"]
        assert info["n_instances"] == alone["stats"]["n_instances"]
        bh, oh = bruno_amd.synth_reads_host(c)
        orc = run_oracle(k, TH, bh, oh, 1, nthreads=THREADS)
        assert u == orc.chain_fasta(), i
        assert b == orc.branch_fasta(), i


def test_config5_full_size_stream_invariants_and_alone():
    cfgs, ks, arrivals = config5_jobs(n_jobs=256, rate=40.0)
    caps = [capacity(c, k) for c, k in zip(cfgs, ks)]
    pool_trim()  # memory the pool kept from earlier tests goes back to the device
    free, _ = mem_info()
    infos, texts, st = serve_stream(cfgs, ks, arrivals, caps, 0.85 * free)
    failed = [(i, x["status"], caps[i], ks[i]) for i, x in enumerate(infos) if x["status"] != 0]
    assert st["jobs_done"] == 256 and st["jobs_failed"] == 0, (failed, free)
    for i, (c, k) in enumerate(zip(cfgs, ks)):
        info, (du, db, u, b) = infos[i], texts[i]
        assert info["status"] == 0
        assert info["n_instances"] == c.n_reads * (c.read_len - k + 1)
        alone = run_alone(k, c)
        assert (du, db) == alone["digest"], i
        assert info["iterations"] == alone["stats"]["iterations"]
        check_outputs(k, alone["fasta"], alone["branch"], alone["stats"])
