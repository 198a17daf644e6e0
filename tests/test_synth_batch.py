"""The batch generator lays jobs back to back: the same bytes as each job's
own generator, with global offsets (host generator; no GPU)."""
import numpy as np

import bruno_amd
from bruno_amd.configs import config4_jobs
from bruno_amd.pipeline import synth_batch_host
from bruno_amd.planner import plan_groups


def test_batch_host_equals_jobs_back_to_back():
    jobs = config4_jobs(seed=4, n_jobs=5, scale=0.002)
    b, o, first = synth_batch_host(jobs)
    assert first[0] == 0 and first[-1] == sum(j.n_reads for j in jobs)
    for i, j in enumerate(jobs):
        bj, oj = bruno_amd.synth_reads_host(j)
        r0, r1 = int(first[i]), int(first[i + 1])
        assert np.array_equal(o[r0:r1 + 1] - o[r0], oj)
        assert bytes(b[int(o[r0]):int(o[r1])]) == bytes(bj)


def test_plan_groups_respects_budget_and_order():
    jobs = config4_jobs(seed=4, n_jobs=64)
    from bruno_amd.planner import footprint
    budget = 40e9
    groups = plan_groups(jobs, 31, budget)
    assert sorted(i for g in groups for i in g) == list(range(64))
    for g in groups:
        assert g == sorted(g)
        assert len(g) == 1 or sum(footprint(jobs[i], 31) for i in g) <= budget


def test_plan_groups_runs_region_sized_jobs_alone():
    """Jobs whose table reaches the solo size get a launch group of
    their own, in arrival order; the others fill groups within the budget
    past them."""
    from bruno_amd.planner import SOLO_SLOTS, capacity, footprint
    jobs = config4_jobs(seed=4, n_jobs=64)
    big = [capacity(j, 31) >= SOLO_SLOTS for j in jobs]
    assert 0 < sum(big) < len(jobs)
    budget = 40e9
    groups = plan_groups(jobs, 31, budget)
    assert sorted(i for g in groups for i in g) == list(range(64))
    for g in groups:
        assert g == sorted(g)
        assert len(g) == 1 or (not any(big[i] for i in g) and sum(footprint(jobs[i], 31) for i in g) <= budget)
    assert sum(len(g) == 1 and big[g[0]] for g in groups) == sum(big)
    # the small jobs need as many groups as their footprint asks, not one per gap between large jobs
    small = [footprint(j, 31) for j, b in zip(jobs, big) if not b]
    assert sum(len(g) > 1 or not big[g[0]] for g in groups) <= 2 * sum(small) / budget + 2
    # solo_slots=0: budget packing only, consecutive
    flat = plan_groups(jobs, 31, budget, solo_slots=0)
    assert [i for g in flat for i in g] == list(range(64))


def test_short_first_orders_groups_by_instances_per_job():
    from bruno_amd.planner import instances, short_first
    jobs = config4_jobs(seed=4, n_jobs=64)
    groups = plan_groups(jobs, 31, 40e9)
    order = short_first(groups, jobs, 31)
    assert sorted(map(tuple, order)) == sorted(map(tuple, groups))
    w = [sum(instances(jobs[i], 31) for i in g) / len(g) for g in order]
    assert w == sorted(w)
