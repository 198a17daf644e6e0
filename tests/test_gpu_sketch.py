"""Distinct-k-mer estimate (HyperLogLog on a hash sample) against the exact
count, and jobs created without a capacity sized from it."""
import numpy as np
import pytest

import bruno_amd
from bruno_amd.configs import config2, config3
from bruno_amd.device import DeviceBuffer, synchronize
from bruno_amd.pipeline import estimate_distinct, synth_reads_device
from helpers import pack, random_reads, thresholds_L

pytestmark = pytest.mark.gpu


def device_reads(s):
    bases, offs = DeviceBuffer(s.n_bases), DeviceBuffer((s.n_reads + 1) * 8)
    synth_reads_device(s, bases.ptr, offs.ptr)
    synchronize()
    return bases, offs


@pytest.mark.parametrize("cfg", [config2(scale=0.02), config2(scale=0.2), config3(scale=0.1)])
def test_estimate_close_to_exact(cfg):
    s, k = cfg["synth"], cfg["k"]
    bases, offs = device_reads(s)
    est, cap = estimate_distinct(k, bases.ptr, offs.ptr, s.n_reads)
    job = bruno_amd.GraphJob(k, cfg["thresholds"], capacity=cap)
    job.parse_count_device(bases.ptr, offs.ptr, s.n_reads)
    exact = job.stats()["n_distinct"]
    job.close()
    bases.free()
    offs.free()
    assert abs(est - exact) / exact < 0.05, (est, exact)
    assert cap >= 2 * exact  # the margin keeps the load at or under 1/2


def test_small_inputs_get_room():
    """Tiny inputs (a handful of sampled k-mers) still get a working table:
    at least one slot per instance at load 1/2 when the estimate is noisy."""
    for n in (1, 10, 300):
        _, reads = random_reads(n, 2000, n, 100, 0.01)
        b, off = pack(reads)
        db, do = DeviceBuffer.from_numpy(np.frombuffer(b, np.uint8)), DeviceBuffer.from_numpy(off)
        est, cap = estimate_distinct(31, db.ptr, do.ptr, len(reads))
        assert cap >= 4096
        job = bruno_amd.GraphJob(31, thresholds_L(1))  # capacity 0: sized from the sketch
        job.parse_count_device(db.ptr, do.ptr, len(reads))
        st = job.stats()
        assert st["n_distinct"] <= st["capacity"]
        job.close()
        db.free()
        do.free()


def test_capacity_zero_job_matches_sized_job():
    cfg = config2(scale=0.3)   # 30 Mbp, 10 M reads: a table sized from the bases would need 3 G slots
    s, k = cfg["synth"], cfg["k"]
    bases, offs = device_reads(s)
    outs = []
    for cap in (0, None):
        if cap is None:
            from bruno_amd.planner import capacity
            cap = capacity(s, k)
        job = bruno_amd.GraphJob(k, cfg["thresholds"], capacity=cap)
        job.parse_count_device(bases.ptr, offs.ptr, s.n_reads)
        job.run(1)
        outs.append((job.unitigs(), job.branch_fasta(), job.stats()))
        job.close()
    bases.free()
    offs.free()
    assert outs[0][0] == outs[1][0] and outs[0][1] == outs[1][1]
    st = outs[0][2]
    assert st["n_distinct"] <= st["capacity"] / 2 * 1.1
