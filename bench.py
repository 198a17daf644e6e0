"""Benchmark: k-mers/s of the full de Bruijn hot path on one MI355X.

A step is one job over one batch of synthetic reads already resident in HBM:
clear the job's table, parse + count, filter, extract chains, list rank,
compact to the unitig / branch FASTA bytes in device memory.  The default
workload is 100 Mbp with 2% repeats (50 families, 1% divergence), 33.3 M
150 bp reads at 50x with 1% substitutions, k=31, -T -L 3.

Prints one JSON line with the metric, a roofline object for the count kernel
(the dominant kernel) and a CPU baseline: the oracle (the C++ restatement of
bruno's algorithm, hash-partitioned over host threads) timed on a bounded
sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def count_bytes(n_bases, n_inst):
    """Algorithmic bytes of the count kernel: every base read once plus one
    random read-modify-write of a 64-byte line per k-mer instance."""
    return n_bases * 1 + n_inst * 2 * 64


def capacity_for(synth, k):
    """Hash slots for a synthetic workload: genome k-mers plus the novel
    k-mers each substitution / indel creates, at load factor ~0.5."""
    L = synth.read_len
    per_read_kmers = max(0, L - k + 1)
    err = synth.sub_rate + 2 * synth.indel_rate
    cover = min(k, L) * per_read_kmers / max(1, L)
    novel = synth.n_reads * L * err * cover
    distinct = synth.genome_len + novel
    return int(distinct / 0.5) + 4096


def cpu_baseline(cfg, seconds_target=15.0, threads=None):
    import numpy as np
    import bruno_amd
    from bruno_amd.pipeline import SynthConfig
    from oracle.oracle import Oracle  # the checker, timed as the baseline
    s = cfg["synth"]
    threads = threads or min(16, os.cpu_count() or 1)
    # reads 0..S-1 of the generator are the first S reads of the full workload
    n = max(1000, min(s.n_reads, int(2_000_000 * seconds_target / 15.0)))
    sample = SynthConfig(**{**s.__dict__, "n_reads": n})
    b, o = bruno_amd.synth_reads_host(sample)
    t0 = time.time()
    orc = Oracle(cfg["k"], np.array(cfg["thresholds"], np.uint64), threads)
    orc.add_reads(b, o)
    orc.run(1)
    dt = time.time() - t0
    inst = orc.stat("n_instances")
    return {"value": inst / dt, "unit": "k-mers/s", "cores": threads, "kind": "port",
            "sample": "first %d of %d reads of the same workload (%d k-mer instances), "
                      "all phases, %.1f s" % (n, s.n_reads, inst, dt)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="config2")
    ap.add_argument("--scale", type=float, default=1.0)
    ap.add_argument("--variant", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="after the timed steps, copy 4 GiB device-to-device once (a plain copy kernel of "
                         "known size, to scale FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path "
              "(--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import bruno_amd
    from bruno_amd.configs import CONFIGS
    from bruno_amd.device import DeviceBuffer, synchronize
    from bruno_amd.pipeline import synth_reads_device

    cfg = CONFIGS[args.config](scale=args.scale)
    s = cfg["synth"]
    k = cfg["k"]
    bases = DeviceBuffer(s.n_bases)
    offs = DeviceBuffer((s.n_reads + 1) * 8)
    synth_reads_device(s, bases.ptr, offs.ptr)
    synchronize()

    cap = capacity_for(s, k)
    job = bruno_amd.GraphJob(k, cfg["thresholds"], capacity=cap)
    times, phases = [], []
    stats = None
    for it in range(args.warmup + args.steps):
        synchronize()
        t0 = time.perf_counter()
        job.reset()
        job.parse_count_device(bases.ptr, offs.ptr, s.n_reads)
        job.run(args.variant)
        synchronize()
        dt = time.perf_counter() - t0
        if it >= args.warmup:
            times.append(dt)
            phases.append(job.phase_ms())
        stats = job.stats()
        print("step %d: %.1f ms  %s" % (it, dt * 1e3, json.dumps({k_: round(v, 2) for k_, v in job.phase_ms().items()})),
              file=sys.stderr, flush=True)
    n_inst = stats["n_instances"]
    total = sum(times)
    value = n_inst * len(times) / total
    pc_ms = float(np.mean([p["parse_count_ms"] for p in phases]))
    achieved = count_bytes(s.n_bases, n_inst) / (pc_ms * 1e-3) / 1e9
    ph = {k_: round(float(np.mean([p[k_] for p in phases])), 3) for k_ in phases[0]}
    out = {
        "metric": "k-mers/s (parse+count+filter+compact)",
        "value": value,
        "unit": "k-mers/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": total / len(times) * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic (seeded generator, reads generated in HBM)",
        "config": {
            "workload": "%s: %.0f Mbp genome, %d x %d bp reads, sub %.3g, indel %.3g, repeats %d families %.0f%%, k=%d, -T -L 3"
                        % (args.config, s.genome_len / 1e6, s.n_reads, s.read_len, s.sub_rate, s.indel_rate,
                           s.n_families, 100 * s.repeat_fraction, k),
            "k": k, "n_reads": s.n_reads, "read_len": s.read_len, "genome_bp": s.genome_len,
            "k_mer_instances": n_inst, "list_rank": "min_update" if args.variant else "list_rank",
        },
        "latency_ms": {"p50": float(np.percentile(times, 50) * 1e3), "p99": float(np.percentile(times, 99) * 1e3)},
        "phases_ms": ph,
        "graph": {kk: stats[kk] for kk in ("n_distinct", "n_k2", "n_overflow", "n_nodes", "n_branches", "n_chain_nodes",
                                           "n_unitigs", "n_cycle_nodes", "iterations", "capacity", "device_bytes")},
        "roofline": {
            "kernel": "k_parse (parse + k2 count)",
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
        },
    }
    job.close()
    bases.free()
    offs.free()
    if args.calibrate_copy:
        import ctypes
        nb = 4 << 30
        src, dst = DeviceBuffer(nb), DeviceBuffer(nb)
        bruno_amd.lib().bruno_dev_copy(ctypes.c_void_p(dst.ptr), ctypes.c_void_p(src.ptr), nb, 2)
        synchronize()
        out["calibration_copy_bytes"] = nb
        src.free()
        dst.free()
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(cfg, args.cpu_seconds)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out))


if __name__ == "__main__":
    main()
