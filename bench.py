"""Benchmark: k-mers/s of the full de Bruijn hot path on one MI355X.

A step is one job over one batch of synthetic reads already resident in HBM:
clear the job's table, parse + count, filter, extract chains, list rank,
compact to the unitig / branch FASTA bytes in device memory.  The default
workload is 100 Mbp with 2% repeats (50 families, 1% divergence), 33.3 M
150 bp reads at 50x with 1% substitutions, k=31, -T -L 3.

--config config4 runs the batched workload instead: 256 independent jobs of
uneven size and read length, packed by estimated HBM footprint into launch
groups, each group one batch (one set of launches for all its jobs), timed
against the same jobs run one job per launch.

--config config5 serves jobs drawn from the config4 distribution (k = 31 or
63, 80/20) arriving as a Poisson process: the request loop admits them by
estimated HBM footprint and batches whatever is waiting on two streams; the
same arrivals are replayed against a one-job-at-a-time server.

Prints one JSON line with the metric, a roofline object for the count (the
dominant phase: on k <= 31 region tables a record parser, per-bucket radix
sorts and on-chip region builds, timed together with device events) and a
CPU baseline: the oracle (the C++ restatement of bruno's algorithm,
hash-partitioned over host threads) timed on a bounded sample of the same
workload.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
POOL_FRACTION = 0.9


def filter_bytes(stats, k):
    """Algorithmic bytes of the filter: every distinct (k-mer, biedge) pair
    read once and every kept k-mer's node record written once."""
    r, w = (16, 32) if k <= 31 else (24, 48)
    return stats["n_k2"] * r + stats["n_nodes"] * w


def count_bytes(n_bases, n_inst):
    """Algorithmic bytes of the count kernel: every base read once plus one
    random read-modify-write of a 64-byte line per k-mer instance."""
    return n_bases * 1 + n_inst * 2 * 64


def measured_traffic(config, scale, sized_from, kernel):
    """HBM bytes per launch of a kernel on this workload, from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE runs (profiles/hbm_traffic.json), or
    None when no counter run covers the workload (same table sizing)."""
    try:
        with open(os.path.join(ROOT, "profiles", "hbm_traffic.json")) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    if (t.get("workload") != config or float(t.get("scale", -1)) != float(scale)
            or t.get("table_sized_from", "model") != sized_from):
        return None
    return t["bytes_per_launch"].get(kernel)


def capacity_for(synth, k, load_factor=0.5):
    """Hash slots for a synthetic workload: genome k-mers plus the novel
    k-mers each substitution / indel creates, at the given load factor."""
    from bruno_amd.planner import capacity
    return capacity(synth, k, load_factor)


def fastq_bytes_on_device(bases, s):
    """The workload's reads as a FASTQ file ("@\n", read, "\n+\n", 'I' x
    read length, "\n" per record), built on the host and copied to HBM."""
    import numpy as np
    from bruno_amd.device import DeviceBuffer
    L, n = s.read_len, s.n_reads
    rec = np.empty((n, 2 * L + 6), np.uint8)
    rec[:, 0], rec[:, 1] = ord("@"), ord("\n")
    rec[:, 2:2 + L] = bases.to_numpy(np.uint8, n * L).reshape(n, L)
    rec[:, 2 + L], rec[:, 3 + L], rec[:, 4 + L] = ord("\n"), ord("+"), ord("\n")
    rec[:, 5 + L:5 + 2 * L] = ord("I")
    rec[:, 5 + 2 * L] = ord("\n")
    return DeviceBuffer.from_numpy(rec.reshape(-1))


def host_cpu():
    """The host the CPU baseline ran on: CPU model, logical CPUs, and the CPUs
    this process may use."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        usable = os.cpu_count()
    return {"model": model, "logical_cpus": os.cpu_count(), "usable_cpus": usable}


def cpu_baseline(cfg, seconds_target=15.0, threads=None, edge_filter=False):
    """The oracle (bruno's algorithm restated in C++, hash-partitioned over
    host threads like MPI ranks) on the first reads of the same workload:
    with 4 threads (the reference scripts' mpirun -np 4) and with the host
    share (16 threads), each on a sample sized for ~seconds_target."""
    import numpy as np
    import bruno_amd
    from bruno_amd.pipeline import SynthConfig
    from oracle.oracle import Oracle  # the checker, timed as the baseline
    s = cfg["synth"]
    share = threads or min(16, os.cpu_count() or 1)
    runs = []
    for t in sorted({min(4, share), share}):
        # reads 0..S-1 of the generator are the first S reads of the full
        # workload; ~0.8 M reads per 15 s of work on 16 host threads
        n = max(1000, min(s.n_reads, int(800_000 * seconds_target / 15.0 * t / 16)))
        sample = SynthConfig(**{**s.__dict__, "n_reads": n})
        b, o = bruno_amd.synth_reads_host(sample)
        t0 = time.time()
        orc = Oracle(cfg["k"], np.array(cfg["thresholds"], np.uint64), t, edge_filter=edge_filter)
        orc.add_reads(b, o)
        orc.run(1)
        dt = time.time() - t0
        inst = orc.stat("n_instances")
        runs.append({"threads": t, "value": inst / dt, "reads": n, "k_mer_instances": inst, "seconds": dt})
        del orc
    best = runs[-1]
    return {"value": best["value"], "unit": "k-mers/s", "cores": best["threads"], "kind": "port",
            "sample": "first %d of %d reads of the same workload (%d k-mer instances), all phases, %.1f s"
                      % (best["reads"], s.n_reads, best["k_mer_instances"], best["seconds"]),
            "runs": runs, "host": host_cpu()}


def copy_peak_gbs(nbytes=4 << 30, reps=5):
    """Achievable HBM bandwidth on this device: the best of `reps` device to
    device copies of nbytes (read + write bytes over the copy's time)."""
    import ctypes
    import bruno_amd
    from bruno_amd.device import DeviceBuffer, synchronize
    src, dst = DeviceBuffer(nbytes), DeviceBuffer(nbytes)
    best = None
    try:
        for _ in range(reps):
            synchronize()
            t0 = time.perf_counter()
            bruno_amd.lib().bruno_dev_copy(ctypes.c_void_p(dst.ptr), ctypes.c_void_p(src.ptr), nbytes, 2)
            synchronize()
            dt = time.perf_counter() - t0
            best = dt if best is None else min(best, dt)
    finally:
        src.free()
        dst.free()
    return 2 * nbytes / best / 1e9


def time_edge_outputs(job, bases, offs, s, fq):
    """One untimed step that keeps its table, then the two getters of the
    non-benchmark outputs (device work + copy to the host), wall clock."""
    from bruno_amd.device import synchronize
    job.set_recycle(False)
    job.reset()
    if fq is not None:
        job.parse_count_fastq_device(fq.ptr, fq.nbytes)
    else:
        job.parse_count_device(bases.ptr, offs.ptr, s.n_reads)
    job.run(1)
    synchronize()
    t0 = time.perf_counter()
    bn = job.branch_nodes()
    t1 = time.perf_counter()
    ce = job.chain_edges()
    t2 = time.perf_counter()
    return {"branch_nodes_ms": (t1 - t0) * 1e3, "n_branch_nodes": int(len(bn)),
            "chain_edges_ms": (t2 - t1) * 1e3, "n_unitigs": int(len(ce)),
            "chain_nodes_summarised": int(ce["n_nodes"].sum()) if len(ce) else 0}


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
    ap.add_argument("--load-factor", type=float, default=0.5)
    ap.add_argument("--budget-gb", type=float, default=0.0,
                    help="HBM budget per launch group for config4 (0 = 80%% of free HBM)")
    ap.add_argument("--n-jobs", type=int, default=256, help="config4: number of jobs")
    ap.add_argument("--verbose", action="store_true")
    ap.add_argument("--rate", type=float, default=40.0, help="config5: arrivals per second")
    ap.add_argument("--workers", type=int, default=2, help="config5: batches in flight")
    ap.add_argument("--autotune", action="store_true",
                    help="tune load factor and launch knobs on this workload first (untimed)")
    ap.add_argument("--tuning", default="", help="apply a tuning saved by bruno_amd.autotune")
    ap.add_argument("--knob", action="append", default=[],
                    help="set one tuning knob before the run, name=value (bruno_set_tuning)")
    ap.add_argument("--fastq", action="store_true",
                    help="feed the reads as FASTQ bytes in HBM (split on the device inside each step)")
    ap.add_argument("--edge-filter", action="store_true",
                    help="filter by canonical k+1-mer counts as the block-construct driver does (-L 3 on "
                         "k+1-mer windows) instead of local_insert_by_freqencies")
    ap.add_argument("--edge-outputs", action="store_true",
                    help="after the timed steps, run one more step without table recycling and time the "
                         ".branch.edges / .chain.edges getters (non-benchmark outputs)")
    ap.add_argument("--size-from", choices=("model", "sketch"), default="sketch",
                    help="size the hash table from the generator's error model of the distinct k-mers, or from "
                         "the library's HyperLogLog sketch of the reads (bruno_estimate_distinct, untimed), as a "
                         "caller without a model would")
    ap.add_argument("--fuse-filter", action="store_true",
                    help="run the filter inside the region build's write-back (bruno_job_set_fused_filter) "
                         "instead of as a pass of its own after the count")
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
    from bruno_amd.device import DeviceBuffer, mem_info, pool_reserve, synchronize
    from bruno_amd.pipeline import synth_reads_device

    # one pool chunk for everything the run allocates: no HIP allocator
    # calls inside the timed steps
    free0, _ = mem_info()
    pool_reserve(int(POOL_FRACTION * free0))
    if args.knob:
        from bruno_amd.autotune import set_tuning
        set_tuning(**{kv.split("=")[0]: int(kv.split("=")[1]) for kv in args.knob})
    if args.config == "config4":
        return run_batched(args, free0)
    if args.config == "config5":
        return run_serving(args, free0)
    cfg = CONFIGS[args.config](scale=args.scale)
    s = cfg["synth"]
    k = cfg["k"]
    bases = DeviceBuffer(s.n_bases)
    offs = DeviceBuffer((s.n_reads + 1) * 8)
    synth_reads_device(s, bases.ptr, offs.ptr)
    synchronize()
    fq = None
    if args.fastq:
        fq = fastq_bytes_on_device(bases, s)

    lf = args.load_factor
    tuned = None
    if args.tuning:
        from bruno_amd.autotune import apply
        tuned = apply(args.tuning)
        lf = tuned["load_factor"]
    if args.autotune:
        from bruno_amd.autotune import tune
        tuned = tune(k, cfg["thresholds"], bases.ptr, offs.ptr, s.n_reads, lambda x: capacity_for(s, k, x),
                     log=lambda *a: print("autotune:", *a, file=sys.stderr, flush=True))
        lf = tuned["load_factor"]
    if args.size_from == "sketch":
        from bruno_amd.pipeline import estimate_distinct
        est, cap = estimate_distinct(k, bases.ptr, offs.ptr, s.n_reads, lf)
    else:
        est, cap = None, capacity_for(s, k, lf)
    job = bruno_amd.GraphJob(k, cfg["thresholds"], capacity=cap, edge_filter=args.edge_filter)
    # steps repeat one job: its table is cleared beside list ranking and
    # output of each step, so the next step's reset has nothing left to clear
    job.set_recycle(True)
    # optionally the filter fused into the count (the job's one batch)
    job.set_fused_filter(args.fuse_filter)
    times, phases = [], []
    stats = None
    for it in range(args.warmup + args.steps):
        synchronize()
        t0 = time.perf_counter()
        job.reset()
        if fq is not None:
            job.parse_count_fastq_device(fq.ptr, fq.nbytes)
        else:
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
    fused = bool(stats["filter_fused"])
    unfused = None
    if fused:
        # one more (untimed) step with the filter as a pass of its own, for
        # the filter's own roofline beside the fused one
        job.set_fused_filter(False)
        job.reset()
        if fq is not None:
            job.parse_count_fastq_device(fq.ptr, fq.nbytes)
        else:
            job.parse_count_device(bases.ptr, offs.ptr, s.n_reads)
        job.run(args.variant)
        synchronize()
        up = job.phase_ms()
        unfused = {"parse_count_ms": round(up["parse_count_ms"], 3), "filter_ms": round(up["filter_ms"], 3),
                   "filter_achieved": filter_bytes(stats, k) / (up["filter_ms"] * 1e-3) / 1e9,
                   "filter_frac": filter_bytes(stats, k) / (up["filter_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS}
    total = sum(times)
    value = n_inst * len(times) / total
    pc_ms = float(np.mean([p["parse_count_ms"] for p in phases]))
    achieved = count_bytes(s.n_bases, n_inst) / (pc_ms * 1e-3) / 1e9
    payload_bytes = s.n_bases + n_inst * 2 * 16
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
            "workload": "%s: %.0f Mbp genome, %d x %d bp reads, sub %.3g, indel %.3g, repeats %d families %.0f%%, k=%d, %s"
                        % (args.config, s.genome_len / 1e6, s.n_reads, s.read_len, s.sub_rate, s.indel_rate,
                           s.n_families, 100 * s.repeat_fraction, k,
                           "-T -L 3 on canonical k+1-mer counts (edge-frequency filter)" if args.edge_filter
                           else "-T -L 3"),
            "k": k, "n_reads": s.n_reads, "read_len": s.read_len, "genome_bp": s.genome_len,
            "k_mer_instances": n_inst, "list_rank": "min_update" if args.variant else "list_rank",
            "load_factor": lf,
            "table_sized_from": args.size_from if est is None else "sketch (%.4g distinct estimated)" % est,
            "input": "FASTQ bytes in HBM (%.1f GB), split on the device in each step" % (fq.nbytes / 1e9)
                     if fq is not None else "concatenated bases + offsets in HBM",
        },
        "latency_ms": {"p50": float(np.percentile(times, 50) * 1e3), "p99": float(np.percentile(times, 99) * 1e3)},
        "phases_ms": ph,
        "graph": {kk: stats[kk] for kk in ("n_distinct", "n_k2", "n_overflow", "n_nodes", "n_branches", "n_chain_nodes",
                                           "n_unitigs", "n_cycle_nodes", "iterations", "capacity", "device_bytes",
                                           "peak_device_bytes")},
        "roofline": {
            "kernel": ("count phase: k_parse_records + k_split1 + k_split2 + k_region_build%s + k_spill_flush "
                       "(device events around the phase)" % (" (with the filter fused into its write-back)"
                                                              if fused else "")) if stats["count_by_regions"] else
                      "k_parse (parse + atomic k2 count)",
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": measured_traffic(args.config, args.scale, args.size_from, "count_phase")
                       if lf == 0.5 and stats["count_by_regions"] and fq is None else None,
            "traffic_note": "FETCH_SIZE x 2.00 + WRITE_SIZE x 1.00 (factors from a 4 GiB copy): bytes the L2 "
                            "caches fetched and wrote back, Infinity Cache (MALL) hits included, so not DRAM "
                            "bytes alone",
            "bytes_per_unit": "1 B per base + 128 B per k-mer instance (one 64-byte slot read and write)",
            "payload_only": {
                "achieved": payload_bytes / (pc_ms * 1e-3) / 1e9,
                "bytes_per_unit": "1 B per base + 32 B per k-mer instance (16-byte record read and write)",
            },
        },
        "roofline_filter": {
            "kernel": "filter phase (candidate bits or table scan, then candidate and spilled-slot passes)",
            "bound": "hbm",
            "achieved": filter_bytes(stats, k) / (ph["filter_ms"] * 1e-3) / 1e9,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": filter_bytes(stats, k) / (ph["filter_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "traffic": measured_traffic(args.config, args.scale, args.size_from, "filter_phase")
                       if lf == 0.5 and stats["count_by_regions"] and fq is None else None,
            "bytes_per_unit": "16 B per distinct k2-mer read (24 B for k > 31) + 32 B per kept k-mer written (48 B)",
        },
    }
    if fused:
        # the filter runs inside the count's region build: its work is timed
        # with the count, so the count + filter phases are measured together
        both = count_bytes(s.n_bases, n_inst) + filter_bytes(stats, k)
        both_ms = pc_ms + ph["filter_ms"]
        out["roofline_filter"] = {
            "kernel": "filter fused into k_region_build's write-back (local_insert_by_freqencies + classification "
                      "in LDS), then bruno_filter's remainder (list holes, k_filter_bits, k_filter_spilled): "
                      "count + filter phases together",
            "bound": "hbm",
            "achieved": both / (both_ms * 1e-3) / 1e9,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": both / (both_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "bytes_per_unit": "count (1 B per base + 128 B per k-mer instance) + filter (16 B per distinct k2-mer "
                              "+ 32 B per node; 24 B and 48 B for k > 31), over the two phases' time",
            "remainder_ms": ph["filter_ms"],
            "unfused": unfused,
        }
    if args.edge_outputs:
        out["edge_outputs"] = time_edge_outputs(job, bases, offs, s, fq)
    if tuned is not None:
        from bruno_amd.autotune import get_tuning
        out["tuning"] = {"load_factor": lf, "knobs": get_tuning(),
                         "source": "autotuned on this workload" if args.autotune else args.tuning}
    job.close()
    bases.free()
    offs.free()
    cp = copy_peak_gbs()
    for key in ("roofline", "roofline_filter"):
        out[key]["copy_peak"] = cp
        out[key]["frac_of_copy_peak"] = out[key]["achieved"] / cp
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
        out["cpu_baseline"] = cpu_baseline(cfg, args.cpu_seconds, edge_filter=args.edge_filter)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out))


def run_batched(args, free):
    """The config4 workload: batches vs one job per launch."""
    import numpy as np
    import bruno_amd
    from bruno_amd.configs import config4
    from bruno_amd.device import DeviceBuffer, synchronize
    from bruno_amd.pipeline import synth_batch_device, synth_reads_device
    from bruno_amd.planner import capacity, instances, plan_groups, short_first

    cfg = config4(scale=args.scale, n_jobs=args.n_jobs)
    k, th, jobs = cfg["k"], cfg["thresholds"], cfg["jobs"]
    lf = args.load_factor
    budget = args.budget_gb * 1e9 if args.budget_gb > 0 else 0.80 * free
    groups = short_first(plan_groups(jobs, k, budget, lf), jobs, k)
    log = lambda *a: print(*a, file=sys.stderr, flush=True)  # noqa: E731
    log("config4: %d jobs, %d launch groups, budget %.1f GB" % (len(jobs), len(groups), budget / 1e9))

    def run_group(idx):
        cfgs = [jobs[i] for i in idx]
        nb, nr = sum(c.n_bases for c in cfgs), sum(c.n_reads for c in cfgs)
        bases, offs = DeviceBuffer(nb), DeviceBuffer((nr + 1) * 8)
        first = synth_batch_device(cfgs, bases.ptr, offs.ptr)
        synchronize()
        t0 = time.perf_counter()
        bt = bruno_amd.GraphBatch(k, th, [capacity(c, k, lf) for c in cfgs])
        bt.parse_count_device(bases.ptr, offs.ptr, nr, first)
        bt.run(args.variant)
        synchronize()
        dt = time.perf_counter() - t0
        per_job = [bt.job(q).stats() for q in range(len(cfgs))]
        st, ph = bt.stats(), bt.phase_ms()
        bt.close()
        bases.free()
        offs.free()
        return dt, st, ph, per_job, nb

    def run_alone(i):
        c = jobs[i]
        bases, offs = DeviceBuffer(c.n_bases), DeviceBuffer((c.n_reads + 1) * 8)
        synth_reads_device(c, bases.ptr, offs.ptr)
        synchronize()
        t0 = time.perf_counter()
        job = bruno_amd.GraphJob(k, th, capacity=capacity(c, k, lf))
        job.parse_count_device(bases.ptr, offs.ptr, c.n_reads)
        job.run(args.variant)
        synchronize()
        dt = time.perf_counter() - t0
        st = job.stats()
        ph = job.phase_ms()
        job.close()
        bases.free()
        offs.free()
        return dt, st, ph

    # warm-up: the first two jobs, batched and alone
    for _ in range(args.warmup):
        run_group([0, 1])
        run_alone(0)

    passes = []
    for step in range(max(1, args.steps)):
        clock, lat, inst, pc_ms, pc_bytes, peak = 0.0, [], 0, 0.0, 0, 0
        results = {}
        for gi, idx in enumerate(groups):
            dt, st, ph, per_job, nb = run_group(idx)
            clock += dt
            inst += st["n_instances"]
            pc_ms += ph["parse_count_ms"]
            pc_bytes += count_bytes(nb, st["n_instances"])
            peak = max(peak, st["peak_device_bytes"])
            for q, i in enumerate(idx):
                lat.append(clock)
                results[i] = (per_job[q]["n_unitigs"], per_job[q]["n_unitig_nodes"], per_job[q]["iterations"])
            log("group %d: %d jobs, %.0f ms, %.3g k-mers, peak %.1f GB, phases %s"
                % (gi, len(idx), dt * 1e3, st["n_instances"], st["peak_device_bytes"] / 1e9,
                   json.dumps({a: round(b, 1) for a, b in ph.items()})))
        passes.append((clock, lat, inst, pc_ms, pc_bytes, peak, results))
    clock, lat, inst, pc_ms, pc_bytes, peak, results = min(passes, key=lambda p: p[0])

    clock1, lat1, inst1, same = 0.0, [], 0, True
    ph1 = {}
    # one job per launch, in the same order: fewest k-mer instances first
    for i in sorted(range(len(jobs)), key=lambda i: (instances(jobs[i], k), i)):
        dt, st, ph = run_alone(i)
        clock1 += dt
        inst1 += st["n_instances"]
        lat1.append(clock1)
        same &= results[i] == (st["n_unitigs"], st["n_unitig_nodes"], st["iterations"])
        for a, b in ph.items():
            ph1[a] = ph1.get(a, 0.0) + b
        if args.verbose:
            log("job %d: G %d L %d reads %d: %.1f ms %s" % (i, jobs[i].genome_len, jobs[i].read_len, jobs[i].n_reads,
                                                          dt * 1e3, json.dumps({a: round(b, 2) for a, b in ph.items()})))
    log("one job per launch: %.0f ms, phases %s" % (clock1 * 1e3, json.dumps({a: round(b, 1) for a, b in ph1.items()})))

    achieved = pc_bytes / (pc_ms * 1e-3) / 1e9
    G = sum(c.genome_len for c in jobs)
    out = {
        "metric": "k-mers/s (parse+count+filter+compact)",
        "value": inst / clock,
        "unit": "k-mers/s",
        "n_gpus": 1,
        "steps": max(1, args.steps),
        "warmup": args.warmup,
        "ms_per_step": clock * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic (seeded generator, reads generated in HBM)",
        "config": {
            "workload": "config4: %d independent jobs, genomes log-uniform 50 kbp-20 Mbp (total %.0f Mbp), "
                        "reads {100,150,250} bp, 20-60x, 1%% substitutions, k=31, -T -L 3, batched"
                        % (len(jobs), G / 1e6),
            "n_jobs": len(jobs), "launch_groups": len(groups), "k_mer_instances": inst,
            "bases": sum(c.n_bases for c in jobs), "load_factor": lf, "budget_gb": budget / 1e9,
        },
        "latency_ms": {"p50": float(np.percentile(lat, 50) * 1e3), "p99": float(np.percentile(lat, 99) * 1e3),
                       "note": "all jobs submitted at t=0; launch groups run fewest k-mer instances per job first; "
                               "a job completes with its launch group"},
        "one_job_per_launch": {"value": inst1 / clock1, "ms_total": clock1 * 1e3,
                               "latency_ms": {"p50": float(np.percentile(lat1, 50) * 1e3),
                                              "p99": float(np.percentile(lat1, 99) * 1e3)}},
        "speedup_vs_one_job_per_launch": clock1 / clock,
        "per_job_results_equal": bool(same and inst1 == inst),
        "peak_device_gb": peak / 1e9,
        "roofline": {
            "kernel": "count phase, all groups (a batch: k_parse, the atomic count; a large job alone: the region count)",
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
        },
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline_jobs(k, th, jobs, args.cpu_seconds)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out))


def serve(cfgs, ks, arrivals, th, lf, workers, max_batch, budget):
    from bruno_amd.planner import capacity
    from bruno_amd.server import Server
    srv = Server(hbm_budget=int(budget), n_workers=workers, max_batch_jobs=max_batch, reserve_pool=False)
    t0 = time.perf_counter()
    tickets = []
    for c, k, a in zip(cfgs, ks, arrivals):
        d = t0 + a - time.perf_counter()
        if d > 0:
            time.sleep(d)
        tickets.append(srv.submit_synth(k, th, c, capacity=capacity(c, k, lf)))
    infos = [srv.wait(t) for t in tickets]
    st = srv.stats()
    for t in tickets:
        srv.release(t)
    srv.close()
    return infos, st


def run_serving(args, free):
    import numpy as np
    from bruno_amd.pipeline import thresholds_L
    th = thresholds_L(3)
    from bruno_amd.configs import config5_jobs
    cfgs, ks, arrivals = config5_jobs(5, args.n_jobs, args.rate, args.scale)
    budget = 0.85 * free
    log = lambda *a: print(*a, file=sys.stderr, flush=True)  # noqa: E731
    runs = {}
    for name, workers, max_batch in (("continuous_batching", args.workers, 64), ("one_job_at_a_time", 1, 1)):
        infos, st = serve(cfgs, ks, arrivals, th, args.load_factor, workers, max_batch, budget)
        bad = [i for i in infos if i["status"] != 0]
        lat = np.array([i["end_s"] - i["submit_s"] for i in infos])
        inst = sum(i["n_instances"] for i in infos)
        span = max(i["end_s"] for i in infos) - min(i["submit_s"] for i in infos)
        runs[name] = {"value": inst / span, "span_s": span, "k_mer_instances": inst, "failed": len(bad),
                      "latency_ms": {"p50": float(np.percentile(lat, 50) * 1e3),
                                     "p99": float(np.percentile(lat, 99) * 1e3),
                                     "mean": float(lat.mean() * 1e3)},
                      "batches": st["batches"], "max_batch_jobs": st["max_batch_jobs"], "retries": st["retries"],
                      "peak_reserved_gb": st["peak_reserved"] / 1e9,
                      "unitigs": [i["n_unitigs"] for i in infos]}
        log("%s: %.3g k-mers/s over %.2f s, p50 %.1f ms p99 %.1f ms, %d batches (max %d jobs), %d retries"
            % (name, inst / span, span, runs[name]["latency_ms"]["p50"], runs[name]["latency_ms"]["p99"],
               st["batches"], st["max_batch_jobs"], st["retries"]))
    cb, seq = runs["continuous_batching"], runs["one_job_at_a_time"]
    same = cb.pop("unitigs") == seq.pop("unitigs")
    out = {
        "metric": "k-mers/s (parse+count+filter+compact)",
        "value": cb["value"],
        "unit": "k-mers/s",
        "n_gpus": 1,
        "steps": 1,
        "warmup": 0,
        "ms_per_step": cb["span_s"] * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic (seeded generator, reads generated in HBM per request)",
        "config": {
            "workload": "config5: %d jobs from the config4 distribution, k 31/63 (80/20), Poisson arrivals at "
                        "%.0f/s, -T -L 3, served with admission by HBM footprint" % (len(cfgs), args.rate),
            "n_jobs": len(cfgs), "rate_per_s": args.rate, "k63_jobs": int(sum(k == 63 for k in ks)),
            "workers": args.workers, "budget_gb": budget / 1e9, "load_factor": args.load_factor,
        },
        "latency_ms": cb["latency_ms"],
        "continuous_batching": cb,
        "one_job_at_a_time": seq,
        "per_job_results_equal": bool(same),
    }
    out["cpu_baseline"] = None
    print(json.dumps(out))


def cpu_baseline_jobs(k, th, jobs, seconds_target=15.0, threads=None):
    """The oracle on the first jobs of a job list, each job alone, until the
    sample's k-mer budget is used (the last job sampled by its first reads)."""
    import numpy as np
    import bruno_amd
    from bruno_amd.pipeline import SynthConfig
    from oracle.oracle import Oracle  # the checker, timed as the baseline
    threads = threads or min(16, os.cpu_count() or 1)
    budget = int(96_000_000 * seconds_target / 15.0)
    inst, dt, used = 0, 0.0, []
    for c in jobs:
        if inst >= budget:
            break
        per = max(1, c.read_len - k + 1)
        n = max(1, min(c.n_reads, (budget - inst) // per + 1))
        sample = SynthConfig(**{**c.__dict__, "n_reads": n})
        b, o = bruno_amd.synth_reads_host(sample)
        t0 = time.time()
        orc = Oracle(k, np.array(th, np.uint64), threads)
        orc.add_reads(b, o)
        orc.run(1)
        dt += time.time() - t0
        inst += orc.stat("n_instances")
        used.append("%d/%d" % (n, c.n_reads))
    return {"value": inst / dt, "unit": "k-mers/s", "cores": threads, "kind": "port",
            "sample": "jobs 0..%d of the same workload, each alone (reads used per job: %s), "
                      "%d k-mer instances, all phases, %.1f s" % (len(used) - 1, ", ".join(used), inst, dt),
            "host": host_cpu()}


if __name__ == "__main__":
    main()
