"""Autotuner for the table load factor and the launch parameters.

The knobs change speed only, never results:
  load factor   hash slots per distinct k-mer (count probes vs. table size,
                which the filter scan and the table reset stream);
  region_slots  region size of a region-layout table (k <= 31 plain jobs);
  count_groups  sub-bucket groups of the region count's split / build
                pipeline (0: automatic);
  parse_block   workgroup size of the atomic parse + count kernel;
  parse_grid    waves of the atomic parse + count launch;
  scan_grid     workgroups of the filter's table scan (atomic-count tables);
  ruler_sample  list-ranking ruler density (walk length vs. contracted size).
The region-count path has no launch knobs of the atomic kernel, so those
trials run only when the workload counts with atomics.

tune() runs coordinate descent on the caller's own reads (already in device
memory): each knob in turn, the others fixed at their best so far, each
candidate timed over `reps` runs of the phase it affects (device events on
the job's stream), best-of.  The result is applied with bruno_set_tuning and
can be saved as JSON and re-applied with apply().
"""
import ctypes
import json

from . import _lib
from ._lib import check, lib

LOAD_FACTORS = (0.4, 0.5, 0.6, 0.7, 0.8)
PARSE_BLOCKS = (128, 256, 512)
PARSE_GRIDS = (16384, 32768, 65536, 131072, 262144)
SCAN_GRIDS = (4096, 8192, 16384, 32768)
RULERS = (16, 32, 64, 128, 256)
COUNT_GROUPS = (0, 1, 4, 8, 16)
REGION_SLOTS = (2560, 1280)


def get_tuning():
    t = _lib.Tuning()
    check(lib().bruno_get_tuning(ctypes.byref(t)), "bruno_get_tuning")
    return {f: int(getattr(t, f)) for f, _ in _lib.Tuning._fields_}


def set_tuning(**kw):
    cur = get_tuning()
    cur.update({k: int(v) for k, v in kw.items() if k in cur})
    t = _lib.Tuning(**cur)
    check(lib().bruno_set_tuning(ctypes.byref(t)), "bruno_set_tuning")
    return cur


def apply(path):
    """Apply a saved tuning; returns it (load_factor included)."""
    with open(path) as f:
        res = json.load(f)
    set_tuning(**res["knobs"])
    return res


def tune(k, thresholds, bases_ptr, offsets_ptr, n_reads, capacity_for_lf, reps=2, log=None, save=None):
    """capacity_for_lf(lf) -> hash slots for the workload at load factor lf."""
    from .pipeline import GraphJob
    from .device import synchronize
    say = log or (lambda *a: None)
    trials = []

    regions = {"on": False}

    def run(lf, phases):
        job = GraphJob(k, thresholds, capacity=capacity_for_lf(lf))
        best = None
        try:
            for _ in range(reps):
                job.reset()
                job.parse_count_device(bases_ptr, offsets_ptr, n_reads)
                if phases != ("parse_count_ms",):
                    job.run(1)
                synchronize()
                ph = job.phase_ms()
                v = sum(ph[p] for p in phases) if phases else sum(ph.values())
                best = v if best is None else min(best, v)
            regions["on"] = bool(job.stats()["count_by_regions"])
        finally:
            job.close()
        return best

    # 1. load factor: the whole step (the reset streams the table too)
    res = {}
    for lf in LOAD_FACTORS:
        res[lf] = run(lf, ())
        trials.append({"knob": "load_factor", "value": lf, "ms": res[lf]})
        say("load factor %.2f: %.1f ms" % (lf, res[lf]))
    lf = min(res, key=res.get)
    run(lf, ("parse_count_ms",))
    if regions["on"]:
        # 2a. region size of the region count (the table layout)
        res = {}
        for rs in REGION_SLOTS:
            set_tuning(region_slots=rs)
            res[rs] = run(lf, ("parse_count_ms",))
            trials.append({"knob": "region_slots", "value": rs, "ms": res[rs]})
            say("region slots %d: %.1f ms" % (rs, res[rs]))
        set_tuning(region_slots=min(res, key=res.get))
        res = {}
        for cg in COUNT_GROUPS:
            set_tuning(count_groups=cg)
            res[cg] = run(lf, ("parse_count_ms",))
            trials.append({"knob": "count_groups", "value": cg, "ms": res[cg]})
            say("count groups %d: %.1f ms" % (cg, res[cg]))
        set_tuning(count_groups=min(res, key=res.get))
    else:
        # 2b. launch shape of the atomic count
        res = {}
        for b in PARSE_BLOCKS:
            for g in PARSE_GRIDS:
                set_tuning(parse_block=b, parse_grid=g)
                res[(b, g)] = run(lf, ("parse_count_ms",))
                trials.append({"knob": "parse", "value": [b, g], "ms": res[(b, g)]})
                say("parse block %d grid %d: %.1f ms" % (b, g, res[(b, g)]))
        b, g = min(res, key=res.get)
        set_tuning(parse_block=b, parse_grid=g)
        # 3. filter scan grid (a region count leaves the scan's bits instead)
        res = {}
        for sg in SCAN_GRIDS:
            set_tuning(scan_grid=sg)
            res[sg] = run(lf, ("filter_ms",))
            trials.append({"knob": "scan_grid", "value": sg, "ms": res[sg]})
            say("scan grid %d: %.1f ms" % (sg, res[sg]))
        set_tuning(scan_grid=min(res, key=res.get))
    # 4. list-ranking rulers
    res = {}
    for r in RULERS:
        set_tuning(ruler_sample=r)
        res[r] = run(lf, ("list_rank_ms",))
        trials.append({"knob": "ruler_sample", "value": r, "ms": res[r]})
        say("ruler sample %d: %.1f ms" % (r, res[r]))
    set_tuning(ruler_sample=min(res, key=res.get))
    out = {"load_factor": lf, "knobs": get_tuning(), "trials": trials}
    if save:
        with open(save, "w") as f:
            json.dump(out, f, indent=1)
    return out
