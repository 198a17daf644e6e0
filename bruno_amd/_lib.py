"""ctypes binding of libbruno_hip.so (the C ABI in include/bruno_hip.h).

The library is the product: there is no CPU fallback.  If it is missing or
fails to load, every entry point raises.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libbruno_hip.so")

OK, BAD_INPUT, OOM, CAPACITY, INTERNAL, STATE, NO_DEVICE = range(7)


class JobDesc(ctypes.Structure):
    _fields_ = [("k", ctypes.c_int), ("thresholding", ctypes.c_int),
                ("thresholds", ctypes.c_uint64 * 6), ("capacity", ctypes.c_uint64),
                ("load_factor", ctypes.c_double), ("stream", ctypes.c_void_p)]


class PhaseMs(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in
                ("parse_count_ms", "filter_ms", "chains_ms", "list_rank_ms", "output_ms")]


STAT_FIELDS = ("n_reads", "n_reads_dropped", "n_instances", "n_distinct", "n_k2", "n_overflow",
               "n_nodes", "n_branches", "n_chain_nodes", "n_cycle_nodes", "n_isolated",
               "n_unitig_nodes", "n_unitigs", "iterations", "n_conflicts", "capacity",
               "device_bytes", "peak_device_bytes", "output_by_chains", "count_by_regions",
               "filter_fused")


class Stats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint64) for n in STAT_FIELDS]


class CleanStats(ctypes.Structure):
    _fields_ = [("rounds", ctypes.c_uint32), ("deadends", ctypes.c_uint64), ("bubble_chains", ctypes.c_uint64),
                ("edges_erased", ctypes.c_uint64)]


class SynthDesc(ctypes.Structure):
    _fields_ = [("seed", ctypes.c_uint64), ("genome_len", ctypes.c_uint64),
                ("n_families", ctypes.c_uint32), ("family_min_len", ctypes.c_uint32),
                ("family_max_len", ctypes.c_uint32), ("repeat_fraction", ctypes.c_double),
                ("family_divergence", ctypes.c_double), ("n_reads", ctypes.c_uint64),
                ("read_len", ctypes.c_uint32), ("sub_rate", ctypes.c_double),
                ("indel_rate", ctypes.c_double)]


class Tuning(ctypes.Structure):
    _fields_ = [("parse_block", ctypes.c_uint32), ("parse_grid", ctypes.c_uint32),
                ("scan_grid", ctypes.c_uint32), ("ruler_sample", ctypes.c_uint32),
                ("count_regions", ctypes.c_uint32), ("region_slots", ctypes.c_uint32),
                ("fuse_list_cap", ctypes.c_uint32), ("count_groups", ctypes.c_uint32)]


class ServerDesc(ctypes.Structure):
    _fields_ = [("hbm_budget", ctypes.c_uint64), ("n_workers", ctypes.c_uint32),
                ("max_batch_jobs", ctypes.c_uint32), ("variant", ctypes.c_int), ("reserve_pool", ctypes.c_int)]


class Request(ctypes.Structure):
    _fields_ = [("job", JobDesc), ("bases", ctypes.c_void_p), ("offsets", ctypes.c_void_p),
                ("n_reads", ctypes.c_uint64), ("synth", ctypes.POINTER(SynthDesc))]


RESULT_FIELDS = [("status", ctypes.c_int32), ("batch_jobs", ctypes.c_uint32), ("batch_id", ctypes.c_uint64),
                 ("n_reads", ctypes.c_uint64), ("n_instances", ctypes.c_uint64), ("n_unitigs", ctypes.c_uint64),
                 ("n_unitig_nodes", ctypes.c_uint64), ("n_branches", ctypes.c_uint64),
                 ("unitig_bytes", ctypes.c_uint64), ("branch_bytes", ctypes.c_uint64),
                 ("iterations", ctypes.c_uint32), ("attempts", ctypes.c_uint32),
                 ("submit_s", ctypes.c_double), ("start_s", ctypes.c_double), ("end_s", ctypes.c_double)]


class ResultInfo(ctypes.Structure):
    _fields_ = RESULT_FIELDS


SERVER_STAT_FIELDS = ("submitted", "jobs_done", "jobs_failed", "batches", "retries", "max_batch_jobs",
                      "queued", "reserved_now", "peak_reserved", "hbm_budget")


class ServerStats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint64) for n in SERVER_STAT_FIELDS]


NODE_DTYPE = np.dtype([("kmer", "<u8", (2,)), ("counts", "<u2", (9,)), ("pad", "<u2", (3,))])
CHAIN_DTYPE = np.dtype([("k", "<u8", (2,)), ("rep", "<u8", (2,)), ("dist", "<i4"), ("pad", "<i4")])
CHAIN_EDGE_DTYPE = np.dtype([("left", "<u8", (2,)), ("right", "<u8", (2,)), ("sum", "<u8"), ("n_nodes", "<u4"),
                             ("in", "<u2", (4,)), ("out", "<u2", (4,)), ("min", "<u2"), ("max", "<u2")])
assert NODE_DTYPE.itemsize == 40 and CHAIN_DTYPE.itemsize == 40 and CHAIN_EDGE_DTYPE.itemsize == 64

# every entry point declared in include/bruno_hip.h
EXPORTS = (
    "bruno_status_string", "bruno_job_create", "bruno_job_destroy", "bruno_job_reset", "bruno_parse_count",
    "bruno_parse_count_host", "bruno_filter", "bruno_build_chains", "bruno_list_rank",
    "bruno_compact", "bruno_run_pipeline", "bruno_get_stats", "bruno_get_phase_ms",
    "bruno_get_counts", "bruno_get_k2", "bruno_get_nodes", "bruno_get_branches",
    "bruno_get_chain_nodes", "bruno_get_unitigs", "bruno_get_branch_fasta",
    "bruno_parse_records", "bruno_synth_reads_device", "bruno_synth_reads_host",
    "bruno_dev_alloc", "bruno_dev_free", "bruno_dev_copy", "bruno_dev_sync", "bruno_dev_mem_info",
    "bruno_batch_create", "bruno_batch_parse_count", "bruno_batch_get_stats", "bruno_batch_get_counts",
    "bruno_batch_get_k2", "bruno_batch_get_nodes", "bruno_batch_get_branches", "bruno_batch_get_chain_nodes",
    "bruno_batch_get_unitigs", "bruno_batch_get_branch_fasta", "bruno_synth_batch_device",
    "bruno_synth_batch_host", "bruno_pool_reserve", "bruno_pool_trim", "bruno_pool_stats",
    "bruno_server_create", "bruno_server_submit", "bruno_server_poll", "bruno_server_wait",
    "bruno_server_get_unitigs", "bruno_server_get_branch_fasta", "bruno_server_release",
    "bruno_server_get_stats", "bruno_server_destroy", "bruno_set_tuning", "bruno_get_tuning",
    "bruno_fastq_split", "bruno_parse_count_fastq", "bruno_parse_count_fastq_host", "bruno_estimate_distinct",
    "bruno_job_set_recycle", "bruno_job_set_fused_filter", "bruno_get_branch_nodes", "bruno_get_chain_edges", "bruno_batch_get_branch_nodes",
    "bruno_batch_get_chain_edges", "bruno_fastq_split_noalloc", "bruno_clean",
)

_LIB = None


class BrunoError(RuntimeError):
    def __init__(self, code, what):
        self.code = code
        super().__init__("%s: %s (status %d)" % (what, status_string(code), code))


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libbruno_hip.so is not built (run __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        vp, u64, i = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int
        for name in EXPORTS:
            getattr(L, name).restype = ctypes.c_int
        L.bruno_status_string.restype = ctypes.c_char_p
        L.bruno_status_string.argtypes = [i]
        L.bruno_job_create.argtypes = [ctypes.POINTER(JobDesc), ctypes.POINTER(vp)]
        L.bruno_job_destroy.argtypes = [vp]
        L.bruno_job_reset.argtypes = [vp]
        L.bruno_parse_count.argtypes = [vp, vp, vp, u64]
        L.bruno_parse_count_host.argtypes = [vp, vp, vp, u64]
        for n in ("bruno_filter", "bruno_build_chains", "bruno_compact"):
            getattr(L, n).argtypes = [vp]
        L.bruno_run_pipeline.argtypes = [vp, i]
        L.bruno_clean.argtypes = [vp, i, u64, ctypes.POINTER(CleanStats)]
        L.bruno_list_rank.argtypes = [vp, i, ctypes.POINTER(ctypes.c_uint32)]
        L.bruno_get_stats.argtypes = [vp, ctypes.POINTER(Stats)]
        L.bruno_get_phase_ms.argtypes = [vp, ctypes.POINTER(PhaseMs)]
        u64p = ctypes.POINTER(ctypes.c_uint64)
        for n in ("bruno_get_counts", "bruno_get_nodes", "bruno_get_chain_nodes",
                  "bruno_get_unitigs", "bruno_get_branch_fasta", "bruno_get_branches"):
            getattr(L, n).argtypes = [vp, vp, u64p]
        L.bruno_get_k2.argtypes = [vp, vp, vp, vp, u64p]
        L.bruno_parse_records.argtypes = [i, vp, vp, u64, vp, vp, u64p, vp]
        L.bruno_synth_reads_device.argtypes = [ctypes.POINTER(SynthDesc), vp, vp, vp]
        L.bruno_synth_reads_host.argtypes = [ctypes.POINTER(SynthDesc), vp, vp]
        u32 = ctypes.c_uint32
        L.bruno_batch_create.argtypes = [ctypes.POINTER(JobDesc), u32, vp, ctypes.POINTER(vp)]
        L.bruno_batch_parse_count.argtypes = [vp, vp, vp, u64, vp]
        L.bruno_batch_get_stats.argtypes = [vp, u32, ctypes.POINTER(Stats)]
        for n in ("bruno_batch_get_counts", "bruno_batch_get_nodes", "bruno_batch_get_chain_nodes",
                  "bruno_batch_get_unitigs", "bruno_batch_get_branch_fasta", "bruno_batch_get_branches"):
            getattr(L, n).argtypes = [vp, u32, vp, u64p]
        L.bruno_batch_get_k2.argtypes = [vp, u32, vp, vp, vp, u64p]
        L.bruno_synth_batch_device.argtypes = [vp, u32, vp, vp, vp]
        L.bruno_synth_batch_host.argtypes = [vp, u32, vp, vp]
        L.bruno_pool_reserve.argtypes = [u64]
        L.bruno_pool_trim.argtypes = []
        L.bruno_pool_stats.argtypes = [u64p, u64p, u64p]
        L.bruno_server_create.argtypes = [ctypes.POINTER(ServerDesc), ctypes.POINTER(vp)]
        L.bruno_server_submit.argtypes = [vp, ctypes.POINTER(Request), u64p]
        L.bruno_server_poll.argtypes = [vp, u64, ctypes.POINTER(ctypes.c_int)]
        L.bruno_server_wait.argtypes = [vp, u64, ctypes.POINTER(ResultInfo)]
        L.bruno_server_get_unitigs.argtypes = [vp, u64, vp, u64p]
        L.bruno_server_get_branch_fasta.argtypes = [vp, u64, vp, u64p]
        L.bruno_server_release.argtypes = [vp, u64]
        L.bruno_server_get_stats.argtypes = [vp, ctypes.POINTER(ServerStats)]
        L.bruno_server_destroy.argtypes = [vp]
        L.bruno_set_tuning.argtypes = [ctypes.POINTER(Tuning)]
        L.bruno_fastq_split.argtypes = [vp, u64, vp, vp, u64p, u64p, vp]
        L.bruno_fastq_split_noalloc.argtypes = [vp, u64, vp, vp, u64p, u64p, vp]
        L.bruno_parse_count_fastq.argtypes = [vp, vp, u64]
        L.bruno_parse_count_fastq_host.argtypes = [vp, vp, u64]
        L.bruno_job_set_recycle.argtypes = [vp, i]
        L.bruno_job_set_fused_filter.argtypes = [vp, i]
        L.bruno_estimate_distinct.argtypes = [i, vp, vp, u64, ctypes.POINTER(ctypes.c_double), u64p,
                                              ctypes.c_double, vp]
        L.bruno_get_tuning.argtypes = [ctypes.POINTER(Tuning)]
        L.bruno_dev_alloc.argtypes = [u64, ctypes.POINTER(vp)]
        L.bruno_dev_free.argtypes = [vp]
        L.bruno_dev_copy.argtypes = [vp, vp, u64, i]
        L.bruno_dev_sync.argtypes = []
        L.bruno_dev_mem_info.argtypes = [ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64)]
        _LIB = L
    return _LIB


def status_string(code):
    try:
        return lib().bruno_status_string(code).decode()
    except Exception:  # pragma: no cover - only when the library is absent
        return "status %d" % code


def check(code, what):
    if code != OK:
        raise BrunoError(code, what)
