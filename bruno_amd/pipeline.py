"""Host-side driver of one de Bruijn job over the C ABI.

GraphJob keeps the reference driver's phase order
(test/test/compact_dbg_freq_construct_clean.cpp:431-840):

    parse_count  -> compute_biedge_freqencies / insert      (:318-328)
    filter       -> local_insert_by_freqencies              (:338)
    build_chains -> chainmap.extract_chains(idx)            (:549)
    list_rank    -> list_rank_min_update() / list_rank()    (:568-571)
    compact      -> to_ranked_chain_nodes + print_chain_string (:737-775)
"""
import ctypes
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import CHAIN_DTYPE, CHAIN_EDGE_DTYPE, NODE_DTYPE, check, lib

DEFAULT_THRESH = 65536


def thresholds_L(t):
    """-T -L t (compact_dbg_freq_construct_clean.cpp:940-945)."""
    return [0, DEFAULT_THRESH, min(int(t), DEFAULT_THRESH), DEFAULT_THRESH, 0, DEFAULT_THRESH]


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


class _Views:
    """Getters of one job's results; `_call(name, *args)` binds them to a
    whole job or to one job of a batch."""
    words = 1

    def _two_call(self, name, dtype=None, text=False):
        n = ctypes.c_uint64(0)
        self._call(name, None, ctypes.byref(n))
        if text:
            buf = ctypes.create_string_buffer(max(1, n.value))
            self._call(name, buf, ctypes.byref(n))
            return buf.raw[:n.value]
        out = np.zeros(n.value, dtype)
        if n.value:
            self._call(name, _ptr(out), ctypes.byref(n))
        return out[:n.value]

    def counts(self):
        """P1 table (before filter)."""
        return self._two_call("counts", NODE_DTYPE)

    def nodes(self):
        """Filtered graph (after filter)."""
        return self._two_call("nodes", NODE_DTYPE)

    def chain_nodes(self):
        return self._two_call("chain_nodes", CHAIN_DTYPE)

    def unitigs(self):
        return self._two_call("unitigs", text=True)

    def branch_nodes(self):
        """Branch nodes sorted by k-mer, with counts (after compact)."""
        return self._two_call("branch_nodes", NODE_DTYPE)

    def chain_edges(self):
        """Per unitig: end k-mers L, R, L's in-edge and R's out-edge counts,
        self-count sum / n_nodes / min / max (after compact)."""
        return self._two_call("chain_edges", CHAIN_EDGE_DTYPE)

    def branch_edges_text(self):
        """The .branch.edges bytes (print_graph_node lines)."""
        return format_branch_edges(self.branch_nodes(), self.k)

    def chain_edges_text(self):
        """The .chain.edges bytes (print_chain_frequencies lines)."""
        return format_chain_edges(self.chain_edges(), self.k)

    def branch_fasta(self):
        return self._two_call("branch_fasta", text=True)

    def branches(self):
        n = ctypes.c_uint64(0)
        self._call("branches", None, ctypes.byref(n))
        out = np.zeros(n.value * self.words, np.uint64)
        if n.value:
            self._call("branches", _ptr(out), ctypes.byref(n))
        return out.reshape(-1, self.words)

    def k2(self):
        n = ctypes.c_uint64(0)
        self._call("k2", None, None, None, ctypes.byref(n))
        km = np.zeros(n.value * self.words, np.uint64)
        e = np.zeros(n.value, np.uint8)
        c = np.zeros(n.value, np.uint32)
        if n.value:
            self._call("k2", _ptr(km), _ptr(e), _ptr(c), ctypes.byref(n))
        return km.reshape(-1, self.words), e, c


def kmer_text(words, k):
    """A k-mer (two 64-bit words, low first) as bases, first base most
    significant."""
    v = int(words[0]) | (int(words[1]) << 64)
    return "".join("ACGT"[(v >> (2 * (k - 1 - i))) & 3] for i in range(k))


def format_branch_edges(nodes, k):
    """print_graph_node lines (debruijn_graph_operations.hpp:58-78): k-mer,
    in A C G T, out A C G T, self."""
    out = []
    for r in nodes:
        c = [int(x) for x in r["counts"]]
        out.append("\t".join([kmer_text(r["kmer"], k)] + [str(x) for x in c[4:8] + c[0:4] + [c[8]]]) + "\n")
    return "".join(out).encode()


def format_chain_edges(edges, k):
    """print_chain_frequencies lines (compact_dbg_stats.hpp:1036-1055): L, R,
    L's in-edge counts, R's out-edge counts, mean, min, max, each followed by
    a tab; the mean is float32(sum) / float32(n_nodes) converted to the
    16-bit count type (truncated)."""
    out = []
    for r in edges:
        mean = int(np.float32(int(r["sum"])) / np.float32(int(r["n_nodes"])))
        vals = [int(x) for x in r["in"]] + [int(x) for x in r["out"]] + [mean, int(r["min"]), int(r["max"])]
        out.append(kmer_text(r["left"], k) + "\t" + kmer_text(r["right"], k) + "\t" +
                   "".join("%d\t" % v for v in vals) + "\n")
    return "".join(out).encode()


def job_desc(k, thresholds=None, edge_filter=False):
    """bruno_job_desc for k and the -T thresholds (None: unthresholded).
    edge_filter selects the block-construct driver's filter: an edge is kept
    when its canonical k+1-mer occurs lower <= count < upper times over the
    reads, lower/upper = thresholds[2]/[3]
    (compact_debruijn_graph_block_construct.cpp:415-480)."""
    if edge_filter and thresholds is None:
        raise ValueError("the edge-frequency filter needs thresholds")
    d = _lib.JobDesc()
    d.k = int(k)
    d.thresholding = 0 if thresholds is None else (2 if edge_filter else 1)
    th = thresholds if thresholds is not None else [DEFAULT_THRESH] * 6
    for i in range(6):
        d.thresholds[i] = int(th[i])
    return d


class GraphJob(_Views):
    def __init__(self, k, thresholds=None, capacity=0, load_factor=0.0, stream=None, edge_filter=False):
        d = job_desc(k, thresholds, edge_filter)
        d.capacity = int(capacity)
        d.load_factor = float(load_factor)
        d.stream = stream
        h = ctypes.c_void_p()
        check(lib().bruno_job_create(ctypes.byref(d), ctypes.byref(h)), "bruno_job_create")
        self.h = h
        self.k = int(k)
        self.words = 1 if k <= 31 else 2

    def close(self):
        if getattr(self, "h", None):
            lib().bruno_job_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def reset(self):
        """Clear the job for reuse, keeping its hash-table allocation."""
        check(lib().bruno_job_reset(self.h), "bruno_job_reset")

    def set_recycle(self, on=True):
        """Clear the table in the background once the chains are built (the
        next reset then has nothing to clear; nodes() is unavailable after
        build_chains)."""
        check(lib().bruno_job_set_recycle(self.h, 1 if on else 0), "bruno_job_set_recycle")

    def set_fused_filter(self, on=True):
        """Filter fused into the job's first count (bruno_job_set_fused_filter):
        the region build also filters and classifies, bruno_filter finishes
        the rest; the job then takes no further batch before a reset."""
        check(lib().bruno_job_set_fused_filter(self.h, 1 if on else 0), "bruno_job_set_fused_filter")

    # ---- phases ------------------------------------------------------
    def parse_count(self, bases, offsets):
        """Host arrays (bytes / uint8 and uint64 offsets, n_reads+1)."""
        b = np.frombuffer(bases, np.uint8) if isinstance(bases, (bytes, bytearray)) else np.ascontiguousarray(bases, np.uint8)
        o = np.ascontiguousarray(offsets, np.uint64)
        check(lib().bruno_parse_count_host(self.h, _ptr(b), _ptr(o), o.size - 1), "bruno_parse_count_host")

    def parse_count_device(self, bases_ptr, offsets_ptr, n_reads):
        check(lib().bruno_parse_count(self.h, ctypes.c_void_p(bases_ptr), ctypes.c_void_p(offsets_ptr),
                                      int(n_reads)), "bruno_parse_count")

    def parse_count_fastq(self, data):
        """A whole FASTQ file's bytes (host), split into reads on the device."""
        b = np.frombuffer(data, np.uint8) if isinstance(data, (bytes, bytearray)) else np.ascontiguousarray(data, np.uint8)
        check(lib().bruno_parse_count_fastq_host(self.h, _ptr(b) if b.size else None, b.size),
              "bruno_parse_count_fastq_host")

    def parse_count_fastq_device(self, ptr, nbytes):
        check(lib().bruno_parse_count_fastq(self.h, ctypes.c_void_p(ptr), int(nbytes)), "bruno_parse_count_fastq")

    def filter(self):
        check(lib().bruno_filter(self.h), "bruno_filter")

    def build_chains(self):
        check(lib().bruno_build_chains(self.h), "bruno_build_chains")

    def list_rank(self, variant=1):
        r = ctypes.c_uint32()
        check(lib().bruno_list_rank(self.h, int(variant), ctypes.byref(r)), "bruno_list_rank")
        return r.value

    def compact(self):
        check(lib().bruno_compact(self.h), "bruno_compact")

    def run(self, variant=1):
        check(lib().bruno_run_pipeline(self.h, int(variant)), "bruno_run_pipeline")

    def clean(self, variant=1, k1_lo=65536):
        """Deadend / bubble cleaning after run(); the getters then return
        the cleaned graph's outputs.  k1_lo: the driver's lower k+1-mer
        threshold (65536 without -L)."""
        st = _lib.CleanStats()
        check(lib().bruno_clean(self.h, int(variant), int(k1_lo), ctypes.byref(st)), "bruno_clean")
        return {f: int(getattr(st, f)) for f, _ in _lib.CleanStats._fields_}

    # ---- getters -----------------------------------------------------
    def _call(self, name, *args):
        fn = getattr(lib(), "bruno_get_" + name)
        return check(fn(self.h, *args), "bruno_get_" + name)

    def stats(self):
        s = _lib.Stats()
        check(lib().bruno_get_stats(self.h, ctypes.byref(s)), "bruno_get_stats")
        return {f: int(getattr(s, f)) for f in _lib.STAT_FIELDS}

    def phase_ms(self):
        p = _lib.PhaseMs()
        check(lib().bruno_get_phase_ms(self.h, ctypes.byref(p)), "bruno_get_phase_ms")
        return {f: float(getattr(p, f)) for f, _ in _lib.PhaseMs._fields_}


class BatchJobView(_Views):
    """Results of job `idx` of a GraphBatch."""

    def __init__(self, batch, idx):
        self.batch = batch
        self.idx = int(idx)
        self.k = batch.k
        self.words = batch.words

    def _call(self, name, *args):
        fn = getattr(lib(), "bruno_batch_get_" + name)
        return check(fn(self.batch.h, self.idx, *args), "bruno_batch_get_" + name)

    def stats(self):
        s = _lib.Stats()
        check(lib().bruno_batch_get_stats(self.batch.h, self.idx, ctypes.byref(s)), "bruno_batch_get_stats")
        return {f: int(getattr(s, f)) for f in _lib.STAT_FIELDS}


class GraphBatch(GraphJob):
    """Many independent jobs (same k and thresholds) through one set of
    launches; job i owns capacities[i] hash slots.  Results per job via
    job(i), bit-identical to running each job alone."""

    def __init__(self, k, thresholds, capacities, stream=None, edge_filter=False):
        d = job_desc(k, thresholds, edge_filter)
        d.stream = stream
        caps = np.ascontiguousarray(capacities, np.uint64)
        h = ctypes.c_void_p()
        check(lib().bruno_batch_create(ctypes.byref(d), int(caps.size), _ptr(caps), ctypes.byref(h)),
              "bruno_batch_create")
        self.h = h
        self.k = int(k)
        self.words = 1 if k <= 31 else 2
        self.n_jobs = int(caps.size)

    def parse_count_device(self, bases_ptr, offsets_ptr, n_reads, job_first_read):
        f = np.ascontiguousarray(job_first_read, np.uint64)
        if f.size != self.n_jobs + 1:
            raise ValueError("job_first_read needs n_jobs + 1 entries")
        check(lib().bruno_batch_parse_count(self.h, ctypes.c_void_p(bases_ptr), ctypes.c_void_p(offsets_ptr),
                                            int(n_reads), _ptr(f)), "bruno_batch_parse_count")

    def parse_count(self, bases, offsets, job_first_read):
        """Host arrays: copied to device memory, then parsed as one batch."""
        from .device import DeviceBuffer
        b = np.frombuffer(bases, np.uint8) if isinstance(bases, (bytes, bytearray)) else np.ascontiguousarray(bases, np.uint8)
        o = np.ascontiguousarray(offsets, np.uint64)
        n = o.size - 1
        db = DeviceBuffer.from_numpy(b[int(o[0]):int(o[-1])])
        do = DeviceBuffer.from_numpy(o - o[0])
        try:
            self.parse_count_device(db.ptr, do.ptr, n, job_first_read)
        finally:
            db.free()
            do.free()

    def job(self, idx):
        if not 0 <= idx < self.n_jobs:
            raise IndexError(idx)
        return BatchJobView(self, idx)


@dataclass
class SynthConfig:
    seed: int
    genome_len: int
    n_reads: int
    read_len: int
    sub_rate: float
    indel_rate: float = 0.0
    n_families: int = 0
    family_min_len: int = 0
    family_max_len: int = 0
    repeat_fraction: float = 0.0
    family_divergence: float = 0.0

    def desc(self):
        d = _lib.SynthDesc()
        for f in ("seed", "genome_len", "n_families", "family_min_len", "family_max_len",
                  "repeat_fraction", "family_divergence", "n_reads", "read_len", "sub_rate",
                  "indel_rate"):
            setattr(d, f, getattr(self, f))
        return d

    @property
    def n_bases(self):
        return self.n_reads * self.read_len


def synth_reads_host(cfg):
    """Reads of a synthetic configuration on the host (bit-identical to the
    device generator)."""
    b = np.empty(cfg.n_bases, np.uint8)
    o = np.empty(cfg.n_reads + 1, np.uint64)
    d = cfg.desc()
    check(lib().bruno_synth_reads_host(ctypes.byref(d), _ptr(b), _ptr(o)), "bruno_synth_reads_host")
    return b, o


def synth_reads_device(cfg, bases_ptr, offsets_ptr, stream=None):
    d = cfg.desc()
    check(lib().bruno_synth_reads_device(ctypes.byref(d), ctypes.c_void_p(bases_ptr),
                                         ctypes.c_void_p(offsets_ptr), stream), "bruno_synth_reads_device")


def _descs(cfgs):
    arr = (_lib.SynthDesc * max(1, len(cfgs)))()
    for i, c in enumerate(cfgs):
        arr[i] = c.desc()
    return arr


def synth_batch_host(cfgs):
    """Reads of several synthetic workloads back to back (one batch), with
    global offsets and the job_first_read array of the batch."""
    nb = sum(c.n_bases for c in cfgs)
    nr = sum(c.n_reads for c in cfgs)
    b = np.empty(max(1, nb), np.uint8)
    o = np.empty(nr + 1, np.uint64)
    check(lib().bruno_synth_batch_host(_descs(cfgs), len(cfgs), _ptr(b), _ptr(o)), "bruno_synth_batch_host")
    return b[:nb], o, job_first_read(cfgs)


def synth_batch_device(cfgs, bases_ptr, offsets_ptr, stream=None):
    check(lib().bruno_synth_batch_device(_descs(cfgs), len(cfgs), ctypes.c_void_p(bases_ptr),
                                         ctypes.c_void_p(offsets_ptr), stream), "bruno_synth_batch_device")
    return job_first_read(cfgs)


def job_first_read(cfgs):
    f = np.zeros(len(cfgs) + 1, np.uint64)
    f[1:] = np.cumsum([c.n_reads for c in cfgs])
    return f


def fastq_split_device(fastq_ptr, nbytes, stream=None, noalloc=False):
    """Split FASTQ bytes in device memory into (bases, offsets) host arrays
    (through device buffers; for tests).  noalloc: through
    bruno_fastq_split_noalloc (the output buffers are still the caller's)."""
    from .device import DeviceBuffer
    L = lib()
    fn, name = ((L.bruno_fastq_split_noalloc, "bruno_fastq_split_noalloc") if noalloc
                else (L.bruno_fastq_split, "bruno_fastq_split"))
    nr, nb = ctypes.c_uint64(0), ctypes.c_uint64(0)
    check(fn(ctypes.c_void_p(fastq_ptr), int(nbytes), None, None, ctypes.byref(nr), ctypes.byref(nb), stream), name)
    db, do = DeviceBuffer(nb.value + 1), DeviceBuffer((nr.value + 1) * 8)
    try:
        check(fn(ctypes.c_void_p(fastq_ptr), int(nbytes), ctypes.c_void_p(db.ptr), ctypes.c_void_p(do.ptr),
                 ctypes.byref(nr), ctypes.byref(nb), stream), name)
        return db.to_numpy(np.uint8, nb.value), do.to_numpy(np.uint64, nr.value + 1)
    finally:
        db.free()
        do.free()


def estimate_distinct(k, bases_ptr, offsets_ptr, n_reads, load_factor=0.5, stream=None):
    """(estimated distinct canonical k-mers, hash slots needed) of reads in
    device memory, from the library's sketch."""
    est, cap = ctypes.c_double(0), ctypes.c_uint64(0)
    check(lib().bruno_estimate_distinct(int(k), ctypes.c_void_p(bases_ptr), ctypes.c_void_p(offsets_ptr),
                                        int(n_reads), ctypes.byref(est), ctypes.byref(cap), float(load_factor),
                                        stream), "bruno_estimate_distinct")
    return est.value, cap.value
