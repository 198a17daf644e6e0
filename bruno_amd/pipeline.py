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
from ._lib import CHAIN_DTYPE, NODE_DTYPE, check, lib

DEFAULT_THRESH = 65536


def thresholds_L(t):
    """-T -L t (compact_dbg_freq_construct_clean.cpp:940-945)."""
    return [0, DEFAULT_THRESH, min(int(t), DEFAULT_THRESH), DEFAULT_THRESH, 0, DEFAULT_THRESH]


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


class GraphJob:
    def __init__(self, k, thresholds=None, capacity=0, load_factor=0.0, stream=None):
        d = _lib.JobDesc()
        d.k = int(k)
        d.thresholding = 0 if thresholds is None else 1
        th = thresholds if thresholds is not None else [DEFAULT_THRESH] * 6
        for i in range(6):
            d.thresholds[i] = int(th[i])
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

    # ---- phases ------------------------------------------------------
    def parse_count(self, bases, offsets):
        """Host arrays (bytes / uint8 and uint64 offsets, n_reads+1)."""
        b = np.frombuffer(bases, np.uint8) if isinstance(bases, (bytes, bytearray)) else np.ascontiguousarray(bases, np.uint8)
        o = np.ascontiguousarray(offsets, np.uint64)
        check(lib().bruno_parse_count_host(self.h, _ptr(b), _ptr(o), o.size - 1), "bruno_parse_count_host")

    def parse_count_device(self, bases_ptr, offsets_ptr, n_reads):
        check(lib().bruno_parse_count(self.h, ctypes.c_void_p(bases_ptr), ctypes.c_void_p(offsets_ptr),
                                      int(n_reads)), "bruno_parse_count")

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

    # ---- getters -----------------------------------------------------
    def _two_call(self, fn, dtype=None, text=False):
        n = ctypes.c_uint64(0)
        check(fn(self.h, None, ctypes.byref(n)), fn.__name__)
        if text:
            buf = ctypes.create_string_buffer(max(1, n.value))
            check(fn(self.h, buf, ctypes.byref(n)), fn.__name__)
            return buf.raw[:n.value]
        out = np.zeros(n.value, dtype)
        if n.value:
            check(fn(self.h, _ptr(out), ctypes.byref(n)), fn.__name__)
        return out

    def counts(self):
        """P1 table (before filter)."""
        return self._two_call(lib().bruno_get_counts, NODE_DTYPE)

    def nodes(self):
        """Filtered graph (after filter)."""
        return self._two_call(lib().bruno_get_nodes, NODE_DTYPE)

    def chain_nodes(self):
        return self._two_call(lib().bruno_get_chain_nodes, CHAIN_DTYPE)

    def unitigs(self):
        return self._two_call(lib().bruno_get_unitigs, text=True)

    def branch_fasta(self):
        return self._two_call(lib().bruno_get_branch_fasta, text=True)

    def branches(self):
        n = ctypes.c_uint64(0)
        check(lib().bruno_get_branches(self.h, None, ctypes.byref(n)), "bruno_get_branches")
        out = np.zeros(n.value * self.words, np.uint64)
        if n.value:
            check(lib().bruno_get_branches(self.h, _ptr(out), ctypes.byref(n)), "bruno_get_branches")
        return out.reshape(-1, self.words)

    def k2(self):
        n = ctypes.c_uint64(0)
        L = lib()
        check(L.bruno_get_k2(self.h, None, None, None, ctypes.byref(n)), "bruno_get_k2")
        km = np.zeros(n.value * self.words, np.uint64)
        e = np.zeros(n.value, np.uint8)
        c = np.zeros(n.value, np.uint32)
        if n.value:
            check(L.bruno_get_k2(self.h, _ptr(km), _ptr(e), _ptr(c), ctypes.byref(n)), "bruno_get_k2")
        return km.reshape(-1, self.words), e, c

    def stats(self):
        s = _lib.Stats()
        check(lib().bruno_get_stats(self.h, ctypes.byref(s)), "bruno_get_stats")
        return {f: int(getattr(s, f)) for f in _lib.STAT_FIELDS}

    def phase_ms(self):
        p = _lib.PhaseMs()
        check(lib().bruno_get_phase_ms(self.h, ctypes.byref(p)), "bruno_get_phase_ms")
        return {f: float(getattr(p, f)) for f, _ in _lib.PhaseMs._fields_}


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
