"""ctypes binding of the CPU parity oracle (test infrastructure only).

Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg
import this module, and only to check or to time the CPU restatement beside
the HIP path.  The product never calls it.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


class OracleNode(ctypes.Structure):
    _fields_ = [("kmer_lo", ctypes.c_uint64), ("kmer_hi", ctypes.c_uint64),
                ("counts", ctypes.c_uint16 * 9)]


class OracleChainNode(ctypes.Structure):
    _fields_ = [("k_lo", ctypes.c_uint64), ("k_hi", ctypes.c_uint64),
                ("rep_lo", ctypes.c_uint64), ("rep_hi", ctypes.c_uint64),
                ("dist", ctypes.c_int32)]


NODE_DTYPE = np.dtype([("kmer_lo", "<u8"), ("kmer_hi", "<u8"), ("counts", "<u2", (9,))], align=True)
CHAIN_DTYPE = np.dtype([("k_lo", "<u8"), ("k_hi", "<u8"), ("rep_lo", "<u8"), ("rep_hi", "<u8"),
                        ("dist", "<i4")], align=True)


def build():
    import subprocess
    subprocess.check_call(["make", "-s", "-C", _HERE, "-B"])


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        vp, u64, u64p = ctypes.c_void_p, ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint64)
        L.oracle_create.restype = vp
        L.oracle_create.argtypes = [ctypes.c_int, vp, ctypes.c_int]
        L.oracle_destroy.argtypes = [vp]
        L.oracle_set_edge_filter.argtypes = [vp, ctypes.c_int]
        L.oracle_add_reads.argtypes = [vp, ctypes.c_char_p, vp, u64]
        L.oracle_run.argtypes = [vp, ctypes.c_int]
        for name in ["n_instances", "n_k2", "n_nodes", "n_branches", "n_chain_nodes",
                     "n_cycle_nodes", "n_isolated", "iterations", "chain_edge_misses"]:
            f = getattr(L, "oracle_" + name)
            f.restype = u64
            f.argtypes = [vp]
        L.oracle_get_records.restype = u64
        L.oracle_get_records.argtypes = [vp, vp, vp, vp, u64]
        L.oracle_get_k2.restype = u64
        L.oracle_get_k2.argtypes = [vp, vp, vp, vp, vp, u64]
        for name in ["get_unfiltered", "get_nodes", "get_chain_nodes",
                     "get_branch_fasta", "get_chain_fasta", "get_chain_components",
                     "get_branch_edges", "get_chain_edges"]:
            f = getattr(L, "oracle_" + name)
            f.restype = u64
            f.argtypes = [vp, vp, u64]
        L.oracle_phase_seconds.argtypes = [vp, vp]
        L.oracle_set_order_seed.argtypes = [vp, u64]
        L.oracle_run_clean.argtypes = [vp, ctypes.c_int]
        for name in ["clean_rounds", "clean_deadends", "clean_bubbles", "n_clean_chain_nodes"]:
            f = getattr(L, "oracle_" + name)
            f.restype = u64
            f.argtypes = [vp]
        L.oracle_get_clean_output.restype = u64
        L.oracle_get_clean_output.argtypes = [vp, ctypes.c_int, vp, u64]
        L.oracle_get_clean_nodes.restype = u64
        L.oracle_get_clean_nodes.argtypes = [vp, vp, u64]
        L.oracle_get_clean_chain_nodes.restype = u64
        L.oracle_get_clean_chain_nodes.argtypes = [vp, vp, u64]
        _LIB = L
    return _LIB


class Oracle:
    """One oracle job: add reads, run, read the parity objects back."""

    def __init__(self, k, thresholds=None, nthreads=1, edge_filter=False):
        L = lib()
        self._th = None
        if thresholds is not None:
            self._th = np.ascontiguousarray(thresholds, dtype=np.uint64)
            assert self._th.size == 6
        self.h = L.oracle_create(k, self._th.ctypes.data if self._th is not None else None, nthreads)
        if not self.h:
            raise ValueError("bad k")
        self.k = k
        if edge_filter:
            L.oracle_set_edge_filter(self.h, 1)

    def __del__(self):
        if getattr(self, "h", None):
            lib().oracle_destroy(self.h)
            self.h = None

    def add_reads(self, bases, offsets):
        """bases: bytes/np.uint8 array; offsets: n_reads+1 uint64."""
        b = np.ascontiguousarray(np.frombuffer(bases, dtype=np.uint8) if isinstance(bases, (bytes, bytearray)) else bases,
                                 dtype=np.uint8)
        o = np.ascontiguousarray(offsets, dtype=np.uint64)
        lib().oracle_add_reads(self.h, b.ctypes.data_as(ctypes.c_char_p), o.ctypes.data, o.size - 1)

    def set_order_seed(self, seed):
        lib().oracle_set_order_seed(self.h, seed)

    def run(self, lr_variant=1):
        rc = lib().oracle_run(self.h, lr_variant)
        if rc != 0:
            raise RuntimeError("oracle_run failed %d" % rc)

    def stat(self, name):
        return int(getattr(lib(), "oracle_" + name)(self.h))

    def records(self):
        L = lib()
        n = L.oracle_get_records(self.h, None, None, None, 0)
        lo = np.empty(n, np.uint64); hi = np.empty(n, np.uint64); e = np.empty(n, np.uint8)
        L.oracle_get_records(self.h, lo.ctypes.data, hi.ctypes.data, e.ctypes.data, n)
        return lo, hi, e

    def k2(self):
        L = lib()
        n = L.oracle_get_k2(self.h, None, None, None, None, 0)
        lo = np.empty(n, np.uint64); hi = np.empty(n, np.uint64)
        e = np.empty(n, np.uint8); c = np.empty(n, np.uint32)
        L.oracle_get_k2(self.h, lo.ctypes.data, hi.ctypes.data, e.ctypes.data, c.ctypes.data, n)
        return lo, hi, e, c

    def _nodes(self, fn):
        f = getattr(lib(), fn)
        n = f(self.h, None, 0)
        out = np.zeros(n, NODE_DTYPE)
        f(self.h, out.ctypes.data, n)
        return out

    def unfiltered(self):
        return self._nodes("oracle_get_unfiltered")

    def nodes(self):
        return self._nodes("oracle_get_nodes")

    def chain_nodes(self):
        f = lib().oracle_get_chain_nodes
        n = f(self.h, None, 0)
        out = np.zeros(n, CHAIN_DTYPE)
        f(self.h, out.ctypes.data, n)
        return out

    def _str(self, fn):
        f = getattr(lib(), fn)
        n = f(self.h, None, 0)
        buf = ctypes.create_string_buffer(max(n, 1))
        f(self.h, buf, n)
        return buf.raw[:n]

    def branch_fasta(self):
        return self._str("oracle_get_branch_fasta")

    def chain_fasta(self):
        return self._str("oracle_get_chain_fasta")

    def chain_components(self):
        return self._str("oracle_get_chain_components")

    def branch_edges(self):
        return self._str("oracle_get_branch_edges")

    def chain_edges(self):
        return self._str("oracle_get_chain_edges")

    def run_clean(self, lr_variant=1):
        """Graph cleaning after run(): the .clean.* outputs."""
        rc = lib().oracle_run_clean(self.h, lr_variant)
        if rc != 0:
            raise RuntimeError("oracle_run_clean failed %d" % rc)

    CLEAN_OUTPUTS = ("chain_fasta", "branch_fasta", "chain_components", "branch_edges", "chain_edges")

    def clean_output(self, name):
        which = self.CLEAN_OUTPUTS.index(name)
        f = lib().oracle_get_clean_output
        n = f(self.h, which, None, 0)
        buf = ctypes.create_string_buffer(max(n, 1))
        f(self.h, which, buf, n)
        return buf.raw[:n]

    def clean_nodes(self):
        return self._nodes("oracle_get_clean_nodes")

    def clean_chain_nodes(self):
        f = lib().oracle_get_clean_chain_nodes
        n = f(self.h, None, 0)
        out = np.zeros(n, CHAIN_DTYPE)
        f(self.h, out.ctypes.data, n)
        return out

    def phase_seconds(self):
        out = np.zeros(6, np.float64)
        lib().oracle_phase_seconds(self.h, out.ctypes.data)
        return out
