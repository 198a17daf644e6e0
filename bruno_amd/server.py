"""Host binding of the request loop (bruno_server_* in include/bruno_hip.h).

Server.submit queues one assembly job (host reads or a synthetic workload
generated on the device) and returns a ticket; the library's worker threads
admit queued jobs by estimated HBM footprint and run them in batches on
their own streams.  wait(ticket) blocks until the job is done and returns
its result record; unitigs / branch_fasta return its output bytes.

ctypes releases the GIL during each call, so wait() in one Python thread
does not hold up submit() in another.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import check, lib
from .pipeline import DEFAULT_THRESH


class Server:
    def __init__(self, hbm_budget=0, n_workers=2, max_batch_jobs=64, variant=1, reserve_pool=True):
        d = _lib.ServerDesc(int(hbm_budget), int(n_workers), int(max_batch_jobs), int(variant),
                            1 if reserve_pool else 0)
        h = ctypes.c_void_p()
        check(lib().bruno_server_create(ctypes.byref(d), ctypes.byref(h)), "bruno_server_create")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            check(lib().bruno_server_destroy(self.h), "bruno_server_destroy")
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @staticmethod
    def _job(k, thresholds, capacity):
        j = _lib.JobDesc()
        j.k = int(k)
        j.thresholding = 0 if thresholds is None else 1
        th = thresholds if thresholds is not None else [DEFAULT_THRESH] * 6
        for i in range(6):
            j.thresholds[i] = int(th[i])
        j.capacity = int(capacity)
        return j

    def submit_reads(self, k, thresholds, bases, offsets, capacity=0):
        b = np.frombuffer(bases, np.uint8) if isinstance(bases, (bytes, bytearray)) else np.ascontiguousarray(bases, np.uint8)
        o = np.ascontiguousarray(offsets, np.uint64)
        q = _lib.Request()
        q.job = self._job(k, thresholds, capacity)
        q.bases = b.ctypes.data if b.size else None
        q.offsets = o.ctypes.data
        q.n_reads = o.size - 1
        t = ctypes.c_uint64()
        check(lib().bruno_server_submit(self.h, ctypes.byref(q), ctypes.byref(t)), "bruno_server_submit")
        return t.value

    def submit_synth(self, k, thresholds, synth, capacity=0):
        q = _lib.Request()
        q.job = self._job(k, thresholds, capacity)
        sd = synth.desc()
        q.synth = ctypes.pointer(sd)
        t = ctypes.c_uint64()
        check(lib().bruno_server_submit(self.h, ctypes.byref(q), ctypes.byref(t)), "bruno_server_submit")
        return t.value

    def done(self, ticket):
        d = ctypes.c_int()
        check(lib().bruno_server_poll(self.h, int(ticket), ctypes.byref(d)), "bruno_server_poll")
        return bool(d.value)

    def wait(self, ticket):
        r = _lib.ResultInfo()
        check(lib().bruno_server_wait(self.h, int(ticket), ctypes.byref(r)), "bruno_server_wait")
        return {f: getattr(r, f) for f, _ in _lib.RESULT_FIELDS}

    def _text(self, fn, ticket):
        n = ctypes.c_uint64()
        check(fn(self.h, int(ticket), None, ctypes.byref(n)), fn.__name__)
        buf = ctypes.create_string_buffer(max(1, n.value))
        check(fn(self.h, int(ticket), buf, ctypes.byref(n)), fn.__name__)
        return buf.raw[:n.value]

    def unitigs(self, ticket):
        return self._text(lib().bruno_server_get_unitigs, ticket)

    def branch_fasta(self, ticket):
        return self._text(lib().bruno_server_get_branch_fasta, ticket)

    def release(self, ticket):
        check(lib().bruno_server_release(self.h, int(ticket)), "bruno_server_release")

    def stats(self):
        s = _lib.ServerStats()
        check(lib().bruno_server_get_stats(self.h, ctypes.byref(s)), "bruno_server_get_stats")
        return {f: int(getattr(s, f)) for f in _lib.SERVER_STAT_FIELDS}
