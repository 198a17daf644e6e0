"""The allocation-free FASTQ split (bruno_fastq_split_noalloc): the same
outputs as bruno_fastq_split on consecutive calls (the second smaller than
the first, so stale scratch from the first call would show), no device
memory taken from the pool during the call, and the stated limits rejected
with BRUNO_CAPACITY."""
import ctypes

import numpy as np
import pytest

from bruno_amd import _lib
from bruno_amd.device import DeviceBuffer, pool_stats
from bruno_amd.pipeline import fastq_split_device
from helpers import random_reads

pytestmark = pytest.mark.gpu


def to_fastq(seqs, crlf=False):
    nl = "\r\n" if crlf else "\n"
    return "".join("@r%d%s%s%s+%s%s%s" % (i, nl, s, nl, nl, "I" * len(s), nl) for i, s in enumerate(seqs)).encode()


def split_into(fn, dev_fastq, nbytes, cap_reads, cap_bases):
    """One call of fn with caller-owned outputs; returns (status, bases, offsets)."""
    ob, oo = DeviceBuffer(cap_bases + 1), DeviceBuffer((cap_reads + 1) * 8)
    try:
        nr, nb = ctypes.c_uint64(cap_reads), ctypes.c_uint64(cap_bases)
        before = pool_stats()
        st = fn(ctypes.c_void_p(dev_fastq), int(nbytes), ctypes.c_void_p(ob.ptr), ctypes.c_void_p(oo.ptr),
                ctypes.byref(nr), ctypes.byref(nb), None)
        after = pool_stats()
        if st != 0:
            return st, None, None, (before, after)
        return st, ob.to_numpy(np.uint8, nb.value), oo.to_numpy(np.uint64, nr.value + 1), (before, after)
    finally:
        ob.free()
        oo.free()


def test_noalloc_split_equals_split_on_two_calls():
    L = _lib.lib()
    inputs = []
    for seed, n, rl, crlf in ((1, 20000, (60, 250), False), (2, 3000, 100, True)):  # second call smaller
        _, reads = random_reads(seed, 50000, n, rl, 0.01)
        inputs.append(to_fastq(reads, crlf=crlf))
    assert len(inputs[1]) < len(inputs[0])
    for data in inputs:
        buf = DeviceBuffer.from_numpy(np.frombuffer(data, np.uint8))
        try:
            ref_b, ref_o = fastq_split_device(buf.ptr, len(data))
            st, b, o, (before, after) = split_into(L.bruno_fastq_split_noalloc, buf.ptr, len(data), len(ref_o) - 1,
                                                   len(ref_b))
            assert st == 0
            assert before == after  # no device memory taken during the call
            assert np.array_equal(o, ref_o)
            assert np.array_equal(b, ref_b)
            # the size query form (both outputs NULL) gives the same sizes
            nb_, no_ = fastq_split_device(buf.ptr, len(data), noalloc=True)
            assert np.array_equal(no_, ref_o) and np.array_equal(nb_, ref_b)
        finally:
            buf.free()


def test_noalloc_split_rejects_inputs_beyond_its_limits():
    L = _lib.lib()
    # 2^21 + 1 one-base records (16 MiB of FASTQ): more records than the scratch holds
    n = (1 << 21) + 1
    data = b"@\nA\n+\nI\n" * n
    buf = DeviceBuffer.from_numpy(np.frombuffer(data, np.uint8))
    try:
        st, _, _, _ = split_into(L.bruno_fastq_split_noalloc, buf.ptr, len(data), n, n)
        assert st == 3
        ref_b, ref_o = fastq_split_device(buf.ptr, len(data))  # the pooled split takes it
        assert len(ref_o) == n + 1 and len(ref_b) == n
    finally:
        buf.free()
    # more than 1 GiB of bytes: rejected before any device work (the buffer
    # is that large, so nothing could read past it either way)
    big = DeviceBuffer((1 << 30) + 1)
    try:
        nr, nb = ctypes.c_uint64(0), ctypes.c_uint64(0)
        assert L.bruno_fastq_split_noalloc(ctypes.c_void_p(big.ptr), (1 << 30) + 1, None, None, ctypes.byref(nr),
                                           ctypes.byref(nb), None) == 3
    finally:
        big.free()
    # and the entry point still works after the rejections
    data = to_fastq(["ACGTACGTAC", "GGGTTTAAAC"])
    buf = DeviceBuffer.from_numpy(np.frombuffer(data, np.uint8))
    try:
        b, o = fastq_split_device(buf.ptr, len(data), noalloc=True)
        assert bytes(b) == b"ACGTACGTACGGGTTTAAAC" and list(o) == [0, 10, 20]
    finally:
        buf.free()
