"""Device buffers through the library's own HIP runtime.

The GPU processes of this project (tests, benchmark) never initialise a
second HIP runtime (torch bundles its own), so device memory for reads and
records is allocated and copied here."""
import ctypes

import numpy as np

from ._lib import check, lib


class DeviceBuffer:
    def __init__(self, nbytes):
        p = ctypes.c_void_p()
        check(lib().bruno_dev_alloc(int(nbytes), ctypes.byref(p)), "bruno_dev_alloc")
        self.ptr = p.value
        self.nbytes = int(nbytes)

    @classmethod
    def from_numpy(cls, a):
        a = np.ascontiguousarray(a)
        b = cls(max(1, a.nbytes))
        if a.nbytes:
            check(lib().bruno_dev_copy(ctypes.c_void_p(b.ptr), ctypes.c_void_p(a.ctypes.data), a.nbytes, 0),
                  "bruno_dev_copy")
        return b

    def to_numpy(self, dtype, count):
        out = np.empty(count, dtype)
        if out.nbytes:
            check(lib().bruno_dev_copy(ctypes.c_void_p(out.ctypes.data), ctypes.c_void_p(self.ptr), out.nbytes, 1),
                  "bruno_dev_copy")
        return out

    def free(self):
        if self.ptr:
            lib().bruno_dev_free(ctypes.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        self.free()


def synchronize():
    check(lib().bruno_dev_sync(), "bruno_dev_sync")


def mem_info():
    f, t = ctypes.c_uint64(), ctypes.c_uint64()
    check(lib().bruno_dev_mem_info(ctypes.byref(f), ctypes.byref(t)), "bruno_dev_mem_info")
    return f.value, t.value


def pool_reserve(nbytes):
    """Reserve one pool chunk of nbytes so later allocations make no HIP
    allocator calls."""
    check(lib().bruno_pool_reserve(int(nbytes)), "bruno_pool_reserve")


def pool_trim():
    check(lib().bruno_pool_trim(), "bruno_pool_trim")


def pool_stats():
    r, u, f = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
    check(lib().bruno_pool_stats(ctypes.byref(r), ctypes.byref(u), ctypes.byref(f)), "bruno_pool_stats")
    return {"reserved": r.value, "in_use": u.value, "largest_free": f.value}
