"""Device memory for the tests of the entry points that write into a caller's device buffer."""
import ctypes

import numpy as np

from phoenix_amd import capi


class DeviceBuffer:
    """n FreqOut entries of device memory from the HIP runtime libphx.so itself uses (torch ships
    a runtime of its own, and two runtimes in one process do not share devices), filled with NaN."""

    def __init__(self, n):
        capi.lib()  # libphx.so first: its libamdhip64.so.7 is the one loaded below
        self.hip = ctypes.CDLL("libamdhip64.so.7")
        self.n = n
        self.ptr = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(self.ptr), ctypes.c_size_t(n * 48)) == 0
        self.clear()

    def clear(self):
        fill = np.full((self.n, 6), np.nan)
        assert self.hip.hipMemcpy(self.ptr, fill.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(self.n * 48), 1) == 0

    def get(self):
        out = np.empty((self.n, 6))
        assert self.hip.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), self.ptr, ctypes.c_size_t(self.n * 48), 2) == 0
        return out

    def free(self):
        if self.ptr:
            self.hip.hipFree(self.ptr)
            self.ptr = None
