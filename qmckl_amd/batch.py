"""Several contexts evaluated from one HIP graph (qmckl_amd_batch_* of include/qmckl.h)."""

import ctypes

import numpy as np

from . import QmcklError, lib


class Batch:
    """Records the device step of every context (coordinates from coord_bufs[i]; MO VGL,
    Jastrow value and GL into the optional output buffers) and replays it as one graph."""

    def __init__(self, contexts, coord_bufs, mo_bufs=None, value_bufs=None, gl_bufs=None):
        L = lib()
        n = len(contexts)
        self._ctx = np.asarray(contexts, np.int64)
        self._arrays = []
        ptrs = []
        for bufs in (coord_bufs, mo_bufs, value_bufs, gl_bufs):
            if bufs is None:
                ptrs.append(None)
            else:
                a = np.asarray([b.ptr for b in bufs], np.uint64)
                self._arrays.append(a)
                ptrs.append(a.ctypes.data)
        h = ctypes.c_int64(0)
        rc = L.qmckl_amd_batch_record(self._ctx.ctypes.data, n, None, *ptrs, ctypes.addressof(h))
        if rc != 0:
            raise QmcklError(rc, "qmckl_amd_batch_record")
        self.handle = h.value

    def replay(self):
        rc = lib().qmckl_amd_batch_replay(self.handle)
        if rc != 0:
            raise QmcklError(rc, "qmckl_amd_batch_replay")

    def synchronize(self):
        rc = lib().qmckl_amd_batch_synchronize(self.handle)
        if rc != 0:
            raise QmcklError(rc, "qmckl_amd_batch_synchronize")

    def destroy(self):
        if self.handle:
            lib().qmckl_amd_batch_destroy(self.handle)
            self.handle = 0
