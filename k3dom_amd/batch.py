"""Many independent maps updated together (k3dom_batch_* in include/k3dom.h;
no counterpart in the reference, which runs one map per process).

Every map's prediction, sort and candidate lists run concurrently on the map
streams, the kernel-inference gather of all maps runs in one persistent launch
over a joint work queue, then every map's tail runs concurrently.  Each map's
results are identical to those of the same map updated alone.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import Params
from .dom import DOM


class Batch:
    def __init__(self, params_list):
        self._lib = _lib.load()
        self.params = [p if isinstance(p, Params) else Params.from_buffer_copy(bytes(p)) for p in params_list]
        arr = (Params * len(self.params))(*self.params)
        h = C.c_void_p()
        st = self._lib.k3dom_batch_create(arr, len(self.params), C.byref(h))
        if st != 0:
            raise _lib.K3domError(f"k3dom_batch_create failed with status {st}: "
                                  f"{self._lib.k3dom_last_error(None).decode()}")
        self._h = h
        self.n = len(self.params)
        self.maps = []
        for i in range(self.n):
            mh = C.c_void_p()
            self._check(self._lib.k3dom_batch_map(self._h, i, C.byref(mh)), "batch_map")
            self.maps.append(DOM(self.params[i], _handle=mh, _owner=self))

    def _check(self, st, what):
        if st != 0:
            msg = self._lib.k3dom_batch_last_error(self._h)
            raise _lib.K3domError(f"{what} failed with status {st}: {msg.decode() if msg else ''}")

    def update(self, t, poses, clouds):
        """poses: n sensor positions; clouds: n (xs, ys, zs) host arrays."""
        assert len(poses) == self.n and len(clouds) == self.n
        keep = [tuple(np.ascontiguousarray(a, np.float32) for a in c) for c in clouds]
        pz = np.ascontiguousarray(np.asarray(poses, np.float32).reshape(-1))
        ptr = lambda k: (C.c_void_p * self.n)(*[c[k].ctypes.data for c in keep])
        n = np.array([len(c[0]) for c in keep], np.int32)
        self._check(self._lib.k3dom_batch_update(self._h, float(t), pz.ctypes.data_as(C.POINTER(C.c_float)), ptr(0),
                                                 ptr(1), ptr(2), n.ctypes.data_as(C.POINTER(C.c_int32))),
                    "batch_update")

    def set_graph(self, on=True):
        self._check(self._lib.k3dom_batch_set_graph(self._h, 1 if on else 0), "batch_set_graph")

    def autotune(self):
        c = C.c_int32()
        self._check(self._lib.k3dom_batch_autotune(self._h, C.byref(c)), "batch_autotune")
        return c.value

    def last_ms(self):
        ms = C.c_float()
        self._check(self._lib.k3dom_batch_last_ms(self._h, C.byref(ms)), "batch_last_ms")
        return ms.value

    def close(self):
        if getattr(self, "_h", None):
            for m in self.maps:
                m._h = None
            self._lib.k3dom_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
