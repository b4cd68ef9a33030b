"""ctypes front end of the continuous-batching server (rppa_server_* in include/rpp_audio.h)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import api


class ServeConfig(C.Structure):
    _fields_ = [("maxBatch", C.c_int), ("deadlineUs", C.c_int), ("memBudgetBytes", C.c_longlong),
                ("numSlots", C.c_int), ("eager", C.c_int), ("keepOutputs", C.c_int), ("growSlots", C.c_int)]


class ServeStats(C.Structure):
    _fields_ = [("submitted", C.c_longlong), ("completed", C.c_longlong), ("refused", C.c_longlong),
                ("batches", C.c_longlong), ("p50Us", C.c_double), ("p90Us", C.c_double), ("p99Us", C.c_double),
                ("meanUs", C.c_double), ("maxUs", C.c_double), ("meanBatchClips", C.c_double),
                ("audioSamples", C.c_double), ("gpuBusySeconds", C.c_double), ("slotCapacitySamples", C.c_longlong),
                ("maxQueueUs", C.c_double), ("maxServiceUs", C.c_double), ("maxLatencySamples", C.c_longlong),
                ("maxBatchGpuUs", C.c_double), ("maxLoopUs", C.c_double), ("maxLaunchUs", C.c_double),
                ("maxCopyInCallUs", C.c_double), ("maxKernelCallUs", C.c_double), ("maxCopyOutCallUs", C.c_double),
                ("arenaSubmits", C.c_longlong), ("stagedSubmits", C.c_longlong), ("slotGrowths", C.c_longlong)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_BOUND = False


def _bind():
    global _BOUND
    l = api.lib()
    if not _BOUND:
        P = C.c_void_p
        l.rppa_server_create.argtypes = [C.POINTER(api.RpptLogMelParams), C.c_int, C.POINTER(ServeConfig),
                                         C.POINTER(P)]
        l.rppa_server_destroy.argtypes = [P]
        l.rppa_server_submit.argtypes = [P, P, C.c_int, C.POINTER(C.c_longlong)]
        l.rppa_server_wait.argtypes = [P, C.c_longlong, P, C.c_longlong, C.POINTER(C.c_int), C.POINTER(C.c_double)]
        l.rppa_server_drain.argtypes = [P, C.c_int]
        l.rppa_server_stats.argtypes = [P, C.POINTER(ServeStats)]
        _BOUND = True
    return l


class Server:
    def __init__(self, params: api.RpptLogMelParams, precision="fp16x3", max_batch=64, deadline_us=2000,
                 mem_budget=1 << 30, slots=4, eager=True, keep_outputs=True, grow_slots=True):
        l = _bind()
        self.params = params
        self.cfg = ServeConfig(max_batch, deadline_us, mem_budget, slots, int(eager), int(keep_outputs),
                               int(grow_slots))
        self.s = C.c_void_p()
        api.check("rppa_server_create", l.rppa_server_create(C.byref(params), api.PRECISION[precision],
                                                            C.byref(self.cfg), C.byref(self.s)))

    def submit(self, samples: np.ndarray, check_status=True):
        x = np.ascontiguousarray(samples, np.float32)
        rid = C.c_longlong()
        st = _bind().rppa_server_submit(self.s, x.ctypes.data, int(x.size), C.byref(rid))
        if check_status:
            api.check("rppa_server_submit", st)
        return rid.value if st == 0 else st

    def wait(self, rid: int, num_filter: int, max_frames: int):
        out = np.empty(num_filter * max_frames, np.float32)
        fr = C.c_int()
        lat = C.c_double()
        api.check("rppa_server_wait", _bind().rppa_server_wait(self.s, rid, out.ctypes.data, out.size, C.byref(fr),
                                                              C.byref(lat)))
        return out[:num_filter * fr.value].reshape(num_filter, fr.value), lat.value

    def wait_timing(self, rid: int):
        """Waits without collecting outputs; returns (frames, latency_us), (-1, -1.0) for a request the
        server already forgot (outputs not kept)."""
        fr = C.c_int()
        lat = C.c_double()
        api.check("rppa_server_wait", _bind().rppa_server_wait(self.s, rid, None, 0, C.byref(fr), C.byref(lat)))
        return fr.value, lat.value

    def drain(self, timeout_ms=60000):
        api.check("rppa_server_drain", _bind().rppa_server_drain(self.s, timeout_ms))

    def stats(self) -> dict:
        st = ServeStats()
        api.check("rppa_server_stats", _bind().rppa_server_stats(self.s, C.byref(st)))
        return st.as_dict()

    def close(self):
        if self.s:
            _bind().rppa_server_destroy(self.s)
            self.s = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *e):
        self.close()
