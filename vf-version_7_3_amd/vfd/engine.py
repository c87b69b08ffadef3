"""Engine binding: resident runs, batched ticks (optionally replayed as a
hipGraph), the plain per-field form, and the continuous-batching server.

ctypes over the engine section of include/vfd.h.  A run holds the fields of
one ABM instance resident in HBM; the reference instead copies every field
host->device and back each tick (src/Diffusion/Diffuser.cpp:300-421).
"""
import ctypes

from . import F32, F64, _p, _i, _d, _sz, _stream, check, lib

_u64 = ctypes.c_uint64


class RunDesc(ctypes.Structure):
    _fields_ = [("nz", _i), ("ny", _i), ("nx", _i), ("nfields", _i), ("precision", _i),
                ("sigma2", ctypes.POINTER(_d)), ("decay", ctypes.POINTER(_d))]


class EngineStats(ctypes.Structure):
    _fields_ = [("budget_bytes", _sz), ("used_bytes", _sz), ("runs", _i), ("shapes", _i),
                ("batches", _u64), ("ticks", _u64), ("requests_done", _u64), ("requests_pending", _u64),
                ("peak_used_bytes", _sz), ("runs_admitted_after_wait", _u64)]


lib.vfd_engine_create.argtypes = [_sz, ctypes.POINTER(_p)]
lib.vfd_engine_destroy.argtypes = [_p]
lib.vfd_engine_destroy.restype = None
lib.vfd_run_bytes.argtypes = [ctypes.POINTER(RunDesc)]
lib.vfd_run_bytes.restype = _sz
lib.vfd_run_create.argtypes = [_p, ctypes.POINTER(RunDesc), ctypes.POINTER(_i)]
lib.vfd_run_create_wait.argtypes = [_p, ctypes.POINTER(RunDesc), _d, ctypes.POINTER(_i)]
lib.vfd_run_destroy.argtypes = [_p, _i]
lib.vfd_run_field_ptr.argtypes = [_p, _i, _i, ctypes.POINTER(_p)]
lib.vfd_run_write_field.argtypes = [_p, _i, _i, _p, _p]
lib.vfd_run_read_field.argtypes = [_p, _i, _i, _p, _p]
lib.vfd_step.argtypes = [_p, ctypes.POINTER(_i), _i, _i, _i, _p]
lib.vfd_step_plain.argtypes = [_p, ctypes.POINTER(_i), _i, _i, _p]
lib.vfd_serve_start.argtypes = [_p, _i, _i]
lib.vfd_serve_stop.argtypes = [_p]
lib.vfd_submit.argtypes = [_p, _i, ctypes.POINTER(_u64)]
lib.vfd_poll.argtypes = [_p, _u64, ctypes.POINTER(_i), ctypes.POINTER(_d)]
lib.vfd_wait.argtypes = [_p, _u64, ctypes.POINTER(_d)]
lib.vfd_engine_get_stats.argtypes = [_p, ctypes.POINTER(EngineStats)]


def _desc(shape, sigma2, decay, precision):
    n = len(sigma2)
    assert len(decay) == n
    s2 = (_d * n)(*[float(x) for x in sigma2])
    dc = (_d * n)(*[float(x) for x in decay])
    d = RunDesc(shape[0], shape[1], shape[2], n, precision,
                ctypes.cast(s2, ctypes.POINTER(_d)), ctypes.cast(dc, ctypes.POINTER(_d)))
    d._keep = (s2, dc)
    return d


def run_bytes(shape, nfields, precision=F32):
    return lib.vfd_run_bytes(ctypes.byref(_desc(shape, [1.0] * nfields, [1.0] * nfields, precision)))


class Engine:
    """HBM-budgeted engine on the current device."""

    def __init__(self, budget_bytes):
        h = ctypes.c_void_p()
        check(lib.vfd_engine_create(int(budget_bytes), ctypes.byref(h)))
        self.h = h
        self._shape = {}

    def close(self):
        if getattr(self, "h", None) and lib is not None:
            lib.vfd_engine_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ---- runs ----
    def create_run(self, shape, sigma2, decay, precision=F32):
        d = _desc(shape, sigma2, decay, precision)
        rid = _i()
        check(lib.vfd_run_create(self.h, ctypes.byref(d), ctypes.byref(rid)))
        self._shape[rid.value] = (tuple(shape), len(sigma2), precision)
        return rid.value

    def create_run_wait(self, shape, sigma2, decay, timeout_s, precision=F32):
        """create_run that waits up to timeout_s seconds for memory freed by
        runs destroyed on other threads (ctypes releases the GIL meanwhile)."""
        d = _desc(shape, sigma2, decay, precision)
        rid = _i()
        check(lib.vfd_run_create_wait(self.h, ctypes.byref(d), float(timeout_s), ctypes.byref(rid)))
        self._shape[rid.value] = (tuple(shape), len(sigma2), precision)
        return rid.value

    def destroy_run(self, run):
        check(lib.vfd_run_destroy(self.h, int(run)))
        self._shape.pop(run, None)

    def field_ptr(self, run, field):
        p = ctypes.c_void_p()
        check(lib.vfd_run_field_ptr(self.h, int(run), int(field), ctypes.byref(p)))
        return p.value

    def write_field(self, run, field, tensor, stream=None):
        """Copy a contiguous torch tensor (host or device, dtype of the run) in."""
        self._check_tensor(run, tensor)
        check(lib.vfd_run_write_field(self.h, int(run), int(field), ctypes.c_void_p(tensor.data_ptr()),
                                      _stream(stream)))

    def read_field(self, run, field, out=None, device="cuda", stream=None):
        import torch
        shape, _, prec = self._shape[run]
        if out is None:
            out = torch.empty(shape, dtype=torch.float32 if prec == F32 else torch.float64, device=device)
        self._check_tensor(run, out)
        check(lib.vfd_run_read_field(self.h, int(run), int(field), ctypes.c_void_p(out.data_ptr()),
                                     _stream(stream)))
        return out

    def _check_tensor(self, run, t):
        import torch
        shape, _, prec = self._shape[run]
        want = torch.float32 if prec == F32 else torch.float64
        if t.dtype != want or tuple(t.shape) != shape or not t.is_contiguous():
            raise ValueError(f"field tensor must be contiguous {want} of shape {shape}")

    # ---- ticks ----
    @staticmethod
    def _ids(runs):
        runs = list(runs)
        return (_i * max(1, len(runs)))(*runs), len(runs)

    def step(self, runs, ticks=1, graph=False, stream=None):
        ids, n = self._ids(runs)
        check(lib.vfd_step(self.h, ids, n, int(ticks), 1 if graph else 0, _stream(stream)))

    def step_plain(self, runs, ticks=1, stream=None):
        ids, n = self._ids(runs)
        check(lib.vfd_step_plain(self.h, ids, n, int(ticks), _stream(stream)))

    # ---- serving ----
    def serve_start(self, slots=2, max_batch_fields=256):
        check(lib.vfd_serve_start(self.h, int(slots), int(max_batch_fields)))

    def serve_stop(self):
        check(lib.vfd_serve_stop(self.h))

    def submit(self, run):
        t = _u64()
        check(lib.vfd_submit(self.h, int(run), ctypes.byref(t)))
        return t.value

    def poll(self, ticket):
        done, lat = _i(), _d()
        check(lib.vfd_poll(self.h, _u64(ticket), ctypes.byref(done), ctypes.byref(lat)))
        return bool(done.value), lat.value

    def wait(self, ticket):
        lat = _d()
        check(lib.vfd_wait(self.h, _u64(ticket), ctypes.byref(lat)))
        return lat.value

    def stats(self):
        s = EngineStats()
        check(lib.vfd_engine_get_stats(self.h, ctypes.byref(s)))
        return {k: getattr(s, k) for k, _ in EngineStats._fields_}


__all__ = ["Engine", "RunDesc", "EngineStats", "run_bytes", "F32", "F64"]
