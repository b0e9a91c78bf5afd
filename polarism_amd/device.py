"""Device memory, streams and events on top of libpgpe_hip.so.

``DeviceArray`` is the array type the solvers hand back in ``state.field`` and
``reservoir.get_state()``: a typed, shaped view of a device allocation.  It
converts to NumPy on demand (``np.asarray(a)`` or ``a.get()``), the analogue
of the reference's ``compute_engine.to_cpu`` for CuPy arrays
(polarism/compute_engine.py:33-97).
"""
from __future__ import annotations

import ctypes as C
import gc
import threading

import numpy as np

from . import _native as N

_init_lock = threading.Lock()
_initialised = False


def init(device: int = 0) -> None:
    """Select and initialise the GPU once per process."""
    global _initialised
    with _init_lock:
        if _initialised:
            return
        lib = N.lib()
        count = C.c_int(0)
        rc = lib.pgpe_device_count(C.byref(count))
        if rc != 0 or count.value == 0:
            raise RuntimeError("no HIP device available: the pgpe solvers run only on an AMD GPU "
                               f"({N.last_error() or 'device count is 0'})")
        N.call("pgpe_device_init", device)
        _initialised = True


def available() -> bool:
    """True when the library loads and a HIP device is visible."""
    try:
        lib = N.lib()
    except RuntimeError:
        return False
    count = C.c_int(0)
    return lib.pgpe_device_count(C.byref(count)) == 0 and count.value > 0


# Graph capture in progress on this thread: hipFree is not allowed inside a capture (it would
# invalidate it), so buffers released meanwhile are freed when the capture ends.
_capture_depth = 0
_deferred_frees: list[int] = []


def _release(ptr: int) -> None:
    if _capture_depth:
        _deferred_frees.append(ptr)
    elif N._lib is not None:
        N._lib.pgpe_free(ptr)


class DeviceBuffer:
    """Owning device allocation (freed when garbage collected)."""

    __slots__ = ("ptr", "nbytes", "__weakref__")

    def __init__(self, nbytes: int):
        init()
        p = C.c_void_p()
        N.call("pgpe_malloc", C.byref(p), int(nbytes))
        self.ptr = p.value
        self.nbytes = int(nbytes)

    def __del__(self):
        ptr = getattr(self, "ptr", None)
        if ptr:
            _release(ptr)
            self.ptr = None


class DeviceArray:
    """A C-contiguous typed view into a device buffer."""

    def __init__(self, shape, dtype, buffer: DeviceBuffer | None = None, offset: int = 0):
        self.shape = tuple(int(s) for s in shape)
        self.dtype = np.dtype(dtype)
        nbytes = int(np.prod(self.shape, dtype=np.int64)) * self.dtype.itemsize
        if buffer is None:
            buffer = DeviceBuffer(max(nbytes, 16))
        if offset + nbytes > buffer.nbytes:
            raise ValueError("view exceeds its buffer")
        self.buffer = buffer
        self.offset = int(offset)
        self.nbytes = nbytes

    # -- construction -----------------------------------------------------
    @classmethod
    def empty(cls, shape, dtype):
        return cls(shape, dtype)

    @classmethod
    def zeros(cls, shape, dtype, stream=None):
        a = cls(shape, dtype)
        a.fill_zero(stream)
        return a

    @classmethod
    def from_numpy(cls, host, dtype=None, stream=None):
        host = np.ascontiguousarray(host, dtype=dtype)
        a = cls(host.shape, host.dtype)
        a.upload(host, stream)
        return a

    # -- properties -------------------------------------------------------
    @property
    def ptr(self) -> int:
        return self.buffer.ptr + self.offset

    @property
    def size(self) -> int:
        return int(np.prod(self.shape, dtype=np.int64))

    @property
    def ndim(self) -> int:
        return len(self.shape)

    def __len__(self):
        return self.shape[0]

    def __repr__(self):
        return f"DeviceArray(shape={self.shape}, dtype={self.dtype})"

    # -- transfers --------------------------------------------------------
    def upload(self, host, stream=None) -> None:
        host = np.ascontiguousarray(host, dtype=self.dtype)
        if host.size != self.size:
            raise ValueError(f"size mismatch: {host.shape} vs {self.shape}")
        N.call("pgpe_memcpy_h2d", self.ptr, host.ctypes.data, self.nbytes, _sp(stream))

    def get(self, stream=None) -> np.ndarray:
        out = np.empty(self.shape, dtype=self.dtype)
        N.call("pgpe_memcpy_d2h", out.ctypes.data, self.ptr, self.nbytes, _sp(stream))
        return out

    def __array__(self, dtype=None, copy=None):
        a = self.get()
        return a if dtype is None else a.astype(dtype)

    def copy_from(self, other: "DeviceArray", stream=None) -> None:
        if other.nbytes != self.nbytes:
            raise ValueError("copy_from size mismatch")
        N.call("pgpe_memcpy_d2d", self.ptr, other.ptr, self.nbytes, _sp(stream))

    def copy(self, stream=None) -> "DeviceArray":
        out = DeviceArray(self.shape, self.dtype)
        out.copy_from(self, stream)
        sync(stream)
        return out

    def fill_zero(self, stream=None) -> None:
        N.call("pgpe_fill_zero", self.ptr, self.nbytes, _sp(stream))

    def reshape(self, *shape) -> "DeviceArray":
        if len(shape) == 1 and isinstance(shape[0], (tuple, list)):
            shape = tuple(shape[0])
        shape = tuple(shape)
        if -1 in shape:
            known = int(np.prod([s for s in shape if s != -1]))
            shape = tuple(self.size // known if s == -1 else s for s in shape)
        if int(np.prod(shape)) != self.size:
            raise ValueError("reshape changes the size")
        return DeviceArray(shape, self.dtype, self.buffer, self.offset)

    def view_range(self, start_elem: int, shape) -> "DeviceArray":
        """Sub-view of ``shape`` starting ``start_elem`` elements into this array."""
        return DeviceArray(shape, self.dtype, self.buffer, self.offset + int(start_elem) * self.dtype.itemsize)


def _sp(stream):
    if stream is None:
        return None
    return stream.handle if isinstance(stream, Stream) else stream


class Stream:
    """A non-blocking HIP stream at the default priority."""

    def __init__(self):
        init()
        s = C.c_void_p()
        N.call("pgpe_stream_create", C.byref(s))
        self.handle = s.value

    def synchronize(self) -> None:
        N.call("pgpe_stream_sync", self.handle)

    def __del__(self):
        h = getattr(self, "handle", None)
        if h and N._lib is not None:
            N._lib.pgpe_stream_destroy(h)
            self.handle = None


def sync(stream=None) -> None:
    N.call("pgpe_stream_sync", _sp(stream))


def device_name(device: int = 0) -> str:
    """Marketing name of the device (part of the autotuner's cache key)."""
    init()
    buf = C.create_string_buffer(256)
    N.call("pgpe_device_name", device, buf, C.c_size_t(256))
    return buf.value.decode(errors="replace")


class Event:
    def __init__(self):
        init()
        e = C.c_void_p()
        N.call("pgpe_event_create", C.byref(e))
        self.handle = e.value

    def record(self, stream=None) -> None:
        N.call("pgpe_event_record", self.handle, _sp(stream))

    def synchronize(self) -> None:
        N.call("pgpe_event_sync", self.handle)

    def elapsed_ms(self, later: "Event") -> float:
        ms = C.c_float()
        N.call("pgpe_event_elapsed_ms", self.handle, later.handle, C.byref(ms))
        return float(ms.value)

    def __del__(self):
        h = getattr(self, "handle", None)
        if h and N._lib is not None:
            N._lib.pgpe_event_destroy(h)
            self.handle = None


class PinnedBuffer:
    """Page-locked host memory (hipHostMalloc) for copies the host waits on through an event."""

    def __init__(self, nbytes: int):
        init()
        p = C.c_void_p()
        N.call("pgpe_host_alloc", C.byref(p), int(nbytes))
        self.ptr = p.value
        self.nbytes = int(nbytes)

    def fetch(self, src: "DeviceArray", offset: int, stream: "Stream") -> None:
        """Queue a copy of ``src`` into this buffer at ``offset`` bytes; returns at once."""
        if offset + src.nbytes > self.nbytes:
            raise ValueError("pinned buffer too small")
        N.call("pgpe_memcpy_d2h_async", self.ptr + offset, src.ptr, src.nbytes, stream.handle)

    def read(self, offset: int, nbytes: int) -> bytes:
        return C.string_at(self.ptr + offset, nbytes)

    def __del__(self):
        ptr = getattr(self, "ptr", None)
        if ptr and N._lib is not None:
            N._lib.pgpe_host_free(ptr)
            self.ptr = None


class Graph:
    """An instantiated HIP graph of launches captured on one stream."""

    def __init__(self, exec_handle: int):
        self.handle = exec_handle

    @staticmethod
    def capture(stream: Stream, body) -> "Graph":
        """Capture the launches ``body`` queues on ``stream``.  Python's cyclic collector is held
        off and buffers released meanwhile are freed after the capture: a hipFree (or a graph
        destruction) inside it would invalidate the capture."""
        global _capture_depth
        collecting = gc.isenabled()
        gc.disable()
        _capture_depth += 1
        try:
            N.call("pgpe_graph_begin", stream.handle)
            try:
                body()
            except BaseException:
                out = C.c_void_p()
                N.lib().pgpe_graph_end(stream.handle, C.byref(out))
                if out.value:
                    N.lib().pgpe_graph_destroy(out.value)
                raise
            out = C.c_void_p()
            N.call("pgpe_graph_end", stream.handle, C.byref(out))
            return Graph(out.value)
        finally:
            _capture_depth -= 1
            if collecting:
                gc.enable()
            while not _capture_depth and _deferred_frees:
                _release(_deferred_frees.pop())

    def launch(self, stream: Stream) -> None:
        N.call("pgpe_graph_launch", self.handle, stream.handle)

    def __del__(self):
        h = getattr(self, "handle", None)
        if h and N._lib is not None:
            N._lib.pgpe_graph_destroy(h)
            self.handle = None


def to_cpu(x):
    """Return a NumPy array for a DeviceArray, pass NumPy through."""
    if isinstance(x, DeviceArray):
        return x.get()
    return np.asarray(x)
