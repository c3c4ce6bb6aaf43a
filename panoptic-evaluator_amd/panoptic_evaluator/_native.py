"""ctypes binding of libpanoptic_hip.so (the gfx950 kernels' C ABI).

The library is built in-tree by ``panoptic-evaluator_amd/build.py``.  There is
no fallback: if the library is missing or the tensors are not on a HIP device,
the call fails loudly.
"""

from __future__ import annotations

import collections
import contextlib
import ctypes
import pathlib
import threading
import weakref

import torch

LIBRARY_PATH = pathlib.Path(__file__).with_name("libpanoptic_hip.so")
ABI_VERSION = 3

_i64, _int, _ptr, _dbl = ctypes.c_int64, ctypes.c_int, ctypes.c_void_p, ctypes.c_double

# name: (restype, argtypes)
_SIGNATURES = {
    "pe_last_error": (ctypes.c_char_p, []),
    "pe_abi_version": (_int, []),
    "pe_profile_enable": (_int, [_int]),
    "pe_profile_reset": (_int, []),
    "pe_profile_read": (_int, [ctypes.c_char_p, ctypes.POINTER(_dbl), ctypes.POINTER(_i64)]),
    "pe_tune_count": (_int, [ctypes.c_char_p]),
    "pe_tune_set": (_int, [ctypes.c_char_p, _int]),
    "pe_tune_get": (_int, [ctypes.c_char_p]),
    "pe_tune_describe": (ctypes.c_char_p, [ctypes.c_char_p, _int]),
    "pe_update_workspace_bytes": (_i64, [_i64] * 5),
    "pe_update": (_int, [_ptr, _int, _ptr, _int, _i64, _i64, _i64, _ptr, _ptr, _ptr, _ptr, _i64, _i64, _i64,
                         _ptr, _ptr, _ptr, _ptr, _ptr, _ptr, _ptr, _ptr, _ptr, _ptr, _ptr, _ptr,
                         _ptr, _ptr, _ptr, _i64, _int, _int, _int, _ptr]),
    "pe_compute": (_int, [_ptr, _ptr, _ptr, _ptr, _ptr, _ptr, _i64, _i64, _ptr, _ptr, _ptr, _ptr,
                          _ptr, _ptr, _ptr, _ptr, _ptr, _ptr]),
    "pe_compute_with_ap": (_int, [_ptr, _ptr, _ptr, _ptr, _ptr, _ptr, _i64, _i64, _ptr, _ptr, _ptr, _ptr,
                                  _ptr, _ptr, _ptr, _ptr, _ptr,
                                  _i64, _ptr, _ptr, _ptr, _ptr, _ptr, _i64, _ptr, _ptr, _ptr, _ptr, _i64, _ptr]),
    "pe_compute_with_ap_match": (_int, [_ptr, _ptr, _ptr, _ptr, _ptr, _ptr, _i64, _i64, _ptr, _ptr, _ptr, _ptr,
                                        _ptr, _ptr, _ptr, _ptr, _ptr,
                                        _i64, _ptr, _ptr, _ptr, _ptr, _ptr, _i64, _ptr, _ptr, _ptr, _ptr, _i64,
                                        _i64, _i64, _ptr, _ptr, _i64, _ptr, _ptr, _ptr, _i64, _ptr, _ptr, _ptr]),
    "pe_image_pq": (_int, [_ptr, _ptr, _ptr, _ptr, _i64, _i64, _ptr, _ptr, _ptr]),
    "pe_pack": (_int, [_ptr, _int, _i64, _i64, _ptr, _ptr, _i64, _ptr, _ptr, _ptr]),
    "pe_stream_read": (_int, [_ptr, _i64, _int, _ptr, _ptr]),
    "pe_event_create": (_int, [ctypes.POINTER(_ptr)]),
    "pe_event_destroy": (_int, [_ptr]),
    "pe_stream_wait": (_int, [_ptr, _ptr, _ptr]),
    "pe_stream_set_concurrent": (_int, [_ptr, _int]),
}

_lock = threading.Lock()
_library = None


def register(name, restype, argtypes):
    """Declare an additional entry point (used by the other modules)."""
    _SIGNATURES[name] = (restype, argtypes)
    if _library is not None:
        _bind(_library, name)


def _bind(lib, name):
    fn = getattr(lib, name)
    fn.restype, fn.argtypes = _SIGNATURES[name]


def library():
    """Load and bind the HIP library (once per process)."""
    global _library
    if _library is None:
        with _lock:
            if _library is None:
                if not LIBRARY_PATH.exists():
                    raise RuntimeError(
                        f"{LIBRARY_PATH.name} is not built; run python panoptic-evaluator_amd/build.py")
                lib = ctypes.CDLL(str(LIBRARY_PATH))
                for name in _SIGNATURES:
                    _bind(lib, name)
                if lib.pe_abi_version() != ABI_VERSION:
                    raise RuntimeError("libpanoptic_hip.so ABI version mismatch; rebuild it")
                _library = lib
    return _library


_tables: "collections.OrderedDict[tuple, tuple[torch.Tensor, int, set]]" = collections.OrderedDict()
_TABLE_CACHE = 256
_pins: "list | None" = None   # tables handed out while a capture is recorded


@contextlib.contextmanager
def pinned_tables():
    """Collect every table device_table hands out inside the block (cache
    hits included) in the yielded list.  A HIP graph captured inside the
    block bakes those tables' addresses into its kernels, so its owner keeps
    the list for the graph's lifetime: the cache may evict the entries, but
    the caching allocator cannot reuse their memory while the list holds
    them."""
    global _pins
    with _lock:
        if _pins is not None:
            raise RuntimeError("pinned_tables blocks do not nest")
        _pins = []
        pins = _pins
    try:
        yield pins
    finally:
        with _lock:
            _pins = None


def device_table(values, device) -> torch.Tensor:
    """Read-only int64 device copy of a short host list (offsets, pointer
    tables), cached by content: repeated batch layouts upload nothing, so an
    update issued after a warm-up call performs no host-to-device copy (as HIP
    graph capture requires).  Callers must never write to the result.

    A table may be handed out on several streams; each entry remembers the
    streams other than its allocation stream that requested it, and an
    evicted entry is marked as used by them, so the caching allocator reuses
    its memory only after their work (the allocation stream is ordered by the
    allocator itself)."""
    device = torch.device(device)
    if device.index is None:
        device = torch.device(device.type, torch.cuda.current_device())
    key = (device.index, tuple(int(v) for v in values))
    current = stream(device)
    with _lock:
        entry = _tables.get(key)
        if entry is not None:
            _tables.move_to_end(key)
            if current != entry[1]:
                entry[2].add(current)
            if _pins is not None:
                _pins.append(entry[0])
            return entry[0]
    table = torch.tensor(key[1], dtype=torch.int64).to(device)
    evicted = []
    with _lock:
        if _pins is not None:
            _pins.append(table)
        _tables[key] = (table, current, set())
        while len(_tables) > _TABLE_CACHE:
            evicted.append(_tables.popitem(last=False)[1])
    for old, _, streams in evicted:
        for handle in streams:
            old.record_stream(torch.cuda.ExternalStream(handle, device=old.device))
    return table


class SideStream:
    """A second stream of default priority on ``device`` with two reusable
    events: ``fork(main)`` orders it after the work issued so far on the raw
    stream handle ``main``; ``join(main)`` makes ``main`` wait for it.  Both
    are plain event record/wait pairs, so they also fork and join a side
    branch inside HIP graph capture."""

    def __init__(self, device):
        self.stream = torch.cuda.Stream(device)
        self.handle = self.stream.cuda_stream
        lib = library()
        self._events = []
        for _ in range(2):
            event = _ptr()
            check(lib.pe_event_create(ctypes.byref(event)))
            self._events.append(event.value)
        # Its latency-bound launches use narrow workgroups, which fit beside
        # the main stream's kernel (torch reuses pooled stream handles, so the
        # mark is dropped with this object).
        check(lib.pe_stream_set_concurrent(self.handle, 1))
        weakref.finalize(self, _destroy_events, list(self._events), self.handle)

    def fork(self, main: int):
        call("pe_stream_wait", self.handle, main, self._events[0])

    def join(self, main: int):
        call("pe_stream_wait", main, self.handle, self._events[1])


def _destroy_events(events, handle=None):
    if _library is not None:
        for event in events:
            _library.pe_event_destroy(event)
        if handle is not None:
            _library.pe_stream_set_concurrent(handle, 0)


def ptr(tensor):
    return None if tensor is None else tensor.data_ptr()


_raw_stream = getattr(torch._C, "_cuda_getCurrentRawStream", None)


def stream(device=None) -> int:
    """The current HIP stream of `device` as an integer handle (every launch
    asks, so this avoids building a torch Stream object when torch exposes
    the raw handle)."""
    if _raw_stream is None:
        return torch.cuda.current_stream(device).cuda_stream
    if isinstance(device, torch.device) and device.index is not None:
        index = device.index
    elif isinstance(device, int):
        index = device
    elif device is None or torch.device(device).index is None:
        index = torch.cuda.current_device()
    else:
        index = torch.device(device).index
    return _raw_stream(index)


def check(code: int) -> int:
    """Raise on negative status; return validation bits (>= 0) unchanged."""
    if code < 0:
        message = library().pe_last_error().decode(errors="replace")
        raise RuntimeError(f"panoptic HIP kernels: {message}")
    return code


def call(name, *args) -> int:
    return check(getattr(library(), name)(*args))


def profile_read(kernel: str):
    """(total milliseconds, launches) of a kernel since pe_profile_reset()."""
    total, count = _dbl(), _i64()
    call("pe_profile_read", kernel.encode(), ctypes.byref(total), ctypes.byref(count))
    return total.value, count.value


def require_gpu(device: torch.device) -> torch.device:
    """The product path evaluates HIP device tensors only; no CPU fallback."""
    device = torch.device(device)
    if device.type != "cuda":
        raise ValueError(
            "this panoptic_evaluator build evaluates on AMD GPUs (device='cuda'); "
            f"got device {device}")
    if not torch.cuda.is_available():
        raise RuntimeError("no HIP device is available")
    return device
