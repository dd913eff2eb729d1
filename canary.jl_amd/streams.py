"""Streams owned by the batched and served runtimes.

Each one is created by the library (hipStreamCreateWithFlags, non-blocking,
default priority) and wrapped as a torch ExternalStream, so every runtime
stream is distinct: streams drawn from torch's shared pool wrap around after
32 and can alias a stream that still carries other work, which would then
join a graph capture.

A collected OwnedStream is not destroyed by its finalizer: CPython runs a
finalizer in whichever thread triggers the collection -- a host worker
thread of the server, say -- which could then call hipStreamSynchronize /
hipStreamDestroy while the main thread captures or launches a graph.  The
finalizer only queues the stream; reap() destroys the queued ones, and the
runtimes call it where nothing else is in flight (before a server starts,
after it drains, when a batch is built).
"""
import ctypes as C

import torch

from ._lib import check, lib

_graveyard = []  # streams of collected OwnedStream objects (list.append is atomic)


def reap():
    """Destroy the streams queued by collected OwnedStream objects.  Call it
    from the thread that launches work, with no graph capture in progress."""
    n = len(_graveyard)
    ptrs = _graveyard[:n]
    del _graveyard[:n]
    for p in ptrs:
        lib().canary_stream_destroy(C.c_void_p(p))
    return len(ptrs)


class OwnedStream:
    def __init__(self, device="cuda"):
        p = C.c_void_p()
        check(lib().canary_stream_create(C.byref(p)), "stream_create")
        self.ptr = p.value
        self.stream = torch.cuda.ExternalStream(self.ptr, device=torch.device(device))

    def __del__(self):
        ptr = getattr(self, "ptr", None)
        if ptr:
            self.ptr = None
            _graveyard.append(ptr)
