"""Streams owned by the batched and served runtimes.

Each one is created by the library (hipStreamCreateWithFlags, non-blocking,
default priority) and wrapped as a torch ExternalStream, so every runtime
stream is distinct: streams drawn from torch's shared pool wrap around after
32 and can alias a stream that still carries other work, which would then
join a graph capture.  The stream is released when the object is collected.
"""
import ctypes as C

import torch

from ._lib import check, lib


class OwnedStream:
    def __init__(self, device="cuda"):
        p = C.c_void_p()
        check(lib().canary_stream_create(C.byref(p)), "stream_create")
        self.ptr = p.value
        self.stream = torch.cuda.ExternalStream(self.ptr, device=torch.device(device))

    def __del__(self):
        if getattr(self, "ptr", None):
            lib().canary_stream_destroy(C.c_void_p(self.ptr))
            self.ptr = None
