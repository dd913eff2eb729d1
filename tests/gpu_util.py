"""Host <-> device helpers for the GPU tests (flat tensors in the reference
column-major memory order)."""
import numpy as np
import torch


def dev(a, dtype=None):
    a = np.asarray(a)
    flat = np.ascontiguousarray(a.reshape(-1, order="F"))
    t = torch.from_numpy(flat)
    if dtype is not None:
        t = t.to(dtype)
    return t.to("cuda")


def host(t, shape):
    return t.detach().cpu().numpy().reshape(shape, order="F")


class DeviceSetup:
    """Device copies of a problems.Setup's kernel inputs."""

    def __init__(self, s):
        self.s = s
        self.vgeo = dev(s.vgeo)
        self.sgeo = dev(s.sgeo)
        self.D = dev(s.D)
        self.vmapM = dev(s.vmapM)
        self.vmapP = dev(s.vmapP)
        self.bndy = dev(s.elemtobndy)
