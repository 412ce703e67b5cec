"""A minimal device array of "another library" for the interchange tests: it owns a torch device
tensor and exposes only DLPack (as CuPy does), and the module has from_dlpack (as cupy does)."""


class DevArray:
    def __init__(self, t):
        self._t = t
        self.shape = tuple(t.shape)

    def __dlpack__(self, stream=None):
        return self._t.__dlpack__()

    def __dlpack_device__(self):
        return self._t.__dlpack_device__()

    def torch(self):
        return self._t


def from_dlpack(x):
    import torch

    return DevArray(torch.from_dlpack(x))
