"""GPU-resident batches: a ``Layout`` buffer, its descriptor array and the
argument records, all in GPU memory allocated through torch (used here only
as a device allocator and for host<->device copies)."""
import numpy as np

from . import DMAT


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("blasfeo_amd: no GPU visible to torch")
    return torch


def to_device(arr):
    """numpy array -> GPU uint8 tensor holding the same bytes."""
    torch = _torch()
    a = np.ascontiguousarray(arr)
    return torch.from_numpy(a.view(np.uint8).reshape(-1).copy()).to("cuda")


class DeviceBatch:
    """Matrices of a Layout in one GPU buffer plus their GPU descriptors."""

    def __init__(self, layout, host_buf):
        torch = _torch()
        self.layout = layout
        self.buf = torch.from_numpy(np.ascontiguousarray(host_buf)).to("cuda")
        self.desc_host = layout.build(self.buf.data_ptr())
        self.desc = to_device(self.desc_host)

    def desc_addr(self, i):
        return self.desc.data_ptr() + int(i) * DMAT.itemsize

    def download(self):
        return self.buf.cpu().numpy()

    def download_desc(self):
        return self.desc.cpu().numpy().view(DMAT)


class HostBatch:
    """The same Layout in host memory, for the CPU oracle."""

    def __init__(self, layout, host_buf):
        self.layout = layout
        self.buf = np.array(host_buf, copy=True)
        self.desc = layout.build(self.buf.ctypes.data)

    def desc_addr(self, i):
        return self.desc.ctypes.data + int(i) * DMAT.itemsize
