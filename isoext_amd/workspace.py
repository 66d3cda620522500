"""Device workspaces reused across calls.

Each (calling thread, device, stream, kind) gets one zero-initialised byte
buffer that grows on demand; the kernels keep their intermediate state in it
(inside masks, count arrays, special-vertex lists, scan partials) and reset
what must start at zero themselves.  A call runs in phases with host
synchronisations between them (counts, then outputs; kernels, then the weld
fix-up), so a buffer must not be shared by two host threads even when they
use the same stream: the key includes the thread, and a thread's buffers are
freed when the thread ends.
"""

from __future__ import annotations

import threading

import torch

_local = threading.local()


def _buffers() -> dict:
    d = getattr(_local, "buffers", None)
    if d is None:
        d = _local.buffers = {}
    return d


def get(kind: str, nbytes: int, device: torch.device, stream: int | None = None) -> torch.Tensor:
    """stream: the current stream's handle on `device`, when the caller has
    it already."""
    if stream is None:
        stream = torch.cuda.current_stream(device).cuda_stream
    key = (kind, device.index, stream)
    bufs = _buffers()
    buf = bufs.get(key)
    if buf is None or buf.numel() < nbytes:
        size = max(int(nbytes), 1 << 20)
        if buf is not None:
            size = max(size, int(buf.numel() * 1.5))
        buf = torch.zeros(size, dtype=torch.uint8, device=device)
        bufs[key] = buf
    return buf


def release_all() -> None:
    """Drop the calling thread's workspaces."""
    _buffers().clear()
