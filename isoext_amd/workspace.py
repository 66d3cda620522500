"""Device workspaces reused across calls.

Each (device, stream, kind) gets one zero-initialised byte buffer that grows
on demand; the kernels keep their intermediate state in it (inside masks,
count arrays, special-vertex lists, scan partials) and reset what must start
at zero themselves.  Calls sharing one buffer are serialised by their
stream.
"""

from __future__ import annotations

import threading

import torch

_lock = threading.Lock()
_buffers: dict = {}


def get(kind: str, nbytes: int, device: torch.device) -> torch.Tensor:
    stream = torch.cuda.current_stream(device).cuda_stream
    key = (kind, device.index, stream)
    with _lock:
        buf = _buffers.get(key)
        if buf is None or buf.numel() < nbytes:
            size = max(int(nbytes), 1 << 20)
            if buf is not None:
                size = max(size, int(buf.numel() * 1.5))
            buf = torch.zeros(size, dtype=torch.uint8, device=device)
            _buffers[key] = buf
        return buf


def release_all() -> None:
    with _lock:
        _buffers.clear()
