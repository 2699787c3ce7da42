"""Device-memory plumbing: torch ROCm tensors carry every buffer the C ABI sees.

torch is used only for allocation, host<->device copies and stream queries;
all arithmetic on the path runs in the library's HIP kernels.
"""

from __future__ import annotations

from typing import Any

import numpy as np

_TORCH_COMPLEX = {np.dtype(np.complex64): "complex64", np.dtype(np.complex128): "complex128"}


def torch():
    import torch as _torch

    return _torch


def is_tensor(x: Any) -> bool:
    try:
        import torch as _torch
    except ImportError:  # pragma: no cover - torch is a hard dependency
        return False
    return isinstance(x, _torch.Tensor)


def device():
    t = torch()
    if not t.cuda.is_available():
        raise RuntimeError("rcabeam MI355X kernels need a HIP GPU; none is visible")
    return t.device("cuda", t.cuda.current_device())


# Host arrays at least this large go through the library's pinned staging (rcabeam_upload_host).
STAGED_UPLOAD_BYTES = 32 << 20


def upload_threads() -> int:
    """Host copy threads for staged uploads: the CPUs this process may run on, at most 16."""
    import os

    return max(1, min(16, len(os.sched_getaffinity(0))))


def upload(arr: np.ndarray, dtype) -> Any:
    """Device tensor of `dtype` holding the host array `arr`, copied through pinned staging buffers
    (rcabeam_upload_host: threaded host copies overlapped with the host-to-device DMA)."""
    from rcabeam_amd import _lib

    t = torch()
    host = np.ascontiguousarray(arr, dtype=t.empty((), dtype=dtype).numpy().dtype)
    out = t.empty(host.shape, dtype=dtype, device=device())
    _lib.call("rcabeam_upload_host", host.ctypes.data, out.data_ptr(), host.nbytes, upload_threads())
    return out


def to_device(x: Any, dtype) -> Any:
    """Contiguous device tensor of `dtype` from a NumPy array or a tensor (copied only if needed)."""
    t = torch()
    dev = device()
    if isinstance(x, t.Tensor):
        y = x.to(device=dev, dtype=dtype)
        return y.contiguous()
    arr = np.asarray(x)
    if arr.nbytes >= STAGED_UPLOAD_BYTES and arr.dtype.kind in "fc":
        return upload(arr, dtype)
    arr = np.ascontiguousarray(arr)
    return t.from_numpy(arr).to(device=dev, dtype=dtype, non_blocking=False).contiguous()


def empty(shape, dtype) -> Any:
    t = torch()
    return t.empty(shape, dtype=dtype, device=device())


def ptr(x: Any) -> int:
    """Device address of a tensor (0 for None or an empty tensor)."""
    if x is None:
        return 0
    return int(x.data_ptr()) if x.numel() else 0


def synchronize() -> None:
    torch().cuda.synchronize()


def on_default_stream() -> bool:
    t = torch()
    stream = t.cuda.current_stream()
    return stream == t.cuda.default_stream(stream.device)


def require_default_stream(message: str) -> None:
    """The C ABI launches on the null stream, so callers must be on torch's default stream."""
    if not on_default_stream():
        raise ValueError(message)


def host(x: Any) -> np.ndarray:
    return x.detach().cpu().numpy()
