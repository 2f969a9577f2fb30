"""Device context: one ``mgpbd_ctx`` per HIP device, bound to torch's current stream.

torch supplies device memory (tensors) and the stream; every computation is a libmgpbd
kernel. Without a visible GPU the constructors raise -- there is no CPU fallback.
"""

from __future__ import annotations

import numpy as np
import torch

from . import _lib
from ._lib import MGPBDError, check

_CTX = {}


class Context:
    def __init__(self, device: int):
        self.lib = _lib.load()
        self.device = torch.device("cuda", device)
        with torch.cuda.device(self.device):
            self._stream = torch.cuda.current_stream(self.device).cuda_stream
            self.handle = self.lib.mgpbd_ctx_create(self._stream)
        if not self.handle:
            raise MGPBDError("mgpbd_ctx_create: " + self.lib.mgpbd_last_error().decode())
        self.scalars = torch.zeros(16, dtype=torch.float64, device=self.device)

    def bind_stream(self) -> None:
        s = torch.cuda.current_stream(self.device).cuda_stream
        if s != self._stream:
            check(self.lib.mgpbd_ctx_set_stream(self.handle, s), "mgpbd_ctx_set_stream")
            self._stream = s

    def __call__(self, name: str, *args) -> None:
        check(getattr(self.lib, name)(self.handle, *args), name)

    def sync(self) -> None:
        check(self.lib.mgpbd_ctx_sync(self.handle), "mgpbd_ctx_sync")

    def __del__(self):
        try:
            if getattr(self, "handle", None):
                self.lib.mgpbd_ctx_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


def require_gpu() -> None:
    if not torch.cuda.is_available():
        raise MGPBDError("controllable_pbd_3d_amd needs a HIP GPU (torch.cuda.is_available() "
                         "is False); there is no CPU fallback")


def new_context(device: int | None = None) -> Context:
    """A private context: its multigrid levels, PCG state and captured PCG graph belong to one
    solver (the per-device shared context serves the stateless kernels)."""
    require_gpu()
    dev = torch.cuda.current_device() if device is None else int(device)
    ctx = Context(dev)
    ctx.bind_stream()
    return ctx


def get_context(device: int | None = None) -> Context:
    require_gpu()
    dev = torch.cuda.current_device() if device is None else int(device)
    ctx = _CTX.get(dev)
    if ctx is None:
        ctx = Context(dev)
        _CTX[dev] = ctx
    ctx.bind_stream()
    return ctx


def to_device(a, dtype, device, shape=None, copy=False):
    """numpy / torch -> contiguous torch tensor on ``device`` (no copy when it already is
    one, unless ``copy``)."""
    if isinstance(a, torch.Tensor):
        t = a
        if t.device != device or t.dtype != dtype or not t.is_contiguous() or copy:
            t = t.to(device=device, dtype=dtype).contiguous()
            if copy and t.data_ptr() == a.data_ptr():
                t = t.clone()
    else:
        t = torch.as_tensor(np.ascontiguousarray(np.asarray(a)), dtype=dtype).to(device)
    if shape is not None:
        t = t.reshape(shape)
    return t


def device_bytes(*objs, seen=None) -> int:
    """Bytes of the distinct device tensors reachable from ``objs`` (attributes, dicts, lists),
    each storage counted once; storages whose address is in ``seen`` are skipped, and the
    ones counted are added to it."""
    seen_obj, total = set(), 0
    seen_mem = set() if seen is None else seen
    stack = list(objs)
    while stack:
        o = stack.pop()
        if id(o) in seen_obj or o is None:
            continue
        seen_obj.add(id(o))
        if isinstance(o, torch.Tensor):
            if o.is_cuda:
                st = o.untyped_storage()
                if st.data_ptr() not in seen_mem:
                    seen_mem.add(st.data_ptr())
                    total += st.nbytes()
        elif isinstance(o, dict):
            stack.extend(o.values())
        elif isinstance(o, (list, tuple)):
            stack.extend(o)
        elif hasattr(o, "__dict__") and not isinstance(o, (type, Context)):
            stack.extend(vars(o).values())
    return total
