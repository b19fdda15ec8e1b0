"""PyTorch-ROCm caller: MRG8 fills of torch HIP tensors through the C ABI.

* ``Generator``: one MRG8 stream, seeded like the reference's ``mrg8_cuda``
  (cuda/mrg8_cuda.cu:22-63); ``rand``/``randint`` allocate a tensor and
  ``rand_``/``randint_`` fill a given one, on torch's current HIP stream, and
  successive calls continue the sequence (the reference's d_state hand-off,
  cuda/mrg8_cuda.cu:424).
* ``torch.ops.mrg8.fill_(out, state, offset, fmt)``: a registered operator
  with no hidden state.  It writes stream positions offset .. offset+numel-1
  of the stream whose 8-word state is ``state`` into the contiguous tensor
  ``out``, on the current stream.  ``offset`` is a schema ``int`` (signed
  64-bit), so it reaches 2^63 - 1; larger offsets go through ``Generator``.

Formats: ``F64`` (x / p in (0, 1], original/mrg8.cpp:233-236) and ``F64_01``
((x - 1) / p in [0, 1), cuda/mrg8_cuda.cu:363) fill float64 tensors; ``U32``
(the raw word x in [1, p]) fills int32 or uint32 tensors (p = 2^31 - 1 fits
either exactly).  There is no CPU path: CPU tensors are rejected.
"""

import threading
from typing import List

import torch

from . import F64, F64_01, U32, mrg8_cuda

_DTYPES = {F64: (torch.float64,), F64_01: (torch.float64,), U32: (torch.int32, torch.uint32)}

_lock = threading.Lock()
_handles: dict = {}


def _check_out(out: torch.Tensor, fmt: int) -> None:
    if fmt not in _DTYPES:
        raise ValueError(f"unknown format {fmt}")
    if out.device.type != "cuda":
        raise ValueError("MRG8 fills HIP device tensors only")
    if out.dtype not in _DTYPES[fmt]:
        raise TypeError(f"format {fmt} writes {[str(d) for d in _DTYPES[fmt]]}, got {out.dtype}")
    if not out.is_contiguous():
        raise ValueError("output tensor must be contiguous")


def _handle(device: torch.device) -> mrg8_cuda:
    idx = device.index if device.index is not None else torch.cuda.current_device()
    h = _handles.get(idx)
    if h is None:
        with torch.cuda.device(idx):
            h = mrg8_cuda(0)
        _handles[idx] = h
    return h


@torch.library.custom_op("mrg8::fill_", mutates_args=("out",))
def fill_(out: torch.Tensor, state: List[int], offset: int, fmt: int) -> None:
    """Writes positions offset .. offset+numel-1 of the stream `state` into `out`."""
    _check_out(out, fmt)
    if len(state) != 8:
        raise ValueError("state must have 8 words")
    if offset < 0:
        raise ValueError("offset must be non-negative")
    n = out.numel()
    if n == 0:
        return
    with _lock, torch.cuda.device(out.device):
        h = _handle(out.device)
        h.set_state(state)
        h.jump_ahead(offset)
        h.generate(out, n, fmt, stream=torch.cuda.current_stream(out.device))


@fill_.register_fake
def _(out, state, offset, fmt):
    return None


class Generator:
    """One MRG8 stream producing torch tensors (see the module docstring)."""

    def __init__(self, seed: int = 0, device=None):
        self.device = torch.device(device if device is not None else "cuda")
        if self.device.type != "cuda":
            raise ValueError("MRG8 generates on HIP devices only")
        if self.device.index is None:
            self.device = torch.device("cuda", torch.cuda.current_device())
        with torch.cuda.device(self.device):
            self._g = mrg8_cuda(seed)

    def get_state(self):
        return self._g.get_state()

    def set_state(self, state) -> None:
        self._g.set_state(state)

    def jump_ahead(self, step: int) -> None:
        self._g.jump_ahead(step)

    def _fill(self, out: torch.Tensor, fmt: int) -> torch.Tensor:
        _check_out(out, fmt)
        if out.device != self.device:
            raise ValueError(f"tensor on {out.device}, generator on {self.device}")
        if out.numel():
            with torch.cuda.device(self.device):
                self._g.generate(out, out.numel(), fmt, stream=torch.cuda.current_stream(self.device))
        return out

    def rand_(self, out: torch.Tensor, fmt: int = F64) -> torch.Tensor:
        """Fills a float64 tensor with the next numel() outputs as doubles."""
        return self._fill(out, fmt)

    def randint_(self, out: torch.Tensor) -> torch.Tensor:
        """Fills an int32/uint32 tensor with the next numel() raw words."""
        return self._fill(out, U32)

    def rand(self, *size, fmt: int = F64) -> torch.Tensor:
        return self._fill(torch.empty(*size, dtype=torch.float64, device=self.device), fmt)

    def randint(self, *size, dtype=torch.int32) -> torch.Tensor:
        return self._fill(torch.empty(*size, dtype=dtype, device=self.device), U32)


__all__ = ["Generator", "fill_", "F64", "F64_01", "U32"]
