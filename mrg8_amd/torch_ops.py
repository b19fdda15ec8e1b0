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

* ``torch.ops.mrg8.advance_(states, offsets)`` jumps rows of an (n, 8) state
  tensor by up to 256-bit offsets on the device; ``substream_states(seed, n)``
  builds n substreams spaced 2^100 apart with it.

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


@torch.library.custom_op("mrg8::advance_", mutates_args=("states",))
def advance_(states: torch.Tensor, offsets: torch.Tensor) -> None:
    """Jumps each row of `states` ((n, 8) int32/uint32 words, newest first) ahead
    by the matching row of `offsets` ((n, words) int64/uint64, little-endian
    64-bit words, words 1..4: up to 256-bit offsets), in place, on the current
    stream (the bulk device jump of mrg8_advance_states)."""
    if states.device.type != "cuda" or offsets.device != states.device:
        raise ValueError("states and offsets must be HIP tensors on one device")
    if states.dtype not in (torch.int32, torch.uint32) or states.dim() != 2 or states.shape[1] != 8:
        raise TypeError("states must be an (n, 8) int32 or uint32 tensor")
    if offsets.dtype not in (torch.int64, torch.uint64) or offsets.dim() != 2 or offsets.shape[0] != states.shape[0]:
        raise TypeError("offsets must be an (n, words) int64 or uint64 tensor")
    if not (states.is_contiguous() and offsets.is_contiguous()):
        raise ValueError("states and offsets must be contiguous")
    n, words = states.shape[0], offsets.shape[1]
    if n == 0:
        return
    from . import advance_states

    advance_states(states, offsets, words, states, n, stream=torch.cuda.current_stream(states.device),
                   device=states.device.index if states.device.index is not None else torch.cuda.current_device())


@advance_.register_fake
def _(states, offsets):
    return None


def substream_states(seed: int, n: int, log2_spacing: int = 100, device=None) -> torch.Tensor:
    """(n, 8) int32 states: row i is the stream of `seed` (seeded like the
    reference's mrg8(seed)) advanced by i * 2^log2_spacing positions, computed
    on the device -- n independent substreams for n workers."""
    from . import seed_state

    device = torch.device(device if device is not None else "cuda")
    if n < 0 or log2_spacing < 0 or max(n - 1, 0).bit_length() + log2_spacing > 256:
        raise ValueError("need n >= 0 and n * 2^log2_spacing below 2^256")
    if n == 0:
        return torch.empty((0, 8), dtype=torch.int32, device=device)
    words = max(1, ((max(n - 1, 1)).bit_length() + log2_spacing + 63) // 64)
    base = torch.tensor(seed_state(seed).astype("int64"), dtype=torch.int64).to(torch.int32)
    states = base.repeat(n, 1).to(device)
    offs = [[((i << log2_spacing) >> (64 * w)) & (2**64 - 1) for w in range(words)] for i in range(n)]
    offsets = torch.tensor(offs, dtype=torch.uint64).to(device)
    torch.ops.mrg8.advance_(states, offsets)
    return states


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


__all__ = ["Generator", "fill_", "advance_", "substream_states", "F64", "F64_01", "U32"]
