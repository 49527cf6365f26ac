"""Device buffers and on-device frame generation through the C ABI."""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from .ldpc import Context, LdpcCode


class DeviceBuffer:
    """A hipMalloc'd buffer owned by a Context (gf2l_device_alloc)."""

    def __init__(self, ctx: Context, nbytes: int):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p()
        _lib.check(ctx.handle, _lib.load().gf2l_device_alloc(ctx.handle, self.nbytes, ctypes.byref(p)))
        self.ptr = p.value

    @classmethod
    def from_array(cls, ctx: Context, arr: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(arr)
        b = cls(ctx, a.nbytes)
        b.upload(a)
        return b

    def upload(self, arr: np.ndarray) -> None:
        a = np.ascontiguousarray(arr)
        if a.nbytes > self.nbytes:
            raise ValueError("array larger than the buffer")
        _lib.check(self.ctx.handle, _lib.load().gf2l_copy_to_device(
            self.ctx.handle, self.ptr, a.ctypes.data, a.nbytes))

    def download(self, dtype, shape, offset: int = 0) -> np.ndarray:
        """Copies np.prod(shape) elements starting `offset` bytes into the buffer."""
        out = np.empty(shape, dtype=dtype)
        if offset < 0 or offset + out.nbytes > self.nbytes:
            raise ValueError("requested more bytes than the buffer holds")
        _lib.check(self.ctx.handle, _lib.load().gf2l_copy_to_host(
            self.ctx.handle, out.ctypes.data, self.ptr + offset, out.nbytes))
        return out

    def free(self) -> None:
        if self.ptr:
            _lib.load().gf2l_device_free(self.ctx.handle, self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def generate_frames(code: LdpcCode, seed: int, first_frame: int, frames: int, ebn0_lo: float,
                    ebn0_hi: float = None, with_codewords: bool = True):
    """Frames generated on the device (gf2l_frames_generate).  Returns
    (llr DeviceBuffer [frames, n] float32, codeword DeviceBuffer
    [frames, ceil(n/64)] uint64 or None)."""
    ctx = code.context
    if ebn0_hi is None:
        ebn0_hi = ebn0_lo
    llr = DeviceBuffer(ctx, frames * code.n() * 4)
    cw = DeviceBuffer(ctx, frames * ((code.n() + 63) // 64) * 8) if with_codewords else None
    _lib.check(ctx.handle, _lib.load().gf2l_frames_generate(
        ctx.handle, code.handle, seed, first_frame, frames, ebn0_lo, ebn0_hi, llr.ptr,
        cw.ptr if cw else None))
    ctx.synchronize()
    return llr, cw


def generate_zero_frames(code: LdpcCode, seed: int, first_frame: int, frames: int, ebn0_lo: float,
                         ebn0_hi: float = None) -> DeviceBuffer:
    """All-zero codeword frames of any code on the device
    (gf2l_frames_generate_zero).  Returns the LLR DeviceBuffer [frames, n]."""
    ctx = code.context
    if ebn0_hi is None:
        ebn0_hi = ebn0_lo
    llr = DeviceBuffer(ctx, frames * code.n() * 4)
    _lib.check(ctx.handle, _lib.load().gf2l_frames_generate_zero(
        ctx.handle, code.handle, seed, first_frame, frames, ebn0_lo, ebn0_hi, llr.ptr))
    ctx.synchronize()
    return llr
