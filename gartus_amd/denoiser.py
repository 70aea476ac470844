"""Python host for the learned denoiser (include/gartus_denoise.h).

Images are [n][h][w][3] float32: numpy arrays (copied to the device) or
torch CUDA tensors (used in place; torch's stream is synchronised first).
The network runs on the HIP library; nothing here computes.
"""
from __future__ import annotations

import ctypes as C
import json
from pathlib import Path
from typing import Optional, Sequence

import numpy as np

from . import abi
from .api import Context, default_context


class _DeviceArrays:
    """Device copies of host arrays, reused across calls of one shape."""

    def __init__(self, ctx: Context):
        self.ctx = ctx
        self.bufs: dict[int, tuple[int, C.c_void_p]] = {}

    def put(self, slot: int, a: np.ndarray) -> C.c_void_p:
        a = np.ascontiguousarray(a, np.float32)
        have = self.bufs.get(slot)
        if have is None or have[0] < a.nbytes:
            if have is not None:
                self.ctx.lib.gt_device_free(self.ctx.handle, have[1])
            p = C.c_void_p()
            abi.check(self.ctx.lib.gt_device_malloc(self.ctx.handle, max(a.nbytes, 4), C.byref(p)))
            self.bufs[slot] = have = (a.nbytes, p)
        abi.check(self.ctx.lib.gt_memcpy_to_device(self.ctx.handle, have[1], a.ctypes.data_as(C.c_void_p), a.nbytes))
        return have[1]

    def get(self, slot: int, nbytes: int) -> C.c_void_p:
        have = self.bufs.get(slot)
        if have is None or have[0] < nbytes:
            if have is not None:
                self.ctx.lib.gt_device_free(self.ctx.handle, have[1])
            p = C.c_void_p()
            abi.check(self.ctx.lib.gt_device_malloc(self.ctx.handle, max(nbytes, 4), C.byref(p)))
            self.bufs[slot] = have = (nbytes, p)
        return have[1]

    def close(self):
        for _, p in self.bufs.values():
            self.ctx.lib.gt_device_free(self.ctx.handle, p)
        self.bufs.clear()


def _is_torch(a) -> bool:
    return type(a).__module__.startswith("torch")


def _is_device(a) -> bool:
    """A library-allocated device buffer (denoise_train.DeviceArray)."""
    return hasattr(a, "ptr") and hasattr(a, "shape") and not _is_torch(a)


class Denoiser:
    """U-Net denoiser: `levels` resolutions with `channels[l]` features."""

    def __init__(self, context: Optional[Context] = None, levels: int = 5,
                 channels: Sequence[int] = (32, 48, 64, 96, 128), seed: int = 1,
                 loss: int = abi.GT_DENOISE_LOSS_REL_L2, halves: bool = False):
        self.ctx = context or default_context()
        cfg = abi.GtDenoiserConfig()
        cfg.levels = levels
        for i, c in enumerate(channels[:levels]):
            cfg.channels[i] = c
        cfg.seed = seed
        cfg.loss = loss
        cfg.beauty_halves = 1 if halves else 0
        self.loss, self.halves = loss, bool(halves)
        self.levels, self.channels = levels, tuple(channels[:levels])
        h = C.c_void_p()
        abi.check(self.ctx.lib.gt_denoiser_create(self.ctx.handle, C.byref(cfg), C.byref(h)))
        self.ctx._adopt(self)
        self.handle = h
        n = C.c_uint64()
        abi.check(self.ctx.lib.gt_denoiser_param_count(h, C.byref(n)))
        self.param_count = n.value
        self._dev = _DeviceArrays(self.ctx)

    # parameters ----------------------------------------------------------
    def get_params(self) -> np.ndarray:
        out = np.zeros(self.param_count, np.float32)
        abi.check(self.ctx.lib.gt_denoiser_get_params(self.handle, out.ctypes.data_as(abi.F)))
        return out

    def set_params(self, p: np.ndarray) -> None:
        p = np.ascontiguousarray(p, np.float32)
        if p.size != self.param_count:
            raise abi.InvalidArgument(abi.GT_EINVAL, "parameter count mismatch")
        abi.check(self.ctx.lib.gt_denoiser_set_params(self.handle, p.ctypes.data_as(abi.F)))

    def config(self) -> dict:
        return {"levels": self.levels, "channels": list(self.channels), "beauty_halves": self.halves,
                "loss": "rel-l2" if self.loss == abi.GT_DENOISE_LOSS_REL_L2 else "l1-tonemap"}

    def save(self, path: str, **info) -> None:
        """The parameters as .npy and, beside them (same stem, .json), the
        network's configuration with `info` (how it was trained)."""
        np.save(path, self.get_params(), allow_pickle=False)
        Path(path).with_suffix(".json").write_text(json.dumps(dict(self.config(), **info), indent=1) + "\n")

    @classmethod
    def from_file(cls, path: str, context: Optional[Context] = None) -> "Denoiser":
        """A network built from the configuration saved beside `path` (the
        standard one when there is none) with the parameters in `path`."""
        meta = Path(path).with_suffix(".json")
        cfg = json.loads(meta.read_text()) if meta.exists() else {}
        loss = abi.GT_DENOISE_LOSS_L1_TONEMAP if cfg.get("loss") == "l1-tonemap" else abi.GT_DENOISE_LOSS_REL_L2
        dn = cls(context, levels=int(cfg.get("levels", 5)), channels=tuple(cfg.get("channels", (32, 48, 64, 96, 128))),
                 loss=loss, halves=bool(cfg.get("beauty_halves", False)))
        dn.load(path)
        return dn

    def load(self, path: str) -> None:
        self.set_params(np.load(path, allow_pickle=False))

    # inputs --------------------------------------------------------------
    def _inputs(self, beauty, albedo, normal, target=None):
        """A halves network takes beauty as [2][n][h][w][3] (or [2][h][w][3]):
        two renders of half the samples each."""
        arrays = [beauty, albedo, normal] + ([target] if target is not None else [])
        shape = tuple(beauty.shape)
        if self.halves:
            if len(shape) < 4 or shape[0] != 2:
                raise abi.InvalidArgument(abi.GT_EINVAL, "a halves denoiser takes beauty as [2][n][h][w][3]")
            shape = shape[1:]
        if len(shape) == 3:
            shape = (1,) + shape
        if len(shape) != 4 or shape[3] != 3:
            raise abi.InvalidArgument(abi.GT_EINVAL, "images must be [n][h][w][3]")
        for a in arrays[1:]:
            if tuple(a.shape)[-3:] != shape[1:] or (len(a.shape) == 4 and a.shape[0] != shape[0]):
                raise abi.InvalidArgument(abi.GT_EINVAL, "image shapes differ")
        if _is_device(beauty):
            ptrs = [a.ptr for a in arrays]
        elif _is_torch(beauty):
            import torch
            for a in arrays:
                if not (a.is_cuda and a.dtype == torch.float32 and a.is_contiguous()):
                    raise abi.InvalidArgument(abi.GT_EINVAL, "torch inputs must be contiguous float32 CUDA tensors")
            torch.cuda.synchronize()
            ptrs = [C.c_void_p(a.data_ptr()) for a in arrays]
        else:
            ptrs = [self._dev.put(i, np.asarray(a)) for i, a in enumerate(arrays)]
        return shape, ptrs

    # calls ---------------------------------------------------------------
    def infer(self, beauty, albedo, normal, out=None):
        """Denoised radiance, same shape as beauty (numpy in, numpy out;
        torch in, torch out unless `out` is given; device arrays in, `out`
        required)."""
        shape, ptrs = self._inputs(beauty, albedo, normal)
        n, h, w, _ = shape
        if _is_device(beauty):
            if out is None:
                raise abi.InvalidArgument(abi.GT_EINVAL, "device-array inputs need a device-array out")
            abi.check(self.ctx.lib.gt_denoiser_infer(self.handle, n, h, w, *ptrs, out.ptr))
            return out
        if _is_torch(beauty):
            import torch
            res = out if out is not None else torch.empty((n, h, w, 3), dtype=torch.float32, device=beauty.device)
            dst = C.c_void_p(res.data_ptr())
        else:
            res = np.zeros((n, h, w, 3), np.float32)
            dst = self._dev.get(10, res.nbytes)
        abi.check(self.ctx.lib.gt_denoiser_infer(self.handle, n, h, w, *ptrs, dst))
        if not _is_torch(beauty):
            abi.check(self.ctx.lib.gt_memcpy_to_host(self.ctx.handle, res.ctypes.data_as(C.c_void_p), dst, res.nbytes))
            if len(albedo.shape) == 3:
                res = res[0]
        return res

    def loss_and_gradients(self, beauty, albedo, normal, target):
        shape, ptrs = self._inputs(beauty, albedo, normal, target)
        n, h, w, _ = shape
        loss = C.c_float()
        g = np.zeros(self.param_count, np.float32)
        abi.check(self.ctx.lib.gt_denoiser_loss_and_gradients(self.handle, n, h, w, *ptrs, C.byref(loss),
                                                              g.ctypes.data_as(abi.F)))
        return float(loss.value), g

    def train_step(self, beauty, albedo, normal, target, lr: float = 1e-4) -> float:
        shape, ptrs = self._inputs(beauty, albedo, normal, target)
        n, h, w, _ = shape
        loss = C.c_float()
        abi.check(self.ctx.lib.gt_denoiser_train_step(self.handle, n, h, w, *ptrs, C.c_float(lr), C.byref(loss)))
        return float(loss.value)

    def close(self) -> None:
        self._release()

    def _release(self) -> None:
        """Frees the network and its buffers while the context is alive (a
        closed context has released them already)."""
        if getattr(self, "handle", None) and self.ctx.handle:
            self._dev.close()
            self.ctx.lib.gt_denoiser_destroy(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
