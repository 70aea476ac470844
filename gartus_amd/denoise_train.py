"""Training data, training loop and error measurement for the learned denoiser.

Data. Seeded scenes in the style of the 1080p GGX/environment workload
(scenes.c3_style_scene), rendered by this library's own path tracer:
  - make_crop_dataset (--crops): one random 256x256 crop of a 1920x1080 frame
    per scene, inputs at 1-4096 spp with their first-hit albedo and normal
    AOVs, target at 16384 spp with another seed;
  - make_dataset: whole 256x256 renders, a 4-spp input and a 1024-spp target.
Input and target noise are independent. The default loss, relative L2 with
the prediction's magnitude held in its denominator, is minimised by the
targets' mean, so noisy but unbiased targets teach the expected radiance.
All images stay in device memory as torch tensors; torch supplies the memory
and the batch gathers only, and the network runs on the HIP library. A process
that uses torch must import it before the library loads (abi.load), which
main() does.

Error. relMSE(x, r) = mean over pixels and channels of (x - r)^2 / (r^2 + eps),
with eps = 1e-2, against a high-spp reference accumulated in FP64.

Usage, on a GPU (writes the trained parameters as .npy):
    python -m gartus_amd.denoise_train --crops --reject-ggx-fallback --scenes 320 --spp-ref 16384 \
        --steps 6000 --out weights.npy
"""
from __future__ import annotations

import argparse
import ctypes as C
import sys
import time

import numpy as np

from . import abi, api, scenes
from .denoiser import Denoiser

RELMSE_EPS = 1e-2
TRIM = 1e-3  # fraction of pixels with the largest error left out of the robust relMSE


def _pixel_relmse(x, ref, eps):
    """Per-pixel relMSE (mean over the 3 channels) as a float64 torch or numpy array."""
    if type(x).__module__.startswith("torch"):
        d = x.double() - ref.double()
        return (d * d / (ref.double() ** 2 + eps)).mean(dim=-1).flatten()
    x = np.asarray(x, np.float64)
    ref = np.asarray(ref, np.float64)
    return ((x - ref) ** 2 / (ref ** 2 + eps)).mean(axis=-1).ravel()


def relmse(x, ref, eps: float = RELMSE_EPS) -> float:
    """Mean of (x - ref)^2 / (ref^2 + eps) over pixels and channels (torch or numpy)."""
    e = _pixel_relmse(x, ref, eps)
    return float(e.mean())


def relmse_trimmed(x, ref, eps: float = RELMSE_EPS, trim: float = TRIM) -> float:
    """relMSE without the `trim` fraction of pixels with the largest error: the
    usual firefly-robust form, since a single heavy-tailed sample can dominate
    the plain mean."""
    e = _pixel_relmse(x, ref, eps)
    n = e.shape[0]
    keep = n - int(n * trim)
    if type(e).__module__.startswith("torch"):
        import torch
        return float(torch.sort(e).values[:keep].mean())
    return float(np.sort(e)[:keep].mean())


def _ptr(a):
    """Device address of a torch tensor or a DeviceArray."""
    return a.data_ptr() if hasattr(a, "data_ptr") else a.ptr.value


class DeviceArray:
    """A device buffer from the library's own allocator (no torch), with a
    host copy on request."""

    def __init__(self, ctx: api.Context, shape, dtype=np.float32):
        self.ctx, self.shape, self.dtype = ctx, tuple(shape), np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        self.ptr = C.c_void_p()
        abi.check(ctx.lib.gt_device_malloc(ctx.handle, max(self.nbytes, 4), C.byref(self.ptr)))

    def sub(self, i: int) -> "DeviceArray":
        """View of entry i along the first axis (shares the memory)."""
        v = DeviceArray.__new__(DeviceArray)
        v.ctx, v.shape, v.dtype = self.ctx, self.shape[1:], self.dtype
        v.nbytes = self.nbytes // self.shape[0]
        v.ptr = C.c_void_p(self.ptr.value + i * v.nbytes)
        v._owner = self
        return v

    def numpy(self) -> np.ndarray:
        out = np.empty(self.shape, self.dtype)
        abi.check(self.ctx.lib.gt_memcpy_to_host(self.ctx.handle, out.ctypes.data_as(C.c_void_p), self.ptr,
                                                 self.nbytes))
        return out

    def __del__(self):
        if getattr(self, "_owner", None) is None and getattr(self, "ptr", None) is not None and self.ptr.value:
            try:
                self.ctx.lib.gt_device_free(self.ctx.handle, self.ptr)
            except Exception:
                pass


class SceneHandle:
    """A compiled scene with its environment and camera, ready to render."""

    def __init__(self, ctx: api.Context, data, env_rgb, cam_spec, flags: int = 0):
        self.ctx = ctx
        self.scene = api.RayScene(data, ctx).with_bvh()
        self.env = api.EnvironmentLight.from_rgb32f(env_rgb)
        self.sh = self.scene._compile()
        self.eh = self.env._compile(ctx)
        self.cam = scenes.camera_record(cam_spec)
        self.cam.direct_lighting_mode = abi.GT_LIGHTING_NEE
        self.cam.flags = flags
        self.w = self.cam.image_width
        self.h = int(self.ctx.lib.gt_camera_image_height(C.byref(self.cam)))

    def camera(self, spp: int, seed: int):
        c = abi.GtCamera()
        C.memmove(C.byref(c), C.byref(self.cam), C.sizeof(c))
        c.samples_per_pixel = spp
        c.pixel_sample_mode = abi.GT_PIXEL_RANDOM
        c.rng_seed = seed
        return c

    def render(self, out, spp: int, seed: int, double: bool = False) -> None:
        """Renders into a device tensor out [h][w][3] (float32, or float64 with double)."""
        c = self.camera(spp, seed)
        abi.check(self.ctx.lib.gt_render_hdr_device(self.ctx.handle, self.sh, C.byref(c), None, self.eh,
                                                    C.c_void_p(_ptr(out)), 1 if double else 0, None))
        self.ctx.synchronize()

    def render_pixels(self, out, spp: int, seed: int, pixels: np.ndarray) -> None:
        """The listed pixels (y * W + x) into a full-frame FP64 device tensor
        out [h][w][3]; the others are left as they were."""
        c = self.camera(spp, seed)
        px = np.ascontiguousarray(pixels, dtype=np.uint32)
        abi.check(self.ctx.lib.gt_render_pixels_device(self.ctx.handle, self.sh, C.byref(c), None, self.eh,
                                                       px.ctypes.data_as(abi.U32), len(px), C.c_void_p(_ptr(out)),
                                                       1, None))
        self.ctx.synchronize()

    def render_aovs(self, beauty, albedo, normal, spp: int, seed: int) -> None:
        c = self.camera(spp, seed)
        abi.check(self.ctx.lib.gt_render_aovs_device(
            self.ctx.handle, self.sh, C.byref(c), None, self.eh, C.cast(_ptr(beauty), abi.F),
            C.cast(_ptr(albedo), abi.F), C.cast(_ptr(normal), abi.F), None))
        self.ctx.synchronize()


def make_dataset(ctx: api.Context, n_scenes: int, first_seed: int = 10_000, size: int = 256, spp_in: int = 4,
                 spp_ref: int = 1024, flags: int = 0, log=print):
    """Device tensors [n][size][size][3]: beauty, albedo, normal (spp_in) and target (spp_ref)."""
    import torch
    shape = (n_scenes, size, size, 3)
    data = {k: torch.empty(shape, dtype=torch.float32, device="cuda") for k in ("beauty", "albedo", "normal", "target")}
    t0 = time.perf_counter()
    for i in range(n_scenes):
        seed = first_seed + i
        sd, env, cam = scenes.c3_style_scene(seed, size, size)
        sh = SceneHandle(ctx, sd, env, cam, flags)
        torch.cuda.synchronize()
        sh.render_aovs(data["beauty"][i], data["albedo"][i], data["normal"][i], spp_in, 2 * seed)
        sh.render(data["target"][i], spp_ref, 2 * seed + 1)
        if (i + 1) % 50 == 0 or i + 1 == n_scenes:
            log(f"dataset: {i + 1}/{n_scenes} scenes rendered ({time.perf_counter() - t0:.1f} s)")
    return data


CROP_SPP_LEVELS = (1, 4, 16, 64, 256, 1024, 4096)
FULL_FRAME_MAX_SPP = 64  # input levels up to this are cut from full-frame AOV renders


def make_crop_dataset(ctx: api.Context, n_scenes: int, first_seed: int = 20_000, size: int = 256,
                      frame=(1920, 1080), spp_levels=CROP_SPP_LEVELS, spp_ref: int = 16384, flags: int = 0,
                      halves: bool = False, log=print):
    """Random size x size crops of full-resolution frames of seeded scenes in
    the style of the GGX/environment workload (scenes.c3_style_scene at
    `frame`), one crop per scene at a seeded position, as device tensors:
      - beauty, albedo, normal [n][levels][size][size][3]: the crop's inputs at
        each spp of spp_levels. Levels up to 64 spp are cut from full-frame
        gt_render_aovs_device renders; higher levels render only the crop's
        pixels (gt_render_pixels_device, each pixel equal to the full
        render's) and take their albedo and normal from the 64-spp level,
        whose first-hit features are already converged;
      - target [n][size][size][3]: the crop at spp_ref from another seed,
        FP64 accumulation, so input and target noise are independent.
    With halves, beauty is [n][levels][2][size][size][3]: each level's
    samples as two renders of half as many (levels must be even), the
    first with the AOVs."""
    import torch
    spp_levels = tuple(sorted(spp_levels))
    if not spp_levels or spp_levels[0] > FULL_FRAME_MAX_SPP:
        raise ValueError(f"the lowest input spp must be at most {FULL_FRAME_MAX_SPP} (it supplies the AOVs)")
    if halves and any(s % 2 for s in spp_levels):
        raise ValueError("with halves every input spp must be even")
    W, H = frame
    L = len(spp_levels)
    shape = (n_scenes, L, size, size, 3)
    data = {k: torch.empty(shape, dtype=torch.float32, device="cuda") for k in ("albedo", "normal")}
    data["beauty"] = torch.empty((n_scenes, L, 2, size, size, 3) if halves else shape, dtype=torch.float32,
                                 device="cuda")
    data["target"] = torch.empty((n_scenes, size, size, 3), dtype=torch.float32, device="cuda")
    full = {k: torch.empty((H, W, 3), dtype=torch.float32, device="cuda") for k in ("beauty", "albedo", "normal")}
    wide = torch.empty((H, W, 3), dtype=torch.float64, device="cuda")
    t0 = time.perf_counter()
    for i in range(n_scenes):
        seed = first_seed + i
        sd, env, cam = scenes.c3_style_scene(seed, W, H)
        sh = SceneHandle(ctx, sd, env, cam, flags)
        pick = np.random.default_rng(seed)
        x0, y0 = int(pick.integers(0, W - size + 1)), int(pick.integers(0, H - size + 1))
        ys, xs = np.mgrid[y0:y0 + size, x0:x0 + size]
        pix = np.ascontiguousarray((ys * W + xs).ravel().astype(np.uint32))
        crop = (slice(y0, y0 + size), slice(x0, x0 + size))
        torch.cuda.synchronize()
        aov_level = None
        for j, spp in enumerate(spp_levels):
            parts = (spp // 2, spp // 2) if halves else (spp,)
            for half, s in enumerate(parts):
                dst = data["beauty"][i, j, half] if halves else data["beauty"][i, j]
                rseed = 16 * seed + 2 * j + half
                if spp <= FULL_FRAME_MAX_SPP:
                    if half == 0:
                        sh.render_aovs(full["beauty"], full["albedo"], full["normal"], s, rseed)
                        data["albedo"][i, j] = full["albedo"][crop]
                        data["normal"][i, j] = full["normal"][crop]
                        aov_level = j
                    else:
                        sh.render(full["beauty"], s, rseed)
                    dst.copy_(full["beauty"][crop])
                else:
                    sh.render_pixels(wide, s, rseed, pix)
                    dst.copy_(wide[crop].float())
                    data["albedo"][i, j] = data["albedo"][i, aov_level]
                    data["normal"][i, j] = data["normal"][i, aov_level]
        sh.render_pixels(wide, spp_ref, 16 * seed + 15, pix)
        data["target"][i] = wide[crop].float()
        if (i + 1) % 25 == 0 or i + 1 == n_scenes:
            log(f"dataset: {i + 1}/{n_scenes} crops rendered ({time.perf_counter() - t0:.1f} s)")
    return data


def augment(batch, k: int):
    """The k-th of the 8 flips/transposes of a [..][h][w][3] square batch."""
    if k & 1:
        batch = batch.flip(-2)
    if k & 2:
        batch = batch.flip(-3)
    if k & 4:
        batch = batch.transpose(-3, -2)
    return batch.contiguous()


def train(dn: Denoiser, data, steps: int, batch: int = 16, lr: float = 1e-4, seed: int = 0, log=print,
          log_every: int = 100):
    import torch
    n = data["beauty"].shape[0]
    g = torch.Generator(device="cpu")
    g.manual_seed(seed)
    losses = []
    t0 = time.perf_counter()
    # crop datasets hold [n][levels] inputs (albedo [n][levels][h][w][3]; beauty with a halves axis
    # after the level for a halves network); whole-render datasets [n][h][w][3]
    levels = data["albedo"].shape[1] if data["albedo"].dim() == 5 else 0
    dev = data["albedo"].device
    for step in range(steps):
        idx = torch.randint(0, n, (batch,), generator=g).to(dev)
        k = int(torch.randint(0, 8, (1,), generator=g))
        if levels:  # each element at a random input spp of its crop
            lv = torch.randint(0, levels, (batch,), generator=g).to(dev)
            b, a, nr = (augment(data[key][idx, lv], k) for key in ("beauty", "albedo", "normal"))
            if b.dim() == 5:  # halves: [batch][2][h][w][3] -> [2][batch][h][w][3]
                b = b.transpose(0, 1).contiguous()
            t = augment(data["target"].index_select(0, idx), k)
        else:
            b, a, nr, t = (augment(data[key].index_select(0, idx), k)
                           for key in ("beauty", "albedo", "normal", "target"))
        losses.append(dn.train_step(b, a, nr, t, lr=lr))
        if (step + 1) % log_every == 0 or step + 1 == steps:
            recent = float(np.mean(losses[-log_every:]))
            log(f"train: step {step + 1}/{steps} loss {recent:.5f} ({time.perf_counter() - t0:.1f} s)")
    return losses


def main(argv=None) -> int:
    import torch  # noqa: F401  (before the library loads: one HIP runtime per process)
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--scenes", type=int, default=400)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--lr", type=float, default=1e-4)
    ap.add_argument("--spp-ref", type=int, default=1024)
    ap.add_argument("--init", default="", help="start from these parameters (.npy)")
    ap.add_argument("--reject-ggx-fallback", action="store_true",
                    help="render with GT_FLAG_GGX_REJECT_BELOW_HORIZON (the well-posed estimator)")
    ap.add_argument("--crops", action="store_true",
                    help="train on random 256x256 crops of 1920x1080 frames at input spp "
                         f"{CROP_SPP_LEVELS} (make_crop_dataset) instead of whole 256x256 renders at 4 spp")
    ap.add_argument("--loss", choices=["rel-l2", "l1"], default="rel-l2")
    ap.add_argument("--spp-levels", default=",".join(map(str, CROP_SPP_LEVELS)),
                    help="input spp of the crop dataset, comma separated")
    ap.add_argument("--first-seed", type=int, default=20_000)
    ap.add_argument("--halves", action="store_true",
                    help="a network that takes each input as two half-sample renders (beauty_halves)")
    ap.add_argument("--out", required=True)
    args = ap.parse_args(argv)

    def log(m):
        print(m, flush=True)

    ctx = api.default_context()
    loss = abi.GT_DENOISE_LOSS_REL_L2 if args.loss == "rel-l2" else abi.GT_DENOISE_LOSS_L1_TONEMAP
    dn = Denoiser(ctx, loss=loss, halves=args.halves)
    if args.init:
        dn.load(args.init)
    log(f"denoiser: {dn.param_count} parameters, loss {args.loss}")
    flags = abi.GT_FLAG_GGX_REJECT_BELOW_HORIZON if args.reject_ggx_fallback else 0
    if args.crops:
        levels = tuple(int(v) for v in args.spp_levels.split(","))
        data = make_crop_dataset(ctx, args.scenes, first_seed=args.first_seed, spp_levels=levels,
                                 spp_ref=args.spp_ref, flags=flags, halves=args.halves, log=log)
    else:
        data = make_dataset(ctx, args.scenes, spp_ref=args.spp_ref, flags=flags, log=log)
    train(dn, data, args.steps, args.batch, args.lr, log=log)
    dn.save(args.out, training=" ".join(["python -m gartus_amd.denoise_train"] + list(argv or sys.argv[1:])))
    log(f"saved {args.out}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
