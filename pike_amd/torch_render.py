"""Differentiable rendering harness: a torch.autograd.Function over the C ABI.

The forward pass writes the parameter tensors (the RGBA32F atlas and the
environment radiance, both in the PK_PARAM_* layouts) into the scene's
evaluation copies, renders the frame on the context stream and copies the
running-mean image into a torch tensor; the backward pass runs the
path-replay adjoint (pk_adjoint) with the incoming image gradient and returns
the parameter gradients.  Sampling is detached: the sampling atlas and the
environment's importance tables keep the values given at upload (or the
latest parameters written with ``sync_sampling``).

torch has to load the HIP runtime before libpike_amd.so does, so that both
use one runtime and device pointers are shared: this module imports torch
before pike_amd, and a process that uses both must import torch (or this
module) first.

``python -m pike_amd.torch_render --selftest`` runs the harness checks on the
GPU and prints one JSON line.
"""
import torch  # noqa: I001  (must precede pike_amd: one HIP runtime for both)

import argparse
import json
import sys
import time

import numpy as np

import pike_amd
from pike_amd import _ffi, scenes

ATLAS, ENV = _ffi.PK_PARAM_ATLAS, _ffi.PK_PARAM_ENV


class Renderer:
    """A scene, a frame with its camera and the render settings."""

    def __init__(self, ctx, scene, frame, max_depth=4, spp=16, env_theta=0.0):
        self.ctx, self.scene, self.frame = ctx, scene, frame
        self.max_depth, self.spp, self.env_theta = max_depth, spp, env_theta
        # Seeds: the forward renders samples [first_sample, first_sample + spp);
        # the backward replays [first_sample + adjoint_offset, ... + spp).  An
        # offset of spp makes the adjoint's samples independent of the ones
        # the loss gradient was computed from (reusing them biases the
        # gradient of an L2 loss toward lower variance).
        self.first_sample = 0
        self.adjoint_offset = 0
        i = scene.param_info()
        self.atlas_shape = (i["tex_layers"], i["tex_height"], i["tex_width"], 4)
        self.env_shape = (i["env_height"], i["env_width"], 4)

    def write(self, atlas, env, sync_sampling=False):
        torch.cuda.current_stream().synchronize()
        self.scene.write_params(ATLAS, device_ptr=atlas.data_ptr(), count=atlas.numel(), also_sampling=sync_sampling)
        if env.numel():
            self.scene.write_params(ENV, device_ptr=env.data_ptr(), count=env.numel())


class RenderFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, atlas, env, renderer):
        r = renderer
        a = atlas.detach().contiguous().float()
        e = env.detach().contiguous().float()
        assert tuple(a.shape) == r.atlas_shape and tuple(e.shape) == r.env_shape
        r.write(a, e)
        f = r.frame
        f.reset()
        f.render(r.scene, max_depth=r.max_depth, first_sample=r.first_sample, num_frames=r.spp,
                 env_theta=r.env_theta)
        img = torch.empty((f.height, f.width, 4), device=a.device, dtype=torch.float32)
        f.read_image_device(img.data_ptr())
        r.ctx.sync()
        if r.first_sample:
            # the running mean counts from RenderState.samples (accumulate.wgsl),
            # so frames starting at sample k divide by k + spp
            img *= (r.first_sample + r.spp) / r.spp
        ctx.renderer = r
        ctx.first_sample = r.first_sample + r.adjoint_offset
        ctx.save_for_backward(a, e)
        return img

    @staticmethod
    def backward(ctx, grad_img):
        r = ctx.renderer
        a, e = ctx.saved_tensors
        g = grad_img.contiguous().float()
        ga = torch.empty(r.atlas_shape, device=g.device, dtype=torch.float32)
        ge = torch.empty(r.env_shape if e.numel() else (1, 4), device=g.device, dtype=torch.float32)
        r.write(a, e)  # the parameters this forward saw (another forward may have run since)
        r.frame.adjoint_device(r.scene, g.data_ptr(), ga.data_ptr(), ge.data_ptr(), max_depth=r.max_depth,
                               first_sample=ctx.first_sample, num_frames=r.spp, env_theta=r.env_theta)
        r.ctx.sync()
        return ga, (ge if e.numel() else None), None


def render(atlas, env, renderer):
    """Differentiable render: (atlas, env) tensors -> (H, W, 4) image."""
    return RenderFunction.apply(atlas, env, renderer)


def _selftest(args):
    if not torch.cuda.is_available():
        return {"skipped": "torch sees no GPU"}
    dev = torch.device("cuda:0")
    data = scenes.config5(cells=args.cells, num_spheres=8, tex_res=args.tex, env_size=(args.tex, args.tex // 2))
    bvh = pike_amd.BVH(data.positions, data.triangles)
    cam = scenes.normalized_camera(data.camera, bvh.normalizer())
    ctx = pike_amd.Context(0)
    W = H = args.res
    # ground truth and target image
    gt_atlas = torch.from_numpy(data.tex.copy()).to(dev)
    env0 = np.zeros((args.tex // 2, args.tex, 4), np.float32)
    env0[..., :3] = data.env.reshape(args.tex // 2, args.tex, 3)
    gt_env = torch.from_numpy(env0).to(dev)
    scene = pike_amd.Scene(ctx, data, bvh)
    frame = pike_amd.Frame(ctx, W, H)
    frame.set_camera(cam)
    r_target = Renderer(ctx, scene, frame, max_depth=args.depth, spp=args.target_spp)
    with torch.no_grad():
        target = render(gt_atlas, gt_env, r_target).clone()
    # initial parameters: ground truth + N(0, 0.1), clamped; sampling follows them
    g = torch.Generator(device="cpu").manual_seed(5)
    atlas = gt_atlas.clone()
    atlas[0, ..., :3] = (atlas[0, ..., :3] + 0.1 * torch.randn(atlas[0, ..., :3].shape, generator=g).to(dev)).clamp(
        0.02, 1.0)
    atlas[1, ..., 1] = (atlas[1, ..., 1] + 0.1 * torch.randn(atlas[1, ..., 1].shape, generator=g).to(dev)).clamp(
        0.05, 1.0)
    env = (gt_env * (1.0 + 0.1 * torch.randn(gt_env.shape, generator=g).to(dev))).clamp(min=0.0)
    r = Renderer(ctx, scene, frame, max_depth=args.depth, spp=args.spp)
    r.write(atlas, env, sync_sampling=True)
    out = {}
    # 1. autograd plumbing: torch's gradient equals a direct host adjoint call
    a = atlas.clone().requires_grad_(True)
    e = env.clone().requires_grad_(True)
    img = render(a, e, r)
    loss = ((img[..., :3] - target[..., :3]) ** 2).sum()
    loss.backward()
    d_image = np.zeros((H, W, 4), np.float32)
    d_image[..., :3] = (2 * (img[..., :3] - target[..., :3])).detach().cpu().numpy()
    ha, he = frame.adjoint(scene, d_image, max_depth=args.depth, num_frames=args.spp)
    ta, te = a.grad.cpu().numpy(), e.grad.cpu().numpy()
    scale = max(np.abs(ha).max(), np.abs(he).max())
    out["autograd_vs_adjoint_max_abs"] = float(max(np.abs(ta - ha).max(), np.abs(te - he).max()) / scale)
    # 2. central differences through the autograd forward (eps = 1e-3) on the
    #    8 base-color parameters with the largest gradients
    order = np.argsort(np.abs(ta[0, ..., :3]).ravel())[::-1][:8]
    fd_err = []
    with torch.no_grad():
        for k in order:
            t, c = divmod(int(k), 3)
            y, x = divmod(t, args.tex)
            p0 = float(atlas[0, y, x, c])
            vals = []
            for sgn in (1, -1):
                p = atlas.clone()
                p[0, y, x, c] = p0 + sgn * 1e-3
                im = render(p, env, r).double()
                vals.append(float(((im[..., :3] - target[..., :3].double()) ** 2).sum()))
            fd = (vals[0] - vals[1]) / (float(np.float32(p0 + 1e-3)) - float(np.float32(p0 - 1e-3)))
            fd_err.append(abs(fd - float(ta[0, y, x, c])) / abs(fd))
    out["fd_rel_err_max"] = float(max(fd_err))
    # 3. inverse rendering of the base-color texture: Adam from perturbed
    #    values, roughness and environment held at their ground truth
    r.write(gt_atlas, gt_env, sync_sampling=True)
    start = gt_atlas.clone()
    start[0, ..., :3] = atlas[0, ..., :3]
    # well-observed texels: the top quarter of the first gradient's magnitudes
    mag = np.abs(ta[0, ..., :3])
    seen = torch.from_numpy(mag >= np.quantile(mag[mag > 0], 0.75)).to(dev)
    a = start.clone().requires_grad_(True)
    opt = torch.optim.Adam([a], lr=args.lr)
    losses = []
    r.adjoint_offset = r.spp
    t0 = time.perf_counter()
    for it in range(args.iters):
        r.first_sample = 2 * r.spp * it  # fresh, independent samples every iteration
        opt.zero_grad()
        img = render(a, gt_env, r)
        loss = ((img[..., :3] - target[..., :3]) ** 2).sum()
        loss.backward()
        with torch.no_grad():
            mask = torch.zeros_like(a.grad)
            mask[0, ..., :3] = 1
            a.grad *= mask
        opt.step()
        with torch.no_grad():
            a[0, ..., :3].clamp_(0.02, 1.0)
        losses.append(float(loss))
    torch.cuda.synchronize()
    out["iters_per_s"] = args.iters / (time.perf_counter() - t0)

    def eval_loss(p):
        # an independent higher-spp render, so the loss reflects the parameters
        # more than the Monte Carlo noise
        r_eval = Renderer(ctx, scene, frame, max_depth=args.depth, spp=args.eval_spp)
        r_eval.first_sample = 1 << 16
        with torch.no_grad():
            im = render(p, gt_env, r_eval)
            return float(((im[..., :3] - target[..., :3]) ** 2).sum())

    out["loss_first"], out["loss_last"] = eval_loss(start), eval_loss(a.detach())
    out["train_loss_first"], out["train_loss_last"] = losses[0], losses[-1]
    err0 = (start[0, ..., :3] - gt_atlas[0, ..., :3]).abs()[seen]
    err1 = (a.detach()[0, ..., :3] - gt_atlas[0, ..., :3]).abs()[seen]
    out["param_err_first"], out["param_err_last"] = float(err0.mean()), float(err1.mean())
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--selftest", action="store_true")
    ap.add_argument("--res", type=int, default=64)
    ap.add_argument("--tex", type=int, default=64)
    ap.add_argument("--cells", type=int, default=40)
    ap.add_argument("--depth", type=int, default=3)
    ap.add_argument("--spp", type=int, default=64)
    ap.add_argument("--target-spp", type=int, default=1024)
    ap.add_argument("--eval-spp", type=int, default=256)
    ap.add_argument("--iters", type=int, default=40)
    ap.add_argument("--lr", type=float, default=0.01)
    args = ap.parse_args(argv)
    if not args.selftest:
        ap.print_help()
        return 0
    print(json.dumps(_selftest(args)), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
