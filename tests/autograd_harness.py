"""torch.autograd wrappers of the library's forward/VJP pairs (test harness
only) and their checks; run as a child process by tests/test_gpu_autograd.py.

- Project: the cone-beam projector of one view (stt_forward_project) with
  its matched transpose (stt_fp_vjp) as backward.
- Warp: the trilinear semi-Lagrangian warp (stt_warp) with its VJP with
  respect to the volume and the flow (stt_warp_vjp).
- BrightnessConstancy: beta * ||W(u, x_next) - x||^2 of one frame pair.  It
  saves only its inputs and recomputes the warped volume in backward
  (per-frame checkpointing), where the plain composition of Warp and torch
  operations keeps the residual volume of every frame.

The library is imported before torch, so it binds the system HIP runtime;
torch is used for CPU tensors and autograd only and never touches the GPU.
Prints one JSON line of results.
"""
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import spacetimetomography_amd as stt  # noqa: E402  (before torch, see above)
import torch  # noqa: E402


def _np(t):
    return np.ascontiguousarray(t.detach().cpu().numpy(), np.float32).reshape(-1)


class Project(torch.autograd.Function):
    @staticmethod
    def forward(ctx, vol, sctx, frame, view):
        ctx.sctx, ctx.frame, ctx.view = sctx, frame, view
        sctx.set_volume(frame, _np(vol))
        return torch.from_numpy(sctx.forward_project(frame, view)[0].copy())

    @staticmethod
    def backward(ctx, g):
        gv = ctx.sctx.fp_vjp(ctx.frame, ctx.view, _np(g))
        return torch.from_numpy(gv), None, None, None


class Warp(torch.autograd.Function):
    @staticmethod
    def forward(ctx, uvw, src, sctx, dt):
        ctx.sctx, ctx.dt = sctx, dt
        ctx.save_for_backward(uvw, src)
        return torch.from_numpy(sctx.warp(_np(uvw), _np(src), dt, stt.STT_INTERP_TRILINEAR))

    @staticmethod
    def backward(ctx, g):
        uvw, src = ctx.saved_tensors
        gsrc, guvw = ctx.sctx.warp_vjp(_np(uvw), _np(src), ctx.dt, _np(g))
        return torch.from_numpy(guvw), torch.from_numpy(gsrc), None, None


class BrightnessConstancy(torch.autograd.Function):
    """beta * ||W(u, x_next, dt) - x||^2, the warped volume recomputed in backward."""

    @staticmethod
    def forward(ctx, x, x_next, uvw, sctx, dt, beta):
        ctx.sctx, ctx.dt, ctx.beta = sctx, dt, beta
        ctx.save_for_backward(x, x_next, uvw)  # inputs only
        w = sctx.warp(_np(uvw), _np(x_next), dt, stt.STT_INTERP_TRILINEAR)
        r = w.astype(np.float64) - _np(x).astype(np.float64)
        return torch.tensor(beta * float(np.dot(r, r)), dtype=torch.float64)

    @staticmethod
    def backward(ctx, g):
        x, x_next, uvw = ctx.saved_tensors
        w = ctx.sctx.warp(_np(uvw), _np(x_next), ctx.dt, stt.STT_INTERP_TRILINEAR)  # recomputed
        gr = (2.0 * ctx.beta * float(g) * (w - _np(x))).astype(np.float32)
        gsrc, guvw = ctx.sctx.warp_vjp(_np(uvw), _np(x_next), ctx.dt, gr)
        return torch.from_numpy(-gr), torch.from_numpy(gsrc), torch.from_numpy(guvw), None, None, None


def smooth_flow(dims, amp, seed):
    rng = np.random.default_rng(seed)
    nx, ny, nz = dims
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    out = np.zeros((3, nz, ny, nx))
    for c in range(3):
        for _ in range(3):
            k = rng.integers(0, 3, 3)
            ph = rng.uniform(0, 2 * np.pi)
            out[c] += rng.uniform(-1, 1) * np.sin(2 * np.pi * (k[0] * x / nx + k[1] * y / ny + k[2] * z / nz) + ph)
    return (out * amp / max(np.abs(out).max(), 1e-9)).astype(np.float32).reshape(-1)


def main():
    dims, det, F, P = (20, 18, 16), (32, 28), 3, 3
    g = stt.Geometry.make(dims, det, (1, 1, 1), 1.0, 600.0, 1000.0)
    p = stt.params_default()
    p.interp = stt.STT_INTERP_TRILINEAR
    sctx = stt.Context(g, p, [P] * F)
    V, D = int(np.prod(dims)), det[0] * det[1]
    rng = np.random.default_rng(11)
    truth = [stt.synth_phantom_frame(dims, f, n_ellipsoids=4, seed=5, deform_seed=6, amp=1.0) for f in range(F)]
    b = []
    for f in range(F):
        sctx.set_views(f, [30.0 * f + 55.0 * k for k in range(P)])
        sctx.set_volume(f, truth[f])
        b.append([torch.from_numpy(sctx.forward_project(f, k)[0].copy()) for k in range(P)])
    beta, dt = 0.7, -1.0
    flows = [torch.from_numpy(smooth_flow(dims, 1.2, 20 + f)) for f in range(F - 1)]
    xs0 = [torch.from_numpy((0.8 * truth[f] + 0.05 * rng.standard_normal(V)).astype(np.float32)) for f in range(F)]

    def loss(xs, us, checkpointed):
        total = torch.zeros((), dtype=torch.float64)
        for f in range(F):
            for k in range(P):
                r = Project.apply(xs[f], sctx, f, k) - b[f][k]
                total = total + 0.5 * (r.double() ** 2).sum()
        for f in range(F - 1):
            if checkpointed:
                total = total + BrightnessConstancy.apply(xs[f], xs[f + 1], us[f], sctx, dt, beta)
            else:
                r = Warp.apply(us[f], xs[f + 1], sctx, dt).double() - xs[f].double()
                total = total + beta * (r ** 2).sum()
        return total

    out = {}
    grads = {}
    saved = {}
    for mode in (True, False):
        xs = [x.clone().requires_grad_(True) for x in xs0]
        us = [u.clone().requires_grad_(True) for u in flows]
        nbytes = [0]

        def pack(t):
            # bytes of intermediate tensors kept for backward (leaf inputs cost nothing extra)
            if not (t.is_leaf and t.requires_grad):
                nbytes[0] += t.numel() * t.element_size()
            return t

        with torch.autograd.graph.saved_tensors_hooks(pack, lambda t: t):
            L = loss(xs, us, mode)
        L.backward()
        grads[mode] = [x.grad.clone() for x in xs] + [u.grad.clone() for u in us]
        saved[mode] = nbytes[0]
        out["loss_checkpointed" if mode else "loss_plain"] = float(L)
    # the two graphs compute the same gradient (summation orders differ)
    rel = max(float((a.double() - c.double()).norm() / max(c.double().norm(), 1e-30))
              for a, c in zip(grads[True], grads[False]))
    out["grad_rel_diff"] = rel
    out["saved_intermediate_bytes_checkpointed"] = saved[True]
    out["saved_intermediate_bytes_plain"] = saved[False]
    out["residual_volume_bytes"] = 8 * V  # one float64 residual volume per frame pair
    out["pairs"] = F - 1

    # directional derivative against FP64 central differences of the loss
    dirs = [torch.from_numpy(rng.standard_normal(V).astype(np.float32)) for _ in range(F)]
    dirs_u = [torch.from_numpy((0.1 * rng.standard_normal(3 * V)).astype(np.float32)) for _ in range(F - 1)]
    ga = sum(float((gx.double() * d.double()).sum()) for gx, d in zip(grads[True], dirs + dirs_u))
    h = 1e-2
    with torch.no_grad():
        lp = float(loss([x + h * d for x, d in zip(xs0, dirs)], [u + h * d for u, d in zip(flows, dirs_u)], True))
        lm = float(loss([x - h * d for x, d in zip(xs0, dirs)], [u - h * d for u, d in zip(flows, dirs_u)], True))
    fd = (lp - lm) / (2 * h)
    out["directional_autograd"] = ga
    out["directional_fd"] = fd
    out["directional_rel_err"] = abs(ga - fd) / max(abs(fd), 1e-30)
    sctx.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
