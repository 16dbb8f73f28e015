"""Batched fitting with the differentiable form: many independent scenes fitted at
once, every forward and backward of the batch being one launch per kernel
(rvk_diff_forward_batch / rvk_diff_backward_batch through diff.Renderer.batched).

Workload (`fit_workload`): `count` targets, each rendered by the f32 form from its
own seeded scene (`workloads.diff_scene`: `n_paths` closed paths of `segments`
cubics, a 4-stop linear gradient each, then blur σ and a colour matrix).  Each fit
starts from re-seeded random parameters: the paint, σ, matrix and geometry of
another seeded scene; the gradient end points are the scene's (they are scene data,
not parameters).  Fitting (`fit`) runs Adam on a normalised parametrisation (control
points and translations in canvas units, the rest as is), so that one learning
rate moves every kind of parameter by comparable amounts; the loss of a scene is the
mean squared error of its premultiplied RGBA image, and the scenes' losses are
summed (their gradients stay independent).  The result is reported as the PSNR of
the f32 render of the fitted parameters and of the plain u8 render (rvk_render_batch
of the same documents, diff.u8_scene) against the targets.
"""
from __future__ import annotations

import sys
import time

import numpy as np

from . import abi, diff
from .workloads import diff_scene


def paths_from_geometry(geom, k: int):
    """The closed k-cubic paths (verbs, device points) of one scene's geometry
    (paths x GEOM_FLOATS(k)): the affine applied to the control points (an affine
    map of a cubic's control points is the cubic's image)."""
    g = np.asarray(geom, np.float64).reshape(-1, diff.geom_floats(k))
    out = []
    for row in g:
        sx, ky, kx, sy, tx, ty = row[:6]
        pts = row[6:].reshape(-1, 2)
        dev = np.stack([sx * pts[:, 0] + kx * pts[:, 1] + tx, ky * pts[:, 0] + sy * pts[:, 1] + ty], 1)
        out.append(([abi.VERB_MOVE] + [abi.VERB_CUBIC] * k + [abi.VERB_CLOSE], dev.astype(np.float32)))
    return out


def fit_workload(count: int = 64, size: int = 512, n_paths: int = 128, segments: int = 8, seed: int = 0):
    """Targets' scenes and parameters, and the fits' starting parameters."""
    rng = np.random.default_rng(seed)
    scenes, tp, tg, ip, ig = [], [], [], [], []
    for b in range(count):
        sig_t, sig_i = rng.uniform(0.5, 4.0, 2)
        _, _, paths, ends, p_t, _ = diff_scene(seed=seed * 100003 + 2 * b, size=size, n_paths=n_paths,
                                               segments=segments, sigma=float(sig_t))
        _, _, paths_i, _, p_i, _ = diff_scene(seed=seed * 100003 + 2 * b + 1, size=size, n_paths=n_paths,
                                              segments=segments, sigma=float(sig_i))
        scenes.append((paths, ends))
        tp.append(p_t)
        tg.append(diff.geometry(paths, segments))
        ip.append(p_i)
        ig.append(diff.geometry(paths_i, segments))
    return {"size": size, "paths": n_paths, "segments": segments, "scenes": scenes,
            "target_params": np.stack(tp), "target_geom": np.stack(tg),
            "init_params": np.stack(ip), "init_geom": np.stack(ig)}


def _psnr(mse):
    return 10.0 * np.log10(1.0 / np.maximum(mse, 1e-12))


def fit(ctx, wl, steps: int = 500, lr: float = 1e-2, log_every: int = 0):
    """Fits every scene of the workload; returns the timing and quality figures."""
    import torch
    k, size = wl["segments"], wl["size"]
    count = len(wl["scenes"])
    r = diff.Renderer.batched(ctx, size, size, wl["scenes"], cubics=k)
    dev = torch.device("cuda", ctx.device)
    render = diff.render_fn()
    scale = torch.ones(diff.geom_floats(k), device=dev)
    scale[4:] = float(size)  # translations and points in canvas units
    with torch.no_grad():
        target = r.forward(torch.as_tensor(wl["target_params"], device=dev).contiguous(),
                           torch.as_tensor(wl["target_geom"], device=dev).contiguous()).clone()
    p = torch.as_tensor(wl["init_params"], device=dev).clone().requires_grad_(True)
    q = (torch.as_tensor(wl["init_geom"], device=dev) / scale).contiguous().requires_grad_(True)
    opt = torch.optim.Adam([p, q], lr=lr)

    def step():
        opt.zero_grad(set_to_none=True)
        out = render(p, (q * scale).contiguous(), r)
        per_scene = ((out - target) ** 2).mean(dim=(1, 2, 3))
        per_scene.sum().backward()
        opt.step()
        return per_scene

    with torch.no_grad():
        mse0 = ((r.forward(p.detach().contiguous(), (q * scale).contiguous()) - target) ** 2).mean(dim=(1, 2, 3))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(steps):
        per_scene = step()
        if log_every and (i + 1) % log_every == 0:
            print(f"[fit] step {i + 1}: mean mse {per_scene.mean().item():.5f}", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0

    with torch.no_grad():
        gfit = (q * scale).contiguous()
        out = r.forward(p.detach().contiguous(), gfit)
        mse = ((out - target) ** 2).mean(dim=(1, 2, 3)).cpu().numpy().astype(np.float64)
    # the plain u8 render of the fitted documents
    pf, gf = p.detach().cpu().numpy(), gfit.cpu().numpy()
    docs = [diff.u8_scene(size, size, paths_from_geometry(gf[b], k), wl["scenes"][b][1], pf[b])
            for b in range(count)]
    px = torch.zeros((count, size, size, 4), dtype=torch.uint8, device=dev)
    ctx.render_batch(docs, abi.Transform.identity(), [px[b] for b in range(count)])
    mse_u8 = ((px.float() / 255.0 - target) ** 2).mean(dim=(1, 2, 3)).cpu().numpy().astype(np.float64)
    m0 = mse0.cpu().numpy().astype(np.float64)
    del r, target, out, px
    torch.cuda.empty_cache()
    return {"scenes": count, "size": size, "paths": wl["paths"], "cubics_per_path": k, "steps": steps, "lr": lr,
            "steps_per_s": round(steps / wall, 2), "ms_per_step": round(wall / steps * 1e3, 2),
            "scene_steps_per_s": round(count * steps / wall, 1),
            "psnr_start_db": round(float(np.mean(_psnr(m0))), 2),
            "psnr_f32_db": round(float(np.mean(_psnr(mse))), 2),
            "psnr_u8_db": round(float(np.mean(_psnr(mse_u8))), 2),
            "psnr_u8_min_db": round(float(np.min(_psnr(mse_u8))), 2),
            "mse_ratio_end_over_start": round(float(np.mean(mse / np.maximum(m0, 1e-12))), 4)}
