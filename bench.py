#!/usr/bin/env python3
"""Throughput of resvg's pixel pipeline (rvk_render, the resvg::render twin) on one
MI355X, on the 4096x4096 full-filter scene.

One step renders the whole seeded document onto a cleared 4096x4096 RGBA8
premultiplied device pixmap: 2000 paths (3-200 cubic/quad segments, evenodd and
nonzero, 40% also stroked) with 2-8 stop linear/radial gradients, each in its own
opacity group, and 8 filter regions covering blur (IIR and box), morphology,
convolution, colour matrix, component transfer, turbulence, lighting and blend.
The scene tree is built before the timed region (usvg's job); host-side edge
building, uploads and every kernel are inside it.

Prints one JSON line: Mpixel/s of the document, the roofline of the dominant kernel
(algorithmic bytes / its HIP-event time over the timed steps, against 8 TB/s HBM),
the oracle's CPU throughput on a bounded sample of the same document (with the
per-channel mismatch histogram of two seeded documents against the oracle), and under
"batched" the documents/s of rvk_render_batch against a loop of rvk_render calls on a
sample of the batched workload, under "differentiable" the f32 forward and
forward+backward times of the differentiable form against the plain u8 forward,
under "batched_fitting" the steps/s and final PSNR of 64 independent 512x512 fits
run as one batch, under "png_output" the PNG encoding of the rendered document
(the device's demultiply + row-filter kernel against HBM peak, the host's deflate),
and under "clip_mask" a 4096x4096 document of 300 clipped and masked groups (clip
paths and masks at scale; its 1024x1024 quarter is timed on the oracle beside it).
"""
from __future__ import annotations

import argparse
import json
import multiprocessing as mp
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E


def _workload(size, paths, seed):
    from resvg_amd.workloads import full_filter_doc
    return full_filter_doc(seed=seed, size=size, n_paths=paths)


# ---------------------------------------------------------------------------
# CPU baseline: the oracle (a scalar C++ restatement of the reference path), built
# -O3 -march=native on this machine, on a stratified sample of the same document's
# top-level items.  resvg sizes a group's layer by the group's bounds (up to 5x the
# canvas), not by the visible canvas, so a crop of the canvas does not shrink the
# work; a subset of the items does.
# ---------------------------------------------------------------------------
_FE_NAMES = ["blend", "color_matrix", "component_transfer", "composite", "convolve_matrix", "diffuse_lighting",
             "displacement_map", "drop_shadow", "flood", "gaussian_blur", "image", "merge", "morphology", "offset",
             "specular_lighting", "tile", "turbulence"]


class _Keep:
    """Keeps filter group number `fg` (-1: none) and, of the plain items, every
    `stride`-th one, dealt round-robin over `parts` jobs (this is job `part`)."""

    def __init__(self, fg, stride, parts, part):
        self.fg, self.stride, self.parts, self.part = fg, stride, parts, part
        self.nf = self.np = 0

    def __call__(self, item, is_filter):
        if is_filter:
            self.nf += 1
            return self.nf - 1 == self.fg
        self.np += 1
        k = self.np - 1
        return k % self.stride == 0 and (k // self.stride) % self.parts == self.part


def _cpu_init(lib_path):
    os.environ["RVK_ORACLE_LIB"] = lib_path  # before oracle.bindings is imported


def _cpu_job(job):
    size, paths, seed, fg, stride, parts, part = job
    import numpy as np
    from oracle import bindings as orc
    from resvg_amd import abi
    from resvg_amd.workloads import full_filter_doc
    sc, _ = full_filter_doc(seed=seed, size=size, n_paths=paths, keep=_Keep(fg, stride, parts, part))
    s = sc.to_abi()
    px = np.zeros((size, size, 4), np.uint8)
    orc.prof_enable(True)
    t = time.perf_counter()
    orc.render(s.abi, abi.Transform.identity(), px)
    dt = time.perf_counter() - t
    prof = orc.prof_read()
    orc.prof_enable(False)
    return dt, len(sc.paths), prof


def _cpu_clip_job(job):
    size, groups, seed = job
    import numpy as np
    from oracle import bindings as orc
    from resvg_amd import abi
    from resvg_amd.workloads import clip_mask_doc
    sc, _ = clip_mask_doc(seed=seed, size=size, n_groups=groups)
    s = sc.to_abi()
    px = np.zeros((size, size, 4), np.uint8)
    t = time.perf_counter()
    orc.render(s.abi, abi.Transform.identity(), px)
    return time.perf_counter() - t


def _cpu_model():
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(size, paths, seed, workers, plain_stride=128, n_filter_groups=8, runs=5, clip_groups=0):
    """The oracle built -O3 -march=native for this host, one document part per core:
    each of the 8 filter groups alone, and every `plain_stride`-th plain path spread
    over the remaining cores.  One warm-up run, then `runs` timed runs; per part the
    median time.  The whole-document time per core is estimated as the sum of the 8
    filter-group medians + plain_stride x the sum of the plain parts' medians."""
    import statistics
    import tempfile
    from oracle import bindings as orc
    lib_dir = tempfile.mkdtemp(prefix="rvk_oracle_")
    t0 = time.perf_counter()
    lib_path = str(orc.build_native(lib_dir))
    t_build = time.perf_counter() - t0
    parts = max(1, workers - n_filter_groups)
    jobs = [(size, paths, seed, fg, 1 << 30, 1, 1) for fg in range(n_filter_groups)]  # the filter groups alone
    jobs += [(size, paths, seed, -1, plain_stride, parts, p) for p in range(parts)]
    ctx = mp.get_context("spawn")
    times = [[] for _ in jobs]
    profs = None
    t0 = time.perf_counter()
    with ctx.Pool(min(workers, len(jobs)), initializer=_cpu_init, initargs=(lib_path,)) as pool:
        for r in range(runs + 1):
            res = pool.map(_cpu_job, jobs)
            print(f"[bench] cpu baseline run {r + 1}/{runs + 1} ({'warm-up' if r == 0 else 'timed'}): "
                  f"{max(x[0] for x in res):.1f} s longest part", file=sys.stderr, flush=True)
            if r == 0:
                continue
            for i, x in enumerate(res):
                times[i].append(x[0])
            if profs is None:
                profs = [x[2] for x in res]
        # the clip/mask document at 1024x1024 (a quarter of the groups), one core,
        # after a warm-up render
        clip_s = None
        if clip_groups > 0:
            cjob = (1024, max(1, clip_groups // 4), seed)
            pool.apply(_cpu_clip_job, (cjob,))
            clip_s = pool.apply(_cpu_clip_job, (cjob,))
    wall = time.perf_counter() - t0
    med = [statistics.median(t) for t in times]
    n_paths = [len_ for _, len_, _ in res]
    t_filter = sum(med[:n_filter_groups])
    t_plain = sum(med[n_filter_groups:])
    t_doc = t_filter + plain_stride * t_plain
    # per-operation split of one document per core (ms), from the first timed run:
    # filter groups count once, the plain sample plain_stride times
    split = {}
    for i, pr in enumerate(profs):
        wgt = 1.0 if i < n_filter_groups else float(plain_stride)
        for k, name in enumerate(_FE_NAMES):
            if pr[k] > 0:
                split[name] = split.get(name, 0.0) + wgt * pr[k]
        for name, k in orc.PROF_SLOTS.items():
            split[name] = split.get(name, 0.0) + wgt * pr[k]
    split = {k: round(v / 1e3, 2) for k, v in sorted(split.items(), key=lambda kv: -kv[1])}
    return {
        "value": round(workers * size * size / 1e6 / t_doc, 4),
        "unit": "Mpixel/s",
        "cores": workers,
        "kind": "port",
        "cpu_model": _cpu_model(),
        "nproc": os.cpu_count(),
        "sample": f"oracle (C++ restatement of the reference path, not the reference binary; built -O3 "
                  f"-march=native -ffp-contract=off on this host in {t_build:.0f} s), one document part per core: "
                  f"the {size}x{size} {paths}-path document's {n_filter_groups} filter groups each alone "
                  f"({sum(n_paths[:n_filter_groups])} paths, {t_filter:.1f} s) and every {plain_stride}th plain "
                  f"path ({sum(n_paths[n_filter_groups:])} paths over {parts} parts, {t_plain:.1f} s); median of "
                  f"{runs} runs after 1 warm-up; estimated {t_doc:.0f} s per document per core; wall {wall:.1f} s",
        "seconds_per_document_per_core": round(t_doc, 2),
        "split_s_per_document": split,
        "clip_mask": None if clip_s is None else {
            "workload": f"clip/mask 1024x1024, {max(1, clip_groups // 4)} groups, seed {seed}",
            "seconds": round(clip_s, 2), "mpixel_per_s_per_core": round(1024 * 1024 / 1e6 / clip_s, 4)},
    }


def cpu_parity(ctx, seed):
    """Part of the CPU-baseline leg: the per-channel mismatch histogram (|GPU - oracle|
    = 0, 1, > 1) of two seeded documents rendered by rvk_render and by the oracle: the
    512x512 blur -> offset -> composite document, and the full-filter document at
    1024x1024 with 250 paths (the oracle renders it in about 10 s on one core)."""
    import numpy as np
    import torch
    from oracle import bindings as orc
    from resvg_amd import abi
    from resvg_amd.workloads import blur_offset_composite_doc, full_filter_doc
    out = []
    for name, (sc, (w, h)) in (("blur_offset_composite 512x512", blur_offset_composite_doc(seed=seed, size=512)),
                               ("full_filter 1024x1024, 250 paths", full_filter_doc(seed=seed, size=1024,
                                                                                    n_paths=250))):
        s = sc.to_abi()
        ref = np.zeros((h, w, 4), np.uint8)
        orc.render(s.abi, abi.Transform.identity(), ref)
        px = torch.zeros((h, w, 4), dtype=torch.uint8, device="cuda")
        ctx.render(s, abi.Transform.identity(), px)
        d = np.abs(px.cpu().numpy().astype(np.int16) - ref.astype(np.int16))
        out.append({"document": name, "channels": int(d.size), "equal": int((d == 0).sum()),
                    "off_by_1": int((d == 1).sum()), "off_by_more": int((d > 1).sum())})
        print(f"[bench] parity {name}: {out[-1]}", file=sys.stderr, flush=True)
    return out


def _build_doc(spec):
    from resvg_amd.workloads import batch_doc
    return batch_doc(*spec)


def batched(ctx, n_docs, seed, steps):
    """The batched form: `n_docs` documents of the batched distribution (sides
    log-uniform 64-2048, aspect 0.5-2, 5-200 paths, a 1-6 primitive filter chain each)
    rendered by one rvk_render_batch call, against the same documents rendered by a
    loop of rvk_render calls (each ends with its own wait).  Documents are built before
    the timed region; both forms composite onto the same resident device pixmaps."""
    import torch
    from resvg_amd import abi
    from resvg_amd.workloads import document_specs
    t0 = time.perf_counter()
    specs = document_specs(n_docs, seed=seed)
    workers = max(1, min(16, len(os.sched_getaffinity(0))))
    with mp.get_context("spawn").Pool(workers) as pool:
        built = pool.map(_build_doc, specs, chunksize=max(1, n_docs // (4 * workers)))
    docs = [(sc, (sp[1], sp[2])) for sc, sp in zip(built, specs)]
    scenes = [sc.to_abi() for sc, _ in docs]
    pxs = [torch.zeros((h, w, 4), dtype=torch.uint8, device="cuda") for _, (w, h) in docs]
    print(f"[bench] batched: {n_docs} documents built in {time.perf_counter() - t0:.1f} s", file=sys.stderr,
          flush=True)
    ts = abi.Transform.identity()

    def loop():
        for sc, px in zip(scenes, pxs):
            ctx.render(sc, ts, px)

    def batch():
        ctx.render_batch(scenes, ts, pxs)

    out = {}
    for name, fn in (("loop", loop), ("batch", batch)):
        fn()  # warm-up: scratch buffers and staging reach their working sizes
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        out[name] = (time.perf_counter() - t) / steps
        print(f"[bench] batched: {name} {out[name] * 1e3:.1f} ms per pass", file=sys.stderr, flush=True)
    mpix = sum(w * h for _, (w, h) in docs) / 1e6
    return {"docs": n_docs, "seed": seed, "mpixels": round(mpix, 2),
            "paths": sum(len(sc.paths) for sc, _ in docs),
            "docs_per_s": round(n_docs / out["batch"], 1), "loop_docs_per_s": round(n_docs / out["loop"], 1),
            "speedup_vs_loop": round(out["loop"] / out["batch"], 3),
            "mpixel_per_s": round(mpix / out["batch"], 2), "ms_per_pass": round(out["batch"] * 1e3, 2),
            "loop_ms_per_pass": round(out["loop"] * 1e3, 2), "passes": steps,
            "sample": f"{n_docs} documents of the batched distribution (seed {seed})"}


def differentiable(ctx, size, n_paths, seed, reps):
    """The differentiable f32 form at `size`² with `n_paths` paths: forward and
    forward+backward of rvk_diff_forward / rvk_diff_backward against the plain u8
    forward (rvk_render) of the same document; in the planes mode (coverage computed
    once, before the timed region) and in the analytic mode (coverage from the
    geometry on every forward, geometry gradients in the backward)."""
    import torch
    from resvg_amd import abi, diff
    from resvg_amd.workloads import diff_scene
    t0 = time.perf_counter()
    segments = 8
    w, h, paths, ends, p0, pt = diff_scene(seed=seed, size=size, n_paths=n_paths, segments=segments)
    r = diff.Renderer(ctx, w, h, paths, ends)
    params = torch.as_tensor(p0, device="cuda")
    target = r.forward(torch.as_tensor(pt, device="cuda")).clone()
    u8 = diff.u8_scene(w, h, paths, ends, p0).to_abi()
    px = torch.zeros((h, w, 4), dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    print(f"[bench] differentiable: scene and coverage ready in {time.perf_counter() - t0:.1f} s", file=sys.stderr,
          flush=True)

    def fwd():
        return r.forward(params)

    def fwd_bwd():
        out = r.forward(params)
        return r.backward(params, 2.0 * (out - target) / out.numel())

    def plain():
        px.zero_()
        ctx.render(u8, abi.Transform.identity(), px)

    res = {}
    for name, fn in (("fwd", fwd), ("fwd_bwd", fwd_bwd), ("plain", plain)):
        fn()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        res[name] = (time.perf_counter() - t) / reps * 1e3
    loss = float(((r.forward(params) - target) ** 2).mean())
    ctx.profile_begin()
    fwd_bwd()
    kern = [{"name": k["name"], "launches": k["launches"], "ms": round(k["total_ms"], 3)} for k in ctx.profile_end()]
    del r
    torch.cuda.empty_cache()

    # the analytic mode: the same paths as closed chains of cubics whose control points
    # and affine are parameters too; the coverage is rebuilt from the geometry on every
    # forward and the backward also returns dL/d(geometry)
    ra = diff.Renderer(ctx, w, h, paths, ends, cubics=segments)
    geom = ra.geom0.clone()
    target_a = ra.forward(torch.as_tensor(pt, device="cuda"), geom).clone()

    def fwd_a():
        return ra.forward(params, geom)

    def fwd_bwd_a():
        out = ra.forward(params, geom)
        return ra.backward(params, 2.0 * (out - target_a) / out.numel(), geom)

    for name, fn in (("fwd_a", fwd_a), ("fwd_bwd_a", fwd_bwd_a)):
        fn()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        res[name] = (time.perf_counter() - t) / reps * 1e3
    ctx.profile_begin()
    fwd_bwd_a()
    kern_a = [{"name": k["name"], "launches": k["launches"], "ms": round(k["total_ms"], 3)} for k in ctx.profile_end()]
    analytic = {"cubics_per_path": segments, "geometry_params": int(geom.numel()),
                "f32_forward_ms": round(res["fwd_a"], 3), "f32_forward_backward_ms": round(res["fwd_bwd_a"], 3),
                "fwd_bwd_over_plain": round(res["fwd_bwd_a"] / res["plain"], 3),
                "gradient_flows_to": "paint, opacity, blur sigma, colour matrix, control points, affine",
                "kernels_fwd_bwd": kern_a[:10]}
    del ra
    torch.cuda.empty_cache()
    return {"size": size, "paths": n_paths, "params": int(params.numel()), "f32_forward_ms": round(res["fwd"], 3),
            "f32_forward_backward_ms": round(res["fwd_bwd"], 3), "plain_u8_forward_ms": round(res["plain"], 3),
            "fwd_bwd_over_plain": round(res["fwd_bwd"] / res["plain"], 3), "loss": loss, "reps": reps,
            "gradient_flows_to": "paint, opacity, blur sigma, colour matrix (coverage planes fixed)",
            "kernels_fwd_bwd": kern[:8], "analytic_geometry": analytic}


def fitting(ctx, count, size, n_paths, steps, seed):
    """Batched fitting: `count` independent size² f32 targets (each its own seeded
    `n_paths`-path scene of 8-cubic closed paths with blur and colour matrix), fitted
    from re-seeded random parameters (paint, σ, matrix, control points, affines) by
    Adam (lr 1e-2) for `steps` steps; each step's forward and backward for all scenes
    is one launch per kernel.  Reports steps/s and the PSNR against the targets of the
    f32 render and of the plain u8 render of the fitted parameters."""
    from resvg_amd import fit
    t0 = time.perf_counter()
    wl = fit.fit_workload(count=count, size=size, n_paths=n_paths, segments=8, seed=seed)
    print(f"[bench] fitting: {count} scenes built in {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    res = fit.fit(ctx, wl, steps=steps, lr=1e-2, log_every=max(1, steps // 5))
    print(f"[bench] fitting: {res['steps_per_s']} steps/s, PSNR u8 {res['psnr_u8_db']} dB", file=sys.stderr,
          flush=True)
    return res


def png_output(ctx, px, reps):
    """PNG output of the rendered document (tiny-skia Pixmap::encode_png): the device's
    demultiply + per-row filter kernel, timed with HIP events over `reps` launches
    (8 algorithmic bytes per pixel), then the host's deflate (zlib level 6, in parallel
    1 MiB pieces) of its rows and the whole Context.encode_png call."""
    import torch

    h, w = px.shape[:2]
    rows = torch.empty(h * (4 * w + 1), dtype=torch.uint8, device=px.device)
    ctx.png_filter_rows(px, rows)  # warm-up
    ctx.synchronize()
    ctx.profile_begin()
    for _ in range(reps):
        ctx.png_filter_rows(px, rows)
    ks = {k["name"]: k for k in ctx.profile_end()}
    k = ks["png_rows"]
    ms = k["total_ms"] / k["launches"]
    gbs = k["bytes"] / k["launches"] / (ms / 1e3) / 1e9
    host_rows = rows.cpu().numpy().tobytes()
    from resvg_amd import png
    t0 = time.perf_counter()
    deflated = png.deflate(host_rows, 6)
    deflate_ms = (time.perf_counter() - t0) * 1e3
    t0 = time.perf_counter()
    data = ctx.encode_png(px)
    encode_ms = (time.perf_counter() - t0) * 1e3
    return {"rows_kernel_ms": round(ms, 4), "rows_GBps": round(gbs, 1), "rows_frac": round(gbs / HBM_PEAK_GBS, 4),
            "deflate_ms": round(deflate_ms, 1), "encode_png_ms": round(encode_ms, 1),
            "file_MB": round(len(data) / 1e6, 2), "idat_MB": round(len(deflated) / 1e6, 2)}


def clip_mask(ctx, size, groups, seed, steps):
    """The clip/mask document (clip paths and masks at scale): `steps` renders timed
    with HIP events on the context stream after one warm-up render, with the per-kernel
    split; the same document at 1024x1024 with a quarter of the groups is what the
    CPU baseline times (`cpu_baseline.clip_mask`)."""
    import torch
    from resvg_amd import abi
    from resvg_amd.workloads import clip_mask_doc
    sc, (w, h) = clip_mask_doc(seed=seed, size=size, n_groups=groups)
    scene = sc.to_abi()
    ts = abi.Transform.identity()
    px = torch.zeros((h, w, 4), dtype=torch.uint8, device="cuda:0")
    small = torch.zeros((1024, 1024, 4), dtype=torch.uint8, device="cuda:0")
    sc_small, _ = clip_mask_doc(seed=seed, size=1024, n_groups=max(1, groups // 4))
    scene_small = sc_small.to_abi()
    ctx.render(scene, ts, px)
    ctx.render(scene_small, ts, small)
    ctx.synchronize()
    stream = torch.cuda.current_stream()
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ctx.profile_begin()
    start.record(stream)
    for _ in range(steps):
        px.zero_()
        ctx.render(scene, ts, px)
    stop.record(stream)
    torch.cuda.synchronize()
    ks = ctx.profile_end()
    st = ctx.stats()
    ms = start.elapsed_time(stop) / steps
    start.record(stream)
    for _ in range(steps):
        small.zero_()
        ctx.render(scene_small, ts, small)
    stop.record(stream)
    torch.cuda.synchronize()
    ms_small = start.elapsed_time(stop) / steps
    per = lambda k: round(k["total_ms"] / steps, 3)
    masks = [k for k in ks if k["name"] in ("apply_mask", "mask_from_pixmap", "mask_invert", "fill_mask")]
    return {"workload": f"clip/mask {w}x{h}, {groups} groups, 16 clip paths, 16 masks, seed {seed}",
            "mpixel_per_s": round(w * h / 1e6 / (ms / 1e3), 2), "ms_per_render": round(ms, 2), "steps": steps,
            "layers": st.layers, "flushes": st.flushes, "kernel_launches": st.kernel_launches,
            "mask_kernels_ms_per_render": round(sum(k["total_ms"] for k in masks) / steps, 3),
            "small_ms_per_render": round(ms_small, 3),
            "kernels": [{"name": k["name"], "launches": k["launches"] // steps, "ms_per_render": per(k),
                         "GB/s": round(k["bytes"] / (k["total_ms"] / 1e3) / 1e9, 1) if k["total_ms"] > 0 else None}
                        for k in ks[:8]]}


def _counter_per_dispatch(csv_dir: str, counter: str) -> dict:
    """The values (KiB) of `counter`, one per dispatch, per kernel name, from the
    *_counter_collection.csv files rocprofv3 wrote under `csv_dir`."""
    import csv
    import glob
    vals: dict = {}
    for f in glob.glob(os.path.join(csv_dir, "**", "*counter_collection.csv"), recursive=True):
        for r in csv.DictReader(open(f)):
            if r.get("Counter_Name") == counter:
                vals.setdefault(r["Kernel_Name"], []).append(float(r["Counter_Value"]))
    return vals


def live_traffic(args):
    """HBM bytes per launch of the compositor kernel, measured in this run: two child
    processes render one step of the same document under rocprofv3 --pmc (FETCH_SIZE,
    then WRITE_SIZE; one counter per pass), after a 1 GiB device copy whose counter
    values give the scale from counter units to bytes.  None when rocprofv3 is not
    available or a pass fails."""
    import shutil
    import subprocess
    import tempfile
    if shutil.which("rocprofv3") is None:
        return None
    child = [sys.executable, str(Path(__file__).resolve()), "--steps", "1", "--warmup", "0", "--no-cpu-baseline",
             "--batch-docs", "0", "--diff-size", "0", "--fit-scenes", "0", "--png-reps", "0", "--clip-groups", "0",
             "--copy-calibration",
             "--traffic", "stored", "--size", str(args.size), "--paths", str(args.paths), "--seed", str(args.seed)]
    env = dict(os.environ, TMPDIR="/tmp")
    per = {}
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        d = tempfile.mkdtemp(prefix="rvk_pmc_", dir="/tmp")
        print(f"[bench] traffic: rocprofv3 --pmc {counter}", file=sys.stderr, flush=True)
        try:
            r = subprocess.run(["rocprofv3", "--pmc", counter, "-d", d, "-o", "pmc", "--output-format", "csv", "--",
                                *child], cwd="/tmp", env=env, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL,
                               timeout=240)
        except subprocess.TimeoutExpired:
            return None
        vals = _counter_per_dispatch(d, counter)
        shutil.rmtree(d, ignore_errors=True)
        # the 1 GiB copy is the largest copy dispatch; the profiler may end with a
        # crash after writing complete results, so the results decide
        cal = max(vals.get("__amd_rocclr_copyBuffer", [0.0]))
        draw = [x for k, v in vals.items() if "k_draw<false>" in k or "k_draw<true>" in k for x in v]
        if cal <= 0 or not draw:
            print(f"[bench] traffic: {counter} pass gave no values (exit {r.returncode})", file=sys.stderr, flush=True)
            return None
        per[counter] = sum(draw) / len(draw) * 1024.0 * ((1 << 20) / cal)  # KiB -> bytes, times GiB / copy value
    return {"bytes_per_launch": round(per["FETCH_SIZE"] + per["WRITE_SIZE"]),
            "fetch_bytes": round(per["FETCH_SIZE"]), "write_bytes": round(per["WRITE_SIZE"])}


def _primitive_rooflines(kernels, steps):
    """Per filter primitive / pipeline stage: its kernels' time per step and the
    achieved fraction of HBM peak on the primitive's algorithmic bytes."""
    groups = {}
    for k in kernels:
        n = k["name"]
        p = ("iir_blur" if n.startswith("iir_") else "box_blur" if n.startswith("box_blur") else
             "morphology" if n.startswith("morphology") else n)
        g = groups.setdefault(p, {"ms_per_step": 0.0, "bytes": 0.0})
        g["ms_per_step"] += k["total_ms"] / steps
        g["bytes"] += k["bytes"] / steps
    out = {}
    for p, g in sorted(groups.items(), key=lambda kv: -kv[1]["ms_per_step"]):
        gbs = g["bytes"] / (g["ms_per_step"] / 1e3) / 1e9 if g["ms_per_step"] > 0 else 0.0
        out[p] = {"ms_per_step": round(g["ms_per_step"], 3), "GB/s": round(gbs, 1),
                  "frac": round(gbs / HBM_PEAK_GBS, 4)}
    return out


# ---------------------------------------------------------------------------
def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--size", type=int, default=4096)
    ap.add_argument("--paths", type=int, default=2000)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--cpu-stride", type=int, default=128, help="CPU sample: every n-th plain path")
    ap.add_argument("--cpu-runs", type=int, default=5, help="CPU sample: timed runs (after 1 warm-up)")
    ap.add_argument("--cpu-workers", type=int, default=0, help="CPU baseline processes (0: min(16, cores))")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--batch-docs", type=int, default=4096, help="documents of the batched leg (0: skip it)")
    ap.add_argument("--batch-steps", type=int, default=1)
    ap.add_argument("--diff-size", type=int, default=1024, help="canvas side of the differentiable leg (0: skip)")
    ap.add_argument("--diff-paths", type=int, default=256)
    ap.add_argument("--fit-scenes", type=int, default=64, help="scenes of the batched fitting leg (0: skip)")
    ap.add_argument("--fit-size", type=int, default=512)
    ap.add_argument("--fit-paths", type=int, default=128)
    ap.add_argument("--fit-steps", type=int, default=500)
    ap.add_argument("--png-reps", type=int, default=3, help="timed launches of the PNG-output leg (0: skip)")
    ap.add_argument("--clip-groups", type=int, default=300, help="groups of the clip/mask leg's document (0: skip)")
    ap.add_argument("--clip-steps", type=int, default=3)
    ap.add_argument("--traffic", choices=["live", "stored"], default="live",
                    help="roofline.traffic: measured by two rocprofv3 --pmc child runs (live) or read from "
                         "profiles/traffic.json (stored)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first copy a 1 GiB device buffer once (the known-size kernel that scales HBM counters)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        return 2

    cpu = None
    if not args.no_cpu_baseline:
        workers = args.cpu_workers or max(1, min(16, len(os.sched_getaffinity(0))))
        print(f"[bench] cpu baseline: {workers} workers", file=sys.stderr, flush=True)
        cpu = cpu_baseline(args.size, args.paths, args.seed, workers, plain_stride=args.cpu_stride, runs=args.cpu_runs,
                           clip_groups=args.clip_groups)
        print(f"[bench] cpu baseline {cpu['value']} Mpixel/s", file=sys.stderr, flush=True)

    import torch
    import resvg_amd
    from resvg_amd import abi

    print(f"[bench] building the {args.size}x{args.size} document ({args.paths} paths)", file=sys.stderr, flush=True)
    sc, (w, h) = _workload(args.size, args.paths, args.seed)
    scene = sc.to_abi()
    ts = abi.Transform.identity()
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    ctx = resvg_amd.Context(0)
    px = torch.zeros((h, w, 4), dtype=torch.uint8, device=dev)
    if args.copy_calibration:
        a = torch.ones(1 << 30, dtype=torch.uint8, device=dev)
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b

    def step():
        px.zero_()
        ctx.render(scene, ts, px)

    for i in range(args.warmup):
        step()
        torch.cuda.synchronize()
        print(f"[bench] warmup {i + 1}/{args.warmup}", file=sys.stderr, flush=True)

    stream = torch.cuda.current_stream(dev)  # the context renders on torch's current stream
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    ctx.profile_begin()
    t_wall = time.perf_counter()
    start.record(stream)
    for i in range(args.steps):
        step()
        print(f"[bench] step {i + 1}/{args.steps} queued", file=sys.stderr, flush=True)
    stop.record(stream)
    torch.cuda.synchronize()
    t_wall = time.perf_counter() - t_wall
    kernels = ctx.profile_end()
    st = ctx.stats()  # of the last render call
    if cpu is not None:
        cpu["parity"] = cpu_parity(ctx, args.seed)
    pngo = png_output(ctx, px, args.png_reps) if args.png_reps > 0 else None
    clipm = clip_mask(ctx, args.size, args.clip_groups, args.seed, args.clip_steps) if args.clip_groups > 0 else None
    bat = batched(ctx, args.batch_docs, args.seed, args.batch_steps) if args.batch_docs > 0 else None
    dif = differentiable(ctx, args.diff_size, args.diff_paths, args.seed, 5) if args.diff_size > 0 else None
    fitr = fitting(ctx, args.fit_scenes, args.fit_size, args.fit_paths, args.fit_steps, args.seed) \
        if args.fit_scenes > 0 else None
    host = {"build_ms": round(st.host_build_ms, 1), "wait_ms": round(st.host_wait_ms, 1), "flushes": st.flushes,
            "edge_piece_slots": st.edge_pieces, "upload_MB": round(st.upload_bytes / 1e6, 1),
            "kernel_launches": st.kernel_launches, "layers": st.layers}
    ms = start.elapsed_time(stop) / args.steps
    mpix = w * h / 1e6
    value = mpix / (ms / 1e3)

    dom = kernels[0] if kernels else None
    roofline = None
    if dom:
        achieved = dom["bytes"] / (dom["total_ms"] / 1e3) / 1e9
        traffic, source = None, None
        tf = ROOT / "profiles" / "traffic.json"
        if tf.exists():
            rec = json.loads(tf.read_text()).get(dom["name"])
            if rec and rec.get("workload") == f"full-filter {args.size} {args.paths} {args.seed}":
                traffic, source = rec.get("bytes_per_launch"), "profiles/traffic.json"
        if args.traffic == "live":
            live = live_traffic(args)
            if live:
                traffic, source = live["bytes_per_launch"], "rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE in this run"
        roofline = {
            "bound": "hbm",
            "kernel": dom["name"],
            "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "traffic_source": source,
            "algorithmic_bytes_per_launch": round(dom["bytes"] / dom["launches"], 1),
            "avg_launch_ms": round(dom["total_ms"] / dom["launches"], 4),
            "kernel_share_of_step": round(dom["total_ms"] / (ms * args.steps), 4),
        }
        if traffic:
            # the bytes the kernel actually moves (the rocprofv3 counters above) at its
            # live launch time: far below `achieved`, since the fused compositor reads
            # and writes each target pixel once per flush rather than once per path
            moved = traffic / (dom["total_ms"] / dom["launches"] / 1e3) / 1e9
            roofline["moved_GBps"] = round(moved, 2)
            roofline["moved_frac"] = round(moved / HBM_PEAK_GBS, 4)

    line = {
        "metric": "Mpixel/s at 4096x4096 full-filter scene (resvg::render pixel pipeline)",
        "value": round(value, 3),
        "unit": "Mpixel/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic (seeded document generator, resvg_amd/workloads.py)",
        "config": {"workload": f"full-filter scene {args.size}x{args.size}, {args.paths} paths, 8 filter regions",
                   "size": args.size, "paths": args.paths, "seed": args.seed},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "kernels": [{k: (round(v, 3) if isinstance(v, float) else v) for k, v in d.items()} for d in kernels[:8]],
        "primitives": _primitive_rooflines(kernels, args.steps),
        "wall_s_per_step": round(t_wall / args.steps, 3),
        "host_last_step": host,
        "png_output": pngo,
        "clip_mask": clipm,
        "batched": bat,
        "differentiable": dif,
        "batched_fitting": fitr,
    }
    print(json.dumps(line), flush=True)
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
