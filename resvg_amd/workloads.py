"""Seeded synthetic documents for the three plain-method workloads.

* ``blur_offset_composite_doc`` — 512x512: one closed nonzero path of 64 cubics with
  control points uniform in the canvas, filled with a 4-stop linear gradient, inside
  a group filtered by feGaussianBlur(4) -> feOffset(3,3) -> feComposite(over
  SourceGraphic) with the default filter region.
* ``full_filter_doc`` — 4096x4096 by default: N paths (3-200 cubic/quad segments,
  points uniform in the canvas, evenodd/nonzero 50/50, 40% also stroked 0.5-8 px with
  round/miter/bevel joins), each with a 2-8 stop linear or radial gradient and in its
  own group with opacity U[0.2, 1]; 8 filter groups with userSpaceOnUse regions
  covering blur (IIR and box), morphology, 3x3/5x5 convolution, colour matrix,
  component transfer, turbulence, distant/point/spot lighting and multiply/screen.
* ``document_batch`` — documents with log-uniform sides 64-2048 (aspect 0.5-2), 5-200
  paths of the full-filter distributions and a 1-6 primitive filter chain each.
* ``clip_mask_doc`` — 4096x4096 by default: 300 groups of 1-3 such paths, clipped
  (60%) and masked (40%) by 16 clip paths and 16 masks that nest and clip their own
  children.

Random numbers come from numpy's PCG64 seeded per document; every float is rounded
to f32 before use, so the same document is produced on every machine.
"""
from __future__ import annotations

import numpy as np

from . import abi
from .filters import SOURCE_ALPHA, SOURCE_GRAPHIC, FilterGraph, distant_light, point_light, spot_light
from .scene import PaintSpec, Scene

F32 = np.float32


def _f(v):
    return float(F32(v))


def _stops(rng, n):
    offs = np.sort(rng.uniform(0.0, 1.0, size=n)).astype(F32)
    cols = rng.integers(0, 256, size=(n, 3))
    return [(_f(o), tuple(int(c) for c in col), 1.0) for o, col in zip(offs, cols)]


def _cubic_extrema_bounds(p0, p1, p2, p3):
    xs, ys = [p0[0], p3[0]], [p0[1], p3[1]]
    for k in range(2):
        a = -p0[k] + 3 * p1[k] - 3 * p2[k] + p3[k]
        b = 2 * (p0[k] - 2 * p1[k] + p2[k])
        c = p1[k] - p0[k]
        roots = []
        if abs(a) < 1e-12:
            if abs(b) > 1e-12:
                roots.append(-c / b)
        else:
            disc = b * b - 4 * a * c
            if disc >= 0:
                s = np.sqrt(disc)
                roots += [(-b + s) / (2 * a), (-b - s) / (2 * a)]
        for t in roots:
            if 0 < t < 1:
                u = 1 - t
                v = u ** 3 * p0[k] + 3 * u * u * t * p1[k] + 3 * u * t * t * p2[k] + t ** 3 * p3[k]
                (xs if k == 0 else ys).append(v)
    return min(xs), min(ys), max(xs), max(ys)


def blur_offset_composite_doc(seed: int = 0, size: int = 512):
    """The 512x512 single-path document (returns the Scene and its canvas size)."""
    rng = np.random.default_rng(seed)
    pts = rng.uniform(0, size, size=(1 + 64 * 3, 2)).astype(F32)
    pts[-1] = pts[0]  # p64 = p0
    verbs = [abi.VERB_MOVE] + [abi.VERB_CUBIC] * 64 + [abi.VERB_CLOSE]
    # tight bounding box (usvg's objectBoundingBox) for the default filter region
    bx0 = by0 = np.inf
    bx1 = by1 = -np.inf
    for i in range(64):
        q = [tuple(map(float, pts[3 * i + k])) for k in range(4)]
        a, b, c, d = _cubic_extrema_bounds(*q)
        bx0, by0, bx1, by1 = min(bx0, a), min(by0, b), max(bx1, c), max(by1, d)
    bw, bh = bx1 - bx0, by1 - by0
    region = (_f(bx0 - 0.1 * bw), _f(by0 - 0.1 * bh), _f(1.2 * bw), _f(1.2 * bh))
    g = FilterGraph(region)
    blur = g.gaussian_blur(4.0)
    off = g.offset(3.0, 3.0, in1=blur)
    g.composite(abi.COMPOSITE_OVER, in1=off, in2=SOURCE_GRAPHIC)
    e = rng.uniform(0, size, size=4).astype(F32)
    paint = PaintSpec.linear(_f(e[0]), _f(e[1]), _f(e[2]), _f(e[3]), _stops(rng, 4))
    sc = Scene(size, size)
    grp = sc.group(sc.root, filters=[g])
    sc.path(grp, verbs, pts, fill=paint, fill_rule=abi.FILL_NONZERO)
    return sc, (size, size)


def _random_path(rng, size, nseg):
    verbs = [abi.VERB_MOVE]
    kinds = rng.integers(0, 2, size=nseg)
    npts = 1 + int(sum(3 if k else 2 for k in kinds))
    pts = rng.uniform(0, size, size=(npts, 2)).astype(F32)
    for k in kinds:
        verbs.append(abi.VERB_CUBIC if k else abi.VERB_QUAD)
    verbs.append(abi.VERB_CLOSE)
    return verbs, pts


def _random_paint(rng, size, nstops):
    stops = _stops(rng, nstops)
    if rng.random() < 0.5:
        e = rng.uniform(0, size, size=4).astype(F32)
        return PaintSpec.linear(_f(e[0]), _f(e[1]), _f(e[2]), _f(e[3]), stops)
    c = rng.uniform(0, size, size=2).astype(F32)
    r = _f(rng.uniform(0.1, 0.6) * size)
    return PaintSpec.radial(_f(c[0]), _f(c[1]), r, stops)


def _filter_chains(rng, size):
    """Eight filter graphs whose primitives together cover the full-filter workload."""
    out = []

    def region():
        w, h = rng.uniform(0.2, 0.45, size=2) * size
        x, y = rng.uniform(0, size - w), rng.uniform(0, size - h)
        return (_f(x), _f(y), _f(w), _f(h))

    g = FilterGraph(region())  # IIR blur + multiply
    b = g.gaussian_blur(0.5)
    g.blend("multiply", in1=b, in2=SOURCE_GRAPHIC)
    out.append(g)
    g = FilterGraph(region())  # IIR blur + colour matrix
    b = g.gaussian_blur(1.5)
    g.color_matrix(abi.COLOR_MATRIX_MATRIX, rng.uniform(-1, 1, 20).astype(F32).tolist(), in1=b)
    out.append(g)
    g = FilterGraph(region())  # box blur + dilate
    b = g.gaussian_blur(6.0)
    g.morphology(abi.MORPH_DILATE, _f(rng.uniform(1, 8)), in1=b)
    out.append(g)
    g = FilterGraph(region())  # box blur sigma 24 + screen
    b = g.gaussian_blur(24.0)
    g.blend("screen", in1=b, in2=SOURCE_GRAPHIC)
    out.append(g)
    g = FilterGraph(region())  # erode + 3x3 convolution
    m = g.morphology(abi.MORPH_ERODE, _f(rng.uniform(1, 8)))
    g.convolve_matrix(3, 3, rng.integers(-4, 5, 9).astype(float), in1=m)
    out.append(g)
    g = FilterGraph(region())  # 5x5 convolution + component transfer
    cv = g.convolve_matrix(5, 5, rng.integers(-4, 5, 25).astype(float))
    g.component_transfer([("table", rng.uniform(0, 1, 5).astype(F32).tolist()), ("gamma", 1.2, 0.8, 0.05),
                          ("gamma", 0.9, 2.2, 0.0), None], in1=cv)
    out.append(g)
    g = FilterGraph(region())  # fractal turbulence composited into the source
    t = g.turbulence(_f(rng.uniform(0.01, 0.1)), _f(rng.uniform(0.01, 0.1)), 4, int(rng.integers(0, 1000)),
                     False, True)
    g.composite(abi.COMPOSITE_IN, in1=t, in2=SOURCE_GRAPHIC)
    out.append(g)
    x, y, w, h = region()
    g = FilterGraph((x, y, w, h))  # distant / point / spot lighting on the source alpha
    d = g.lighting(False, distant_light(_f(rng.uniform(0, 360)), _f(rng.uniform(10, 80))), 2.0, 1.0,
                   in1=SOURCE_ALPHA)
    s = g.lighting(True, point_light(_f(x + w / 2), _f(y + h / 2), _f(0.2 * w)), 2.0, 1.0, 12.0, in1=SOURCE_ALPHA)
    ar = g.composite(abi.COMPOSITE_ARITHMETIC, k=(0.0, 1.0, 1.0, 0.0), in1=d, in2=s)
    sp = g.lighting(False, spot_light(_f(x), _f(y), _f(0.3 * w), _f(x + w), _f(y + h), 0.0, 2.0, 35.0), 2.0, 1.0,
                    in1=SOURCE_ALPHA)
    g.blend("multiply", in1=sp, in2=ar)
    out.append(g)
    return out


def full_filter_doc(seed: int = 0, size: int = 4096, n_paths: int = 2000, paths_per_filter: int = 25,
                    max_segments: int = 200, keep=None):
    """The full-filter document.  `keep(item, is_filter_group) -> bool`, when given,
    selects top-level items (a filter group with its paths, or one path in its opacity
    group, numbered in document order) without changing the random stream, so a kept
    item is identical to the same item of the whole document."""
    rng = np.random.default_rng(seed)
    sc = Scene(size, size)
    chains = _filter_chains(rng, size)
    n_filtered = min(n_paths, paths_per_filter * len(chains))
    # filter groups are spread evenly through the document order
    starts = {int(i * n_paths / len(chains)): k for k, i in enumerate(range(len(chains)))}
    i = 0
    item = -1
    while i < n_paths:
        parent = sc.root
        count = 1
        item += 1
        is_filter = i in starts and n_filtered > 0
        kept = keep is None or keep(item, is_filter)
        if is_filter:
            if kept:
                parent = sc.group(sc.root, filters=[chains[starts[i]]])
            count = min(paths_per_filter, n_paths - i)
        for _ in range(count):
            nseg = int(rng.integers(3, max_segments + 1))
            verbs, pts = _random_path(rng, size, nseg)
            fill = _random_paint(rng, size, int(rng.integers(2, 9)))
            rule = abi.FILL_EVENODD if rng.random() < 0.5 else abi.FILL_NONZERO
            stroke, width, join = None, 1.0, abi.JOIN_MITER
            if rng.random() < 0.4:
                stroke = _random_paint(rng, size, int(rng.integers(2, 9)))
                width = _f(rng.uniform(0.5, 8.0))
                join = [abi.JOIN_ROUND, abi.JOIN_MITER, abi.JOIN_BEVEL][int(rng.integers(3))]
            opacity = _f(rng.uniform(0.2, 1.0))
            if kept:
                grp = sc.group(parent, opacity=opacity)
                sc.path(grp, verbs, pts, fill=fill, fill_rule=rule, stroke=stroke, stroke_width=width,
                        stroke_join=join)
            i += 1
    return sc, (size, size)


def clip_mask_doc(seed: int = 0, size: int = 4096, n_groups: int = 300, n_clips: int = 16, n_masks: int = 16,
                  max_segments: int = 40):
    """The clip/mask document: `n_groups` top-level groups with opacity U[0.5, 1], 60%
    clipped by one of `n_clips` clip paths and 40% masked by one of `n_masks` masks
    (independently), each holding 1-3 paths of the full-filter distributions (3 to
    `max_segments` segments, a 2-8 stop gradient, 30% also stroked).  A clip path holds
    1-3 black paths (even-odd or nonzero), half of them under a transform of its own,
    a third nesting the clip path before it, and 30% of its children sit in a group
    clipped again by an earlier clip path; a mask covers a random rect, luminance (60%)
    or alpha, a third nesting the mask before it, with 1-3 gradient-filled paths."""
    rng = np.random.default_rng(seed)
    sc = Scene(size, size)

    def rand_ts():
        a, k = _f(rng.uniform(0.6, 1.3)), _f(rng.uniform(-0.3, 0.3))
        return (a, k, -k, a, _f(rng.uniform(-0.1, 0.1) * size), _f(rng.uniform(-0.1, 0.1) * size))

    def paths_into(parent, count, paint):
        for _ in range(count):
            verbs, pts = _random_path(rng, size, int(rng.integers(3, max_segments + 1)))
            stroke = _random_paint(rng, size, int(rng.integers(2, 9))) if paint and rng.random() < 0.3 else None
            sc.path(parent, verbs, pts,
                    fill=_random_paint(rng, size, int(rng.integers(2, 9))) if paint else PaintSpec.solid((0, 0, 0)),
                    fill_rule=abi.FILL_EVENODD if rng.random() < 0.5 else abi.FILL_NONZERO,
                    stroke=stroke, stroke_width=_f(rng.uniform(0.5, 8.0)))

    clips = []
    for _ in range(n_clips):
        nested = clips[-1] if clips and rng.random() < 1 / 3 else -1
        c, root = sc.clip_path(transform=rand_ts() if rng.random() < 0.5 else (1.0, 0.0, 0.0, 1.0, 0.0, 0.0),
                               clip_path=nested)
        for _ in range(int(rng.integers(1, 4))):
            parent = root
            if clips and rng.random() < 0.3:
                parent = sc.group(root, clip_path=clips[int(rng.integers(len(clips)))])
            paths_into(parent, 1, False)
        clips.append(c)
    masks = []
    for _ in range(n_masks):
        nested = masks[-1] if masks and rng.random() < 1 / 3 else -1
        x, y = rng.uniform(-0.1 * size, 0.5 * size, 2)
        wm, hm = rng.uniform(0.3 * size, size, 2)
        m, root = sc.mask((_f(x), _f(y), _f(wm), _f(hm)),
                          kind=abi.MASK_LUMINANCE if rng.random() < 0.6 else abi.MASK_ALPHA, mask=nested)
        paths_into(root, int(rng.integers(1, 4)), True)
        masks.append(m)
    for _ in range(n_groups):
        clip = clips[int(rng.integers(len(clips)))] if rng.random() < 0.6 else -1
        mask = masks[int(rng.integers(len(masks)))] if rng.random() < 0.4 else -1
        g = sc.group(sc.root, opacity=_f(rng.uniform(0.5, 1.0)), clip_path=clip, mask=mask)
        paths_into(g, int(rng.integers(1, 4)), True)
    return sc, (size, size)


def _sampled_chain(rng, w, h, n_prims):
    """A filter of `n_prims` primitives drawn from the full-filter primitive set, each
    reading the previous one's result, over a region covering 40-100% of each side."""
    rw, rh = rng.uniform(0.4, 1.0, size=2) * (w, h)
    x, y = rng.uniform(0, w - rw), rng.uniform(0, h - rh)
    g = FilterGraph((_f(x), _f(y), _f(rw), _f(rh)))
    prev = SOURCE_GRAPHIC
    for _ in range(n_prims):
        k = int(rng.integers(0, 10))
        if k == 0:
            prev = g.gaussian_blur(float(rng.choice([0.5, 1.5, 6.0, 24.0])), in1=prev)
        elif k == 1:
            op = abi.MORPH_ERODE if rng.random() < 0.5 else abi.MORPH_DILATE
            prev = g.morphology(op, _f(rng.uniform(1, 8)), in1=prev)
        elif k == 2:
            n = 3 if rng.random() < 0.5 else 5
            prev = g.convolve_matrix(n, n, rng.integers(-4, 5, n * n).astype(float), in1=prev)
        elif k == 3:
            prev = g.color_matrix(abi.COLOR_MATRIX_MATRIX, rng.uniform(-1, 1, 20).astype(F32).tolist(), in1=prev)
        elif k == 4:
            prev = g.component_transfer([("table", rng.uniform(0, 1, 5).astype(F32).tolist()),
                                         ("gamma", 1.2, 0.8, 0.05), ("gamma", 0.9, 2.2, 0.0), None], in1=prev)
        elif k == 5:
            t = g.turbulence(_f(rng.uniform(0.01, 0.1)), _f(rng.uniform(0.01, 0.1)), 4,
                             int(rng.integers(0, 1000)), False, True)
            prev = g.composite(abi.COMPOSITE_IN, in1=t, in2=prev)
        elif k == 6:
            light = distant_light(_f(rng.uniform(0, 360)), _f(rng.uniform(10, 80)))
            prev = g.lighting(False, light, 2.0, 1.0, in1=prev)
        elif k == 7:
            light = point_light(_f(x + rw / 2), _f(y + rh / 2), _f(0.2 * rw))
            prev = g.lighting(True, light, 2.0, 1.0, 12.0, in1=prev)
        elif k == 8:
            light = spot_light(_f(x), _f(y), _f(0.3 * rw), _f(x + rw), _f(y + rh), 0.0, 2.0, 35.0)
            prev = g.lighting(False, light, 2.0, 1.0, in1=prev)
        else:
            prev = g.blend("multiply" if rng.random() < 0.5 else "screen", in1=prev, in2=SOURCE_GRAPHIC)
    return g


def batch_doc(seed: int, w: int, h: int, n_paths: int, n_prims: int):
    """One document of the batched workload: `n_paths` paths of the full-filter
    distributions (control points and gradient geometry uniform in the w x h canvas),
    each in its own opacity group; the paths of the middle third sit in a group with a
    `n_prims`-primitive filter chain."""
    rng = np.random.default_rng(seed)
    ext = np.array([w, h], dtype=np.float64)
    ext4 = np.array([w, h, w, h], dtype=np.float64)
    sc = Scene(w, h)
    fgroup = sc.group(sc.root, filters=[_sampled_chain(rng, w, h, n_prims)])
    lo, hi = n_paths // 3, max(n_paths // 3 + 1, 2 * n_paths // 3)
    for i in range(n_paths):
        parent = fgroup if lo <= i < hi else sc.root
        nseg = int(rng.integers(3, 201))
        verbs, pts = _random_path(rng, ext, nseg)
        stops = _stops(rng, int(rng.integers(2, 9)))

        def paint():
            if rng.random() < 0.5:
                e = rng.uniform(0, ext4).astype(F32)
                return PaintSpec.linear(_f(e[0]), _f(e[1]), _f(e[2]), _f(e[3]), stops)
            c = rng.uniform(0, ext).astype(F32)
            return PaintSpec.radial(_f(c[0]), _f(c[1]), _f(rng.uniform(0.1, 0.6) * min(w, h)), stops)

        fill = paint()
        rule = abi.FILL_EVENODD if rng.random() < 0.5 else abi.FILL_NONZERO
        stroke, width, join = None, 1.0, abi.JOIN_MITER
        if rng.random() < 0.4:
            stops = _stops(rng, int(rng.integers(2, 9)))
            stroke = paint()
            width = _f(rng.uniform(0.5, 8.0))
            join = [abi.JOIN_ROUND, abi.JOIN_MITER, abi.JOIN_BEVEL][int(rng.integers(3))]
        grp = sc.group(parent, opacity=_f(rng.uniform(0.2, 1.0)))
        sc.path(grp, verbs, pts, fill=fill, fill_rule=rule, stroke=stroke, stroke_width=width, stroke_join=join)
    return sc


def document_specs(n_docs: int, seed: int = 0, min_side: int = 64, max_side: int = 2048, min_paths: int = 5,
                   max_paths: int = 200):
    """The batched workload's documents as batch_doc arguments (seed, w, h, paths,
    filter primitives): sides log-uniform in [min_side, max_side], aspect 0.5-2,
    5-200 paths and a 1-6 primitive filter chain each."""
    rng = np.random.default_rng(seed)
    specs = []
    for _ in range(n_docs):
        side = float(np.exp(rng.uniform(np.log(min_side), np.log(max_side))))
        aspect = float(np.exp(rng.uniform(np.log(0.5), np.log(2.0))))
        w = int(max(min_side, min(max_side, round(side * np.sqrt(aspect)))))
        h = int(max(min_side, min(max_side, round(side / np.sqrt(aspect)))))
        n = int(rng.integers(min_paths, max_paths + 1))
        k = int(rng.integers(1, 7))
        specs.append((int(rng.integers(0, 2**31)), w, h, n, k))
    return specs


def document_batch(n_docs: int, seed: int = 0, min_side: int = 64, max_side: int = 2048, min_paths: int = 5,
                   max_paths: int = 200):
    """The batched workload (document_specs, built).  Returns [(Scene, (w, h))]."""
    return [(batch_doc(*sp), (sp[1], sp[2]))
            for sp in document_specs(n_docs, seed, min_side, max_side, min_paths, max_paths)]


def diff_scene(seed: int = 0, size: int = 1024, n_paths: int = 256, segments: int = 8, sigma: float = 4.0):
    """The differentiable workload: `n_paths` closed paths of `segments` cubics with
    points uniform in the canvas, each with a linear gradient between two uniform
    points.  Returns (width, height, paths, ends, params, target_params): parameters
    in resvg_amd.diff's layout (tint, opacity, 4 stops, 4 sorted offsets per path,
    then σ and a colour matrix near the identity), and the target's parameters
    perturbed by N(0, 0.05)."""
    from .diff import PATH_PARAMS
    rng = np.random.default_rng(seed)
    paths, ends, params = [], [], []
    for _ in range(n_paths):
        pts = rng.uniform(0, size, size=(1 + 3 * segments, 2)).astype(F32)
        pts[-1] = pts[0]
        verbs = [abi.VERB_MOVE] + [abi.VERB_CUBIC] * segments + [abi.VERB_CLOSE]
        paths.append((verbs, pts))
        ends.append(rng.uniform(0, size, size=4).astype(F32))
        q = np.empty(PATH_PARAMS, np.float64)
        q[0:3] = rng.uniform(0.5, 1.0, 3)
        q[3] = rng.uniform(0.3, 1.0)
        q[4:20] = rng.uniform(0.1, 0.9, 16)
        q[20:24] = np.sort(rng.uniform(0.05, 0.95, 4))
        params.append(q)
    m = np.eye(4, 5) + rng.normal(0.0, 0.1, (4, 5))
    flat = np.concatenate(params + [np.array([sigma]), m.reshape(-1)]).astype(np.float32)
    target = (flat.astype(np.float64) + rng.normal(0.0, 0.05, flat.shape)).astype(np.float32)
    for i in range(n_paths):  # keep the target's offsets ascending
        o = target[i * PATH_PARAMS + 20:i * PATH_PARAMS + 24]
        target[i * PATH_PARAMS + 20:i * PATH_PARAMS + 24] = np.sort(o)
    return size, size, paths, ends, flat, target


def main(argv=None) -> int:
    """Writes a seeded workload document as SVG (so that the reference can render it):
    ``python -m resvg_amd.workloads full-filter --size 4096 --paths 2000 --seed 0 -o c2.svg``;
    ``batch`` writes documents ``--first`` .. ``--first + --count - 1`` of the batched
    workload into the directory given by ``-o``."""
    import argparse
    from pathlib import Path

    from .svg_export import scene_to_svg

    ap = argparse.ArgumentParser(prog="python -m resvg_amd.workloads", description=main.__doc__)
    ap.add_argument("doc", choices=["blur-offset-composite", "full-filter", "batch"])
    ap.add_argument("-o", "--out", required=True)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--size", type=int, default=None)
    ap.add_argument("--paths", type=int, default=2000)
    ap.add_argument("--docs", type=int, default=4096, help="batch: documents in the workload")
    ap.add_argument("--first", type=int, default=0)
    ap.add_argument("--count", type=int, default=1)
    a = ap.parse_args(argv)
    if a.doc == "batch":
        out = Path(a.out)
        out.mkdir(parents=True, exist_ok=True)
        specs = document_specs(a.docs, a.seed)
        for i in range(a.first, min(a.first + a.count, len(specs))):
            (out / f"doc_{i:04d}.svg").write_text(scene_to_svg(batch_doc(*specs[i])))
        return 0
    if a.doc == "blur-offset-composite":
        sc, _ = blur_offset_composite_doc(a.seed, a.size or 512)
    else:
        sc, _ = full_filter_doc(a.seed, a.size or 4096, a.paths)
    Path(a.out).write_text(scene_to_svg(sc))
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
