"""The differentiable f32 form on the device: forward against the FP64 oracle,
analytic gradients against the oracle's FP64 central differences and against
central differences (h = 1e-3) of the device's own f32 forward."""
import numpy as np
import pytest

from oracle import diff_ref as D
from resvg_amd import diff
from resvg_amd.workloads import diff_scene

pytestmark = pytest.mark.gpu


def _setup(ctx, seed, size=40, n=6, sigma=1.3):
    import torch
    w, h, paths, ends, p0, _ = diff_scene(seed=seed, size=size, n_paths=n, segments=4, sigma=sigma)
    r = diff.Renderer(ctx, w, h, paths, ends)
    params = torch.as_tensor(p0, device="cuda")
    cov = r.coverage.cpu().numpy()
    return r, params, cov, np.asarray(ends, np.float32), w, h


@pytest.mark.parametrize("seed", [0, 1])
def test_forward_matches_fp64_oracle(ctx, seed):
    r, params, cov, ends, w, h = _setup(ctx, seed)
    out = r.forward(params).cpu().numpy()
    ref = D.forward(cov, ends, params.cpu().numpy(), w, h)
    assert np.abs(out - ref).max() < 2e-5
    assert cov.max() == 255 and (cov > 0).mean() > 0.05  # the paths cover part of the canvas


@pytest.mark.parametrize("seed", [0, 1])
def test_gradient_matches_fp64_finite_differences(ctx, seed):
    import torch
    r, params, cov, ends, w, h = _setup(ctx, seed)
    rng = np.random.default_rng(seed)
    go = rng.normal(0.0, 1.0, (h, w, 4)).astype(np.float32)
    r.forward(params)
    g = r.backward(params, torch.as_tensor(go, device="cuda")).cpu().numpy().astype(np.float64)
    ref = D.grad_fd(cov, ends, params.cpu().numpy(), w, h, go)
    err = np.linalg.norm(g - ref) / np.linalg.norm(ref)
    assert err < 1e-3, f"relative gradient error {err:.2e}"
    # every parameter group, not only the largest entries
    n = cov.shape[0]
    P = diff.PATH_PARAMS
    groups = {"tint": [i * P + k for i in range(n) for k in range(3)], "opacity": [i * P + 3 for i in range(n)],
              "stops": [i * P + k for i in range(n) for k in range(4, 20)],
              "offsets": [i * P + k for i in range(n) for k in range(20, 24)], "sigma": [n * P],
              "matrix": list(range(n * P + 1, n * P + 21))}
    for name, idx in groups.items():
        e = np.linalg.norm(g[idx] - ref[idx]) / max(np.linalg.norm(ref[idx]), 1e-12)
        assert e < 1e-3, f"{name}: relative error {e:.2e}"


def test_gradient_matches_f32_central_differences(ctx):
    # Directional derivatives against central differences of the f32 forward with
    # h = 1e-3.  The stop offsets are left out of the directions: a step of 1e-3
    # moves pixels across stop intervals, where the model has kinks.  The tolerance
    # (2%) covers the f32 rounding of the forward and the colour matrix's clamps;
    # the strict check is the FP64 one above.
    import torch
    r, params, cov, ends, w, h = _setup(ctx, 3, size=48, n=8, sigma=2.2)
    rng = np.random.default_rng(3)
    go = torch.as_tensor(rng.normal(0.0, 1.0, (h, w, 4)).astype(np.float32), device="cuda")
    r.forward(params)
    g = r.backward(params, go)
    n = cov.shape[0]
    smooth = torch.ones(params.numel(), device="cuda")
    for i in range(n):
        smooth[i * diff.PATH_PARAMS + 20:(i + 1) * diff.PATH_PARAMS] = 0.0
    for k in range(4):
        v = torch.as_tensor(rng.normal(0.0, 1.0, params.numel()).astype(np.float32), device="cuda") * smooth
        v[-21:] *= 0.1  # keep σ and the matrix inside their smooth ranges
        hstep = 1e-3
        fp = (r.forward(params + hstep * v) * go).double().sum().item()
        fm = (r.forward(params - hstep * v) * go).double().sum().item()
        fd = (fp - fm) / (2 * hstep)
        an = float((g * v).double().sum())
        assert abs(fd - an) <= 2e-2 * max(abs(an), 1.0), f"direction {k}: fd {fd} analytic {an}"


def test_autograd_function_and_fit_step(ctx):
    import torch
    w, h, paths, ends, p0, pt = diff_scene(seed=5, size=48, n_paths=8, segments=4, sigma=2.0)
    r = diff.Renderer(ctx, w, h, paths, ends)
    render = diff.render_fn()
    target = render(torch.as_tensor(pt, device="cuda"), None, r).detach()
    params = torch.as_tensor(p0, device="cuda").clone().requires_grad_(True)
    loss = ((render(params, None, r) - target) ** 2).mean()
    loss.backward()
    out = r.forward(params.detach())
    direct = r.backward(params.detach(), 2.0 * (out - target) / out.numel())
    torch.testing.assert_close(params.grad, direct, rtol=1e-5, atol=1e-8)
    # a few Adam steps lower the loss
    opt = torch.optim.Adam([params], lr=1e-2)
    first = loss.item()
    for _ in range(20):
        opt.zero_grad()
        loss = ((render(params, None, r) - target) ** 2).mean()
        loss.backward()
        opt.step()
    assert loss.item() < first


@pytest.mark.parametrize("analytic", [False, True])
def test_backward_after_a_second_forward(ctx, analytic):
    # two forwards on one renderer (two parameter sets in one loss), then the first
    # one's backward: its gradients must equal those of a lone forward + backward,
    # although the second forward replaced the workspace's intermediates
    import torch
    k = 3 if analytic else 0
    w, h, paths, ends, p0, pt = diff_scene(seed=9, size=40, n_paths=5, segments=3, sigma=1.5)
    r = diff.Renderer(ctx, w, h, paths, ends, cubics=k)
    render = diff.render_fn()
    geom = r.geom0.clone() if analytic else None
    target = torch.zeros((h, w, 4), device="cuda")
    pa = torch.as_tensor(p0, device="cuda").clone().requires_grad_(True)
    pb = torch.as_tensor(pt, device="cuda").clone().requires_grad_(True)
    ga = geom.clone().requires_grad_(True) if analytic else None
    gb = (geom * 1.01).contiguous().requires_grad_(True) if analytic else None
    out_a = render(pa, ga, r)
    out_b = render(pb, gb, r)
    ((out_a - target) ** 2).sum().backward()  # only the first forward's graph
    lone_p = torch.as_tensor(p0, device="cuda").clone().requires_grad_(True)
    lone_g = geom.clone().requires_grad_(True) if analytic else None
    ((render(lone_p, lone_g, r) - target) ** 2).sum().backward()
    assert torch.equal(pa.grad, lone_p.grad)
    if analytic:
        assert torch.equal(ga.grad, lone_g.grad)
    ((out_b - target) ** 2).sum().backward()
    assert pb.grad is not None and torch.isfinite(pb.grad).all()


# ---- analytic mode: coverage from the geometry, gradients to control points and
# transforms -------------------------------------------------------------------------
def _setup_geom(ctx, seed, size=40, n=5, k=3, sigma=1.3):
    import torch
    w, h, paths, ends, p0, _ = diff_scene(seed=seed, size=size, n_paths=n, segments=k, sigma=sigma)
    r = diff.Renderer(ctx, w, h, paths, ends, cubics=k)
    params = torch.as_tensor(p0, device="cuda")
    geom = r.geom0.clone()
    # a non-trivial affine per path: a small rotation and scale about the canvas centre
    rng = np.random.default_rng(seed)
    for i in range(n):
        a = rng.uniform(-0.1, 0.1)
        sc = rng.uniform(0.9, 1.0)
        c, s = sc * np.cos(a), sc * np.sin(a)
        cx = cy = size / 2
        geom[i, :6] = torch.tensor([c, s, -s, c, cx - c * cx + s * cy, cy - s * cx - c * cy], device="cuda")
    return r, params, geom.contiguous(), np.asarray(ends, np.float32), w, h, k


@pytest.mark.parametrize("seed", [0, 1])
def test_analytic_forward_matches_fp64_oracle(ctx, seed):
    r, params, geom, ends, w, h, k = _setup_geom(ctx, seed)
    out = r.forward(params, geom).cpu().numpy()
    ref = D.forward_geom(geom.cpu().numpy(), k, ends, params.cpu().numpy(), w, h)
    assert np.abs(out - ref).max() < 1e-4


@pytest.mark.parametrize("seed", [0, 1])
def test_analytic_gradient_matches_fp64_finite_differences(ctx, seed):
    import torch
    r, params, geom, ends, w, h, k = _setup_geom(ctx, seed)
    rng = np.random.default_rng(10 + seed)
    go = rng.normal(0.0, 1.0, (h, w, 4)).astype(np.float32)
    r.forward(params, geom)
    gp, gg = r.backward(params, torch.as_tensor(go, device="cuda"), geom)
    gg = gg.cpu().numpy().astype(np.float64).reshape(-1)
    gnp = geom.cpu().numpy()
    idx = sorted(rng.choice(gg.size, size=50, replace=False).tolist() + list(range(6)))  # every path's affine too
    ref = D.grad_fd_geom(gnp, k, ends, params.cpu().numpy(), w, h, go, idx)
    e = np.linalg.norm(gg[idx] - ref[idx]) / np.linalg.norm(ref[idx])
    assert e < 1e-3, f"geometry: relative error {e:.2e}"
    # the paint gradients with the analytic coverage
    cov = D.analytic_coverage(gnp, k, w, h) * 255.0
    refp = D.grad_fd(cov, ends, params.cpu().numpy(), w, h, go)
    gp = gp.cpu().numpy().astype(np.float64)
    e = np.linalg.norm(gp - refp) / np.linalg.norm(refp)
    assert e < 1e-3, f"paint: relative error {e:.2e}"


def test_analytic_paths_leaving_the_canvas(ctx):
    """Paths shifted partly off the canvas (left and bottom for some, right and top
    for others): the coverage is still the exact winding integral over each canvas
    pixel, and the geometry gradient is its boundary term inside the canvas."""
    import torch
    r, params, geom, ends, w, h, k = _setup_geom(ctx, 2)
    shift = torch.tensor([[-14.0, 11.0], [15.0, -12.0], [-9.0, -9.0], [12.0, 13.0], [0.0, 0.0]], device="cuda")
    geom[:, 4:6] += shift
    out = r.forward(params, geom).cpu().numpy()
    gnp = geom.cpu().numpy()
    ref = D.forward_geom(gnp, k, ends, params.cpu().numpy(), w, h)
    assert np.abs(out - ref).max() < 1e-4
    rng = np.random.default_rng(12)
    go = rng.normal(0.0, 1.0, (h, w, 4)).astype(np.float32)
    _, gg = r.backward(params, torch.as_tensor(go, device="cuda"), geom)
    gg = gg.cpu().numpy().astype(np.float64).reshape(-1)
    idx = sorted(rng.choice(gg.size, size=50, replace=False).tolist() + list(range(6)))
    fd = D.grad_fd_geom(gnp, k, ends, params.cpu().numpy(), w, h, go, idx)
    e = np.linalg.norm(gg[idx] - fd[idx]) / np.linalg.norm(fd[idx])
    assert e < 1e-3, f"geometry: relative error {e:.2e}"


def test_analytic_gradient_matches_f32_central_differences(ctx):
    import torch
    r, params, geom, ends, w, h, k = _setup_geom(ctx, 4, size=48, n=6, k=4, sigma=2.2)
    rng = np.random.default_rng(4)
    go = torch.as_tensor(rng.normal(0.0, 1.0, (h, w, 4)).astype(np.float32), device="cuda")
    r.forward(params, geom)
    _, gg = r.backward(params, go, geom)
    for t in range(4):
        v = torch.as_tensor(rng.normal(0.0, 1.0, geom.shape).astype(np.float32), device="cuda")
        v[:, :4] *= 0.01  # the linear part of the affine moves points by up to the canvas size
        hstep = 1e-3
        fp = (r.forward(params, geom + hstep * v) * go).double().sum().item()
        fm = (r.forward(params, geom - hstep * v) * go).double().sum().item()
        fd = (fp - fm) / (2 * hstep)
        an = float((gg * v).double().sum())
        assert abs(fd - an) <= 2e-2 * max(abs(an), 1.0), f"direction {t}: fd {fd} analytic {an}"


def test_analytic_fit_moves_the_geometry(ctx):
    import torch
    r, params, geom, ends, w, h, k = _setup_geom(ctx, 6, size=48, n=6, k=4, sigma=1.5)
    render = diff.render_fn()
    target_geom = geom.clone()
    target_geom[:, 4:6] += 1.5  # the target is the same scene shifted by 1.5 px
    target = render(params, target_geom, r).detach()
    g = geom.clone().requires_grad_(True)
    opt = torch.optim.Adam([g], lr=0.05)
    first = None
    for _ in range(30):
        opt.zero_grad()
        loss = ((render(params, g, r) - target) ** 2).mean()
        loss.backward()
        first = loss.item() if first is None else first
        opt.step()
    assert loss.item() < 0.5 * first


# ---- the batch form: one launch per kernel for a stack of scenes -------------------
@pytest.mark.parametrize("cubics", [0, 3])
def test_batch_matches_one_call_per_scene(ctx, cubics):
    """rvk_diff_forward_batch / rvk_diff_backward_batch give, for every scene of the
    stack, exactly what rvk_diff_forward / rvk_diff_backward give for it alone (the
    same kernels, per-scene reductions in the same order)."""
    import torch
    size, n, k = 40, 5, (cubics or 4)
    scenes, ps, gs = [], [], []
    for b in range(3):
        w, h, paths, ends, p0, _ = diff_scene(seed=20 + b, size=size, n_paths=n, segments=k, sigma=1.0 + 0.6 * b)
        scenes.append((paths, ends))
        ps.append(p0)
        gs.append(diff.geometry(paths, k))
    rb = diff.Renderer.batched(ctx, size, size, scenes, cubics=cubics)
    P = torch.as_tensor(np.stack(ps), device="cuda").contiguous()
    G = torch.as_tensor(np.stack(gs), device="cuda").contiguous() if cubics else None
    if cubics:
        G[:, :, 4:6] += torch.tensor([0.7, -0.4], device="cuda")  # a non-trivial translation
    rng = np.random.default_rng(7)
    go = torch.as_tensor(rng.normal(0.0, 1.0, (3, size, size, 4)).astype(np.float32), device="cuda")
    out_b = rb.forward(P, G).clone()
    grads_b = rb.backward(P, go, G)
    for b in range(3):
        r1 = diff.Renderer(ctx, size, size, scenes[b][0], scenes[b][1], cubics=cubics)
        g1 = G[b].contiguous() if cubics else None
        out1 = r1.forward(P[b].contiguous(), g1)
        assert torch.equal(out1, out_b[b]), f"scene {b}: forward differs"
        grads1 = r1.backward(P[b].contiguous(), go[b], g1)
        if cubics:
            assert torch.equal(grads1[0], grads_b[0][b]) and torch.equal(grads1[1], grads_b[1][b]), f"scene {b}"
        else:
            assert torch.equal(grads1, grads_b[b]), f"scene {b}: parameter gradients differ"


def test_batch_fit_lowers_every_scene_loss(ctx):
    from resvg_amd import fit
    wl = fit.fit_workload(count=4, size=64, n_paths=6, segments=3, seed=3)
    res = fit.fit(ctx, wl, steps=60, lr=1e-2)
    assert res["mse_ratio_end_over_start"] < 0.8, res
    assert res["psnr_f32_db"] > res["psnr_start_db"] + 1.0, res
    assert np.isfinite(res["psnr_u8_db"]) and res["psnr_u8_db"] > 5.0, res


def test_analytic_mode_is_deterministic(ctx):
    """The winding-area planes are accumulated in fixed point, so repeated forwards
    and backwards of a busy scene (many edges per row) are bit-identical."""
    import torch
    w, h, paths, ends, p0, _ = diff_scene(seed=8, size=256, n_paths=48, segments=8, sigma=2.0)
    r = diff.Renderer(ctx, w, h, paths, ends, cubics=8)
    params = torch.as_tensor(p0, device="cuda")
    go = torch.as_tensor(np.random.default_rng(8).normal(0.0, 1.0, (h, w, 4)).astype(np.float32), device="cuda")
    runs = []
    for _ in range(3):
        out = r.forward(params, r.geom0).clone()
        gp, gg = r.backward(params, go, r.geom0)
        runs.append((out, gp.clone(), gg.clone()))
    for out, gp, gg in runs[1:]:
        assert torch.equal(out, runs[0][0]) and torch.equal(gp, runs[0][1]) and torch.equal(gg, runs[0][2])


def test_scene_limits_are_rejected(ctx):
    """Scenes whose vertex or prefix-row counts exceed the kernels' 32-bit indexing
    are refused with RVK_ERR_INVALID_ARGUMENT before anything is launched."""
    import ctypes as C
    import torch
    from resvg_amd import abi
    L = diff._bind()
    buf = torch.zeros(64, dtype=torch.float32, device="cuda")
    p = C.c_void_p(buf.data_ptr())
    for n, k, h in ((140000, 1000, 16), (3_000_000, 0, 1024)):
        sc = diff.DiffScene(16, h, n, k, p if k == 0 else None, p)
        st = L.rvk_diff_forward_batch(ctx.h, C.byref(sc), 1, p, p if k else None, p, p)
        assert st == abi.ERR_INVALID_ARGUMENT, (n, k, h, st)


def _styled_scene(size=48, seed=3):
    """Path entries of every coverage kind the plane mode takes: an even-odd star
    (its centre left empty), the same star nonzero (centre filled), a round-joined
    stroke of a random cubic chain and a dashed square stroke."""
    from resvg_amd import abi
    rng = np.random.default_rng(seed)
    c = size / 2.0
    star = [(c + 0.45 * size * np.cos(np.pi / 2 + k * 4 * np.pi / 5), c - 0.45 * size * np.sin(np.pi / 2 + k * 4 * np.pi / 5))
            for k in range(5)]
    sv = [abi.VERB_MOVE] + [abi.VERB_LINE] * 4 + [abi.VERB_CLOSE]
    cub = rng.uniform(4, size - 4, (1 + 3 * 3, 2)).astype(np.float32)
    cv = [abi.VERB_MOVE] + [abi.VERB_CUBIC] * 3
    sq = [(8.0, 8.0), (size - 8.0, 8.0), (size - 8.0, size - 8.0), (8.0, size - 8.0)]
    qv = [abi.VERB_MOVE] + [abi.VERB_LINE] * 3 + [abi.VERB_CLOSE]
    paths = [(sv, star, {"fill_rule": abi.FILL_EVENODD}), (sv, star, None),
             (cv, cub, {"stroke_width": 3.0, "join": abi.JOIN_ROUND, "cap": abi.CAP_ROUND}),
             (qv, sq, {"stroke_width": 2.0, "dasharray": [5.0, 3.0], "dashoffset": 1.0})]
    n = len(paths)
    ends = rng.uniform(0, size, (n, 4)).astype(np.float32)
    q = []
    for _ in range(n):
        v = np.empty(diff.PATH_PARAMS)
        v[0:3], v[3] = rng.uniform(0.5, 1.0, 3), rng.uniform(0.3, 1.0)
        v[4:20], v[20:24] = rng.uniform(0.1, 0.9, 16), np.sort(rng.uniform(0.05, 0.95, 4))
        q.append(v)
    m = np.eye(4, 5) + rng.normal(0.0, 0.1, (4, 5))
    params = np.concatenate(q + [np.array([1.1]), m.reshape(-1)]).astype(np.float32)
    return size, paths, ends, params


def test_planes_of_evenodd_fills_and_strokes(ctx):
    """The coverage-plane mode with even-odd fills, strokes and dashes: each plane is
    the u8 rasterizer's coverage of the entry (bit-exact against the oracle), the
    forward matches the FP64 oracle and the gradient its finite differences."""
    import torch
    from oracle import bindings as orc
    from resvg_amd import abi
    size, paths, ends, params = _styled_scene()
    r = diff.Renderer(ctx, size, size, paths, ends)
    cov = r.coverage.cpu().numpy()
    for i, path in enumerate(paths):
        px = np.zeros((size, size, 4), np.uint8)
        orc.render(diff.one_path_scene((size, size), *diff._entry(path)).to_abi().abi, abi.Transform.identity(), px)
        assert np.array_equal(cov[i], px[..., 3]), f"plane {i}"
    c = size // 2
    assert cov[0][c, c] == 0 and cov[1][c, c] == 255  # even-odd leaves the star's centre empty
    assert 0 < (cov[2] > 0).mean() < 0.5 and 0 < (cov[3] > 0).mean() < 0.3  # strokes: thin outlines
    p = torch.as_tensor(params, device="cuda")
    out = r.forward(p).cpu().numpy()
    assert np.abs(out - D.forward(cov, ends, params, size, size)).max() < 2e-5
    go = np.random.default_rng(5).normal(0.0, 1.0, (size, size, 4)).astype(np.float32)
    g = r.backward(p, torch.as_tensor(go, device="cuda")).cpu().numpy().astype(np.float64)
    ref = D.grad_fd(cov, ends, params, size, size, go)
    assert np.linalg.norm(g - ref) / np.linalg.norm(ref) < 1e-3
