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
    target = render(torch.as_tensor(pt, device="cuda"), r).detach()
    params = torch.as_tensor(p0, device="cuda").clone().requires_grad_(True)
    loss = ((render(params, r) - target) ** 2).mean()
    loss.backward()
    out = r.forward(params.detach())
    direct = r.backward(params.detach(), 2.0 * (out - target) / out.numel())
    torch.testing.assert_close(params.grad, direct, rtol=1e-5, atol=1e-8)
    # a few Adam steps lower the loss
    opt = torch.optim.Adam([params], lr=1e-2)
    first = loss.item()
    for _ in range(20):
        opt.zero_grad()
        loss = ((render(params, r) - target) ** 2).mean()
        loss.backward()
        opt.step()
    assert loss.item() < first
