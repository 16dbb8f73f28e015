"""rvk_box_blur_noalloc (the entry point that allocates nothing) against rvk_box_blur,
through the C-ABI: bit-exact on two calls in a row, the second on a smaller image, and
the size limit rejected before any work."""
import numpy as np
import pytest

from resvg_amd import abi

pytestmark = pytest.mark.gpu


def _premul(h, w, seed):
    rng = np.random.default_rng(seed)
    a = rng.integers(0, 256, (h, w, 1), dtype=np.uint8)
    return np.concatenate([np.minimum(rng.integers(0, 256, (h, w, 3), dtype=np.uint8), a), a], -1)


@pytest.mark.parametrize("cases", [
    [((768, 1024), (6.0, 6.0)), ((200, 300), (2.5, 4.0))],
    [((512, 640), (40.0, 3.0)), ((96, 64), (24.0, 0.0))],
])
def test_noalloc_box_blur_matches_box_blur(ctx, cases):
    import torch
    for i, ((h, w), (sx, sy)) in enumerate(cases):
        img = _premul(h, w, 100 + i)
        ref = torch.from_numpy(img.copy()).cuda()
        ctx.box_blur(ref, sx, sy)
        got = torch.from_numpy(img.copy()).cuda()
        assert ctx.box_blur_noalloc(got, sx, sy) == abi.OK
        ctx.synchronize()
        assert torch.equal(got, ref), f"call {i}: {w}x{h} sigma ({sx}, {sy})"
        assert not torch.equal(got, torch.from_numpy(img).cuda()), "the blur changed nothing"


def test_noalloc_box_blur_rejects_images_beyond_its_array(ctx):
    import torch
    t = torch.zeros((1, 1, 4), dtype=torch.uint8, device="cuda")
    # 4097 x 4096 x 4 bytes > 64 MiB: refused on the host before anything is launched
    from resvg_amd import lib
    assert lib().rvk_box_blur_noalloc(ctx.h, t.data_ptr(), 4097, 4096, 2.0, 2.0) == abi.ERR_TOO_LARGE
    assert lib().rvk_status_string(abi.ERR_TOO_LARGE).startswith(b"input larger")
