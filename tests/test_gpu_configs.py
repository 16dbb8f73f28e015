"""The workload configurations at their full sizes on the device.

* The 4096x4096 full-filter document (2000 paths of 3-200 segments, 8 filter
  regions): each of the 8 filter groups, with a distinct slice of the plain paths,
  rendered at the full canvas from the real 2000-path stream and compared with the
  oracle bit for bit; and the whole unmodified document checked by properties that
  hold at any size (a render repeated is the same, a small flush budget draws the
  same pixels as the default one, rendering onto a canvas twice composites twice).
* The batched workload at its real sizes: documents of 1024-2048 px sides with
  100-200 paths, in one rvk_render_batch call, each compared with the oracle.
* The clip/mask document: its generator at 512x512 against the oracle, and the
  4096x4096 document by the repeat and flush-budget properties.
"""
import numpy as np
import pytest

from resvg_amd import abi
from resvg_amd.workloads import full_filter_doc

import oracle_pool
from test_gpu_render import _assert_same

pytestmark = pytest.mark.gpu

SIZE, PATHS, SEED = 4096, 2000, 0
STRIDE = 128  # every 128th plain path per part: 14-15 paths beside each filter group
PARTS = [(SIZE, PATHS, SEED, fg, STRIDE, fg) for fg in range(8)]


@pytest.fixture(scope="module")
def oracle_parts():
    print("\n[configs] oracle renders of the 8 full-filter parts (4096x4096) in a process pool", flush=True)
    return oracle_pool.oracle_full_filter_parts(PARTS)


@pytest.mark.timeout(900)
@pytest.mark.parametrize("part", range(8))
def test_full_filter_4096_part_matches_oracle(ctx, oracle_parts, part):
    import torch
    sc, (w, h) = oracle_pool.full_filter_part(PARTS[part])
    assert (w, h) == (SIZE, SIZE) and len(sc.groups) > 1
    px = torch.zeros((h, w, 4), dtype=torch.uint8, device="cuda")
    ctx.render(sc, abi.Transform.identity(), px)
    ctx.synchronize()
    got = px.cpu().numpy()
    exp = oracle_parts[part]
    _assert_same(got, exp, f"full-filter 4096 filter group {part} + plain slice")
    assert exp[..., 3].any()


@pytest.fixture(scope="module")
def whole_doc():
    sc, (w, h) = full_filter_doc(seed=SEED, size=SIZE, n_paths=PATHS)
    return sc.to_abi(), (w, h)


def _render(ctx, sc, w, h, times=1):
    import torch
    px = torch.zeros((h, w, 4), dtype=torch.uint8, device="cuda")
    for _ in range(times):
        ctx.render(sc, abi.Transform.identity(), px)
    ctx.synchronize()
    return px


@pytest.mark.timeout(300)
def test_full_filter_4096_whole_document_properties(ctx, monkeypatch, whole_doc):
    import torch
    sc, (w, h) = whole_doc
    a = _render(ctx, sc, w, h)
    st = ctx.stats()
    assert st.paths >= PATHS and st.layers >= PATHS
    b = _render(ctx, sc, w, h)
    assert torch.equal(a, b), "repeated renders differ"
    monkeypatch.setenv("RVK_FLUSH_BYTES", str(256 << 20))
    c = _render(ctx, sc, w, h)
    assert ctx.stats().flushes > st.flushes
    assert torch.equal(a, c), "a 256 MiB flush budget changes pixels"
    assert a[..., 3].float().mean() > 200  # 2000 overlapping paths cover the canvas
    # resvg::render composites and never clears: a second render over the first
    # changes the canvas exactly where the document is not yet opaque
    twice = _render(ctx, sc, w, h, times=2)
    opaque = a[..., 3] == 255
    assert torch.equal(twice[opaque][:, 3], a[opaque][:, 3])


BATCH_KW = dict(min_side=1024, max_side=2048, min_paths=100, max_paths=200)
BATCH_DOCS, BATCH_SEED = 6, 21


@pytest.mark.timeout(900)
def test_batched_documents_at_full_size(ctx):
    import torch
    docs = oracle_pool.batch_docs(BATCH_DOCS, BATCH_SEED, **BATCH_KW)
    assert max(max(w, h) for _, (w, h) in docs) > 1500 and max(len(sc.paths) for sc, _ in docs) >= 150
    print(f"\n[configs] oracle renders of {len(docs)} batched documents", flush=True)
    refs = oracle_pool.oracle_batch_docs(BATCH_DOCS, BATCH_SEED, list(range(len(docs))), **BATCH_KW)
    pxs = [torch.zeros((h, w, 4), dtype=torch.uint8, device="cuda") for _, (w, h) in docs]
    st = ctx.render_batch([sc for sc, _ in docs], abi.Transform.identity(), pxs)
    ctx.synchronize()
    assert st == [abi.OK] * len(docs)
    for i, (px, ref) in enumerate(zip(pxs, refs)):
        _assert_same(px.cpu().numpy(), ref, f"batched document {i} {docs[i][1]}")



def test_clip_mask_document_matches_oracle(ctx):
    """The clip/mask workload's generator at 512x512 (40 groups, 8 clip paths, 8
    masks): nested clip paths, clip-path'd clip children, nested luminance and alpha
    masks, bit-exact against the oracle."""
    from oracle import bindings as orc
    from resvg_amd.workloads import clip_mask_doc
    sc, (w, h) = clip_mask_doc(seed=3, size=512, n_groups=40, n_clips=8, n_masks=8)
    s = sc.to_abi()
    got = _render(ctx, s, w, h).cpu().numpy()
    exp = np.zeros((h, w, 4), np.uint8)
    orc.render(s.abi, abi.Transform.identity(), exp)
    assert (exp[..., 3] > 0).mean() > 0.3
    _assert_same(got, exp, "clip/mask 512")


def test_clip_mask_document_full_size_is_deterministic(ctx, monkeypatch):
    """The 4096x4096 clip/mask document (300 groups) renders the same pixels twice and
    with a 64 MiB flush budget."""
    from resvg_amd.workloads import clip_mask_doc
    sc, (w, h) = clip_mask_doc(seed=0)
    s = sc.to_abi()
    a = _render(ctx, s, w, h).cpu().numpy()
    assert (a[..., 3] > 0).mean() > 0.3
    assert np.array_equal(a, _render(ctx, s, w, h).cpu().numpy())
    monkeypatch.setenv("RVK_FLUSH_BYTES", str(64 << 20))
    assert np.array_equal(a, _render(ctx, s, w, h).cpu().numpy())
