"""The #[repr(C)] mirror of the scene structs that INTEGRATION.md gives a Rust binding
is laid out like the C header: every struct's size and every field's offset, computed
from the Rust declarations with repr(C) rules, equal the ctypes mirror's (which
tests/test_abi.py checks against a C compiler)."""
import ctypes as C
import re
from pathlib import Path

import pytest

from resvg_amd import abi

DOC = Path(__file__).resolve().parents[1] / "INTEGRATION.md"
PRIM = {"u8": (1, 1), "u32": (4, 4), "i32": (4, 4), "f32": (4, 4), "u64": (8, 8)}


def _rust_items():
    text = DOC.read_text()
    code = "\n".join(re.findall(r"```rust\n(.*?)```", text, re.S))
    code = re.sub(r"//[^\n]*", "", code)
    items = {}
    for kind, name, body in re.findall(r"pub (struct|union) (\w+)\s*\{(.*?)\}", code, re.S):
        fields = re.findall(r"pub (\w+):\s*([^,]+?)\s*(?:,|$)", body.strip() + ",")
        items[name] = (kind, [(f, t.strip()) for f, t in fields])
    return items


def _layout(items, t, memo):
    """(size, align) of Rust type `t` under repr(C)."""
    t = t.strip()
    m = re.fullmatch(r"\[(.+);\s*(\d+)\]", t)
    if m:
        s, a = _layout(items, m.group(1), memo)
        return s * int(m.group(2)), a
    if t.startswith("*const") or t.startswith("*mut"):
        return 8, 8
    if t in PRIM:
        return PRIM[t]
    if t in memo:
        return memo[t][:2]
    kind, fields = items[t]
    off, align, offsets = 0, 1, {}
    size = 0
    for f, ft in fields:
        s, a = _layout(items, ft, memo)
        align = max(align, a)
        if kind == "union":
            offsets[f] = 0
            size = max(size, s)
        else:
            off = (off + a - 1) // a * a
            offsets[f] = off
            off += s
            size = off
    size = (size + align - 1) // align * align
    memo[t] = (size, align, offsets)
    return size, align


MIRRORS = {
    "rvk_transform": abi.Transform, "rvk_rect": abi.Rect, "rvk_stop": abi.Stop, "rvk_paint": abi.Paint,
    "rvk_path": abi.Path, "rvk_group": abi.Group, "rvk_node_ref": abi.NodeRef, "rvk_image": abi.Image,
    "rvk_clip_path": abi.ClipPath, "rvk_mask": abi.Mask, "rvk_light_source": abi.LightSource,
    "rvk_primitive": abi.Primitive, "rvk_filter": abi.Filter, "rvk_filter_data": abi.FilterData,
    "rvk_scene": abi.Scene,
}


@pytest.mark.parametrize("name", sorted(MIRRORS))
def test_rust_mirror_matches_the_c_layout(name):
    items = _rust_items()
    assert name in items, f"{name} missing from INTEGRATION.md"
    memo = {}
    size, _ = _layout(items, name, memo)
    ct = MIRRORS[name]
    assert size == C.sizeof(ct), (name, size, C.sizeof(ct))
    offsets = memo[name][2]
    for fname, *_ in ct._fields_:
        assert fname in offsets, f"{name}.{fname} missing from the Rust mirror"
        assert offsets[fname] == getattr(ct, fname).offset, (name, fname, offsets[fname], getattr(ct, fname).offset)
    assert len(offsets) == len(ct._fields_), (name, sorted(offsets), [f for f, *_ in ct._fields_])
