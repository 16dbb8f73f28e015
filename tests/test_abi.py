"""The C-ABI library loads here (no GPU) and exports every entry point the header declares."""
import ctypes as C
import re
from pathlib import Path

import resvg_amd
from resvg_amd import abi

HDR = Path(__file__).resolve().parent.parent / "include" / "resvg_amd.h"


def declared():
    return sorted(set(re.findall(r"\b(rvk_[a-z0-9_]+)\s*\(", HDR.read_text())))


def test_library_exports_every_declared_symbol():
    L = resvg_amd.lib()
    missing = [n for n in declared() if not hasattr(L, n)]
    assert not missing, missing


def test_status_strings_and_no_device_status():
    L = resvg_amd.lib()
    assert L.rvk_status_string(abi.OK) == b"ok"
    assert L.rvk_status_string(abi.ERR_UNSUPPORTED) == b"unsupported feature"


def test_invalid_arguments_are_rejected_without_a_device():
    L = resvg_amd.lib()
    # a null context is rejected before any HIP call
    assert L.rvk_box_blur(None, None, 1, 1, 2.0, 2.0) == abi.ERR_INVALID_ARGUMENT
    assert L.rvk_filter_apply(None, None, None, abi.Transform.identity(), None, 1, 1, None) == \
        abi.ERR_INVALID_ARGUMENT


def test_struct_sizes_match_header_layout():
    # rvk_primitive: 4+4+16+12 bytes of header + 192-byte union
    assert C.sizeof(abi.Primitive) == 4 + 4 + 16 + 4 * 3 + 192
    assert C.sizeof(abi.Transform) == 24
    assert C.sizeof(abi.LightSource) == 48


# ---------------------------------------------------------------------------
# Layout of every rvk_* structure, as a C compiler lays it out from the header,
# against the ctypes mirror (the layout an FFI crate's #[repr(C)] mirror must match)
# ---------------------------------------------------------------------------
def _layout_structs():
    from resvg_amd.diff import DiffScene
    return {
        "rvk_transform": abi.Transform, "rvk_rect": abi.Rect, "rvk_transfer_fn": abi.TransferFn,
        "rvk_light_source": abi.LightSource, "rvk_primitive": abi.Primitive, "rvk_filter": abi.Filter,
        "rvk_filter_data": abi.FilterData, "rvk_stop": abi.Stop, "rvk_paint": abi.Paint, "rvk_path": abi.Path,
        "rvk_group": abi.Group, "rvk_node_ref": abi.NodeRef, "rvk_clip_path": abi.ClipPath, "rvk_mask": abi.Mask,
        "rvk_scene": abi.Scene, "rvk_path_geometry": abi.PathGeometry, "rvk_stroke_style": abi.StrokeStyle,
        "rvk_stats": abi.Stats, "rvk_kernel_stat": abi.KernelStat, "rvk_diff_scene": DiffScene,
    }


def _members(ctype, prefix=""):
    """(designator, ctypes offset, ctypes size) of every field, nested structs and unions included."""
    out = []
    for f in ctype._fields_:
        name, ft = f[0], f[1]
        desc = getattr(ctype, name)
        out.append((prefix + name, desc.offset, C.sizeof(ft)))
        if isinstance(ft, type) and issubclass(ft, (C.Structure, C.Union)):
            out += [(d, desc.offset + o, sz) for d, o, sz in _members(ft, prefix + name + ".")]
    return out


def test_struct_layouts_match_a_c_compiler(tmp_path):
    import shutil
    import subprocess
    cc = shutil.which("gcc") or shutil.which("cc")
    assert cc, "a C compiler is needed to check the header's layouts"
    structs = _layout_structs()
    lines = ["#include <stdio.h>", "#include <stddef.h>", f'#include "{HDR}"', "int main(void) {"]
    expect = {}
    for cname, ct in structs.items():
        lines.append(f'printf("%s size %zu\\n", "{cname}", sizeof({cname}));')
        expect[(cname, "size")] = C.sizeof(ct)
        for d, off, sz in _members(ct):
            if d.endswith("pad_"):
                continue
            lines.append(f'printf("%s %s %zu %zu\\n", "{cname}", "{d}", offsetof({cname}, {d}), '
                         f'sizeof((({cname} *)0)->{d}));')
            expect[(cname, d)] = (off, sz)
    lines += ["return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    r = subprocess.run([cc, "-std=c11", "-Wall", "-o", str(exe), str(src)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    seen, bad = 0, []
    for ln in filter(None, out):
        parts = ln.split()
        if parts[1] == "size":
            got, exp = int(parts[2]), expect[(parts[0], "size")]
        else:
            got, exp = (int(parts[2]), int(parts[3])), expect[(parts[0], parts[1])]
        seen += 1
        if got != exp:
            bad.append((parts[0], parts[1], got, exp))
    assert not bad, bad
    assert seen == len(expect) and seen > 200
