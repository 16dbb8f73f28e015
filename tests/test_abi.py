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
