"""The product's host geometry (raster_host.cpp, edge_clip.cpp, stroker.cpp, dash.cpp,
hairline.cpp) built with AddressSanitizer and UndefinedBehaviorSanitizer (g++, host
code only) and run over seeded random scenes by tests/native/geom_sanitize.cpp, which
also checks every prepared coverage and hairline job against the invariants the
device kernels index by.  render.cpp (the walk and the flush packing) needs the HIP
runtime; its packed blocks are checked by the RVK_VALIDATE index validator in the
GPU suite instead."""
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
CSRC = ROOT / "resvg_amd" / "csrc"
SOURCES = ["raster_host.cpp", "edge_clip.cpp", "stroker.cpp", "dash.cpp", "hairline.cpp"]


@pytest.mark.skipif(shutil.which("g++") is None or not Path("/opt/rocm/include").is_dir(),
                    reason="needs g++ and the ROCm headers")
def test_host_geometry_under_asan_and_ubsan(tmp_path):
    exe = tmp_path / "geom_sanitize"
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
           "-ffp-contract=off", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
           str(ROOT / "tests" / "native" / "geom_sanitize.cpp"), *(str(CSRC / s) for s in SOURCES), "-o", str(exe)]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-4000:]
    r = subprocess.run([str(exe), "150"], capture_output=True, text=True, timeout=600,
                       env={"ASAN_OPTIONS": "detect_leaks=1:abort_on_error=0", "PATH": "/usr/bin:/bin"})
    assert r.returncode == 0 and r.stdout.startswith("ok "), (r.stdout[-2000:], r.stderr[-4000:])
    jobs, hair = (int(v) for v in r.stdout.split()[1:3])
    assert jobs > 500 and hair > 20
