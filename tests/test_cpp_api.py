"""The header-only C++ host API (include/gartus.hpp) over the C ABI.

A small program, tests/cpp/cornell_demo.cpp, builds the Cornell box
(scenes.rs:226-294) with RayScene / SamplingTargetList / RayCamera /
PathTracer exactly as a reference user would. The CPU test compiles it and
runs its host-only checks: camera defaults and heights, panics on invalid
input, the file headers. The GPU test renders through it (which also checks
the device encoders on known values) and requires the image to be
bit-identical to the same render made through the Python API, so both host
layers hand the library identical arguments.
"""
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

from gartus_amd import abi, api, scenes

ROOT = Path(__file__).resolve().parents[1]
LIBDIR = ROOT / "gartus_amd"


@pytest.fixture(scope="module")
def demo(tmp_path_factory):
    if not (LIBDIR / "libgartus_hip.so").exists():
        pytest.fail("libgartus_hip.so is not built")
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    exe = tmp_path_factory.mktemp("cpp") / "cornell_demo"
    cmd = [gxx, "-O2", "-std=c++17", "-Wall", "-Werror", f"-I{ROOT / 'include'}",
           str(ROOT / "tests" / "cpp" / "cornell_demo.cpp"), f"-L{LIBDIR}", "-lgartus_hip",
           f"-Wl,-rpath,{LIBDIR}", "-o", str(exe)]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    return exe


def _read_pfm(path: Path) -> np.ndarray:
    raw = path.read_bytes()
    parts = raw.split(b"\n", 3)
    assert parts[0] == b"PF"
    w, h = map(int, parts[1].split())
    assert float(parts[2]) < 0  # little-endian
    img = np.frombuffer(parts[3], "<f4").reshape(h, w, 3)
    return img[::-1]  # rows are stored bottom-up (display.rs:370)


def test_cpp_host_checks(demo):
    out = subprocess.run([str(demo), "host"], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stderr + out.stdout


@pytest.mark.gpu
def test_cpp_render_equals_python_render(demo, tmp_path, gpu_context):
    pfm = tmp_path / "cornell.pfm"
    out = subprocess.run([str(demo), "render", str(pfm)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr + out.stdout
    cpp = _read_pfm(pfm)

    data, lights, cam = scenes.cornell()
    cam.update(width=64, grid=2)
    rs = api.RayScene(data).with_bvh()
    L = api.SamplingTargetList()
    for kind, q, w in lights:
        L._add(kind, q, w)
    tracer = api.PathTracer(api.RayCamera.__new__(api.RayCamera))
    tracer.camera.raw = scenes.camera_record(cam)
    assert tracer.camera.raw.direct_lighting_mode == abi.GT_LIGHTING_NEE
    py = tracer.render_with_lights_hdr_image(rs, L).pixels
    assert cpp.shape == py.shape == (64, 64, 3)
    assert np.isfinite(cpp).all() and cpp.mean() > 0.05
    np.testing.assert_array_equal(cpp, py.astype(np.float32))
