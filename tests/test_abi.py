"""The C-ABI library loads and exports every entry point include/*.h
declares, and the Python ctypes mirror declares the same set (no device
calls). CPU only."""
import ctypes as C
import re
from pathlib import Path

from gartus_amd import abi

ROOT = Path(__file__).resolve().parent.parent


def declared_functions():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\**\s+(gt_[a-z_0-9]+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return names


def test_header_declares_the_boundary():
    names = declared_functions()
    for must in ("gt_render_hdr", "gt_render_aovs", "gt_intersect", "gt_scene_build", "gt_env_create_rgb32f",
                 "gt_lights_add_quad", "gt_scene_emissive_targets", "gt_render_batch"):
        assert must in names


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(str(abi.LIB_PATH))
    missing = [n for n in sorted(declared_functions()) if not hasattr(lib, n)]
    assert not missing, missing


def test_python_mirror_covers_the_header():
    assert declared_functions() == set(abi.PROTOTYPES)


def test_struct_layouts_match_header_sizes():
    # sizes of the plain-data records shared across the boundary
    assert C.sizeof(abi.GtRay) == 32
    assert C.sizeof(abi.GtHit) == 32
    assert C.sizeof(abi.GtTraversalStats) == 56


STRUCTS = {"gt_camera": "GtCamera", "gt_tuning": "GtTuning", "gt_render_stats": "GtRenderStats",
           "gt_ray": "GtRay", "gt_hit": "GtHit", "gt_traversal_stats": "GtTraversalStats",
           "gt_scene_info": "GtSceneInfo", "gt_tone_map": "GtToneMap", "gt_denoiser_config": "GtDenoiserConfig"}


def test_struct_field_offsets_match_the_c_compiler(tmp_path):
    """Every field offset and struct size of the ctypes mirror equals what the
    C compiler lays out from the headers."""
    import shutil
    import subprocess
    gcc = shutil.which("gcc")
    if gcc is None:
        import pytest
        pytest.skip("gcc not available")
    lines = ['#include <stddef.h>', '#include <stdio.h>', '#include "gartus_hip.h"', '#include "gartus_denoise.h"',
             'int main(void) {']
    for cname, pyname in STRUCTS.items():
        st = getattr(abi, pyname)
        lines.append(f'  printf("{cname} size %zu\\n", sizeof({cname}));')
        for field, _ in st._fields_:
            lines.append(f'  printf("{cname}.{field} %zu\\n", offsetof({cname}, {field}));')
    lines += ['  return 0;', '}']
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run([gcc, "-std=c99", f"-I{ROOT / 'include'}", str(src), "-o", str(exe)], check=True,
                   capture_output=True, text=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    want = dict(line.rsplit(" ", 1) for line in out if line)
    for cname, pyname in STRUCTS.items():
        st = getattr(abi, pyname)
        assert int(want[f"{cname} size"]) == C.sizeof(st), cname
        for field, _ in st._fields_:
            assert int(want[f"{cname}.{field}"]) == getattr(st, field).offset, f"{cname}.{field}"
