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
    assert C.sizeof(abi.GtCamera) == 4 + 4 + 8 + 4 * 4 + 8 + 8 + 72 + 8 * 4 + 4 + 4 + 72 + 8 + 4 + 4
    assert C.sizeof(abi.GtTraversalStats) == 56
