"""The C-ABI library loads and exports every entry point include/dmsh.h declares
(no compute calls: this runs without a GPU)."""

import ctypes
import pathlib
import re

ROOT = pathlib.Path(__file__).resolve().parent.parent


def declared_functions():
    text = (ROOT / "include" / "dmsh.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(dmsh_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared_functions()
    for must in ["dmsh_vtx2morton_u32", "dmsh_argsort_u32", "dmsh_bvhnodes_from_sorted_mortons_u32",
                 "dmsh_bvhnode2aabb_trimesh3", "dmsh_build_trimesh3", "dmsh_pix2tri",
                 "dmsh_pix2depth", "dmsh_rasterized_edge_gradient_bwd"]:
        assert must in names


def test_library_exports_every_declared_symbol():
    from del_msh_dlpack import _lib

    path = _lib.library_path()
    assert path.exists(), f"{path} not built"
    handle = ctypes.CDLL(str(path))
    missing = [n for n in declared_functions() if not hasattr(handle, n)]
    assert not missing, missing


def test_python_prototypes_cover_header():
    from del_msh_dlpack import _lib

    assert set(declared_functions()) <= set(_lib._PROTOTYPES), \
        set(declared_functions()) - set(_lib._PROTOTYPES)


def test_host_side_validation_without_gpu():
    """Size queries and argument validation run on the host only."""
    from del_msh_dlpack import _lib

    L = _lib.lib()
    assert L.dmsh_abi_version() == 1
    assert L.dmsh_status_string(3) == b"workspace too small"
    assert L.dmsh_argsort_workspace_bytes(1000, 4) > 8 * 1000
    assert L.dmsh_build_workspace_bytes(1 << 20, 63) > L.dmsh_build_workspace_bytes(1 << 20, 30)
    assert L.dmsh_pix2tri_workspace_bytes(1000, 64, 64) >= 4 * 64 * 64 + 64 * 999 + 48 * 1000
    assert L.dmsh_set_tuning(b"pix2tri.stack", 24) == 0
    assert L.dmsh_get_tuning(b"pix2tri.stack") == 24
    assert L.dmsh_set_tuning(b"pix2tri.stack", 7) == 1
    assert L.dmsh_get_tuning(b"no.such.key") == -1
    # invalid arguments are rejected before anything is enqueued
    assert L.dmsh_vtx2morton_u32(None, 10, None, 3, None, None) == 1
    assert L.dmsh_build_trimesh3(None, 10, None, 10, None, 31, None, None, None, 0) == 1
    assert L.dmsh_argsort_u32(None, 10, 1, 1, 32, None, 0) == 3
