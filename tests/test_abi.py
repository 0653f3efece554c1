"""The C ABI without a GPU: the library loads, exports every function that
include/sspt.h declares, and the ctypes mirrors in seesharp_amd.abi have the
header's struct layouts (compiled from the header itself with gcc)."""
import ctypes as C
import os
import re
import subprocess

import pytest

from seesharp_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "sspt.h")

STRUCTS = {
    "sspt_mesh_desc": abi.MeshDesc,
    "sspt_material_desc": abi.MaterialDesc,
    "sspt_texture_desc": abi.TextureDesc,
    "sspt_emitter_desc": abi.EmitterDesc,
    "sspt_envmap_desc": abi.EnvmapDesc,
    "sspt_camera_desc": abi.CameraDesc,
    "sspt_scene_desc": abi.SceneDesc,
    "sspt_pt_settings": abi.PtSettings,
    "sspt_ray_stats": abi.RayStats,
    "sspt_hit": abi.Hit,
}


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(sspt_\w+)\s*\(", text)
    return sorted(set(names))


def test_header_declares_the_render_path():
    names = declared_functions()
    for n in ("sspt_ctx_create", "sspt_scene_create", "sspt_render", "sspt_framebuffer_read",
              "sspt_trace_closest", "sspt_trace_occluded", "sspt_render_batch", "sspt_render_adjoint"):
        assert n in names


def test_library_exports_every_declared_function():
    lib = abi.load_library()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_ctypes_table_matches_header():
    assert sorted(abi.SYMBOLS) == declared_functions()


def test_abi_version_without_device():
    lib = abi.load_library()
    declared = int(re.search(r"#define SSPT_ABI_VERSION (\d+)", open(HEADER).read()).group(1))
    assert abi.ABI_VERSION == declared
    assert lib.sspt_abi_version() == declared


def test_null_context_is_rejected_without_device():
    lib = abi.load_library()
    settings = abi.PtSettings()
    assert lib.sspt_render(None, None, C.byref(settings), 0, 1, None) == abi.INVALID_ARGUMENT
    assert lib.sspt_set_launch_config(None, 0, 0, 1) == abi.INVALID_ARGUMENT


def test_struct_layouts_match_header(tmp_path):
    lines = ["#include <stddef.h>", "#include <stdio.h>", f'#include "{HEADER}"', "int main(void) {"]
    for cname, py in STRUCTS.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    r = subprocess.run(["gcc", "-std=c11", "-o", str(exe), str(src)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    got = {}
    for line in out:
        if line:
            cname, field, value = line.split()
            got[(cname, field)] = int(value)
    for cname, py in STRUCTS.items():
        assert got[(cname, "size")] == C.sizeof(py), cname
        for fname, _ in py._fields_:
            assert got[(cname, fname)] == getattr(py, fname).offset, (cname, fname)
