"""The C++ facade (include/edxraster.hpp): reference-style host code compiles
and links against libedxraster.so here, and on a GPU renders exactly what the
Python path renders."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "facade_demo.cpp")


def build_demo(tmp_path):
    exe = str(tmp_path / "facade_demo")
    lib = os.path.join(ROOT, "edxraster_amd")
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"), SRC, "-o", exe,
                    "-L", lib, "-ledxraster", "-Wl,-rpath," + lib], check=True)
    return exe


def test_facade_compiles_and_links(built, tmp_path):
    assert os.path.exists(build_demo(tmp_path))


@pytest.mark.gpu
def test_facade_matches_python_path(built, tmp_path):
    from edxraster_amd import configs, Renderer
    exe = build_demo(tmp_path)
    out = tmp_path / "bb.rgba"
    r = subprocess.run([exe, str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    got = np.fromfile(out, np.uint8).reshape(600, 800, 4)
    w = configs.small_sphere()
    rr = Renderer(w.width, w.height)
    w.apply(rr)
    rr.render_mesh(w.mesh)
    np.testing.assert_array_equal(got, rr.get_back_buffer())
    rr.close()


@pytest.mark.gpu
def test_facade_loads_textured_obj_like_python_path(built, tmp_path):
    """Mesh::LoadMesh with its material slots and the textured shader through
    the facade equals the Python path's render of the same OBJ."""
    from edxraster_amd import configs, Renderer, Mesh
    from edxraster_amd import _lib as L
    from tests.test_io import write_textured_obj, TEXTURED_OBJ_CAMERA
    exe = build_demo(tmp_path)
    path = write_textured_obj(tmp_path)
    out = tmp_path / "obj.rgba"
    r = subprocess.run([exe, str(out), path], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    got = np.fromfile(out, np.uint8).reshape(240, 320, 4)
    m = Mesh.load_obj(path)
    w = configs.Workload("obj_textured", 320, 240, m, *configs.camera(320, 240, *TEXTURED_OBJ_CAMERA),
                         shader=L.SHADER_TEXTURED_LAMBERT, shininess=0.0, textures=Mesh.obj_texture_slots(path))
    rr = Renderer(w.width, w.height)
    w.apply(rr)
    rr.render_mesh(w.mesh)
    np.testing.assert_array_equal(got, rr.get_back_buffer())
    assert (got[:, :, :3] > 0).any()
    rr.close()
