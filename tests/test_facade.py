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
