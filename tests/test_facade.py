"""The C++ facade (include/edxraster.hpp): reference-style host code compiles
and links against libedxraster.so here, and on a GPU renders exactly what the
Python path renders."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "facade_demo.cpp")
VIEWER = os.path.join(ROOT, "tests", "cpp", "viewer_sequence.cpp")


def build_demo(tmp_path, src=SRC):
    exe = str(tmp_path / os.path.splitext(os.path.basename(src))[0])
    lib = os.path.join(ROOT, "edxraster_amd")
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), src,
                    "-o", exe, "-L", lib, "-ledxraster", "-Wl,-rpath," + lib], check=True)
    return exe


def test_facade_compiles_and_links(built, tmp_path):
    assert os.path.exists(build_demo(tmp_path))


def test_viewer_sequence_compiles_and_links(built, tmp_path):
    """The reference viewer's globals, OnInit, OnRender and Load Scene code
    (RealtimeViewer/Main.cpp:18-131 without window, GL and GUI) compile
    against the facade."""
    assert os.path.exists(build_demo(tmp_path, VIEWER))


def test_load_sphere_and_plane_declared_forms(built):
    """CPU: Mesh::LoadSphere / LoadPlane with placement (pos, scl, rot):
    vertex counts, texture coordinates, white colours, slot 0, the placement
    T(pos) Rz Ry Rx S(scl) and unit normals."""
    from edxraster_amd import Mesh
    s = Mesh.load_sphere((1.0, 2.0, 3.0), (2.0, 1.0, 1.0), (0.0, 90.0, 0.0), 1.5, 16, 8)
    assert (s.num_vertices, s.num_triangles) == (17 * 9, 2 * 16 * 8)
    assert s.tu is not None and s.tex_ids is not None and (s.tex_ids == 0).all()
    assert (s.cr == 1).all() and (s.cg == 1).all() and (s.cb == 1).all()
    np.testing.assert_allclose(np.hypot(np.hypot(s.nx, s.ny), s.nz), 1.0, rtol=1e-6)
    # the unplaced sphere, scaled by 2 in x, rotated 90 degrees about y, moved to (1, 2, 3)
    u = Mesh.load_sphere((0.0, 0.0, 0.0), (1.0, 1.0, 1.0), (0.0, 0.0, 0.0), 1.5, 16, 8)
    x, y, z = 2.0 * u.px.astype(np.float64), u.py.astype(np.float64), u.pz.astype(np.float64)
    # Ry(90): (x, y, z) -> (z, y, -x)
    np.testing.assert_allclose(s.px, z + 1.0, atol=1e-5)
    np.testing.assert_allclose(s.py, y + 2.0, atol=1e-5)
    np.testing.assert_allclose(s.pz, -x + 3.0, atol=1e-5)
    np.testing.assert_allclose(u.tu.reshape(9, 17)[3], np.arange(17) / 16.0, atol=1e-7)
    np.testing.assert_allclose(u.tv.reshape(9, 17)[:, 5], np.arange(9) / 8.0, atol=1e-7)
    lo, hi = u.get_bounds()
    np.testing.assert_allclose(hi, [1.5, 1.5, 1.5], atol=1e-6)
    p = Mesh.load_plane((0.0, 0.0, 0.0), (1.0, 1.0, 1.0), (-90.0, 0.0, 0.0), 1.2)
    assert (p.num_vertices, p.num_triangles) == (4, 2)
    # Rx(-90) turns the +y normal to -z (toward an eye on the -z axis)
    np.testing.assert_allclose(np.stack([p.nx, p.ny, p.nz], 1), [[0, 0, -1]] * 4, atol=1e-7)
    np.testing.assert_allclose(np.abs(p.px), 0.6, atol=1e-7)
    np.testing.assert_allclose(np.abs(p.py), 0.6, atol=1e-7)


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


def viewer_python_path(w=1280, h=720):
    """The viewer's first frame through the Python mirror: LoadSphere(0, 1, 0,
    1.2), camera at -5 z looking at the origin, fov 65, near 0.01, far 1000,
    the reference's Lambertian shader with the sphere's 0.9-white slot."""
    from edxraster_amd import Mesh, Renderer, look_at, perspective, raster_matrix
    from edxraster_amd import _lib as L
    m = Mesh.load_sphere((0, 0, 0), (1, 1, 1), (0, 0, 0), 1.2)
    mv = look_at((0.0, 0.0, -5.0), (0.0, 0.0, 0.0))
    p = perspective(65.0, w / h, 0.01, 1000.0)
    r = raster_matrix(w, h)
    return m, mv, p, r


@pytest.mark.gpu
def test_viewer_sequence_matches_python_path_and_oracle(built, tmp_path):
    from edxraster_amd import configs, Renderer, Mesh
    from edxraster_amd import _lib as L
    from oracle.oracle import Oracle
    from tests.test_io import write_textured_obj
    exe = build_demo(tmp_path, VIEWER)
    # a textured wall 100 units wide, 200 in front of the origin (2 after the
    # viewer's 0.01 scale), facing the camera at the origin
    wall = lambda P: np.stack([100.0 * P[:, 0], 100.0 * P[:, 2], 200.0 - 100.0 * P[:, 1]], 1)
    obj = write_textured_obj(tmp_path, wall)
    out1, out2 = tmp_path / "sphere.rgba", tmp_path / "scene.rgba"
    res = subprocess.run([exe, str(out1), obj, str(out2), "1"], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    assert "Triangle Count: %d" % (2 * 64 * 64) in res.stdout
    got = np.fromfile(out1, np.uint8).reshape(720, 1280, 4)
    m, mv, p, r = viewer_python_path()
    w = configs.Workload("viewer_sphere", 1280, 720, m, mv, p, r, shader=L.SHADER_TEXTURED_LAMBERT,
                         shininess=0.0, textures=[(0.9, 0.9, 0.9)])
    rr = Renderer(1280, 720)
    w.apply(rr)
    rr.render_mesh(m)
    np.testing.assert_array_equal(got, rr.get_back_buffer())
    assert (got[:, :, 3] == 255).sum() > 50000  # the sphere covers about 57,000 pixels
    o = Oracle()
    o.render_workload(w)
    assert np.abs(got.astype(int) - o.color().astype(int)).max() <= 1
    # The Load Scene branch: LoadMesh at 0.01 scale, camera fitted to the
    # bounds, bilinear filtering chosen through SetTextureFilter.
    got2 = np.fromfile(out2, np.uint8).reshape(720, 1280, 4)
    m2 = Mesh.load_obj(obj, (0, 0, 0), (0.01, 0.01, 0.01), (0, 0, 0))
    rad = np.float32(float.fromhex(res.stdout.split("scene radius ")[1].split()[0]))
    lo, hi = m2.get_bounds()
    np.testing.assert_allclose(rad, np.linalg.norm(hi - 0.5 * (lo + hi)), rtol=1e-6)
    from edxraster_amd import look_at, perspective, raster_matrix
    w2 = configs.Workload("viewer_obj", 1280, 720, m2, look_at((0, 0, 0), (0, 0, 1)),
                          perspective(65.0, 1280 / 720, float(rad * np.float32(0.02)), float(rad * np.float32(5.0))),
                          raster_matrix(1280, 720), shader=L.SHADER_TEXTURED_LAMBERT, shininess=0.0,
                          textures=Mesh.obj_texture_slots(obj), tex_filter=L.FILTER_LINEAR)
    w2.apply(rr)
    rr.render_mesh(m2)
    np.testing.assert_array_equal(got2, rr.get_back_buffer())
    assert (got2[:, :, 3] == 255).sum() > 200000
    rr.close()
