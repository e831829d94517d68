"""C-ABI checks that need no GPU: the library loads, exports every function
declared in include/edxraster.h, host-side helpers work, and the product
package never reaches the oracle."""
import ctypes as C
import os
import re

import numpy as np

from edxraster_amd import _lib as L
from edxraster_amd import Mesh, look_at, perspective, raster_matrix, default_material

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "edxraster.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(edxr_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol(built):
    lib = L.lib()
    names = declared_functions()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # and the ctypes signature table covers the same set
    assert sorted(n for n, _, _ in L.SIGNATURES) == names


def test_create_fails_cleanly_without_gpu(built):
    import torch
    if torch.cuda.is_available():
        return
    ctx = C.c_void_p()
    assert L.lib().edxr_create(C.byref(ctx), 64, 64) == L.EDXR_E_HIP
    assert not ctx.value


def test_null_context_errors(built):
    lib = L.lib()
    assert lib.edxr_resize(None, 64, 64) == L.EDXR_E_INVALID
    assert lib.edxr_render_mesh(None, None) == L.EDXR_E_INVALID
    assert lib.edxr_last_error(None) == b"null context"


def test_default_materials(built):
    m = default_material(L.SHADER_BLINN_PHONG_REF)
    assert (m.ambient, m.kd, m.ks, m.shininess) == (np.float32(0.2), 3.0, 3.0, 200.0)
    m = default_material(L.SHADER_LAMBERT_ALBEDO)
    assert tuple(m.albedo) == (np.float32(0.9),) * 3


def test_camera_matrices(built):
    mv = look_at((0, 0, -3), (0, 0, 0))
    np.testing.assert_allclose(mv @ np.array([0, 0, 0, 1]), [0, 0, 3, 1], atol=1e-6)
    p = perspective(60.0, 4 / 3, 0.1, 100.0)
    for z, expect in ((0.1, 0.0), (100.0, 1.0)):
        c = p @ np.array([0, 0, z, 1.0])
        assert abs(c[2] / c[3] - expect) < 1e-5
    r = raster_matrix(800, 600)
    np.testing.assert_allclose(r @ np.array([-1, 1, 0, 1.0]), [0, 0, 0, 1])
    np.testing.assert_allclose(r @ np.array([1, -1, 0, 1.0]), [800, 600, 0, 1])


def test_scenes_are_deterministic(built):
    a = Mesh.load_heightfield(33, 0.2, 3, 33)
    b = Mesh.load_heightfield(33, 0.2, 3, 33)
    for x, y in zip(a.planes() + [a.indices], b.planes() + [b.indices]):
        assert np.array_equal(x, y)
    c = Mesh.load_heightfield(33, 0.2, 4, 33)
    assert not np.array_equal(a.py, c.py)
    assert np.abs(a.py).max() <= 0.2 + 1e-6
    n = np.sqrt(a.nx ** 2 + a.ny ** 2 + a.nz ** 2)
    np.testing.assert_allclose(n, 1.0, atol=1e-5)
    assert a.cr.min() >= 0 and a.cr.max() < 1


def test_product_never_uses_the_oracle():
    """Only tests/, __graft_entry__.smoke() and bench.py's CPU baseline may
    reach the oracle; the product package must not."""
    pkg = os.path.join(ROOT, "edxraster_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".cpp", ".hip", ".h")):
                text = open(os.path.join(dirpath, f)).read()
                assert "oracle" not in text.lower(), os.path.join(dirpath, f)
