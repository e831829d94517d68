"""edxr_render_mesh_noalloc: the allocation-free frame against edxr_render_mesh.

The entry point renders a device-resident mesh into caller-owned device
buffers with every intermediate buffer carved from the library's static
device scratch.  Its output must equal edxr_render_mesh's bit for bit
(triangle IDs, depth, colour), on two calls in a row, the second with a
smaller mesh, so that the scratch's leftovers from the first call cannot
leak into the second.
"""
import ctypes as C

import numpy as np
import pytest

from edxraster_amd import configs, Renderer
from edxraster_amd import _lib as L

pytestmark = pytest.mark.gpu


def device_mesh(mesh):
    import torch
    planes = [torch.from_numpy(a.copy()).cuda() for a in mesh.planes()]
    idx = torch.from_numpy(mesh.indices.view(np.int32).copy()).cuda()
    ptr = lambda t: C.cast(C.c_void_p(t.data_ptr()), L.F)
    m = L.Mesh(*[ptr(t) for t in planes], C.cast(C.c_void_p(idx.data_ptr()), L.U32), mesh.num_vertices,
               mesh.num_triangles, L.MEM_DEVICE)
    return m, (planes, idx)


def noalloc_frame(r, w, dmesh):
    """(ids, depth, rgba) of one allocation-free render, bottom-up rows."""
    import torch
    color = torch.full((w.height, w.width), 0x5A5A5A5A, dtype=torch.int64).to(torch.int32).cuda()
    depth = torch.full((w.height, w.width), -7.0, dtype=torch.float32).cuda()
    ids = torch.full((w.height, w.width), 12345, dtype=torch.int32).cuda()
    torch.cuda.synchronize()
    rc = L.lib().edxr_render_mesh_noalloc(r.ctx, C.byref(dmesh), C.c_void_p(color.data_ptr()),
                                          C.c_void_p(depth.data_ptr()), C.c_void_p(ids.data_ptr()))
    assert rc == 0, L.lib().edxr_last_error(r.ctx).decode()
    rgba = color.cpu().numpy().view(np.uint8).reshape(w.height, w.width, 4)
    return ids.cpu().numpy().view(np.uint32), depth.cpu().numpy(), rgba


def plain_frame(r, w):
    r.render_mesh(w.mesh)
    return r.read_triangle_ids(), r.read_depth(), r.get_back_buffer()


def assert_equal_frames(a, b, what):
    for x, y, n in zip(a, b, ("ids", "depth", "colour")):
        assert np.array_equal(np.ascontiguousarray(x).view(np.uint8), np.ascontiguousarray(y).view(np.uint8)), \
            "%s: %s differ" % (what, n)


def test_noalloc_equals_render_mesh_twice_second_smaller(built):
    big = configs.small_sphere(800, 600)                       # 4,096 triangles
    small = configs.small_sphere(800, 600, slices=24, stacks=12)  # 576 triangles
    r = Renderer(800, 600)
    big.apply(r)
    outs = []
    for w in (big, small):
        dm, keep = device_mesh(w.mesh)
        got = noalloc_frame(r, w, dm)
        st = r.stats()
        assert st["triangles_in"] == w.mesh.num_triangles and st["retries"] == 0
        want = plain_frame(r, w)
        assert (want[0] != 0xFFFFFFFF).sum() > 1000
        assert_equal_frames(got, want, "%d triangles" % w.mesh.num_triangles)
        outs.append(got)
    assert not np.array_equal(outs[0][0], outs[1][0])
    r.close()


def test_noalloc_clipped_scene_and_shaders(built):
    """Clipping (clip queue, clip-created vertices and records in the scratch)
    and the three untextured shaders."""
    w = configs.sphere_grid_1080p(640, 360, grid=4, segments=24)
    r = Renderer(w.width, w.height)
    w.apply(r)
    dm, keep = device_mesh(w.mesh)
    for kind in (L.SHADER_VC_BLINN_PHONG, L.SHADER_LAMBERT_ALBEDO, L.SHADER_BLINN_PHONG_REF):
        w.shader = kind
        w.shininess = {0: 0.0, 1: 200.0, 2: 32.0}[kind]
        w.apply(r)
        got = noalloc_frame(r, w, dm)
        assert r.stats()["clipped"] > 0
        assert_equal_frames(got, plain_frame(r, w), "clipped scene, shader %d" % kind)
    r.close()


def test_noalloc_rejects_what_it_cannot_do(built):
    w = configs.small_sphere(64, 48)
    r = Renderer(w.width, w.height)
    w.apply(r)
    import torch
    lib = L.lib()
    out = torch.zeros(1920 * 1080, dtype=torch.int32).cuda()  # large enough for every call below
    buf = C.c_void_p(out.data_ptr())
    host = w.mesh.c_struct()
    assert lib.edxr_render_mesh_noalloc(r.ctx, C.byref(host), buf, buf, buf) == L.EDXR_E_INVALID
    assert b"device" in lib.edxr_last_error(r.ctx)
    dm, keep = device_mesh(w.mesh)
    assert lib.edxr_render_mesh_noalloc(r.ctx, C.byref(dm), None, buf, buf) == L.EDXR_E_INVALID
    r.set_msaa_mode(2)
    assert lib.edxr_render_mesh_noalloc(r.ctx, C.byref(dm), buf, buf, buf) == L.EDXR_E_STATE
    r.set_msaa_mode(0)
    # A scene larger than the scratch: a clear limit error, nothing launched.
    big = configs.sphere_grid_1080p(1920, 1080, grid=8, segments=64)   # 524,288 triangles at 1080p
    r.resize(1920, 1080)
    big.apply(r)
    bm, keep2 = device_mesh(big.mesh)
    assert lib.edxr_render_mesh_noalloc(r.ctx, C.byref(bm), buf, buf, buf) == L.EDXR_E_LIMIT
    assert b"scratch" in lib.edxr_last_error(r.ctx)
    r.close()
