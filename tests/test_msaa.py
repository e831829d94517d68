"""GPU parity of multisampling (Renderer::SetMSAAMode, Core/Renderer.cpp:94-98)
against the oracle's per-sample restatement (Core/Rasterizer.h:202-300,355-415,
Core/FrameBuffer.cpp:54-87).

Bar: per-sample triangle IDs bit-exact, per-sample depth within 1 ulp (exact
is expected), resolved 8-bit colour within 1 LSB.
"""
import numpy as np
import pytest

from edxraster_amd import configs, Renderer
from edxraster_amd import _lib as L

from tests.test_gpu_parity import ulp_diff, assert_frame_parity

pytestmark = pytest.mark.gpu


def render_ms(w, level, hierarchical=True, renderer=None):
    from oracle.oracle import Oracle
    r = renderer or Renderer(w.width, w.height)
    r.set_msaa_mode(level)
    w.apply(r)
    r.set_hierarchical_rasterize(hierarchical)
    r.render_mesh(w.mesh)
    o = Oracle()
    o.render_workload(w, hierarchical=1 if hierarchical else 0, threads=16, msaa_log2=level)
    return r, o


def assert_ms_parity(r, o, name):
    d_g, ids_g = r.read_sample_buffers()
    ids_o, d_o = o.sample_ids(), o.sample_depth()
    assert ids_g.shape == ids_o.shape
    mism = np.count_nonzero(ids_g != ids_o)
    assert mism == 0, "%s: %d per-sample triangle-ID mismatches" % (name, mism)
    du = ulp_diff(d_g, d_o)
    assert du.max() <= 1, "%s: per-sample depth differs by %d ulp" % (name, du.max())
    c_g, c_o = r.get_back_buffer(), o.color()
    dc = np.abs(c_g.astype(np.int16) - c_o.astype(np.int16))
    assert dc.max() <= 1, "%s: resolved colour differs by %d LSB at %d pixels" % (name, dc.max(),
                                                                                  np.count_nonzero(dc))
    # sample 0 is what the single-sample readers return
    np.testing.assert_array_equal(r.read_triangle_ids(), ids_g[0])
    # pixels whose samples have different winners, so the resolve mixes colours
    return int(np.count_nonzero((ids_o != ids_o[:1]).any(axis=0)))


@pytest.mark.parametrize("level", [1, 2, 3, 4, 5])
def test_msaa_levels_parity(built, level):
    for name, w in (("sphere", configs.small_sphere(320, 240, 48, 24)),
                    ("heightfield", configs.heightfield_small(64, 320, 240))):
        r, o = render_ms(w, level)
        edge = assert_ms_parity(r, o, "%s level %d" % (name, level))
        assert edge > 0
        r.close()


def test_msaa_hierarchical_off_parity(built):
    w = configs.heightfield_small(64, 320, 240)
    r, o = render_ms(w, 2, hierarchical=False)
    assert_ms_parity(r, o, "hierarchical-off 4x")
    r.close()


def test_msaa_clipped_scene_parity(built):
    """Near-plane clipped pieces (stored records) under 8x multisampling."""
    w = configs.sphere_grid_1080p(640, 360, grid=4, segments=24)
    r, o = render_ms(w, 3)
    assert_ms_parity(r, o, "clipped 8x")
    assert r.stats()["clipped"] > 0
    r.close()


def test_msaa_hd_frame_parity(built):
    w = configs.heightfield_small(200, 1280, 720, seed=11)
    r, o = render_ms(w, 2)
    assert_ms_parity(r, o, "hd 4x")
    r.close()


def test_msaa_full_size_parity(built):
    """The 2,097,152-triangle sphere grid at 1920x1080 with 4 samples per pixel."""
    w = configs.sphere_grid_1080p()
    r, o = render_ms(w, 2)
    assert_ms_parity(r, o, "grid1m 4x")
    r.close()


def test_msaa_mode_switch_and_graph(built):
    """SetMSAAMode resizes the buffers; the captured graph follows the mode,
    and returning to one sample gives the single-sample frame again."""
    w = configs.heightfield_small(64, 320, 240)
    r = Renderer(w.width, w.height)
    r.set_graph(True)
    for level in (2, 0, 3, 3):
        _, o = render_ms(w, level, renderer=r)
        if level:
            assert_ms_parity(r, o, "graph level %d" % level)
        else:
            assert_frame_parity(r, o, "graph level 0")
    r.close()


def test_msaa_rejects_bad_level_and_backward(built):
    w = configs.heightfield_small(64, 160, 120)
    r = Renderer(w.width, w.height)
    with pytest.raises(RuntimeError):
        r.set_msaa_mode(6)
    with pytest.raises(RuntimeError):
        r.set_msaa_mode(-1)
    r.set_msaa_mode(1)
    w.apply(r)
    assert L.lib().edxr_set_state(r.ctx, L.STATE_FLOAT_COLOR, 1) == 0
    r.render_mesh(w.mesh)
    import ctypes as C
    import torch
    d_rgb = torch.zeros(w.height * w.width * 3, device="cuda")
    g = (L.Grads * 1)()
    rc = L.lib().edxr_backward(r.ctx, C.cast(d_rgb.data_ptr(), L.F), g, 1)
    assert rc == L.EDXR_E_STATE
    assert b"multisampled" in L.lib().edxr_last_error(r.ctx)
    r.close()
