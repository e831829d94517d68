"""The textured Lambertian path: LambertianAlbedoPixelShader with texture slots
(Core/Shader.h:209-244), per-triangle texture IDs (Utils/Mesh.cpp:24-31,
Core/Clipper.h:95) and the filters of RenderStates::TexFilter.

EDXUtil's Texture2D (sampling, mip chain, filters) is absent: its semantics
are the ones include/edxraster.h declares, so the oracle's sampling is pinned
by known answers that follow from those declarations and from the shading
formula -- a constant slot equals the constant-albedo shader, nearest and
bilinear reproduce exact texel values and ramps, trilinear minification by 4
lands on the twice box-filtered level, coordinates wrap -- and the GPU path is
compared with the oracle (IDs and depth exact, colour within 1 LSB).
"""
import numpy as np
import pytest

from edxraster_amd import configs, Mesh
from edxraster_amd import _lib as L

from tests.test_oracle import SCENES, mesh_from_pixels, identity_render

F = np.float32


def textured_rect(x0, y0, x1, y1, w, h, su, sv, slot=0):
    """Pixel-aligned rectangle with u running 0..su across and v 0..sv down."""
    tris = [[(x0, y0), (x1, y0), (x0, y1)], [(x1, y0), (x1, y1), (x0, y1)]]
    m = mesh_from_pixels(tris, w, h)
    pts = np.asarray(tris, np.float64).reshape(-1, 2)
    m.tu = ((pts[:, 0] - x0) / (x1 - x0) * su).astype(F)
    m.tv = ((pts[:, 1] - y0) / (y1 - y0) * sv).astype(F)
    m.tex_ids = np.full(2, slot, np.uint32)
    return m


def flat_diffuse():
    """The Lambertian term with the light perpendicular to the normal: the
    ambient part only, in the shader's float order."""
    return F(F(F(0.0) + F(0.2)) * F(3.0)) * F(1.0 / np.pi)


def to_byte(c):
    return np.floor(np.clip(c, 0, 1).astype(F) * F(255.0) + F(0.5)).astype(np.int32)


def render_tex(mesh, w, h, textures, tex_filter):
    from oracle.oracle import Oracle, SHADER_TEXTURED_LAMBERT
    o = Oracle()
    identity_render(o, mesh, w, h, shader=SHADER_TEXTURED_LAMBERT, light=(1.0, 0.0, 0.0), textures=textures,
                    tex_filter=tex_filter, shininess=0.0)
    return o.color()[::-1, :, :3].astype(np.int32)  # raster rows


def test_constant_slot_equals_constant_albedo_shader(built):
    from oracle.oracle import Oracle, SHADER_LAMBERT_ALBEDO, SHADER_TEXTURED_LAMBERT
    for name, mesh, w, h, mv, p, r in SCENES[:3]:
        mesh.tu = np.zeros(mesh.num_vertices, F)
        mesh.tv = np.zeros(mesh.num_vertices, F)
        mesh.tex_ids = np.zeros(mesh.num_triangles, np.uint32)
        a, b = Oracle(), Oracle()
        a.render(mesh, w, h, mv, p, r, shader=SHADER_LAMBERT_ALBEDO, albedo=(0.3, 0.6, 0.9), shininess=0.0)
        for f in (0, 1, 2):
            b.render(mesh, w, h, mv, p, r, shader=SHADER_TEXTURED_LAMBERT, textures=[(0.3, 0.6, 0.9)],
                     tex_filter=f, shininess=0.0)
            np.testing.assert_array_equal(a.color(), b.color())
            np.testing.assert_array_equal(a.ids(), b.ids())


def test_nearest_one_texel_per_pixel_and_wrap(built):
    rng = np.random.default_rng(3)
    tex = rng.integers(0, 256, (16, 16, 4), dtype=np.uint8)
    w = h = 32
    # u, v run 0..2 over 32 pixels: pixel x samples texel x mod 16
    got = render_tex(textured_rect(0, 0, 32, 32, w, h, 2.0, 2.0), w, h, [tex], 0)
    yy, xx = np.mgrid[0:32, 0:32]
    want = to_byte(flat_diffuse() * (tex[yy % 16, xx % 16, :3].astype(F) / F(255.0)))
    np.testing.assert_array_equal(got, want)


def test_bilinear_magnified_ramp(built):
    # texel i holds 8 i in red: 2x magnification samples at texel x = (px + 0.5) / 2 - 0.5
    tex = np.zeros((4, 16, 4), np.uint8)
    tex[:, :, 0] = (8 * np.arange(16))[None, :]
    tex[:, :, 3] = 255
    w, h = 32, 8
    got = render_tex(textured_rect(0, 0, 32, 8, w, h, 1.0, 1.0), w, h, [tex], 1)
    xt = (np.arange(32) + 0.5) / 2.0 - 0.5
    interior = (xt >= 0) & (xt <= 15)
    want = to_byte(flat_diffuse() * (F(8.0) * xt.astype(F) / F(255.0)))
    assert np.abs(got[:, interior, 0] - want[None, interior]).max() <= 1
    # the edge column wraps: x = -0.25 takes a quarter of texel 15 (120) and the rest of texel 0 (0)
    assert abs(int(got[0, 0, 0]) - int(to_byte(flat_diffuse() * F(0.25 * 120 / 255.0)))) <= 1


def test_trilinear_minified_by_four_is_level_two(built):
    rng = np.random.default_rng(4)
    tex = rng.integers(0, 256, (64, 64, 4), dtype=np.uint8)
    w = h = 16
    # 64 texels over 16 pixels: differentials of 4 texels, lod = 2, level 2 is 16 x 16
    got = render_tex(textured_rect(0, 0, 16, 16, w, h, 1.0, 1.0), w, h, [tex], 2)
    l0 = tex.astype(F) / F(255.0)

    def box(a):
        return F(0.25) * (((a[0::2, 0::2] + a[0::2, 1::2]) + a[1::2, 0::2]) + a[1::2, 1::2])

    want = to_byte(flat_diffuse() * box(box(l0))[:, :, :3])
    assert np.abs(got - want).max() <= 1
    assert np.mean(got == want) > 0.95
    # the bilinear filter on level 0 differs (no prefiltering)
    lin = render_tex(textured_rect(0, 0, 16, 16, w, h, 1.0, 1.0), w, h, [tex], 1)
    assert np.abs(lin - want).max() > 8


def test_oracle_rejects_bad_texture_slots(built):
    from oracle.oracle import Oracle
    m = textured_rect(0, 0, 8, 8, 8, 8, 1.0, 1.0, slot=2)
    with pytest.raises(RuntimeError):
        identity_render(Oracle(), m, 8, 8, shader=3, textures=[(1, 1, 1)])


# ---- GPU parity (through the C ABI) ----

def render_gpu_and_oracle(w, msaa=0, hierarchical=True):
    from oracle.oracle import Oracle
    from edxraster_amd import Renderer
    r = Renderer(w.width, w.height)
    r.set_msaa_mode(msaa)
    w.apply(r)
    r.set_hierarchical_rasterize(hierarchical)
    r.render_mesh(w.mesh)
    o = Oracle()
    o.render_workload(w, threads=8, msaa_log2=msaa, hierarchical=1 if hierarchical else 0)
    return r, o


@pytest.mark.gpu
@pytest.mark.parametrize("tex_filter", [0, 1, 2])
def test_gpu_textured_parity(built, tex_filter):
    from tests.test_gpu_parity import assert_frame_parity
    w = configs.textured_scene(tex_filter=tex_filter)
    r, o = render_gpu_and_oracle(w)
    info = assert_frame_parity(r, o, "textured filter %d" % tex_filter)
    assert info["covered"] > 10000
    r.close()


@pytest.mark.gpu
def test_gpu_textured_clipped_hierarchy_off_and_msaa(built):
    from tests.test_gpu_parity import assert_frame_parity
    from tests.test_msaa import assert_ms_parity
    w = configs.textured_scene(320, 240, n=48)
    w.mv, w.proj, w.raster = configs.camera(320, 240, (0.0, 0.05, -0.6), (0.0, 0.0, 0.5))  # near plane cuts
    r, o = render_gpu_and_oracle(w, hierarchical=False)
    assert_frame_parity(r, o, "textured close-up")
    assert r.stats()["near_clipped"] > 0
    r.close()
    r, o = render_gpu_and_oracle(configs.textured_scene(320, 240, n=48), msaa=2)
    assert_ms_parity(r, o, "textured 4x")
    r.close()


@pytest.mark.gpu
def test_gpu_textured_batch_and_errors(built):
    from edxraster_amd import Renderer
    w = configs.textured_scene(256, 192, n=40)
    r = Renderer(w.width, w.height)
    w.apply(r)
    r.render_mesh(w.mesh)
    single = r.get_back_buffer().copy()
    jobs = [w.job(), w.job()]
    r.render_batch(jobs)
    rgba, depth, ids = r.batch_read(1)
    np.testing.assert_array_equal(rgba, single)
    # a texture slot beyond the set is an index error
    bad = Mesh(*w.mesh.planes(), w.mesh.indices, w.mesh.tu, w.mesh.tv, w.mesh.tex_ids.copy())
    bad.tex_ids[5] = 3
    with pytest.raises(RuntimeError):
        r.render_mesh(bad)
    # no backward through the textured shader
    import ctypes as C
    import torch
    assert L.lib().edxr_set_state(r.ctx, L.STATE_FLOAT_COLOR, 1) == 0
    r.render_mesh(w.mesh)
    d_rgb = torch.zeros(w.height * w.width * 3, device="cuda")
    g = (L.Grads * 1)()
    assert L.lib().edxr_backward(r.ctx, C.cast(d_rgb.data_ptr(), L.F), g, 1) == L.EDXR_E_STATE
    r.close()
