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


def band_ramp_texture():
    """128 x 64 RGBA8: red in bands of 4 rows (255 on rows with (j // 4) even,
    0 otherwise), green a ramp of 2 per texel column (2 i), blue 128."""
    tex = np.zeros((64, 128, 4), np.uint8)
    tex[:, :, 0] = np.where((np.arange(64) // 4) % 2 == 0, 255, 0)[:, None]
    tex[:, :, 1] = (2 * np.arange(128))[None, :]
    tex[:, :, 2] = 128
    tex[:, :, 3] = 255
    return tex


def band_profile(level, v):
    """FP64 red of the band texture at mip level `level`, bilinear in v: the
    level's rows are 2x2 box means (pairs of rows for this u-constant red),
    texel centres at (j + 0.5) / h, rows wrap."""
    rows = np.where((np.arange(64) // 4) % 2 == 0, 1.0, 0.0)
    for _ in range(level):
        rows = 0.5 * (rows[0::2] + rows[1::2])
    h = rows.shape[0]
    y = v * h - 0.5
    j0 = np.floor(y).astype(np.int64)
    f = y - j0
    return (1.0 - f) * rows[j0 % h] + f * rows[(j0 + 1) % h]


def aniso_rect(w, h, x0, y0, x1, y1, u0, du, dv):
    """Screen-aligned rectangle with u = u0 + du * (x - x0) and v = dv * (y - y0)
    in texture coordinates per pixel (affine: identity camera, w = 1)."""
    tris = [[(x0, y0), (x1, y0), (x0, y1)], [(x1, y0), (x1, y1), (x0, y1)]]
    m = mesh_from_pixels(tris, w, h)
    pts = np.asarray(tris, np.float64).reshape(-1, 2)
    m.tu = (u0 + (pts[:, 0] - x0) * du).astype(F)
    m.tv = ((pts[:, 1] - y0) * dv).astype(F)
    m.tex_ids = np.zeros(2, np.uint32)
    return m


def render_lit(mesh, w, h, textures, tex_filter):
    """Oracle render with the light along the normal (diffuse = 1.2 * 3 / pi)."""
    from oracle.oracle import Oracle, SHADER_TEXTURED_LAMBERT
    o = Oracle()
    identity_render(o, mesh, w, h, shader=SHADER_TEXTURED_LAMBERT, light=(0.0, 0.0, -1.0), textures=textures,
                    tex_filter=tex_filter, shininess=0.0)
    return o.color()[::-1, :, :3].astype(np.int32)


LIT = 1.2 * 3.0 / np.pi


@pytest.mark.parametrize("tex_filter", [3, 4, 5], ids=["aniso4", "aniso8", "aniso16"])
def test_anisotropic_keeps_the_minor_axis_detail(built, tex_filter):
    """Footprint 4 texels along u, 1/2 texel along v (ratio 8:1 at most).
    Anisotropic filtering samples at the minor axis' lod 0 (raised to 4/N
    texels, still <= 1): the red bands come out at level 0, where trilinear
    (lod 2 from the major axis) blurs them; the N taps are centred on the
    pixel's coordinate, so the green ramp reads its value there."""
    w, h, x0, y0, x1, y1 = 40, 64, 4, 8, 28, 56
    du, dv = 4.0 / 128, 0.5 / 64
    m = aniso_rect(w, h, x0, y0, x1, y1, 16.0 / 128, du, dv)
    tex = band_ramp_texture()
    got = render_lit(m, w, h, [tex], tex_filter)
    ys, xs = np.mgrid[y0:y1, x0:x1]
    u = 16.0 / 128 + (xs + 0.5 - x0) * du
    v = (ys + 0.5 - y0) * dv
    red = band_profile(0, v)
    green = 2.0 * (u * 128 - 0.5) / 255.0          # the ramp is linear: taps centred on u average to it
    want_r = np.trunc(np.clip(LIT * red, 0, 1) * 255 + 0.5)
    want_g = np.trunc(np.clip(LIT * green, 0, 1) * 255 + 0.5)
    assert np.abs(got[y0:y1, x0:x1, 0] - want_r).max() <= 1
    assert np.abs(got[y0:y1, x0:x1, 1] - want_g).max() <= 1
    # trilinear blurs the bands (lod 2) ...
    tri = render_lit(m, w, h, [tex], 2)
    assert np.abs(tri[y0:y1, x0:x1, 0] - want_r).max() > 50
    # ... and taps placed from the pixel's coordinate forward (k / N instead of
    # (2k + 1 - N) / 2N of the differential) would read a ramp 1.5 texels on
    N = 4 << (tex_filter - 3)
    shifted = np.trunc(np.clip(LIT * 2.0 * ((u + (N - 1) / (2.0 * N) * du) * 128 - 0.5) / 255.0, 0, 1) * 255 + 0.5)
    assert np.abs(shifted - want_g).max() >= 3


@pytest.mark.parametrize("tex_filter,level", [(3, 2), (4, 1), (5, 0)], ids=["aniso4", "aniso8", "aniso16"])
def test_anisotropic_ratio_clamp_sets_the_level(built, tex_filter, level):
    """Footprint 16 texels along u, 1/2 along v (32:1): N:1 filtering raises
    the minor axis to 16 / N texels, so 4x reads level 2, 8x level 1 and 16x
    level 0 of the (u-constant) red bands."""
    w, h, x0, y0, x1, y1 = 40, 64, 4, 8, 36, 56
    du, dv = 16.0 / 128, 0.5 / 64
    m = aniso_rect(w, h, x0, y0, x1, y1, 0.0, du, dv)
    got = render_lit(m, w, h, [band_ramp_texture()], tex_filter)
    ys, xs = np.mgrid[y0:y1, x0:x1]
    v = (ys + 0.5 - y0) * dv
    want = np.trunc(np.clip(LIT * band_profile(level, v), 0, 1) * 255 + 0.5)
    assert np.abs(got[y0:y1, x0:x1, 0] - want).max() <= 1
    others = [l for l in (0, 1, 2) if l != level]
    for l in others:  # the neighbouring levels differ
        alt = np.trunc(np.clip(LIT * band_profile(l, v), 0, 1) * 255 + 0.5)
        assert np.abs(got[y0:y1, x0:x1, 0] - alt).max() > 20


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
@pytest.mark.parametrize("tex_filter", [0, 1, 2, 3, 4, 5])
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
