"""Synthetic workloads: seeded scenes, cameras and resolutions.

Every scene is generated on the host with SplitMix64 seeds, so every consumer
receives identical bytes.  Projection: D3D-style
(z in [0,1]), left-handed, vertical fov 60 degrees, near 0.1, far 100.  Light
(1,1,-1) (EDXRaster/Core/Renderer.cpp:290), vertex-colour Blinn-Phong with the
reference constants (ambient 0.2, diffuse 3/pi, specular x3) and shininess 32.
"""
from dataclasses import dataclass, field

import numpy as np

from . import _lib as L
from .renderer import Mesh, look_at, perspective, raster_matrix, default_material

FOV, NEAR, FAR = 60.0, 0.1, 100.0


@dataclass
class Workload:
    name: str
    width: int
    height: int
    mesh: Mesh
    mv: np.ndarray
    proj: np.ndarray
    raster: np.ndarray
    light: tuple = (1.0, 1.0, -1.0)
    shader: int = L.SHADER_VC_BLINN_PHONG
    shininess: float = 32.0
    info: dict = field(default_factory=dict)
    textures: list = None               # texture slots (textured shader)
    tex_filter: int = L.FILTER_TRILINEAR

    def material(self):
        m = default_material(self.shader)
        m.shininess = self.shininess
        return m

    def job(self, mesh=None, hierarchical=True):
        """edxr_job of this workload for a batch (mesh: Mesh or edxr_mesh)."""
        from .renderer import Renderer
        return Renderer.make_job(self.mesh if mesh is None else mesh, self.width, self.height, self.mv, self.proj,
                                 self.raster, self.light, self.shader, self.material(), hierarchical)

    def apply(self, renderer):
        renderer.set_transform(self.mv, self.proj, self.raster)
        renderer.set_light(self.light)
        renderer.set_pixel_shader(self.shader, self.material())
        if self.textures is not None:
            renderer.set_textures(self.textures)
            renderer.set_texture_filter(self.tex_filter)


def camera(width, height, eye, at, up=(0.0, 1.0, 0.0)):
    return look_at(eye, at, up), perspective(FOV, width / height, NEAR, FAR), raster_matrix(width, height)


def small_sphere(width=800, height=600, slices=64, stacks=32):
    """UV sphere 64x32 (V=2,145, T=4,096), r=1 at the origin, eye (0,0,-3)."""
    mesh = Mesh.sphere(slices, stacks, 1.0, (0, 0, 0), color_seed=1)
    mv, p, r = camera(width, height, (0.0, 0.0, -3.0), (0.0, 0.0, 0.0))
    return Workload("sphere_64x32_800x600", width, height, mesh, mv, p, r)


def sphere_grid_1080p(width=1920, height=1080, grid=16, segments=64):
    """16x16 UV spheres (64x64 segments, T=2,097,152), radii U[0.5,2]; the eye sits
    inside sphere (8,8), 0.05 from its surface on the -z side, and looks along -x
    (slightly down) across the grid, so the enclosing sphere's surface crosses the
    near plane and is clipped (its triangles then face away and are culled)."""
    mesh = Mesh.sphere_grid(grid, segments, seed=2)
    k = (grid // 2) * grid + grid // 2
    vs = (segments + 1) ** 2
    pole = k * vs  # vertex 0 of a sphere is its north pole (c.x, c.y + r, c.z)
    c = np.array([mesh.px[pole], 0.0, mesh.pz[pole]], np.float64)
    radius = float(mesh.py[pole])
    eye = c + (radius - 0.05) * np.array([0.0, 0.0, -1.0])
    at = eye + 50.0 * np.array([-1.0, -0.08, 0.0])
    mv, p, r = camera(width, height, eye, at)
    return Workload("sphere_grid_16x16x64_1920x1080", width, height, mesh, mv, p, r,
                    info={"eye_sphere_radius": radius})


def heightfield_4k(width=3840, height=2160, n=2236):
    """2236x2236 heightfield (T=9,990,450), 8-octave value noise amplitude 0.2,
    eye (0,1.2,-1.6) looking at the origin, 3840x2160."""
    mesh = Mesh.load_heightfield(n, 0.2, height_seed=3, color_seed=33)
    mv, p, r = camera(width, height, (0.0, 1.2, -1.6), (0.0, 0.0, 0.0))
    return Workload("heightfield_2236_3840x2160", width, height, mesh, mv, p, r)


def heightfield_4k_textured(width=3840, height=2160, n=2236, tex_filter=L.FILTER_TRILINEAR):
    """The 10M-triangle heightfield with the textured Lambertian shader:
    coordinates repeating 32 times across, alternating between two 1024^2
    checker textures by triangle pair; trilinear filtering by default."""
    w = heightfield_4k(width, height, n)
    m = w.mesh
    m.tu = ((m.px + 1.0) * 16.0).astype(np.float32)
    m.tv = ((m.pz + 1.0) * 16.0).astype(np.float32)
    m.tex_ids = ((np.arange(m.num_triangles) // 2) % 2).astype(np.uint32)
    w.name = "heightfield_2236_textured_3840x2160"
    w.shader, w.shininess = L.SHADER_TEXTURED_LAMBERT, 0.0
    w.textures = [checker_texture(1024, 16, 21), checker_texture(1024, 32, 22)]
    w.tex_filter = tex_filter
    return w


def heightfield_small(n=64, width=320, height=240, seed=7):
    mesh = Mesh.load_heightfield(n, 0.2, height_seed=seed, color_seed=seed + 30)
    mv, p, r = camera(width, height, (0.0, 1.2, -1.6), (0.0, 0.0, 0.0))
    return Workload("heightfield_%d_%dx%d" % (n, width, height), width, height, mesh, mv, p, r)


def checker_texture(size, cells, seed):
    """RGBA8 checkerboard of `cells` x `cells` squares in seeded colours, with
    a seeded per-texel noise of +-16 so that mip levels differ."""
    rng = np.random.default_rng(seed)
    pal = rng.integers(0, 256, (2, 3))
    yy, xx = np.mgrid[0:size, 0:size]
    sq = ((xx * cells // size) + (yy * cells // size)) % 2
    rgb = pal[sq] + rng.integers(-16, 17, (size, size, 3))
    img = np.empty((size, size, 4), np.uint8)
    img[:, :, :3] = np.clip(rgb, 0, 255)
    img[:, :, 3] = 255
    return img


def textured_scene(width=640, height=480, n=96, tex_filter=L.FILTER_TRILINEAR, sphere=(48, 24)):
    """LambertianAlbedoPixelShader with texture slots: a heightfield with its
    texture coordinates repeating 3 times across, alternating between two
    checker textures (256^2 and 100x60) by triangle pair, and a UV sphere
    (u = longitude, v = latitude) on a constant-colour slot, seen obliquely so
    that the trilinear filter spans several mip levels."""
    hf = Mesh.load_heightfield(n, 0.15, height_seed=5, color_seed=6)
    hf.tu = ((hf.px + 1.0) * 1.5).astype(np.float32)
    hf.tv = ((hf.pz + 1.0) * 1.5).astype(np.float32)
    T = hf.num_triangles
    hf.tex_ids = ((np.arange(T) // 2) % 2).astype(np.uint32)
    sp = Mesh.sphere(sphere[0], sphere[1], 0.35, (0.1, 0.45, 0.2), color_seed=8)
    d = np.stack([sp.px - 0.1, sp.py - 0.45, sp.pz - 0.2], 1) / 0.35
    sp.tu = (np.arctan2(d[:, 2], d[:, 0]) / (2 * np.pi) + 0.5).astype(np.float32)
    sp.tv = (np.arccos(np.clip(d[:, 1], -1, 1)) / np.pi).astype(np.float32)
    sp.tex_ids = np.full(sp.num_triangles, 2, np.uint32)
    mesh = Mesh.merge(hf, sp)
    mv, p, r = camera(width, height, (0.0, 0.9, -1.9), (0.0, 0.0, 0.3))
    tex = [checker_texture(256, 8, 11), checker_texture(64, 4, 12)[:60, :, :].copy(), (0.8, 0.55, 0.3)]
    tex[1] = np.ascontiguousarray(np.concatenate([tex[1], tex[1][:, :36]], axis=1))  # 100 x 60
    return Workload("textured_%d_%dx%d" % (n, width, height), width, height, mesh, mv, p, r,
                    shader=L.SHADER_TEXTURED_LAMBERT, shininess=0.0, textures=tex, tex_filter=tex_filter)


def _front_facing(v, mvp):
    """Orders each triangle of v (T, 3, 3) clockwise on screen (the winding
    that survives the det > 0 cull) when all its vertices are in front of the
    eye; other triangles are left as given."""
    P = np.concatenate([v, np.ones(v.shape[:2] + (1,))], 2) @ np.asarray(mvp, np.float64).T
    w = P[..., 3]
    sx, sy = P[..., 0] / w, -P[..., 1] / w  # y down on screen
    det = (sx[:, 1] - sx[:, 0]) * (sy[:, 2] - sy[:, 0]) - (sy[:, 1] - sy[:, 0]) * (sx[:, 2] - sx[:, 0])
    flip = (det < 0) & (w > 0).all(1)
    v[flip] = v[flip][:, [0, 2, 1]]
    return v


def clip_scene(kind, width=160, height=120, seed=0):
    """Scenes whose triangles straddle the clip planes (Core/Clipper.h:71-288).

    kind "lateral": large triangles at depth 1.5..4.5 in front of an eye at
    (0, 0, -3), spreading beyond the left, right, bottom and top planes, no
    vertex near the near or far plane.  "far": the same camera with the far
    plane at view depth 6 and triangles from depth 1.5 to 9, so they cross it.
    "all": four triangles around the eye that cross the near, far and side
    planes at once (clip polygons of seven and eight vertices), in both
    windings, with the far plane at view depth 30.  Vertex normals and colours are seeded per vertex, and
    front faces are made clockwise where a triangle lies in front of the eye.
    Returns a Workload."""
    rng = np.random.Generator(np.random.PCG64(9100 + seed + {"lateral": 0, "far": 10, "all": 20}[kind]))
    eye, at = (0.0, 0.0, -3.0), (0.0, 0.0, 0.0)
    far = {"lateral": FAR, "far": 6.0, "all": 30.0}[kind]
    mv = look_at(eye, at)
    p = perspective(FOV, width / height, NEAR, far)
    r = raster_matrix(width, height)
    if kind == "lateral":
        n = 24
        c = np.stack([rng.uniform(-2.5, 2.5, n), rng.uniform(-2.0, 2.0, n), rng.uniform(-1.3, 1.3, n)], 1)
        v = c[:, None, :] + rng.normal(0.0, 1.6, (n, 3, 3)) * np.array([1.0, 1.0, 0.25])
        v[..., 2] = np.clip(v[..., 2], -1.5, 1.5)
    elif kind == "far":
        n = 24
        c = np.stack([rng.uniform(-1.2, 1.2, n), rng.uniform(-0.9, 0.9, n), rng.uniform(1.0, 4.0, n)], 1)
        v = c[:, None, :] + rng.normal(0.0, 1.0, (n, 3, 3)) * np.array([1.0, 1.0, 2.5])
        v[..., 2] = np.clip(v[..., 2], -1.5, 6.0)
    else:
        # Each crosses all six planes (clip polygons of 7-8 vertices); both
        # windings are included, so one of each pair survives the cull.
        v = np.array([[[9.074, 0.241, 9.2], [-1.0, 0.63, -4.97], [-3.408, -20.731, 33.659]],
                      [[-2.068, 0.011, -4.996], [19.598, -34.443, 39.561], [21.456, 15.514, 24.759]],
                      [[-20.488, -12.724, 29.12], [0.161, -0.044, -3.223], [2.515, 16.773, 18.458]],
                      [[-24.562, -8.192, 31.236], [16.651, -12.3, 26.239], [0.108, 0.888, -5.33]]])
        v = np.concatenate([v, v[:, [0, 2, 1]]])
    mvp = np.asarray(p, np.float64) @ np.asarray(mv, np.float64)
    v = _front_facing(v, mvp).reshape(-1, 3).astype(np.float32)
    nrm = rng.normal(0.0, 1.0, v.shape).astype(np.float32)
    nrm[:, 2] = -np.abs(nrm[:, 2]) - 0.5  # mostly toward the eye
    col = rng.uniform(0.0, 1.0, v.shape).astype(np.float32)
    mesh = Mesh(v[:, 0], v[:, 1], v[:, 2], nrm[:, 0], nrm[:, 1], nrm[:, 2], col[:, 0], col[:, 1], col[:, 2],
                np.arange(v.shape[0], dtype=np.uint32))
    return Workload("clip_%s_%dx%d" % (kind, width, height), width, height, mesh, mv, p, r,
                    info={"far": far})


def _unit(rng):
    v = rng.normal(size=3)
    return v / np.linalg.norm(v)


def batch_jobs(n_jobs=256, seed_base=4000, tmin=1e3, tmax=5e5, resolutions=(256, 512, 1024)):
    """Independent render jobs of uneven sizes.  Job j draws from a PCG64
    stream seeded seed_base + j: a UV sphere (r = 1, slices = 2s, stacks = s,
    T = 4 s^2) or an n x n heightfield (T = 2 (n-1)^2) with p = 0.5, the target
    triangle count log-uniform in [tmin, tmax], a square resolution from
    `resolutions`, the eye uniform on the radius-3 sphere looking at the
    origin and the light direction uniform on S^2."""
    jobs = []
    for j in range(n_jobs):
        rng = np.random.Generator(np.random.PCG64(seed_base + j))
        sphere = rng.random() < 0.5
        target = float(np.exp(rng.uniform(np.log(tmin), np.log(tmax))))
        res = int(resolutions[rng.integers(len(resolutions))])
        d = _unit(rng)
        light = tuple(float(x) for x in _unit(rng))
        if sphere:
            s = max(2, int(round(np.sqrt(target / 4.0))))
            mesh = Mesh.sphere(2 * s, s, 1.0, (0.0, 0.0, 0.0), color_seed=seed_base + j)
        else:
            n = max(3, int(round(1.0 + np.sqrt(target / 2.0))))
            mesh = Mesh.load_heightfield(n, 0.2, height_seed=seed_base + j, color_seed=seed_base + 1000 + j)
        up = (0.0, 1.0, 0.0) if abs(d[1]) < 0.99 else (0.0, 0.0, 1.0)
        mv, p, r = camera(res, res, 3.0 * d, (0.0, 0.0, 0.0), up)
        jobs.append(Workload("job%d_%s_%dx%d" % (j, "sphere" if sphere else "heightfield", res, res), res, res,
                             mesh, mv, p, r, light=light, info={"seed": seed_base + j}))
    return jobs


def diff_scene(n_views=64, res=512, hf_n=71, sphere_segments=(80, 64), shininess=32.0, light=(1.0, 1.0, -1.0)):
    """Sphere above a heightfield, seen from a ring of views.  A 71x71
    heightfield (9,800 triangles, amplitude 0.2, seeds 5/55) and an 80x64 UV
    sphere (10,240 triangles, r = 0.35, centre (0, 0.55, 0), colour seed 56)
    merged into one mesh; view k looks at (0, 0.15, 0) from distance 2.4 at
    elevation 30 degrees and azimuth 5.625 k degrees, res x res pixels."""
    hf = Mesh.load_heightfield(hf_n, 0.2, height_seed=5, color_seed=55)
    sph = Mesh.sphere(sphere_segments[0], sphere_segments[1], 0.35, (0.0, 0.55, 0.0), color_seed=56)
    mesh = Mesh.merge(hf, sph)
    views = []
    el = np.radians(30.0)
    target = np.array([0.0, 0.15, 0.0])
    for k in range(n_views):
        az = np.radians(5.625 * k)
        eye = target + 2.4 * np.array([np.cos(el) * np.sin(az), np.sin(el), -np.cos(el) * np.cos(az)])
        mv, p, r = camera(res, res, eye, target)
        views.append(Workload("diff_view%d_%dx%d" % (k, res, res), res, res, mesh, mv, p, r, light=light,
                              shininess=shininess))
    return mesh, views


def rotate(v, axis, deg):
    """Rodrigues rotation of v about axis by deg degrees."""
    v = np.asarray(v, np.float64)
    k = np.asarray(axis, np.float64)
    k = k / np.linalg.norm(k)
    t = np.radians(deg)
    return v * np.cos(t) + np.cross(k, v) * np.sin(t) + k * np.dot(k, v) * (1.0 - np.cos(t))


WORKLOADS = {
    "small": small_sphere,
    "grid1m": sphere_grid_1080p,
    "hf10m": heightfield_4k,
    "hf10m_tex": heightfield_4k_textured,
}
