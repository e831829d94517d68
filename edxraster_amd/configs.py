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

    def material(self):
        m = default_material(self.shader)
        m.shininess = self.shininess
        return m

    def apply(self, renderer):
        renderer.set_transform(self.mv, self.proj, self.raster)
        renderer.set_light(self.light)
        renderer.set_pixel_shader(self.shader, self.material())


def camera(width, height, eye, at, up=(0.0, 1.0, 0.0)):
    return look_at(eye, at, up), perspective(FOV, width / height, NEAR, FAR), raster_matrix(width, height)


def small_sphere(width=800, height=600, slices=64, stacks=32):
    """UV sphere 64x32 (V=2,145, T=4,096), r=1 at the origin, eye (0,0,-3)."""
    mesh = Mesh.load_sphere(slices, stacks, 1.0, (0, 0, 0), color_seed=1)
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


def heightfield_small(n=64, width=320, height=240, seed=7):
    mesh = Mesh.load_heightfield(n, 0.2, height_seed=seed, color_seed=seed + 30)
    mv, p, r = camera(width, height, (0.0, 1.2, -1.6), (0.0, 0.0, 0.0))
    return Workload("heightfield_%d_%dx%d" % (n, width, height), width, height, mesh, mv, p, r)


WORKLOADS = {
    "small": small_sphere,
    "grid1m": sphere_grid_1080p,
    "hf10m": heightfield_4k,
}
