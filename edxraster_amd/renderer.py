"""Python mirror of EDXRaster's host API over the C ABI.

``Renderer`` follows ``EDX::RasterRenderer::Renderer`` (EDXRaster/Core/Renderer.h:37-50):
``initialize``, ``resize``, ``set_transform``, ``render_mesh``, ``get_back_buffer``,
``set_hierarchical_rasterize``; the pixel-shader object (Core/Shader.h:173-282) is
selected with ``set_pixel_shader`` and the light with ``set_light``.  ``Mesh``
follows ``EDX::RasterRenderer::Mesh`` (EDXRaster/Utils/Mesh.h:17-69) with SoA
fp32 arrays.  Every render goes through libedxraster.so (HIP, gfx950).
"""
import ctypes as C
import os

import numpy as np

from . import _lib as L


def _check(ctx, rc):
    if rc != L.EDXR_OK:
        msg = L.lib().edxr_last_error(ctx).decode() if ctx else ""
        raise L.EdxrError(rc, msg)


def _mat(m):
    a = np.ascontiguousarray(np.asarray(m, dtype=np.float32).reshape(16))
    return a


def _load_rgba(fn, path):
    w, h = C.c_uint32(), C.c_uint32()
    p = C.POINTER(C.c_uint8)()
    _check(None, fn(os.fsencode(path), C.byref(w), C.byref(h), C.byref(p)))
    try:
        return np.ctypeslib.as_array(p, shape=(h.value, w.value, 4)).copy()
    finally:
        L.lib().edxr_free(p)


def load_bmp(path):
    """(H, W, 4) uint8 RGBA of an uncompressed 24/32-bit BMP, rows bottom-up (v = 0 first)."""
    return _load_rgba(L.lib().edxr_load_bmp, path)


def load_image(path):
    """(H, W, 4) uint8 RGBA of a BMP, TGA (types 2 and 10) or binary PPM
    diffuse map, rows bottom-up (v = 0 first)."""
    return _load_rgba(L.lib().edxr_load_image, path)


class Mesh:
    """SoA mesh: positions, normals, RGB colours (float32, V each) and uint32
    indices (3T); optionally texture coordinates tu, tv (V each) and a texture
    slot per triangle (tex_ids, T) for the textured shader."""

    def __init__(self, px, py, pz, nx, ny, nz, cr, cg, cb, indices, tu=None, tv=None, tex_ids=None):
        f = lambda a: np.ascontiguousarray(a, dtype=np.float32)
        self.px, self.py, self.pz = f(px), f(py), f(pz)
        self.nx, self.ny, self.nz = f(nx), f(ny), f(nz)
        self.cr, self.cg, self.cb = f(cr), f(cg), f(cb)
        self.indices = np.ascontiguousarray(indices, dtype=np.uint32).reshape(-1)
        self.tu = None if tu is None else f(tu)
        self.tv = None if tv is None else f(tv)
        self.tex_ids = None if tex_ids is None else np.ascontiguousarray(tex_ids, dtype=np.uint32).reshape(-1)

    @property
    def num_vertices(self):
        return int(self.px.shape[0])

    @property
    def num_triangles(self):
        return int(self.indices.shape[0] // 3)

    def planes(self):
        return [self.px, self.py, self.pz, self.nx, self.ny, self.nz, self.cr, self.cg, self.cb]

    def positions(self):
        return np.stack([self.px, self.py, self.pz], axis=1)

    def c_struct(self):
        """edxr_mesh over this mesh's host arrays (keep the Mesh alive meanwhile)."""
        return L.Mesh(*[L.fptr(a) for a in self.planes()], L.uptr(self.indices),
                      self.num_vertices, self.num_triangles, L.MEM_HOST,
                      L.fptr(self.tu), L.fptr(self.tv), L.uptr(self.tex_ids))

    @staticmethod
    def _from_host(hm):
        V, T = hm.num_vertices, hm.num_triangles
        planes = []
        for name in ("px", "py", "pz", "nx", "ny", "nz", "cr", "cg", "cb"):
            p = getattr(hm, name)
            planes.append(np.ctypeslib.as_array(p, shape=(V,)).copy() if V else np.zeros(0, np.float32))
        idx = np.ctypeslib.as_array(hm.indices, shape=(3 * T,)).copy() if T else np.zeros(0, np.uint32)
        extra = {}
        for name, n in (("tu", V), ("tv", V), ("tex_ids", T)):
            p = getattr(hm, name)
            if p:
                extra[name] = np.ctypeslib.as_array(p, shape=(n,)).copy() if n else \
                    np.zeros(0, np.uint32 if name == "tex_ids" else np.float32)
        L.lib().edxr_scene_free(C.byref(hm))
        return Mesh(*planes, idx, **extra)

    # A UV sphere with seeded vertex colours (the synthetic benchmark scenes).
    @staticmethod
    def sphere(slices=64, stacks=64, radius=1.0, center=(0.0, 0.0, 0.0), color_seed=1):
        hm = L.HostMesh()
        c = np.asarray(center, dtype=np.float32)
        rc = L.lib().edxr_scene_sphere(C.byref(hm), slices, stacks, radius, L.fptr(c), color_seed)
        _check(None, rc)
        return Mesh._from_host(hm)

    # Mesh::LoadSphere (EDXRaster/Utils/Mesh.h:39-44, Utils/Mesh.cpp:53-70):
    # texture coordinates, white vertex colours, texture slot 0 everywhere.
    @staticmethod
    def load_sphere(pos, scl, rot, radius, slices=64, stacks=64):
        hm = L.HostMesh()
        p, s, r = (np.asarray(v, dtype=np.float32) for v in (pos, scl, rot))
        _check(None, L.lib().edxr_load_sphere(L.fptr(p), L.fptr(s), L.fptr(r), radius, slices, stacks, C.byref(hm)))
        return Mesh._from_host(hm)

    # Mesh::LoadPlane (EDXRaster/Utils/Mesh.h:34-37, Utils/Mesh.cpp:36-51).
    @staticmethod
    def load_plane(pos, scl, rot, length):
        hm = L.HostMesh()
        p, s, r = (np.asarray(v, dtype=np.float32) for v in (pos, scl, rot))
        _check(None, L.lib().edxr_load_plane(L.fptr(p), L.fptr(s), L.fptr(r), length, C.byref(hm)))
        return Mesh._from_host(hm)

    def get_bounds(self):
        """Mesh::GetBounds: (min, max) corners of the vertex positions."""
        P = self.positions()
        return P.min(0), P.max(0)

    # Mesh::LoadPlane stand-in: a noise heightfield (EDXRaster/Utils/Mesh.cpp:36-51)
    @staticmethod
    def load_heightfield(n, amplitude=0.2, height_seed=3, color_seed=33):
        hm = L.HostMesh()
        _check(None, L.lib().edxr_scene_heightfield(C.byref(hm), n, amplitude, height_seed, color_seed))
        return Mesh._from_host(hm)

    # Mesh::LoadMesh (EDXRaster/Utils/Mesh.cpp:11-34)
    @staticmethod
    def load_obj(path, pos=(0.0, 0.0, 0.0), scl=(1.0, 1.0, 1.0), rot=(0.0, 0.0, 0.0)):
        hm = L.HostMesh()
        p, s, r = (np.asarray(v, dtype=np.float32) for v in (pos, scl, rot))
        _check(None, L.lib().edxr_load_obj(os.fsencode(path), L.fptr(p), L.fptr(s), L.fptr(r), C.byref(hm)))
        return Mesh._from_host(hm)

    # The material list Mesh::LoadMesh builds (Utils/Mesh.cpp:22-31): per
    # material (name, Kd, diffuse map path or "").
    @staticmethod
    def obj_materials(path):
        n = L.lib().edxr_obj_materials(os.fsencode(path), None, 0)
        _check(None, n if n < 0 else 0)
        arr = (L.ObjMaterial * max(1, n))()
        L.lib().edxr_obj_materials(os.fsencode(path), arr, n)
        return [(m.name.decode(), tuple(m.kd), m.map_kd.decode()) for m in arr[:n]]

    # The texture slots of an OBJ's triangles (tex_ids): ImageTexture for a
    # material whose diffuse map reads as BMP, TGA or PPM, ConstantTexture2D(Kd)
    # otherwise, and a last 0.9-grey slot for faces without a material.
    @staticmethod
    def obj_texture_slots(path):
        slots = []
        for name, kd, mp in Mesh.obj_materials(path):
            img = None
            if mp and os.path.exists(mp):
                try:
                    img = load_image(mp)
                except L.EdxrError:
                    img = None
            slots.append(kd if img is None else img)
        return slots + [(0.9, 0.9, 0.9)]

    @staticmethod
    def sphere_grid(grid=16, segments=64, seed=2):
        hm = L.HostMesh()
        _check(None, L.lib().edxr_scene_sphere_grid(C.byref(hm), grid, segments, seed))
        return Mesh._from_host(hm)

    @staticmethod
    def merge(a, b):
        planes = [np.concatenate([x, y]) for x, y in zip(a.planes(), b.planes())]
        idx = np.concatenate([a.indices, b.indices + np.uint32(a.num_vertices)])
        extra = {}
        for n, size in (("tu", "num_vertices"), ("tv", "num_vertices"), ("tex_ids", "num_triangles")):
            x, y = getattr(a, n), getattr(b, n)
            if x is not None or y is not None:
                z = lambda m, v: v if v is not None else np.zeros(getattr(m, size), np.uint32 if n == "tex_ids"
                                                                  else np.float32)
                extra[n] = np.concatenate([z(a, x), z(b, y)])
        return Mesh(*planes, idx, **extra)


def look_at(eye, at, up=(0.0, 1.0, 0.0)):
    out = np.zeros(16, np.float32)
    e, a, u = (np.asarray(v, np.float32) for v in (eye, at, up))
    L.lib().edxr_look_at(L.fptr(e), L.fptr(a), L.fptr(u), L.fptr(out))
    return out.reshape(4, 4)


def perspective(fov_y_deg, aspect, znear, zfar):
    out = np.zeros(16, np.float32)
    L.lib().edxr_perspective(fov_y_deg, aspect, znear, zfar, L.fptr(out))
    return out.reshape(4, 4)


def raster_matrix(width, height):
    out = np.zeros(16, np.float32)
    L.lib().edxr_raster_matrix(width, height, L.fptr(out))
    return out.reshape(4, 4)


def default_material(kind):
    m = L.Material()
    L.lib().edxr_default_material(kind, C.byref(m))
    return m


class Renderer:
    """EDX::RasterRenderer::Renderer on one MI355X (one context, one HIP stream)."""

    def __init__(self, width=None, height=None):
        self._ctx = None
        if width is not None:
            self.initialize(width, height)

    # Renderer::Initialize (Core/Renderer.cpp:22-62)
    def initialize(self, width, height):
        self.close()
        ctx = C.c_void_p()
        rc = L.lib().edxr_create(C.byref(ctx), width, height)
        if rc != L.EDXR_OK:
            raise L.EdxrError(rc, "edxr_create failed (is a gfx950 GPU visible?)")
        self._ctx = ctx
        self.width, self.height = int(width), int(height)

    def close(self):
        if self._ctx:
            L.lib().edxr_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def ctx(self):
        return self._ctx

    # Renderer::Resize (Core/Renderer.cpp:64-83)
    def resize(self, width, height):
        _check(self._ctx, L.lib().edxr_resize(self._ctx, width, height))
        self.width, self.height = int(width), int(height)

    # Renderer::SetTransform (Core/Renderer.cpp:85-92)
    def set_transform(self, model_view, proj, to_raster):
        mv, p, r = _mat(model_view), _mat(proj), _mat(to_raster)
        _check(self._ctx, L.lib().edxr_set_transform(self._ctx, L.fptr(mv), L.fptr(p), L.fptr(r)))

    def set_light(self, direction):
        d = np.ascontiguousarray(direction, dtype=np.float32)
        _check(self._ctx, L.lib().edxr_set_light(self._ctx, L.fptr(d)))

    def set_pixel_shader(self, kind, material=None):
        _check(self._ctx, L.lib().edxr_set_pixel_shader(self._ctx, kind, C.byref(material) if material else None))

    # Renderer::SetHierarchicalRasterize (Core/Renderer.h:49)
    def set_hierarchical_rasterize(self, on):
        _check(self._ctx, L.lib().edxr_set_state(self._ctx, L.STATE_HIERARCHICAL, 1 if on else 0))

    # Renderer::SetWriteFrames (Core/Renderer.h:50); frames go to <directory>/Frames/
    def set_write_frames(self, on, directory=None):
        if directory is not None:
            _check(self._ctx, L.lib().edxr_set_frame_directory(self._ctx, os.fsencode(directory)))
        _check(self._ctx, L.lib().edxr_set_state(self._ctx, L.STATE_WRITE_FRAMES, 1 if on else 0))

    # Renderer::WriteFrameToFile (Core/Renderer.cpp:352-358)
    def write_frame_to_file(self):
        _check(self._ctx, L.lib().edxr_write_frame_to_file(self._ctx))

    @property
    def frame_count(self):
        return int(L.lib().edxr_frame_count(self._ctx))

    # Mesh::GetTextures bound as RenderStates::TextureSlots: each slot an
    # (H, W, 4) uint8 image (rows from v = 0) or an RGB colour triple.
    def set_textures(self, textures):
        arr = (L.Texture * max(1, len(textures)))()
        self._textures = []
        for k, t in enumerate(textures):
            if isinstance(t, np.ndarray):
                img = np.ascontiguousarray(t, dtype=np.uint8)
                assert img.ndim == 3 and img.shape[2] == 4
                self._textures.append(img)
                arr[k].width, arr[k].height, arr[k].rgba = img.shape[1], img.shape[0], img.ctypes.data
            else:
                arr[k].color[:] = [float(c) for c in t]
        _check(self._ctx, L.lib().edxr_set_textures(self._ctx, arr, len(textures)))

    # Renderer::SetTextureFilter (RenderStates::TexFilter): L.FILTER_*
    def set_texture_filter(self, f):
        _check(self._ctx, L.lib().edxr_set_state(self._ctx, L.STATE_TEXTURE_FILTER, int(f)))

    # Renderer::SetMSAAMode (Core/Renderer.cpp:94-98)
    def set_msaa_mode(self, sample_count_log2):
        _check(self._ctx, L.lib().edxr_set_state(self._ctx, L.STATE_MSAA, int(sample_count_log2)))
        self.samples = 1 << int(sample_count_log2)

    def read_sample_buffers(self):
        """(depth, ids), each (samples, H, W), rows as read_depth."""
        S = getattr(self, "samples", 1)
        depth = np.empty((S, self.height, self.width), np.float32)
        ids = np.empty((S, self.height, self.width), np.uint32)
        _check(self._ctx, L.lib().edxr_read_sample_buffers(self._ctx, depth.ctypes.data, ids.ctypes.data, depth.size))
        return depth, ids

    def set_bin_chunks(self, n):
        _check(self._ctx, L.lib().edxr_set_state(self._ctx, L.STATE_BIN_CHUNKS, int(n)))

    # Renderer::RenderMesh (Core/Renderer.cpp:100-118)
    def render_mesh(self, mesh):
        if isinstance(mesh, Mesh):
            cm = mesh.c_struct()
        else:
            cm = mesh  # an edxr_mesh structure (e.g. device pointers)
        _check(self._ctx, L.lib().edxr_render_mesh(self._ctx, C.byref(cm)))

    # Renderer::GetBackBuffer (Core/Renderer.cpp:360-363): RGBA8, bottom-up rows.
    def get_back_buffer(self):
        out = np.empty((self.height, self.width, 4), np.uint8)
        _check(self._ctx, L.lib().edxr_read_color(self._ctx, out.ctypes.data, out.nbytes))
        return out

    def read_depth(self):
        out = np.empty((self.height, self.width), np.float32)
        _check(self._ctx, L.lib().edxr_read_depth(self._ctx, out.ctypes.data, out.size))
        return out

    def read_triangle_ids(self):
        out = np.empty((self.height, self.width), np.uint32)
        _check(self._ctx, L.lib().edxr_read_triangle_ids(self._ctx, out.ctypes.data, out.size))
        return out

    def set_profiling(self, on):
        _check(self._ctx, L.lib().edxr_set_state(self._ctx, L.STATE_PROFILE, 1 if on else 0))

    def profile(self, reset=False):
        """Summed per-stage milliseconds (HIP events) and the frame count."""
        ms = (C.c_double * len(L.STAGES))()
        n = C.c_uint64()
        _check(self._ctx, L.lib().edxr_profile_read(self._ctx, ms, C.byref(n)))
        if reset:
            _check(self._ctx, L.lib().edxr_profile_reset(self._ctx))
        return dict(zip(L.STAGES, list(ms))), n.value

    def device_buffers(self):
        c, d, i = C.c_void_p(), C.c_void_p(), C.c_void_p()
        _check(self._ctx, L.lib().edxr_device_buffers(self._ctx, C.byref(c), C.byref(d), C.byref(i)))
        return c.value, d.value, i.value

    def stats(self):
        s = L.Stats()
        _check(self._ctx, L.lib().edxr_get_stats(self._ctx, C.byref(s)))
        return {name: getattr(s, name) for name, _ in L.Stats._fields_}

    # ---- batched rendering (edxr_render_batch) ----
    @staticmethod
    def make_job(mesh, width, height, mv, proj, raster, light=(1.0, 1.0, -1.0), shader=L.SHADER_VC_BLINN_PHONG,
                 material=None, hierarchical=True):
        """edxr_job for a batch; mesh is a Mesh (host arrays, kept alive by the
        caller) or an edxr_mesh structure (e.g. device pointers)."""
        j = L.Job()
        j.mesh = mesh.c_struct() if isinstance(mesh, Mesh) else mesh
        j.width, j.height = int(width), int(height)
        j.mv[:] = [float(x) for x in _mat(mv)]
        j.proj[:] = [float(x) for x in _mat(proj)]
        j.raster[:] = [float(x) for x in _mat(raster)]
        j.light[:] = [float(x) for x in light]
        j.shader = shader
        j.material = material if material is not None else default_material(shader)
        j.hierarchical = 1 if hierarchical else 0
        return j

    def render_batch(self, jobs):
        """Renders a list (or ctypes array) of edxr_job in one launch sequence."""
        arr = jobs if isinstance(jobs, C.Array) else (L.Job * len(jobs))(*jobs)
        self._batch_sizes = [(j.width, j.height) for j in arr]
        _check(self._ctx, L.lib().edxr_render_batch(self._ctx, arr, len(arr)))

    def batch_read(self, k):
        """(rgba (H,W,4) uint8, depth (H,W) float32, ids (H,W) uint32) of job k."""
        w, h = self._batch_sizes[k]
        rgba = np.empty((h, w, 4), np.uint8)
        depth = np.empty((h, w), np.float32)
        ids = np.empty((h, w), np.uint32)
        _check(self._ctx, L.lib().edxr_batch_read(self._ctx, k, rgba.ctypes.data, depth.ctypes.data,
                                                  ids.ctypes.data))
        return rgba, depth, ids

    def batch_device_buffers(self):
        n = len(self._batch_sizes)
        c, d, i = C.c_void_p(), C.c_void_p(), C.c_void_p()
        offs = (C.c_uint64 * (n + 1))()
        _check(self._ctx, L.lib().edxr_batch_device_buffers(self._ctx, C.byref(c), C.byref(d), C.byref(i), offs))
        return c.value, d.value, i.value, list(offs)

    def set_graph(self, on):
        _check(self._ctx, L.lib().edxr_set_state(self._ctx, L.STATE_GRAPH, 1 if on else 0))

    def set_tuning(self, tris_per_chunk=0, raster_blocks_per_cu=0):
        _check(self._ctx, L.lib().edxr_set_state(self._ctx, L.STATE_TRIS_PER_CHUNK, int(tris_per_chunk)))
        _check(self._ctx, L.lib().edxr_set_state(self._ctx, L.STATE_RASTER_BLOCKS_PER_CU, int(raster_blocks_per_cu)))

    def autotune(self, jobs):
        arr = jobs if isinstance(jobs, C.Array) else (L.Job * len(jobs))(*jobs)
        t = L.Tuning()
        _check(self._ctx, L.lib().edxr_autotune(self._ctx, arr, len(arr), C.byref(t)))
        self._batch_sizes = [(j.width, j.height) for j in arr]
        return {"tris_per_chunk": t.tris_per_chunk, "raster_blocks_per_cu": t.raster_blocks_per_cu,
                "frame_ms": t.frame_ms}

    def debug_records(self):
        """Setup records in submission order, (N, 16) uint32 (see edxr_debug_read_records)."""
        n = L.lib().edxr_debug_read_records(self._ctx, None, 0)
        if n < 0:
            _check(self._ctx, int(n))
        out = np.zeros((max(n, 1), 16), np.uint32)
        n2 = L.lib().edxr_debug_read_records(self._ctx, out.ctypes.data, n)
        if n2 < 0:
            _check(self._ctx, int(n2))
        return out[:n]

    def debug_bins(self):
        """(offsets[num_tiles+1], refs (N,3) = tri_id, sub, flags)."""
        ntiles = ((self.width + 31) // 32) * ((self.height + 31) // 32)
        offs = np.zeros(ntiles + 1, np.uint64)
        n = L.lib().edxr_debug_read_bins(self._ctx, offs.ctypes.data, None, 0)
        if n < 0:
            _check(self._ctx, int(n))
        refs = np.zeros((max(n, 1), 3), np.uint32)
        n2 = L.lib().edxr_debug_read_bins(self._ctx, offs.ctypes.data, refs.ctypes.data, n)
        if n2 < 0:
            _check(self._ctx, int(n2))
        return offs, refs[:n]
