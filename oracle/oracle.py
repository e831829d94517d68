"""ctypes binding of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, never by the product package.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

F = C.POINTER(C.c_float)
U32 = C.POINTER(C.c_uint32)

SHADER_LAMBERT_ALBEDO, SHADER_BLINN_PHONG_REF, SHADER_VC_BLINN_PHONG, SHADER_TEXTURED_LAMBERT = 0, 1, 2, 3
FILTER_NEAREST, FILTER_LINEAR, FILTER_TRILINEAR = 0, 1, 2
FILTER_ANISO_4X, FILTER_ANISO_8X, FILTER_ANISO_16X = 3, 4, 5


class TextureC(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("rgba", C.c_void_p), ("color", C.c_float * 3)]


class Params(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("mv", C.c_float * 16), ("proj", C.c_float * 16),
                ("raster", C.c_float * 16), ("light", C.c_float * 3), ("shader", C.c_int32),
                ("shininess", C.c_float), ("ambient", C.c_float), ("kd", C.c_float), ("ks", C.c_float),
                ("albedo", C.c_float * 3), ("clamp_specular", C.c_int32), ("hierarchical", C.c_int32),
                ("num_threads", C.c_int32), ("mutation", C.c_int32), ("msaa_log2", C.c_int32),
                ("textures", C.POINTER(TextureC)), ("num_textures", C.c_uint32), ("tex_filter", C.c_int32)]


class MeshC(C.Structure):
    _fields_ = [("px", F), ("py", F), ("pz", F), ("nx", F), ("ny", F), ("nz", F), ("cr", F), ("cg", F),
                ("cb", F), ("indices", U32), ("num_vertices", C.c_uint32), ("num_triangles", C.c_uint32),
                ("tu", F), ("tv", F), ("tex_ids", U32)]


class Stats(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in ("triangles_in", "trivially_rejected", "clipped", "near_clipped",
                                            "clip_dropped", "setup_records", "culled", "refs", "fragments",
                                            "clip_vertices")] + \
               [(n, C.c_double) for n in ("ms_vs", "ms_clip", "ms_bin", "ms_raster", "ms_shade", "ms_write",
                                          "ms_total")] + [("threads", C.c_int32)]


class Record(C.Structure):
    _fields_ = [("x", C.c_int32 * 3), ("y", C.c_int32 * 3), ("inv_det", C.c_float), ("z", C.c_float * 3),
                ("inv_w", C.c_float * 3), ("tri_id", C.c_uint32), ("sub", C.c_uint32), ("corners", C.c_uint32)]


class Ref(C.Structure):
    _fields_ = [("record", C.c_uint32), ("flags", C.c_uint32)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle/liboracle.so is not built (make -C oracle)")
        l = C.CDLL(LIB_PATH)
        l.oracle_create.restype = C.c_void_p
        l.oracle_destroy.argtypes = [C.c_void_p]
        l.oracle_render.argtypes = [C.c_void_p, C.POINTER(Params), C.POINTER(MeshC), C.POINTER(Stats)]
        for n in ("oracle_read_color", "oracle_read_depth", "oracle_read_ids", "oracle_read_sample_depth",
                  "oracle_read_sample_ids"):
            getattr(l, n).argtypes = [C.c_void_p, C.c_void_p, C.c_uint64]
        l.oracle_num_records.argtypes = [C.c_void_p]
        l.oracle_num_records.restype = C.c_uint64
        l.oracle_read_records.argtypes = [C.c_void_p, C.c_void_p, C.c_uint64]
        l.oracle_num_tiles.argtypes = [C.c_void_p]
        l.oracle_num_tiles.restype = C.c_uint64
        l.oracle_read_bins.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64]
        l.oracle_mat_mul.argtypes = [F, F, F]
        l.oracle_mat_inverse.argtypes = [F, F]
        l.oracle_transform_point4.argtypes = [F, F, F]
        _lib = l
    return _lib


def _f(a):
    return np.ascontiguousarray(a, dtype=np.float32)


class Oracle:
    """One oracle renderer (Renderer::RenderMesh restated on the CPU)."""

    def __init__(self):
        self._ctx = C.c_void_p(lib().oracle_create())
        self._keep = None

    def close(self):
        if self._ctx:
            lib().oracle_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def render(self, mesh, width, height, mv, proj, raster, light=(1.0, 1.0, -1.0), shader=SHADER_VC_BLINN_PHONG,
               shininess=32.0, ambient=0.2, kd=3.0, ks=3.0, albedo=(0.9, 0.9, 0.9), clamp_specular=0,
               hierarchical=1, threads=0, mutation=0, msaa_log2=0, textures=(), tex_filter=FILTER_TRILINEAR):
        p = Params()
        p.width, p.height = width, height
        for name, m in (("mv", mv), ("proj", proj), ("raster", raster)):
            getattr(p, name)[:] = [float(v) for v in np.asarray(m, np.float32).reshape(16)]
        p.light[:] = [float(v) for v in light]
        p.shader, p.shininess, p.ambient, p.kd, p.ks = shader, shininess, ambient, kd, ks
        p.albedo[:] = [float(v) for v in albedo]
        p.clamp_specular, p.hierarchical, p.num_threads, p.mutation = clamp_specular, hierarchical, threads, mutation
        p.msaa_log2 = msaa_log2
        # texture slots: (H, W, 4) uint8 images or RGB colour triples
        tex = (TextureC * max(1, len(textures)))()
        imgs = []
        for k, t in enumerate(textures):
            if isinstance(t, np.ndarray):
                img = np.ascontiguousarray(t, dtype=np.uint8)
                imgs.append(img)
                tex[k].width, tex[k].height, tex[k].rgba = img.shape[1], img.shape[0], img.ctypes.data
            else:
                tex[k].color[:] = [float(c) for c in t]
        p.textures, p.num_textures, p.tex_filter = tex, len(textures), tex_filter
        planes = [_f(a) for a in (mesh.px, mesh.py, mesh.pz, mesh.nx, mesh.ny, mesh.nz, mesh.cr, mesh.cg, mesh.cb)]
        idx = np.ascontiguousarray(mesh.indices, dtype=np.uint32)
        uv = [None if getattr(mesh, n, None) is None else _f(getattr(mesh, n)) for n in ("tu", "tv")]
        tid = getattr(mesh, "tex_ids", None)
        tid = None if tid is None else np.ascontiguousarray(tid, dtype=np.uint32)
        self._keep = (planes, idx, uv, tid, tex, imgs)
        m = MeshC(*[a.ctypes.data_as(F) for a in planes], idx.ctypes.data_as(U32), planes[0].shape[0],
                  idx.shape[0] // 3, *[None if a is None else a.ctypes.data_as(F) for a in uv],
                  None if tid is None else tid.ctypes.data_as(U32))
        st = Stats()
        rc = lib().oracle_render(self._ctx, C.byref(p), C.byref(m), C.byref(st))
        if rc != 0:
            raise RuntimeError("oracle_render failed: %d" % rc)
        self.width, self.height, self.samples = width, height, 1 << msaa_log2
        return {n: getattr(st, n) for n, _ in Stats._fields_}

    def render_workload(self, w, **kw):
        m = w.material()
        if getattr(w, "textures", None) is not None:
            kw.setdefault("textures", w.textures)
            kw.setdefault("tex_filter", w.tex_filter)
        return self.render(w.mesh, w.width, w.height, w.mv, w.proj, w.raster, light=w.light, shader=w.shader,
                           shininess=m.shininess, ambient=m.ambient, kd=m.kd, ks=m.ks, albedo=tuple(m.albedo),
                           clamp_specular=m.clamp_specular, **kw)

    def color(self):
        out = np.empty((self.height, self.width, 4), np.uint8)
        assert lib().oracle_read_color(self._ctx, out.ctypes.data, out.nbytes) == 0
        return out

    def depth(self):
        out = np.empty((self.height, self.width), np.float32)
        assert lib().oracle_read_depth(self._ctx, out.ctypes.data, out.size) == 0
        return out

    def ids(self):
        out = np.empty((self.height, self.width), np.uint32)
        assert lib().oracle_read_ids(self._ctx, out.ctypes.data, out.size) == 0
        return out

    def sample_depth(self):
        """(samples, H, W) per-sample depth."""
        out = np.empty((self.samples, self.height, self.width), np.float32)
        assert lib().oracle_read_sample_depth(self._ctx, out.ctypes.data, out.size) == 0
        return out

    def sample_ids(self):
        """(samples, H, W) per-sample winning triangle IDs."""
        out = np.empty((self.samples, self.height, self.width), np.uint32)
        assert lib().oracle_read_sample_ids(self._ctx, out.ctypes.data, out.size) == 0
        return out

    def records(self):
        n = lib().oracle_num_records(self._ctx)
        arr = (Record * max(n, 1))()
        assert lib().oracle_read_records(self._ctx, arr, n) == 0
        return arr[:n]

    def records_array(self):
        """(N, 16) uint32 in the layout of edxr_debug_read_records (vertex ids omitted: 0)."""
        recs = self.records()
        out = np.zeros((len(recs), 16), np.uint32)
        for i, r in enumerate(recs):
            out[i, 0:6] = np.array([r.x[0], r.y[0], r.x[1], r.y[1], r.x[2], r.y[2]], np.int32).view(np.uint32)
            out[i, 6:10] = np.array([r.inv_det, r.z[0], r.z[1], r.z[2]], np.float32).view(np.uint32)
            out[i, 10], out[i, 11], out[i, 12] = r.tri_id, r.sub, r.corners
        return out

    def bins(self):
        """(offsets[num_tiles+1], refs (N,3) = tri_id, sub, flags) in the layout of edxr_debug_read_bins."""
        nt = lib().oracle_num_tiles(self._ctx)
        recs = self.records()
        offs = np.zeros(nt + 1, np.uint64)
        # capacity query: read offsets with a large buffer after sizing
        total_guess = 1 << 16
        while True:
            refs = (Ref * total_guess)()
            if lib().oracle_read_bins(self._ctx, offs.ctypes.data, refs, total_guess) == 0:
                break
            total_guess *= 4
        n = int(offs[-1])
        out = np.zeros((n, 3), np.uint32)
        for i in range(n):
            r = refs[i]
            rec = recs[r.record]
            out[i] = (rec.tri_id, rec.sub, r.flags)
        return offs, out


def mat_mul(a, b):
    out = np.zeros(16, np.float32)
    lib().oracle_mat_mul(_f(a).reshape(16).ctypes.data_as(F), _f(b).reshape(16).ctypes.data_as(F),
                         out.ctypes.data_as(F))
    return out.reshape(4, 4)
