"""ctypes binding of libedxraster.so (the C ABI declared in include/edxraster.h).

The library is loaded from this package directory only; a missing library is
an error (there is no CPU fallback for any render entry point).
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libedxraster.so")

EDXR_OK = 0
EDXR_E_INVALID = -1
EDXR_E_HIP = -2
EDXR_E_NOMEM = -3
EDXR_E_INDEX = -4
EDXR_E_LIMIT = -5
EDXR_E_STATE = -6
EDXR_E_IO = -7
EDXR_E_FORMAT = -8

SHADER_LAMBERT_ALBEDO = 0
SHADER_BLINN_PHONG_REF = 1
SHADER_VC_BLINN_PHONG = 2
SHADER_TEXTURED_LAMBERT = 3

FILTER_NEAREST = 0
FILTER_LINEAR = 1
FILTER_TRILINEAR = 2
FILTER_ANISO_4X = 3
FILTER_ANISO_8X = 4
FILTER_ANISO_16X = 5

STATE_HIERARCHICAL = 1
STATE_BIN_CHUNKS = 2
STATE_PROFILE = 3
STATE_GRAPH = 4
STATE_TRIS_PER_CHUNK = 5
STATE_RASTER_BLOCKS_PER_CU = 6
STATE_FLOAT_COLOR = 7
STATE_MSAA = 8
STATE_WRITE_FRAMES = 9
STATE_TEXTURE_FILTER = 10
STATE_DEBUG_CAPACITY = 11
STAGES = ("setup", "bin_count", "bin_scan", "bin_scatter", "raster")

MEM_HOST = 0
MEM_DEVICE = 1

F = C.POINTER(C.c_float)
U32 = C.POINTER(C.c_uint32)


class Material(C.Structure):
    _fields_ = [("ambient", C.c_float), ("kd", C.c_float), ("ks", C.c_float), ("shininess", C.c_float),
                ("albedo", C.c_float * 3), ("clamp_specular", C.c_int32)]


class Mesh(C.Structure):
    _fields_ = [("px", F), ("py", F), ("pz", F), ("nx", F), ("ny", F), ("nz", F),
                ("cr", F), ("cg", F), ("cb", F), ("indices", U32),
                ("num_vertices", C.c_uint32), ("num_triangles", C.c_uint32), ("memory", C.c_int32),
                ("tu", F), ("tv", F), ("tex_ids", U32)]


class Texture(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("rgba", C.c_void_p), ("color", C.c_float * 3)]


class Stats(C.Structure):
    _fields_ = [("triangles_in", C.c_uint64), ("trivially_rejected", C.c_uint64), ("clipped", C.c_uint64),
                ("near_clipped", C.c_uint64), ("setup_records", C.c_uint64), ("refs", C.c_uint64),
                ("clip_vertices", C.c_uint64), ("retries", C.c_uint32), ("bin_chunks", C.c_uint32)]


class Job(C.Structure):
    _fields_ = [("mesh", Mesh), ("width", C.c_uint32), ("height", C.c_uint32),
                ("mv", C.c_float * 16), ("proj", C.c_float * 16), ("raster", C.c_float * 16),
                ("light", C.c_float * 3), ("shader", C.c_int32), ("material", Material),
                ("hierarchical", C.c_int32)]


class Grads(C.Structure):
    _fields_ = [(n, F) for n in ("d_px", "d_py", "d_pz", "d_nx", "d_ny", "d_nz", "d_cr", "d_cg", "d_cb",
                                 "d_light", "d_material")]


class Tuning(C.Structure):
    _fields_ = [("tris_per_chunk", C.c_uint32), ("raster_blocks_per_cu", C.c_uint32), ("frame_ms", C.c_double)]


class HostMesh(C.Structure):
    _fields_ = [("px", F), ("py", F), ("pz", F), ("nx", F), ("ny", F), ("nz", F),
                ("cr", F), ("cg", F), ("cb", F), ("indices", U32),
                ("num_vertices", C.c_uint32), ("num_triangles", C.c_uint32),
                ("tu", F), ("tv", F), ("tex_ids", U32)]


class ObjMaterial(C.Structure):
    _fields_ = [("name", C.c_char * 128), ("kd", C.c_float * 3), ("map_kd", C.c_char * 512)]


# (name, restype, argtypes) for every symbol of include/edxraster.h.
SIGNATURES = [
    ("edxr_create", C.c_int, [C.POINTER(C.c_void_p), C.c_uint32, C.c_uint32]),
    ("edxr_destroy", None, [C.c_void_p]),
    ("edxr_resize", C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32]),
    ("edxr_set_transform", C.c_int, [C.c_void_p, F, F, F]),
    ("edxr_set_light", C.c_int, [C.c_void_p, F]),
    ("edxr_set_pixel_shader", C.c_int, [C.c_void_p, C.c_int32, C.POINTER(Material)]),
    ("edxr_default_material", None, [C.c_int32, C.POINTER(Material)]),
    ("edxr_set_state", C.c_int, [C.c_void_p, C.c_int32, C.c_int32]),
    ("edxr_render_mesh", C.c_int, [C.c_void_p, C.POINTER(Mesh)]),
    ("edxr_render_mesh_noalloc", C.c_int, [C.c_void_p, C.POINTER(Mesh), C.c_void_p, C.c_void_p, C.c_void_p]),
    ("edxr_read_color", C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    ("edxr_read_depth", C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    ("edxr_read_triangle_ids", C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    ("edxr_read_sample_buffers", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t]),
    ("edxr_device_buffers", C.c_int, [C.c_void_p, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p),
                                      C.POINTER(C.c_void_p)]),
    ("edxr_get_stats", C.c_int, [C.c_void_p, C.POINTER(Stats)]),
    ("edxr_profile_read", C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_uint64)]),
    ("edxr_profile_reset", C.c_int, [C.c_void_p]),
    ("edxr_stream", C.c_void_p, [C.c_void_p]),
    ("edxr_render_generation", C.c_uint64, [C.c_void_p]),
    ("edxr_last_error", C.c_char_p, [C.c_void_p]),
    ("edxr_render_batch", C.c_int, [C.c_void_p, C.POINTER(Job), C.c_uint32]),
    ("edxr_batch_read", C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("edxr_batch_device_buffers", C.c_int, [C.c_void_p, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p),
                                            C.POINTER(C.c_void_p), C.POINTER(C.c_uint64)]),
    ("edxr_autotune", C.c_int, [C.c_void_p, C.POINTER(Job), C.c_uint32, C.POINTER(Tuning)]),
    ("edxr_float_color", C.c_int, [C.c_void_p, C.POINTER(C.c_void_p)]),
    ("edxr_copy_float_color", C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    ("edxr_backward", C.c_int, [C.c_void_p, F, C.POINTER(Grads), C.c_uint32]),
    ("edxr_debug_read_records", C.c_int64, [C.c_void_p, C.c_void_p, C.c_uint64]),
    ("edxr_debug_read_bins", C.c_int64, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64]),
    ("edxr_scene_sphere", C.c_int, [C.POINTER(HostMesh), C.c_uint32, C.c_uint32, C.c_float, F, C.c_uint64]),
    ("edxr_scene_heightfield", C.c_int, [C.POINTER(HostMesh), C.c_uint32, C.c_float, C.c_uint64, C.c_uint64]),
    ("edxr_scene_sphere_grid", C.c_int, [C.POINTER(HostMesh), C.c_uint32, C.c_uint32, C.c_uint64]),
    ("edxr_scene_merge", C.c_int, [C.POINTER(HostMesh), C.POINTER(HostMesh)]),
    ("edxr_load_sphere", C.c_int, [F, F, F, C.c_float, C.c_uint32, C.c_uint32, C.POINTER(HostMesh)]),
    ("edxr_load_plane", C.c_int, [F, F, F, C.c_float, C.POINTER(HostMesh)]),
    ("edxr_scene_free", None, [C.POINTER(HostMesh)]),
    ("edxr_load_obj", C.c_int, [C.c_char_p, F, F, F, C.POINTER(HostMesh)]),
    ("edxr_write_bmp", C.c_int, [C.c_char_p, C.c_void_p, C.c_uint32, C.c_uint32]),
    ("edxr_obj_materials", C.c_int64, [C.c_char_p, C.POINTER(ObjMaterial), C.c_uint32]),
    ("edxr_load_bmp", C.c_int, [C.c_char_p, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                                C.POINTER(C.POINTER(C.c_uint8))]),
    ("edxr_load_tga", C.c_int, [C.c_char_p, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                                C.POINTER(C.POINTER(C.c_uint8))]),
    ("edxr_load_ppm", C.c_int, [C.c_char_p, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                                C.POINTER(C.POINTER(C.c_uint8))]),
    ("edxr_load_image", C.c_int, [C.c_char_p, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                                  C.POINTER(C.POINTER(C.c_uint8))]),
    ("edxr_free", None, [C.c_void_p]),
    ("edxr_set_frame_directory", C.c_int, [C.c_void_p, C.c_char_p]),
    ("edxr_set_textures", C.c_int, [C.c_void_p, C.POINTER(Texture), C.c_uint32]),
    ("edxr_write_frame_to_file", C.c_int, [C.c_void_p]),
    ("edxr_frame_count", C.c_int64, [C.c_void_p]),
    ("edxr_look_at", None, [F, F, F, F]),
    ("edxr_perspective", None, [C.c_float, C.c_float, C.c_float, C.c_float, F]),
    ("edxr_raster_matrix", None, [C.c_uint32, C.c_uint32, F]),
]

_lib = None


def _share_torch_runtime():
    """PyTorch wheels bundle a HIP runtime whose soname (libamdhip64.so.7) is
    the one this library links against.  Importing torch first makes the
    library bind to that copy, so the process holds a single HIP runtime; the
    other order makes torch load a second runtime, which then sees no device."""
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


def lib():
    """Loads libedxraster.so once; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                "libedxraster.so is not built (run __graft_entry__.build() or python edxraster_amd/build.py)")
        _share_torch_runtime()
        l = C.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            fn = getattr(l, name)
            fn.restype = res
            fn.argtypes = args
        _lib = l
    return _lib


class EdxrError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("edxraster error %d: %s" % (code, msg))
        self.code = code


def fptr(a):
    return a.ctypes.data_as(F) if a is not None else None


def uptr(a):
    return a.ctypes.data_as(U32) if a is not None else None
