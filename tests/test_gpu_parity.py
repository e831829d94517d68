"""GPU parity: the HIP pipeline (through the C ABI) against the CPU oracle.

Bar: triangle IDs, coverage, setup records and per-tile bins bit-exact; depth
bit-exact (the contract allows 1 ulp); 8-bit colour within 1 LSB (exact is
expected).  Inputs are the seeded synthetic scenes of edxraster_amd.configs at
sizes the oracle renders in seconds.
"""
import numpy as np
import pytest

from edxraster_amd import configs, Renderer
from edxraster_amd import _lib as L

pytestmark = pytest.mark.gpu


def render_both(w, hierarchical=True, threads=0):
    from oracle.oracle import Oracle
    r = Renderer(w.width, w.height)
    w.apply(r)
    r.set_hierarchical_rasterize(hierarchical)
    r.render_mesh(w.mesh)
    o = Oracle()
    ost = o.render_workload(w, hierarchical=1 if hierarchical else 0, threads=threads)
    return r, o, ost


def ulp_diff(a, b):
    ai = a.view(np.int32).astype(np.int64)
    bi = b.view(np.int32).astype(np.int64)
    ai = np.where(ai < 0, -(ai & 0x7FFFFFFF), ai)
    bi = np.where(bi < 0, -(bi & 0x7FFFFFFF), bi)
    return np.abs(ai - bi)


def assert_frame_parity(r, o, name):
    ids_g, ids_o = r.read_triangle_ids(), o.ids()
    mism = np.count_nonzero(ids_g != ids_o)
    assert mism == 0, "%s: %d triangle-ID mismatches" % (name, mism)
    d_g, d_o = r.read_depth(), o.depth()
    du = ulp_diff(d_g, d_o)
    assert du.max() <= 1, "%s: depth differs by %d ulp" % (name, du.max())
    c_g, c_o = r.get_back_buffer(), o.color()
    dc = np.abs(c_g.astype(np.int16) - c_o.astype(np.int16))
    assert dc.max() <= 1, "%s: colour differs by %d LSB at %d pixels" % (name, dc.max(), np.count_nonzero(dc))
    return {"covered": int(np.count_nonzero(ids_o != 0xFFFFFFFF)), "depth_exact": bool(du.max() == 0),
            "color_exact": bool(dc.max() == 0)}


WORKLOADS = [
    ("sphere_64x32", lambda: configs.small_sphere()),
    ("heightfield_64", lambda: configs.heightfield_small(64, 320, 240)),
    ("heightfield_200_hd", lambda: configs.heightfield_small(200, 1280, 720, seed=11)),
    ("sphere_grid_4x4_inside", lambda: configs.sphere_grid_1080p(640, 360, grid=4, segments=24)),
]


@pytest.mark.parametrize("name,make", WORKLOADS, ids=[w[0] for w in WORKLOADS])
def test_frame_parity(built, name, make):
    w = make()
    r, o, _ = render_both(w)
    info = assert_frame_parity(r, o, name)
    assert info["covered"] > 0
    r.close()


@pytest.mark.parametrize("name,make", WORKLOADS[:2], ids=[w[0] for w in WORKLOADS[:2]])
def test_records_and_bins_parity(built, name, make):
    w = make()
    r, o, _ = render_both(w)
    gr, orr = r.debug_records(), o.records_array()
    assert gr.shape[0] == orr.shape[0], "record count %d vs %d" % (gr.shape[0], orr.shape[0])
    np.testing.assert_array_equal(gr[:, :13], orr[:, :13])
    goffs, grefs = r.debug_bins()
    ooffs, orefs = o.bins()
    np.testing.assert_array_equal(goffs, ooffs)
    np.testing.assert_array_equal(grefs, orefs)
    r.close()


def test_hierarchical_off_parity(built):
    w = configs.heightfield_small(64, 320, 240)
    r, o, _ = render_both(w, hierarchical=False)
    assert_frame_parity(r, o, "hierarchical-off")
    r.close()


def test_shader_kinds_parity(built):
    w = configs.small_sphere(400, 300)
    for kind in (L.SHADER_LAMBERT_ALBEDO, L.SHADER_BLINN_PHONG_REF, L.SHADER_VC_BLINN_PHONG):
        w.shader = kind
        w.shininess = {0: 0.0, 1: 200.0, 2: 32.0}[kind]
        r, o, _ = render_both(w)
        assert_frame_parity(r, o, "shader %d" % kind)
        r.close()


FULL = [
    ("grid1m_1920x1080", lambda: configs.sphere_grid_1080p()),
    ("hf10m_3840x2160", lambda: configs.heightfield_4k()),
]


@pytest.mark.parametrize("name,make", FULL, ids=[w[0] for w in FULL])
def test_full_size_parity(built, name, make):
    """The benchmark workloads at full size against the oracle (16 host threads)."""
    w = make()
    r, o, ost = render_both(w, threads=16)
    info = assert_frame_parity(r, o, name)
    st = r.stats()
    assert st["setup_records"] == ost["setup_records"]
    assert st["refs"] == ost["refs"]
    assert st["clipped"] == ost["clipped"] and st["near_clipped"] == ost["near_clipped"]
    print(name, info, st)
    r.close()


def test_clipping_stats_and_parity(built):
    w = configs.sphere_grid_1080p(640, 360, grid=4, segments=24)
    r, o, ost = render_both(w)
    st = r.stats()
    assert st["clipped"] == ost["clipped"]
    assert st["near_clipped"] == ost["near_clipped"]
    assert st["setup_records"] == ost["setup_records"]
    assert st["refs"] == ost["refs"]
    assert st["clipped"] > 0
    r.close()


def test_empty_frame_after_a_full_one(built):
    """A frame without triangles is cleared tile by tile by the raster pass,
    over the previous frame's pixels: colour, depth, IDs and float colour."""
    import ctypes as C
    from edxraster_amd import Mesh
    w = configs.small_sphere(96, 64)
    r = Renderer(w.width, w.height)
    w.apply(r)
    assert L.lib().edxr_set_state(r.ctx, L.STATE_FLOAT_COLOR, 1) == 0
    r.render_mesh(w.mesh)
    assert (r.read_triangle_ids() != 0xFFFFFFFF).sum() > 1000
    r.render_mesh(Mesh(*w.mesh.planes(), w.mesh.indices[:0]))
    assert (r.read_triangle_ids() == 0xFFFFFFFF).all()
    assert (r.read_depth() == np.float32(1.0)).all()
    assert (r.get_back_buffer() == 0).all()
    f = np.full((w.height, w.width, 3), 7.0, np.float32)
    assert L.lib().edxr_copy_float_color(r.ctx, C.c_void_p(f.ctypes.data), f.nbytes) == 0
    assert (f == 0.0).all()
    r.close()


CLIP = ["lateral", "far", "all"]


@pytest.mark.parametrize("kind", CLIP + ["near_floor"])
def test_clip_scene_parity(built, kind):
    """Scenes straddling the side planes, the far plane (view depth 6 against
    geometry from 1.5 to 9), and all six planes at once (clip polygons of 7-8
    vertices): frame, setup records (clipped pieces included) and clip counts
    equal the oracle's."""
    import test_oracle_clip as TC
    w = dict(TC.WORKLOADS)[kind]
    r, o, ost = render_both(w)
    info = assert_frame_parity(r, o, "clip_" + kind)
    assert info["covered"] > 200
    st = r.stats()
    assert st["clipped"] == ost["clipped"] > 0
    assert st["near_clipped"] == ost["near_clipped"]
    assert st["setup_records"] == ost["setup_records"]
    assert st["clip_vertices"] == ost["clip_vertices"]
    gr, orr = r.debug_records(), o.records_array()
    assert gr.shape == orr.shape
    np.testing.assert_array_equal(gr[:, :13], orr[:, :13])
    r.close()


def wide_triangles(n=1200, width=1280, height=720, seed=23):
    """n large, overlapping triangles in front of the camera, each over many
    tiles: one binning chunk holds more wide records than the count and
    scatter kernels list for their waves, so the rest are enumerated by one
    thread each."""
    from edxraster_amd import Mesh
    rng = np.random.default_rng(seed)
    c = rng.uniform(-1.0, 1.0, (n, 1, 2))
    v = c + rng.uniform(-0.9, 0.9, (n, 3, 2))
    z = rng.uniform(-0.5, 0.5, (n, 3))
    px, py, pz = v[..., 0].ravel(), v[..., 1].ravel(), z.ravel()
    V = 3 * n
    nrm = np.tile(np.array([0.0, 0.0, -1.0], np.float32), (V, 1))
    col = rng.uniform(0.0, 1.0, (V, 3))
    idx = np.arange(V, dtype=np.uint32)
    mesh = Mesh(px, py, pz, nrm[:, 0], nrm[:, 1], nrm[:, 2], col[:, 0], col[:, 1], col[:, 2], idx)
    mv, p, r = configs.camera(width, height, (0.0, 0.0, -3.0), (0.0, 0.0, 0.0))
    return configs.Workload("wide_triangles_%d" % n, width, height, mesh, mv, p, r)


def test_wide_triangles_bins_parity(built):
    """Hundreds of wide records in one chunk: frame, records and bins equal
    the oracle's, at the default chunk size and with one chunk per 64 slots."""
    w = wide_triangles()
    r, o, ost = render_both(w)
    assert_frame_parity(r, o, "wide triangles")
    goffs, grefs = r.debug_bins()
    ooffs, orefs = o.bins()
    np.testing.assert_array_equal(goffs, ooffs)
    np.testing.assert_array_equal(grefs, orefs)
    wide = (orefs[:, 2] & 16) != 0
    assert wide.sum() > 8000, "too few tile-level references: %d" % wide.sum()
    assert r.stats()["setup_records"] > 512  # more wide records than one step lists
    r.set_tuning(64, 4)
    r.render_mesh(w.mesh)
    assert_frame_parity(r, o, "wide triangles, 64-slot chunks")
    goffs, grefs = r.debug_bins()
    np.testing.assert_array_equal(grefs, orefs)
    r.close()
