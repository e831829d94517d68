"""Failure handling: capacity overflow with re-run, bad indices, stale
backward, input checks of the differentiable wrapper.

The capacity path is forced from the host with EDXR_STATE_DEBUG_CAPACITY,
which starts every frame with room for only a few clip-queue entries,
clip-created records and vertices, and tile references: the kernels raise
their overflow flags, the host grows the buffers and re-runs the frame, and
the result must still equal the oracle and a normal render bit for bit.
"""
import ctypes as C

import numpy as np
import pytest

from edxraster_amd import configs, Mesh, Renderer
from edxraster_amd import _lib as L


def test_check_inputs_rejects_wrong_dtypes_and_devices():
    """CPU: the autograd wrapper refuses tensors the kernels would misread
    through raw pointers (int64 indices, float64 planes, host tensors)."""
    torch = pytest.importorskip("torch")
    from edxraster_amd.diff import check_inputs
    planes = [torch.zeros(6, dtype=torch.float32) for _ in range(9)]
    with pytest.raises(TypeError, match="int32"):
        check_inputs(torch.zeros(6, dtype=torch.int64), planes)
    with pytest.raises(ValueError, match="3 entries"):
        check_inputs(torch.zeros(5, dtype=torch.int32), planes)
    bad = list(planes)
    bad[4] = torch.zeros(6, dtype=torch.float64)
    with pytest.raises(TypeError, match="ny must be float32"):
        check_inputs(torch.zeros(6, dtype=torch.int32), bad)
    bad[4] = torch.zeros(5, dtype=torch.float32)
    with pytest.raises(ValueError, match="ny must be a flat tensor of 6"):
        check_inputs(torch.zeros(6, dtype=torch.int32), bad)
    with pytest.raises(ValueError, match="CUDA"):
        check_inputs(torch.zeros(6, dtype=torch.int32), planes)


def test_debug_capacity_key_range():
    """CPU: the state key is declared and its range is part of the ABI."""
    assert L.STATE_DEBUG_CAPACITY == 11
    src = open(__import__("os").path.join(__import__("os").path.dirname(__file__), "..", "include",
                                          "edxraster.h")).read()
    assert "EDXR_STATE_DEBUG_CAPACITY = 11" in src and "edxr_render_generation" in src


def _frame(r):
    return r.read_triangle_ids(), r.read_depth(), r.get_back_buffer()


def _assert_same(a, b, what):
    for x, y, n in zip(a, b, ("ids", "depth", "colour")):
        assert np.array_equal(x.view(np.uint8), y.view(np.uint8)), "%s: %s differ" % (what, n)


SCENES = [
    ("sphere_grid_4x4_inside", lambda: configs.sphere_grid_1080p(640, 360, grid=4, segments=24)),
    ("clip_all", lambda: configs.clip_scene("all")),
    ("heightfield_200_hd", lambda: configs.heightfield_small(200, 1280, 720, seed=11)),
]


@pytest.mark.gpu
@pytest.mark.parametrize("name,make", SCENES, ids=[s[0] for s in SCENES])
@pytest.mark.parametrize("graph", [0, 1], ids=["eager", "graph"])
def test_capacity_overflow_retries_equal_normal_render(built, name, make, graph):
    from oracle.oracle import Oracle
    import test_gpu_parity as G
    w = make()
    r = Renderer(w.width, w.height)
    w.apply(r)
    r.render_mesh(w.mesh)
    normal = _frame(r)
    assert r.stats()["retries"] == 0
    assert L.lib().edxr_set_state(r.ctx, L.STATE_GRAPH, graph) == 0
    assert L.lib().edxr_set_state(r.ctx, L.STATE_DEBUG_CAPACITY, 16) == 0
    for rep in range(2):  # the second frame starts from tiny capacities again
        r.render_mesh(w.mesh)
        st = r.stats()
        assert st["retries"] > 0, "no overflow was forced"
        _assert_same(_frame(r), normal, "%s frame %d" % (name, rep))
    o = Oracle()
    ost = o.render_workload(w)
    G.assert_frame_parity(r, o, name + " after retries")
    assert st["refs"] == ost["refs"] and st["setup_records"] == ost["setup_records"]
    # Back to automatic capacities: no retries.
    assert L.lib().edxr_set_state(r.ctx, L.STATE_DEBUG_CAPACITY, 0) == 0
    r.render_mesh(w.mesh)
    assert r.stats()["retries"] == 0
    _assert_same(_frame(r), normal, name + " after reset")
    r.close()


@pytest.mark.gpu
def test_capacity_overflow_in_a_batch(built):
    """Two jobs of a batch with the capacity key: the re-run batch equals the
    batch rendered with automatic capacities."""
    jobs = [configs.clip_scene("far"), configs.sphere_grid_1080p(320, 180, grid=4, segments=16)]
    r = Renderer(64, 64)
    ej = [w.job() for w in jobs]
    r.render_batch(ej)
    ref = [r.batch_read(k) for k in range(2)]
    assert L.lib().edxr_set_state(r.ctx, L.STATE_DEBUG_CAPACITY, 8) == 0
    r.render_batch(ej)
    assert r.stats()["retries"] > 0
    for k in range(2):
        got = r.batch_read(k)
        for a, b in zip(got, ref[k]):
            assert np.array_equal(a, b)
    r.close()


@pytest.mark.gpu
def test_out_of_range_vertex_index_is_reported(built):
    """An index >= num_vertices: EDXR_E_INDEX with a message, and the context
    renders the next valid mesh normally."""
    w = configs.small_sphere(160, 120)
    r = Renderer(w.width, w.height)
    w.apply(r)
    bad_idx = w.mesh.indices.copy()
    bad_idx[3 * 100 + 1] = w.mesh.num_vertices  # one past the end
    bad = Mesh(*w.mesh.planes(), bad_idx)
    with pytest.raises(L.EdxrError) as e:
        r.render_mesh(bad)
    assert e.value.code == L.EDXR_E_INDEX
    assert "index" in str(e.value)
    bad_idx[3 * 100 + 1] = 0xFFFFFFFF
    with pytest.raises(L.EdxrError) as e:
        r.render_mesh(Mesh(*w.mesh.planes(), bad_idx))
    assert e.value.code == L.EDXR_E_INDEX
    r.render_mesh(w.mesh)
    from oracle.oracle import Oracle
    import test_gpu_parity as G
    o = Oracle()
    o.render_workload(w)
    G.assert_frame_parity(r, o, "after a bad index")
    r.close()


@pytest.mark.gpu
def test_backward_after_another_render_is_refused(built):
    """Two forwards on one renderer, then backward of the first image: the
    wrapper raises instead of differentiating the second frame."""
    import torch
    from edxraster_amd import diff
    mesh, views = configs.diff_scene(n_views=2, res=64, hf_n=9, sphere_segments=(8, 6))
    idx, planes = diff.mesh_tensors(mesh)
    for p in planes:
        p.requires_grad_(True)
    light = torch.tensor([1.0, 1.0, -1.0], requires_grad=True)
    mat = torch.tensor([0.2, 3.0, 3.0, 32.0], requires_grad=True)
    r = Renderer(64, 64)
    img1 = diff.render_views(r, views, idx, *planes, light, mat)
    img2 = diff.render_views(r, views[:1], idx, *planes, light, mat)
    with pytest.raises(RuntimeError, match="rendered again"):
        img1.sum().backward()
    img2.sum().backward()  # the last render differentiates normally
    assert planes[6].grad is not None and torch.isfinite(planes[6].grad).all()
    r.close()


@pytest.mark.gpu
def test_render_generation_counts_renders(built):
    w = configs.small_sphere(64, 48)
    r = Renderer(w.width, w.height)
    w.apply(r)
    g0 = L.lib().edxr_render_generation(r.ctx)
    r.render_mesh(w.mesh)
    r.render_mesh(w.mesh)
    assert L.lib().edxr_render_generation(r.ctx) == g0 + 2
    r.close()
