"""Batched rendering (edxr_render_batch): many independent jobs of uneven mesh
sizes, resolutions and cameras in one launch sequence.

GPU bar: every job of a batch equals the CPU oracle's render of that job
alone (IDs and depth bit-exact, colour within 1 LSB), and the full 256-job
batch equals 256 single-job renders bit for bit and, on nine sampled jobs,
the oracle.  The CPU part checks the
job generator.
"""
import numpy as np
import pytest

from edxraster_amd import configs, Renderer


def small_batch():
    """Eleven jobs (spheres and heightfields, 300-30k triangles, 3 sizes) plus
    an empty mesh and a camera inside a sphere (near-plane clipping)."""
    jobs = configs.batch_jobs(10, seed_base=777, tmin=300, tmax=3e4, resolutions=(96, 256, 320))
    inside = configs.sphere_grid_1080p(200, 120, grid=2, segments=16)
    inside.name = "inside_sphere_200x120"
    jobs.insert(3, inside)
    empty = configs.small_sphere(64, 48)
    empty.mesh.indices = empty.mesh.indices[:0]
    empty.name = "empty_64x48"
    jobs.insert(6, empty)
    return jobs


def test_batch_generator_is_deterministic():
    a = configs.batch_jobs(6, seed_base=4000)
    b = configs.batch_jobs(6, seed_base=4000)
    for x, y in zip(a, b):
        assert x.name == y.name and x.width == y.width
        assert np.array_equal(x.mesh.indices, y.mesh.indices)
        assert np.array_equal(x.mesh.px, y.mesh.px) and np.array_equal(x.mv, y.mv)
    sizes = {w.width for w in configs.batch_jobs(40, seed_base=4000)}
    assert sizes == {256, 512, 1024}


@pytest.mark.gpu
def test_batch_matches_oracle_per_job(built):
    from oracle.oracle import Oracle
    jobs = small_batch()
    r = Renderer(64, 64)
    r.render_batch([w.job() for w in jobs])
    o = Oracle()
    for k, w in enumerate(jobs):
        rgba, depth, ids = r.batch_read(k)
        o.render_workload(w)
        assert np.array_equal(ids, o.ids()), "%s: triangle IDs differ" % w.name
        assert np.array_equal(depth.view(np.uint32), o.depth().view(np.uint32)), "%s: depth differs" % w.name
        dc = np.abs(rgba.astype(np.int16) - o.color().astype(np.int16)).max()
        assert dc <= 1, "%s: colour differs by %d LSB" % (w.name, dc)
    st = r.stats()
    assert st["triangles_in"] == sum(w.mesh.num_triangles for w in jobs)
    assert st["clipped"] > 0
    r.close()


@pytest.mark.gpu
def test_batch_graph_replays_follow_camera_changes(built):
    jobs = small_batch()
    eager = Renderer(64, 64)
    graph = Renderer(64, 64)
    graph.set_graph(True)
    for step in range(3):
        # a new light and a moved eye per step: the replay reads the job table
        for k, w in enumerate(jobs):
            ang = 0.3 * step + 0.1 * k
            w.light = (float(np.cos(ang)), 1.0, float(np.sin(ang)))
            eye = 3.0 * np.array([np.cos(ang), 0.6, np.sin(ang)]) / np.sqrt(1.36)
            w.mv, w.proj, w.raster = configs.camera(w.width, w.height, eye, (0.0, 0.0, 0.0))
        jl = [w.job() for w in jobs]
        eager.render_batch(jl)
        graph.render_batch(jl)
        for k, w in enumerate(jobs):
            a, b = eager.batch_read(k), graph.batch_read(k)
            for x, y in zip(a, b):
                assert np.array_equal(x, y), "%s step %d: graph replay differs" % (w.name, step)
    eager.close()
    graph.close()


@pytest.mark.gpu
def test_single_job_graph_replays_follow_camera_changes(built):
    """A single frame replayed from a captured graph takes each call's job
    (camera, light; the graph's first kernel gets it by value, its arguments
    updated between replays): each replay equals the eager frame."""
    w = configs.sphere_grid_1080p(320, 240, grid=4, segments=24)
    eager = Renderer(w.width, w.height)
    graph = Renderer(w.width, w.height)
    graph.set_graph(True)
    for step in range(4):
        ang = 0.4 * step
        w.light = (float(np.cos(ang)), 1.0, float(np.sin(ang)))
        eye = (float(2.5 * np.sin(ang)), 1.5, float(-6.0 + 0.5 * step))
        w.mv, w.proj, w.raster = configs.camera(w.width, w.height, eye, (0.0, 0.0, 0.0))
        for r in (eager, graph):
            w.apply(r)
            r.render_mesh(w.mesh)
        assert np.array_equal(eager.get_back_buffer(), graph.get_back_buffer()), "step %d: colour" % step
        assert np.array_equal(eager.read_depth(), graph.read_depth()), "step %d: depth" % step
        assert np.array_equal(eager.read_triangle_ids(), graph.read_triangle_ids()), "step %d: IDs" % step
    eager.close()
    graph.close()


@pytest.mark.gpu
def test_single_job_graph_replays_follow_mesh_changes(built):
    """Two device meshes of the same size (same launch layout, so the graph
    is not captured again) alternated under graph replay: the first kernel's
    by-value job carries the mesh pointers, and each replay equals the eager
    frame of the mesh it was given."""
    import torch
    from bench import device_mesh
    from edxraster_amd.renderer import Mesh
    w = configs.sphere_grid_1080p(320, 240, grid=4, segments=24)
    a = w.mesh
    b = Mesh(a.px + 0.35, a.py - 0.2, a.pz, a.nx, a.ny, a.nz, a.cb, a.cr, a.cg, a.indices)
    da, ka = device_mesh(a)
    db, kb = device_mesh(b)
    eager = Renderer(w.width, w.height)
    graph = Renderer(w.width, w.height)
    graph.set_graph(True)
    for r in (eager, graph):
        w.apply(r)
    depths = []
    for step, m in enumerate((da, db, da, db)):
        eager.render_mesh(m)
        graph.render_mesh(m)
        assert np.array_equal(eager.read_triangle_ids(), graph.read_triangle_ids()), "step %d: IDs" % step
        assert np.array_equal(eager.read_depth(), graph.read_depth()), "step %d: depth" % step
        assert np.array_equal(eager.get_back_buffer(), graph.get_back_buffer()), "step %d: colour" % step
        depths.append(eager.read_depth())
    # the two meshes give different frames, so a stale mesh would be caught
    assert not np.array_equal(depths[0], depths[1])
    eager.close()
    graph.close()
    torch.cuda.synchronize()


@pytest.mark.gpu
def test_autotune_picks_a_candidate_and_keeps_results(built):
    jobs = small_batch()
    r = Renderer(64, 64)
    jl = [w.job() for w in jobs]
    r.render_batch(jl)
    before = [r.batch_read(k) for k in range(len(jobs))]
    t = r.autotune(jl)
    assert t["tris_per_chunk"] in (512, 1024, 2048, 4096)
    assert t["raster_blocks_per_cu"] in (4, 8, 16)
    assert t["frame_ms"] > 0
    r.render_batch(jl)
    for k in range(len(jobs)):
        for x, y in zip(before[k], r.batch_read(k)):
            assert np.array_equal(x, y)
    r.close()


@pytest.mark.gpu
def test_full_batch_equals_single_job_renders(built):
    """The 256-job batch (19.2M triangles) against 256 plain renders, and
    nine of its jobs (every 32nd and the one with the most triangles)
    against the oracle."""
    import torch
    from bench import device_mesh
    from oracle.oracle import Oracle
    jobs = configs.batch_jobs()
    keep, dm = [], []
    for w in jobs:
        m, k = device_mesh(w.mesh)
        dm.append(m)
        keep.append(k)
    r = Renderer(64, 64)
    r.render_batch([w.job(mesh=m) for w, m in zip(jobs, dm)])
    single = Renderer(64, 64)
    for k, w in enumerate(jobs):
        single.resize(w.width, w.height)
        w.apply(single)
        single.render_mesh(dm[k])
        rgba, depth, ids = r.batch_read(k)
        assert np.array_equal(ids, single.read_triangle_ids()), w.name
        assert np.array_equal(depth.view(np.uint32), single.read_depth().view(np.uint32)), w.name
        assert np.array_equal(rgba, single.get_back_buffer()), w.name
    o = Oracle()
    biggest = max(range(len(jobs)), key=lambda k: jobs[k].mesh.num_triangles)
    for k in sorted(set(range(0, len(jobs), 32)) | {biggest}):
        w = jobs[k]
        rgba, depth, ids = r.batch_read(k)
        o.render_workload(w)
        assert np.array_equal(ids, o.ids()), "%s: triangle IDs differ from the oracle" % w.name
        assert np.array_equal(depth.view(np.uint32), o.depth().view(np.uint32)), "%s: depth differs" % w.name
        dc = np.abs(rgba.astype(np.int16) - o.color().astype(np.int16)).max()
        assert dc <= 1, "%s: colour differs from the oracle by %d LSB" % (w.name, dc)
    r.close()
    single.close()
    torch.cuda.synchronize()
