"""Differentiable path: the hand-written HIP adjoint (edxr_backward) against
an FP64 CPU reference gradient (torch float64 reverse mode of the same
fixed-visibility image function, tests/adjoint_ref.py), which is itself
pinned by FP64 central finite differences (h = 1e-4).

Bar: per parameter group (positions, normals, colours, light, material) the
relative L2 error of the GPU gradient against the FP64 reference is <= 1e-3.
"""
import numpy as np
import pytest
import torch

from edxraster_amd import configs
from tests import adjoint_ref as R

MATERIAL = (0.2, 3.0, 3.0, 32.0)
GROUPS = {"positions": ("px", "py", "pz"), "normals": ("nx", "ny", "nz"), "colours": ("cr", "cg", "cb"),
          "light": ("light",), "material": ("material",)}


def small_scene(n_views=2, res=96):
    mesh, views = configs.diff_scene(n_views, res, hf_n=12, sphere_segments=(12, 8))
    return mesh, views


def oracle_ids(views):
    from oracle.oracle import Oracle
    o = Oracle()
    out = []
    for v in views:
        o.render_workload(v)
        out.append(o.ids())
    return out


def upstream(views, seed=9):
    rng = np.random.default_rng(seed)
    return [rng.normal(size=(v.height, v.width, 3)) for v in views]


def reference_grads(mesh, views, ids, g):
    th = R.theta_from(mesh, views[0].light, MATERIAL)
    R.loss_fp64(th, mesh.indices, views, ids, g).backward()
    return th, {k: v.grad.numpy() for k, v in th.items()}


def test_fp64_reference_matches_central_differences(built):
    mesh, views = small_scene()
    ids = oracle_ids(views)
    g = upstream(views)
    th, grads = reference_grads(mesh, views, ids, g)
    rng = np.random.default_rng(3)
    h = 1e-4
    checked = 0
    for name in ("px", "py", "pz", "nx", "ny", "nz", "cr", "cg", "cb", "light", "material"):
        n = th[name].numel()
        for i in rng.choice(n, size=min(n, 6), replace=False):
            def f(delta):
                t2 = {k: v.detach().clone() for k, v in th.items()}
                t2[name][i] += delta
                return float(R.loss_fp64(t2, mesh.indices, views, ids, g))
            fd = (f(h) - f(-h)) / (2 * h)
            an = grads[name].reshape(-1)[i]
            assert abs(fd - an) <= 1e-5 + 1e-3 * max(abs(fd), abs(an)), (name, i, fd, an)
            checked += 1
    assert checked >= 50


def test_fp64_reference_colour_is_the_oracle_shading(built):
    """The reference image function evaluates the oracle's shader (up to the
    28.4 snap of the barycentrics): 8-bit colours agree closely."""
    from oracle.oracle import Oracle
    mesh, views = small_scene(1, 128)
    o = Oracle()
    o.render_workload(views[0])
    th = R.theta_from(mesh, views[0].light, MATERIAL, requires_grad=False)
    col, (rows, xs) = R.image_fp64(th, mesh.indices, views[0], o.ids())
    ref = o.color()[rows, xs, :3].astype(np.float64) / 255.0
    err = np.abs(col.numpy() - ref)
    assert err.mean() < 0.01 and np.median(err) < 0.004, (err.mean(), np.median(err))


def gpu_grads(mesh, views, g):
    from edxraster_amd import Renderer, diff
    idx, planes = diff.mesh_tensors(mesh)
    for t in planes:
        t.requires_grad_(True)
    light = torch.tensor(views[0].light, dtype=torch.float32, requires_grad=True)
    material = torch.tensor(MATERIAL, dtype=torch.float32, requires_grad=True)
    r = Renderer(64, 64)
    img = diff.render_views(r, views, idx, *planes, light, material)
    gt = torch.from_numpy(np.stack(g).astype(np.float32)).cuda()
    (img * gt).sum().backward()
    out = {k: t.grad.detach().cpu().numpy().astype(np.float64) for k, t in
           zip(("px", "py", "pz", "nx", "ny", "nz", "cr", "cg", "cb"), planes)}
    out["light"] = light.grad.numpy().astype(np.float64)
    out["material"] = material.grad.numpy().astype(np.float64)
    ids = []
    for k in range(len(views)):
        ids.append(r.batch_read(k)[2])
    fimg = img.detach().cpu().numpy()
    r.close()
    return out, ids, fimg


def rel_err(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


@pytest.mark.gpu
@pytest.mark.parametrize("n_views,res,hf_n,seg,stride", [(2, 96, 12, (12, 8), 1), (4, 256, 71, (80, 64), 1),
                                                         (64, 512, 71, (80, 64), 8)],
                         ids=["tiny", "c5_scene_4views_256", "c5_scene_8of64views_512"])
def test_gpu_adjoint_matches_fp64_reference(built, n_views, res, hf_n, seg, stride):
    """GPU gradients against the FP64 reference, within 1e-3 relative L2 error
    per parameter group; the last case is the differentiable workload's own
    scene and resolution (every 8th of its 64 views at 512 x 512)."""
    mesh, views = configs.diff_scene(n_views, res, hf_n=hf_n, sphere_segments=seg)
    views = views[::stride]
    g = upstream(views)
    gg, ids, fimg = gpu_grads(mesh, views, g)
    # visibility of the GPU forward equals the oracle's
    for a, b in zip(ids, oracle_ids(views)):
        assert np.array_equal(a, b)
    _, ref = reference_grads(mesh, views, ids, g)
    for group, names in GROUPS.items():
        a = np.concatenate([gg[n].reshape(-1) for n in names])
        b = np.concatenate([ref[n].reshape(-1) for n in names])
        assert rel_err(a, b) <= 1e-3, (group, rel_err(a, b))


@pytest.mark.gpu
def test_gpu_float_image_is_the_byte_image(built):
    """The float RGB output quantises to the 8-bit colour buffer."""
    from edxraster_amd import Renderer
    from edxraster_amd import _lib as L
    mesh, views = configs.diff_scene(2, 128)
    r = Renderer(64, 64)
    assert L.lib().edxr_set_state(r.ctx, L.STATE_FLOAT_COLOR, 1) == 0
    r.render_batch([v.job() for v in views])
    f = np.empty((2, 128, 128, 3), np.float32)
    assert L.lib().edxr_copy_float_color(r.ctx, f.ctypes.data, f.nbytes) == 0
    for k in range(2):
        rgba = r.batch_read(k)[0]
        q = (f[k] * np.float32(255.0) + np.float32(0.5)).astype(np.int32)
        assert np.array_equal(q, rgba[..., :3].astype(np.int32))
    r.close()
