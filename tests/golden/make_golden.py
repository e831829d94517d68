"""Writes the golden fixtures of tests/test_golden.py (test infrastructure).

Each fixture is one small seeded scene rendered by the CPU oracle: the
scene's inputs are regenerated from the seeds in `CASES` by the tests, and
the fixture stores a digest of those inputs beside the expected outputs
(triangle IDs, depth, RGBA8 in GetBackBuffer layout; per-sample IDs and
depth for multisampled cases).  Run from the repository root after a
deliberate change of the oracle's semantics:

    python tests/golden/make_golden.py
"""
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))


def sphere_small():
    from edxraster_amd import configs
    return configs.small_sphere(200, 150)


def sphere_tiny():
    from edxraster_amd import configs
    return configs.small_sphere(120, 90)


def clip_all():
    from edxraster_amd import configs
    return configs.clip_scene("all", 160, 120)


def heightfield_blinn_s200():
    from edxraster_amd import configs
    from edxraster_amd import _lib as L
    w = configs.heightfield_small(40, 160, 120, seed=21)
    w.shader, w.shininess = L.SHADER_BLINN_PHONG_REF, 200.0
    return w


def textured_aniso8():
    from edxraster_amd import configs
    return configs.textured_scene(160, 120, n=24, tex_filter=4, sphere=(16, 8))


# name -> (workload builder, render keywords)
CASES = {
    "sphere_64x32_200x150": (sphere_small, {}),
    "clip_all_160x120": (clip_all, {}),
    "heightfield_40_blinn_s200_160x120": (heightfield_blinn_s200, {}),
    "sphere_64x32_120x90_msaa4": (sphere_tiny, {"msaa_log2": 2}),
    "textured_24_aniso8_160x120": (textured_aniso8, {}),
}


def input_digest(w):
    """SHA-256 of the scene's bytes (mesh planes, indices, uv, slots,
    matrices, light, shader, texture slots, filter)."""
    h = hashlib.sha256()
    m = w.mesh
    for a in m.planes() + [m.indices, m.tu, m.tv, m.tex_ids]:
        if a is not None:
            h.update(np.ascontiguousarray(a).tobytes())
    for a in (w.mv, w.proj, w.raster):
        h.update(np.ascontiguousarray(a, np.float32).tobytes())
    h.update(repr((tuple(w.light), w.shader, float(w.shininess), w.width, w.height, w.tex_filter)).encode())
    for t in (w.textures or []):
        h.update(np.ascontiguousarray(t).tobytes() if isinstance(t, np.ndarray) else repr(tuple(t)).encode())
    return h.hexdigest()


def render(name):
    from oracle.oracle import Oracle
    make, kw = CASES[name]
    w = make()
    o = Oracle()
    o.render_workload(w, threads=4, **kw)
    out = {"digest": np.array(input_digest(w)), "ids": o.ids(), "depth": o.depth(), "color": o.color()}
    if kw.get("msaa_log2"):
        out["sample_ids"] = o.sample_ids()
        out["sample_depth"] = o.sample_depth()
    return w, out


def main():
    for name in CASES:
        w, out = render(name)
        path = os.path.join(HERE, name + ".npz")
        np.savez_compressed(path, **out)
        print("%s: %d covered pixels, %d bytes" % (name, int((out["ids"] != 0xFFFFFFFF).sum()),
                                                   os.path.getsize(path)))


if __name__ == "__main__":
    main()
