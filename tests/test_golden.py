"""Committed golden fixtures (tests/golden/*.npz, written by
tests/golden/make_golden.py from the oracle).

- CPU: the scene generators still produce the inputs the fixtures were made
  from (SHA-256 digest), and the oracle still renders them bit for bit, so a
  change of the oracle's semantics cannot pass unnoticed.
- GPU: the HIP pipeline renders every fixture to the contract's bar (IDs
  exact, depth within 1 ulp, colour within 1 LSB) without the live oracle.
"""
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import make_golden as G  # noqa: E402

NAMES = sorted(G.CASES)


def load(name):
    with np.load(os.path.join(HERE, "golden", name + ".npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.mark.parametrize("name", NAMES)
def test_fixture_inputs_regenerate(built, name):
    make, _ = G.CASES[name]
    assert str(load(name)["digest"]) == G.input_digest(make()), "scene generation changed for " + name


@pytest.mark.parametrize("name", NAMES)
def test_oracle_reproduces_fixture(built, name):
    ref = load(name)
    w, out = G.render(name)
    for k in ("ids", "depth", "color") + (("sample_ids", "sample_depth") if "sample_ids" in ref else ()):
        assert np.array_equal(out[k].view(np.uint8), ref[k].view(np.uint8)), "%s: %s differs" % (name, k)
    assert (ref["ids"] != 0xFFFFFFFF).sum() > 1000


@pytest.mark.gpu
@pytest.mark.parametrize("name", NAMES)
def test_gpu_matches_fixture(built, name):
    from edxraster_amd import Renderer
    from test_gpu_parity import ulp_diff
    ref = load(name)
    make, kw = G.CASES[name]
    w = make()
    r = Renderer(w.width, w.height)
    r.set_msaa_mode(kw.get("msaa_log2", 0))
    w.apply(r)
    r.render_mesh(w.mesh)
    assert np.array_equal(r.read_triangle_ids(), ref["ids"]), name + ": triangle IDs differ"
    assert ulp_diff(r.read_depth(), ref["depth"]).max() <= 1, name + ": depth differs"
    dc = np.abs(r.get_back_buffer().astype(np.int16) - ref["color"].astype(np.int16))
    assert dc.max() <= 1, "%s: colour differs by %d LSB" % (name, dc.max())
    if "sample_ids" in ref:
        d, i = r.read_sample_buffers()
        assert np.array_equal(i, ref["sample_ids"]) and ulp_diff(d, ref["sample_depth"]).max() <= 1
    r.close()
