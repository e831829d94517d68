"""Build recipe for the gfx950 library.

``build_library()`` compiles ``csrc/*.hip`` and ``csrc/*.cpp`` with hipcc for
gfx950 into ``edxraster_amd/libedxraster.so`` (in-tree, so the built library
travels with the source tree).
"""
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "libedxraster.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# -ffp-contract=off: no FMA contraction (the reference's arithmetic has none).
# -fwrapv: int32 edge functions wrap like the reference's SSE integer ops.
# The packed-fp32 feature is disabled: v_pk_*_f32 are never emitted.
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-fwrapv",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def sources():
    out = []
    for name in sorted(os.listdir(CSRC)):
        if name.endswith((".hip", ".cpp")):
            out.append(os.path.join(CSRC, name))
    return out


def build_library(verbose=False):
    tmp = LIB + ".tmp"
    cmd = [HIPCC] + HIP_FLAGS + ["-I", os.path.join(ROOT, "include"), "-o", tmp] + sources() + ["-pthread"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build_library(verbose=True)
    print("built", LIB, file=sys.stderr)
