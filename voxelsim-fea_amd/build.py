"""Build ``lib/libvxfea.so`` for gfx950 with hipcc (no GPU needed to build).

    python voxelsim-fea_amd/build.py [--debug]
"""

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("VX_OFFLOAD_ARCH", "gfx950")

SOURCES = [os.path.join(HERE, "csrc", f) for f in ("vxfea.hip", "post.hip", "net.hip", "inside.hip")]
DEPS = [os.path.join(HERE, "csrc", f) for f in ("kernels.h", "elem.h", "parts.h")] + [
    os.path.join(ROOT, "include", "vxfea.h")]
OUT = os.path.join(HERE, "lib", "libvxfea.so")

FLAGS = [
    f"--offload-arch={ARCH}", "-std=c++17", "-fPIC", "-shared",
    # packed FP32 VALU ops stay off: scalar/float2 code must not be turned
    # into v_pk_* instructions on this target
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(HERE, "csrc"),
]


def build(debug=False, verbose=True):
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    cmd = [HIPCC] + FLAGS + (["-O1", "-g"] if debug else ["-O3"]) + SOURCES + ["-o", OUT + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    noise = "packed-fp32-ops"
    msgs = [l for l in (res.stdout + res.stderr).splitlines() if noise not in l and l.strip()]
    if res.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + "\n".join(msgs))
    if verbose and msgs:
        print("\n".join(msgs))
    os.replace(OUT + ".tmp", OUT)
    return OUT


if __name__ == "__main__":
    print(build(debug="--debug" in sys.argv))
