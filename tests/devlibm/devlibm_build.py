"""TEST INFRASTRUCTURE ONLY: builds tests/devlibm/libdevlibm.so (the core's libm entry points on the GPU, devlibm.hip) with
the flags of the product's own HIP build (rsim/build.py), so the functions compile exactly as they do inside the tick."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
LIB = os.path.join(HERE, "libdevlibm.so")


def build():
    pkg = os.path.join(ROOT, "rocketsim-cuda_amd")
    if pkg not in sys.path:
        sys.path.insert(0, pkg)
    from rsim import build as rb

    tmp = LIB + ".tmp"
    cmd = [rb._hipcc()] + rb.COMMON + rb.HIP_FLAGS + ["-x", "hip", "-shared", os.path.join(HERE, "devlibm.hip"), "-o", tmp]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("devlibm build failed:\n" + " ".join(cmd) + "\n" + r.stdout[-8000:])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build())
