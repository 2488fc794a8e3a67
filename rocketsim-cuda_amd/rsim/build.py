"""Builds librsim.so in-tree: the gfx950 kernels and the C ABI (hipcc) plus the host-side world builder (g++).

    python -m rsim.build            (from rocketsim-cuda_amd/), or rsim.build.build() from Python

Float policy of both compilers: -ffp-contract=off (no fused multiply-add), no fast-math, fp32 denormals kept, HIP's correctly
rounded float divide and sqrt. All HIP code is compiled without packed-fp32 instructions, and with a larger budget for
keeping small private arrays in registers. Everything is compiled with
-fno-strict-aliasing: the states are copied between buffers and structs through u32 word views.
"""
import concurrent.futures
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")
OUT_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "librsim.so")
EXAMPLES = {"kickoff": os.path.join(PKG, "examples", "kickoff.cpp")}
# user kernels on the device API (include/rsim_device.h), each built into its own library beside librsim.so
DEVICE_EXAMPLES = {"libobserve.so": os.path.join(PKG, "examples", "observe.hip")}

ARCH = "gfx950"
# -fno-strict-aliasing: states are copied in and out through u32 word views of the structs
COMMON = ["-std=c++17", "-O3", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-fno-strict-aliasing", "-I" + CSRC, "-I" + INCLUDE]
HIP_FLAGS = [
    "--offload-arch=" + ARCH,
    "-fno-gpu-flush-denormals-to-zero",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # small private arrays indexed at run time (the island builder's sort keys, the box-box clipping buffers) kept in
    # vector registers instead of private memory: the default budget leaves them in memory (measured 1.2% faster)
    "-mllvm", "-amdgpu-promote-alloca-to-vector-limit=1024",
    "-mllvm", "-amdgpu-promote-alloca-to-vector-max-regs=64",
    "-mllvm", "-amdgpu-promote-alloca-to-vector-vgpr-ratio=2",
    # loops over private arrays fully unrolled up to a larger size, so more indices become constants and more of those
    # arrays stay in registers (measured 1.5% faster; a larger budget still is slower)
    "-mllvm", "-amdgpu-unroll-threshold-private=6000",
    "-mllvm", "-amdgpu-unroll-max-block-to-analyze=64",
    "-Wno-unused-command-line-argument",
]
HIP_SOURCES = [os.path.join(CSRC, "device", "sim_kernels.hip"), os.path.join(CSRC, "device", "cost_model.hip"),
               os.path.join(CSRC, "host", "rsim.cpp")]
HOST_SOURCES = [os.path.join(CSRC, "host", "world.cpp"), os.path.join(CSRC, "host", "v3wire.cpp")]


def _hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found: librsim.so needs the ROCm toolchain to build")


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build step failed (%d):\n%s\n%s" % (r.returncode, " ".join(cmd), r.stdout[-20000:]))
    return r.stdout


def build(verbose=False):
    """Compiles every source and links librsim.so (always rebuilds)."""
    os.makedirs(OUT_DIR, exist_ok=True)
    hipcc = _hipcc()
    objs, steps = [], []
    for src in HIP_SOURCES:
        obj = os.path.join(OUT_DIR, os.path.basename(src) + ".o")
        lang = ["-x", "hip"] if src.endswith(".hip") else []
        steps.append([hipcc] + COMMON + HIP_FLAGS + lang + ["-c", src, "-o", obj])
        objs.append(obj)
    for src in HOST_SOURCES:
        obj = os.path.join(OUT_DIR, os.path.basename(src) + ".o")
        steps.append(["g++"] + COMMON + ["-c", src, "-o", obj])
        objs.append(obj)
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(len(steps), 8)) as pool:  # the compiles are independent
        list(pool.map(_run, steps))
    tmp = LIB + ".tmp"
    _run([hipcc, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    # the C++ host examples, linked against librsim.so beside them
    for name, src in EXAMPLES.items():
        out = os.path.join(HERE, name)
        _run(["g++", "-std=c++17", "-O2", "-I" + INCLUDE, src, "-o", out + ".tmp", "-L" + HERE, "-lrsim",
              "-Wl,-rpath,$ORIGIN"])
        os.replace(out + ".tmp", out)
    for name, src in DEVICE_EXAMPLES.items():
        out = os.path.join(HERE, name)
        _run([hipcc] + COMMON + HIP_FLAGS + ["-x", "hip", "-shared", src, "-o", out + ".tmp"])
        os.replace(out + ".tmp", out)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
