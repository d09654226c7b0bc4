"""Build libyafgpu.so (hipcc, gfx950) and the yafgpu-render scene-file
renderer in-tree.

Usage: python -m yafray_amd.build  (or __graft_entry__.build()).
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libyafgpu.so")
CLI = os.path.join(HERE, "yafgpu-render")

SOURCES = ["capi.cpp", "qmc.cpp", "bvh_build.cpp", "bvh_gpu.hip", "autotune.cpp", "trace.hip", "render.hip",
           "photon.hip", "gphoton.hip",
           "../host/yafgpu_host.cpp", "../host/yafgpu_xml.cpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("YG_OFFLOAD_ARCH", "gfx950")

FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=" + ARCH,
    # all arithmetic is evaluated as written: device results must equal a
    # strict-IEEE evaluation of the reference expressions
    "-ffp-contract=off",
    # no packed FP32 instructions in generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def _compile(src, obj_dir=OBJ, defines=()):
    out = os.path.join(obj_dir, os.path.splitext(os.path.basename(src))[0] + ".o")
    cmd = [HIPCC] + FLAGS + ["-D" + d for d in defines] + ["-x", "hip", "-c", os.path.join(CSRC, src), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    return out


def build(verbose=True, defines=(), lib=LIB, obj_dir=OBJ):
    """Compile every source and link `lib`.  `defines` (NAME=VALUE strings)
    override the tuning constants (stack depths, block size) for a variant
    build; the default build uses none."""
    os.makedirs(obj_dir, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(lambda s: _compile(s, obj_dir, tuple(defines)), srcs))
    tmp = lib + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs + ["-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s" % (r.stdout, r.stderr))
    os.replace(tmp, lib)
    if lib == LIB:
        # the command-line renderer, linked against the library beside it
        cmd = ["g++", "-std=c++17", "-O2", "-I", os.path.join(HERE, "host"),
               os.path.join(HERE, "host", "yafgpu_render_main.cpp"), "-L", HERE, "-lyafgpu",
               "-Wl,-rpath,$ORIGIN", "-Wl,-rpath,/opt/rocm/lib", "-o", CLI + ".tmp"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("link failed: %s\n%s" % (r.stdout, r.stderr))
        os.replace(CLI + ".tmp", CLI)
    if verbose:
        print("built", lib)
    return lib


if __name__ == "__main__":
    build()
    sys.exit(0)
