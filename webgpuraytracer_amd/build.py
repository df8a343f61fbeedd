"""Builds the native libraries in-tree.

    python -m webgpuraytracer_amd.build [--only pt|scene|addon]

Outputs (git-ignored, shipped with the tree):
    webgpuraytracer_amd/lib/libpt.so       HIP kernels + the pt_* C ABI (gfx950)
    webgpuraytracer_amd/lib/libptscene.so  host scene builder (ptgen_* C ABI)
    webgpuraytracer_amd/js/build/pt_addon.node   Node-API addon over libpt
"""
import concurrent.futures
import os
import shutil
import subprocess
import sys
import tempfile

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
INC = os.path.join(ROOT, "include")
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "lib")
ARCH = "gfx950"

HIP_FLAGS = [
    "--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC",
    # one rounding per WGSL operation: no FMA contraction, no fast-math
    "-ffp-contract=off", "-fno-fast-math",
    # packed FP32 ops are disabled for every HIP translation unit
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + INC, "-I" + CSRC,
]
# Per-source additions.  The learned denoiser's kernels keep their MFMA
# accumulators in VGPRs: with the compiler's default (AGPRs) they were shuffled
# between the two register files around every loop, ~15% of the convolution's
# VALU, and the level-1 convolution needed 192 instead of 154 registers.
SRC_FLAGS = {"nn_kernels.hip": ["-mllvm", "--amdgpu-mfma-vgpr-form"]}


def _run(cmd):
    print("+ " + " ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if out.strip():
        print(out)
    if r.returncode != 0:
        raise RuntimeError("command failed (%d): %s" % (r.returncode, " ".join(cmd)))


def _hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def build_scene():
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "libptscene.so")
    tmp = out + ".tmp"
    _run(["g++", "-std=c++17", "-O2", "-fPIC", "-shared", "-I" + INC,
          os.path.join(CSRC, "scene", "scene_gen.cpp"), os.path.join(CSRC, "scene", "assets.cpp"), "-o", tmp])
    os.replace(tmp, out)
    return out


def build_pt():
    os.makedirs(LIB, exist_ok=True)
    hipcc = _hipcc()
    out = os.path.join(LIB, "libpt.so")
    with tempfile.TemporaryDirectory(prefix="ptbuild") as td:
        srcs = ("pt_kernels.hip", "pt_wavefront.hip", "nn_kernels.hip", "pt_build.hip", "pt_api.cpp", "pt_layout.cpp",
                "pt_unet.cpp", "pt_batch.cpp", "pt_noalloc.cpp")
        objs = [os.path.join(td, src.split(".")[0] + ".o") for src in srcs]
        with concurrent.futures.ThreadPoolExecutor(max_workers=len(srcs)) as ex:
            jobs = [ex.submit(_run, [hipcc] + HIP_FLAGS + SRC_FLAGS.get(src, []) +
                              ["-x", "hip", "-c", os.path.join(CSRC, src), "-o", o])
                    for src, o in zip(srcs, objs)]
            for j in jobs:
                j.result()
        tmp = out + ".tmp"
        _run([hipcc, "--offload-arch=" + ARCH, "-shared", "-fPIC"] + objs + ["-o", tmp])
        os.replace(tmp, out)
    return out


def build_addon():
    """Node-API addon (plain g++ against the system node headers)."""
    node_inc = "/usr/include/node"
    src = os.path.join(PKG, "js", "pt_addon.cc")
    if not os.path.exists(src):
        print("Node-API addon source absent; skipping")
        return None
    if not os.path.exists(os.path.join(node_inc, "node_api.h")):
        print("node headers not found; skipping the Node-API addon")
        return None
    jsdir = os.path.join(PKG, "js")
    outdir = os.path.join(jsdir, "build")
    os.makedirs(outdir, exist_ok=True)
    out = os.path.join(outdir, "pt_addon.node")
    tmp = out + ".tmp"
    _run(["g++", "-std=c++17", "-O2", "-fPIC", "-shared", "-DNAPI_VERSION=8",
          "-I" + node_inc, "-I" + INC, os.path.join(jsdir, "pt_addon.cc"),
          "-L" + LIB, "-lpt", "-Wl,-rpath,$ORIGIN/../../lib", "-o", tmp])
    os.replace(tmp, out)
    return out


def build_all(only=None):
    outs = []
    if only in (None, "scene"):
        outs.append(build_scene())
    if only in (None, "pt"):
        outs.append(build_pt())
    if only in (None, "addon"):
        outs.append(build_addon())
    return outs


if __name__ == "__main__":
    only = None
    if "--only" in sys.argv:
        only = sys.argv[sys.argv.index("--only") + 1]
    for o in build_all(only):
        print("built", o)
