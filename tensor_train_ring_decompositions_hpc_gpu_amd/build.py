"""Build the ttx HIP library in-tree: ``libttx.so`` (engine + kernels, C ABI of ``include/ttx.h``)
and the two reference-name shim libraries ``libttx_ttdec6.so`` / ``libttx_trdec6.so``.

Every HIP translation unit is compiled by hipcc for gfx950 only, with packed-FP32 instruction
selection disabled; objects go to ``_build/`` beside this file and the libraries next to it.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
ROOT = os.path.dirname(HERE)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = ["--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-I" + os.path.join(ROOT, "include")]

SOURCES = ["ttx_stream.hip", "ttx_dense.hip", "ttx_rng.hip", "ttx_gemm.hip", "ttx_permute.hip",
           "ttx_mlp.hip", "ttx_noalloc.hip", "ttx_engine.cpp"]
LIB = os.path.join(HERE, "libttx.so")
SHIMS = {"libttx_ttdec6.so": "shim_ttdec6.cpp", "libttx_trdec6.so": "shim_trdec6.cpp"}


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    # hipcc's host pass warns about the device-only target feature; that line is noise
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError("build failed: %s\n%s" % (" ".join(cmd), out))
    return out


def build(verbose=False, jobs=8):
    os.makedirs(OBJ, exist_ok=True)
    objs = []
    cmds = []
    for src in SOURCES:
        o = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
        objs.append(o)
        cmds.append([HIPCC] + FLAGS + ["-c", os.path.join(CSRC, src), "-o", o])
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        for out in ex.map(_run, cmds):
            if verbose and out.strip():
                print(out)
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    for lib, src in SHIMS.items():
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        out = os.path.join(HERE, lib)
        tmp = out + ".tmp"
        _run([HIPCC] + FLAGS + ["-shared", "-o", tmp, path, "-L" + HERE, "-lttx",
                                "-Wl,-rpath,$ORIGIN"])
        os.replace(tmp, out)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)
