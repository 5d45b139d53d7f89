"""Builds the engine in-tree for gfx950: opengpop_amd/libogp.so (C ABI of include/ogp.h) and the
opengpop command-line tool. Run: python -m opengpop_amd.build  (or __graft_entry__.build()).
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
OBJ = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libogp.so")
CLI = os.path.join(PKG, "opengpop")
ARCH = os.environ.get("OGP_OFFLOAD_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
COMMON = ["-O3", "-fPIC", "-std=c++17", "-Wall", f"--offload-arch={ARCH}",
          # no packed FP32 instructions (v_pk_fma_f32 and kin): some of their forms have given
          # wrong results on this hardware without an error
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-I", os.path.join(ROOT, "include")]

SOURCES = ["datasets.cpp", "kernels.hip", "propagate.hip", "conjunction.hip", "adjoint.hip",
           "perturb.hip", "capi.cpp"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + r.stdout)
    # keep compiler warnings visible, minus the host pass's note about the GPU-only feature
    lines = [l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l]
    if any(l.strip() for l in lines):
        sys.stderr.write("\n".join(lines) + "\n")
    return r


def _compile(src):
    path = os.path.join(CSRC, src)
    out = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC] + COMMON + ["-x", "hip", "-c", path, "-o", out]
    _run(cmd)
    return out


def build(verbose=True):
    os.makedirs(OBJ, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    _run([HIPC_LINK := HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    cli_src = os.path.join(CSRC, "cli.cpp")
    if os.path.exists(cli_src):
        _run([HIPCC, "-O2", "-std=c++17", "-x", "c++", cli_src, "-I", os.path.join(ROOT, "include"),
              "-o", CLI + ".tmp", "-L", PKG, "-logp", "-Wl,-rpath,$ORIGIN"])
        os.replace(CLI + ".tmp", CLI)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build()
