"""The reference-side C++ adapter (tests/integration/ogp_adapter.hpp: the classes a maintainer
of the reference would put in front of the C ABI) builds
against include/ogp.h with a plain host compiler and links libogp.so (CPU), and its driver
(tests/integration/adapter_test.cpp) passes on the GPU: stepwise GpuPropagator on the caller's
device buffers equal to the bulk propagation, the IForceModel, GpuStartup and Predictor forms,
and ClosestApproach with ClosestApproachData::Get."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "opengpop_amd")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")


def _build(out):
    cxx = shutil.which("g++") or shutil.which("c++")
    if cxx is None:
        pytest.skip("no host C++ compiler")
    cmd = [cxx, "-std=c++17", "-O1", "-Wall", "-Werror", "-D__HIP_PLATFORM_AMD__",
           "-I", os.path.join(ROOT, "include"), "-I", os.path.join(ROCM, "include"),
           os.path.join(ROOT, "tests", "integration", "adapter_test.cpp"), "-o", out,
           "-L", PKG, "-logp", "-L", os.path.join(ROCM, "lib"), "-lamdhip64",
           "-Wl,-rpath," + PKG]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    return out


def test_adapter_builds_against_the_header(tmp_path):
    assert os.path.exists(os.path.join(PKG, "libogp.so"))
    _build(str(tmp_path / "adapter_test"))


@pytest.mark.gpu
def test_adapter_drives_the_library(tmp_path):
    exe = _build(str(tmp_path / "adapter_test"))
    r = subprocess.run([exe, os.path.join(ROOT, "data")], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0 and "adapter ok" in r.stdout, r.stdout + r.stderr
