"""Build recipe of the native parts (run from any directory).

Outputs go to multi-vector-retrieval-gpu_amd/lib/:
  libmvr.so                      HIP kernels + C ABI (hipcc, gfx950)
  IGPGPUPP, IGPGPUBasic          command-line drop-ins (g++, link libmvr.so)
  GPGPUBuildIndexPackGPU*.so     pybind11 index-build module (g++)
  BruteForceProgressive*.so      pybind11 exhaustive-MaxSim module (g++, calls libmvr)

Usage: python multi-vector-retrieval-gpu_amd/build.py [lib|cli|pymod|all]
"""
from __future__ import annotations

import os
import subprocess
import sys
import sysconfig

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "lib")
INCLUDE = os.path.join(ROOT, "include")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# -packed-fp32-ops: no v_pk_*_f32 (unreliable on these boxes);
# -ffp-contract=off: only explicit fmaf() fuses, matching the CPU oracle.
HIP_FLAGS = ["-O3", "--offload-arch=gfx950", "-std=c++17", "-fPIC", "-ffp-contract=off",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wall",
             "-Wno-unused-command-line-argument"]
HIP_SOURCES = ["mvr_kernels.hip", "mvr_capi.hip", "mvr_mlp.hip", "mvr_plaid.hip", "mvr_exhaustive.hip"]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_lib() -> str:
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "libmvr.so")
    objs = []
    for src in HIP_SOURCES:
        obj = os.path.join(LIB, src.replace(".hip", ".o"))
        _run([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    tmp = out + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", tmp])
    os.replace(tmp, out)
    return out


def build_cli() -> list[str]:
    outs = []
    for name, variant in (("IGPGPUPP", "MVR_VARIANT_PP"), ("IGPGPUBasic", "MVR_VARIANT_BASIC")):
        out = os.path.join(LIB, name)
        _run(["g++", "-O2", "-std=c++17", f"-D{variant}=1", "-I", INCLUDE,
              os.path.join(CSRC, "igpgpu_cli.cpp"), "-o", out, "-L", LIB, "-lmvr",
              "-Wl,-rpath,$ORIGIN"])
        outs.append(out)
    return outs


def build_pymod() -> str:
    import pybind11
    suffix = sysconfig.get_config_var("EXT_SUFFIX")
    out = os.path.join(LIB, "GPGPUBuildIndexPackGPU" + suffix)
    _run(["g++", "-O3", "-std=c++17", "-shared", "-fPIC", "-fopenmp", "-ffp-contract=off",
          "-I", pybind11.get_include(), "-I", sysconfig.get_paths()["include"],
          os.path.join(CSRC, "build_index_module.cpp"), "-o", out, "-ldl"])
    return out


def build_bfmodule() -> str:
    """BruteForceProgressive pybind module (exhaustive MaxSim over libmvr)."""
    import pybind11
    suffix = sysconfig.get_config_var("EXT_SUFFIX")
    out = os.path.join(LIB, "BruteForceProgressive" + suffix)
    _run(["g++", "-O3", "-std=c++17", "-shared", "-fPIC", "-fopenmp",
          "-I", pybind11.get_include(), "-I", sysconfig.get_paths()["include"],
          os.path.join(CSRC, "bruteforce_module.cpp"), "-o", out, "-ldl"])
    return out


def build_probe() -> str:
    """Test-only MFMA layout probe (tests/probes/libmvr_probe.so)."""
    src = os.path.join(ROOT, "tests", "probes", "mfma_probe.hip")
    out = os.path.join(ROOT, "tests", "probes", "libmvr_probe.so")
    _run([HIPCC, *HIP_FLAGS, "-shared", src, "-o", out])
    return out


def build_all():
    build_lib()
    build_cli()
    build_pymod()
    build_bfmodule()
    build_probe()


if __name__ == "__main__":
    what = sys.argv[1] if len(sys.argv) > 1 else "all"
    {"lib": build_lib, "cli": build_cli, "pymod": build_pymod, "bf": build_bfmodule, "probe": build_probe,
     "all": build_all}[what]()
