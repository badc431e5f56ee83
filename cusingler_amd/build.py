"""Build every native part of the package in place.

    python -m cusingler_amd.build

- ``cusingler_amd/libcusingler_amd.so``: the device engine (HIP, gfx950) and
  the host preprocessing / pipeline (C++), exporting the C-ABI of
  ``include/cusingler_amd.h`` and ``include/cusingler_prep.h``;
- ``cusingler/cusingler<EXT_SUFFIX>``: the pybind11 module ``cusingler`` with
  the reference's ``run`` entry point;
- ``cusingler_amd/libcsg_synth.so``: the seeded synthetic data sampler.

Device code is compiled with packed FP32 instructions disabled.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
import sysconfig

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
ARCH = os.environ.get("CSG_OFFLOAD_ARCH", "gfx950")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")

HIP_FLAGS = ["-std=c++17", "-O3", "-fPIC", f"--offload-arch={ARCH}",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-ffp-contract=off", "-Wno-unused-result"]
CXX_FLAGS = ["-std=c++17", "-O2", "-fPIC", "-ffp-contract=off", "-Wall", "-Wno-sign-compare",
             "-Wno-unused-function", "-I/opt/rocm/include", "-D__HIP_PLATFORM_AMD__"]


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout)
        raise RuntimeError("build step failed: " + " ".join(cmd))
    return r.stdout


def _fresh(target, sources):
    if not os.path.exists(target):
        return False
    t = os.path.getmtime(target)
    return all(os.path.getmtime(s) <= t for s in sources)


def build_engine(force=True):
    os.makedirs(BUILD, exist_ok=True)
    hdrs = [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".h")]
    hdrs += [os.path.join(ROOT, "include", f) for f in os.listdir(os.path.join(ROOT, "include"))]
    objs = []
    for src in ["csg_engine.hip"]:
        s = os.path.join(CSRC, src)
        o = os.path.join(BUILD, src + ".o")
        if force or not _fresh(o, [s] + hdrs):
            _run([HIPCC] + HIP_FLAGS + ["-c", s, "-o", o])
        objs.append(o)
    for src in ["csg_prep.cpp", "csg_pipeline.cpp"]:
        s = os.path.join(CSRC, src)
        o = os.path.join(BUILD, src + ".o")
        if force or not _fresh(o, [s] + hdrs):
            _run([CXX] + CXX_FLAGS + ["-c", s, "-o", o])
        objs.append(o)
    lib = os.path.join(PKG, "libcusingler_amd.so")
    tmp = lib + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs + ["-lpthread"])
    os.replace(tmp, lib)
    return lib


def build_pymodule():
    import pybind11
    out_dir = os.path.join(ROOT, "cusingler")
    os.makedirs(out_dir, exist_ok=True)
    suffix = sysconfig.get_config_var("EXT_SUFFIX")
    out = os.path.join(out_dir, "cusingler" + suffix)
    tmp = out + ".tmp"
    _run([CXX, "-shared", "-std=c++17", "-O2", "-fPIC", "-fvisibility=hidden",
          "-I" + pybind11.get_include(), "-I" + sysconfig.get_paths()["include"],
          os.path.join(CSRC, "csg_pymodule.cpp"), "-o", tmp, "-L" + PKG, "-lcusingler_amd",
          "-Wl,-rpath,$ORIGIN/../cusingler_amd"])
    os.replace(tmp, out)
    return out


def build_synth():
    out = os.path.join(PKG, "libcsg_synth.so")
    tmp = out + ".tmp"
    _run([CXX, "-std=c++17", "-O3", "-fPIC", "-shared", "-fopenmp",
          os.path.join(CSRC, "synth.cpp"), "-o", tmp])
    os.replace(tmp, out)
    return out


def build_shim():
    """The reference's device API over the C-ABI (shim/cusingler_shim.cpp)
    and a driver that runs the reference's pipeline order through it."""
    shim = os.path.join(PKG, "shim")
    out = os.path.join(shim, "shim_driver")
    tmp = out + ".tmp"
    _run([CXX, "-std=c++17", "-O2", "-Wall", os.path.join(shim, "cusingler_shim.cpp"),
          os.path.join(shim, "shim_driver.cpp"), "-o", tmp, "-L" + PKG, "-lcusingler_amd",
          "-Wl,-rpath,$ORIGIN/.."])
    os.replace(tmp, out)
    return out


def build_all(force=True):
    outs = [build_engine(force), build_pymodule(), build_synth(), build_shim()]
    return outs


if __name__ == "__main__":
    for p in build_all(force="--incremental" not in sys.argv):
        print("built", os.path.relpath(p, ROOT))
