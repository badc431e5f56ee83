"""The shim of the reference's device API (cusingler_amd/shim): it compiles
against the reference's own src/types.h and src/cusingler.cuh declarations
(where the reference sources are present), and its driver is built and
linked against the library."""

import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SHIM = os.path.join(ROOT, "cusingler_amd", "shim")
REF_SRC = "/root/reference/src"


@pytest.mark.skipif(not os.path.isdir(REF_SRC), reason="reference sources not present")
def test_shim_compiles_against_reference_types():
    # the shim's definitions must match the reference's declarations of the
    # seven functions: compile a unit that includes both
    unit = ('#include "cusingler.cuh"\n'
            '#include "' + os.path.join(SHIM, "cusingler_shim.cpp") + '"\n')
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-DCSG_SHIM_REFERENCE_TYPES",
                        "-I" + REF_SRC, "-I" + SHIM, "-x", "c++", "-"], input=unit,
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_shim_driver_is_built_and_linked():
    drv = os.path.join(SHIM, "shim_driver")
    assert os.path.exists(drv), "run python -m cusingler_amd.build"
    r = subprocess.run(["ldd", drv], capture_output=True, text=True)
    assert "libcusingler_amd.so" in r.stdout and "not found" not in r.stdout, r.stdout
