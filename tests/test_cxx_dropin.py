"""include/dream_hip.hpp compiles against the reference's own dream.h (the
header its gateways include) and links against libdream_hip.so; on the GPU
the drop-in classes give the oracle's results."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_INC = os.path.join(os.environ.get("DREAM_REFERENCE", "/root/reference"), "include")
SRC = os.path.join(ROOT, "tests", "cxx", "dropin_check.cc")


def _build(tmp_path):
    exe = str(tmp_path / "dropin_check")
    cmd = ["g++", "-std=c++14", "-O1", "-I" + REF_INC, "-I" + os.path.join(ROOT, "include"),
           "-I/opt/rocm/include", SRC, "-o", exe, "-L" + os.path.join(ROOT, "dream_amd"),
           "-ldream_hip", "-Wl,-rpath," + os.path.join(ROOT, "dream_amd")]
    subprocess.run(cmd, check=True)
    return exe


@pytest.mark.skipif(not os.path.isdir(REF_INC), reason="reference headers not present")
def test_dropin_header_compiles_and_links(tmp_path):
    exe = _build(tmp_path)
    out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout
    assert out.startswith("created") or out.startswith("no-device")
