"""The drop-in binding (rast-3dmhd-cpp_amd/shim/mhd3d_shim.h) compiles inside
the reference's own translation unit: a file that includes the reference's
headers, as its 3dmhd.cpp does, then the shim, and calls the three shim
functions, passes the compiler's syntax and type checks (-fsyntax-only, the
reference's MPI compiler wrapper).  Only where the reference tree and an MPI
compiler exist (the build container); nothing is built or run from the
reference."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference"


def test_shim_compiles_against_the_reference_headers(tmp_path):
    if not os.path.isfile(os.path.join(REF, "3dmhddata.h")):
        pytest.skip("the reference tree is not here")
    mpicxx = shutil.which("mpic++") or "/opt/conda/bin/mpic++"
    if not os.path.exists(mpicxx):
        pytest.skip("no MPI C++ compiler")
    src = tmp_path / "use_shim.cpp"
    src.write_text('#include "3dmhddata.h"\n#include "3dmhdutils.h"\n#include "mhd3d_shim.h"\n'
                   "int main() { gpu_init(); gpu_step(); gpu_sync_host(); return 0; }\n")
    env = dict(os.environ, MPICH_CXX="g++")
    cmd = [mpicxx, "-std=c++20", "-fsyntax-only", "-I", REF, "-I", os.path.join(ROOT, "include"),
           "-I", os.path.join(ROOT, "rast-3dmhd-cpp_amd", "shim"), str(src)]
    out = subprocess.run(cmd, env=env, capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr[-3000:]
