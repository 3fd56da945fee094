"""CPU check of the streaming Ωξ kernel's indexing (zandpack_amd/csrc/zh_omgxi.hpp):
a host program runs the kernel's own address functions (the same __host__ __device__
code the kernel calls) over every workgroup, chunk, load and lane of several shapes,
emulates the direct-to-LDS loads into an LDS image poisoned with NaN before every chunk
and the matrix-core operand reads from it, and checks that every global and LDS access
is in bounds, every LDS slot of a chunk is written exactly once, the Ω entries the
kernel must never read (the unstored lower triangle, the skipped passive block) are
never read, and the summed partial products equal the direct product over the lead's
modes (both precisions, Ω in full and as its triangle, with and without the passive
block, 1-4 column groups, Md not a multiple of 64, one to three leads)."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "zandpack_amd", "csrc")


@pytest.mark.skipif(not os.path.exists("/opt/rocm/bin/hipcc"), reason="needs hipcc")
def test_xi_stream_indexing(tmp_path):
    exe = str(tmp_path / "xi_emul")
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    r = subprocess.run([hipcc, "-x", "hip", "--offload-arch=gfx950", "-O2", "-std=c++17",
                        "-I", CSRC, "-o", exe, os.path.join(HERE, "xi_stream_emulation.cpp")],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0 and "all ok" in r.stdout, r.stdout + r.stderr
