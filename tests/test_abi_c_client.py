"""The C ABI from plain C: tests/abi_c_client.c is compiled with gcc against
include/masx.h, linked to libmasx.so and run (host-only analysis path)."""
import os
import shutil
import subprocess

import pytest

from agipc_learning_amd import native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not found")
def test_c_client_links_and_analyzes(tmp_path):
    native.load()  # builds the library if it is missing
    lib = os.path.join(ROOT, "agipc_learning_amd", "lib")
    exe = str(tmp_path / "abi_c_client")
    subprocess.run(["gcc", "-std=c11", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "abi_c_client.c"), "-L", lib, "-lmasx", f"-Wl,-rpath,{lib}",
                    "-o", exe], check=True)
    out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout.split()
    vals = dict(zip(out[0:14:2], out[1:14:2]))
    # 100 chain rows padded to 128; 100 diagonal + 99 off-diagonal upper blocks
    assert vals["rows"] == "128" and vals["nb"] == "199" and vals["nnz"] == str(100 + 2 * 99)
    # level-1: one component per occupied 32-slot bank (100 rows -> 4 banks)
    assert out[out.index("counts") + 1] == "4"
