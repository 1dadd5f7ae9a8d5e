"""The file tier's storage and staging layout (libxm_amd/csrc/filestore.cpp),
compiled for the host: file pointers, the backing file's first-fit
allocator against a brute-force page map, pread/pwrite round trips, file
growth and removal, libxm's creation semantics (open with O_CREAT, an
existing file truncated, a bad path refused) and the merging of staged
ranges (tests/filestore_check.cpp)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "libxm_amd", "csrc")


def test_filestore(tmp_path):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("no host C++ compiler")
    exe = str(tmp_path / "fscheck")
    subprocess.check_call([gxx, "-O2", "-std=c++17", "-Wall", "-I" + CSRC,
                           "-o", exe, os.path.join(CSRC, "filestore.cpp"),
                           os.path.join(ROOT, "tests", "filestore_check.cpp")])
    work = tmp_path / "files"
    work.mkdir()
    out = subprocess.run([exe, str(work)], capture_output=True, text=True,
                         timeout=120)
    print(out.stdout, out.stderr)
    assert out.returncode == 0 and out.stdout.strip().endswith("OK")
    assert os.listdir(work) == []       # every backing file removed
