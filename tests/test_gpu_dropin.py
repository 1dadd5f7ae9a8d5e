"""Drop-in check: libxm's own test driver (test.c) and tutorial
(example.c), compiled unmodified from libxm's sources against include/ and
libxm_amd/libxm.so (oracle/Makefile, target `dropin`), run on the GPU.

test.c checks every operation against brute-force element loops for 4
scalar types and both allocator kinds: 712 "success" lines
(libxm test.c:1961-2073).  The binaries exist only where libxm's sources
were present at build time; the tests skip otherwise.
"""
import os
import subprocess

import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu

REF = os.path.join(ROOT, "oracle", "_ref")


def _bin(name):
    p = os.path.join(REF, name)
    if not os.path.exists(p):
        pytest.skip("%s not built (libxm sources absent at build time)" % name)
    return p


def test_libxm_test_driver_passes(gpulib):
    r = subprocess.run([_bin("test_dropin")], capture_output=True, text=True,
                       timeout=600, cwd="/tmp")
    assert r.returncode == 0, r.stderr[-2000:]
    lines = r.stdout.splitlines()
    assert sum(1 for x in lines if x.endswith("success")) == 712
    assert sum(1 for x in lines if x.startswith("contract test")) == 8 * 46


def test_libxm_example(gpulib):
    r = subprocess.run([_bin("example_dropin")], capture_output=True,
                       text=True, timeout=120, cwd="/tmp")
    assert r.returncode == 0, r.stderr
    out = r.stdout.split("tensor c = 2*a*b + 1")[1].split()
    assert [float(x) for x in out] == [43, 11, -1, 65, 9, 1, 17, 3, 1,
                                       39, 1, 3]
