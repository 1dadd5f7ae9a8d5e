"""Shared fixtures.  GPU tests are marked `gpu`; everything else runs on a
CPU-only machine."""
import os
import sys

# BLAS used by the CPU checkers: sequential inside their OpenMP threads
# (libxm's release build links sequential MKL, libxm Makefile:7).
os.environ.setdefault("MKL_THREADING_LAYER", "SEQUENTIAL")
os.environ.setdefault("OMP_NUM_THREADS", "4")

import pytest  # noqa: E402

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from libxm_amd import xm  # noqa: E402

ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libxmref.so")
RECORDER_SO = os.path.join(ROOT, "oracle", "libgemmrec.so")
LIBXM_SO = os.path.join(ROOT, "libxm_amd", "libxm.so")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")
    config.addinivalue_line("markers", "slow: long-running CPU test")


@pytest.fixture(scope="session")
def recorder():
    """?gemm_ recorder, loaded globally so it sits in front of the
    reference library's BLAS calls (must load before the reference)."""
    if not os.path.exists(RECORDER_SO):
        pytest.skip("oracle/libgemmrec.so not built")
    import ctypes
    rec = ctypes.CDLL(RECORDER_SO, mode=ctypes.RTLD_GLOBAL)
    rec.gemm_recorder_count.restype = ctypes.c_size_t
    rec.gemm_recorder_get.argtypes = [ctypes.c_size_t,
                                      ctypes.POINTER(ctypes.c_double)]
    return rec


@pytest.fixture(scope="session")
def oracle():
    if not os.path.exists(ORACLE_SO):
        pytest.fail("oracle/liboracle.so is not built (make -C oracle)")
    return xm.XmLib(ORACLE_SO)


@pytest.fixture(scope="session")
def ref(recorder):
    """libxm itself, compiled from its sources into oracle/_ref (only where
    the sources are present)."""
    if not os.path.exists(REF_SO):
        pytest.skip("reference build oracle/_ref/libxmref.so absent")
    return xm.XmLib(REF_SO)


@pytest.fixture(scope="session")
def gpulib():
    """The MI355X library.  Fails loudly if it is not built."""
    return xm.load(LIBXM_SO)


@pytest.fixture(autouse=True)
def _restore_precision(request):
    """Tests that use the MI355X library leave its precision settings as
    they found them (the library defaults, for the tests that check
    them)."""
    if "gpulib" not in request.fixturenames:
        yield
        return
    lib = request.getfixturevalue("gpulib")
    saved = (lib.lib.xm_get_contract_precision(),
             lib.lib.xm_get_split_slices(), lib.lib.xm_get_split_moduli())
    yield
    lib.set_precision(*saved)
