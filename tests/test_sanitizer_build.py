"""The sanitizer build is live: under `make -C vulkan-light-bakery_amd
asan-test` (ASan + UBSan on the host code and the oracle) a heap overflow
planted in a call to the oracle is reported.  Skipped in ordinary runs."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

asan = "_asan" in os.environ.get("VLB_ORC_LIB", "") and "asan" in os.environ.get("LD_PRELOAD", "")


@pytest.mark.skipif(not asan, reason="runs under the sanitizer build (make asan-test)")
def test_planted_overflow_is_reported():
    code = (
        "import sys, ctypes, numpy as np; sys.path.insert(0, 'tests'); import orc\n"
        "src = np.zeros(300, np.uint8); out = np.zeros(100, np.uint8)\n"
        "orc.L.orc_base64(src.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), 300,\n"
        "                 out.ctypes.data_as(ctypes.c_char_p))\n"   # 400 chars into 100 bytes
    )
    r = subprocess.run([sys.executable, "-c", code], cwd=ROOT, capture_output=True, text=True)
    assert r.returncode != 0
    assert "heap-buffer-overflow" in r.stderr
