"""A plain C program calls the drop-in API with host arrays, as the
reference's own callers do, and checks forward, inverse and convolution
results against FP64 direct sums (tests/c/dropin_host.c)."""
import os
import subprocess

import pytest

from tests.gpu_util import torch_cuda

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(HERE, "c", "dropin_host")


def test_c_caller_with_host_arrays():
    if torch_cuda() is None:
        pytest.skip("no GPU")
    if not os.path.exists(BIN):
        pytest.fail("tests/c/dropin_host is not built (run __graft_entry__.build())")
    print("running", BIN, flush=True)
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "all ok" in r.stdout
