"""libgbg's host code under AddressSanitizer and LeakSanitizer, on a CPU
stand-in for the HIP runtime (tests/hostsan): every C-ABI entry point on
fused, composed, permuted-layout and chirp-z grids, the batch and the server,
the argument refusals, and each scenario again with its k-th device
allocation, stream, event or graph instantiation made to fail, for every k it
reaches.  A failed call must return an error status (or complete where the
library declines an optional resource), release every device allocation it
made and leave host memory intact.  Kernels do not run here (no GPU): this
checks the host logic only."""

import os
import shutil
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "hostsan"))


def _toolchain():
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    return (shutil.which("hipcc") or os.path.exists(os.path.join(rocm, "bin", "hipcc"))) and \
        os.path.exists(os.path.join(rocm, "llvm", "bin", "clang++"))


@pytest.mark.skipif(not _toolchain(), reason="hipcc and the ROCm clang++ are needed to build the harness")
def test_host_code_under_sanitizers_with_failure_injection():
    import build as hostsan_build
    drive = hostsan_build.build()
    env = dict(os.environ)
    env["ASAN_OPTIONS"] = "detect_leaks=1:abort_on_error=0"
    r = subprocess.run([drive], capture_output=True, text=True, env=env, timeout=600)
    print(r.stdout)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-6000:]
    assert "hostsan: ok" in r.stdout
    # every scenario reached its resource calls and refused failures there
    for line in r.stdout.splitlines()[:-1]:
        assert "resource calls" in line and " 0 resource calls" not in line, line
