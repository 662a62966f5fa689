"""bench.py without a GPU: the figures its roofline reads from the rocprofv3
summaries under profiles/ are present for both matrix-core GEMMs, and
consistent with the summaries' own tables; more than one GPU is refused
with a message."""
import os
import subprocess
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, REPO)


def _bench():
    import importlib
    return importlib.import_module("bench")


@pytest.mark.parametrize("kernel,busy,clock", [("corr_f6_kernel", (60, 90), (1.3, 2.4)),
                                               ("corr_mfma_kernel", (70, 95), (1.3, 2.4))])
def test_profile_figures_present(kernel, busy, clock):
    b = _bench()
    traffic = b.from_profile(kernel, "traffic_bytes_per_launch")
    mfma = b.from_profile(kernel, "mfma_busy_pct")
    ghz = b.from_profile(kernel, "clock_ghz")
    assert traffic is not None and mfma is not None and ghz is not None
    # a config-2 GEMM launch reads at least its planes once (int8 4.29 GB,
    # FP6 3.23 GB) and writes the visibilities (0.27 GB)
    assert 3.2e9 < traffic < 3.0e10
    assert busy[0] < mfma < busy[1]
    assert clock[0] < ghz < clock[1]


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(REPO, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
