"""The flag-compatible C host (chime-x_amd/correlator_test, the reference's
main_wrapper.c command line over libchx) with its own CPU check (-c), for
every method and both conventions, on the GPU: its timed loop double-buffers
two stages on a copy queue and a compute queue (main_wrapper.c:613-785), and
both stages' visibilities must agree."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CLI = os.path.join(HERE, "..", "chime-x_amd", "correlator_test")


def _cli():
    if not os.path.exists(CLI):
        pytest.skip("correlator_test not built (make -C chime-x_amd)")
    return CLI


def test_help_runs_without_a_device():
    r = subprocess.run([_cli(), "-h"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "--kernel_batch" in r.stdout


@pytest.mark.gpu
@pytest.mark.parametrize("U", [1, 0])
@pytest.mark.parametrize("k,ta", [(0, 256), (1, 2048), (2, 256), (3, 256)])
def test_cli_matches_its_cpu_check(k, ta, U):
    args = [_cli(), "-e", "256", "-f", "2", "-T", "4096", "-t", str(ta), "-i", "2", "-p",
            "-c", "-k", str(k), "-U", str(U)]
    print("  " + " ".join(args[1:]), flush=True)
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "CPU matches GPU" in r.stdout
    assert "Stages 0 and 1 agree" in r.stdout


@pytest.mark.gpu
@pytest.mark.parametrize("k", [0, 3])
@pytest.mark.parametrize("w", [False, True])
def test_cli_double_buffered_loop(k, w):
    # five passes alternate the two stages; -w loads both stages once
    args = [_cli(), "-e", "128", "-f", "3", "-T", "2048", "-t", "256", "-i", "5", "-p",
            "-c", "-k", str(k)] + (["-w"] if w else [])
    print("  " + " ".join(args[1:]), flush=True)
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "CPU matches GPU" in r.stdout and "Stages 0 and 1 agree" in r.stdout
