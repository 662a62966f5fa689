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


# Reference-generator fixtures the C host can reproduce from its command
# line (T a multiple of 32): params are (T, F, N, gen_type, seed,
# default_real, default_imag, initial_real, initial_imag, single_freq,
# no_repeat), the arguments of generate_char_data_set (input_generator.c:17).
CLI_CASES = [("c1_repeat", [0, 1, 2, 3]), ("c1_continuous", [0, 3]),
             ("rand_cont_s7", [3]), ("single_freq_1", [1, 3])]


@pytest.mark.gpu
@pytest.mark.parametrize("case,methods", CLI_CASES)
def test_cli_end_to_end_matches_reference_fixtures(golden, case, methods, tmp_path):
    # The C host's own generator and the GPU correlation, written as the
    # packed upper triangle (-o), equal the reference CPU correlator's output
    # for the same generator arguments, both conventions.
    import numpy as np

    T, F, N, gen, seed, dre, dim, ire, iim, q, p = (int(v) for v in golden[f"{case}__params"])
    for k in methods:
        for U in (1, 0):
            ta = 256 if T % 256 == 0 else 32
            out = tmp_path / f"{case}_{k}_{U}.bin"
            args = [_cli(), "-e", str(N), "-f", str(F), "-T", str(T), "-t", str(ta), "-i", "1",
                    "-g", str(gen), "-r", str(seed), "-x", str(dre), "-y", str(dim),
                    "-X", str(ire), "-Y", str(iim), "-q", str(q), "-k", str(k), "-U", str(U),
                    "-o", str(out)] + (["-p"] if p else [])
            print("  " + " ".join(args[1:]), flush=True)
            r = subprocess.run(args, capture_output=True, text=True, timeout=120)
            assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
            got = np.fromfile(out, dtype=np.int32).reshape(F, N * (N + 1) // 2, 2)
            np.testing.assert_array_equal(got, golden[f"{case}__tri_U{U}"],
                                          err_msg=f"{case} -k {k} -U {U}")
