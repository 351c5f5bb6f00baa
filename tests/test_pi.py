"""π by Chudnovsky binary splitting (the reference's examples/pi.cpp workload).

The GPU run (bigint_amd/_build/pi: level-wise batched splitting on the engine) must print the
same digits as the CPU run (oracle/pi_cpu: the same host code on the threaded AVX2 restatement of
the reference's multiply), and both must agree with mpmath's π up to the final rounded digit."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PI_GPU = os.path.join(ROOT, "bigint_amd", "_build", "pi")
PI_CPU = os.path.join(ROOT, "oracle", "pi_cpu")


PLANNER = os.path.join(ROOT, "bigint_amd", "plans", "planner.bin")


def run(binary, digits, path, mode=1, planner=None):
    args = [binary, str(digits), str(mode), path] + ([planner] if planner else [])
    out = subprocess.run(args, capture_output=True, text=True, timeout=600, check=True)
    with open(path) as f:
        return f.read().strip(), out.stdout


def mp_pi(digits):
    mpmath = pytest.importorskip("mpmath")
    mpmath.mp.dps = digits + 30
    return mpmath.nstr(mpmath.pi, digits + 20, strip_zeros=False)


def check_digits(text, digits):
    ref = mp_pi(digits)
    assert text.startswith("3.") and len(text) == digits + 2
    assert text[:-2] == ref[:digits]  # the last digit is rounded (RoundHalfUp)


def test_pi_cpu_small(tmp_path):
    if not os.path.exists(PI_CPU):
        pytest.skip("oracle/pi_cpu not built")
    text, _ = run(PI_CPU, 3000, str(tmp_path / "pi.txt"))
    check_digits(text, 3000)


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [0, 1, 2])
def test_pi_gpu_matches_cpu_run(tmp_path, mode):
    """Plain, table and learned plans (the committed planner file, which must load)."""
    digits = 200000
    gpu, log = run(PI_GPU, digits, str(tmp_path / "pi_gpu.txt"), mode, PLANNER if mode else None)
    assert '"gpu_products": 0' not in log
    if mode:
        assert '"planner_loaded": true' in log
    cpu, _ = run(PI_CPU, digits, str(tmp_path / "pi_cpu.txt"))
    assert gpu == cpu
    check_digits(gpu[:20002], 20000)
