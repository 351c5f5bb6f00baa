"""The host library (bigint_amd/host) under AddressSanitizer + UndefinedBehaviorSanitizer and under
ThreadSanitizer (CPU only).

tests/cpp/Makefile builds bigint_calc over the host BigInt/BigFloat code linked to the CPU product
backend of oracle/, so every code path of the host library runs without a GPU: products on the
init_thread_pool threads (batches), the threaded decimal conversion, DeviceInt batches, Newton
division and BigFloat bookkeeping.  Each result is checked against Python ints and the run must
leave no sanitizer report.
"""
import os
import random
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BUILD = os.path.join(ROOT, "tests", "cpp", "_build")

if hasattr(__import__("sys"), "set_int_max_str_digits"):
    __import__("sys").set_int_max_str_digits(0)


def hx(x: int) -> str:
    return ("-" if x < 0 else "") + format(abs(x), "x")


def unhx(s: str) -> int:
    return -int(s[1:], 16) if s.startswith("-") else int(s, 16)


def trunc_div(a, b):
    q = abs(a) // abs(b)
    return q if (a >= 0) == (b >= 0) else -q


def _session():
    """Commands and checks covering the host library's threaded and single-threaded paths."""
    rng = random.Random(404)
    cmds, checks = ["poolinit 4"], [lambda r: r == "ok"]

    def add(cmd, check):
        cmds.append(cmd)
        checks.append(check)

    for bits in [100, 5000, 40000, 300000, 1500000]:
        a = rng.getrandbits(bits) | 1
        b = -(rng.getrandbits(bits // 2 + 7) | 1)
        add(f"mul {hx(a)} {hx(b)}", lambda r, v=a * b: unhx(r) == v)
        add(f"sq {hx(b)}", lambda r, v=b * b: unhx(r) == v)
        add(f"add {hx(a)} {hx(b)}", lambda r, v=a + b: unhx(r) == v)
        add(f"sub {hx(b)} {hx(a)}", lambda r, v=b - a: unhx(r) == v)
        if bits <= 300000:
            add(f"divmod {hx(a * 3 + 11)} {hx(b)} 0",
                lambda r, n=a * 3 + 11, d=b: tuple(map(unhx, r.split())) == (trunc_div(n, d), n - d * trunc_div(n, d)))
    big = rng.getrandbits(400000) | 1
    add(f"todec {hx(big)}", lambda r, v=big: r == str(v))
    add(f"fromdec {big}", lambda r, v=big: unhx(r) == v)
    add("frecip " + hx(rng.getrandbits(3000) | 1) + " 40 200", lambda r: r.startswith("+ "))
    add("fmul " + hx(rng.getrandbits(20000)) + " 3 " + hx(rng.getrandbits(9000)) + " 5 300",
        lambda r: r[0] in "+-")
    a, b, c, d = (rng.getrandbits(200000) | 1 for _ in range(4))
    add(f"devmuladd {hx(a)} {hx(b)} {hx(-c)} {hx(d)}",
        lambda r, s=a * b - c * d, q=a * a: tuple(map(unhx, r.split())) == (s, q))
    add("batch 3000 91 6000", lambda r: r.isdigit())
    return cmds, checks


@pytest.mark.parametrize("kind", ["asan", "tsan"])
def test_host_library_under_sanitizer(kind):
    exe = os.path.join(BUILD, f"bigint_calc_{kind}")
    if not os.path.exists(exe):
        pytest.skip(f"{exe} not built (make -C tests/cpp)")
    cmds, checks = _session()
    env = dict(os.environ)
    env["ASAN_OPTIONS"] = "detect_leaks=1:abort_on_error=0:halt_on_error=1"
    env["UBSAN_OPTIONS"] = "print_stacktrace=1:halt_on_error=1"
    env["TSAN_OPTIONS"] = "halt_on_error=1:second_deadlock_stack=1"
    out = subprocess.run([exe], input="\n".join(cmds) + "\n", capture_output=True, text=True,
                         timeout=600, env=env)
    assert "Sanitizer" not in out.stderr and "runtime error" not in out.stderr, out.stderr[-4000:]
    assert out.returncode == 0, out.stderr[-4000:]
    lines = out.stdout.strip().split("\n")
    assert len(lines) == len(cmds)
    for cmd, check, line in zip(cmds, checks, lines):
        assert not line.startswith("EXC"), (cmd[:40], line)
        assert check(line), (cmd[:40], line[:80])
