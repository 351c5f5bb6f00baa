"""Timing of the CPU restatement (oracle/libcpu_mt.so) for the CPU baseline: one process per
thread count, because the pool follows init_thread_pool's first-call-wins rule
(reference ntt_multithread.cpp:407-409).  Test/benchmark infrastructure only.

  python -m oracle.cpu_time --threads 8 --bits 28          # 2^28 x 2^28, warm-up + timed runs
  python -m oracle.cpu_time --threads 16 --batch 4096 --stride 16
  python -m oracle.cpu_time --threads 8 --c1                # 2^22 x 2^22 and the square of A

--batch times a sample of the batched product-tree workload (every `stride`-th pair of the
4096-pair log-uniform draw) in two ways: a sequential loop of multiplies on the reference's routes
(brute and FFT single-threaded, NTT on the threaded pool: the reference's `operator*` loop), and
pairs spread across `threads` workers each running a one-thread multiply.  Prints one JSON line.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import threading
import time

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

from bigint_amd import limbs as L  # noqa: E402
from oracle import oracle  # noqa: E402


def _mul(M, a: np.ndarray, b: np.ndarray, out: np.ndarray) -> None:
    n = ctypes.c_size_t()
    p = ctypes.POINTER(ctypes.c_uint32)
    M.cpu_mul_digits(a.ctypes.data_as(p), a.size, b.ctypes.data_as(p), b.size, 0,
                     out.ctypes.data_as(p), ctypes.byref(n))


def balanced(threads: int, log_bits: int, max_s: float) -> dict:
    M = oracle.mt_lib()
    M.cpu_init_thread_pool(threads)
    bits = 1 << log_bits
    a = L.random_operand(bits, L.SEED)
    b = L.random_operand(bits, L.SEED + 1)
    out = np.zeros(a.size + b.size + 4, np.uint32)
    t0 = time.perf_counter()
    _mul(M, a, b, out)
    warm = time.perf_counter() - t0
    times = []
    while not times or (sum(times) + warm < max_s and len(times) < 3):
        t0 = time.perf_counter()
        _mul(M, a, b, out)
        times.append(time.perf_counter() - t0)
    med = float(np.median(times))
    return {"threads": int(M.cpu_pool_threads()), "bits": bits, "ms": med * 1e3, "runs": len(times),
            "Gbit_s": 2 * bits / med / 1e9}


ROUTES = ("brute", "fft", "ntt", "ssa")


def batch_ops(count: int, stride: int):
    """Every `stride`-th pair of the batched workload: the same SplitMix64 operands the GPU runs
    (limbs.batch_bits, limbs.batch_seeds)."""
    draw = L.batch_bits(count)
    seeds = L.batch_seeds(count)
    ops = []
    for i in range(0, count, stride):
        (ba, bb), (sa, sb) = draw[i], seeds[i]
        a, b = L.random_operand(ba, sa), L.random_operand(bb, sb)
        ops.append((a, b, np.zeros(a.size + b.size + 4, np.uint32)))
    return draw, ops


def batch(threads: int, count: int, stride: int) -> dict:
    """The sample as a sequential loop of multiplies (the reference's operator* loop) on the
    reference's routes: brute and FFT single-threaded, the NTT on the threaded pool."""
    M = oracle.mt_lib()
    M.cpu_init_thread_pool(threads)
    draw, ops = batch_ops(count, stride)
    bits = sum(ba + bb for ba, bb in draw[::stride])
    routes = {r: 0 for r in ROUTES}
    for a, b, _ in ops:
        routes[ROUTES[M.cpu_route(a.size, b.size)]] += 1
    for a, b, o in ops[:8]:  # warm-up: pool, pages
        _mul(M, a, b, o)
    t0 = time.perf_counter()
    for a, b, o in ops:
        _mul(M, a, b, o)
    seq = time.perf_counter() - t0
    return {"threads": int(M.cpu_pool_threads()), "pairs_timed": len(ops), "stride": stride,
            "routes": routes, "sample_bits": int(bits), "sequential_ms": seq * 1e3,
            "sequential_Gbit_s": bits / seq / 1e9, "full_batch_bits": int(sum(ba + bb for ba, bb in draw))}


def c1(threads: int, reps: int = 5) -> dict:
    """2^22 x 2^22 bits (A seed, B seed + 1) and the square of A, as the reference routes them
    (its FFT: 8,388,608 total bits), median of `reps` after a warm-up."""
    M = oracle.mt_lib()
    M.cpu_init_thread_pool(threads)
    a = L.random_operand(1 << 22, L.SEED)
    b = L.random_operand(1 << 22, L.SEED + 1)
    out = np.zeros(a.size + b.size + 4, np.uint32)
    res = {"threads": int(M.cpu_pool_threads()), "route": ROUTES[M.cpu_route(a.size, b.size)]}
    for name, y, same in (("product", b, 0), ("square", a, 1)):
        n = ctypes.c_size_t()
        p = ctypes.POINTER(ctypes.c_uint32)
        times = []
        for r in range(reps + 1):
            t0 = time.perf_counter()
            M.cpu_mul_digits(a.ctypes.data_as(p), a.size, y.ctypes.data_as(p), y.size, same,
                             out.ctypes.data_as(p), ctypes.byref(n))
            if r:
                times.append(time.perf_counter() - t0)
        res[name + "_ms"] = float(np.median(times)) * 1e3
    return res


def batch_across(threads: int, count: int, stride: int) -> dict:
    """Pairs spread over `threads` Python threads, each multiply on a one-thread pool (the
    ctypes call releases the GIL)."""
    M = oracle.mt_lib()
    M.cpu_init_thread_pool(1)
    draw, ops = batch_ops(count, stride)
    bits = sum(ba + bb for ba, bb in draw[::stride])
    nxt = [0]
    lock = threading.Lock()

    def work():
        while True:
            with lock:
                i = nxt[0]
                nxt[0] += 1
            if i >= len(ops):
                return
            _mul(M, *ops[i])

    t0 = time.perf_counter()
    ts = [threading.Thread(target=work) for _ in range(threads)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    dt = time.perf_counter() - t0
    return {"workers": threads, "pairs_timed": len(ops), "across_ms": dt * 1e3, "across_Gbit_s": bits / dt / 1e9}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, default=1)
    ap.add_argument("--bits", type=int, default=28, help="log2 bits of each balanced operand")
    ap.add_argument("--max-s", type=float, default=25.0)
    ap.add_argument("--batch", type=int, default=0, help="time the batched workload of this many pairs")
    ap.add_argument("--stride", type=int, default=16)
    ap.add_argument("--across", action="store_true", help="pairs across threads (one-thread pool)")
    ap.add_argument("--c1", action="store_true", help="2^22 x 2^22 bits and the square of A")
    args = ap.parse_args()
    if args.c1:
        r = c1(args.threads)
    elif args.batch and args.across:
        r = batch_across(args.threads, args.batch, args.stride)
    elif args.batch:
        r = batch(args.threads, args.batch, args.stride)
    else:
        r = balanced(args.threads, args.bits, args.max_s)
    print(json.dumps(r), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
