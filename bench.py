"""Benchmark of the MI355X NTT multiply engine on its headline workload.

Workload: one 2^28-bit x 2^28-bit multiply (9,586,981 x 9,586,981 limbs, NTT length 2^25) of
SplitMix64 operands (A seed 0x9e3779b97f4a7c15, B seed+1), operands and product resident in HBM.
A step is one such multiply through the C ABI (bi_hip_mul_device).  `value` is input bits per
second of the whole timed region (Gbit/s).  Each timed product is checked exactly (residues
mod six primes) outside the timed region.

Prints one JSON line.  With --gpus N > 1 it exits: the engine runs on one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

METRIC = "2^28-bit mul ms + batched product-tree Gbit/s, learned vs plain, bit-exact"
HBM_PEAK_GBS = 8000.0


def cpu_info() -> dict:
    info = {"nproc": os.cpu_count(), "model": platform.processor()}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            v = v.strip()
            if k == "Model name":
                info["model"] = v
            elif k == "Socket(s)":
                info["sockets"] = int(v)
            elif k == "Core(s) per socket":
                info["cores_per_socket"] = int(v)
            elif k == "Thread(s) per core":
                info["threads_per_core"] = int(v)
    except OSError:
        pass
    return info


def cpu_baseline(a: np.ndarray, b: np.ndarray, bits: int, max_s: float = 25.0) -> dict:
    """The threaded AVX2 restatement of the reference's NTT path (oracle/libcpu_mt.so) on the
    host cores, timed around mul_digits with steady-clock semantics; bounded sample."""
    import ctypes
    from oracle import oracle
    info = cpu_info()
    phys = info.get("sockets", 1) * info.get("cores_per_socket", os.cpu_count() or 1)
    threads = max(1, min(phys, 16))
    M = oracle.mt_lib()
    M.cpu_init_thread_pool(threads)
    out = np.zeros(a.size + b.size + 4, np.uint32)
    n = ctypes.c_size_t()
    pa, pb, po = (x.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)) for x in (a, b, out))
    t0 = time.perf_counter()
    M.cpu_mul_digits(pa, a.size, pb, b.size, 0, po, ctypes.byref(n))  # warm-up: pool + pages
    warm = time.perf_counter() - t0
    times = []
    while not times or (sum(times) + warm < max_s and len(times) < 3):
        t0 = time.perf_counter()
        M.cpu_mul_digits(pa, a.size, pb, b.size, 0, po, ctypes.byref(n))
        times.append(time.perf_counter() - t0)
    med = float(np.median(times))
    return {"value": 2 * bits / med / 1e9, "unit": "Gbit/s", "cores": int(M.cpu_pool_threads()),
            "kind": "port", "ms": med * 1e3,
            "sample": f"{len(times)} timed 2^{bits.bit_length() - 1}x2^{bits.bit_length() - 1}-bit "
                      f"multiplies after 1 warm-up, median; threaded AVX2 restatement of the reference "
                      f"NTT path; host {info.get('model')}, {info.get('sockets')} socket(s) x "
                      f"{info.get('cores_per_socket')} cores x {info.get('threads_per_core')} threads"}


def batch_cpu_baseline(pairs_count: int, cores: int, stride: int = 16) -> dict:
    """The batched workload on the CPU restatement, on the reference's routes (brute and FFT
    single-threaded, NTT on the threaded pool already created by cpu_baseline): every 16th pair
    of the same size draw, SplitMix64 operands, multiplied one after another as the reference's
    operator* loop does; a bounded sample (about 10 s)."""
    from oracle import cpu_time
    r = cpu_time.batch(cores, pairs_count, stride)
    return {"value": r["sequential_Gbit_s"], "unit": "Gbit/s", "cores": r["threads"], "kind": "port",
            "ms": r["sequential_ms"], "routes": r["routes"],
            "sample": f"{r['pairs_timed']} of {pairs_count} pairs (every {stride}th, {r['sample_bits']} "
                      f"input bits), sequential loop on the reference's routes"}


def c1_cpu() -> dict:
    """2^22 x 2^22 bits and the square of A on the CPU restatement (the reference's FFT route)."""
    from oracle import cpu_time
    return cpu_time.c1(0)


def batch_inputs(hip, L, pairs_count: int):
    """Operands of the batched workload in one device pool: log-uniform bit lengths in
    [2^12, 2^24] (SplitMix64 from the fixed seed), uniform random limbs from a seeded generator
    with the exact top bit set; pair i's operands at offa[i], offb[i] (its product's slot in an
    output pool of the same layout starts at offa[i]).  Also returns the generator, for samples
    drawn after the operands."""
    bits = L.batch_bits(pairs_count)
    la = np.array([-(-ba // 28) for ba, _ in bits], np.int64)
    lb = np.array([-(-bb // 28) for _, bb in bits], np.int64)
    tot_in = int(la.sum() + lb.sum())
    rng = np.random.default_rng(L.SEED & 0xFFFFFFFF)
    host = rng.integers(0, 1 << 28, tot_in, dtype=np.uint32)
    offa = np.concatenate([[0], np.cumsum(la + lb)[:-1]])
    offb = offa + la
    for i, (ba, bb) in enumerate(bits):  # exact bit lengths
        for off, nb, n in ((offa[i], ba, la[i]), (offb[i], bb, lb[i])):
            r = nb - 28 * (n - 1)
            host[off + n - 1] = (host[off + n - 1] & ((1 << r) - 1)) | (1 << (r - 1))
    pool = hip.Buffer.from_array(host)
    return bits, la, lb, offa, offb, host, pool, tot_in, rng


def batch_only(hip, engine, L, pairs_count: int, mode: str, steps: int) -> int:
    """The batched workload in one plan mode only (one warm-up batch, then `steps` timed
    batches): the command that kernel traces of the batch are taken with."""
    from bigint_amd import planner
    engine.init(0)
    if os.path.exists(planner.DEFAULT_PATH):
        planner.load()
    m = {"plain": engine.PLAN_FIXED, "table": engine.PLAN_TABLE, "learned": engine.PLAN_LEARNED}[mode]
    bits, la, lb, offa, offb, host, pool, tot_in, _ = batch_inputs(hip, L, pairs_count)
    out = hip.Buffer(4 * tot_in)
    stream = hip.Stream()
    pairs = [(pool.ptr + 4 * int(offa[i]), int(la[i]), pool.ptr + 4 * int(offb[i]), int(lb[i]),
              out.ptr + 4 * int(offa[i]), int(la[i] + lb[i])) for i in range(pairs_count)]
    engine.mul_batch_device(pairs, m, stream.handle)
    stream.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        engine.mul_batch_device(pairs, m, stream.handle)
    stream.synchronize()
    dt = (time.perf_counter() - t0) / steps
    print(json.dumps({"batch_mode": mode, "pairs": pairs_count, "ms": round(dt * 1e3, 3),
                      "Gbit_s": round(sum(a + b for a, b in bits) / dt / 1e9, 2)}), flush=True)
    return 0


def valu_rates() -> int:
    """Issue rate of v_add_u32, v_mul_lo_u32, v_mul_hi_u32 and v_mad_u64_u32 (libbigint_rates.so:
    eight independent chains per thread, eight waves per SIMD), in wave-instructions per CU per
    clock: the peak the kernels' VALU issue fractions are stated against."""
    import ctypes
    lib = ctypes.CDLL(os.path.join(ROOT, "bigint_amd", "libbigint_rates.so"))
    cus, khz = ctypes.c_int(), ctypes.c_int()
    if lib.bi_rates_device(ctypes.byref(cus), ctypes.byref(khz)):
        raise RuntimeError("bi_rates_device failed")
    blocks, iters = 16 * cus.value, 20000
    forms = [("v_add_u32", 8), ("v_mul_lo_u32", 8), ("v_mul_hi_u32 + v_xor_b32", 16),
             ("v_mad_u64_u32 + v_xor_b32", 16)]
    rows = []
    for op, (name, per_iter) in enumerate(forms):
        ms = ctypes.c_float()
        rc = lib.bi_rates_run(op, blocks, iters, ctypes.byref(ms))
        if rc:
            raise RuntimeError(f"bi_rates_run({op}) failed: {rc}")
        per_ns = blocks * 4 * iters * per_iter / (ms.value * 1e-3) / cus.value / 1e9
        rows.append({"form": name, "ms": round(ms.value, 3), "wave_instr_per_CU_per_ns": round(per_ns, 3)})
    print(json.dumps({"valu_rates": rows, "CUs": cus.value, "peak_clock_MHz": khz.value / 1e3,
                      "blocks": blocks, "threads_per_block": 256, "iters": iters}), flush=True)
    return 0


def batch_leg(hip, engine, L, pairs_count: int, steps: int) -> dict:
    """The batched product-tree workload: `pairs_count` independent products whose operand sizes
    are log-uniform in [2^12, 2^24] bits (SplitMix64 from the fixed seed), uniform random limbs
    from a seeded generator with the exact top bit set, resident in HBM.  Timed per plan mode;
    every mode's products must equal the plain mode's limb for limb, and a sample is checked
    against residues on the host."""
    bits, la, lb, offa, offb, host, pool, tot_in, rng = batch_inputs(hip, L, pairs_count)
    outs = {m: hip.Buffer(4 * tot_in) for m in (0, 1, 2)}
    stream = hip.Stream()
    res = {}
    for mode, name in ((engine.PLAN_FIXED, "plain"), (engine.PLAN_TABLE, "table"),
                       (engine.PLAN_LEARNED, "learned")):
        base, ob = pool.ptr, outs[mode].ptr
        pairs = [(base + 4 * int(offa[i]), int(la[i]), base + 4 * int(offb[i]), int(lb[i]),
                  ob + 4 * int(offa[i]), int(la[i] + lb[i])) for i in range(pairs_count)]
        engine.mul_batch_device(pairs, mode, stream.handle)  # warm-up (plans, workspace)
        stream.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            engine.mul_batch_device(pairs, mode, stream.handle)
        stream.synchronize()
        dt = (time.perf_counter() - t0) / steps
        res[name] = {"ms": round(dt * 1e3, 3), "Gbit_s": round(sum(a + b for a, b in bits) / dt / 1e9, 2)}
        if mode != engine.PLAN_FIXED:
            # latency of one batch on its own (host preparation included, nothing to overlap it
            # with); the timed loop above queues batches back to back, where a batch's host
            # preparation overlaps the previous batch's device work
            lat = []
            for _ in range(3):
                t1 = time.perf_counter()
                engine.mul_batch_device(pairs, mode, stream.handle)
                stream.synchronize()
                lat.append(time.perf_counter() - t1)
            res[name]["single_batch_ms"] = round(min(lat) * 1e3, 3)
    # the learned mode again with stream launches instead of one captured graph per batch
    engine.lib().bi_hip_set_graphs(0)
    try:
        base, ob = pool.ptr, outs[engine.PLAN_LEARNED].ptr
        lp = [(base + 4 * int(offa[i]), int(la[i]), base + 4 * int(offb[i]), int(lb[i]),
               ob + 4 * int(offa[i]), int(la[i] + lb[i])) for i in range(pairs_count)]
        engine.mul_batch_device(lp, engine.PLAN_LEARNED, stream.handle)
        stream.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            engine.mul_batch_device(lp, engine.PLAN_LEARNED, stream.handle)
        stream.synchronize()
        dt = (time.perf_counter() - t0) / steps
        res["learned_stream_launches"] = {"ms": round(dt * 1e3, 3),
                                          "Gbit_s": round(sum(a + b for a, b in bits) / dt / 1e9, 2)}
    finally:
        engine.lib().bi_hip_set_graphs(1)
    # per-kernel device time of one learned-mode batch (engine's event timing)
    engine.profile(True)
    base, ob = pool.ptr, outs[engine.PLAN_LEARNED].ptr
    engine.mul_batch_device([(base + 4 * int(offa[i]), int(la[i]), base + 4 * int(offb[i]), int(lb[i]),
                              ob + 4 * int(offa[i]), int(la[i] + lb[i])) for i in range(pairs_count)],
                            engine.PLAN_LEARNED, stream.handle)
    stream.synchronize()
    prof = engine.profile_read()
    engine.profile(False)
    res["learned_kernels"] = {k: {"ms": round(v["ms"], 3), "launches": v["launches"]}
                              for k, v in prof.items() if v["launches"]}
    o = {m: outs[m].download(np.uint32, tot_in) for m in (0, 1, 2)}
    same = bool(np.array_equal(o[0], o[1]) and np.array_equal(o[0], o[2]))
    ok = True
    for i in rng.choice(pairs_count, size=min(48, pairs_count), replace=False):
        a = host[offa[i]:offa[i] + la[i]]
        b = host[offb[i]:offb[i] + lb[i]]
        p = o[0][offa[i]:offa[i] + la[i] + lb[i]]
        for q in L.RESIDUE_PRIMES[:3]:
            ok &= L.residue_mod(p, q) == L.residue_mod(a, q) * L.residue_mod(b, q) % q
    res["pairs"] = pairs_count
    res["input_bits"] = int(sum(a + b for a, b in bits))
    res["learned_vs_plain"] = round(res["plain"]["ms"] / res["learned"]["ms"], 3)
    res["table_vs_plain"] = round(res["plain"]["ms"] / res["table"]["ms"], 3)
    res["exact"] = bool(same and ok)
    return res


def profiled_traffic(kernel_class: str):
    """HBM bytes per launch (FETCH_SIZE + WRITE_SIZE, scaled) and VALU issue fraction of a kernel
    class at the 2^28-bit workload, from the committed rocprofv3 summary; None if absent."""
    path = os.path.join(ROOT, "profiles", "ntt_mul_2e28_traffic.txt")
    if not os.path.exists(path):
        return None
    rows = []
    with open(path) as f:
        for line in f:
            if line.startswith("#") or not line.strip():
                continue
            cls, kern, fb, wb, vi, frac, us = line.split()
            if cls == kernel_class:
                rows.append((kern, int(fb) + int(wb), float(frac), float(us)))
    if not rows:
        return None
    kern, traffic, frac, us = max(rows, key=lambda r: r[3])
    return {"kernel": kern, "traffic": traffic, "valu_issue_frac": frac}


def roofline(dom: str, achieved: float, bytes_per_launch: float, avg_ms: float, prof_t) -> dict:
    """Roofline of the dominant kernel class.  `achieved` is its algorithmic HBM bytes per launch
    over its live event-timed average launch; `frac` is that against the 8 TB/s HBM peak.  The
    kernels are integer modular arithmetic, so the competing bound is the VALU: `valu_issue_frac`
    is VALU wave-instructions per CU-cycle from the committed rocprofv3 counters of the same
    command (one wave-instruction per CU-cycle is the issue peak).  `binding` names the resource
    closer to its peak."""
    hbm_frac = achieved / HBM_PEAK_GBS
    valu = prof_t["valu_issue_frac"] if prof_t else None
    return {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(hbm_frac, 4),
            "traffic": prof_t["traffic"] if prof_t else None,
            "traffic_source": "profiles/ntt_mul_2e28_traffic.txt (rocprofv3 FETCH_SIZE/WRITE_SIZE)"
                              if prof_t else None,
            "valu_issue_frac": valu,
            "binding": ("valu" if valu is not None and valu > hbm_frac else "hbm"),
            "kernel_name": prof_t["kernel"] if prof_t else None,
            "alg_bytes_per_launch": bytes_per_launch, "avg_launch_ms": round(avg_ms, 4)}


def size_sweep(hip, engine, L) -> int:
    """Per size: device-resident time (operands and product in HBM, bi_hip_mul_device) and
    end-to-end time through bi_hip_mul (host limbs in, host limbs out, copies included); median
    of 5 after 2 warm-ups; every product checked against residues."""
    engine.init(0)
    from bigint_amd import planner
    if os.path.exists(planner.DEFAULT_PATH):
        planner.load()
    rows = []
    for name, ba, bb in [("2^22 x 2^22", 22, 22), ("2^22 squared", 22, None), ("2^24 x 2^24", 24, 24),
                         ("2^26 x 2^26", 26, 26), ("2^28 x 2^28", 28, 28), ("2^28 x 2^20", 28, 20)]:
        a = L.random_operand(1 << ba, L.SEED)
        b = a if bb is None else L.random_operand(1 << bb, L.SEED + 1)
        da = hip.Buffer.from_array(a)
        db = da if bb is None else hip.Buffer.from_array(b)
        dout = hip.Buffer(4 * (a.size + b.size))
        st = hip.Stream()
        e0, e1 = hip.Event(), hip.Event()
        dev = []
        for r in range(20):  # 10 warm-ups bring the GPU to its steady clock; median of 10
            e0.record(st)
            engine.mul_device(da.ptr, a.size, db.ptr, b.size, bb is None, dout.ptr, a.size + b.size, st.handle)
            e1.record(st)
            e1.synchronize()
            if r >= 10:
                dev.append(e0.elapsed_ms(e1))
        p = dout.download(np.uint32, a.size + b.size)
        ok = all(L.residue_mod(p, q) == L.residue_mod(a, q) * L.residue_mod(b, q) % q
                 for q in L.RESIDUE_PRIMES[:3])
        e2e = []
        for r in range(7):
            t0 = time.perf_counter()
            h = engine.mul(a, None if bb is None else b)
            if r >= 2:
                e2e.append((time.perf_counter() - t0) * 1e3)
        ok = ok and np.array_equal(L.strip(p), h)
        bits = (1 << ba) + (1 << (ba if bb is None else bb))
        dm, em = float(np.median(dev)), float(np.median(e2e))
        row = {"workload": name, "transform_log2": engine.transform_log2(a.size, b.size),
               "device_ms": round(dm, 4), "device_Gbit_s": round(bits / dm / 1e6, 2),
               "end_to_end_ms": round(em, 4), "end_to_end_Gbit_s": round(bits / em / 1e6, 2), "exact": bool(ok)}
        rows.append(row)
        print(json.dumps(row), file=sys.stderr, flush=True)
        for buf in (da, db, dout):
            buf.free()
    copies = host_copy_breakdown(hip, engine, L)
    print(json.dumps({"size_sweep": rows, "host_copies_2e28": copies}), flush=True)
    return 0 if all(r["exact"] for r in rows) else 1


def host_copy_breakdown(hip, engine, L) -> dict:
    """Where the end-to-end time of the 2^28-bit host-pointer product goes: both operands up
    (bi_hip_copy_to_device, one call), the product down (bi_hip_copy_to_host) into memory already
    faulted in and into fresh pages, and bi_hip_mul with a faulted-in output; medians of 5."""
    import ctypes
    Lb = engine.lib()
    a = L.random_operand(1 << 28, L.SEED)
    b = L.random_operand(1 << 28, L.SEED + 1)
    n = a.size + b.size
    u32p = ctypes.POINTER(ctypes.c_uint32)
    dbuf = hip.Buffer(4 * n)
    out = np.ones(n, np.uint32)

    def med(f, reps=5):
        t = []
        for _ in range(reps):
            t0 = time.perf_counter()
            f()
            t.append((time.perf_counter() - t0) * 1e3)
        return round(float(np.median(t)), 3)

    def check(rc):
        if rc:
            raise engine.EngineError(rc, "host copy breakdown")

    srcs = (ctypes.c_void_p * 2)(a.ctypes.data, b.ctypes.data)
    sz = (ctypes.c_size_t * 2)(a.size * 4, b.size * 4)
    up = med(lambda: check(Lb.bi_hip_copy_to_device(dbuf.ptr, srcs, sz, 2)))
    dst = (ctypes.c_void_p * 1)(out.ctypes.data)
    dsz = (ctypes.c_size_t * 1)(n * 4)
    down = med(lambda: check(Lb.bi_hip_copy_to_host(dst, dsz, 1, dbuf.ptr)))

    def fresh_down():
        f = np.zeros(n, np.uint32)
        d = (ctypes.c_void_p * 1)(f.ctypes.data)
        t0 = time.perf_counter()
        check(Lb.bi_hip_copy_to_host(d, dsz, 1, dbuf.ptr))
        return (time.perf_counter() - t0) * 1e3
    down_fresh = round(float(np.median([fresh_down() for _ in range(5)])), 3)
    k = ctypes.c_size_t()
    for _ in range(2):
        check(Lb.bi_hip_mul(a.ctypes.data_as(u32p), a.size, b.ctypes.data_as(u32p), b.size, 0,
                            out.ctypes.data_as(u32p), out.size, ctypes.byref(k)))
    mul_ms = med(lambda: check(Lb.bi_hip_mul(a.ctypes.data_as(u32p), a.size, b.ctypes.data_as(u32p), b.size,
                                             0, out.ctypes.data_as(u32p), out.size, ctypes.byref(k))))
    dbuf.free()
    return {"bytes_each_way": 4 * n, "to_device_ms": up, "to_host_ms": down,
            "to_host_fresh_pages_ms": down_fresh, "bi_hip_mul_faulted_in_output_ms": mul_ms,
            "to_device_GB_s": round(4 * n / up / 1e6, 1), "to_host_GB_s": round(4 * n / down / 1e6, 1)}


def cpu_sweep() -> int:
    """CPU restatement timings, one process per thread count (the pool's first call wins)."""
    info = cpu_info()
    phys = info.get("sockets", 1) * info.get("cores_per_socket", os.cpu_count() or 1)
    runs = []
    def run(args):
        out = subprocess.run([sys.executable, "-m", "oracle.cpu_time"] + args, cwd=ROOT,
                             capture_output=True, text=True, timeout=900)
        if out.returncode:
            raise RuntimeError(out.stderr)
        r = json.loads(out.stdout.strip().splitlines()[-1])
        r["args"] = " ".join(args)
        runs.append(r)
        print(json.dumps(r), file=sys.stderr, flush=True)
    for t in sorted({1, 8, min(phys, 16)}):
        run(["--threads", str(t), "--bits", "28", "--max-s", "60"])
    run(["--threads", str(min(phys, 16)), "--batch", "4096", "--stride", "16"])
    run(["--threads", str(min(phys, 16)), "--batch", "4096", "--stride", "16", "--across"])
    print(json.dumps({"cpu_sweep": runs, "host": info}), flush=True)
    return 0


def copy_calibration(hip) -> int:
    """A plain 1 GiB device-to-device copy, three times: rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE on
    this command gives the counters' reading for a known byte count (1 GiB read + 1 GiB written)."""
    import ctypes
    src = hip.Buffer(1 << 30)
    dst = hip.Buffer(1 << 30)
    src.zero()
    for _ in range(3):
        rc = hip.rt().hipMemcpy(ctypes.c_void_p(dst.ptr), ctypes.c_void_p(src.ptr), 1 << 30, hip.D2D)
        if rc:
            raise RuntimeError(f"hipMemcpy D2D failed: {rc}")
    hip.synchronize()
    print(json.dumps({"copy_bytes_each_way": 1 << 30, "launches": 3}), flush=True)
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--bits", type=int, default=28, help="log2 of the operand size in bits")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--batch-pairs", type=int, default=4096,
                    help="products in the batched product-tree leg (0 skips it)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only a 1 GiB device copy (reference point for rocprofv3 byte counters)")
    ap.add_argument("--sweep", action="store_true",
                    help="run only the size sweep: 2^22 (and its square), 2^24, 2^26, 2^28 and "
                         "2^28 x 2^20, device-resident and end-to-end through the host API")
    ap.add_argument("--cpu-sweep", action="store_true",
                    help="run only the CPU restatement: 2^28 x 2^28 at 1, 8 and all physical cores, "
                         "and a 1/16 sample of the batched workload, sequential and across threads")
    ap.add_argument("--batch-only", choices=("plain", "table", "learned"), default=None,
                    help="run only the batched workload in one plan mode, --steps batches")
    ap.add_argument("--valu-rates", action="store_true",
                    help="run only the issue-rate microbenchmark of the integer VALU forms")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this engine runs on one GPU; --gpus {args.gpus} is not supported "
              "(a product-tree level would shard by pairs, one huge product by four-step column "
              "blocks with an all-to-all per transform; neither is built)", file=sys.stderr)
        return 2

    if args.cpu_sweep:
        return cpu_sweep()

    from bigint_amd import engine, hip
    from bigint_amd import limbs as L

    if args.copy_calibration:
        return copy_calibration(hip)
    if args.valu_rates:
        return valu_rates()
    if args.batch_only:
        return batch_only(hip, engine, L, args.batch_pairs, args.batch_only, args.steps)
    if args.sweep:
        return size_sweep(hip, engine, L)

    bits = 1 << args.bits
    engine.init(0)
    from bigint_amd import planner
    planner_loaded = os.path.exists(planner.DEFAULT_PATH)
    if planner_loaded:
        planner.load()
    a = L.random_operand(bits, L.SEED)
    b = L.random_operand(bits, L.SEED + 1)
    da = hip.Buffer.from_array(a)
    db = hip.Buffer.from_array(b)
    dout = hip.Buffer(4 * (a.size + b.size))
    stream = hip.Stream()

    def step():
        engine.mul_device(da.ptr, a.size, db.ptr, b.size, False, dout.ptr, a.size + b.size,
                          stream.handle)

    # Warm-up: at least three unprofiled steps first (the first launches pay code-object loads and
    # the GPU has not reached its steady clock), then up to five steps that bracket every launch
    # with events: the per-kernel breakdown and the dominant kernel class, from warm steps only.
    # The timed steps bracket only that class, so the other launches run without event records.
    # The unprofiled part also lasts at least 0.3 s: with only a handful of 1-ms steps the GPU is
    # still raising its clock and the first timed steps read up to 10 % slow.
    n_prof = max(1, min(5, args.warmup - 3))
    n_plain = max(3, args.warmup - n_prof)
    t_warm = time.perf_counter()
    for _ in range(n_plain):
        step()
    hip.synchronize()
    while time.perf_counter() - t_warm < 0.3:
        for _ in range(16):
            step()
        n_plain += 16
        hip.synchronize()
    prof_steps = n_prof
    engine.profile(True)
    for _ in range(n_prof):
        step()
    hip.synchronize()
    prof_all = engine.profile_read()
    engine.profile(False)
    dom = max((k for k in prof_all if prof_all[k]["launches"]), key=lambda k: prof_all[k]["ms"])
    engine.profile_kinds([dom])
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    hip.synchronize()
    wall = time.perf_counter() - t0
    prof = engine.profile_read()
    engine.profile(False)

    # exactness of the timed product (identical every step), outside the timed region
    p = dout.download(np.uint32, a.size + b.size)
    exact = all(L.residue_mod(p, q) == L.residue_mod(a, q) * L.residue_mod(b, q) % q
                for q in L.RESIDUE_PRIMES)

    ms = wall / args.steps * 1e3
    value = args.steps * 2 * bits / wall / 1e9
    d = prof[dom]  # timed steps: only the dominant class is bracketed
    avg_ms = d["ms"] / d["launches"]
    bytes_per_launch = d["bytes"] / d["launches"]
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9
    prof_t = profiled_traffic(dom) if args.bits == 28 else None
    n = 1 << engine.transform_log2(a.size, b.size)
    plan_id = planner.choose([(int(a.size), int(b.size), 1)], engine.PLAN_TABLE)[0]
    kernels = {k: {"avg_ms": v["ms"] / v["launches"], "launches_per_step": v["launches"] / prof_steps,
                   "alg_GBs": v["bytes"] / (v["ms"] * 1e-3) / 1e9 if v["ms"] else None}
               for k, v in prof_all.items() if v["launches"]}
    line = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Gbit/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {"workload": f"2^{args.bits}-bit x 2^{args.bits}-bit multiply, one product per step, "
                               f"device-resident", "limbs": [int(a.size), int(b.size)],
                   "plain_ntt_length": n, "plan": plan_id, "exact": bool(exact)},
        "roofline": roofline(dom, achieved, bytes_per_launch, avg_ms, prof_t),
        "kernels": kernels,
        "kernels_source": f"{prof_steps} warm-up steps after {n_plain} unprofiled ones, every launch "
                          f"bracketed by HIP events",
        "warmup_run": n_plain + n_prof,
        "mul_ms": round(ms, 4),
    }
    if args.batch_pairs:
        line["batch"] = batch_leg(hip, engine, L, args.batch_pairs, 2)
        line["batch"]["planner_loaded"] = planner_loaded
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(a, b, bits)
        line["speedup_vs_cpu"] = round(value / line["cpu_baseline"]["value"], 2)
        if args.batch_pairs:
            line["batch"]["cpu_baseline"] = batch_cpu_baseline(args.batch_pairs, line["cpu_baseline"]["cores"])
            cb = line["batch"]["cpu_baseline"]["value"]
            line["batch"]["learned_vs_cpu"] = round(line["batch"]["learned"]["Gbit_s"] / cb, 1) if cb else None
        line["cpu_c1"] = c1_cpu()
    print(json.dumps(line), flush=True)
    return 0 if exact and line.get("batch", {}).get("exact", True) else 1


if __name__ == "__main__":
    sys.exit(main())
