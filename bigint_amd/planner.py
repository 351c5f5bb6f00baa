"""Plan selection for the NTT multiply engine: timing collection, autotuned table, and the MLP
cost model trained on the GPU with the engine's own HIP kernels (bi_hip_mlp_*).

Plan ids (include/bigint_hip.h): 0/1/2 unchunked with a 4096/2048/1024-point block pass;
3/4/5 longer operand cut into 2^c / 2^(c+1) / 2^(c+2)-limb chunks; 6/7 longer operand cut into 2 / 3
chunks.  Chunked plans transform the shorter operand once and reuse it.  The reference picks its
algorithm with fixed size thresholds (include/bigint/mul.h:80, 153, 192); here the choice is
measured (table) or predicted (MLP), and every plan computes the same exact product.

Features of a product: x = (log2 bits of a, log2 bits of b, log2 batch), where batch is the
number of products of that transform length launched together.  The MLP predicts log2 of the
per-product time for each plan; the chosen plan is the argmin over the plans allowed.
"""
from __future__ import annotations

import ctypes
import json
import math
import os
import time

import numpy as np

from . import engine

K = 8  # plan ids
TABLE_SHAPE = (19, 19, 13)
HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_PATH = os.path.join(HERE, "plans", "planner.json")


def allowed(la: int, lb: int, square: bool = False) -> list[int]:
    return [k for k in range(K) if engine.lib().bi_hip_plan_allowed(la, lb, int(square), k)]


def features(bits_a: float, bits_b: float, batch: float) -> list[float]:
    return [math.log2(bits_a), math.log2(bits_b), math.log2(max(1.0, batch))]


# ---- timing collection (GPU) -------------------------------------------------------------------

class Timer:
    """Times planned device batches of k identical-shape products on a dedicated stream."""

    def __init__(self, max_limbs: int = 1 << 25, seed: int = 1):
        from . import hip
        self.hip = hip
        rng = np.random.default_rng(seed)
        self.pool = hip.Buffer.from_array(rng.integers(1, 1 << 28, max_limbs, dtype=np.uint32))
        self.out_words = 2 * max_limbs + 64
        self.out = hip.Buffer(4 * self.out_words)
        self.max_limbs = max_limbs
        self.stream = hip.Stream()
        self.e0, self.e1 = hip.Event(), hip.Event()
        engine.lib().bi_hip_mul_batch_device_planned.argtypes = [
            ctypes.POINTER(engine.DPair), ctypes.c_size_t, ctypes.POINTER(ctypes.c_int32), ctypes.c_void_p]

    def run(self, la: int, lb: int, k: int, plan: int, reps: int = 2) -> float:
        """Median per-product device time (us) of k products la x lb under `plan`."""
        if k * (la + lb) > self.out_words or la > self.max_limbs or lb > self.max_limbs:
            raise ValueError("shape too large for the timing pools")
        base, obase = self.pool.ptr, self.out.ptr
        pairs = []
        for i in range(k):
            offa = (i * 7919) % max(1, self.max_limbs - la)
            offb = (i * 104729 + 17) % max(1, self.max_limbs - lb)
            pairs.append(engine.DPair(base + 4 * offa, la, base + 4 * offb, lb,
                                      obase + 4 * i * (la + lb), la + lb))
        arr = (engine.DPair * k)(*pairs)
        ids = (ctypes.c_int32 * k)(*([plan] * k))
        L = engine.lib()
        times = []
        for r in range(reps + 1):
            self.e0.record(self.stream)
            rc = L.bi_hip_mul_batch_device_planned(arr, k, ids, self.stream.handle)
            self.e1.record(self.stream)
            if rc:
                raise engine.EngineError(rc, "planned batch")
            self.e1.synchronize()
            if r:
                times.append(self.e0.elapsed_ms(self.e1) * 1e3 / k)
        return float(np.median(times))


def sample_shapes(count: int, seed: int, lo: float = 10.0, hi: float = 28.0, max_batch_log: int = 10,
                  mem_limit_bytes: float = 24e9):
    """Log-uniform operand sizes in [2^lo, 2^hi] bits and log-uniform batch sizes up to
    2^max_batch_log (product-tree levels put hundreds of products of one length in a launch, where
    the GPU is saturated and the best plan differs from a lone product's), with the batch capped
    so the residues of one launch stay within mem_limit_bytes."""
    rng = np.random.default_rng(seed)
    out = []
    while len(out) < count:
        ba = float(2 ** rng.uniform(lo, hi))
        bb = float(2 ** rng.uniform(lo, hi))
        la, lb = max(1, int(ba) // 28 + 1), max(1, int(bb) // 28 + 1)
        lg = engine.transform_log2(la, lb)
        if lg < 0:
            continue
        k = int(2 ** rng.uniform(0, max_batch_log))
        while k > 1 and k * 24 * (1 << lg) > mem_limit_bytes:
            k //= 2
        out.append((la, lb, k))
    return out


def collect(shapes, timer: Timer, log=None, reps: int = 2):
    """Time every allowed plan for every shape; returns records."""
    recs = []
    t0 = time.time()
    for i, (la, lb, k) in enumerate(shapes):
        for p in allowed(la, lb):
            us = timer.run(la, lb, k, p, reps)
            recs.append({"la": la, "lb": lb, "k": k, "plan": p, "us": us})
        if log and (i % 200 == 0):
            log(f"collected {i}/{len(shapes)} shapes, {len(recs)} runs, {time.time() - t0:.0f}s")
    return recs


# ---- table -------------------------------------------------------------------------------------

def group_best(recs):
    """{(la, lb, k): {plan: us}} from records."""
    d = {}
    for r in recs:
        d.setdefault((r["la"], r["lb"], r["k"]), {})[r["plan"]] = r["us"]
    return d


def build_table(recs, heuristic: dict | None = None) -> np.ndarray:
    """Bucket by rounded log2 features; each cell takes the plan with the lowest mean log-time
    ratio to the per-shape best among shapes in the cell (plans not allowed for a shape count as
    2^10 times slower).  With `heuristic` ({(la, lb, k): the plan id the engine's built-in cost
    heuristic picks}), the heuristic itself is a candidate too, stored as -1: the engine answers
    -1 with the heuristic's exact-size choice, so a cell whose shapes straddle a power-of-two
    boundary keeps it.  Empty cells copy the nearest filled cell."""
    shapes = group_best(recs)
    score = {}
    for key, plans in shapes.items():
        la, lb, k = key
        best = min(plans.values())
        f = features(28 * la, 28 * lb, k)
        cell = (min(18, max(0, round(f[0]) - 10)), min(18, max(0, round(f[1]) - 10)),
                min(12, max(0, round(f[2]))))
        s = score.setdefault(cell, {})
        for p in range(K):
            s.setdefault(p, []).append(math.log2(plans[p] / best) if p in plans else 10.0)
        h = heuristic.get(key) if heuristic is not None else None
        s.setdefault(-1, []).append(math.log2(plans[h] / best) if h in plans else 10.0)
    table = np.zeros(TABLE_SHAPE, dtype=np.int8)
    filled = np.zeros(TABLE_SHAPE, dtype=bool)
    for cell, s in score.items():
        cand = [p for p in s if p >= 0 or heuristic is not None]
        table[cell] = min(cand, key=lambda p: (np.mean(s[p]), p))
        filled[cell] = True
    idx_filled = np.argwhere(filled)
    if idx_filled.size:
        for idx in np.argwhere(~filled):
            d = np.abs(idx_filled - idx).sum(axis=1)
            table[tuple(idx)] = table[tuple(idx_filled[int(np.argmin(d))])]
    return table


# ---- MLP ----------------------------------------------------------------------------------------

def plan_cost(la: int, lb: int, plan: int, square: bool = False) -> float:
    """The engine's transform-count cost of a plan (bi_hip_plan_cost); -1 if it does not apply."""
    L = engine.lib()
    L.bi_hip_plan_cost.restype = ctypes.c_double
    L.bi_hip_plan_cost.argtypes = [ctypes.c_size_t, ctypes.c_size_t, ctypes.c_int, ctypes.c_int]
    return float(L.bi_hip_plan_cost(la, lb, int(square), plan))


def dataset(recs):
    """Inputs (log2 bits a, log2 bits b, log2 batch); targets log2(time / transform-count cost)
    per plan, so the model learns the smooth part (efficiency by size, batch and plan shape)
    while the engine adds back the cost's steps at powers of two when it chooses."""
    shapes = group_best(recs)
    X, T, M = [], [], []
    for (la, lb, k), plans in shapes.items():
        X.append(features(28 * la, 28 * lb, k))
        t = np.zeros(K, np.float32)
        m = np.zeros(K, np.float32)
        for p, us in plans.items():
            t[p] = math.log2(us) - math.log2(plan_cost(la, lb, p))
            m[p] = 1.0
        T.append(t)
        M.append(m)
    return np.array(X, np.float32), np.array(T, np.float32), np.array(M, np.float32)


def init_params(seed: int = 0) -> np.ndarray:
    rng = np.random.default_rng(seed)
    H = 64
    parts = [rng.normal(0, math.sqrt(2 / 3), (3, H)), np.zeros(H),
             rng.normal(0, math.sqrt(2 / H), (H, H)), np.zeros(H),
             rng.normal(0, math.sqrt(1 / H), (H, K)), np.zeros(K)]
    return np.concatenate([p.reshape(-1) for p in parts]).astype(np.float32)


def train_gpu(X: np.ndarray, T: np.ndarray, M: np.ndarray, epochs: int = 3000, lr: float = 3e-3,
              seed: int = 0, log=None, lr_end: float | None = None):
    """Full-batch Adam on the GPU through bi_hip_mlp_grad / bi_hip_mlp_adam.  Inputs are
    normalised; the learning rate falls from lr to lr_end on a cosine (constant when lr_end is
    None).  Returns (params, norm6, loss history)."""
    from . import hip
    mean = X.mean(axis=0)
    std = X.std(axis=0) + 1e-6
    Xn = ((X - mean) / std).astype(np.float32)
    L = engine.lib()
    NP = L.bi_hip_mlp_param_count(K)
    p0 = init_params(seed)
    assert p0.size == NP
    p = hip.Buffer.from_array(p0)
    m = hip.Buffer.from_array(np.zeros(NP, np.float32))
    v = hip.Buffer.from_array(np.zeros(NP, np.float32))
    g = hip.Buffer.from_array(np.zeros(NP + 1, np.float32))
    x = hip.Buffer.from_array(Xn)
    t = hip.Buffer.from_array(T.astype(np.float32))
    mk = hip.Buffer.from_array(M.astype(np.float32))
    st = hip.Stream()
    hist = []
    for ep in range(1, epochs + 1):
        rc = L.bi_hip_mlp_grad(x.ptr, t.ptr, mk.ptr, Xn.shape[0], p.ptr, K, g.ptr, st.handle)
        if rc:
            raise engine.EngineError(rc, "bi_hip_mlp_grad")
        lr_ep = lr if lr_end is None else lr_end + 0.5 * (lr - lr_end) * (1 + math.cos(math.pi * (ep - 1) / epochs))
        rc = L.bi_hip_mlp_adam(p.ptr, g.ptr, m.ptr, v.ptr, NP, lr_ep, 0.9, 0.999, 1e-8, ep, st.handle)
        if rc:
            raise engine.EngineError(rc, "bi_hip_mlp_adam")
        if ep % 100 == 0 or ep == 1:
            st.synchronize()
            hist.append((ep, float(g.download(np.float32, NP + 1)[NP])))
            if log and ep % 500 == 0:
                log(f"epoch {ep} loss {hist[-1][1]:.5f}")
    st.synchronize()
    return p.download(np.float32, NP), np.concatenate([mean, std]).astype(np.float32), hist


def install(table: np.ndarray | None, params: np.ndarray | None, norm6: np.ndarray | None) -> None:
    L = engine.lib()
    L.bi_hip_set_planner.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
    tb = np.ascontiguousarray(table, dtype=np.int8) if table is not None else None
    pp = np.ascontiguousarray(params, dtype=np.float32) if params is not None else None
    nn = np.ascontiguousarray(norm6, dtype=np.float32) if norm6 is not None else None
    rc = L.bi_hip_set_planner(tb.ctypes.data if tb is not None else None,
                              pp.ctypes.data if pp is not None else None, K,
                              nn.ctypes.data if nn is not None else None)
    if rc:
        raise engine.EngineError(rc, "bi_hip_set_planner")


def choose(shapes, mode: int) -> list[int]:
    L = engine.lib()
    L.bi_hip_choose_plans.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int,
                                      ctypes.c_void_p]
    sh = np.array([[la, lb, k] for la, lb, k in shapes], dtype=np.uint32).reshape(-1)
    ids = np.zeros(len(shapes), dtype=np.int32)
    rc = L.bi_hip_choose_plans(sh.ctypes.data, None, len(shapes), mode, ids.ctypes.data)
    if rc:
        raise engine.EngineError(rc, "bi_hip_choose_plans")
    return [int(i) for i in ids]


def save(path: str, table: np.ndarray, params: np.ndarray, norm6: np.ndarray, meta: dict) -> None:
    """JSON (with the training metadata) plus the binary form bi_hip_load_planner reads:
    b"BIPL", int32 K, int8 table[19*19*13], int32 parameter count, float32 params, float32 norm[6]."""
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "w") as f:
        json.dump({"K": K, "table": table.reshape(-1).tolist(), "params": params.tolist(),
                   "norm": norm6.tolist(), "meta": meta}, f)
    write_binary(os.path.splitext(path)[0] + ".bin", table, params, norm6)


def write_binary(path: str, table: np.ndarray, params: np.ndarray, norm6: np.ndarray) -> None:
    with open(path, "wb") as f:
        f.write(b"BIPL")
        f.write(np.int32(K).tobytes())
        f.write(np.ascontiguousarray(table, np.int8).reshape(-1).tobytes())
        f.write(np.int32(params.size).tobytes())
        f.write(np.ascontiguousarray(params, np.float32).tobytes())
        f.write(np.ascontiguousarray(norm6, np.float32).tobytes())


def load(path: str = DEFAULT_PATH, do_install: bool = True):
    with open(path) as f:
        d = json.load(f)
    table = np.array(d["table"], np.int8).reshape(TABLE_SHAPE)
    params = np.array(d["params"], np.float32)
    norm = np.array(d["norm"], np.float32)
    if do_install:
        install(table, params, norm)
    return table, params, norm, d.get("meta", {})
