"""Planner MLP: the float64 CPU reference is pinned by finite differences; the GPU kernels'
forward pass and gradient match it (fp32 tolerances stated per check); the learned choice
never selects a plan that is not allowed; every plan computes the exact product."""
import ctypes

import numpy as np
import pytest

from oracle import mlp_oracle as MO

K = 8


def rand_problem(B=96, seed=0):
    rng = np.random.default_rng(seed)
    NP = 3 * 64 + 64 + 64 * 64 + 64 + 64 * K + K
    p = rng.normal(0, 0.3, NP)
    x = rng.normal(0, 1, (B, 3))
    t = rng.normal(0, 1, (B, K))
    mask = (rng.random((B, K)) < 0.6).astype(np.float64)
    mask[0, 0] = 1
    return p, x, t, mask


def test_reference_gradient_matches_finite_differences():
    p, x, t, mask = rand_problem()
    _, g = MO.loss_grad(p, x, t, mask, K)
    rng = np.random.default_rng(1)
    idx = rng.choice(p.size, 60, replace=False)
    h = 1e-6
    for i in idx:
        e = np.zeros_like(p)
        e[i] = h
        lp, _ = MO.loss_grad(p + e, x, t, mask, K)
        lm, _ = MO.loss_grad(p - e, x, t, mask, K)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])


def test_table_keeps_the_heuristic_where_it_is_better():
    """A table cell holding shapes on both sides of a power-of-two boundary: plan 0 is best for
    one, plan 3 for the other, plan 1 is the best single plan for both.  Without the heuristic
    the cell takes plan 1; offered the heuristic's per-shape choices (each shape's best), the cell
    stores -1, which the engine answers with the heuristic's exact-size choice."""
    from bigint_amd import planner as P
    recs = []
    for la, times in [(1000, {0: 10.0, 1: 12.0, 3: 20.0}), (1100, {0: 30.0, 1: 12.0, 3: 10.0})]:
        recs += [{"la": la, "lb": 1000, "k": 1, "plan": p, "us": us} for p, us in times.items()]
    cell = (round(np.log2(28 * 1000)) - 10, round(np.log2(28 * 1000)) - 10, 0)
    assert P.build_table(recs)[cell] == 1
    t = P.build_table(recs, {(1000, 1000, 1): 0, (1100, 1000, 1): 3})
    assert t[cell] == -1
    assert (t == -1).all()  # the one filled cell fills every other
    # a heuristic that picks the wrong plan everywhere loses to plan 1
    assert P.build_table(recs, {(1000, 1000, 1): 3, (1100, 1000, 1): 0})[cell] == 1


@pytest.mark.gpu
def test_gpu_forward_and_gradient_match_reference():
    from bigint_amd import engine, hip
    engine.init(0)
    L = engine.lib()
    p, x, t, mask = rand_problem(B=1000, seed=3)
    NP = L.bi_hip_mlp_param_count(K)
    assert NP == p.size
    tp = hip.Buffer.from_array(p.astype(np.float32))
    tx = hip.Buffer.from_array(x.astype(np.float32))
    tt = hip.Buffer.from_array(t.astype(np.float32))
    tm = hip.Buffer.from_array(mask.astype(np.float32))
    ty = hip.Buffer(4 * x.shape[0] * K)
    tg = hip.Buffer(4 * (NP + 1))
    st = hip.Stream()
    assert L.bi_hip_mlp_forward(tx.ptr, x.shape[0], tp.ptr, K, ty.ptr, st.handle) == 0
    st.synchronize()
    # reference run in float64 on the fp32-rounded parameters and inputs
    p32 = p.astype(np.float32).astype(np.float64)
    x32 = x.astype(np.float32).astype(np.float64)
    y_ref, _ = MO.forward(p32, x32, K)
    y = ty.download(np.float32, x.shape[0] * K).reshape(-1, K)
    # fp32 tolerance: |dy| <= 1e-4 * (1 + |y|)
    assert np.all(np.abs(y - y_ref) <= 1e-4 * (1 + np.abs(y_ref)))
    assert L.bi_hip_mlp_grad(tx.ptr, tt.ptr, tm.ptr, x.shape[0], tp.ptr, K, tg.ptr, st.handle) == 0
    st.synchronize()
    loss_ref, g_ref = MO.loss_grad(p32, x32, t.astype(np.float32), mask, K)
    g = tg.download(np.float32, NP + 1)
    # fp32 tolerance on the gradient: 1e-4 relative to the gradient's scale
    scale = np.abs(g_ref).max()
    assert np.max(np.abs(g[:NP] - g_ref)) <= 1e-4 * scale
    assert abs(g[NP] - loss_ref) <= 1e-4 * abs(loss_ref)


@pytest.mark.gpu
def test_training_reduces_loss_and_choices_are_allowed():
    """A few hundred Adam steps on a synthetic cost surface lower the loss; the learned choice
    only ever returns allowed plans."""
    from bigint_amd import engine, planner
    engine.init(0)
    rng = np.random.default_rng(5)
    X = rng.uniform([10, 10, 0], [28, 28, 6], (400, 3)).astype(np.float32)
    T = np.stack([X[:, 0] + X[:, 1] * 0.5 + k * 0.1 * (X[:, 2] - 3) for k in range(K)], 1).astype(np.float32)
    M = np.ones_like(T)
    params, norm, hist = planner.train_gpu(X, T, M, epochs=400, lr=3e-3)
    assert hist[-1][1] < 0.2 * hist[0][1]
    table = np.zeros(planner.TABLE_SHAPE, np.int8)
    planner.install(table, params, norm)
    shapes = [(50000, 900, 1), (3000, 3000, 4), (9586981, 37450, 1), (10, 10, 2)]
    for (la, lb, k), c in zip(shapes, planner.choose(shapes, engine.PLAN_LEARNED)):
        assert c in planner.allowed(la, lb)


@pytest.mark.gpu
def test_every_plan_is_exact():
    from bigint_amd import engine, hip
    from bigint_amd import limbs as L
    from oracle import oracle
    from bigint_amd import planner
    engine.init(0)
    engine.lib().bi_hip_mul_batch_device_planned.argtypes = [
        ctypes.POINTER(engine.DPair), ctypes.c_size_t, ctypes.POINTER(ctypes.c_int32), ctypes.c_void_p]
    for la, lb in [(50000, 900), (3000, 3000), (70000, 20000), (20000, 70000), (9000, 130),
                   (200000, 150000)]:
        a = L.random_operand(28 * la, la)
        b = L.random_operand(28 * lb, lb)
        exp = oracle.mul_limbs(a, b)
        da, db = hip.Buffer.from_array(a), hip.Buffer.from_array(b)
        for plan in planner.allowed(a.size, b.size):
            out = hip.Buffer(4 * (a.size + b.size))
            out.zero()
            pair = (engine.DPair * 1)(engine.DPair(da.ptr, a.size, db.ptr, b.size, out.ptr, a.size + b.size))
            ids = (ctypes.c_int32 * 1)(plan)
            assert engine.lib().bi_hip_mul_batch_device_planned(pair, 1, ids, None) == 0
            engine.synchronize()
            got = L.strip(out.download(np.uint32, a.size + b.size))
            assert np.array_equal(got, exp), (la, lb, plan)


def test_plan_allowed_rules():
    import os
    from bigint_amd import engine
    if not os.path.exists(engine.LIB_PATH):
        pytest.skip("library not built")
    from bigint_amd import planner
    assert planner.allowed(10, 10) == [0]
    assert 0 in planner.allowed(9586981, 9586981) and 3 not in planner.allowed(9586981, 9586981)
    al = planner.allowed(9586981, 37450)
    assert {0, 1, 2, 3, 4, 5} <= set(al) | {1, 2}
    assert 3 in al and 4 in al
    assert not engine.lib().bi_hip_plan_allowed(9586981, 9586981, 1, 3)  # squares never chunk
    # balanced 2^28-bit operands fit one 2^24-point transform: no chunk plan is shorter
    assert planner.allowed(9586981, 9586981) == [0, 1, 2]
    assert set(planner.allowed(70000, 20000)) >= {0, 6, 7}
    assert set(planner.allowed(200000, 150000)) >= {0, 6}
    # beyond one 2^26-point transform: no unchunked plan; a short operand still allows chunks,
    # two long ones run the engine's both-sides split (no plan id)
    assert min(planner.allowed(38347925, 38347925)) >= 6  # A cut in 2 or 3, B whole
    al = planner.allowed(80_000_000, 1_000_000)
    assert al and min(al) >= 3
    assert planner.allowed(58_000_000, 58_000_000) == []  # only the both-sides split fits
    assert engine.lib().bi_hip_plan_allowed(58_000_000, 58_000_000, 0, 8)
    assert engine.lib().bi_hip_plan_allowed(1000, 1000, 1, 8)  # explicit id: any product
    assert not engine.lib().bi_hip_plan_allowed(76695851, 76695851, 0, 8)  # beyond 3 primes


def _planned(pairs_ids):
    """Run [(a, b, plan)] as ONE planned device batch; returns the products."""
    from bigint_amd import engine, hip
    from bigint_amd import limbs as L
    engine.lib().bi_hip_mul_batch_device_planned.argtypes = [
        ctypes.POINTER(engine.DPair), ctypes.c_size_t, ctypes.POINTER(ctypes.c_int32), ctypes.c_void_p]
    keep, dp = [], []
    for a, b, _ in pairs_ids:
        da, db = hip.Buffer.from_array(a), hip.Buffer.from_array(b)
        out = hip.Buffer(4 * (a.size + b.size))
        out.zero()
        keep.append((da, db, out))
        dp.append(engine.DPair(da.ptr, a.size, db.ptr, b.size, out.ptr, a.size + b.size))
    arr = (engine.DPair * len(dp))(*dp)
    ids = (ctypes.c_int32 * len(dp))(*[p for _, _, p in pairs_ids])
    assert engine.lib().bi_hip_mul_batch_device_planned(arr, len(dp), ids, None) == 0
    engine.synchronize()
    return [L.strip(o.download(np.uint32, a.size + b.size)) for (a, b, _), (_, _, o) in zip(pairs_ids, keep)]


@pytest.mark.gpu
def test_mixed_plan_batch_is_exact():
    """One batch holding plain, re-split and chunked products (shared B transforms of several
    lengths, two- and three-way chunk overlaps, swapped operands) against the oracle."""
    from bigint_amd import engine, planner
    from bigint_amd import limbs as L
    from oracle import oracle
    engine.init(0)
    shapes = [(50000, 900), (900, 50000), (70000, 20000), (20000, 70000), (90000, 60000),
              (3000, 3000), (9000, 130), (200000, 150000), (33000, 4096), (114285, 68571)]
    jobs = []
    for i, (la, lb) in enumerate(shapes):
        for p in planner.allowed(la, lb):
            jobs.append((L.random_operand(28 * la, 100 + 17 * i + p), L.random_operand(28 * lb, 500 + 13 * i + p), p))
    assert {p for _, _, p in jobs} == set(range(planner.K))
    for (a, b, p), got in zip(jobs, _planned(jobs)):
        assert np.array_equal(got, oracle.mul_limbs(a, b)), (a.size, b.size, p)


@pytest.mark.gpu
def test_three_chunk_overlap_all_ones():
    """All-ones operands maximise every coefficient; 100000 x 60000 words with 50176-word chunks
    on 2^17-point transforms: three chunk products overlap in words [100352, 110175)."""
    from bigint_amd import engine, planner
    from bigint_amd import limbs as L
    engine.init(0)
    la, lb = 114285, 68571
    assert 6 in planner.allowed(la, lb)
    x, y = (1 << (28 * la)) - 1, (1 << (28 * lb)) - 1
    (got,) = _planned([(L.int_to_limbs(x), L.int_to_limbs(y), 6)])
    assert L.limbs_to_int(got) == x * y


@pytest.mark.gpu
def test_balanced_2e28_every_plan_agrees():
    """2^28-bit x 2^28-bit (8,388,609 words each, one coefficient wrapping on a 2^24-point
    transform): every allowed plan (4096/2048/1024-point block passes) gives the same product,
    and that product satisfies the residue checks."""
    from bigint_amd import engine, planner
    from bigint_amd import limbs as L
    engine.init(0)
    a = L.random_operand(1 << 28, L.SEED)
    b = L.random_operand(1 << 28, L.SEED + 1)
    plans = planner.allowed(a.size, b.size)
    assert plans == [0, 1, 2]
    outs = _planned([(a, b, p) for p in plans])
    for q in L.RESIDUE_PRIMES:
        assert L.residue_mod(outs[0], q) == L.residue_mod(a, q) * L.residue_mod(b, q) % q
    for p, o in zip(plans[1:], outs[1:]):
        assert np.array_equal(o, outs[0]), p


@pytest.mark.gpu
def test_trained_planner_gains_over_plain():
    """The committed planner (table + MLP trained on this engine's timings) on fresh shapes it was
    not trained on: time every allowed plan, then compare the total time of the learned choices
    with plan 0 (the plain path), with the autotuned table's choices (the fallback) and with the
    best plan, and report the mean regret (time of the choice over the best time, minus one) of
    the learned and table choices."""
    import os
    from bigint_amd import engine, planner
    engine.init(0)
    if not os.path.exists(planner.DEFAULT_PATH):
        pytest.skip("no tuned planner committed")
    planner.load()
    timer = planner.Timer(max_limbs=1 << 23)
    shapes = [s for s in planner.sample_shapes(120, seed=424242, hi=27.0, max_batch_log=8)
              if s[2] * (s[0] + s[1]) <= timer.out_words - 64 and max(s[0], s[1]) <= timer.max_limbs]
    learned = planner.choose(shapes, engine.PLAN_LEARNED)
    table = planner.choose(shapes, engine.PLAN_TABLE)
    t_plain = t_learned = t_table = t_best = 0.0
    reg_l, reg_t = [], []
    for (la, lb, k), c, tc in zip(shapes, learned, table):
        times = {p: timer.run(la, lb, k, p, reps=2) * k for p in planner.allowed(la, lb)}
        assert c in times and tc in times
        best = min(times.values())
        t_plain += times[0]
        t_learned += times[c]
        t_table += times[tc]
        t_best += best
        reg_l.append(times[c] / best - 1.0)
        reg_t.append(times[tc] / best - 1.0)
    ml, mt = float(np.mean(reg_l)), float(np.mean(reg_t))
    print(f"{len(shapes)} shapes: plain {t_plain:.0f} us, table {t_table:.0f} us, learned {t_learned:.0f} us, "
          f"best {t_best:.0f} us; mean regret learned {100 * ml:.1f} %, table {100 * mt:.1f} %; "
          f"learned vs plain {t_plain / t_learned:.3f}x, vs table {t_table / t_learned:.3f}x")
    # the learned choices keep at least half of what the best plans gain over plan 0, stay within
    # 15 % of the best in total, and are no worse than the table they fall back on (3 % timing
    # slack in total, one point of mean regret)
    assert t_plain - t_learned >= 0.5 * (t_plain - t_best)
    assert t_learned / t_best < 1.15
    assert t_learned <= 1.03 * t_table
    assert ml <= mt + 0.01


@pytest.mark.gpu
def test_split_both_plan_is_exact():
    """Plan 8 (both operands in 2^25-word chunks, diagonal sums of chunk spectra) on products
    that also have ordinary plans: 2^25 + 5 words times 2^25 + 3 words (two chunks each, the
    second ones tiny), all-ones operands, against plan 0's product and the residue checks."""
    from bigint_amd import engine
    from bigint_amd import limbs as L
    engine.init(0)
    def limbs_for_words(w):
        la = (8 * w) // 7
        while (la * 7 + 7) // 8 > w:
            la -= 1
        while ((la + 1) * 7 + 7) // 8 == w:
            la += 1
        return la
    la, lb = limbs_for_words((1 << 25) + 5), limbs_for_words((1 << 25) + 3)
    a = L.random_operand(28 * la, 91)
    b = L.random_operand(28 * lb - 11, 92)
    p8, p6 = _planned([(a, b, 8), (a, b, 6)])
    assert np.array_equal(p8, p6)
    for q in L.RESIDUE_PRIMES[:3]:
        assert L.residue_mod(p8, q) == L.residue_mod(a, q) * L.residue_mod(b, q) % q
    x, y = (1 << (28 * 300000)) - 1, (1 << (28 * 200000)) - 1
    (p,) = _planned([(L.int_to_limbs(x), L.int_to_limbs(y), 8)])
    assert L.limbs_to_int(p) == x * y
