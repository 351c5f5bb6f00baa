"""Planner MLP: the float64 CPU reference is pinned by finite differences; the GPU kernels'
forward pass and gradient match it (fp32 tolerances stated per check); the learned choice
never selects a plan that is not allowed; every plan computes the exact product."""
import ctypes

import numpy as np
import pytest

from oracle import mlp_oracle as MO

K = 6


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


@pytest.mark.gpu
def test_gpu_forward_and_gradient_match_reference():
    import torch
    from bigint_amd import engine
    engine.init(0)
    L = engine.lib()
    dev = torch.device("cuda:0")
    p, x, t, mask = rand_problem(B=1000, seed=3)
    NP = L.bi_hip_mlp_param_count(K)
    assert NP == p.size
    tp = torch.tensor(p, dtype=torch.float32, device=dev)
    tx = torch.tensor(x, dtype=torch.float32, device=dev)
    tt = torch.tensor(t, dtype=torch.float32, device=dev)
    tm = torch.tensor(mask, dtype=torch.float32, device=dev)
    ty = torch.zeros(x.shape[0], K, device=dev)
    tg = torch.zeros(NP + 1, device=dev)
    st = torch.cuda.Stream(dev)
    f = ctypes.c_void_p
    assert L.bi_hip_mlp_forward(f(tx.data_ptr()), x.shape[0], f(tp.data_ptr()), K, f(ty.data_ptr()),
                                f(st.cuda_stream)) == 0
    st.synchronize()
    # reference run in float64 on the fp32-rounded parameters and inputs
    p32 = p.astype(np.float32).astype(np.float64)
    x32 = x.astype(np.float32).astype(np.float64)
    y_ref, _ = MO.forward(p32, x32, K)
    # fp32 tolerance: |dy| <= 1e-4 * (1 + |y|)
    assert np.all(np.abs(ty.cpu().numpy() - y_ref) <= 1e-4 * (1 + np.abs(y_ref)))
    assert L.bi_hip_mlp_grad(f(tx.data_ptr()), f(tt.data_ptr()), f(tm.data_ptr()), x.shape[0],
                             f(tp.data_ptr()), K, f(tg.data_ptr()), f(st.cuda_stream)) == 0
    st.synchronize()
    loss_ref, g_ref = MO.loss_grad(p32, x32, t.astype(np.float32), mask, K)
    g = tg.cpu().numpy()
    # fp32 tolerance on the gradient: 1e-4 relative to the gradient's scale
    scale = np.abs(g_ref).max()
    assert np.max(np.abs(g[:NP] - g_ref)) <= 1e-4 * scale
    assert abs(g[NP] - loss_ref) <= 1e-4 * abs(loss_ref)


@pytest.mark.gpu
def test_every_plan_is_exact():
    from bigint_amd import engine
    from bigint_amd import limbs as L
    from oracle import oracle
    from bigint_amd import planner
    import torch
    engine.init(0)
    dev = torch.device("cuda:0")
    for la, lb in [(50000, 900), (3000, 3000), (70000, 20000), (9000, 130)]:
        a = L.random_operand(28 * la, la)
        b = L.random_operand(28 * lb, lb)
        exp = oracle.mul_limbs(a, b)
        da = torch.from_numpy(a.view(np.int32)).to(dev)
        db = torch.from_numpy(b.view(np.int32)).to(dev)
        for plan in planner.allowed(a.size, b.size):
            out = torch.zeros(a.size + b.size, dtype=torch.int32, device=dev)
            pair = (engine.DPair * 1)(engine.DPair(da.data_ptr(), a.size, db.data_ptr(), b.size,
                                                   out.data_ptr(), out.numel()))
            ids = (ctypes.c_int32 * 1)(plan)
            engine.lib().bi_hip_mul_batch_device_planned.argtypes = [
                ctypes.POINTER(engine.DPair), ctypes.c_size_t, ctypes.POINTER(ctypes.c_int32), ctypes.c_void_p]
            assert engine.lib().bi_hip_mul_batch_device_planned(pair, 1, ids, None) == 0
            engine.synchronize()
            got = L.strip(out.cpu().numpy().view(np.uint32))
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
