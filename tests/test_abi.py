"""The C-ABI library loads and exports every symbol include/bigint_hip.h declares (CPU only:
no compute calls are made without a GPU)."""
import os
import re

import pytest

from bigint_amd import engine

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    with open(os.path.join(ROOT, "include", "bigint_hip.h")) as f:
        text = f.read()
    return sorted(set(re.findall(r"\b(bi_hip_[a-z0-9_]+)\s*\(", text)))


def test_header_matches_binding_list():
    assert declared_symbols() == sorted(engine.EXPORTS)


def test_library_exports_every_symbol():
    if not os.path.exists(engine.LIB_PATH):
        pytest.skip("libbigint_hip.so not built")
    L = engine.lib()
    for s in declared_symbols():
        assert hasattr(L, s), s


def test_host_only_entry_points():
    if not os.path.exists(engine.LIB_PATH):
        pytest.skip("libbigint_hip.so not built")
    assert engine.strerror(0) == "ok"
    # transforms run over 32-bit words (7 per 8 limbs); n >= max(wa + wb - 2, wa, wb), the top
    # coefficient wrapping onto coefficient 0 when wa + wb - 1 = n + 1
    assert engine.transform_log2(1, 1) == 4
    assert engine.transform_log2(4096, 1) == 12     # 3584 words
    assert engine.transform_log2(4096, 2) == 12     # 3584 + 2 - 2
    assert engine.transform_log2(4682, 4682) == 13  # 4097 + 4097 - 2 = 8192 words
    assert engine.transform_log2(2341, 2341) == 12  # 2049 + 2049 - 2 = 4096: one wrapped coefficient
    assert engine.transform_log2(1 << 25, 1 << 25) == 26
    assert engine.transform_log2(38347923, 38347923) == 26  # 2^25 + 1 words each: the largest
    assert engine.transform_log2(38347923, 38347924) == -1
    # the 2^28 x 2^28 workload: 9,586,981 limbs = 8,388,609 words each -> n = 2^24
    assert engine.transform_log2(9586981, 9586981) == 24


def test_planner_file_validation(tmp_path):
    """bi_hip_load_planner accepts table cells of -1 (the cell defers to the built-in heuristic)
    and rejects ids outside [-1, K).  Validation precedes device set-up, so without a GPU a valid
    file gets past it (returning the device error) while an invalid one is BI_HIP_EINVAL."""
    if not os.path.exists(engine.LIB_PATH):
        pytest.skip("libbigint_hip.so not built")
    import numpy as np
    from bigint_amd import planner as P
    L = engine.lib()
    L.bi_hip_load_planner.argtypes = [__import__("ctypes").c_char_p]
    params = np.zeros(L.bi_hip_mlp_param_count(P.K), np.float32)
    norm = np.array([0, 0, 0, 1, 1, 1], np.float32)
    table = np.zeros(P.TABLE_SHAPE, np.int8)
    table[3, 4, 5] = -1
    good = tmp_path / "good.bin"
    P.write_binary(str(good), table, params, norm)
    rc_good = L.bi_hip_load_planner(str(good).encode())
    table[3, 4, 5] = -2
    bad = tmp_path / "bad.bin"
    P.write_binary(str(bad), table, params, norm)
    assert L.bi_hip_load_planner(str(bad).encode()) == engine.EINVAL
    table[3, 4, 5] = P.K
    P.write_binary(str(bad), table, params, norm)
    assert L.bi_hip_load_planner(str(bad).encode()) == engine.EINVAL
    assert rc_good != engine.EINVAL
