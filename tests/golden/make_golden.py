"""Writes tests/golden/products.json: products of seeded operands computed with Python's own
int multiplication (CPython, independent of the reference and of this project's code).

Each case records the operand recipe (bits and SplitMix64 seed, see bigint_amd/limbs.py), the
limb count and SHA-256 of the product's 28-bit limbs (uint32 little-endian bytes); cases up to
64 limbs also store the full product limbs.

Run:  python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import numpy as np  # noqa: E402

from bigint_amd import limbs as L  # noqa: E402

S = L.SEED
CASES = [
    # name, bits_a, seed_a, bits_b, seed_b (None = square of A)
    ("1x1", 5, S, 3, S + 1),
    ("28x28", 28, S, 28, S + 1),
    ("29x57", 29, S, 57, S + 1),
    ("100x200", 100, S, 200, S + 1),
    ("sq_300", 300, S, None, None),
    ("1000x1000", 1000, S, 1000, S + 1),
    ("1792x1792", 1792, S, 1792, S + 1),
    ("2^12x2^12", 1 << 12, S, 1 << 12, S + 1),
    ("2^16x2^16", 1 << 16, S, 1 << 16, S + 1),
    ("2^16x2^10", 1 << 16, S, 1 << 10, S + 1),
    ("sq_2^16", 1 << 16, S, None, None),
    ("2^17+5x2^17-3", (1 << 17) + 5, S + 7, (1 << 17) - 3, S + 8),
    ("2^20x2^20", 1 << 20, S, 1 << 20, S + 1),
    ("2^20x2^12", 1 << 20, S, 1 << 12, S + 1),
    ("2^22x2^22", 1 << 22, S, 1 << 22, S + 1),
    ("sq_2^22", 1 << 22, S, None, None),
    ("2^22x2^16", 1 << 22, S, 1 << 16, S + 1),
    ("2^24x2^24", 1 << 24, S, 1 << 24, S + 1),
    ("2^24x2^20", 1 << 24, S, 1 << 20, S + 1),
    ("sq_2^24", 1 << 24, S, None, None),
]


def sha(limbs: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(limbs, dtype="<u4").tobytes()).hexdigest()


def main() -> None:
    out = []
    for name, ba, sa, bb, sb in CASES:
        t = time.time()
        a = L.random_operand(ba, sa)
        ia = L.limbs_to_int(a)
        if bb is None:
            prod = ia * ia
        else:
            prod = ia * L.limbs_to_int(L.random_operand(bb, sb))
        p = L.int_to_limbs(prod)
        case = {"name": name, "bits_a": ba, "seed_a": sa, "bits_b": bb, "seed_b": sb,
                "n_limbs": int(p.size), "sha256": sha(p),
                "head": [int(x) for x in p[:4]], "tail": [int(x) for x in p[-4:]]}
        if p.size <= 64:
            case["limbs"] = [int(x) for x in p]
        out.append(case)
        print(f"{name}: {p.size} limbs, {time.time() - t:.2f}s", flush=True)
    with open(os.path.join(HERE, "products.json"), "w") as f:
        json.dump({"generator": "CPython int multiplication", "cases": out}, f, indent=1)


if __name__ == "__main__":
    main()
