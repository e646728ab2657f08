"""Writes tests/golden/rng.json: PCG32 (rng.rs:9-38) output sequences, next_f64 sequences,
seed_for (rng.rs:41-47) values and Morton keys (render.rs:284-299) produced by the C oracle.

The reference itself (Rust) cannot be built here (no cargo/rustc, crates not vendored), so
these vectors come from the oracle; tests/test_oracle_rng.py checks them against an
independent pure-Python restatement and a jump-ahead closed form of the LCG.

    python tests/golden/make_golden.py
"""
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402

SEEDS = [0, 1, 7, 42, 0xDEADBEEF, 0x9E3779B97F4A7C15, 0xFFFFFFFFFFFFFFFF]


def main():
    out = {"u32": {}, "f64": {}, "seed_for": [], "morton2": []}
    for s in SEEDS:
        out["u32"][str(s)] = [int(x) for x in oracle.rng_u32(s, 32)]
        out["f64"][str(s)] = [float(x).hex() for x in oracle.rng_f64(s, 16)]
    for px, py, s in [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (511, 511, 255), (1919, 1079, 4095),
                      (123, 456, 789), (4294967295, 4294967295, 4294967295)]:
        out["seed_for"].append([px, py, s, oracle.seed_for(px, py, s)])
    for tx, ty, nx, ny in [(0, 0, 1, 1), (1, 0, 4, 4), (0, 1, 4, 4), (3, 3, 4, 4), (5, 2, 4, 4), (119, 67, 120, 68),
                           (1000, 1000, 2, 2)]:
        out["morton2"].append([tx, ty, nx, ny, oracle.morton2(tx, ty, nx, ny)])
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "rng.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
