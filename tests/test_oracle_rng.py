"""Oracle RNG (rng.rs), seeding and tile order (render.rs:284-299) against the committed
golden vectors, an independent pure-Python restatement, and closed forms."""
import json
import os

import numpy as np

import oracle

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "rng.json")))
M64 = (1 << 64) - 1
A, CINC = 6364136223846793005, 1442695040888963407


def py_pcg(seed, n):
    """PCG32 XSH-RR (O'Neill 2014) with the reference's fixed increment and seed-0 rule."""
    s = 0x9E3779B97F4A7C15 if seed == 0 else seed
    out = []
    for _ in range(n):
        old = s
        s = (old * A + CINC) & M64
        xs = (((old >> 18) ^ old) >> 27) & 0xFFFFFFFF
        rot = old >> 59
        out.append(((xs >> rot) | (xs << ((32 - rot) & 31))) & 0xFFFFFFFF)
    return out, s


def lcg_jump(seed, n):
    """State after n steps by squaring the affine map x -> A x + C (Brown 1994)."""
    acc_a, acc_c, a, c = 1, 0, A, CINC
    while n:
        if n & 1:
            acc_a, acc_c = (acc_a * a) & M64, (acc_c * a + c) & M64
        a, c = (a * a) & M64, (c * (a + 1)) & M64
        n >>= 1
    return (acc_a * seed + acc_c) & M64


def test_u32_sequences_match_golden_and_python():
    for s, seq in GOLD["u32"].items():
        s = int(s)
        assert [int(x) for x in oracle.rng_u32(s, len(seq))] == seq
        assert py_pcg(s, len(seq))[0] == seq


def test_lcg_jump_ahead_closed_form():
    for s in (1, 12345, 0xFFFFFFFFFFFFFFFF):
        for n in (1, 2, 31, 1000):
            assert py_pcg(s, n)[1] == lcg_jump(s, n)


def test_next_f64_is_53_bits_of_two_outputs():
    for s, seq in GOLD["f64"].items():
        s = int(s)
        got = oracle.rng_f64(s, len(seq))
        assert [float(x).hex() for x in got] == seq
        u, _ = py_pcg(s, 2 * len(seq))
        ref = [((u[2 * i] << 21) | (u[2 * i + 1] & ((1 << 21) - 1))) * 2.0 ** -53 for i in range(len(seq))]
        assert list(got) == ref
        assert all(0.0 <= x < 1.0 for x in got)


def test_next_f64_uniformity():
    x = oracle.rng_f64(2024, 200000)
    hist = np.histogram(x, bins=20, range=(0, 1))[0]
    assert abs(x.mean() - 0.5) < 0.003
    assert hist.min() > 9500 and hist.max() < 10500


def py_seed_for(px, py, s):
    h = 0xD1B54A32D192ED03
    h ^= (px * 0x9E3779B185EBCA87) & M64
    h ^= (py * 0xC2B2AE3D27D4EB4F) & M64
    h ^= (s * 0x165667B19E3779F9) & M64
    return h ^ (h >> 32)


def test_seed_for_golden_and_python():
    for px, py, s, v in GOLD["seed_for"]:
        assert oracle.seed_for(px, py, s) == v
        assert py_seed_for(px, py, s) == v


def py_interleave(v):
    out = 0
    for b in range(32):
        out |= ((v >> b) & 1) << (2 * b)
    return out


def test_morton2_golden_and_bitwise():
    for tx, ty, nx, ny, v in GOLD["morton2"]:
        assert oracle.morton2(tx, ty, nx, ny) == v
        x, y = min(tx, max(nx - 1, 0)), min(ty, max(ny - 1, 0))
        assert py_interleave(x) | (py_interleave(y) << 1) == v
