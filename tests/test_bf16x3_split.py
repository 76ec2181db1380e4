"""Split-BF16x3 arithmetic of the matrix-core kernels, emulated on the CPU (no GPU).

The BF16x3 kernels split each FP32 operand as x = hi + lo with hi = bf16(x) and
lo = bf16(x - hi) (round to nearest even) and accumulate a b as lo hi' + hi lo' + hi hi'.  The
Legendre kernel feeds four consecutive k-steps of its FP32 16x16x4 form to one 16x16x16 bf16
instruction: lane (row i, group g) holds A[i][4 (4 q + e) + g] as element e, the B lanes the
matching B[4 (4 q + e) + g][j]; the instruction pairs element e of A with element e of B in the
same lane group, so the sum covers the 16 k's of the four steps exactly once.
"""

from __future__ import annotations

import numpy as np


def bf16(x):
    """Round float32 values to bfloat16 (nearest, ties to even), returned as float32."""
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    r = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return r.astype(np.uint32).view(np.float32)


def split(x):
    x = np.asarray(x, np.float32)
    hi = bf16(x)
    lo = bf16(x - hi)  # x - hi is exact in FP32
    return hi, lo


def test_bf16_rounding_known_values():
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: even mantissa (1) wins; 1 + 3 * 2^-8 rounds up.
    assert bf16(np.float32(1 + 2 ** -8)) == np.float32(1.0)
    assert bf16(np.float32(1 + 3 * 2 ** -8)) == np.float32(1 + 2 ** -6)
    assert bf16(np.float32(-1 - 2 ** -9)) == np.float32(-1.0)
    assert bf16(np.float32(3.0)) == np.float32(3.0)


def test_split_error_bound():
    rng = np.random.default_rng(0)
    x = (rng.standard_normal(100000) * 10.0 ** rng.uniform(-20, 20, 100000)).astype(np.float32)
    hi, lo = split(x)
    err = np.abs(x.astype(np.float64) - hi.astype(np.float64) - lo.astype(np.float64))
    assert np.all(err <= 2.0 ** -16 * np.abs(x.astype(np.float64)))


def test_three_product_error_bound():
    """lo hi' + hi lo' + hi hi' against the exact product: |error| <= 2^-14 |a b| (the dropped
    lo lo' term and the two splits)."""
    rng = np.random.default_rng(1)
    a = rng.standard_normal(50000).astype(np.float32)
    b = rng.standard_normal(50000).astype(np.float32)
    ah, al = split(a)
    bh, bl = split(b)
    f = lambda u, v: u.astype(np.float64) * v.astype(np.float64)  # noqa: E731  (bf16 products are exact)
    approx = f(al, bh) + f(ah, bl) + f(ah, bh)
    exact = f(a, b)
    assert np.all(np.abs(approx - exact) <= 2.0 ** -14 * np.abs(exact) + 1e-300)


def _mfma_16x16x16(a_lanes, b_lanes):
    """D[i][j] = sum over lane group g and element e of A(lane i + 16 g)[e] * B(lane j + 16 g)[e]."""
    d = np.zeros((16, 16))
    for g in range(4):
        for e in range(4):
            d += np.outer(a_lanes[16 * g: 16 * g + 16, e], b_lanes[16 * g: 16 * g + 16, e])
    return d


def test_four_k_steps_in_one_bf16_instruction():
    """The kernel's lane feed: element e of lane (i, g) holds k = 4 (4 q + e) + g of group q;
    one emulated 16x16x16 instruction per group equals the FP32 16x16x4 k-steps' sum A @ B."""
    rng = np.random.default_rng(2)
    K = 32  # two groups of four k-steps
    A = rng.standard_normal((16, K))
    B = rng.standard_normal((K, 16))
    d = np.zeros((16, 16))
    for q in range(K // 16):
        a_l = np.zeros((64, 4))
        b_l = np.zeros((64, 4))
        for lane in range(64):
            i, g = lane % 16, lane // 16
            for e in range(4):
                k = 4 * (4 * q + e) + g
                a_l[lane, e] = A[i, k]
                b_l[lane, e] = B[k, i]  # lane's column j = lane % 16
        d += _mfma_16x16x16(a_l, b_l)
    np.testing.assert_allclose(d, A @ B, rtol=1e-13, atol=1e-13)
