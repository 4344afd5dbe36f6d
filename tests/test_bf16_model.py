"""The numpy model of the BF16x3 products against values fixed by the bf16
format (8 significant bits, FP32 exponent range, round to nearest even)."""
import numpy as np

from bf16 import bf16_rne, bf16_split, bf16x3_matmul


def test_bf16_rounding_known_values():
    x = np.array([1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -8, 1 + 2 ** -7, -(1 + 2 ** -9), 3.0e38, 1e-30, 0.0],
                 dtype=np.float32)
    # 1 + 2^-8 is the midpoint of 1 and 1 + 2^-7: ties to the even significand (1);
    # 1 + 3 * 2^-8 is the midpoint of 1 + 2^-7 and 1 + 2^-6: ties to 1 + 2^-6
    want = np.array([1.0, 1.0, 1 + 2 ** -6, 1 + 2 ** -7, -1.0, 0, 0, 0.0], dtype=np.float32)
    got = bf16_rne(x)
    assert np.array_equal(got[:5], want[:5])
    assert np.all(got.view(np.uint32) & 0xFFFF == 0)
    assert abs(got[5] / x[5] - 1) <= 2 ** -8 and abs(got[6] / x[6] - 1) <= 2 ** -8
    assert got[7] == 0


def test_bf16_split_error_bounds():
    rng = np.random.default_rng(0)
    x = (rng.standard_normal(100000) * np.exp(rng.uniform(-20, 20, 100000))).astype(np.float32)
    hi, lo = bf16_split(x)
    assert np.all(np.abs(hi - x) <= 2.0 ** -8 * np.abs(x))
    # the pair carries 16 significant bits: |x - hi - lo| <= 2^-16 |x| (plus 2^-9 of |lo|)
    err = np.abs(x.astype(np.float64) - hi.astype(np.float64) - lo.astype(np.float64))
    assert np.all(err <= 2.0 ** -16 * np.abs(x.astype(np.float64)))


def test_bf16x3_products_are_near_fp32_accuracy():
    rng = np.random.default_rng(1)
    a = rng.standard_normal((40, 300)).astype(np.float32)
    b = rng.standard_normal((300, 30)).astype(np.float32)
    exact = a.astype(np.float64) @ b.astype(np.float64)
    scale = np.abs(a).astype(np.float64) @ np.abs(b).astype(np.float64)
    e3 = np.abs(bf16x3_matmul(a, b) - exact) / scale
    ah, bh = bf16_rne(a).astype(np.float64), bf16_rne(b).astype(np.float64)
    e1 = np.abs(ah @ bh - exact) / scale
    assert e3.max() < 3 * 2.0 ** -16
    assert e1.max() > 30 * e3.max()
