"""numpy model of the BF16x3 product form (test infrastructure): bf16 rounding
to nearest even, the hi/lo split of an FP32 value, and the three-product sum
A_hi B_hi + A_hi B_lo + A_lo B_hi evaluated exactly (FP64)."""
import numpy as np


def bf16_rne(x):
    """FP32 array -> the nearest bf16 values (ties to even), as FP32."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


def bf16_split(x):
    x = np.asarray(x, dtype=np.float32)
    hi = bf16_rne(x)
    lo = bf16_rne(x - hi)  # x - hi is exact in FP32
    return hi, lo


def bf16x3_matmul(a, b):
    """Exact value of the three bf16 products of FP32 matrices a @ b."""
    ah, al = (v.astype(np.float64) for v in bf16_split(a))
    bh, bl = (v.astype(np.float64) for v in bf16_split(b))
    return ah @ bh + ah @ bl + al @ bh
