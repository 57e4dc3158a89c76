"""Vectorised bf16 / OCP e4m3 encoders for building the oracle's low-precision
copies in tests (checked element-for-element against oracle.c's encoders in
tests/test_oracle.py).  Per-vector fp8 scale = max|x| / 448 in fp32, code =
round-to-nearest-even(x / scale), as the library quantises."""
import numpy as np


def to_bf16(x):
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return u.astype(np.uint16)


def _e4m3_grid():
    c = np.arange(0x7F)
    e, m = (c >> 3) & 15, c & 7
    return np.where(e > 0, np.ldexp(1.0 + m / 8.0, e - 7), m * 2.0 ** -9).astype(np.float32)


GRID = _e4m3_grid()  # codes 0..0x7E, increasing


def to_e4m3(x):
    x = np.asarray(x, dtype=np.float32)
    a = np.abs(x)
    hi = np.clip(np.searchsorted(GRID, a, side="left"), 1, 0x7E)
    lo = hi - 1
    dl, dh = a - GRID[lo], GRID[hi] - a
    c = np.where(dl < dh, lo, np.where(dh < dl, hi, np.where(lo % 2 == 1, hi, lo)))
    c = np.where(a >= 448.0, 0x7E, c)
    c = np.where(a <= GRID[0], 0, c)
    sign = np.signbit(x).astype(np.uint8) << 7
    return (c.astype(np.uint8) | sign).astype(np.uint8)


def quantize_fp8_rows(x):
    x = np.asarray(x, dtype=np.float32)
    amax = np.abs(x).max(axis=1)
    scale = np.where(amax > 0, amax / np.float32(448.0), np.float32(1.0)).astype(np.float32)
    codes = to_e4m3((x / scale[:, None]).astype(np.float32))
    return codes, scale
