"""The gather's index decode divides with reciprocals rounded up by one ulp and truncates, with no
correction step (DecodeConsts in csrc/hip/eri_tensor.hip): trunc(a * next_up(fl(1/d))) must equal
a // d for a < 2^12 with every divisor d <= 4096 in float (point coordinates), and for a < 2^24 with
divisor N and a < 2^48 with divisor N^2 in double (the four-index split of grids with N <= 4096).

float: every numerator and divisor.  double: every divisor, at the numerators where the quotient is
most fragile (remainder 0 and remainder d - 1, the largest quotients below the bound, and a seeded
random sample).  IEEE round-to-nearest products and truncating conversions, as on the GPU.

The enumeration's candidate decode (udiv_magic in csrc/hip/enumerate.hip) divides by an interval
length d <= 31 with one 24-bit multiply-high: v // d = ((v << 9) * ceil(2^23 / d)) >> 32 for v < 2^15,
checked for every numerator and divisor."""
import numpy as np


def _inv_up64(d):
    return np.nextafter(1.0 / d.astype(np.float64), np.inf)


def test_float_reciprocal_quotients_exhaustive():
    a = np.arange(1 << 12, dtype=np.uint32)
    af = a.astype(np.float32)
    for d in range(1, 4097):
        inv = np.nextafter(np.float32(1.0) / np.float32(d), np.float32(np.inf))
        q = (af * inv).astype(np.uint32)   # float32 product, truncating conversion
        assert np.array_equal(q, a // d), d


def _check_double(a, d):
    inv = _inv_up64(d)
    q = (a.astype(np.float64) * inv).astype(np.uint64)
    assert np.array_equal(q, a // d), (a[q != a // d][:4], d if np.isscalar(d) else d[q != a // d][:4])


def test_double_reciprocal_quotients_divisor_N():
    rng = np.random.default_rng(24)
    for d in range(1, 4097):
        dd = np.uint64(d)
        top = np.uint64((1 << 24) - 1)
        qmax = top // dd
        qs = np.concatenate([np.arange(max(0, int(qmax) - 64), int(qmax) + 1, dtype=np.uint64),
                             np.arange(0, min(64, int(qmax) + 1), dtype=np.uint64)])
        a = np.concatenate([qs * dd, qs * dd + (dd - np.uint64(1)), rng.integers(0, 1 << 24, 256, dtype=np.uint64)])
        a = a[a < np.uint64(1 << 24)]
        _check_double(a, np.full(a.shape, d, dtype=np.uint64))


def test_double_reciprocal_quotients_divisor_N2():
    rng = np.random.default_rng(48)
    # N = n_x n_y n_z <= 4096: every N, divisor N^2, numerators below 2^48
    for N in range(1, 4097):
        d = np.uint64(N * N)
        top = np.uint64((1 << 48) - 1)
        qmax = int(top // d)
        qs = np.concatenate([np.arange(max(0, qmax - 16), qmax + 1, dtype=np.uint64),
                             np.arange(0, min(16, qmax + 1), dtype=np.uint64)])
        a = np.concatenate([qs * d, qs * d + (d - np.uint64(1)), rng.integers(0, 1 << 48, 64, dtype=np.uint64)])
        a = a[a < np.uint64(1 << 48)]
        _check_double(a, np.full(a.shape, N * N, dtype=np.uint64))


def test_enumeration_magic_quotients_exhaustive():
    v = np.arange(1 << 15, dtype=np.uint64)
    for d in range(1, 32):
        m = ((1 << 23) + d - 1) // d
        assert m < (1 << 24), d   # a 24-bit multiplier operand
        q = ((v << np.uint64(9)) * np.uint64(m)) >> np.uint64(32)
        assert np.array_equal(q, v // np.uint64(d)), d
