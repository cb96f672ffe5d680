"""The numpy BFP32 model pinned by the properties that define the format:
the exponent rule, the error bound, exact round trips of representable
values, the bit layout, and hand-worked words."""
import numpy as np

import bfp_model as bm


def _fields(rng):
    yield rng.standard_normal((4, 3, 32))
    # segments of very different scales, zeros, signs, powers of two
    u = rng.standard_normal((2, 5, 48)) * np.ldexp(1.0, rng.integers(-60, 60, (2, 5, 48)))
    u[0, 1, :16] = 0.0
    u[1, 2, 16:32] = -np.ldexp(1.0, 7)
    yield u
    yield np.ldexp(rng.standard_normal((3, 64)), -140)  # below FP32's normal range: e clamps at -126
    yield np.sin(np.linspace(0, 6 * np.pi, 256)).reshape(2, 128)


def test_exponent_is_smallest_bound():
    rng = np.random.default_rng(1)
    for u in _fields(rng):
        e = bm.exponents(bm.encode(u))
        m = np.abs(u).reshape(-1, 16).max(axis=1)
        nz = m > 0
        # |v| < 2^e, and 2^(e-1) <= max unless clamped
        assert np.all(m < np.ldexp(1.0, e))
        unclamped = nz & (e > -126)
        assert np.all(np.ldexp(1.0, e[unclamped] - 1) <= m[unclamped])
        assert np.all(e[~nz] == -126)


def test_round_trip_error_bound():
    rng = np.random.default_rng(2)
    for u in _fields(rng):
        w = bm.encode(u)
        v = bm.decode(w)
        e = np.repeat(bm.exponents(w), 16).reshape(u.shape)
        err = np.abs(v - u)
        # half a unit 2^(e-31), one unit where rounding reached 2^30
        assert np.all(err <= np.ldexp(1.0, e - 30))
        big = np.abs(u) < np.ldexp(1.0, e) - np.ldexp(1.0, e - 31)
        assert np.all(err[big] <= np.ldexp(1.0, e[big] - 31))
        # against the segment's largest value: 2^-29 (2^-30 above clamping)
        smax = bm.segment_max(u)
        ok = smax > np.ldexp(1.0, -120)
        assert np.all(err[ok] <= np.ldexp(smax[ok], -29))


def test_representable_values_round_trip_exactly():
    rng = np.random.default_rng(3)
    for u in _fields(rng):
        w = bm.encode(u)
        assert np.array_equal(bm.encode(bm.decode(w)), w)
        assert np.array_equal(bm.encode(-bm.decode(w)) >> 1, -(w.astype(np.int64) >> 1))


def test_hand_worked_words():
    # one segment: max 0.75 -> e = 0, q = v 2^30, bits of e + 128 = 0b10000000
    u = np.zeros(16)
    u[0], u[1], u[7], u[15] = 0.75, -0.5, 0.25, 2.0 ** -30
    w = bm.encode(u).astype(np.int64)
    assert w[0] == (int(0.75 * 2 ** 30) << 1) | 0
    assert w[1] == ((-(2 ** 29)) << 1) | 0
    assert w[7] == ((2 ** 28) << 1) | 1          # bit 7 of 128
    assert w[15] == 1 << 1                        # 2^-30 is one unit
    assert all(w[k] & 1 == 0 for k in range(7))
    # 1.0 needs e = 1; a value within half a unit of 2^e saturates
    u = np.full(16, 1.0)
    w = bm.encode(u).astype(np.int64)
    assert bm.exponents(w)[0] == 1 and (w[0] >> 1) == 2 ** 29
    u = np.full(16, np.nextafter(2.0, 0.0))
    w = bm.encode(u).astype(np.int64)
    assert bm.exponents(w)[0] == 1 and (w[0] >> 1) == 2 ** 30 - 1
