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


def _pair_encode(u):
    """The x-line Poisson kernels' encoder, lane by lane: lane i of a
    segment's 8 holds points 2i, 2i+1; the exponent is taken from the largest
    FP64 exponent field of the 16 (sign bit cleared), and point k's word
    carries bit k of e + 128."""
    hi = (u.view(np.uint64) >> np.uint64(32)).astype(np.int64) & 0x7FFFFFFF
    w = np.zeros(u.size, dtype=np.int64)
    for s0 in range(0, u.size, 16):
        e = min(max(int(hi[s0:s0 + 16].max() >> 20) - 1022, -126), 127)
        for i in range(8):
            for t in range(2):
                k = 2 * i + t
                q = int(np.rint(np.ldexp(u[s0 + k], 30 - e)))
                q = min(max(q, -(2 ** 30)), 2 ** 30 - 1)
                w[s0 + k] = ((q << 1) | (((e + 128) >> k) & 1)) & 0xFFFFFFFF
    return w.astype(np.uint32).view(np.int32)


def _pair_decode(w):
    """Their decoder: the bits of points 0, 2, 4, 6 (the first words of lanes
    0-3) and of points 1, 3, 5, 7 (their second words), interleaved."""
    def spread(x):
        x = (x | (x << 2)) & 0x33
        return (x | (x << 1)) & 0x55
    out = np.zeros(w.size)
    for s0 in range(0, w.size, 16):
        seg = w[s0:s0 + 16].astype(np.int64)
        a = sum((int(seg[2 * i]) & 1) << i for i in range(4))
        b = sum((int(seg[2 * i + 1]) & 1) << i for i in range(4))
        eb = spread(a) | (spread(b) << 1)
        out[s0:s0 + 16] = np.ldexp((seg >> 1).astype(np.float64), eb - 158)
    return out


def test_two_points_per_lane_coder_equals_the_format():
    rng = np.random.default_rng(11)
    for u in _fields(rng):
        flat = np.ascontiguousarray(u).reshape(-1)
        w = bm.encode(u).reshape(-1)
        assert np.array_equal(_pair_encode(flat), w)
        assert np.array_equal(_pair_decode(w), bm.decode(w.reshape(u.shape)).reshape(-1))
