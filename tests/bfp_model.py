"""numpy model of the BFP32 word format (the format of x3d_bfp_encode).

x is cut into 16-point segments; the segment's exponent e is the smallest with
|v| < 2^e over its values (clamped to [-126, 127]); word k of a segment is
(q << 1) | b with q = rint(v 2^(30-e)) clamped to [-2^30, 2^30 - 1] and b =
bit k of e + 128 for k < 8, 0 for k >= 8.  Test infrastructure: the checker
of the library's encode and decode kernels.
"""
import numpy as np

SEG = 16


def encode(u: np.ndarray) -> np.ndarray:
    u = np.asarray(u, dtype=np.float64)
    assert u.shape[-1] % SEG == 0
    seg = u.reshape(-1, SEG)
    m = np.abs(seg).max(axis=1)
    _, e = np.frexp(m)
    e = np.where(m > 0, e, -126)
    e = np.clip(e, -126, 127).astype(np.int64)
    s = np.rint(np.ldexp(seg, (30 - e)[:, None]))
    q = np.clip(s, -2.0 ** 30, 2.0 ** 30 - 1).astype(np.int64)
    k = np.arange(SEG)
    bits = np.where(k < 8, ((e[:, None] + 128) >> k) & 1, 0)
    w = (q << 1) | bits
    return w.astype(np.int32).reshape(u.shape)


def exponents(w: np.ndarray) -> np.ndarray:
    seg = np.asarray(w, dtype=np.int64).reshape(-1, SEG)
    k = np.arange(8)
    return ((seg[:, :8] & 1) << k).sum(axis=1) - 128


def decode(w: np.ndarray) -> np.ndarray:
    w = np.asarray(w)
    seg = w.astype(np.int64).reshape(-1, SEG)
    e = exponents(w)
    return np.ldexp((seg >> 1).astype(np.float64), (e - 30)[:, None]).reshape(w.shape)


def segment_max(u: np.ndarray) -> np.ndarray:
    """max |u| over each point's segment, broadcast to the field's shape."""
    seg = np.abs(np.asarray(u, dtype=np.float64)).reshape(-1, SEG)
    return np.repeat(seg.max(axis=1), SEG).reshape(np.shape(u))
