"""The oracle's y-stretching (istret /= 0), pinned without retyping it.

The reference multiplies the results of dery by ppy(j) (src/x3d_derive.f90:
331-337, 427-433), of deryvp by ppyi(j) and of derypv by ppy(j)
(src/x3d_staggered.f90:628-634, 874-880), each right after the line solve.
The checks:
  * unit vectors change nothing, bit for bit;
  * only the y-derivatives change, each by its own vector, and only there;
  * stretching by 2 equals halving the y length: the y coefficients a/dy
    double exactly, so momentum_rhs, divergence and gradp (whose y
    derivatives the reference scales one call at a time) agree bit for bit
    with the unstretched grid of length yly/2, a path that never calls the
    stretching code;
  * on a mapped grid y = h(eta) the stretched derivative with ppy = 1/h'(eta)
    is the physical derivative df/dy (the meaning of the vectors), and
    without them it is off by the mapping's slope.
"""
import numpy as np

import oracle

Y_OPS_PPY = ("dery", "deryp", "derypv")
Y_OPS_PPYI = ("deryvp",)
UNSCALED = ("derx", "derxp", "derz", "derzp", "intyvp", "intypv", "derxvp", "derzpv", "intxvp", "intzpv")


def _fields(orc, seed):
    return np.random.default_rng(seed).standard_normal(orc.shape)


def test_unit_vectors_change_nothing():
    orc = oracle.Oracle(16, 24, 12, 1.0, 2.0, 1.5)
    ref = oracle.Oracle(16, 24, 12, 1.0, 2.0, 1.5)
    orc.set_stretching(1, np.ones(24), np.ones(24))
    u = _fields(orc, 1)
    for name in Y_OPS_PPY + Y_OPS_PPYI + UNSCALED:
        assert np.array_equal(orc.apply(name, u), ref.apply(name, u)), name


def test_only_y_derivatives_scale_each_by_its_vector():
    ny = 24
    orc = oracle.Oracle(16, ny, 12, 1.0, 2.0, 1.5)
    ref = oracle.Oracle(16, ny, 12, 1.0, 2.0, 1.5)
    rng = np.random.default_rng(2)
    ppy = 1.0 + 0.5 * rng.random(ny)
    ppyi = 0.5 + rng.random(ny)
    orc.set_stretching(1, ppy, ppyi)
    u = _fields(orc, 3)
    for name, s in [(n, ppy) for n in Y_OPS_PPY] + [(n, ppyi) for n in Y_OPS_PPYI]:
        assert np.array_equal(orc.apply(name, u), ref.apply(name, u) * s[None, :, None]), name
    for name in UNSCALED:
        assert np.array_equal(orc.apply(name, u), ref.apply(name, u)), name
    # istret = 0 turns it off again
    orc.set_stretching(0)
    assert np.array_equal(orc.apply("dery", u), ref.apply("dery", u))


def test_stretch_by_two_is_half_the_y_length():
    nx, ny, nz = 16, 24, 12
    orc = oracle.Oracle(nx, ny, nz, 1.0, 2.0, 1.5)
    orc.set_stretching(1, np.full(ny, 2.0), np.full(ny, 2.0))
    half = oracle.Oracle(nx, ny, nz, 1.0, 1.0, 1.5)
    u = orc.fields(1, 7, 1e-3)
    for a, b in zip(orc.momentum_rhs(*u), half.momentum_rhs(*u)):
        assert np.array_equal(a, b)
    assert np.array_equal(orc.divergence(*u), half.divergence(*u))
    p = _fields(orc, 8)
    for a, b in zip(orc.gradp(p), half.gradp(p)):
        assert np.array_equal(a, b)
    # and a non-trivial scaling is not a length change: the results move
    orc.set_stretching(1, np.full(ny, 3.0), np.full(ny, 3.0))
    assert not np.array_equal(orc.divergence(*u), half.divergence(*u))


def test_mapped_grid_derivative_is_physical():
    """y = h(eta) = eta + eps sin(2 pi eta) / (2 pi) maps the periodic unit
    interval onto itself; f(y) = sin(2 pi y) sampled at y_j = h(eta_j).  The
    compact scheme differentiates in eta (uniform, dy = 1/ny); ppy = 1/h'(eta)
    turns that into df/dy at the velocity points, ppyi = 1/h'(eta_{j+1/2})
    into df/dy at the pressure points for deryvp."""
    nx, ny, nz, eps = 8, 64, 8, 0.3
    orc = oracle.Oracle(nx, ny, nz, 1.0, 1.0, 1.0)
    h = lambda e: e + eps * np.sin(2 * np.pi * e) / (2 * np.pi)
    dh = lambda e: 1.0 + eps * np.cos(2 * np.pi * e)
    eta = np.arange(ny) / ny
    etai = (np.arange(ny) + 0.5) / ny
    f = np.broadcast_to(np.sin(2 * np.pi * h(eta))[None, :, None], orc.shape).copy()
    exact = 2 * np.pi * np.cos(2 * np.pi * h(eta))
    exacti = 2 * np.pi * np.cos(2 * np.pi * h(etai))
    plain = orc.apply("dery", f)[0, :, 0]
    orc.set_stretching(1, 1.0 / dh(eta), 1.0 / dh(etai))
    got = orc.apply("dery", f)[0, :, 0]
    goti = orc.apply("deryvp", f)[0, :, 0]
    scale = 2 * np.pi
    assert np.abs(got - exact).max() / scale < 1e-6
    assert np.abs(goti - exacti).max() / scale < 1e-6
    # unscaled, the eta-derivative misses by the slope of the mapping
    assert np.abs(plain - exact).max() / scale > 0.1
