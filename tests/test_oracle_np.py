"""The oracle's non-periodic first derivatives (der?_11/12/21/22 with npaire,
src/x3d_derive.f90:156-275) pinned by what the mathematics fixes:

* free-slip at both ends (1,1): with npaire = 1 the boundary rows are the
  periodic scheme on the even extension of the line (period 2(n-1) points),
  with npaire = 0 on the odd extension.  cos(k x) and sin(k x), k = pi m / L,
  are then exact eigenvectors: the result is -k' sin(k x), resp. k' cos(k x),
  with the modified wavenumber of the sixth-order scheme (Lele 1992, a = 14/9,
  b = 1/9, alfa = 1/3) k' h = (a sin kh + (b/2) sin 2kh) / (1 + 2 alfa cos kh);
* Dirichlet closures (one-sided third order at the last point, fourth-order
  Pade next to it) and free-slip ends on fields of the matching parity are
  exact for polynomials up to degree 3, so the derivative of such a cubic is
  exact to rounding at every point, for every combination of ends;
* the Thomas vectors invert their own rows: (rows) @ D has the stencil's
  sparsity, five diagonals, for D the operator's matrix.
"""
import numpy as np
import pytest

import oracle

L = 2.0


def kprime(k, h):
    a, b, al = 14.0 / 9.0, 1.0 / 9.0, 1.0 / 3.0
    return (a * np.sin(k * h) + 0.5 * b * np.sin(2 * k * h)) / (1 + 2 * al * np.cos(k * h)) / h


def line_field(f, n, axis, other=(3, 2)):
    """f(x) on the grid x_i = i L / (n - 1) along axis, repeated on the others."""
    x = np.arange(n) * (L / (n - 1))
    v = f(x)
    shape = {"x": (other[1], other[0], n), "y": (other[1], n, other[0]), "z": (n, other[1], other[0])}[axis]
    if axis == "x":
        return np.broadcast_to(v, shape).copy(), x
    if axis == "y":
        return np.broadcast_to(v[None, :, None], shape).copy(), x
    return np.broadcast_to(v[:, None, None], shape).copy(), x


def along(t, axis):
    return {"x": t[0, 0, :], "y": t[0, :, 0], "z": t[:, 0, 0]}[axis]


@pytest.mark.parametrize("axis", ["x", "y", "z"])
@pytest.mark.parametrize("n", [9, 16, 33])
def test_free_slip_closed_form(axis, n):
    h = L / (n - 1)
    for m in range(1, n - 1):
        k = np.pi * m / L
        uc, x = line_field(lambda x: np.cos(k * x), n, axis)
        us, _ = line_field(lambda x: np.sin(k * x), n, axis)
        dc = along(oracle.der_np(uc, axis, L, 1, 1, 1), axis)
        ds = along(oracle.der_np(us, axis, L, 1, 1, 0), axis)
        kp = kprime(k, h)
        assert np.abs(dc + kp * np.sin(k * x)).max() < 1e-12 * max(kp, 1.0), (m, "cos")
        assert np.abs(ds - kp * np.cos(k * x)).max() < 1e-12 * max(kp, 1.0), (m, "sin")


# (ncl1, ncln, npaire) -> a cubic with the parity each free-slip end needs
# (even about it for npaire = 1, odd for npaire = 0) and its derivative
CUBICS = {
    (2, 2, 0): (lambda x: 0.3 - 1.1 * x + 0.7 * x ** 2 - 0.4 * x ** 3, lambda x: -1.1 + 1.4 * x - 1.2 * x ** 2),
    (2, 2, 1): (lambda x: 0.3 - 1.1 * x + 0.7 * x ** 2 - 0.4 * x ** 3, lambda x: -1.1 + 1.4 * x - 1.2 * x ** 2),
    (1, 2, 1): (lambda x: 0.5 + 0.9 * x ** 2, lambda x: 1.8 * x),
    (1, 2, 0): (lambda x: 1.3 * x - 0.6 * x ** 3, lambda x: 1.3 - 1.8 * x ** 2),
    (2, 1, 1): (lambda x: 0.5 + 0.9 * (x - L) ** 2, lambda x: 1.8 * (x - L)),
    (2, 1, 0): (lambda x: 1.3 * (x - L) - 0.6 * (x - L) ** 3, lambda x: 1.3 - 1.8 * (x - L) ** 2),
    (1, 1, 1): (lambda x: 0.7 + 0.0 * x, lambda x: 0.0 * x),
}


@pytest.mark.parametrize("axis", ["x", "y", "z"])
@pytest.mark.parametrize("bc", sorted(CUBICS))
def test_parity_cubics_exact(axis, bc):
    ncl1, ncln, npaire = bc
    f, df = CUBICS[bc]
    for n in (8, 13, 32):
        u, x = line_field(f, n, axis)
        d = along(oracle.der_np(u, axis, L, ncl1, ncln, npaire), axis)
        assert np.abs(d - df(x)).max() < 1e-11, (bc, n, np.abs(d - df(x)).max())


def test_wrong_parity_is_not_exact():
    """The checks above can fail: the odd cubic under the even (npaire = 1)
    free-slip closure is not differentiated exactly."""
    n = 16
    u, x = line_field(lambda x: 1.3 * x - 0.6 * x ** 3, n, "x")
    d = along(oracle.der_np(u, "x", L, 1, 2, 1), "x")
    assert np.abs(d - (1.3 - 1.8 * x ** 2)).max() > 1e-3


@pytest.mark.parametrize("ncl1,ncln", [(1, 1), (1, 2), (2, 1), (2, 2)])
@pytest.mark.parametrize("npaire", [0, 1])
def test_thomas_inverts_rows(ncl1, ncln, npaire):
    n = 12
    h = L / (n - 1)
    tb = oracle.np_tables(n, h, ncl1, ncln)
    lo, dg, up = tb["rowsp" if npaire else "rows"]
    A = np.diag(dg) + np.diag(up[:-1], 1) + np.diag(lo[:-1], -1)
    # the operator's matrix, one unit vector at a time
    D = np.zeros((n, n))
    for j in range(n):
        e = np.zeros((1, 1, n))
        e[0, 0, j] = 1.0
        D[:, j] = oracle.der_np(e, "x", L, ncl1, ncln, npaire)[0, 0]
    B = A @ D
    mask = np.abs(np.subtract.outer(np.arange(n), np.arange(n))) <= 2
    assert np.abs(B[~mask]).max() < 1e-12 / h
    # interior rows: the explicit sixth-order stencil, antisymmetric
    k = oracle.np_consts(h)
    for i in range(2, n - 2):
        assert np.allclose(B[i, i - 2:i + 3], [-k[2], -k[1], 0.0, k[1], k[2]], rtol=1e-12, atol=1e-12 / h)
