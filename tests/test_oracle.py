"""Pins the CPU oracle by routes independent of its own code.

The reference ships no tests, fixtures or golden vectors and cannot be built
here (it needs 2decomp-fft and MPI), so every oracle function is checked
against what the mathematics fixes:
  * scheme constants: the Taylor-order conditions of the compact schemes;
  * operators: the exact Fourier symbol of each periodic operator (collocated,
    half-cell shifted) on single modes, in every direction, and a numpy-FFT
    implementation of the same symbols on random fields;
  * Thomas + Sherman-Morrison: dense solves of the circulant [alfa, 1, alfa];
  * FFT: numpy.fft; Poisson: the analytic cancellation of the phase-shift
    chain, single-mode solutions, null modes and the identity
    div(grad p) = -K p;
  * the step: an independent composition of the FFT-symbol operators.
Each check is chosen so that a dropped term, wrong sign, index shift or swapped
operand in the oracle fails it.
"""
import numpy as np
import pytest

from oracle import Oracle, prepare, uniform
import oracle as O

TAU = 2 * np.pi


# ---------------------------------------------------------------------------
# symbols computed from the scheme constants (Fourier transform of the stencils)
# ---------------------------------------------------------------------------
def sym_d(c, w):          # collocated first derivative: i * Lambda
    return (2 * c["afi"] * np.sin(w) + 2 * c["bfi"] * np.sin(2 * w)) / (1 + 2 * c["alfai"] * np.cos(w))


def sym_ds(c, w):         # staggered first derivative (per unit half-shift)
    return (2 * c["aci"] * np.sin(w / 2) + 2 * c["bci"] * np.sin(1.5 * w)) / (1 + 2 * c["alcai"] * np.cos(w))


def sym_i(c, w):          # mid-point interpolation transfer function
    num = 2 * (c["aici"] * np.cos(w / 2) + c["bici"] * np.cos(1.5 * w) + c["cici"] * np.cos(2.5 * w)
               + c["dici"] * np.cos(3.5 * w))
    return num / (1 + 2 * c["ailcai"] * np.cos(w))


def spec_apply(u, axis, n, c, kind):
    """Apply a periodic operator through its Fourier symbol (complex FFT)."""
    U = np.fft.fft(u, axis=axis)
    w = TAU * np.fft.fftfreq(n)            # signed frequencies, rad per grid step
    if kind == "d":
        m = 1j * sym_d(c, w)
    elif kind == "dvp":
        m = 1j * sym_ds(c, w) * np.exp(0.5j * w)
    elif kind == "dpv":
        m = 1j * sym_ds(c, w) * np.exp(-0.5j * w)
    elif kind == "ivp":
        m = sym_i(c, w) * np.exp(0.5j * w)
    else:
        m = sym_i(c, w) * np.exp(-0.5j * w)
    shape = [1, 1, 1]
    shape[axis] = n
    return np.real(np.fft.ifft(U * m.reshape(shape), axis=axis))


def rel(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(b).max(), 1e-300)


@pytest.fixture(scope="module")
def orc():
    return Oracle(16, 24, 32, 1.0, 2.0, 3.0)


# ---------------------------------------------------------------------------
# coefficients
# ---------------------------------------------------------------------------
def test_first_derivative_is_sixth_order(orc):
    for d, h in zip(range(3), (1 / 16, 2 / 24, 3 / 32)):
        c = orc.consts()[d]
        a, b, al = 2 * h * c["afi"], 4 * h * c["bfi"], c["alfai"]
        # Taylor conditions of alfa f'_{i-1} + f'_i + alfa f'_{i+1} = a d1 + b d2
        assert abs(a + b - (1 + 2 * al)) < 1e-14
        assert abs(a + 4 * b - 6 * al) < 1e-14
        assert abs(a + 16 * b - 10 * al) < 1e-14
        # and the symbol error decays like w^7
        e = [abs(sym_d(c, w) * h - w) for w in (0.2, 0.1)]
        assert 100 < e[0] / e[1] < 160


def test_staggered_and_interpolation_orders(orc):
    c = orc.consts()[0]
    h = 1 / 16
    e = [abs(sym_ds(c, w) * h - w) for w in (0.2, 0.1)]
    assert 100 < e[0] / e[1] < 160                     # 6th order: ratio ~ 2^7
    assert abs(sym_i(c, 0.0) - 1.0) < 1e-15              # interpolation preserves constants
    e = [abs(sym_i(c, w) - 1.0) for w in (0.2, 0.1)]
    assert 50 < e[0] / e[1] < 80                         # error ~ w^6 (ipinter = 3)
    assert abs(sym_i(c, np.pi)) < 1e-14                  # transfer vanishes at Nyquist


def test_prepare_and_thomas1d_solve_the_tridiagonal():
    rng = np.random.default_rng(0)
    for n in (8, 13, 40):
        b = rng.uniform(-0.4, 0.4, n); f = rng.uniform(-0.4, 0.4, n)
        c = 1.0 + rng.uniform(0, 0.5, n)
        b[-1] = 0.0; f[-1] = 0.0
        s, w = prepare(b, c, f)
        A = np.diag(c) + np.diag(f[:-1], 1) + np.diag(b[:-1], -1)
        r = rng.standard_normal(n)
        y = r.copy()
        O.lib().orc_thomas1d(O._p(y), O._p(np.ascontiguousarray(f)), O._p(s), O._p(w), n)
        np.testing.assert_allclose(y, np.linalg.solve(A, r), rtol=1e-13, atol=1e-14)


@pytest.mark.parametrize("n", [8, 9, 16, 33, 64])
def test_periodic_thomas_is_the_circulant_solve(n):
    """Thomas sweeps + the periodic correction == dense solve of [alfa, 1, alfa]."""
    o = Oracle(n, 8, 8)
    rng = np.random.default_rng(n)
    for name in ("derx", "derxvp", "intxvp", "intxpv"):
        f, s, w, p, alfa, _ = o.op(name)
        C = np.eye(n) + alfa * (np.roll(np.eye(n), 1, 1) + np.roll(np.eye(n), -1, 1))
        r = rng.standard_normal(o.shape)
        got = o.thomas(name, r)
        ref = np.linalg.solve(C, r.reshape(-1, n).T).T.reshape(o.shape)
        assert rel(got, ref) < 1e-13, name


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_plain_sweeps_solve_the_factored_tridiagonal(axis):
    """xthomas_12 / ythomas_12 / zthomas_12 (src/thomas.f90:61-84, 114-136,
    166-188): with an operator's f, s, w they solve, line by line, the
    tridiagonal system those vectors factor.  prepare() keeps upper f,
    multipliers s_i = b_{i-1}/p_{i-1} and inverted pivots w_i = 1/p_i with
    p_i = c_i - f_{i-1} s_i, so b and c follow back from f, s, w."""
    dims = [8, 8, 8]
    dims[axis] = 24
    o = Oracle(*dims, 1.0, 2.0, 3.0)
    name = ["derx", "derypv", "intzvp"][axis]
    f, s, w, _, _, _ = o.op(name)
    n = dims[axis]
    piv = 1.0 / w
    c = piv.copy()
    b = np.zeros(n)
    for i in range(1, n):
        b[i - 1] = s[i] * piv[i - 1]
        c[i] = piv[i] + f[i - 1] * s[i]
    A = np.diag(c) + np.diag(f[:-1], 1) + np.diag(b[:-1], -1)
    r = np.random.default_rng(axis).standard_normal(o.shape)
    x = o.thomas12(name, r)
    # lines along the axis: move it last, then A acts on the last index
    xl = np.moveaxis(x, 2 - axis, -1)
    rl = np.moveaxis(r, 2 - axis, -1)
    assert rel(xl @ A.T, rl) < 1e-13


def sym_dd(w, h):
    """Fourier symbol of the sixth-order compact second derivative (Lele
    1992, alfa = 2/11, a = 12/11, b = 3/11): the factor it multiplies
    e^{i w j} by."""
    al, a, b = 2 / 11, 12 / 11, 3 / 11
    return -(2 * a * (1 - np.cos(w)) + 0.5 * b * (1 - np.cos(2 * w))) / (h * h * (1 + 2 * al * np.cos(w)))


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_second_derivative_symbol_and_order(axis):
    """derxx / deryy / derzz: the reference declares them and implements no
    scheme (src/x3d_derive.f90:54-59), so the oracle's is pinned by the
    mathematics only.  Single modes give back their closed-form symbol (a
    wrong sign, index or coefficient shows), the symbol matches -w^2/h^2 with
    an error of order w^8 (sixth order), and random fields match the FFT of
    the symbol."""
    dims = [8, 8, 8]
    dims[axis] = 48
    L = [1.0, 2.0, 3.0]
    o = Oracle(*dims, *L)
    n, h = dims[axis], L[axis] / dims[axis]
    name = "der" + "xyz"[axis] * 2
    x = np.arange(n)
    shape = [1, 1, 1]
    shape[2 - axis] = n
    for k in (1, 3, 7, 24):
        w = 2 * np.pi * k / n
        u = np.broadcast_to(np.cos(w * x + 0.3).reshape(shape), o.shape).copy()
        got = o.apply(name, u)
        assert rel(got, sym_dd(w, h) * u) < 1e-12, k
    e = [abs(sym_dd(w, 1.0) + w * w) for w in (0.2, 0.1)]
    assert 230 < e[0] / e[1] < 280  # 2^8
    r = np.random.default_rng(axis).standard_normal(o.shape)
    U = np.fft.fft(r, axis=2 - axis)
    m = sym_dd(2 * np.pi * np.fft.fftfreq(n), h).reshape(shape)
    assert rel(o.apply(name, r), np.real(np.fft.ifft(U * m, axis=2 - axis))) < 1e-12


def test_operator_tables_equal_between_pairs(orc):
    # derX and derXp, intXvp and intXpv, derXvp and derXpv share tables (periodic)
    for a, b in (("derx", "derxp"), ("intxvp", "intxpv"), ("derxvp", "derxpv"), ("derz", "derzp")):
        for x, y in zip(orc.op(a)[:5], orc.op(b)[:5]):
            np.testing.assert_array_equal(x, y)


# ---------------------------------------------------------------------------
# operators
# ---------------------------------------------------------------------------
OPS = [("derx", "d"), ("dery", "d"), ("derz", "d"), ("derxvp", "dvp"), ("deryvp", "dvp"),
       ("derzvp", "dvp"), ("derxpv", "dpv"), ("derypv", "dpv"), ("derzpv", "dpv"),
       ("intxvp", "ivp"), ("intyvp", "ivp"), ("intzvp", "ivp"), ("intxpv", "ipv"),
       ("intypv", "ipv"), ("intzpv", "ipv")]


@pytest.mark.parametrize("name,kind", OPS)
def test_operator_single_modes(orc, name, kind):
    axis = "zyx".index(name[3])             # numpy axis of the operator's direction
    d = "xyz".index(name[3])
    n = orc.shape[axis]
    c = orc.consts()[d]
    idx = np.arange(n)
    for m in (1, 3, n // 2 - 1):
        w = TAU * m / n
        phi = 0.3 + m
        u1 = np.sin(w * idx + phi)
        shape = [1, 1, 1]; shape[axis] = n
        u = np.broadcast_to(u1.reshape(shape), orc.shape).copy()
        t = orc.apply(name, u)
        if kind == "d":
            ex = sym_d(c, w) * np.cos(w * idx + phi)
        elif kind == "dvp":
            ex = sym_ds(c, w) * np.cos(w * (idx + 0.5) + phi)
        elif kind == "dpv":
            ex = sym_ds(c, w) * np.cos(w * (idx - 0.5) + phi)
        elif kind == "ivp":
            ex = sym_i(c, w) * np.sin(w * (idx + 0.5) + phi)
        else:
            ex = sym_i(c, w) * np.sin(w * (idx - 0.5) + phi)
        ex = np.broadcast_to(ex.reshape(shape), orc.shape)
        assert np.abs(t - ex).max() < 1e-12 * max(1.0, np.abs(ex).max()), (name, m)


@pytest.mark.parametrize("name,kind", OPS)
def test_operator_random_fields_vs_fft_symbols(orc, name, kind):
    axis = "zyx".index(name[3])
    d = "xyz".index(name[3])
    u = np.random.default_rng(1).standard_normal(orc.shape)
    ref = spec_apply(u, axis, orc.shape[axis], orc.consts()[d], kind)
    assert rel(orc.apply(name, u), ref) < 1e-12


def test_operator_linearity_and_constants(orc):
    rng = np.random.default_rng(2)
    a, b = rng.standard_normal(orc.shape), rng.standard_normal(orc.shape)
    for name in ("derx", "intyvp", "derzpv"):
        lhs = orc.apply(name, 2.5 * a - 1.5 * b)
        rhs = 2.5 * orc.apply(name, a) - 1.5 * orc.apply(name, b)
        assert rel(lhs, rhs) < 1e-13
    one = np.ones(orc.shape)
    assert np.abs(orc.apply("derx", one)).max() < 1e-13
    assert np.abs(orc.apply("intzpv", one) - 1.0).max() < 1e-14


# ---------------------------------------------------------------------------
# Poisson
# ---------------------------------------------------------------------------
def test_fft_matches_numpy(orc):
    r = np.random.default_rng(3).standard_normal(orc.shape)
    cw = orc.r2c(r)
    ref = np.fft.rfftn(r.transpose(2, 1, 0)).transpose(2, 1, 0)   # z halved
    assert rel(cw, ref) < 1e-14
    assert rel(orc.c2r(cw) / r.size, r) < 1e-14


def test_spectral_kernel_phase_chain_cancels(orc):
    """poisson_000's shifts/flips/conjugations reduce to c -> -c / (K ntot)."""
    rng = np.random.default_rng(4)
    cw = rng.standard_normal((orc.nz // 2 + 1, orc.ny, orc.nx)) + 1j * rng.standard_normal(
        (orc.nz // 2 + 1, orc.ny, orc.nx))
    K, Kim = orc.kxyz()
    np.testing.assert_array_equal(K, Kim)
    ntot = orc.nx * orc.ny * orc.nz
    expect = np.where(K < 1e-16, 0.0, -cw / np.where(K < 1e-16, 1.0, K) / ntot)
    assert rel(orc.spectral(cw), expect) < 1e-13


def test_null_modes(orc):
    K, _ = orc.kxyz()
    nx, ny, nz = orc.nx, orc.ny, orc.nz
    assert K[0, 0, 0] == 0.0
    # any two Nyquist indices: the interpolation transfer vanishes -> K ~ 0
    assert K[nz // 2, ny // 2, 0] < 1e-16 and K[nz // 2, 0, nx // 2] < 1e-16 and K[0, ny // 2, nx // 2] < 1e-16
    # one Nyquist index alone is a regular mode
    assert K[0, 0, nx // 2] > 1.0 and K[nz // 2, 0, 0] > 1.0


def test_div_grad_is_minus_K(orc):
    """For one Fourier mode p, divergence(gradp(p)) = -K p (pins kxyz to the operators)."""
    K, _ = orc.kxyz()
    k, j, i = 3, 5, 2
    z, y, x = np.meshgrid(np.arange(orc.nz), np.arange(orc.ny), np.arange(orc.nx), indexing="ij")
    p = np.cos(TAU * (i * x / orc.nx + j * y / orc.ny + k * z / orc.nz) + 0.4)
    g = orc.gradp(p)
    dd = orc.divergence(*g)
    assert rel(dd, -K[k, j, i] * p) < 1e-12


def test_poisson_single_mode_and_projection(orc):
    K, _ = orc.kxyz()
    k, j, i = 2, 3, 5
    z, y, x = np.meshgrid(np.arange(orc.nz), np.arange(orc.ny), np.arange(orc.nx), indexing="ij")
    rhs = np.sin(TAU * (i * x / orc.nx + j * y / orc.ny + k * z / orc.nz) + 1.1)
    assert rel(orc.poisson(rhs), -rhs / K[k, j, i]) < 1e-12
    # projection: the corrected field is discretely divergence free
    u = orc.fields(1, 17, 1e-3)
    pp, g = orc.solve_poisson(*u)
    d0 = orc.divergence(*u)
    d1 = orc.divergence(*[a - b for a, b in zip(u, g)])
    assert np.abs(d1).max() < 1e-11 * np.abs(d0).max()


# ---------------------------------------------------------------------------
# the step
# ---------------------------------------------------------------------------
def test_momentum_rhs_is_the_skew_symmetric_sum(orc):
    """du_c = sum_j [ d(u_j u_c)/dx_j + u_j du_c/dx_j ] (src/transeq.f90:99-351)."""
    u = orc.fields(2, 5)
    cs = orc.consts()
    D = [lambda f, a=a, d=d: spec_apply(f, a, orc.shape[a], cs[d], "d") for d, a in ((0, 2), (1, 1), (2, 0))]
    ref = []
    for c in range(3):
        acc = 0.0
        for j in range(3):
            acc = acc + D[j](u[j] * u[c]) + u[j] * D[j](u[c])
        ref.append(acc)
    got = orc.momentum_rhs(*u)
    for a, b in zip(got, ref):
        assert rel(a, b) < 1e-12


def test_divergence_and_gradp_compositions(orc):
    cs = orc.consts()
    sh = orc.shape
    A = {0: 2, 1: 1, 2: 0}

    def op(f, d, kind):
        return spec_apply(f, A[d], sh[A[d]], cs[d], kind)
    u = orc.fields(2, 9)
    div = (op(op(op(u[0], 0, "dvp"), 1, "ivp"), 2, "ivp") + op(op(op(u[1], 0, "ivp"), 1, "dvp"), 2, "ivp")
           + op(op(op(u[2], 0, "ivp"), 1, "ivp"), 2, "dvp"))
    assert rel(orc.divergence(*u), div) < 1e-12
    p = np.random.default_rng(6).standard_normal(sh)
    gx = op(op(op(p, 2, "ipv"), 1, "ipv"), 0, "dpv")
    gy = op(op(op(p, 2, "ipv"), 1, "dpv"), 0, "ipv")
    gz = op(op(op(p, 2, "dpv"), 1, "ipv"), 0, "ipv")
    for a, b in zip(orc.gradp(p), (gx, gy, gz)):
        assert rel(a, b) < 1e-12


def test_step_euler_and_correction(orc):
    u0 = orc.fields(1, 3, 1e-3)
    gdt = 1e-3
    du = orc.momentum_rhs(*u0)
    ustar = [a + gdt * b for a, b in zip(u0, du)]
    pp, g = orc.solve_poisson(*ustar)
    u1, du1, pp1, g1 = orc.step(*u0, gdt=gdt)
    for a, b in zip(u1, [a - b for a, b in zip(ustar, g)]):
        assert rel(a, b) < 1e-14
    # dt = 0 (the reference's default): the step is the projection of u
    uz, _, _, _ = orc.step(*u0, gdt=0.0)
    pp0, g0 = orc.solve_poisson(*u0)
    for a, b in zip(uz, [a - b for a, b in zip(u0, g0)]):
        assert rel(a, b) < 1e-14


# ---------------------------------------------------------------------------
# fields and test mode
# ---------------------------------------------------------------------------
def test_generators():
    v = np.array([uniform(7, 1, i) for i in range(20000)])
    assert v.min() >= 0.0 and v.max() < 1.0 and abs(v.mean() - 0.5) < 0.01
    assert uniform(7, 1, 5) != uniform(7, 2, 5) and uniform(7, 1, 5) != uniform(8, 1, 5)
    o = Oracle(8, 8, 8, TAU, TAU, TAU)
    u = o.fields(1, 3, 0.0)
    x = np.arange(8) * TAU / 8
    np.testing.assert_allclose(u[0][1, 2, :], np.sin(x) * np.cos(x[2]) * np.cos(x[1]), atol=1e-15)
    assert np.all(u[2] == 0.0)
    un = o.fields(1, 3, 1e-3)
    assert 0 < np.abs(un[2]).max() <= 1e-3


def test_test_mode_rms_error_closed_form():
    """test_du on vel(): RMS of (Lambda(w) - 2 pi / L) cos cos cos (src/mom.f90:95-96)."""
    o = Oracle(32, 16, 24)
    u = o.fields(0)
    c = o.consts()
    for which, (name, d, n, L) in enumerate([("derx", 0, 32, 1.0), ("dery", 1, 16, 1.0), ("derz", 2, 24, 1.0)]):
        dd = o.apply(name, u[which])
        amp = 2.0 if which == 2 else 1.0
        expect = abs(sym_d(c[d], TAU / n) - TAU / L) * amp * np.sqrt(1 / 8)
        assert abs(o.test_err(dd, which) - expect) < 1e-12 * max(expect, 1e-9) + 1e-14
