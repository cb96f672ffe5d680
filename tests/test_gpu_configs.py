"""The benchmark configurations themselves, on the GPU.

* 64^3 FP64, test mode (config 1): the 2 pi cube with seeded Taylor-Green +
  1e-3 noise.  Every operator, the Thomas solves and Poisson against the oracle;
  one eager step; the test-mode RMS errors of the reference's own field
  (src/mom.f90:59-243).  The y/z single-operator tiling of 64-point lines
  (2-point chunks, 32 per line) and the 64-point inverse y-line Poisson kernel
  run here.
* 256^3 FP64 (config 2): the same fields; every operator, Thomas and Poisson
  against the oracle (x: 4-point chunks, 64 per line; y/z: 8-point chunks,
  32 per line; the 256-point Poisson line kernels), and the hipGraph-captured
  step replayed and compared with the oracle's step.
* 1024 x 768 x 768 FP32 (config 5), where the oracle would take minutes and
  ~100 GB of host memory, through properties that do not depend on the size:
  - closed-form derivatives and interpolations of single Fourier modes along
    y and z (and x), all 18 operators;
  - the FP32 step's momentum right-hand side on separable fields, against
    the same composition evaluated term by term from 1-D oracle operators;
  - the FP32 step's projection on u = grad p0: its divergence and Poisson
    passes must return p0 as the pressure; its gradient and correction
    passes must equal the gradient composition applied to that pressure.

Bars.  FP64 operators, Thomas, Poisson, the velocity and the right-hand side:
1e-12 relative in the max norm.  On the cube the Taylor-Green part is
discretely solenoidal, so the pressure and uz come from the 1e-3 noise alone
and are worse conditioned.  Each output of the step is therefore held to
max(1e-12, 10 s), where s is how far the oracle's own output moves when every
input value is moved by one ulp (measured in the test on the same fields),
and the velocity as a vector to 1e-12 of its largest component.
"""
import numpy as np
import pytest

from oracle import Oracle

pytestmark = pytest.mark.gpu

RTOL = 1e-12
TWO_PI = 2 * np.pi
OPS18 = ["derx", "derxp", "dery", "deryp", "derz", "derzp",
         "derxvp", "deryvp", "derzvp", "derxpv", "derypv", "derzpv",
         "intxvp", "intyvp", "intzvp", "intxpv", "intypv", "intzpv"]


def relerr(a, b):
    a = np.asarray(a); b = np.asarray(b)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def relerr2(a, b):
    a = np.asarray(a); b = np.asarray(b)
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def host(t):
    return t.detach().cpu().numpy()


def ulp_perturbed(u, seed=0):
    """Every value moved by one ulp up or down (seeded)."""
    rng = np.random.default_rng(seed)
    return [np.where(rng.random(a.shape) < 0.5, np.nextafter(a, np.inf), np.nextafter(a, -np.inf))
            for a in u]


@pytest.fixture(scope="module", params=[64, 256], ids=lambda n: f"{n}cube")
def cube(request, torch_cuda):
    from x3div_amd import X3D
    n = request.param
    print(f"\n[config cube {n}^3 FP64]", flush=True)
    ctx = X3D(n, n, n, TWO_PI, TWO_PI, TWO_PI)
    orc = Oracle(n, n, n, TWO_PI, TWO_PI, TWO_PI)
    yield n, ctx, orc
    ctx.close()


@pytest.mark.parametrize("name", OPS18)
def test_cube_operator(cube, name):
    n, ctx, orc = cube
    u = orc.fields(1, 4321, 1e-3)["xyz".index(name[3]) % 2]
    ref = orc.apply(name, u)
    t = ctx.field(fill=np.nan)
    ctx.apply(name, t, dev(u))
    assert relerr(host(t), ref) < RTOL


@pytest.mark.parametrize("axis", ["x", "y", "z"])
def test_cube_thomas(cube, axis):
    n, ctx, orc = cube
    tt = np.random.default_rng(3).standard_normal(orc.shape)
    for name in (f"der{axis}", f"int{axis}vp", f"der{axis}pv"):
        ref = orc.thomas(name, tt)
        t = dev(tt)
        getattr(ctx, f"{axis}thomas")(t, name)
        assert relerr(host(t), ref) < RTOL, name


def test_cube_poisson(cube):
    """Runs the n-point z- and inverse y-line kernels (n = 64, 256)."""
    n, ctx, orc = cube
    rhs = orc.divergence(*orc.fields(1, 17, 1e-3))
    rhs += np.random.default_rng(5).standard_normal(orc.shape)
    ref = orc.poisson(rhs)
    t = dev(rhs)
    ctx.poisson(t)
    assert relerr(host(t), ref) < RTOL


def _check_step(orc, u0, gdt, got):
    """got: the 10 step outputs (u, du, pp3, grad p) from the GPU, as numpy."""
    ref_u, ref_du, ref_pp, ref_g = orc.step(*u0, gdt=gdt)
    pu, pdu, ppp, pg = orc.step(*ulp_perturbed(u0), gdt=gdt)
    ref = ref_u + ref_du + [ref_pp] + ref_g
    pert = pu + pdu + [ppp] + pg
    names = ["ux", "uy", "uz", "dux", "duy", "duz", "pp3", "px", "py", "pz"]
    report = []
    for name, a, b, p in zip(names, got, ref, pert):
        sens = relerr(p, b)
        bar = max(RTOL, 10 * sens)
        e = relerr(a, b)
        report.append(f"{name} {e:.1e} (bar {bar:.1e})")
        assert e < bar, (name, e, bar, report)
    umax = max(np.abs(b).max() for b in ref_u)
    vec = max(np.abs(a - b).max() for a, b in zip(got[:3], ref_u)) / umax
    assert vec < RTOL, vec
    print("  " + ", ".join(report) + f"; |u| vector {vec:.1e}", flush=True)


def test_cube_step(cube):
    """64^3: one eager step.  256^3: the captured step, replayed twice with an
    eager step between, compared with the oracle's first and second steps."""
    import torch
    n, ctx, orc = cube
    gdt = 1e-3
    u0 = orc.fields(1, 2024, 1e-3)
    f = [dev(a) for a in u0] + [ctx.field(fill=np.nan) for _ in range(7)]
    if n == 64:
        ctx.step(*f, gdt=gdt)
        torch.cuda.synchronize()
        _check_step(orc, u0, gdt, [host(t) for t in f])
        return
    ctx.graph_capture_step(*f, gdt=gdt)
    ctx.graph_launch()
    torch.cuda.synchronize()
    got = [host(t) for t in f]
    _check_step(orc, u0, gdt, got)
    # eager work on other arrays between replays, then a second replay from the
    # state the first one left
    junk = ctx.field()
    ctx.derz(junk, f[0])
    torch.cuda.synchronize()
    u1 = [a.copy() for a in got[:3]]
    ctx.graph_launch()
    torch.cuda.synchronize()
    _check_step(orc, u1, gdt, [host(t) for t in f])


def test_cube_test_mode(cube):
    """test=1: RMS error of each derivative of the reference's vel() field
    against its analytic derivative (src/mom.f90:59-243)."""
    n, ctx, orc = cube
    u = [ctx.field() for _ in range(3)]
    ctx.init_fields(*u, kind=0)
    ref_u = orc.fields(0)
    for a, b in zip(u, ref_u):
        np.testing.assert_allclose(host(a), b, rtol=0, atol=1e-14)
    for which, op in enumerate(("derx", "dery", "derz")):
        d = ctx.field()
        ctx.apply(op, d, u[which])
        e_gpu = ctx.test_err(d, which)
        ref_d = orc.apply(op, ref_u[which])
        e_orc = orc.test_err(ref_d, which)
        # |rms(a - f) - rms(b - f)| <= rms(a - b) <= max|a - b|, and the
        # derivative itself is held to 1e-12 of max|b|
        assert abs(e_gpu - e_orc) <= RTOL * np.abs(ref_d).max(), (op, e_gpu, e_orc)
        # sixth order: the RMS error of the n-point derivative is tiny
        assert e_orc < (30.0 / n) ** 6, (op, e_orc)


# ---------------------------------------------------------------------------
# 1024 x 768 x 768, FP32
# ---------------------------------------------------------------------------
LNX, LNY, LNZ = 1024, 768, 768


@pytest.fixture(scope="module")
def large(torch_cuda):
    import torch
    from x3div_amd import X3D
    print(f"\n[config large {LNX}x{LNY}x{LNZ} FP32]", flush=True)
    ctx = X3D(LNX, LNY, LNZ, TWO_PI, TWO_PI, TWO_PI)
    yield ctx
    ctx.close()
    torch.cuda.empty_cache()


def _coords(n):
    import torch
    return torch.arange(n, dtype=torch.float64, device="cuda") * (TWO_PI / n)


def _symbol(kind, w, c):
    """Closed-form symbols (the pinned ones of the oracle's tests): the factor
    an operator multiplies the Fourier mode e^{i w j} by, with the half-cell
    shift of the staggered output.  c: the direction's scheme constants."""
    if kind == "d":
        return 1j * (2 * c["afi"] * np.sin(w) + 2 * c["bfi"] * np.sin(2 * w)) / (1 + 2 * c["alfai"] * np.cos(w))
    if kind in ("dvp", "dpv"):
        s = 1j * (2 * c["aci"] * np.sin(w / 2) + 2 * c["bci"] * np.sin(1.5 * w)) / (1 + 2 * c["alcai"] * np.cos(w))
    else:
        s = (2 * c["aici"] * np.cos(w / 2) + 2 * c["bici"] * np.cos(1.5 * w) + 2 * c["cici"] * np.cos(2.5 * w)
             + 2 * c["dici"] * np.cos(3.5 * w)) / (1 + 2 * c["ailcai"] * np.cos(w))
    shift = 0.5 if kind in ("dvp", "ivp") else -0.5
    return s * np.exp(1j * w * shift)


def _kind(name):
    if name.startswith("der"):
        return "dvp" if name.endswith("vp") else "dpv" if name.endswith("pv") else "d"
    return "ivp" if name.endswith("vp") else "ipv"


@pytest.mark.parametrize("name", OPS18)
def test_large_fp32_single_modes(large, name):
    """Every operator at full size on cos(k s + phi) along its own axis: the
    exact discrete result is Re(sigma(w) e^{i(k s + phi)}), w = 2 pi k / n."""
    import torch
    ctx = large
    axis = "xyz".index(name[3])
    n = (LNX, LNY, LNZ)[axis]
    k, phi = 37, 0.4
    c = ctx.scheme_consts()[axis]
    w = TWO_PI * k / n
    sig = _symbol(_kind(name), w, c)
    s = _coords(n)
    line = torch.cos(k * s + phi)
    expect = sig.real * torch.cos(k * s + phi) - sig.imag * torch.sin(k * s + phi)
    shape = [1, 1, 1]
    shape[2 - axis] = n
    u = line.float().reshape(shape).expand(LNZ, LNY, LNX).contiguous()
    t = ctx.field32(fill=float("nan"))
    ctx.apply_f32(name, t, u)
    torch.cuda.synchronize()
    ref = expect.reshape(shape)
    err = float(((t.double() - ref).abs().max()) / ref.abs().max())
    del u, t
    torch.cuda.empty_cache()
    # FP32 input rounding times the operator's gain on the mode, plus the FP32 solve
    assert err < 2e-5, (name, err)


def _sep_eval(terms):
    """Sum of coef * fx(x) fy(y) fz(z) on the (nz, ny, nx) grid, FP64 on the device."""
    import torch
    out = torch.zeros((LNZ, LNY, LNX), dtype=torch.float64, device="cuda")
    for coef, fx, fy, fz in terms:
        fxy = (coef * fy)[:, None] * fx[None, :]
        out.addcmul_(fz[:, None, None], fxy[None, :, :])
    return out


def test_large_fp32_step_momentum_rhs(large):
    """The FP32 step's du on separable fields (Taylor-Green plus a 1e-3
    separable random perturbation): every product u_j u_c and every
    derivative of a separable term stays separable, so
    du_c = sum_j [D_j(u_j u_c) + u_j D_j u_c] (src/transeq.f90:44-355) is a
    short sum of outer products whose 1-D factors come from the oracle's
    operators on single lines, evaluated in FP64."""
    import torch
    ctx = large
    rng = np.random.default_rng(11)
    xs = [np.arange(n) * (TWO_PI / n) for n in (LNX, LNY, LNZ)]
    # per component: list of (coef, fx, fy, fz) numpy factors
    comps = [
        [(1.0, np.sin(xs[0]), np.cos(xs[1]), np.cos(xs[2]))],
        [(-1.0, np.cos(xs[0]), np.sin(xs[1]), np.cos(xs[2]))],
        [],
    ]
    for cpt in range(3):
        comps[cpt].append((1e-3, *(rng.uniform(-1, 1, n) for n in (LNX, LNY, LNZ))))
    lines = [Oracle(LNX, 8, 8, TWO_PI, 1.0, 1.0), Oracle(8, LNY, 8, 1.0, TWO_PI, 1.0),
             Oracle(8, 8, LNZ, 1.0, 1.0, TWO_PI)]

    def d1(axis, v):
        """The axis's first derivative (derx/dery/derz) of one 1-D factor,
        through the oracle on a line grid (8 copies in each other direction)."""
        shape = [8, 8, 8]  # (nz, ny, nx)
        bshape = [1, 1, 1]
        shape[2 - axis] = bshape[2 - axis] = len(v)
        full = np.ascontiguousarray(np.broadcast_to(v.reshape(bshape), shape))
        out = lines[axis].apply("der" + "xyz"[axis], full)
        idx = [0, 0, 0]
        idx[2 - axis] = slice(None)
        return out[tuple(idx)].copy()

    def deriv(axis, terms):
        out = []
        for coef, *f in terms:
            f = list(f)
            f[axis] = d1(axis, f[axis])
            out.append((coef, *f))
        return out

    def prod(ta, tb):
        return [(ca * cb, *(fa * fb for fa, fb in zip(a, b))) for ca, *a in ta for cb, *b in tb]

    expect = []
    for c in range(3):
        terms = []
        for j in range(3):
            terms += deriv(j, prod(comps[j], comps[c]))
            terms += prod(comps[j], deriv(j, comps[c]))
        expect.append(terms)

    def to_dev(terms):
        return [(float(cf), *(torch.from_numpy(np.ascontiguousarray(f)).cuda() for f in fs)) for cf, *fs in terms]

    f = [ctx.field32() for _ in range(10)]
    for c in range(3):
        f[c].copy_(_sep_eval(to_dev(comps[c])).float())
    ctx.step_f32(*f, gdt=0.0)
    torch.cuda.synchronize()
    errs = []
    for c in range(3):
        ref = _sep_eval(to_dev(expect[c]))
        errs.append(float((f[3 + c].double() - ref).abs().max() / ref.abs().max()))
        del ref
    del f
    torch.cuda.empty_cache()
    print(f"  large fp32 du vs separable closed form: {['%.1e' % e for e in errs]}", flush=True)
    # FP32 fields and FP32 line solves; du_z is an O(1e-3) field made of
    # products with the O(1) field, so its relative error is larger
    assert errs[0] < 5e-5 and errs[1] < 5e-5 and errs[2] < 5e-3, errs


def test_large_fp32_step_projection(large):
    """u = grad p0 (the FP32 gradient composition of gradp, src/navier.f90:
    249-359) with gdt = 0: the step's divergence and Poisson passes must give
    back p0 as the pressure (div grad has the symbol -K that Poisson divides
    by); its gradient passes must equal the same composition applied to that
    pressure, and its correction u - grad p.  The pressure's FP32 rounding is
    white, and a gradient amplifies white noise up to ~1.9/h (about 300 here),
    so grad p is compared with the gradient of the step's own pressure, not
    with grad p0."""
    import torch
    ctx = large
    x, y, z = _coords(LNX), _coords(LNY), _coords(LNZ)
    p0 = torch.zeros((LNZ, LNY, LNX), dtype=torch.float64, device="cuda")
    for a, kx, ky, kz, ph in [(1.0, 1, 2, 3, 0.3), (0.5, 5, 1, 2, 1.1), (0.25, 3, 7, 1, 2.0)]:
        # cos(kx x + ky y + kz z + ph) from separable factors
        cxy = torch.cos(kx * x + ph)[None, :] * torch.cos(ky * y)[:, None] - \
            torch.sin(kx * x + ph)[None, :] * torch.sin(ky * y)[:, None]
        sxy = torch.sin(kx * x + ph)[None, :] * torch.cos(ky * y)[:, None] + \
            torch.cos(kx * x + ph)[None, :] * torch.sin(ky * y)[:, None]
        p0 += a * (cxy[None] * torch.cos(kz * z)[:, None, None] - sxy[None] * torch.sin(kz * z)[:, None, None])
    p0 = p0.float()
    t1, t2 = ctx.field32(), ctx.field32()

    def chain(names, src, out):
        cur = src
        for i, name in enumerate(reversed(names)):
            dst = out if i == len(names) - 1 else (t1 if cur is not t1 else t2)
            ctx.apply_f32(name, dst, cur)
            cur = dst
        return out

    def grad(p, out):
        chain(["derxpv", "intypv", "intzpv"], p, out[0])
        chain(["intxpv", "derypv", "intzpv"], p, out[1])
        chain(["intxpv", "intypv", "derzpv"], p, out[2])

    f = [ctx.field32() for _ in range(10)]
    grad(p0, f[:3])
    g = [t.clone() for t in f[:3]]
    ctx.step_f32(*f, gdt=0.0)
    torch.cuda.synchronize()
    g2 = [ctx.field32() for _ in range(3)]
    grad(f[6], g2)
    torch.cuda.synchronize()
    gmax = max(float(t.abs().max()) for t in g)
    e_p = float((f[6].double() - p0.double()).abs().max() / p0.double().abs().max())
    e_g = max(float((a.double() - b.double()).abs().max()) for a, b in zip(f[7:10], g2)) / gmax
    e_u = max(float((u.double() - (a.double() - b.double())).abs().max())
              for u, a, b in zip(f[:3], g, f[7:10])) / gmax
    e_gp0 = max(float((a.double() - b.double()).abs().max()) for a, b in zip(f[7:10], g)) / gmax
    print(f"  large fp32 projection: pressure {e_p:.1e}, grad p {e_g:.1e}, u - grad p {e_u:.1e}, "
          f"(grad p vs grad p0 {e_gp0:.1e})", flush=True)
    del f, g, g2, p0, t1, t2
    torch.cuda.empty_cache()
    assert e_p < 2e-5 and e_g < 2e-5 and e_u < 1e-6, (e_p, e_g, e_u)
