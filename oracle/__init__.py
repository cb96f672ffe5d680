"""CPU oracle for the periodic x3div Euler-step hot path (ctypes over liborcx3d.so).

TEST INFRASTRUCTURE ONLY.  The oracle is the checker for the HIP library and the
CPU-baseline leg of ``bench.py``; only ``tests/``, ``__graft_entry__.smoke()`` and
``bench.py``'s baseline may import it.  Fields are numpy float64 arrays of shape
``(nz, ny, nx)`` in C order, i.e. the reference's Fortran ``(nx, ny, nz)`` layout
with x fastest.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liborcx3d.so")
# the same source built for speed (-O3, FMA contraction): the CPU baseline only
FAST_LIB_PATH = os.path.join(HERE, "liborcx3d_fast.so")

# operator ids (src/x3d_operator_1d.f90:33-47); identical to the HIP library's
OPS = ["derx", "derxp", "dery", "deryp", "derz", "derzp",
       "derxvp", "deryvp", "derzvp", "derxpv", "derypv", "derzpv",
       "intxvp", "intyvp", "intzvp", "intxpv", "intypv", "intzpv"]
# second derivatives after them (declared and never implemented by the
# reference, src/x3d_derive.f90:54-59)
OPS2 = ["derxx", "deryy", "derzz"]
OP = {name: i for i, name in enumerate(OPS + OPS2)}
# stencil kinds: D, Ds v->p, Ds p->v, I v->p, I p->v, D2
KIND = {"d": 0, "dvp": 1, "dpv": 2, "ivp": 3, "ipv": 4, "dd": 5}


def build(force: bool = False) -> str:
    """Compile the checker (gcc, no FP contraction) and the fast baseline build."""
    if force or not (os.path.exists(LIB_PATH) and os.path.exists(FAST_LIB_PATH)):
        subprocess.run(["make", "-B" if force else "-s", "-C", HERE, "all"],
                       check=True, capture_output=True)
    return LIB_PATH


_libs = {}


def lib(fast: bool = False):
    """The checker library, or with fast=True the speed build (baseline timing only)."""
    if fast not in _libs:
        build()
        L = C.CDLL(FAST_LIB_PATH if fast else LIB_PATH)
        dp = C.POINTER(C.c_double)
        vp = C.c_void_p
        L.orc_new.restype = vp
        L.orc_new.argtypes = [C.c_int] * 3 + [C.c_double] * 3
        L.orc_delete.argtypes = [vp]
        L.orc_get_op.argtypes = [vp, C.c_int, dp, dp, dp, dp, dp, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.orc_get_consts.argtypes = [vp, dp]
        L.orc_get_kxyz.argtypes = [vp, dp, dp]
        L.orc_get_ab.argtypes = [vp, dp]
        L.orc_apply.argtypes = [vp, dp, dp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]
        L.orc_apply.restype = C.c_int
        L.orc_thomas.argtypes = [vp, dp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]
        L.orc_thomas.restype = C.c_int
        for d in "xyz":
            getattr(L, f"orc_{d}thomas_12").argtypes = [dp, dp, dp, dp, C.c_int, C.c_int, C.c_int]
        L.orc_prepare.argtypes = [dp, dp, dp, dp, dp, C.c_int]
        L.orc_thomas1d.argtypes = [dp, dp, dp, dp, C.c_int]
        L.orc_fft_r2c_z.argtypes = [vp, dp, dp]
        L.orc_fft_c2r_z.argtypes = [vp, dp, dp]
        L.orc_poisson_spectral.argtypes = [vp, dp]
        L.orc_poisson.argtypes = [vp, dp]
        L.orc_momentum_rhs.argtypes = [vp] + [dp] * 6
        L.orc_int_time.argtypes = [vp, C.c_double] + [dp] * 6
        L.orc_divergence.argtypes = [vp] + [dp] * 4
        L.orc_gradp.argtypes = [vp] + [dp] * 4
        L.orc_cor_vel.argtypes = [vp] + [dp] * 6
        L.orc_solve_poisson.argtypes = [vp] + [dp] * 7
        L.orc_step.argtypes = [vp, C.c_double] + [dp] * 10
        L.orc_splitmix64.restype = C.c_uint64
        L.orc_splitmix64.argtypes = [C.c_uint64]
        L.orc_uniform.restype = C.c_double
        L.orc_uniform.argtypes = [C.c_uint64, C.c_int, C.c_uint64]
        L.orc_init_fields.argtypes = [vp, C.c_int, C.c_uint64, C.c_double, dp, dp, dp]
        L.orc_test_err.restype = C.c_double
        L.orc_test_err.argtypes = [vp, dp, C.c_int]
        L.orc_div_stats.argtypes = [vp, dp, dp, dp]
        L.orc_np_consts.argtypes = [C.c_double, dp]
        L.orc_np_tables.argtypes = [C.c_int, C.c_double, C.c_int, C.c_int] + [dp] * 8
        L.orc_np_tables.restype = C.c_int
        L.orc_der_np.argtypes = [dp, dp] + [C.c_int] * 4 + [C.c_double] + [C.c_int] * 3 + [dp] * 3
        L.orc_der_np.restype = C.c_int
        L.orc_set_stretching.argtypes = [vp, C.c_int, dp, dp]
        L.orc_set_stretching.restype = C.c_int
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_max_threads.restype = C.c_int
        _libs[fast] = L
    return _libs[fast]


def _p(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_double))


def np_tables(n: int, h: float, ncl1: int, ncln: int):
    """Non-periodic first-derivative tables of first_deriv_imp_
    (src/x3d_operator_1d.f90:395-491): {"f", "s", "w"} of the npaire = 0
    instance (derX), {"fp", "sp", "wp"} of the npaire = 1 one (derXp), and
    the rows (lower, diagonal, upper) each was prepared from."""
    L = lib(False)
    a = {k: np.zeros(n) for k in ("f", "s", "w", "fp", "sp", "wp")}
    rows, rowsp = np.zeros(3 * n), np.zeros(3 * n)
    if L.orc_np_tables(n, h, ncl1, ncln, *(_p(a[k]) for k in ("f", "s", "w", "fp", "sp", "wp")),
                       _p(rows), _p(rowsp)):
        raise ValueError("oracle: bad non-periodic table request")
    a["rows"] = rows.reshape(3, n)
    a["rowsp"] = rowsp.reshape(3, n)
    return a


def np_consts(h: float):
    """alfai afi bfi alfa1 af1 bf1 cf1 alfa2 af2 alfan afn bfn cfn alfam afm."""
    k = np.zeros(15)
    lib(False).orc_np_consts(h, _p(k))
    return k


def der_np(u: np.ndarray, axis: str, L: float, ncl1: int, ncln: int, npaire: int) -> np.ndarray:
    """der{x,y,z}_{ncl1}{ncln} of src/x3d_derive.f90:156-275 (and its y, z
    copies) on a (nz, ny, nx) field, grid spacing L / (n - 1), with the
    npaire instance's tables."""
    u = np.ascontiguousarray(u, dtype=np.float64)
    nz, ny, nx = u.shape
    d = "xyz".index(axis)
    n = (nx, ny, nz)[d]
    h = L / (n - 1)
    tb = np_tables(n, h, ncl1, ncln)
    f, s, w = (tb["fp"], tb["sp"], tb["wp"]) if npaire == 1 else (tb["f"], tb["s"], tb["w"])
    t = np.zeros_like(u)
    if lib(False).orc_der_np(_p(t), _p(u), d, nx, ny, nz, h, ncl1, ncln, npaire, _p(f), _p(s), _p(w)):
        raise ValueError("oracle: bad non-periodic derivative request")
    return t


class Oracle:
    """One periodic grid: coefficients, operators, Poisson and the step."""

    def __init__(self, nx: int, ny: int, nz: int, xlx: float = 1.0, yly: float = 1.0,
                 zlz: float = 1.0, fast: bool = False):
        self.nx, self.ny, self.nz = nx, ny, nz
        self.L = (xlx, yly, zlz)
        self._L = lib(fast)
        self._c = self._L.orc_new(nx, ny, nz, xlx, yly, zlz)
        if not self._c:
            raise ValueError(f"oracle: unsupported grid {nx}x{ny}x{nz}")

    def __del__(self):
        if getattr(self, "_c", None):
            self._L.orc_delete(self._c)
            self._c = None

    @property
    def shape(self):
        return (self.nz, self.ny, self.nx)

    def zeros(self):
        return np.zeros(self.shape)

    def op(self, name: str):
        """(f, s, w, periodic, alfa, npaire) of one operator instance."""
        n = {"x": self.nx, "y": self.ny, "z": self.nz}[_dir_letter(name)]
        f, s, w, p = (np.zeros(n) for _ in range(4))
        alfa = C.c_double(); nn = C.c_int(); npaire = C.c_int()
        rc = self._L.orc_get_op(self._c, OP[name], _p(f), _p(s), _p(w), _p(p),
                              C.byref(alfa), C.byref(nn), C.byref(npaire))
        assert rc == 0 and nn.value == n
        return f, s, w, p, alfa.value, npaire.value

    def consts(self):
        out = np.zeros(33)
        self._L.orc_get_consts(self._c, _p(out))
        names = ["alfai", "afi", "bfi", "alcai", "aci", "bci", "ailcai", "aici", "bici", "cici", "dici"]
        return [dict(zip(names, out[11 * d:11 * d + 11])) for d in range(3)]

    def kxyz(self):
        """(Re, Im) of kxyz on the (nx, ny, nz/2+1) spectral grid."""
        re = np.zeros((self.nz // 2 + 1, self.ny, self.nx)); im = np.zeros_like(re)
        self._L.orc_get_kxyz(self._c, _p(re), _p(im))
        return re, im

    def ab(self):
        """ax, bx, ay, by, az, bz of abxyz."""
        out = np.zeros(2 * (self.nx + self.ny + self.nz))
        self._L.orc_get_ab(self._c, _p(out))
        sizes = [self.nx, self.nx, self.ny, self.ny, self.nz, self.nz]
        parts = np.split(out, np.cumsum(sizes)[:-1])
        return dict(zip(["ax", "bx", "ay", "by", "az", "bz"], parts))

    def set_stretching(self, istret: int, ppy=None, ppyi=None) -> None:
        """istret != 0: y-derivatives scaled by ppy(j) (dery, derypv) or ppyi(j) (deryvp)."""
        if istret:
            a = np.ascontiguousarray(ppy, dtype=np.float64)
            b = np.ascontiguousarray(ppyi, dtype=np.float64)
            assert a.size == self.ny and b.size == self.ny
            rc = self._L.orc_set_stretching(self._c, int(istret), _p(a), _p(b))
        else:
            rc = self._L.orc_set_stretching(self._c, 0, None, None)
        assert rc == 0

    def apply(self, name: str, u: np.ndarray, kind: str | None = None) -> np.ndarray:
        """Apply one operator (der?/der?vp/der?pv/inter?vp/inter?pv) along its axis."""
        d = "xyz".index(_dir_letter(name))
        if kind is None:
            kind = _default_kind(name)
        u = np.ascontiguousarray(u, dtype=np.float64)
        t = np.empty_like(u)
        rc = self._L.orc_apply(self._c, _p(t), _p(u), OP[name], d, KIND[kind], self.nx, self.ny, self.nz)
        if rc:
            raise ValueError(f"orc_apply rc={rc}")
        return t

    def thomas(self, name: str, tt: np.ndarray) -> np.ndarray:
        d = "xyz".index(_dir_letter(name))
        tt = np.array(tt, dtype=np.float64, order="C", copy=True)
        rc = self._L.orc_thomas(self._c, _p(tt), OP[name], d, self.nx, self.ny, self.nz)
        assert rc == 0
        return tt

    def thomas12(self, name: str, tt: np.ndarray) -> np.ndarray:
        """The plain sweeps with an operator's f, s, w (xthomas_12 and its y, z
        forms), without the periodic correction."""
        d = _dir_letter(name)
        f, s, w, _, _, _ = self.op(name)
        tt = np.array(tt, dtype=np.float64, order="C", copy=True)
        getattr(self._L, f"orc_{d}thomas_12")(_p(tt), _p(f), _p(s), _p(w), self.nx, self.ny, self.nz)
        return tt

    def r2c(self, rhs):
        cw = np.zeros((self.nz // 2 + 1, self.ny, self.nx), dtype=np.complex128)
        rhs = np.ascontiguousarray(rhs, dtype=np.float64)
        self._L.orc_fft_r2c_z(self._c, _p(rhs), cw.ctypes.data_as(C.POINTER(C.c_double)))
        return cw

    def c2r(self, cw):
        cw = np.array(cw, dtype=np.complex128, order="C", copy=True)
        out = self.zeros()
        self._L.orc_fft_c2r_z(self._c, cw.ctypes.data_as(C.POINTER(C.c_double)), _p(out))
        return out

    def spectral(self, cw):
        cw = np.array(cw, dtype=np.complex128, order="C", copy=True)
        self._L.orc_poisson_spectral(self._c, cw.ctypes.data_as(C.POINTER(C.c_double)))
        return cw

    def poisson(self, rhs):
        out = np.array(rhs, dtype=np.float64, order="C", copy=True)
        self._L.orc_poisson(self._c, _p(out))
        return out

    def momentum_rhs(self, ux, uy, uz):
        du = [self.zeros() for _ in range(3)]
        self._L.orc_momentum_rhs(self._c, *(_p(a) for a in du), *(_p(np.ascontiguousarray(a)) for a in (ux, uy, uz)))
        return du

    def divergence(self, ux, uy, uz):
        pp = self.zeros()
        self._L.orc_divergence(self._c, _p(pp), *(_p(np.ascontiguousarray(a)) for a in (ux, uy, uz)))
        return pp

    def gradp(self, pp):
        g = [self.zeros() for _ in range(3)]
        self._L.orc_gradp(self._c, *(_p(a) for a in g), _p(np.ascontiguousarray(pp)))
        return g

    def solve_poisson(self, ux, uy, uz):
        pp = self.zeros(); g = [self.zeros() for _ in range(3)]
        self._L.orc_solve_poisson(self._c, _p(pp), *(_p(a) for a in g),
                                *(_p(np.ascontiguousarray(a)) for a in (ux, uy, uz)))
        return pp, g

    def step(self, ux, uy, uz, gdt: float = 0.0):
        """One Euler step; returns (u, du, pp3, grad p) as new arrays."""
        u = [np.array(a, dtype=np.float64, order="C", copy=True) for a in (ux, uy, uz)]
        du = [self.zeros() for _ in range(3)]
        pp = self.zeros(); g = [self.zeros() for _ in range(3)]
        self._L.orc_step(self._c, gdt, *(_p(a) for a in u), *(_p(a) for a in du), _p(pp),
                       *(_p(a) for a in g))
        return u, du, pp, g

    def step_inplace(self, u, du, pp, g, gdt: float = 0.0) -> None:
        """One Euler step on caller-owned arrays (u updated; du, pp, g overwritten)."""
        self._L.orc_step(self._c, gdt, *(_p(a) for a in u), *(_p(a) for a in du), _p(pp),
                         *(_p(a) for a in g))

    def fields(self, kind: int = 1, seed: int = 1234, amp: float = 1e-3):
        u = [self.zeros() for _ in range(3)]
        self._L.orc_init_fields(self._c, kind, seed, amp, *(_p(a) for a in u))
        return u

    def test_err(self, d, which: int) -> float:
        return self._L.orc_test_err(self._c, _p(np.ascontiguousarray(d)), which)

    def div_stats(self, pp):
        mx = C.c_double(); mean = C.c_double()
        self._L.orc_div_stats(self._c, _p(np.ascontiguousarray(pp)), C.byref(mx), C.byref(mean))
        return mx.value, mean.value


def _dir_letter(name: str) -> str:
    # derx, derxp, derxvp, intxvp ... -> the axis letter after the 3-letter stem
    return name[3]


def _default_kind(name: str) -> str:
    if name in OPS2:
        return "dd"
    if name.startswith("der"):
        return "dvp" if name.endswith("vp") else "dpv" if name.endswith("pv") else "d"
    return "ivp" if name.endswith("vp") else "ipv"


def prepare(b, c, f):
    n = len(c)
    s = np.zeros(n); w = np.zeros(n)
    b, c, f = (np.ascontiguousarray(a, dtype=np.float64) for a in (b, c, f))
    lib().orc_prepare(_p(b), _p(c), _p(f), _p(s), _p(w), n)
    return s, w


def set_threads(n: int, fast: bool = False) -> None:
    """OpenMP threads of the oracle's line loops (results do not depend on it)."""
    lib(fast).orc_set_threads(int(n))


def max_threads(fast: bool = False) -> int:
    return lib(fast).orc_max_threads()


def uniform(seed: int, comp: int, idx: int) -> float:
    return lib().orc_uniform(seed, comp, idx)


def splitmix64(x: int) -> int:
    return lib().orc_splitmix64(x)
