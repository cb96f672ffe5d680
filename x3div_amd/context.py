"""Host-side mirror of x3div's hot-path interfaces over libx3dhip.

One :class:`X3D` object is one periodic grid on the current HIP device: the
reference's ``init_xcompact3d`` numerics set-up (operators, derivative
pointers, Poisson tables).  Its methods keep the reference's subroutine names
and argument meaning (output first, then input, then the operator instance)
and operate on device tensors of shape ``(nz, ny, nx)`` -- the Fortran
``(nx, ny, nz)`` layout with x fastest.  torch provides the device memory and
the stream; every computation runs in the library's HIP kernels.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import check

# operator instances, src/x3d_operator_1d.f90:33-47 (x3d_op_<name>)
OPS = ["derx", "derxp", "dery", "deryp", "derz", "derzp",
       "derxvp", "deryvp", "derzvp", "derxpv", "derypv", "derzpv",
       "intxvp", "intyvp", "intzvp", "intxpv", "intypv", "intzpv"]
# second derivatives (derxx/deryy/derzz), after the reference's 18 instances
OPS2 = ["derxx", "deryy", "derzz"]
OP = {name: i for i, name in enumerate(OPS + OPS2)}

FIELD_REFERENCE_VEL, FIELD_TG_NOISE, FIELD_FOURIER8 = 0, 1, 2


def _op_id(op) -> int:
    if isinstance(op, str):
        return OP[op.removeprefix("x3d_op_")]
    return int(op)


def _ptr(t, shape=None, dtype=None) -> int:
    import torch
    dtype = dtype or torch.float64
    if not isinstance(t, torch.Tensor):
        raise TypeError("expected a torch tensor on the GPU")
    if not t.is_cuda:
        raise ValueError("x3div_amd operates on device tensors only (no CPU fallback)")
    if t.dtype != dtype:
        raise TypeError(f"expected {dtype}, got {t.dtype}")
    if not t.is_contiguous():
        raise ValueError("tensor must be contiguous")
    if shape is not None and tuple(t.shape) != tuple(shape):
        raise ValueError(f"shape {tuple(t.shape)} != {tuple(shape)}")
    return t.data_ptr()


def _host_vec(v, n: int, dtype=np.float64) -> np.ndarray:
    """A contiguous host copy of a length-n coefficient vector (the caller's
    array must not be freed while the library reads it)."""
    a = np.ascontiguousarray(np.asarray(v, dtype=dtype).reshape(-1))
    if a.size != n:
        raise ValueError(f"vector of length {a.size}, expected {n}")
    return a


class X3D:
    """A periodic nx*ny*nz grid of lengths (xlx, yly, zlz) on the GPU."""

    def __init__(self, nx: int, ny: int, nz: int, xlx: float = 1.0, yly: float = 1.0,
                 zlz: float = 1.0, stream=None, ncl1=(0, 0, 0), ncln=(0, 0, 0)):
        """ncl1, ncln: boundary codes per direction as the reference's
        nclx1/nclxn...: 0 periodic, 1 free-slip, 2 Dirichlet."""
        import torch
        if not torch.cuda.is_available():
            raise RuntimeError("x3div_amd needs a HIP device (no CPU fallback)")
        self.lib = _lib.load()
        self.nx, self.ny, self.nz = nx, ny, nz
        self.L = (xlx, yly, zlz)
        self.ncl1, self.ncln = tuple(ncl1), tuple(ncln)
        h = C.c_void_p()
        b1 = (C.c_int * 3)(*self.ncl1)
        bn = (C.c_int * 3)(*self.ncln)
        check("x3d_create_bc", self.lib.x3d_create_bc(nx, ny, nz, xlx, yly, zlz, b1, bn, C.byref(h)))
        self._h = h
        self.set_stream(stream if stream is not None else torch.cuda.current_stream())

    # ---- lifetime / plumbing ----------------------------------------------
    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            self.lib.x3d_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_stream(self, stream):
        self.stream = stream
        check("x3d_set_stream", self.lib.x3d_set_stream(self._h, C.c_void_p(stream.cuda_stream)))

    def sync(self):
        check("x3d_sync", self.lib.x3d_sync(self._h))

    @property
    def shape(self):
        return (self.nz, self.ny, self.nx)

    def field(self, fill: float | None = 0.0):
        import torch
        t = torch.empty(self.shape, dtype=torch.float64, device="cuda")
        if fill is not None:
            t.fill_(fill)
        return t

    def _f(self, t):
        return _ptr(t, self.shape)

    def field32(self, fill: float | None = 0.0):
        """An FP32 field for the precision form."""
        import torch
        t = torch.empty(self.shape, dtype=torch.float32, device="cuda")
        if fill is not None:
            t.fill_(fill)
        return t

    def _f32(self, t):
        import torch
        return _ptr(t, self.shape, torch.float32)

    # ---- operator tables ----------------------------------------------------
    def operator(self, op):
        """(f, s, w, periodic, alfa, npaire) of one x3doperator1d instance."""
        o = _op_id(op)
        n = {0: self.nx, 1: self.ny, 2: self.nz}[_axis((OPS + OPS2)[o])]
        arrs = [np.zeros(n) for _ in range(4)]
        alfa = C.c_double(); nn = C.c_int(); npaire = C.c_int()
        hp = C.POINTER(C.c_double)
        check("x3d_operator_get", self.lib.x3d_operator_get(
            self._h, o, *(a.ctypes.data_as(hp) for a in arrs), C.byref(alfa), C.byref(nn),
            C.byref(npaire)))
        return (*arrs, alfa.value, npaire.value)

    def set_operator(self, op, f, s, w, periodic, alfa):
        hp = C.POINTER(C.c_double)
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (f, s, w, periodic)]
        check("x3d_operator_set", self.lib.x3d_operator_set(
            self._h, _op_id(op), *(a.ctypes.data_as(hp) for a in arrs), float(alfa)))

    def scheme_consts(self):
        out = np.zeros(33)
        check("x3d_scheme_consts", self.lib.x3d_scheme_consts(
            self._h, out.ctypes.data_as(C.POINTER(C.c_double))))
        names = ["alfai", "afi", "bfi", "alcai", "aci", "bci", "ailcai", "aici", "bici", "cici", "dici"]
        return [dict(zip(names, out[11 * d:11 * d + 11])) for d in range(3)]

    # ---- Thomas (src/thomas.f90) ------------------------------------------------
    def xthomas(self, tt, op):
        check("x3d_xthomas", self.lib.x3d_xthomas(self._h, self._f(tt), _op_id(op), self.nx, self.ny, self.nz))

    def ythomas(self, tt, op):
        check("x3d_ythomas", self.lib.x3d_ythomas(self._h, self._f(tt), _op_id(op), self.nx, self.ny, self.nz))

    def zthomas(self, tt, op):
        check("x3d_zthomas", self.lib.x3d_zthomas(self._h, self._f(tt), _op_id(op), self.nx, self.ny, self.nz))

    def _thomas_tab(self, fn, tt, f, s, w, periodic, alfa):
        hp = C.POINTER(C.c_double)
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (f, s, w)]
        per = None if periodic is None else np.ascontiguousarray(periodic, dtype=np.float64)
        check(fn, getattr(self.lib, fn)(self._h, self._f(tt), *(a.ctypes.data_as(hp) for a in arrs),
                                        None if per is None else per.ctypes.data_as(hp), float(alfa),
                                        self.nx, self.ny, self.nz))

    def xthomas_tab(self, tt, f, s, w, periodic=None, alfa=0.0):
        """xthomas with the caller's own vectors: xthomas_0 (periodic given)
        or xthomas_12 (periodic None), src/thomas.f90:34-84."""
        self._thomas_tab("x3d_xthomas_tab", tt, f, s, w, periodic, alfa)

    def ythomas_tab(self, tt, f, s, w, periodic=None, alfa=0.0):
        self._thomas_tab("x3d_ythomas_tab", tt, f, s, w, periodic, alfa)

    def zthomas_tab(self, tt, f, s, w, periodic=None, alfa=0.0):
        self._thomas_tab("x3d_zthomas_tab", tt, f, s, w, periodic, alfa)

    # ---- derivatives (src/x3d_derive.f90) ----------------------------------------
    def derx(self, tx, ux, op="derx"):
        check("x3d_derx", self.lib.x3d_derx(self._h, self._f(tx), self._f(ux), _op_id(op), self.nx, self.ny, self.nz))

    # ---- y-stretching (istret, ppy, ppyi: src/module_param.f90:70-71,104) ----------
    def set_stretching(self, istret: int, ppy=None, ppyi=None):
        """istret != 0: y-derivatives (dery, deryvp, derypv and their uses in
        the step) are multiplied by ppy(j), deryvp by ppyi(j); host vectors of
        ny values.  istret = 0 turns it off."""
        if istret:
            self._sy = [_host_vec(ppy, self.ny), _host_vec(ppyi, self.ny)]
            a, b = (v.ctypes.data for v in self._sy)
        else:
            a = b = None
        check("x3d_set_stretching", self.lib.x3d_set_stretching(self._h, int(istret), a, b))

    def dery(self, ty, uy, op="dery", ppy=None):
        v = None if ppy is None else _host_vec(ppy, self.ny)
        check("x3d_dery", self.lib.x3d_dery(self._h, self._f(ty), self._f(uy), _op_id(op),
                                            None if v is None else v.ctypes.data, self.nx, self.ny, self.nz))

    def derz(self, tz, uz, op="derz"):
        check("x3d_derz", self.lib.x3d_derz(self._h, self._f(tz), self._f(uz), _op_id(op), self.nx, self.ny, self.nz))

    # ---- second derivatives (derxx/deryy/derzz, src/x3d_derive.f90:54-59) ----------
    def derxx(self, tx, ux, op="derxx"):
        check("x3d_derxx", self.lib.x3d_derxx(self._h, self._f(tx), self._f(ux), _op_id(op), self.nx, self.ny, self.nz))

    def deryy(self, ty, uy, op="deryy"):
        check("x3d_deryy", self.lib.x3d_deryy(self._h, self._f(ty), self._f(uy), _op_id(op), self.nx, self.ny, self.nz))

    def derzz(self, tz, uz, op="derzz"):
        check("x3d_derzz", self.lib.x3d_derzz(self._h, self._f(tz), self._f(uz), _op_id(op), self.nx, self.ny, self.nz))

    # ---- staggered (src/x3d_staggered.f90) -----------------------------------------
    def derxvp(self, tx, ux, op="derxvp"):
        check("x3d_derxvp", self.lib.x3d_derxvp(self._h, self._f(tx), self._f(ux), _op_id(op), self.nx, self.nx, self.ny, self.nz))

    def interxvp(self, tx, ux, op="intxvp"):
        check("x3d_interxvp", self.lib.x3d_interxvp(self._h, self._f(tx), self._f(ux), _op_id(op), self.nx, self.nx, self.ny, self.nz))

    def derxpv(self, tx, ux, op="derxpv"):
        check("x3d_derxpv", self.lib.x3d_derxpv(self._h, self._f(tx), self._f(ux), _op_id(op), self.nx, self.nx, self.ny, self.nz))

    def interxpv(self, tx, ux, op="intxpv"):
        check("x3d_interxpv", self.lib.x3d_interxpv(self._h, self._f(tx), self._f(ux), _op_id(op), self.nx, self.nx, self.ny, self.nz))

    def interyvp(self, ty, uy, op="intyvp"):
        check("x3d_interyvp", self.lib.x3d_interyvp(self._h, self._f(ty), self._f(uy), _op_id(op), self.nx, self.ny, self.ny, self.nz))

    def deryvp(self, ty, uy, op="deryvp", ppyi=None):
        v = None if ppyi is None else _host_vec(ppyi, self.ny)
        check("x3d_deryvp", self.lib.x3d_deryvp(self._h, self._f(ty), self._f(uy), _op_id(op),
                                                None if v is None else v.ctypes.data, self.nx, self.ny, self.ny, self.nz))

    def interypv(self, ty, uy, op="intypv"):
        check("x3d_interypv", self.lib.x3d_interypv(self._h, self._f(ty), self._f(uy), _op_id(op), self.nx, self.ny, self.ny, self.nz))

    def derypv(self, ty, uy, op="derypv", ppy=None):
        v = None if ppy is None else _host_vec(ppy, self.ny)
        check("x3d_derypv", self.lib.x3d_derypv(self._h, self._f(ty), self._f(uy), _op_id(op),
                                                None if v is None else v.ctypes.data, self.nx, self.ny, self.ny, self.nz))

    def derzvp(self, tz, uz, op="derzvp"):
        check("x3d_derzvp", self.lib.x3d_derzvp(self._h, self._f(tz), self._f(uz), _op_id(op), self.nx, self.ny, self.nz, self.nz))

    def interzvp(self, tz, uz, op="intzvp"):
        check("x3d_interzvp", self.lib.x3d_interzvp(self._h, self._f(tz), self._f(uz), _op_id(op), self.nx, self.ny, self.nz, self.nz))

    def derzpv(self, tz, uz, op="derzpv"):
        check("x3d_derzpv", self.lib.x3d_derzpv(self._h, self._f(tz), self._f(uz), _op_id(op), self.nx, self.ny, self.nz, self.nz))

    def interzpv(self, tz, uz, op="intzpv"):
        check("x3d_interzpv", self.lib.x3d_interzpv(self._h, self._f(tz), self._f(uz), _op_id(op), self.nx, self.ny, self.nz, self.nz))

    def apply(self, name: str, out, inp):
        """Apply an operator by its instance name with the reference's routine."""
        routine = {
            "derx": self.derx, "derxp": self.derx, "dery": self.dery, "deryp": self.dery,
            "derz": self.derz, "derzp": self.derz, "derxvp": self.derxvp, "deryvp": self.deryvp,
            "derzvp": self.derzvp, "derxpv": self.derxpv, "derypv": self.derypv, "derzpv": self.derzpv,
            "intxvp": self.interxvp, "intyvp": self.interyvp, "intzvp": self.interzvp,
            "intxpv": self.interxpv, "intypv": self.interypv, "intzpv": self.interzpv,
            "derxx": self.derxx, "deryy": self.deryy, "derzz": self.derzz,
        }[name]
        routine(out, inp, name)

    # ---- transposes (src/x3d_transpose.f90) ------------------------------------------
    def _transpose(self, fn, a, b):
        eb = a.element_size()
        if b.element_size() != eb or a.numel() != b.numel() or not (a.is_cuda and b.is_cuda):
            raise ValueError("transpose: mismatched arrays")
        check(fn, getattr(self.lib, fn)(self._h, a.data_ptr(), b.data_ptr(), eb, self.nx, self.ny,
                                        a.numel() // (self.nx * self.ny)))

    def transpose_x_to_y(self, a, b):
        self._transpose("x3d_transpose_x_to_y", a, b)

    def transpose_y_to_z(self, a, b):
        self._transpose("x3d_transpose_y_to_z", a, b)

    def transpose_z_to_y(self, a, b):
        self._transpose("x3d_transpose_z_to_y", a, b)

    def transpose_y_to_x(self, a, b):
        self._transpose("x3d_transpose_y_to_x", a, b)

    # ---- Poisson (src/poisson.f90) -----------------------------------------------------
    def poisson(self, rhs):
        check("x3d_poisson", self.lib.x3d_poisson(self._h, self._f(rhs)))

    # ---- captured step ------------------------------------------------------------
    def graph_capture_step(self, ux, uy, uz, dux, duy, duz, pp3, px, py, pz, gdt: float = 0.0):
        """Record one step on these fields into the context's HIP graph."""
        check("x3d_graph_capture_step", self.lib.x3d_graph_capture_step(
            self._h, float(gdt), *(self._f(t) for t in (ux, uy, uz, dux, duy, duz, pp3, px, py, pz))))

    def graph_capture_step_f32(self, ux, uy, uz, dux, duy, duz, pp3, px, py, pz, gdt: float = 0.0):
        check("x3d_graph_capture_step_f32", self.lib.x3d_graph_capture_step_f32(
            self._h, float(gdt), *(self._f32(t) for t in (ux, uy, uz, dux, duy, duz, pp3, px, py, pz))))

    def graph_launch(self):
        """Replay the captured step on the context's stream."""
        check("x3d_graph_launch", self.lib.x3d_graph_launch(self._h))

    def graph_begin(self):
        """Record this context's following calls into its graph (not run)."""
        check("x3d_graph_begin", self.lib.x3d_graph_begin(self._h))

    def graph_end(self):
        check("x3d_graph_end", self.lib.x3d_graph_end(self._h))

    def fill(self, t, value: float):
        """Set every value of an FP64 device field (a kernel, recordable in a graph)."""
        check("x3d_fill", self.lib.x3d_fill(self._h, _ptr(t), t.numel(), float(value)))

    # ---- precision form (FP32 fields) ------------------------------------------------
    def derx_f32(self, tx, ux, op="derx"):
        check("x3d_derx_f32", self.lib.x3d_derx_f32(self._h, self._f32(tx), self._f32(ux), _op_id(op),
                                                    self.nx, self.ny, self.nz))

    def dery_f32(self, ty, uy, op="dery", ppy=None):
        v = None if ppy is None else _host_vec(ppy, self.ny, np.float32)
        check("x3d_dery_f32", self.lib.x3d_dery_f32(self._h, self._f32(ty), self._f32(uy), _op_id(op),
                                                    None if v is None else v.ctypes.data, self.nx, self.ny, self.nz))

    def derz_f32(self, tz, uz, op="derz"):
        check("x3d_derz_f32", self.lib.x3d_derz_f32(self._h, self._f32(tz), self._f32(uz), _op_id(op),
                                                    self.nx, self.ny, self.nz))

    def apply_f32(self, name: str, out, inp):
        """Any of the 18 operators, by instance name, along its own direction, in FP32."""
        check("x3d_apply_f32", self.lib.x3d_apply_f32(self._h, _op_id(name), self._f32(out), self._f32(inp)))

    def apply_circ_mfma(self, name: str, out, inp):
        """The operator as a dense circulant on the matrix cores (split bf16, FP32 sums)."""
        check("x3d_apply_circ_mfma", self.lib.x3d_apply_circ_mfma(self._h, _op_id(name), self._f32(out),
                                                                  self._f32(inp)))

    def poisson_f32(self, rhs):
        check("x3d_poisson_f32", self.lib.x3d_poisson_f32(self._h, self._f32(rhs)))

    def step_f32(self, ux, uy, uz, dux, duy, duz, pp3, px, py, pz, gdt: float = 0.0):
        """The step of the precision form on FP32 fields."""
        check("x3d_step_f32", self.lib.x3d_step_f32(
            self._h, float(gdt), *(self._f32(t) for t in (ux, uy, uz, dux, duy, duz, pp3, px, py, pz))))

    def step_timed_f32(self, ux, uy, uz, dux, duy, duz, pp3, px, py, pz, gdt: float = 0.0):
        """step_timed for the precision form; bytes per point are half the FP64
        figures, with those of transeq_x and transeq_z exchanged (FP32 single
        grids run transeq x first, so its x pass moves 6 fields, its z pass 12)."""
        from ._lib import X3D_NPASS
        ms = (C.c_float * X3D_NPASS)()
        check("x3d_step_timed_f32", self.lib.x3d_step_timed_f32(
            self._h, float(gdt), *(self._f32(t) for t in (ux, uy, uz, dux, duy, duz, pp3, px, py, pz)), ms))
        out = {self.lib.x3d_pass_name(i).decode(): [ms[i], self.lib.x3d_pass_bytes_per_point(i) / 2]
               for i in range(X3D_NPASS)}
        out["transeq_x"][1], out["transeq_z"][1] = out["transeq_z"][1], out["transeq_x"][1]
        return {k: tuple(v) for k, v in out.items()}

    # ---- BFP32 form (block-floating-point int32 fields, FP64 line solves) ----------
    def _bfp(self, t):
        import torch
        return _ptr(t, self.shape, torch.int32)

    def bfp_empty(self):
        """An uninitialised BFP32 field (int32 words) of the grid's shape."""
        import torch
        return torch.empty(self.shape, dtype=torch.int32, device="cuda")

    def bfp_encode(self, q, u):
        """q <- the BFP32 words of the FP64 field u (nx a multiple of 16)."""
        check("x3d_bfp_encode", self.lib.x3d_bfp_encode(self._h, self._bfp(q), _ptr(u, self.shape)))

    def bfp_decode(self, u, q):
        """u <- the FP64 values of the BFP32 field q (exact)."""
        check("x3d_bfp_decode", self.lib.x3d_bfp_decode(self._h, _ptr(u, self.shape), self._bfp(q)))

    def apply_bfp(self, name: str, out, inp):
        """Any operator, by instance name, along its own direction, on BFP32 fields (FP64 arithmetic)."""
        check("x3d_apply_bfp", self.lib.x3d_apply_bfp(self._h, _op_id(name), self._bfp(out), self._bfp(inp)))

    def step_bfp(self, ux, uy, uz, dux, duy, duz, pp3, px, py, pz, gdt: float = 0.0):
        """The step on BFP32 fields."""
        check("x3d_step_bfp", self.lib.x3d_step_bfp(
            self._h, float(gdt), *(self._bfp(t) for t in (ux, uy, uz, dux, duy, duz, pp3, px, py, pz))))

    def step_timed_bfp(self, ux, uy, uz, dux, duy, duz, pp3, px, py, pz, gdt: float = 0.0):
        """step_timed for the BFP32 form; bytes per point are half the FP64 figures
        (the Poisson figure leaves out its FP64 decode and encode)."""
        from ._lib import X3D_NPASS
        ms = (C.c_float * X3D_NPASS)()
        check("x3d_step_timed_bfp", self.lib.x3d_step_timed_bfp(
            self._h, float(gdt), *(self._bfp(t) for t in (ux, uy, uz, dux, duy, duz, pp3, px, py, pz)), ms))
        return {self.lib.x3d_pass_name(i).decode(): (ms[i], self.lib.x3d_pass_bytes_per_point(i) / 2)
                for i in range(X3D_NPASS)}

    def poisson_tables(self):
        hp = C.POINTER(C.c_double)
        a = [np.zeros(n) for n in (self.nx, self.nx, self.ny, self.ny, self.nz, self.nz)]
        k = np.zeros((self.nz // 2 + 1, self.ny, self.nx))
        check("x3d_poisson_tables", self.lib.x3d_poisson_tables(
            self._h, *(x.ctypes.data_as(hp) for x in a), k.ctypes.data_as(hp)))
        names = ["ax", "bx", "ay", "by", "az", "bz"]
        out = dict(zip(names, a))
        out["kxyz"] = k
        return out

    # ---- timestep (src/transeq.f90, navier.f90, time_integrators.f90) -----------------
    def calculate_transeq_rhs(self, dux, duy, duz, ux, uy, uz):
        check("x3d_momentum_rhs", self.lib.x3d_momentum_rhs(
            self._h, *(self._f(t) for t in (dux, duy, duz, ux, uy, uz))))

    momentum_rhs_eq = calculate_transeq_rhs

    def int_time(self, ux, uy, uz, dux, duy, duz, gdt: float = 0.0):
        check("x3d_int_time", self.lib.x3d_int_time(
            self._h, float(gdt), *(self._f(t) for t in (ux, uy, uz, dux, duy, duz))))

    def divergence(self, pp3, ux, uy, uz):
        check("x3d_divergence", self.lib.x3d_divergence(self._h, *(self._f(t) for t in (pp3, ux, uy, uz))))

    def gradp(self, px, py, pz, pp3):
        check("x3d_gradp", self.lib.x3d_gradp(self._h, *(self._f(t) for t in (px, py, pz, pp3))))

    def cor_vel(self, ux, uy, uz, px, py, pz):
        check("x3d_cor_vel", self.lib.x3d_cor_vel(self._h, *(self._f(t) for t in (ux, uy, uz, px, py, pz))))

    def solve_poisson(self, pp3, px, py, pz, ux, uy, uz):
        check("x3d_solve_poisson", self.lib.x3d_solve_poisson(
            self._h, *(self._f(t) for t in (pp3, px, py, pz, ux, uy, uz))))

    def step(self, ux, uy, uz, dux, duy, duz, pp3, px, py, pz, gdt: float = 0.0):
        """One iteration of the reference's step loop (src/xcompact3d.f90:80-95)."""
        check("x3d_step", self.lib.x3d_step(
            self._h, float(gdt), *(self._f(t) for t in (ux, uy, uz, dux, duy, duz, pp3, px, py, pz))))

    def step_noalloc(self, ux, uy, uz, dux, duy, duz, pp3, px, py, pz, gdt: float = 0.0):
        """step() with its scratch in the library's static device array: no
        device allocation (grids of at most X3D_NOALLOC_MAX_POINTS points)."""
        check("x3d_step_noalloc", self.lib.x3d_step_noalloc(
            self._h, float(gdt), *(self._f(t) for t in (ux, uy, uz, dux, duy, duz, pp3, px, py, pz))))

    def step_timed(self, ux, uy, uz, dux, duy, duz, pp3, px, py, pz, gdt: float = 0.0):
        """One step with a HIP event after each pass; waits for it and returns
        {pass name: (milliseconds, algorithmic HBM bytes per grid point)}."""
        from ._lib import X3D_NPASS
        ms = (C.c_float * X3D_NPASS)()
        check("x3d_step_timed", self.lib.x3d_step_timed(
            self._h, float(gdt), *(self._f(t) for t in (ux, uy, uz, dux, duy, duz, pp3, px, py, pz)), ms))
        return {self.lib.x3d_pass_name(i).decode(): (ms[i], self.lib.x3d_pass_bytes_per_point(i))
                for i in range(X3D_NPASS)}

    # ---- test mode / fields (src/mom.f90, src/navier.f90:213-231) --------------------
    def test_err(self, d, which: int) -> float:
        e = C.c_double()
        check("x3d_test_err", self.lib.x3d_test_err(self._h, self._f(d), which, C.byref(e)))
        return e.value

    def test_du(self, du):
        return self.test_err(du, 0)

    def test_dv(self, dv):
        return self.test_err(dv, 1)

    def test_dw(self, dw):
        return self.test_err(dw, 2)

    def div_stats(self, pp3):
        mx = C.c_double(); mean = C.c_double()
        check("x3d_div_stats", self.lib.x3d_div_stats(self._h, self._f(pp3), C.byref(mx), C.byref(mean)))
        return mx.value, mean.value

    def init_fields(self, ux, uy, uz, kind: int = FIELD_TG_NOISE, seed: int = 1234, amp: float = 1e-3):
        check("x3d_init_fields", self.lib.x3d_init_fields(
            self._h, kind, seed, amp, *(self._f(t) for t in (ux, uy, uz))))


def _axis(name: str) -> int:
    return "xyz".index(name[3])


class X3DBatch:
    """Many independent periodic grids stepped together (the batching form).

    dims: list of (nx, ny, nz); lengths: list of (xlx, yly, zlz) or None for
    2*pi; precision: 8 (FP64) or 4 (FP32).  The library owns the fields; they
    are read and written through host copies (get / put)."""

    FIELDS = ["ux", "uy", "uz", "dux", "duy", "duz", "pp3", "px", "py", "pz"]

    def __init__(self, dims, lengths=None, precision: int = 8, stream=None):
        import torch
        if not torch.cuda.is_available():
            raise RuntimeError("x3div_amd needs a HIP device (no CPU fallback)")
        self.lib = _lib.load()
        self.dims = [tuple(int(v) for v in d) for d in dims]
        self.precision = precision
        flat = (C.c_int * (3 * len(self.dims)))(*[v for d in self.dims for v in d])
        lp = None
        if lengths is not None:
            lp = (C.c_double * (3 * len(self.dims)))(*[float(v) for d in lengths for v in d])
        h = C.c_void_p()
        check("x3d_batch_create", self.lib.x3d_batch_create(len(self.dims), flat, lp, precision, C.byref(h)))
        self._h = h
        st = stream if stream is not None else torch.cuda.current_stream()
        check("x3d_batch_set_stream", self.lib.x3d_batch_set_stream(self._h, C.c_void_p(st.cuda_stream)))
        self.stream = st

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            self.lib.x3d_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def total_points(self) -> int:
        t = C.c_longlong()
        check("x3d_batch_sizes", self.lib.x3d_batch_sizes(self._h, None, C.byref(t)))
        return t.value

    def _ptr(self, grid: int, name: str):
        p = C.c_void_p()
        check("x3d_batch_field", self.lib.x3d_batch_field(self._h, grid, self.FIELDS.index(name), C.byref(p)))
        return p

    def _ctx(self, grid: int):
        c = C.c_void_p()
        check("x3d_batch_context", self.lib.x3d_batch_context(self._h, grid, C.byref(c)))
        return c

    def _dtype(self):
        return np.float64 if self.precision == 8 else np.float32

    def get(self, grid: int, name: str) -> np.ndarray:
        nx, ny, nz = self.dims[grid]
        out = np.empty((nz, ny, nx), dtype=self._dtype())
        check("x3d_memcpy_d2h", self.lib.x3d_memcpy_d2h(self._ctx(grid), out.ctypes.data_as(C.c_void_p),
                                                         self._ptr(grid, name), out.nbytes))
        return out

    def put(self, grid: int, name: str, a) -> None:
        nx, ny, nz = self.dims[grid]
        a = np.ascontiguousarray(a, dtype=self._dtype()).reshape(nz, ny, nx)
        check("x3d_memcpy_h2d", self.lib.x3d_memcpy_h2d(self._ctx(grid), self._ptr(grid, name),
                                                         a.ctypes.data_as(C.c_void_p), a.nbytes))

    def init_fields(self, kind: int = 2, seed: int = 1234, amp: float = 0.0):
        check("x3d_batch_init_fields", self.lib.x3d_batch_init_fields(self._h, kind, seed, amp))

    def step(self, gdt: float = 0.0):
        check("x3d_batch_step", self.lib.x3d_batch_step(self._h, float(gdt)))

    def graph_capture(self, gdt: float = 0.0):
        check("x3d_batch_graph_capture", self.lib.x3d_batch_graph_capture(self._h, float(gdt)))

    def graph_launch(self):
        check("x3d_batch_graph_launch", self.lib.x3d_batch_graph_launch(self._h))

    def autotune(self, reps: int = 3) -> str:
        buf = C.create_string_buffer(8192)
        check("x3d_batch_autotune", self.lib.x3d_batch_autotune(self._h, reps, buf, len(buf)))
        return buf.value.decode()
