"""ctypes binding of the CPU oracle (``libchem_oracle.so``).

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, ``__graft_entry__.smoke()`` and the
CPU-baseline leg of ``bench.py`` as the checker.  The product path never imports it.
See ``chem_oracle.h`` for what each function restates and the reference lines it follows.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libchem_oracle.so")
TBL_DIR = os.path.join(os.path.dirname(HERE), "pelelm_amd", "mechanisms")

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


class Opts(C.Structure):
    _fields_ = [("rtol", C.c_double), ("atol", C.c_double), ("atol_vec", C.c_void_p),
                ("max_order", C.c_int), ("max_steps", C.c_long), ("solve_type", C.c_int),
                ("gmres_maxl", C.c_int)]


class CellStats(C.Structure):
    _fields_ = [("nst", C.c_long), ("nfe", C.c_long), ("nje", C.c_long), ("nsetups", C.c_long),
                ("nni", C.c_long), ("ncfn", C.c_long), ("netf", C.c_long), ("status", C.c_int),
                ("qlast", C.c_int), ("hlast", C.c_double), ("nli", C.c_long)]


def build(force: bool = False) -> str:
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-B" if force else "-s", "-C", HERE], check=True)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.orc_load.restype = C.c_void_p
        L.orc_load.argtypes = [C.c_char_p]
        L.orc_free.argtypes = [C.c_void_p]
        for f in ("orc_nspec", "orc_nreac", "orc_nelem"):
            getattr(L, f).argtypes = [C.c_void_p]
            getattr(L, f).restype = C.c_int
        L.orc_gas_constant.argtypes = [C.c_void_p]
        L.orc_gas_constant.restype = C.c_double
        L.orc_molwt.argtypes = [C.c_void_p, _dp]
        L.orc_elem_comp.argtypes = [C.c_void_p, _ip]
        L.orc_thermo.argtypes = [C.c_void_p, C.c_double, _dp, _dp, _dp]
        L.orc_wdot.argtypes = [C.c_void_p, _dp, C.c_double, _dp, C.c_void_p]
        L.orc_rop.argtypes = [C.c_void_p, _dp, C.c_double, _dp, _dp]
        L.orc_hy2t.argtypes = [C.c_void_p, C.c_double, _dp, C.c_double, C.POINTER(C.c_int)]
        L.orc_hy2t.restype = C.c_double
        L.orc_rhs.argtypes = [C.c_void_p, C.c_double, _dp, _dp, C.c_double, C.c_double, _dp]
        L.orc_rhs_fixedT.argtypes = [C.c_void_p, _dp, _dp, C.c_double, _dp]
        L.orc_jac.argtypes = [C.c_void_p, _dp, _dp, C.c_double, _dp]
        L.orc_jac_cs.argtypes = [C.c_void_p, _dp, _dp, C.c_double, _dp]
        L.orc_react.argtypes = [C.c_void_p, C.c_int, _dp, C.c_void_p, _dp, _dp, C.c_void_p, _dp,
                                C.c_void_p, C.c_double, C.POINTER(Opts), C.c_int, C.c_void_p]
        L.orc_react.restype = C.c_int
        L.orc_bdf_linear.argtypes = [C.c_int, _dp, _dp, C.c_double, C.POINTER(Opts),
                                     C.POINTER(CellStats)]
        L.orc_bdf_linear.restype = C.c_int
        if hasattr(L, "orc_radau_react"):
            L.orc_radau_react.argtypes = [C.c_void_p, C.c_int, _dp, C.c_void_p, _dp, _dp,
                                          C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_int,
                                          C.c_void_p]
            L.orc_radau_react.restype = C.c_int
            L.orc_radau_coefficients.argtypes = [_dp]
        L.orc_rk64_react.argtypes = [C.c_void_p, C.c_int, _dp, C.c_void_p, _dp, _dp, C.c_void_p, _dp, C.c_void_p,
                                     C.c_double, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int, C.c_int]
        L.orc_rk64_react.restype = C.c_int
        L.orc_rk64_linear.argtypes = [C.c_int, _dp, _dp, C.c_double, C.c_double, C.c_double, C.c_int, C.c_int,
                                      C.c_int, C.c_void_p]
        L.orc_rk64_linear.restype = C.c_int
        L.orc_rk64_tableau.argtypes = [_dp]
        L.orc_spgmr_dense.argtypes = [C.c_int, _dp, _dp, _dp, C.c_int, C.c_double, _dp, _dp]
        L.orc_spgmr_dense.restype = C.c_int
        _lib = L
    return _lib


SOLVE_TYPES = {"dense": 0, "gmres": 1}


def make_opts(rtol=1e-6, atol=1e-10, atol_vec=None, max_order=4, max_steps=100000, solve="dense", maxl=0):
    o = Opts(rtol, atol, None, max_order, max_steps, SOLVE_TYPES[solve], int(maxl))
    keep = None
    if atol_vec is not None:
        keep = np.ascontiguousarray(atol_vec, dtype=np.float64)
        o.atol_vec = keep.ctypes.data
    return o, keep


class Mech:
    """One mechanism table loaded into the oracle."""

    def __init__(self, name: str):
        path = name if os.path.exists(name) else os.path.join(TBL_DIR, f"{name}.tbl")
        self._h = lib().orc_load(path.encode())
        if not self._h:
            raise RuntimeError(f"cannot load mechanism table {path}")
        self.name = name
        self.nspec = lib().orc_nspec(self._h)
        self.nreac = lib().orc_nreac(self._h)
        self.nelem = lib().orc_nelem(self._h)
        self.W = np.zeros(self.nspec)
        lib().orc_molwt(self._h, self.W)
        self.ecomp = np.zeros((self.nspec, self.nelem), dtype=np.int32)
        lib().orc_elem_comp(self._h, self.ecomp)
        self.Ru = lib().orc_gas_constant(self._h)

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.orc_free(self._h)

    # -- thermo / kinetics -------------------------------------------------
    def thermo(self, T):
        cp, h, s = (np.zeros(self.nspec) for _ in range(3))
        lib().orc_thermo(self._h, float(T), cp, h, s)
        return cp, h, s

    def wdot(self, Cc, T, with_q=False):
        w = np.zeros(self.nspec)
        q = np.zeros(self.nreac) if with_q else None
        lib().orc_wdot(self._h, np.ascontiguousarray(Cc, dtype=np.float64), float(T), w,
                       q.ctypes.data if with_q else None)
        return (w, q) if with_q else w

    def rop(self, Cc, T):
        qf, qr = np.zeros(self.nreac), np.zeros(self.nreac)
        lib().orc_rop(self._h, np.ascontiguousarray(Cc, dtype=np.float64), float(T), qf, qr)
        return qf, qr

    def hy2t(self, h, Y, Tguess):
        it = C.c_int(0)
        T = lib().orc_hy2t(self._h, float(h), np.ascontiguousarray(Y, dtype=np.float64), float(Tguess),
                           C.byref(it))
        return T, it.value

    def rhs(self, t, y, fY=None, rhoH0=0.0, fH=0.0):
        fY = np.zeros(self.nspec) if fY is None else np.ascontiguousarray(fY, dtype=np.float64)
        yd = np.zeros(self.nspec + 1)
        lib().orc_rhs(self._h, float(t), np.ascontiguousarray(y, dtype=np.float64), fY, float(rhoH0),
                      float(fH), yd)
        return yd

    def rhs_fixedT(self, y, fY=None, fH=0.0):
        fY = np.zeros(self.nspec) if fY is None else np.ascontiguousarray(fY, dtype=np.float64)
        yd = np.zeros(self.nspec + 1)
        lib().orc_rhs_fixedT(self._h, np.ascontiguousarray(y, dtype=np.float64), fY, float(fH), yd)
        return yd

    def jac(self, y, fY=None, fH=0.0, complex_step=False):
        n = self.nspec + 1
        fY = np.zeros(self.nspec) if fY is None else np.ascontiguousarray(fY, dtype=np.float64)
        J = np.zeros(n * n)
        f = lib().orc_jac_cs if complex_step else lib().orc_jac
        f(self._h, np.ascontiguousarray(y, dtype=np.float64), fY, float(fH), J)
        return J.reshape(n, n)

    # -- reactor ------------------------------------------------------------
    def react(self, rY, T, rE, dt, rYsrc=None, rEsrc=None, mask=None, rtol=1e-6, atol=1e-10,
              atol_vec=None, max_order=4, max_steps=100000, nthreads=0, stats=False, solve="dense", maxl=0):
        """Advance SoA cell states (rY: [nspec, ncell]) in place-copy; returns new arrays.
        solve: "dense" (LU of I - gamma J) or "gmres" (cvode.solve_type = GMRES, Krylov
        dimension maxl, 0: CVODE's default 5)."""
        rY = np.array(rY, dtype=np.float64, order="C", copy=True)
        T = np.array(T, dtype=np.float64, copy=True)
        rE = np.array(rE, dtype=np.float64, copy=True)
        ncell = T.shape[0]
        FC = np.zeros(ncell)
        src = None if rYsrc is None else np.ascontiguousarray(rYsrc, dtype=np.float64)
        esrc = None if rEsrc is None else np.ascontiguousarray(rEsrc, dtype=np.float64)
        msk = None if mask is None else np.ascontiguousarray(mask, dtype=np.int32)
        o, keep = make_opts(rtol, atol, atol_vec, max_order, max_steps, solve, maxl)
        st = (CellStats * ncell)() if stats else None
        nfail = lib().orc_react(self._h, ncell, rY, None if src is None else src.ctypes.data, T, rE,
                                None if esrc is None else esrc.ctypes.data, FC,
                                None if msk is None else msk.ctypes.data, float(dt), C.byref(o),
                                int(nthreads), None if st is None else C.cast(st, C.c_void_p))
        out = dict(rY=rY, T=T, rE=rE, FC=FC, nfail=nfail)
        if stats:
            out["stats"] = {f: np.array([getattr(s, f) for s in st]) for f, _ in CellStats._fields_}
        return out

    def rk64_react(self, rY, T, rE, dt, rYsrc=None, rEsrc=None, mask=None, rtol=1e-6, atol=1e-10,
                   nsub=(10, 5, 500), nthreads=0):
        """The explicit reactor (ReactorRK64) on SoA cells; nsub = (guess, min, max)."""
        rY = np.array(rY, dtype=np.float64, order="C", copy=True)
        T = np.array(T, dtype=np.float64, copy=True)
        rE = np.array(rE, dtype=np.float64, copy=True)
        ncell = T.shape[0]
        FC = np.zeros(ncell)
        src = None if rYsrc is None else np.ascontiguousarray(rYsrc, dtype=np.float64)
        esrc = None if rEsrc is None else np.ascontiguousarray(rEsrc, dtype=np.float64)
        msk = None if mask is None else np.ascontiguousarray(mask, dtype=np.int32)
        nfail = lib().orc_rk64_react(self._h, ncell, rY, None if src is None else src.ctypes.data, T, rE,
                                     None if esrc is None else esrc.ctypes.data, FC,
                                     None if msk is None else msk.ctypes.data, float(dt), float(rtol), float(atol),
                                     int(nsub[0]), int(nsub[1]), int(nsub[2]), int(nthreads))
        return dict(rY=rY, T=T, rE=rE, FC=FC, nfail=nfail)

    def radau_react(self, rY, T, rE, dt, rtol=1e-12, atol=1e-20, rYsrc=None, rEsrc=None, nthreads=0):
        rY = np.array(rY, dtype=np.float64, order="C", copy=True)
        T = np.array(T, dtype=np.float64, copy=True)
        rE = np.array(rE, dtype=np.float64, copy=True)
        ncell = T.shape[0]
        src = None if rYsrc is None else np.ascontiguousarray(rYsrc, dtype=np.float64)
        esrc = None if rEsrc is None else np.ascontiguousarray(rEsrc, dtype=np.float64)
        nfail = lib().orc_radau_react(self._h, ncell, rY, None if src is None else src.ctypes.data, T, rE,
                                      None if esrc is None else esrc.ctypes.data, float(dt), float(rtol),
                                      float(atol), int(nthreads), None)
        return dict(rY=rY, T=T, rE=rE, nfail=nfail)


def rk64_tableau():
    """(c, A, b4, b5) of the explicit reactor's Cash-Karp pair."""
    out = np.zeros(6 + 30 + 12)
    lib().orc_rk64_tableau(out)
    return out[:6], out[6:36].reshape(6, 5), out[36:42], out[42:48]


def rk64_linear(A, y0, tout, rtol=1e-8, atol=1e-12, nsub=(10, 5, 100000)):
    A = np.ascontiguousarray(A, dtype=np.float64)
    y = np.array(y0, dtype=np.float64, copy=True)
    st = np.zeros(3, dtype=np.int64)
    r = lib().orc_rk64_linear(A.shape[0], A.ravel(), y, float(tout), float(rtol), float(atol), nsub[0], nsub[1],
                              nsub[2], st.ctypes.data)
    return y, r, st


def radau_coefficients():
    """(c, A, e, g0) of the Radau IIA-5 integrator as the oracle derived them."""
    out = np.zeros(16)
    lib().orc_radau_coefficients(out)
    return out[:3], out[3:12].reshape(3, 3), out[12:15], out[15]


def spgmr_dense(A, b, w, maxl, delta):
    """The GMRES core on a dense matrix: (status, x, residual estimate, iterations)."""
    A = np.ascontiguousarray(A, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    w = np.ascontiguousarray(w, dtype=np.float64)
    x = np.zeros(A.shape[0])
    out = np.zeros(2)
    r = lib().orc_spgmr_dense(A.shape[0], A.ravel(), b, w, int(maxl), float(delta), x, out)
    return r, x, out[0], int(out[1])


def bdf_linear(A, y0, tout, rtol=1e-8, atol=1e-12, max_order=5, solve="dense", maxl=0):
    A = np.ascontiguousarray(A, dtype=np.float64)
    y = np.array(y0, dtype=np.float64, copy=True)
    o, _ = make_opts(rtol, atol, None, max_order, 100000, solve, maxl)
    st = CellStats()
    r = lib().orc_bdf_linear(A.shape[0], A.ravel(), y, float(tout), C.byref(o), C.byref(st))
    return y, r, st
