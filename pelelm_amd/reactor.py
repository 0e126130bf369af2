"""Python mirror of PeleLM's reactor interface over the C ABI (``include/pelelm_chem.h``).

PeleLM drives chemistry through ``pele::physics::reactions::ReactorBase``
(``Source/PeleLM.H:16,672-674``): ``ReactorBase::create(name)`` then
``init(reactor_type=2, ncells=1)`` (``Source/PeleLM.cpp:701-709``),
``set_typ_vals_ode`` (``Source/PeleLM.cpp:1497-1516``), ``react(box, rY, rYsrc, T, rE,
rEsrc, FC, mask, dt, time, stream)`` per box (``Source/PeleLM.cpp:5752-5758``) and
``close()`` (``Source/PeleLM.cpp:930``).  This class keeps those names and argument
meanings; arrays are torch tensors on the GPU (CGS units, structure-of-arrays with the
cell index fastest), the work is enqueued on the current torch stream, and the
ParmParse keys the reference decks use are passed through ``set``.

The HIP library must be present: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "lib", "libpelelm_chem.so")

MI_OK, MI_EINVAL, MI_ESTATE, MI_EHIP, MI_ECELLFAIL, MI_ENOMEM = 0, -1, -2, -3, -4, -5


class MiReal4(C.Structure):
    _fields_ = [("p", C.c_void_p), ("begin", C.c_int * 3), ("end", C.c_int * 3), ("jstride", C.c_long),
                ("kstride", C.c_long), ("nstride", C.c_long), ("ncomp", C.c_int)]


class MiInt4(C.Structure):
    _fields_ = [("p", C.c_void_p), ("begin", C.c_int * 3), ("end", C.c_int * 3), ("jstride", C.c_long),
                ("kstride", C.c_long), ("nstride", C.c_long), ("ncomp", C.c_int)]


class MiStats(C.Structure):
    _fields_ = [("ncells", C.c_long), ("nsteps", C.c_long), ("nfe", C.c_long), ("nje", C.c_long),
                ("nsetups", C.c_long), ("nni", C.c_long), ("ncfn", C.c_long), ("netf", C.c_long),
                ("nfail", C.c_long), ("max_steps_cell", C.c_long), ("nlu_refine", C.c_long),
                ("nlu_fallback", C.c_long), ("npred_used", C.c_long), ("npred_fallback", C.c_long),
                ("ms_kernel", C.c_double), ("ntrips", C.c_long), ("nhy2t_iters", C.c_long),
                ("kernel_ms", C.c_double * 6),
                ("kernel_launches", C.c_long * 6), ("kernel_cells", C.c_long * 6),
                ("schedule_used", C.c_long), ("graph_trips_used", C.c_long), ("tune_ms", C.c_double * 4),
                ("nli", C.c_long)]


# Every exported entry point of include/pelelm_chem.h.
EXPORTS = ["mi_chem_create", "mi_chem_init", "mi_chem_set", "mi_chem_set_typ_vals", "mi_chem_react_box",
           "mi_chem_react_flat", "mi_chem_react_flat_noalloc", "mi_chem_reaction_rates_flat", "mi_chem_reaction_rates_box", "mi_chem_advance_box",
           "mi_chem_rhoh_to_temp_box",
           "mi_chem_get_stats", "mi_chem_nspec", "mi_chem_last_error", "mi_chem_close"]

_lib = None


def lib():
    """Load the HIP reactor library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.mi_chem_create.argtypes = [C.c_char_p, C.POINTER(vp)]
        L.mi_chem_init.argtypes = [vp, C.c_int, C.c_int]
        L.mi_chem_set.argtypes = [vp, C.c_char_p, C.c_char_p]
        L.mi_chem_set_typ_vals.argtypes = [vp, C.POINTER(C.c_double), C.c_int]
        L.mi_chem_react_box.argtypes = [vp, C.c_int * 3, C.c_int * 3, MiReal4, MiReal4, MiReal4, MiReal4, MiReal4,
                                        MiReal4, MiInt4, C.POINTER(C.c_double), C.POINTER(C.c_double), vp]
        L.mi_chem_react_flat.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp, C.c_int, C.POINTER(C.c_double),
                                         C.POINTER(C.c_double), vp]
        L.mi_chem_react_flat_noalloc.argtypes = L.mi_chem_react_flat.argtypes
        L.mi_chem_reaction_rates_flat.argtypes = [vp, vp, vp, vp, vp, vp, vp, C.c_int, vp]
        L.mi_chem_reaction_rates_box.argtypes = [vp, C.c_int * 3, C.c_int * 3, MiReal4, MiReal4, MiReal4, MiInt4,
                                                 MiReal4, vp]
        L.mi_chem_advance_box.argtypes = [vp, C.c_int * 3, C.c_int * 3, MiReal4, MiReal4, MiReal4, MiReal4, MiReal4,
                                          MiInt4, C.c_double, vp]
        L.mi_chem_rhoh_to_temp_box.argtypes = [vp, C.c_int * 3, C.c_int * 3, MiReal4, MiReal4, MiReal4, MiReal4, vp]
        L.mi_chem_get_stats.argtypes = [vp, C.POINTER(MiStats)]
        L.mi_chem_nspec.argtypes = [vp]
        L.mi_chem_last_error.argtypes = [vp]
        L.mi_chem_last_error.restype = C.c_char_p
        L.mi_chem_close.argtypes = [vp]
        for f in EXPORTS:
            if f != "mi_chem_last_error":
                getattr(L, f).restype = C.c_int
        _lib = L
    return _lib


def _ptr(t):
    if t is None:
        return None
    if not t.is_cuda:
        raise ValueError("reactor arrays must be device tensors")
    if not t.is_contiguous():
        raise ValueError("reactor arrays must be contiguous")
    return C.c_void_p(t.data_ptr())


def _stream(stream):
    import torch
    s = torch.cuda.current_stream() if stream is None else stream
    return C.c_void_p(s.cuda_stream)


class ReactorError(RuntimeError):
    pass


class CellFailure(ReactorError):
    """A cell failed to integrate under ``mi.on_fail = abort`` (the default), where the
    reference's integrator aborts inside ``react``; failed cells carry FC < 0."""


class ReactorBase:
    """``ReactorBase`` with the reference's method names (create / init / set_typ_vals_ode /
    react / close), backed by the MI355X reactor."""

    def __init__(self, handle, name):
        self._h = handle
        self.name = name

    # -- lifecycle ---------------------------------------------------------
    @staticmethod
    def create(name: str) -> "ReactorBase":
        h = C.c_void_p()
        rc = lib().mi_chem_create(name.encode(), C.byref(h))
        if rc != MI_OK:
            raise ReactorError(f"unknown chemistry integrator {name!r}")
        return ReactorBase(h, name)

    def _check(self, rc):
        if rc != MI_OK:
            msg = lib().mi_chem_last_error(self._h)
            err = CellFailure if rc == MI_ECELLFAIL else ReactorError
            raise err(f"{rc}: {msg.decode() if msg else ''}")
        return rc

    def init(self, reactor_type: int = 2, ncells: int = 1):
        self._check(lib().mi_chem_init(self._h, reactor_type, ncells))

    def set(self, key: str, value) -> None:
        self._check(lib().mi_chem_set(self._h, key.encode(), str(value).encode()))

    def set_typ_vals_ode(self, typ) -> None:
        arr = (C.c_double * len(typ))(*[float(x) for x in typ])
        self._check(lib().mi_chem_set_typ_vals(self._h, arr, len(typ)))

    @property
    def nspec(self) -> int:
        return lib().mi_chem_nspec(self._h)

    def close(self) -> None:
        if self._h:
            lib().mi_chem_close(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- work --------------------------------------------------------------
    def react(self, rY, rYsrc, T, rE, rEsrc, FC, mask, dt: float, time: float = 0.0, stream=None,
              noalloc: bool = False):
        """Flat SoA form: rY [nspec, ncell], rYsrc [nspec, ncell] or None, T/rE/FC [ncell],
        rEsrc [ncell] or None, mask [ncell] int32 or None.  Returns the updated (dt, time).
        noalloc=True calls mi_chem_react_flat_noalloc (persistent schedule in the library's
        fixed device array; nothing is allocated during the call)."""
        ncell = T.shape[0]
        for name, t, n in (("rY", rY, self.nspec), ("rYsrc", rYsrc, self.nspec)):
            if t is not None and tuple(t.shape) != (n, ncell):
                raise ValueError(f"{name} must have shape ({n}, {ncell})")
        for name, t in (("rE", rE), ("rEsrc", rEsrc), ("FC", FC), ("mask", mask)):
            if t is not None and tuple(t.shape) != (ncell,):
                raise ValueError(f"{name} must have shape ({ncell},)")
        d, tm = C.c_double(dt), C.c_double(time)
        fn = lib().mi_chem_react_flat_noalloc if noalloc else lib().mi_chem_react_flat
        self._check(fn(self._h, _ptr(rY), _ptr(rYsrc), _ptr(T), _ptr(rE), _ptr(rEsrc), _ptr(FC), _ptr(mask), ncell,
                       C.byref(d), C.byref(tm), _stream(stream)))
        return d.value, tm.value

    def react_box(self, lo, hi, rY: MiReal4, rYsrc: MiReal4, T: MiReal4, rE: MiReal4, rEsrc: MiReal4,
                  FC: MiReal4, mask: MiInt4, dt: float, time: float = 0.0, stream=None):
        d, tm = C.c_double(dt), C.c_double(time)
        self._check(lib().mi_chem_react_box(self._h, (C.c_int * 3)(*lo), (C.c_int * 3)(*hi), rY, rYsrc, T, rE,
                                            rEsrc, FC, mask, C.byref(d), C.byref(tm), _stream(stream)))
        return d.value, tm.value

    def advance_box(self, lo, hi, s_old: MiReal4, s_new: MiReal4, force: MiReal4, rhoYdot: MiReal4, FC: MiReal4,
                    mask: MiInt4, dt: float, stream=None):
        """PeleLM::advance_chemistry on one box (Source/PeleLM.cpp:5626-5844, per-box part):
        MKS state in s_old (rhoY, rhoH, T from component 0), new state into s_new, I_R into
        rhoYdot, RHS counts into FC."""
        self._check(lib().mi_chem_advance_box(self._h, (C.c_int * 3)(*lo), (C.c_int * 3)(*hi), s_old, s_new, force,
                                              rhoYdot, FC, mask, float(dt), _stream(stream)))

    def rhoh_to_temp_box(self, lo, hi, rho: MiReal4, rhoY: MiReal4, rhoH: MiReal4, T: MiReal4, stream=None):
        """PeleLM::RhoH_to_Temp on one box (getTfromHY, Source/PeleLM_K.H:511-530), MKS."""
        self._check(lib().mi_chem_rhoh_to_temp_box(self._h, (C.c_int * 3)(*lo), (C.c_int * 3)(*hi), rho, rhoY, rhoH,
                                                   T, _stream(stream)))

    def reaction_rates(self, rY, rH, T, mask, rYdot, T_out=None, stream=None):
        """Net production rates rho*wdot of flat CGS cells (reactionRateRhoY,
        Source/PeleLM_K.H:57-98, in CGS): T is the HY2T guess and is not written; the
        re-solved temperatures go to T_out when given."""
        ncell = T.shape[0]
        self._check(lib().mi_chem_reaction_rates_flat(self._h, _ptr(rY), _ptr(rH), _ptr(T), _ptr(mask),
                                                      _ptr(rYdot), _ptr(T_out), ncell, _stream(stream)))

    def reaction_rates_box(self, lo, hi, rhoY: MiReal4, rhoH: MiReal4, T: MiReal4, mask: MiInt4, rhoYdot: MiReal4,
                           stream=None):
        """PeleLM::compute_instantaneous_reaction_rates on one box (reactionRateRhoY,
        Source/PeleLM_K.H:57-98, called at Source/PeleLM.cpp:2046-2055): MKS rhoY, rhoH
        and T (read only) in, MKS rhoYdot out, zero where mask < 0."""
        self._check(lib().mi_chem_reaction_rates_box(self._h, (C.c_int * 3)(*lo), (C.c_int * 3)(*hi), rhoY, rhoH, T,
                                                     mask, rhoYdot, _stream(stream)))

    def stats(self) -> dict:
        st = MiStats()
        rc = lib().mi_chem_get_stats(self._h, C.byref(st))
        out = {f: getattr(st, f) for f, _ in MiStats._fields_}
        for f in ("kernel_ms", "kernel_launches", "kernel_cells", "tune_ms"):
            out[f] = list(out[f])
        out["rc"] = rc
        return out


def array4(t, lo, hi, ncomp: int) -> MiReal4:
    """Array4 view of a contiguous torch tensor shaped [ncomp, nz, ny, nx] covering [lo, hi]."""
    nz, ny, nx = t.shape[-3], t.shape[-2], t.shape[-1]
    v = MiReal4()
    v.p = t.data_ptr() if t is not None else None
    v.begin = (C.c_int * 3)(*lo)
    v.end = (C.c_int * 3)(lo[0] + nx, lo[1] + ny, lo[2] + nz)
    v.jstride, v.kstride, v.nstride = nx, nx * ny, nx * ny * nz
    v.ncomp = ncomp
    return v


def int4(t, lo) -> MiInt4:
    v = MiInt4()
    if t is None:
        v.p = None
        return v
    nz, ny, nx = t.shape[-3], t.shape[-2], t.shape[-1]
    v.p = t.data_ptr()
    v.begin = (C.c_int * 3)(*lo)
    v.end = (C.c_int * 3)(lo[0] + nx, lo[1] + ny, lo[2] + nz)
    v.jstride, v.kstride, v.nstride = nx, nx * ny, nx * ny * nz
    v.ncomp = 1
    return v


def null_real4() -> MiReal4:
    return MiReal4()
