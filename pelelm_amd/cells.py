"""Seeded synthetic reactor inputs (wrapper of ``libpelelm_cells.so``).

The workloads the chemistry benchmarks are quoted on: every cell at 1 atm with zero
forcing (the isolated-reactor case PeleLM runs in its initial divu iterations,
``Source/PeleLM.cpp:2396-2400``), its composition a blend of fresh fuel/air and the
mixture's HP-equilibrium products set by a progress variable c, with T interpolated
between 300 K and the adiabatic temperature plus N(0, 20 K) noise.  Cells are
generated independently from Philox4x32-10 keyed by (seed, cell index).
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "lib", "libpelelm_cells.so")
MECH_DIR = os.path.join(PKG, "mechanisms")

P_ATM = 1.01325e6


class _Spec(C.Structure):
    _fields_ = [("fuel", C.c_char_p), ("o2_per_fuel", C.c_double), ("phi_lo", C.c_double),
                ("phi_hi", C.c_double), ("mode", C.c_int), ("seed", C.c_uint64), ("T_fresh", C.c_double),
                ("T_noise", C.c_double), ("pressure", C.c_double)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        L.mi_cells_generate.argtypes = [C.c_char_p, C.POINTER(_Spec), C.c_long, C.c_void_p, C.c_void_p,
                                        C.c_void_p, C.c_void_p, C.c_void_p, C.c_int]
        L.mi_cells_generate.restype = C.c_int
        L.mi_cells_equilibrium_hp.argtypes = [C.c_char_p, C.c_void_p, C.c_double, C.c_void_p,
                                              C.POINTER(C.c_double)]
        L.mi_cells_equilibrium_hp.restype = C.c_int
        L.mi_cells_philox.argtypes = [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
        _lib = L
    return _lib


def table_path(mech: str) -> str:
    return os.path.join(MECH_DIR, f"{mech}.tbl")


def species_names(mech: str) -> list:
    with open(table_path(mech)) as f:
        return [ln.split()[1] for ln in f if ln.startswith("spec ")]


@dataclass
class Workload:
    mech: str
    fuel: str
    o2_per_fuel: float
    phi: tuple
    mode: int          # 0: c ~ U[0,1]; 1: 60/30/10 cold/near-equilibrium/front split, shuffled
    seed: int
    ncell: int
    dt: float


# The configurations the reactor is measured and checked on.
WORKLOADS = {
    "lidryer_16": Workload("LiDryer", "H2", 0.5, (0.5, 1.5), 0, 101, 16 ** 3, 1e-6),
    "drm19_64": Workload("drm19", "CH4", 2.0, (0.6, 1.2), 0, 102, 64 ** 3, 1e-6),
    "drm19_128_mix": Workload("drm19", "CH4", 2.0, (0.6, 1.2), 1, 104, 128 ** 3, 1e-5),
    "dodecane53_64_mix": Workload("dodecane53", "NC12H26", 18.5, (0.6, 1.2), 1, 105, 64 ** 3, 1e-5),
    "dodecane56_64_mix": Workload("dodecane56", "NC12H26", 18.5, (0.6, 1.2), 1, 106, 64 ** 3, 1e-5),
    "heptane88_64_mix": Workload("heptane88", "NC7H16", 11.0, (0.6, 1.2), 1, 107, 64 ** 3, 1e-5),
}


def generate(mech: str, fuel: str, o2_per_fuel: float, phi: tuple, mode: int, seed: int, ncell: int,
             nthreads: int = 0, T_fresh: float = 300.0, T_noise: float = 20.0, pressure: float = P_ATM):
    """Return dict(rY [nspec, ncell], T, rE, c, cls) as float64 numpy arrays (CGS)."""
    ns = len(species_names(mech))
    rY = np.empty((ns, ncell), dtype=np.float64)
    T = np.empty(ncell)
    rE = np.empty(ncell)
    c = np.empty(ncell)
    cls = np.empty(ncell, dtype=np.int32)
    spec = _Spec(fuel.encode(), o2_per_fuel, phi[0], phi[1], mode, seed, T_fresh, T_noise, pressure)
    r = lib().mi_cells_generate(table_path(mech).encode(), C.byref(spec), ncell, rY.ctypes.data,
                                T.ctypes.data, rE.ctypes.data, c.ctypes.data, cls.ctypes.data, nthreads)
    if r != 0:
        raise RuntimeError(f"cell generation failed ({r})")
    return dict(rY=rY, T=T, rE=rE, c=c, cls=cls)


def generate_workload(name: str, ncell: int | None = None, nthreads: int = 0, seed: int | None = None):
    w = WORKLOADS[name]
    n = w.ncell if ncell is None else ncell
    out = generate(w.mech, w.fuel, w.o2_per_fuel, w.phi, w.mode, w.seed if seed is None else seed, n, nthreads)
    out["dt"] = w.dt
    out["mech"] = w.mech
    return out


def equilibrium_hp(mech: str, X0, T0: float = 300.0):
    X0 = np.ascontiguousarray(X0, dtype=np.float64)
    Y = np.zeros_like(X0)
    T = C.c_double(0.0)
    r = lib().mi_cells_equilibrium_hp(table_path(mech).encode(), X0.ctypes.data, float(T0), Y.ctypes.data,
                                      C.byref(T))
    if r != 0:
        raise RuntimeError(f"equilibrium failed ({r})")
    return Y, T.value


def philox(ctr, key):
    out = (C.c_uint32 * 4)()
    lib().mi_cells_philox((C.c_uint32 * 4)(*ctr), (C.c_uint32 * 2)(*key), out)
    return list(out)
