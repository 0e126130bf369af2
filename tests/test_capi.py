"""The C ABI: every symbol include/pelelm_chem.h declares is exported, and the
host-side argument checking behaves (no GPU call is made by these tests)."""
import ctypes as C
import os
import re

import pytest

from pelelm_amd import reactor

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "pelelm_chem.h")).read()
    return sorted(set(re.findall(r"\b(mi_chem_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = reactor.lib()
    syms = declared_symbols()
    assert len(syms) >= 11
    for s in syms:
        assert hasattr(lib, s), s
    assert sorted(reactor.EXPORTS) == syms


def test_library_exports_every_network_symbol():
    from pelelm_amd import predictor
    text = open(os.path.join(ROOT, "include", "pelelm_mlp.h")).read()
    syms = sorted(set(re.findall(r"\b(mi_(?:mlp|chem)_\w+)\s*\(", text)))
    lib = reactor.lib()
    for s in syms:
        assert hasattr(lib, s), s
    assert sorted(predictor.MLP_EXPORTS) == syms


def test_create_rejects_unknown_integrator():
    with pytest.raises(reactor.ReactorError):
        reactor.ReactorBase.create("ReactorArkode")


@pytest.mark.parametrize("name", ["ReactorCvode", "ReactorBDF", "ReactorRK64", "ReactorNull"])
def test_create_init_close(name):
    r = reactor.ReactorBase.create(name)
    r.init(2, 100)
    r.close()


def test_init_rejects_e_reactor():
    r = reactor.ReactorBase.create("ReactorCvode")
    with pytest.raises(reactor.ReactorError, match="reactor_type"):
        r.init(1, 1)
    r.close()


@pytest.mark.parametrize("key,value", [("ode.rtol", "-1"), ("ode.rtol", "abc"), ("cvode.max_order", "7"),
                                       ("cvode.solve_type", "precGMRES"), ("mi.gmres_maxl", "0"),
                                       ("mi.gmres_maxl", "65"), ("mi.mechanism", "gri30"),
                                       ("mi.schedule", "fifo"), ("mi.graph_trips", "3"), ("mi.graph_trips", "-2"),
                                       ("mi.graph_trips", "66"), ("mi.lu", "fp16"), ("mi.lu_form", "warp"),
                                       ("no.such.key", "1")])
def test_set_rejects_bad_keys(key, value):
    r = reactor.ReactorBase.create("ReactorCvode")
    with pytest.raises(reactor.ReactorError):
        r.set(key, value)
    r.close()


def test_set_accepts_reference_deck_keys():
    r = reactor.ReactorBase.create("ReactorCvode")
    for k, v in [("ode.rtol", "1.0e-6"), ("ode.atol", "1.0e-5"), ("cvode.solve_type", "denseAJ_direct"),
                 ("cvode.max_order", "4"), ("peleLM.use_typ_vals_chem", "1"), ("mi.mechanism", "LiDryer"),
                 ("mi.schedule", "plain"), ("mi.schedule", "queue"), ("mi.schedule", "auto"),
                 ("mi.graph_trips", "8"), ("mi.graph_trips", "0"), ("mi.lu", "fp32_ir"), ("mi.lu", "fp64"),
                 ("mi.lu_form", "lane"), ("mi.lu_form", "coop"), ("mi.lu_form", "mfma"), ("mi.ir_sweeps", "2"),
                 ("cvode.solve_type", "GMRES"), ("mi.gmres_maxl", "8"), ("cvode.solve_type", "dense_direct")]:
        r.set(k, v)
    assert r.nspec == 9
    r.set_typ_vals_ode([1e-3] * 9 + [1500.0])
    with pytest.raises(reactor.ReactorError):
        r.set_typ_vals_ode([1e-3] * 9)
    r.close()


def test_react_before_init_is_a_state_error():
    lib = reactor.lib()
    h = C.c_void_p()
    assert lib.mi_chem_create(b"ReactorCvode", C.byref(h)) == 0
    dt, tm = C.c_double(1e-6), C.c_double(0.0)
    rc = lib.mi_chem_react_flat(h, None, None, None, None, None, None, None, 0, C.byref(dt), C.byref(tm), None)
    assert rc == reactor.MI_ESTATE
    assert b"before init" in lib.mi_chem_last_error(h)
    assert lib.mi_chem_init(h, 2, 1) == 0
    dt.value = -1.0
    rc = lib.mi_chem_react_flat(h, None, None, None, None, None, None, None, 0, C.byref(dt), C.byref(tm), None)
    assert rc == reactor.MI_EINVAL
    assert lib.mi_chem_close(h) == 0


def test_advance_box_argument_errors():
    """The glue entry point checks its views before touching the device: call order,
    coverage of the box, and old and new state sharing memory."""
    lib = reactor.lib()
    h = C.c_void_p()
    assert lib.mi_chem_create(b"ReactorCvode", C.byref(h)) == 0
    buf = (C.c_double * (23 * 8))()
    ibuf = (C.c_int * 8)()

    def view(off, ncomp):
        v = reactor.MiReal4()
        v.p = C.cast(C.byref(buf, 8 * off), C.c_void_p).value
        v.begin = (C.c_int * 3)(0, 0, 0)
        v.end = (C.c_int * 3)(2, 2, 2)
        v.jstride, v.kstride, v.nstride, v.ncomp = 2, 4, 8, ncomp
        return v
    m = reactor.MiInt4()
    m.p = C.cast(ibuf, C.c_void_p).value
    lo, hi = (C.c_int * 3)(0, 0, 0), (C.c_int * 3)(1, 1, 1)
    rc = lib.mi_chem_advance_box(h, lo, hi, view(0, 23), view(0, 23), view(0, 22), view(0, 21), view(0, 1), m, 1e-6,
                                 None)
    assert rc == reactor.MI_ESTATE
    assert lib.mi_chem_init(h, 2, 1) == 0
    # drm19 (21 species): a new-state view with too few components
    rc = lib.mi_chem_advance_box(h, lo, hi, view(0, 23), view(8, 21), view(0, 22), view(0, 21), view(0, 1), m, 1e-6,
                                 None)
    assert rc == reactor.MI_EINVAL and b"cover" in lib.mi_chem_last_error(h)
    # old and new state in the same memory
    rc = lib.mi_chem_advance_box(h, lo, hi, view(0, 23), view(0, 23), view(0, 22), view(0, 21), view(0, 1), m, 1e-6,
                                 None)
    assert rc == reactor.MI_EINVAL and b"share memory" in lib.mi_chem_last_error(h)
    big = (C.c_double * (23 * 8))()
    other = view(0, 23)
    other.p = C.cast(big, C.c_void_p).value
    rc = lib.mi_chem_advance_box(h, lo, hi, view(0, 23), other, view(0, 22), view(0, 21), view(0, 1), m, -1.0, None)
    assert rc == reactor.MI_EINVAL and b"dt" in lib.mi_chem_last_error(h)
    assert lib.mi_chem_close(h) == 0
