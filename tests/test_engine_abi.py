"""The C ABI of the lane engine without a GPU: the library loads and exports
every entry point include/espice_lanes.h and include/espice_mlp.h declare, the template calls (host
only) give the oracle's pattern and column order, argument errors come back
as status codes with a message, and the host tape build gives the oracle's
tape words. No compute call touches a device here."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import egspice_amd as E
import oracle_lib as O
from egspice_amd import workloads as W

INCLUDE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include")
HEADERS = [os.path.join(INCLUDE, h) for h in ("espice_lanes.h", "espice_mlp.h")]


def declared():
    text = "".join(open(h).read() for h in HEADERS)
    return sorted(set(re.findall(r"\bespice_status\s+(espice_(?:lanes|mlp)_\w+)\s*\(", text)))


def test_every_declared_symbol_is_exported(engine):
    names = declared()
    assert len(names) >= 18
    assert sorted(E.EXPORTS) == names
    for name in names:
        assert hasattr(engine, name), name


def test_template_matches_oracle_pattern(oracle, engine):
    for stages in (1, 5, 9):
        t = E.Template()
        W.inverter_chain(t.add, stages)
        t.freeze()
        c = O.Circuit(oracle)
        W.inverter_chain(c.add, stages)
        c.freeze()
        assert (t.n, t.nnz) == (c.n, c.nnz)
        assert np.array_equal(t.col_ptr, c.col_ptr) and np.array_equal(t.row_idx, c.row_idx)
        assert np.array_equal(t.diag, c.diag) and np.array_equal(t.cur_row, c.cur_row)
        s = oracle.or_solver_new(c.n, O.u32(c.col_ptr)[1], O.u32(c.row_idx)[1])
        q = np.ctypeslib.as_array(oracle.or_solver_q(s), (c.n,)).copy()
        oracle.or_solver_free(s)
        assert np.array_equal(t.q, q)
        assert not t.tridiagonal


def test_rc_template_rows_and_frozen_names(oracle, engine):
    """The RC decks number nodes alike in both builders, and a frozen
    template still answers for the names it holds (and only those)."""
    for builder, kw in ((W.rc_pulse, {}), (W.rc_ladder, {"stages": 4})):
        t = E.Template()
        builder(t.add, **kw)
        t.freeze()
        c = O.Circuit(oracle)
        builder(c.add, **kw)
        c.freeze()
        assert np.array_equal(t.col_ptr, c.col_ptr) and np.array_equal(t.row_idx, c.row_idx)
        name = "out" if builder is W.rc_pulse else "n4"
        assert t.node(name) == c.node(name)
        with pytest.raises(E.EngineError, match="TemplateFrozen"):
            t.node("not_there")


def status_of(fn, *args):
    return fn(*args)


def message(engine):
    buf = C.create_string_buffer(256)
    req = C.c_size_t()
    assert engine.espice_lanes_error_message(buf, 256, C.byref(req)) == 0
    return buf.value.decode()


def test_argument_errors_are_reported(engine):
    t = E.Template()
    rows = np.array([t.node("a"), 0], np.uint32)
    prm = np.full(3, np.nan)
    inst = C.c_uint32()
    # a resistor takes one parameter, not three
    st = engine.espice_lanes_template_add(t.h, E.RESISTOR, rows.ctypes.data_as(C.POINTER(C.c_uint32)),
                                          prm.ctypes.data_as(C.POINTER(C.c_double)), 3, C.byref(inst))
    assert st == 1 and message(engine)
    # unknown kind
    st = engine.espice_lanes_template_add(t.h, 17, rows.ctypes.data_as(C.POINTER(C.c_uint32)),
                                          prm.ctypes.data_as(C.POINTER(C.c_double)), 1, C.byref(inst))
    assert st == 1
    # null handles
    assert engine.espice_lanes_template_freeze(None) != 0
    assert engine.espice_lanes_prepare(None) != 0
    with pytest.raises(E.EngineError):
        t.add(E.RESISTOR, ["a", "b"], {0: 1.0, 5: 2.0})


def test_error_message_buffer_protocol(engine):
    t = E.Template()
    engine.espice_lanes_template_freeze(None)
    req = C.c_size_t()
    assert engine.espice_lanes_error_message(None, 0, C.byref(req)) in (0, 2)
    assert req.value > 0
    small = C.create_string_buffer(2)
    assert engine.espice_lanes_error_message(small, 2, C.byref(req)) == 2


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is visible: the device path is covered by the gpu tests")
def test_create_without_a_gpu_fails_loudly(engine):
    t = E.Template()
    W.inverter_chain(t.add)
    t.freeze()
    with pytest.raises(E.EngineError, match="NoGpu"):
        E.Lanes(t, 8)


def test_tab_build_words_equal_oracle(oracle, engine):
    from test_oracle_lu import random_system
    rng = np.random.default_rng(0x7AB)
    built = 0
    for trial in range(60):
        n = 2 + trial % 24
        cp, ri, v = random_system(rng, n, 0.3)
        s = oracle.or_solver_new(n, O.u32(cp)[1], O.u32(ri)[1])
        ok = oracle.or_solver_factor(s, O.f64(v)[1]) == 0 and not oracle.or_is_tridiag(n, O.u32(cp)[1], O.u32(ri)[1])
        if not ok:
            oracle.or_solver_free(s)
            with pytest.raises(E.EngineError):
                E.tab_build(n, cp, ri, v)
            continue
        otab = O.LaneTab()
        ln = C.c_size_t()
        p = oracle.or_build_tab(oracle.or_solver_lu(s), O.u32(cp)[1], oracle.or_solver_q(s), 1e-12, 1e-3,
                                C.byref(otab), C.byref(ln))
        oidx = np.ctypeslib.as_array(p, (ln.value,)).copy()
        oracle.free_(C.cast(p, C.c_void_p))
        oracle.or_solver_free(s)
        tab, idx = E.tab_build(n, cp, ri, v)
        assert np.array_equal(idx, oidx)
        for f, _ in O.LaneTab._fields_:
            if f != "_pad":
                assert getattr(tab, f) == getattr(otab, f), f
        built += 1
    assert built >= 20
