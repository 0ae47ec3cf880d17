"""The deck front end and the points API without a GPU: ngspice's number
reading against the reference's known answers (src/frontend/lines.zig:890-925),
the documented decks read and built (create then stops at the missing GPU),
cards outside the subset refused by name, and the API's argument and
handle checks (include/espice.h:139-156, 325-339)."""
import ctypes as C
import os

import numpy as np
import pytest

import egspice_amd as E
import decks

NO_GPU = not os.path.exists("/dev/kfd")


@pytest.mark.parametrize("text,want", [
    ("10meg", 1e7), ("0.1n", 1e-10), ("1m", 1e-3), ("1ms", 1e-3), ("1k", 1000.0), ("1e3", 1000.0),
    ("1e-3k", 1.0), ("+.5", 0.5), ("-3", -3.0), ("5v", 5.0), ("2a", 2e-18), ("1d3", 1000.0), ("1x", 1.0),
    ("1e23", 1.0000000000000001e23), ("1e-300", 1e-300), ("1e300", 1e300), ("3e45", 3.0 * 1e45),
    ("1e210", 1.0000000000000001e210), ("1e-331", 0.0), ("0e400", 0.0), ("1e400", float("inf")),
])
def test_number_known_answers(engine, text, want):
    assert E.spice_number(text) == want


def test_number_rejects_and_mil(engine):
    for bad in ("", "-", ".", "abc", "e5", "1e+"):
        assert E.spice_number(bad) is None
    assert E.spice_number("1mil") == pytest.approx(25.4e-6, rel=1e-15)
    # one mantissa of every digit, then the power: 0.1n is 1e-10, not
    # float("0.1") * 1e-9
    assert E.spice_number("0.1n") != 0.1 * 1e-9


@pytest.mark.skipif(not NO_GPU, reason="a GPU is visible: the documented decks run in the gpu tests")
@pytest.mark.parametrize("deck", [decks.MEASURE_SP, decks.RC_LOWPASS_SP, decks.CMOS_INVERTER_SP,
                                  decks.MEASURE_PARAM_SP])
def test_documented_decks_read_and_build(engine, deck):
    """Reading, building the template and planning the .meas passes all
    succeed; only the GPU is missing."""
    live = ("r", "c") if "{r}" in deck else ()
    with pytest.raises(E.PointsError) as e:
        E.Problem(deck, live=live)
    assert e.value.status == 6 and e.value.name == "GpuUnavailable"


@pytest.mark.skipif(not NO_GPU, reason="a GPU is visible: the deck runs in the gpu tests")
def test_isource_waveforms_read_and_build(engine):
    """I cards take the V cards' waveforms (models/isource.va): the deck
    reads and builds; only the GPU is missing."""
    with pytest.raises(E.PointsError) as e:
        E.Problem(decks.ISOURCES_SP)
    assert e.value.status == 6 and e.value.name == "GpuUnavailable"


@pytest.mark.parametrize("card,name", [
    ("L1 in out 1u", "UnsupportedElement"),
    ("X1 in out sub", "UnsupportedElement"),
    ("I2 a 0 SFFM(0 1 1k 2 10k)", "UnsupportedWaveform"),
    ("V2 a 0 SFFM(0 1 1k 2 10k 0 0 0 1)", "BadSffm"),
    ("V2 a 0 AM(1 0 1k 10k 0 0 0 1)", "UnsupportedCard"),
    ("V2 a 0 AM(0 1 0.5 1k 10k 0 0 0 1)", "BadAm"),
    ("V2 a 0 AM(0 {r} 0.5 1k 10k)", "UnresolvedParameter"),
    ("V2 a 0 PULSE(0 1 1u 1n 1n 0.5u 1.5u 5)", "PwlTooLong"),
    ("V2 a 0 PULSE(0 1 1u 1n 1n 1u 2u 3 1)", "BadPulse"),
    ("V2 a 0 PWL(0 0 1u 1 2u 0) r=1.5u", "InvalidParameterValue"),
    ("V2 a 0 PWL(0 0 1u 1 2u 0) r=2u", "InvalidParameterValue"),
    ("V2 a 0 EXP(0 1 1n 1n 2n 1n 7)", "BadExp"),
    ("V2 a 0 SIN(0 1 1k 0 0 0 9)", "BadSin"),
    (".subckt sub a b", "UnsupportedCard"),
    (".options reltol=1e-4", "UnsupportedCard"),
    (".ic v(out)=1", "UnsupportedCard"),
    (".include other.sp", "UnsupportedCard"),
    (".meas tran bad FIND v(out) WHEN v(in)=0.5 TD=1u", "UnsupportedMeasure"),
    (".meas tran bad FIND v(out) WHEN v(in)=0.5 RISE=LAST", "UnsupportedMeasure"),
    (".meas ac gain MAX v(out)", "UnsupportedMeasureAnalysis"),
    (".model q1 NPN(BF=100)", "UnsupportedModel"),
    (".model m3 NMOS(LEVEL=3)", "UnsupportedModelLevel"),
])
def test_cards_outside_the_subset_are_refused(engine, card, name):
    text = decks.MEASURE_SP.replace(".end", card + "\n.end")
    with pytest.raises(E.PointsError) as e:
        E.Problem(text)
    assert e.value.status == 1 and e.value.name.startswith(name), e.value.name


def test_tran_uic_and_bad_vectors_are_refused(engine):
    with pytest.raises(E.PointsError, match="UicUnsupported"):
        E.Problem(decks.MEASURE_SP.replace(".tran 10n 10u", ".tran 10n 10u uic"))
    with pytest.raises(E.PointsError, match="NoSuchVector"):
        E.Problem(decks.MEASURE_SP.replace("FIND v(out) AT=3u", "FIND v(nowhere) AT=3u"))
    with pytest.raises(E.PointsError, match="NotAParameter"):
        E.Problem(decks.MEASURE_PARAM_SP, live=("r", "nope"))
    with pytest.raises(E.PointsError, match="NoAnalysis"):
        E.Problem("title\nR1 a 0 1k\n.end\n")


def test_create_argument_checks(engine):
    lib = E._points_lib()
    assert lib.espice_abi_version() == 1
    out = C.c_void_p()
    diag = C.create_string_buffer(64)
    assert lib.espice_create(None, C.byref(out), diag, 64) == 1
    o = E.create_options(decks.MEASURE_SP)
    o.abi_version = 2
    assert lib.espice_create(C.byref(o), C.byref(out), diag, 64) == 7 and not out.value
    o = E.create_options(decks.MEASURE_SP)
    o.struct_size = 40
    assert lib.espice_create(C.byref(o), C.byref(out), diag, 64) == 7
    o = E.create_options(decks.MEASURE_SP)
    o.max_parallel = 0
    assert lib.espice_create(C.byref(o), C.byref(out), diag, 64) == 1
    with pytest.raises(E.PointsError, match="UnsupportedDialect"):
        E.Problem(options=E.create_options(decks.MEASURE_SP, dialect=E.ESPICE_HSPICE))
    with pytest.raises(E.PointsError, match="BackendUnavailable"):
        E.Problem(decks.MEASURE_SP, backend=E.ESPICE_CPU)
    # a deck file that does not exist
    with pytest.raises(E.PointsError, match="FileNotFound"):
        E.Problem(path="/nonexistent/deck.sp")


def test_handles_never_returned_are_invalid(engine):
    lib = E._points_lib()
    assert lib.espice_destroy(None) == 0
    for fake in (1, 0x1001, 0xdeadbeef):
        assert lib.espice_destroy(C.c_void_p(fake)) == 8
        n = C.c_uint32()
        assert lib.espice_query_count(C.c_void_p(fake), C.byref(n)) == 8
        assert lib.espice_run_all(C.c_void_p(fake)) == 8
