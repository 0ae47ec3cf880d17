"""The AC lanes' oracle (oracle/ac.py over the oracle's G and C planes)
against the reference's printed AC rows of docs/examples/rc_lowpass.sp
(docs/learn/analyses.md:52-56), the first-order low-pass's closed form,
and the reference's excitation known answer (frontend/tests/builder.zig:
574-593), the last on the engine's own drive builder too (its template
code runs on the host)."""
import os
import sys

import numpy as np

import egspice_amd as E
import oracle_lib as O

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import ac as A  # noqa: E402

# docs/learn/analyses.md:55-56: index, frequency, i(v1) re/im, v(in) re/im,
# v(out) re/im
RC_AC_ROWS = [
    (1e1, -9.979542742416352e-8, -9.989267655685869e-6, 1.0, 0.0, 9.999002045725759e-1, -9.98926765568587e-3),
    (1.2589254117941673e1, -1.5815586205860516e-7, -1.2575008897885423e-5, 1.0, 0.0, 9.998418441379414e-1,
     -1.2575008897885423e-2),
]


def rc_lowpass(add, r=1e3, c=159e-9):
    add(O.VSRC, ["in", "0"], {O.VSRC_DC: 0.0})
    add(O.RES, ["in", "out"], {0: r})
    add(O.CAP, ["out", "0"], {0: c})


def test_rc_lowpass_printed_rows(oracle):
    c = O.Circuit(oracle)
    rc_lowpass(c.add)
    c.freeze()
    L = O.Lanes(c, 1)
    L.prepare()
    x = np.zeros(c.n)
    G, Cm = L.eval_gc(0, x)
    # rows: in 1, V1's branch 2, out 3; V1 has AC 1
    drive = A.excitation(c.n, [(0, 2, 1.0, 0.0)])
    # to a few ulps of each unknown's scale (1 V; 1 V / 1 kohm for the
    # current, whose real part is a difference of near-equal voltages)
    for f, ir, ii, vir, vii, vor, voi in RC_AC_ROWS:
        X = A.solve(G, Cm, 2 * np.pi * f, drive)[0]
        for got, want, scale in ((X[2].real, ir, 1e-3), (X[2].imag, ii, 1e-3), (X[1].real, vir, 1.0),
                                 (X[1].imag, vii, 1.0), (X[3].real, vor, 1.0), (X[3].imag, voi, 1.0)):
            assert abs(got - want) <= 1e-15 * scale, (f, got, want)


def test_rc_lowpass_closed_form(oracle):
    c = O.Circuit(oracle)
    rc_lowpass(c.add, r=2.2e3, c=47e-9)
    c.freeze()
    L = O.Lanes(c, 1)
    L.prepare()
    G, Cm = L.eval_gc(0, np.zeros(c.n))
    f = np.logspace(0, 6, 25)
    X = A.solve(G, Cm, 2 * np.pi * f, A.excitation(c.n, [(0, 2, 1.0, 0.0)]))
    H = 1.0 / (1.0 + 1j * 2 * np.pi * f * 2.2e3 * 47e-9)
    assert np.allclose(X[:, 3], H, rtol=1e-13, atol=0)
    # the source's current: -(v_in - v_out) / R
    assert np.allclose(X[:, 2], -(1.0 - H) / 2.2e3, rtol=1e-12, atol=0)


def test_excitation_known_answer():
    """V1 b 0 0 AC 1 90, I1 0 a AC 2, R1 a 0 1, R2 b 0 1: 2 A pushed into
    a, V1's phasor j on its branch row, |drive| summing to 3."""
    t = E.Template()
    v = t.add(E.VSOURCE, ["b", "0"], {E.VSOURCE_DC: 0.0})
    i = t.add(E.ISOURCE, ["0", "a"], {0: 0.0})
    t.add(E.RESISTOR, ["a", "0"], {0: 1.0})
    t.add(E.RESISTOR, ["b", "0"], {0: 1.0})
    t.freeze()
    d = t.ac_drive([(E.VSOURCE, v, 1.0, 90.0), (E.ISOURCE, i, 2.0, 0.0)])
    n = t.n
    a, br = t.node("a"), 2  # rows: b 1, V1's branch 2, a 3
    assert d[a] == 2.0
    assert abs(d[br]) <= 1e-15
    assert abs(d[n + br] - 1.0) <= 1e-15
    assert abs(np.abs(d).sum() - 3.0) <= 3e-12
    o = A.excitation(n, [(0, br, 1.0, 90.0), (0, a, 2.0, 0.0)])
    assert np.array_equal(o.real, d[:n]) and np.array_equal(o.imag, d[n:])
