"""AC lanes on the GPU (espice_lanes_ac) against the oracle's
linearization and LAPACK's complex solve (oracle/ac.py): every lane's
phasors at every frequency to a stated tolerance (the device factors in
its own operation order), and the reference's printed rc_lowpass.sp rows
(docs/learn/analyses.md:52-56) on the nominal lane."""
import os
import sys

import numpy as np
import pytest

import egspice_amd as E
import oracle_lib as O
from egspice_amd import workloads as W

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import ac as A  # noqa: E402

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def dec_axis(f0, f1, per_decade):
    k = np.arange(int(round(np.log10(f1 / f0) * per_decade)) + 1)
    return f0 * 10.0 ** (k / per_decade)


def test_ac_rc_lowpass_lanes(oracle, engine):
    N = 64
    say(f"case: AC lanes, rc_lowpass.sp, {N} lanes (R, C spread), 41 frequencies")
    r = np.linspace(0.5e3, 2e3, N)
    cap = np.linspace(100e-9, 200e-9, N)
    r[0], cap[0] = 1e3, 159e-9
    t = E.Template()
    v = t.add(E.VSOURCE, ["in", "0"], {E.VSOURCE_DC: 0.0})
    t.add(E.RESISTOR, ["in", "out"], {0: 1e3})
    t.add(E.CAPACITOR, ["out", "0"], {0: 159e-9})
    t.freeze()
    g = E.Lanes(t, N)
    g.set_param(E.RESISTOR, 0, 0, r)
    g.set_param(E.CAPACITOR, 0, 0, cap)
    g.prepare()
    g.upload_x(np.zeros((t.n, N)))
    freqs = dec_axis(10.0, 100e3, 10)
    freqs[1] = 1.2589254117941673e1
    omega = 2 * np.pi * freqs
    drive = t.ac_drive([(E.VSOURCE, v, 1.0, 0.0)])
    X, ms = g.ac(omega, drive)
    say(f"  kernel {ms:.3f} ms")
    c = O.Circuit(oracle)
    c.add(O.VSRC, ["in", "0"], {O.VSRC_DC: 0.0})
    c.add(O.RES, ["in", "out"], {0: 1e3})
    c.add(O.CAP, ["out", "0"], {0: 159e-9})
    c.freeze()
    o = O.Lanes(c, N)
    o.set_param(O.RES, 0, 0, r)
    o.set_param(O.CAP, 0, 0, cap)
    o.prepare()
    d = drive[:t.n] + 1j * drive[t.n:]
    worst = 0.0
    for l in range(N):
        G, Cm = o.eval_gc(l, np.zeros(c.n))
        Xo = A.solve(G, Cm, omega, d)
        # each unknown to 1e-14 of its row's scale (1 V, 1 V / R)
        scale = np.array([1.0, 1.0, 1.0 / r[l], 1.0])
        err = np.abs(X[:, :, l] - Xo) / scale[None, :]
        worst = max(worst, float(err.max()))
    say(f"  worst difference from the oracle {worst:.2e} of each unknown's scale")
    assert worst <= 1e-14
    # the reference's printed rows on lane 0 (R 1k, C 159n)
    rows = [(-9.979542742416352e-8, -9.989267655685869e-6, 9.999002045725759e-1, -9.98926765568587e-3),
            (-1.5815586205860516e-7, -1.2575008897885423e-5, 9.998418441379414e-1, -1.2575008897885423e-2)]
    for k, (ir, ii, vor, voi) in enumerate(rows):
        assert abs(X[k, 2, 0] - complex(ir, ii)) <= 1e-18
        assert abs(X[k, 3, 0] - complex(vor, voi)) <= 1e-15
        assert X[k, 1, 0] == 1.0


def test_ac_opamp_lanes(oracle, engine):
    """The psp103 op-amp's small-signal response (unity-gain buffer) on 16
    Monte Carlo lanes at their operating points (the ladder), 1 Hz to
    1 GHz: every lane against the oracle to 1e-10 of the largest phasor,
    and near-unity gain at low frequency."""
    N = 16
    say(f"case: AC lanes, psp103 op-amp, {N} lanes, 37 frequencies")
    inputs = W.opamp_inputs(N, seed=4)
    t = E.Template()
    eidx = lambda nm: E.param_index(E.PSP103, nm)
    inst = W.miller_opamp(t.add, eidx)
    t.freeze()
    g = E.Lanes(t, N)
    for kind, name, slot, vals in W.opamp_lane_params(inputs, eidx):
        g.set_param(kind, inst[name], slot, vals)
    g.prepare()
    xg, sg, _, rung, _ = g.op_ladder()
    assert (sg == 1).all()
    omega = 2 * np.pi * dec_axis(1.0, 1e9, 4)
    drive = t.ac_drive([(E.VSOURCE, inst["VINP"], 1.0, 0.0)])
    X, ms = g.ac(omega, drive)
    say(f"  kernel {ms:.3f} ms")
    oidx = lambda nm: oracle.or_va_param_index(O.PSP103, nm.encode())
    c = O.Circuit(oracle)
    oinst = W.miller_opamp(c.add, oidx)
    c.freeze()
    o = O.Lanes(c, N)
    for kind, name, slot, vals in W.opamp_lane_params(inputs, oidx):
        o.set_param(kind, oinst[name], slot, vals)
    o.prepare()
    d = drive[:t.n] + 1j * drive[t.n:]
    out = t.node("out")
    worst = 0.0
    for l in range(N):
        G, Cm = o.eval_gc(l, xg[:, l])
        Xo = A.solve(G, Cm, omega, d)
        err = np.abs(X[:, :, l] - Xo).max() / np.abs(Xo).max()
        worst = max(worst, float(err))
        assert abs(abs(X[0, out, l]) - 1.0) < 0.05
    say(f"  worst difference from the oracle {worst:.2e} of the largest phasor")
    assert worst <= 1e-10
