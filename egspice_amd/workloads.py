"""The synthetic sweep workloads, written as the card sequences a deck would
give, so the engine's template and any other builder with the same
``add(kind, ports, params)`` / ``node(name)`` interface number nodes alike
(first touch, in card order).
"""
import numpy as np

from . import CAPACITOR, DIODE, DIODE_PARAMS, MOS1, MOS1_PARAMS, RESISTOR, VSOURCE, VSOURCE_DC

M = MOS1_PARAMS
D = DIODE_PARAMS

# 5-stage CMOS inverter chain, mos1 (LEVEL=1): NMOS KP 1e-4, VTO 0.7,
# W/L 2u/1u; PMOS KP 4e-5, VTO -0.7, W/L 4u/1u; VDD 1.8 V.
NMOS = {M["TYPE"]: 1.0, M["KP"]: 1e-4, M["VTO"]: 0.7, M["W"]: 2e-6, M["L"]: 1e-6}
PMOS = {M["TYPE"]: -1.0, M["KP"]: 4e-5, M["VTO"]: -0.7, M["W"]: 4e-6, M["L"]: 1e-6}


def inverter_chain(add, stages=5, vdd=1.8):
    """Cards of the chain; returns the instance index of the swept VIN.

        VDD vdd 0 DC 1.8
        VIN in 0 DC 0
        MN<k> n<k> n<k-1> 0 0 nch W=2u L=1u
        MP<k> n<k> n<k-1> vdd vdd pch W=4u L=1u
    """
    add(VSOURCE, ["vdd", "0"], {VSOURCE_DC: vdd})
    vin = add(VSOURCE, ["in", "0"], {VSOURCE_DC: 0.0})
    prev = "in"
    for k in range(1, stages + 1):
        out = f"n{k}"
        add(MOS1, [out, prev, "0", "0"], NMOS)
        add(MOS1, [out, prev, "vdd", "vdd"], PMOS)
        prev = out
    return vin


def dc_sweep_points(n_points=64, vmax=1.8):
    """Input values of the chain's sweep: vin_k = vmax * k / (n_points - 1)."""
    return vmax * np.arange(n_points) / (n_points - 1)


def pulse_params(v1, v2, td, tr, tf, pw, per, dc=None):
    """V source parameters of `PULSE(v1 v2 td tr tf pw per)` (models/vsource.va
    order: dc, wave, v1, v2, td, tr, tf, pw, per); dc None reads as not given."""
    return {0: np.nan if dc is None else dc, 1: 1.0, 2: v1, 3: v2, 4: td, 5: tr, 6: tf, 7: pw, 8: per}


def sin_params(vo, va, freq=None, td=None, theta=None, phase=None, dc=None):
    """V source parameters of `SIN(vo va freq td theta phase)` (models/vsource.va
    slots 9-14); arguments left None read as not given."""
    g = lambda v: np.nan if v is None else v
    return {0: g(dc), 1: 2.0, 9: vo, 10: va, 11: g(freq), 12: g(td), 13: g(theta), 14: g(phase)}


# docs/examples/measure.sp: V1 in 0 PULSE(0 1 1u 1n 1n 5u 20u); R1 in out 1k;
# C1 out 0 1n; .tran 10n 10u
MEASURE_PULSE = (0.0, 1.0, 1e-6, 1e-9, 1e-9, 5e-6, 20e-6)
# docs/examples/rc_lowpass.sp: V1 in 0 DC 0 PULSE(0 1 0 1u 1u 5m 10m);
# R1 in out 1k; C1 out 0 159n; .tran 10u 5m
RC_LOWPASS_PULSE = (0.0, 1.0, 0.0, 1e-6, 1e-6, 5e-3, 10e-3)


def rc_pulse(add, r=1e3, c=1e-9, pulse=MEASURE_PULSE, dc=None):
    """V1 in 0 PULSE(...); R1 in out r; C1 out 0 c, in the order a deck's
    builder adds elements (by letter: v, then c, then r;
    src/frontend/builder.zig:886). Returns the capacitor's instance index."""
    add(VSOURCE, ["in", "0"], pulse_params(*pulse, dc=dc))
    ci = add(CAPACITOR, ["out", "0"], {0: c})
    add(RESISTOR, ["in", "out"], {0: r})
    return ci


def rc_ladder(add, stages=3, r=1e3, c=1e-9, pulse=MEASURE_PULSE):
    """V1 in 0 PULSE(...); R<k> n<k-1> n<k> r; C<k> n<k> 0 c (n0 = in),
    added v, c, r like rc_pulse."""
    add(VSOURCE, ["in", "0"], pulse_params(*pulse))
    for k in range(1, stages + 1):
        add(CAPACITOR, [f"n{k}", "0"], {0: c})
    prev = "in"
    for k in range(1, stages + 1):
        add(RESISTOR, [prev, f"n{k}"], {0: r})
        prev = f"n{k}"


# docs/examples/cmos_inverter.sp: .model nch NMOS(LEVEL=1 VTO=0.45 KP=200u
# LAMBDA=0.05 GAMMA=0.4 PHI=0.8), pch the same with VTO=-0.45 KP=80u.
INV_NCH = {M["TYPE"]: 1.0, M["VTO"]: 0.45, M["KP"]: 200e-6, M["LAMBDA"]: 0.05, M["GAMMA"]: 0.4, M["PHI"]: 0.8}
INV_PCH = {M["TYPE"]: -1.0, M["VTO"]: -0.45, M["KP"]: 80e-6, M["LAMBDA"]: 0.05, M["GAMMA"]: 0.4, M["PHI"]: 0.8}


def cmos_inverter(add, vdd=1.8, wn=1e-6, cl=10e-15):
    """docs/examples/cmos_inverter.sp in builder order (v, c, m):

        VDD vdd 0 {vdd}
        VIN in 0 DC 0 PULSE(0 {vdd} 1n 100p 100p 5n 10n)
        CL out 0 10f
        M1 out in 0 0 nch W={wn} L=180n
        M2 out in vdd vdd pch W={2*wn} L=180n

    Returns the instance indices (vin, cl, m1, m2) per kind."""
    add(VSOURCE, ["vdd", "0"], {VSOURCE_DC: vdd})
    vin = add(VSOURCE, ["in", "0"], pulse_params(0.0, vdd, 1e-9, 100e-12, 100e-12, 5e-9, 10e-9, dc=0.0))
    c = add(CAPACITOR, ["out", "0"], {0: cl})
    m1 = add(MOS1, ["out", "in", "0", "0"], {**INV_NCH, M["W"]: wn, M["L"]: 180e-9})
    m2 = add(MOS1, ["out", "in", "vdd", "vdd"], {**INV_PCH, M["W"]: 2 * wn, M["L"]: 180e-9})
    return vin, c, m1, m2


DIODE_RECT = {D["IS"]: 1e-14, D["N"]: 1.05, D["RS"]: 0.5, D["CJO"]: 10e-12}


def diode_rectifier(add, rload=1e3, cres=100e-6, dmodel=None):
    """docs/examples/diode_rectifier.sp in builder order (v, c, d, r):

        V1 in 0 SIN(0 10 50)
        C1 out 0 {cres}
        D1 in out dmod      .model dmod D(IS=1e-14 N=1.05 RS=0.5 CJO=10p)
        R1 out 0 {rload}

    The diode's RS puts its junction on an internal row. Returns the
    instance indices (v1, c1, d1, r1) per kind."""
    v = add(VSOURCE, ["in", "0"], sin_params(0.0, 10.0, 50.0))
    c = add(CAPACITOR, ["out", "0"], {0: cres})
    d = add(DIODE, ["in", "out"], dict(DIODE_RECT if dmodel is None else dmodel))
    r = add(RESISTOR, ["out", "0"], {0: rload})
    return v, c, d, r
