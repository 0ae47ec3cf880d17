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


# Two-stage Miller op-amp on PSP103 (models/psp103.va, default model cards
# with TYPE = +1/-1), connected as a unity-gain buffer so its operating
# point is defined; source/drain sheet resistance RSH 10 ohm/sq with NRS =
# NRD = 1 gives every transistor its SI and DI rows besides NOI.
#
#     VDD   vdd 0 DC 1.8
#     VINP  inp 0 DC 0.9  (PULSE 0.9 -> 1.0 in the transient)
#     IBIAS vdd nbias DC 20u
#     M8 nbias nbias 0 0    nch W=2u  L=1u
#     M5 tail  nbias 0 0    nch W=4u  L=1u
#     M1 n1    out   tail 0 nch W=4u  L=0.5u   (inverting input: the output)
#     M2 n2    inp   tail 0 nch W=4u  L=0.5u
#     M3 n1    n1    vdd vdd pch W=8u  L=0.5u
#     M4 n2    n1    vdd vdd pch W=8u  L=0.5u
#     M6 out   n2    vdd vdd pch W=16u L=0.5u
#     M7 out   nbias 0 0    nch W=8u  L=1u
#     CC n2 out 1p
#     CL out 0  1p
OPAMP_DEVICES = (  # name, drain, gate, source, bulk, type, W, L, width group
    ("M8", "nbias", "nbias", "0", "0", 1, 2e-6, 1e-6, 3),
    ("M5", "tail", "nbias", "0", "0", 1, 4e-6, 1e-6, 3),
    ("M1", "n1", "out", "tail", "0", 1, 4e-6, 0.5e-6, 0),
    ("M2", "n2", "inp", "tail", "0", 1, 4e-6, 0.5e-6, 0),
    ("M3", "n1", "n1", "vdd", "vdd", -1, 8e-6, 0.5e-6, 1),
    ("M4", "n2", "n1", "vdd", "vdd", -1, 8e-6, 0.5e-6, 1),
    ("M6", "out", "n2", "vdd", "vdd", -1, 16e-6, 0.5e-6, 2),
    ("M7", "out", "nbias", "0", "0", 1, 8e-6, 1e-6, 3),
)
OPAMP_NOMINAL = dict(vdd=1.8, vcm=0.9, ibias=20e-6, cc=1e-12, cl=1e-12, temp_c=27.0)


def miller_opamp(add, index, vcm=0.9, step=None):
    """Cards of the op-amp; `index(name)` gives the slot of a named PSP103
    parameter (the engine's param_index or the oracle's). Returns the
    instance indices {'VDD', 'VINP', 'IBIAS', 'CC', 'CL', 'M1'..'M8'}.
    step: (v2, td, tr, pw) turns VINP into PULSE(vcm v2 td tr tr pw 1)."""
    from . import CAPACITOR, ISOURCE, PSP103
    inst = {}
    inst["VDD"] = add(VSOURCE, ["vdd", "0"], {VSOURCE_DC: OPAMP_NOMINAL["vdd"]})
    if step is None:
        inst["VINP"] = add(VSOURCE, ["inp", "0"], {VSOURCE_DC: vcm})
    else:
        v2, td, tr, pw = step
        p = pulse_params(vcm, v2, td, tr, tr, pw, 1.0, dc=vcm)
        inst["VINP"] = add(VSOURCE, ["inp", "0"], p)
    inst["IBIAS"] = add(ISOURCE, ["vdd", "nbias"], {0: OPAMP_NOMINAL["ibias"]})
    for name, d, g, s, b, ty, w, l, _ in OPAMP_DEVICES:
        p = {index("TYPE"): float(ty), index("W"): w, index("L"): l, index("RSH"): 10.0,
             index("NRS"): 1.0, index("NRD"): 1.0}
        inst[name] = add(PSP103, [d, g, s, b], p)
    inst["CC"] = add(CAPACITOR, ["n2", "out"], {0: OPAMP_NOMINAL["cc"]})
    inst["CL"] = add(CAPACITOR, ["out", "0"], {0: OPAMP_NOMINAL["cl"]})
    return inst


def opamp_inputs(n, seed=1):
    """The 8 Monte Carlo inputs per lane, uniform and seeded: width scale of
    the 4 device groups (diff pair, mirror, second stage, current sinks)
    in [0.9, 1.1], VDD in [1.6, 2.0] V, temperature in [-40, 125] degC, bias
    current scale in [0.8, 1.2], load capacitance in [0.1, 2] pF."""
    rng = np.random.default_rng(seed)
    lo = np.array([0.9, 0.9, 0.9, 0.9, 1.6, -40.0, 0.8, 0.1e-12])
    hi = np.array([1.1, 1.1, 1.1, 1.1, 2.0, 125.0, 1.2, 2.0e-12])
    return lo + (hi - lo) * rng.random((n, 8))


def opamp_lane_params(inputs, index):
    """Per-lane parameter columns of the op-amp for `inputs` (n x 8):
    [(kind, instance name, slot, values)]."""
    from . import CAPACITOR, ISOURCE, PSP103
    out = []
    for name, d, g, s, b, ty, w, l, grp in OPAMP_DEVICES:
        out.append((PSP103, name, index("W"), w * inputs[:, grp]))
        out.append((PSP103, name, index("TRISE"), inputs[:, 5] - 27.0))
    out.append((VSOURCE, "VDD", VSOURCE_DC, inputs[:, 4]))
    out.append((ISOURCE, "IBIAS", 0, OPAMP_NOMINAL["ibias"] * inputs[:, 6]))
    out.append((CAPACITOR, "CL", 0, inputs[:, 7]))
    return out


# A 31-stage ring oscillator on BSIM4 (bsim4va): NMOS W/L 1u/0.1u, PMOS
# 2u/0.1u, VTH0 +-0.3 V, TOXE 3 nm, 1 fF load per stage, VDD 1.2 V; a 20 ps
# 200 uA current pulse into n1 at 50 ps starts it from its (metastable)
# operating point, so that every lane oscillates from the first period.
RING_NOMINAL = dict(vdd=1.2, cload=1e-15, vth0=0.3, toxe=3e-9, wn=1e-6, wp=2e-6, l=0.1e-6,
                    kick=(200e-6, 0.05e-9, 5e-12, 20e-12))
# the period's two crossings: v(n1) rising through VDD/2, the first and the
# second after 0.2 ns (past the kick)
RING_MEAS = dict(val=0.6, td=0.2e-9)


def ring_oscillator(add, index, stages=31):
    """Cards of the ring (`index(name)`: the slot of a named BSIM4 parameter):

        VDD vdd 0 DC 1.2
        IKICK 0 n1 PULSE(0 200u 50p 5p 5p 20p 1)
        MN<k> n<k+1> n<k> 0 0 nch W=1u L=0.1u VTH0=0.3 TOXE=3n
        MP<k> n<k+1> n<k> vdd vdd pch W=2u L=0.1u VTH0=-0.3 TOXE=3n
        C<k> n<k+1> 0 1f                      (n<stages+1> is n1)

    Returns {'VDD', 'IKICK', 'MN1'.., 'MP1'.., 'C1'..} instance indices."""
    from . import BSIM4, ISOURCE
    R = RING_NOMINAL
    inst = {"VDD": add(VSOURCE, ["vdd", "0"], {VSOURCE_DC: R["vdd"]})}
    i, td, tr, pw = R["kick"]
    inst["IKICK"] = add(ISOURCE, ["0", "n1"], pulse_params(0.0, i, td, tr, tr, pw, 1.0, dc=0.0))
    for k in range(1, stages + 1):
        a, y = f"n{k}", f"n{k % stages + 1}"
        for name, ty, w, b in (("MN", 1.0, R["wn"], "0"), ("MP", -1.0, R["wp"], "vdd")):
            p = {index("TYPE"): ty, index("W"): w, index("L"): R["l"], index("VTH0"): ty * R["vth0"],
                 index("TOXE"): R["toxe"]}
            inst[f"{name}{k}"] = add(BSIM4, [y, a, b, b], p)
    for k in range(1, stages + 1):
        inst[f"C{k}"] = add(CAPACITOR, [f"n{k % stages + 1}", "0"], {0: R["cload"]})
    return inst


def ring_lane_params(n, index, stages=31, seed=1):
    """Per-lane Monte Carlo columns: every transistor's VTH0 ~ N(+-0.3 V,
    15 mV) and TOXE ~ N(3 nm, 2%), seeded: [(kind, instance name, slot,
    values)]."""
    from . import BSIM4
    R = RING_NOMINAL
    rng = np.random.default_rng(seed)
    out = []
    for k in range(1, stages + 1):
        for name, ty in (("MN", 1.0), ("MP", -1.0)):
            out.append((BSIM4, f"{name}{k}", index("VTH0"), ty * R["vth0"] + rng.normal(0.0, 0.015, n)))
            out.append((BSIM4, f"{name}{k}", index("TOXE"), R["toxe"] * (1.0 + rng.normal(0.0, 0.02, n))))
    return out
