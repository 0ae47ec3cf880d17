"""The synthetic sweep workloads, written as the card sequences a deck would
give, so the engine's template and any other builder with the same
``add(kind, ports, params)`` / ``node(name)`` interface number nodes alike
(first touch, in card order).
"""
import numpy as np

from . import MOS1, MOS1_PARAMS, VSOURCE, VSOURCE_DC

M = MOS1_PARAMS

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
