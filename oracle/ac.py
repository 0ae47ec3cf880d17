"""CPU oracle of the AC lanes (test infrastructure only: nothing in the
product imports it). src/analysis/ac/ac.zig:40-86: one linearization at
the operating point gives G and C; each frequency solves
(G + j omega C) X = drive, the drive being the deck's AC excitation
(frontend/builder.zig:802-817: a voltage source's mag e^(j phase) on its
branch row, a current source's minus on p and plus on n). The solve here
is LAPACK's (numpy), a checker of the device's own complex LU to a stated
tolerance, not its operation order."""
import numpy as np


def excitation(n, sources):
    """drive (complex n) from [(pos, neg, mag, phase_deg)] rows of the
    reference's acExcitation: pos/neg rows, 0 = ground (skipped)."""
    d = np.zeros(n, dtype=complex)
    for pos, neg, mag, phase in sources:
        rad = phase * (np.pi / 180.0)
        v = mag * np.cos(rad) + 1j * mag * np.sin(rad)
        if pos != 0:
            d[pos] -= v
        if neg != 0:
            d[neg] += v
    return d


def solve(G, C, omega, drive):
    """X (frequencies, n): (G + j omega_k C) X_k = drive for each omega_k."""
    G = np.asarray(G, dtype=float)
    C = np.asarray(C, dtype=float)
    return np.array([np.linalg.solve(G + 1j * w * C, drive) for w in np.atleast_1d(omega)])
