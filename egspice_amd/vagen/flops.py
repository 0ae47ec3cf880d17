"""FP64 operation count of a compiled model's evaluation along the paths a
bias point actually takes (measurement support for the roofline).

`counting_source(header)` rewrites the generated C so that every executed
statement adds its arithmetic to a counter: one per +, -, *, / between
reals, two per fused multiply-add, and a fixed cost per elementary-function
call, the operation counts of the engine's own bodies (egspice_amd/csrc/
emath.h): exp 30 (Cody-Waite reduction 4, degree-13 Horner 26), log 30
(reduction and the atanh-series polynomial), pow 62 (log, multiply, exp and
the special-case tests), sqrt 1. Comparisons, selects and integer work are
not counted.

    python -m egspice_amd.vagen.flops HEADER PREFIX
compiles the counting variant with gcc and prints the count of one static,
one transient and one charge-only evaluation: for psp103 at the biases of
the two-stage op-amp's eight transistors (their nominal operating point),
for bsim4va at an inverter's transistors over its input swing, as the
bench's rooflines use.
"""
import os
import re
import subprocess
import sys
import tempfile

COST = {"VA_EXP(": 30, "VA_LOG(": 30, "VA_LOG10(": 31, "VA_POW(": 62, "VA_SQRT(": 1, "VA_FMA(": 2,
        "VA_TANH(": 40, "VA_ATAN(": 40, "VA_SINH(": 40, "VA_COSH(": 40, "VA_SIN(": 40, "VA_COS(": 40}
_binop = re.compile(r"(?<=[\w\])]) [-+*/] (?=[\w(-])")


def statement_flops(line):
    s = line.strip()
    if not s or s.startswith(("#", "//", "{", "}", "for", "if", "double ", "int ")) and "=" not in s.split("(")[0]:
        return 0
    if s.startswith("const int") or s.startswith("int "):
        return 0
    n = 0
    for k, c in COST.items():
        n += c * s.count(k)
    body = s.split("=", 1)[1] if "=" in s else s
    n += len(_binop.findall(body))
    if re.search(r"\[\d+\] [-+]= ", s):
        n += 1
    return n


def counting_source(text):
    out = []
    for ln in text.split("\n"):
        out.append(ln)
        s = ln.strip()
        if s.endswith(";") and not s.startswith(("return", "#")):
            k = statement_flops(ln)
            if k:
                out.append(f"va_flops += {k};")
    return "\n".join(out)


DRIVER = r'''
#include <math.h>
#include <string.h>
static unsigned long long va_flops;
#define VA_FN static
#define VA_TABLE static const
#define VA_HOST_TABLE static const
#define VA_RESTRICT
#define VA_EXP exp
#define VA_LOG log
#define VA_LOG10 log10
#define VA_POW pow
#define VA_SQRT sqrt
#define VA_FABS fabs
#define VA_FLOOR floor
#define VA_CEIL ceil
#define VA_FMA fma
#define VA_TANH tanh
#define VA_ATAN atan
#define VA_SINH sinh
#define VA_COSH cosh
#define VA_SIN sin
#define VA_COS cos
#define VA_INF INFINITY
#define VA_IDIV(a,b) ((b) ? (a)/(b) : 0)
#define VA_IMOD(a,b) ((b) ? (a)%(b) : 0)
#define VA_RTOI(x) ((int)round(x))
#define VA_FETLIM(a, b, c) (a)
#define VA_LIMVDS(a, b) (a)
#define VA_PNJLIM(a, b, c, d, e) (a)
#define VA_SNAP(a, b, c, d, e)
#include "COUNTED"
/* mode 0 static, 1 transient, 2 charges alone */
int count(const double *u, const double *x, int mode, unsigned long long *out)
{
    double prep[PFX_NPREP], res[PFX_NU], jac[PFX_NJ + 1], q[PFX_NU], cj[PFX_NCJ + 1], site[PFX_NSITE + 1];
    int rc = pfx_setup(u, 300.15, 1e-12, prep);
    if (rc) return rc;
    va_flops = 0;
    if (mode == 1) pfx_eval_tran(prep, x, res, jac, q, cj, site);
    else if (mode == 2) pfx_eval_charge(prep, x, q, site);
    else pfx_eval(prep, x, res, jac);
    *out = va_flops;
    return 0;
}
'''


def build(header, prefix, workdir):
    with open(header) as f:
        text = f.read()
    counted = os.path.join(workdir, "counted.h")
    with open(counted, "w") as f:
        f.write(counting_source(text))
    src = os.path.join(workdir, "driver.c")
    with open(src, "w") as f:
        f.write(DRIVER.replace("COUNTED", counted).replace("PFX_", prefix.upper() + "_").replace("pfx_", prefix + "_"))
    lib = os.path.join(workdir, "libcount.so")
    subprocess.run(["gcc", "-O0", "-shared", "-fPIC", "-o", lib, src, "-lm"], check=True)
    return lib


def _names(header, prefix):
    text = open(header).read()
    names = re.search(r"char \*const " + prefix + r"_param_names\[\d+\] = \{(.*?)\};", text, re.S).group(1)
    names = [x.strip().strip('"') for x in names.replace("\n", " ").split(",") if x.strip()]
    return {n.upper(): i for i, n in enumerate(names)}, len(names)


def _mean_counts(lib, cases, nparam):
    """cases: [(user values {slot: v}, unknowns x)] -> mean (static, transient, charge) counts."""
    import ctypes as C
    import numpy as np
    acc = [[], [], []]
    for user, x in cases:
        u = np.full(nparam, np.nan)
        for k, v in user.items():
            u[k] = v
        x = np.ascontiguousarray(x, dtype=np.float64)
        for mode in range(3):
            out = C.c_ulonglong()
            rc = lib.count(u.ctypes.data_as(C.c_void_p), x.ctypes.data_as(C.c_void_p), mode, C.byref(out))
            assert rc == 0
            acc[mode].append(out.value)
    return tuple(sum(a) / len(a) for a in acc)


def ring_counts(header, prefix="bsim4va"):
    """Mean counts over an inverter's two BSIM4 transistors (the ring
    oscillator's devices, every internal unknown on its port) at 7 input
    voltages across the 1.2 V swing, the output its complement."""
    import ctypes as C
    root = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    sys.path.insert(0, root)
    from egspice_amd import workloads as W
    R = W.RING_NOMINAL
    with tempfile.TemporaryDirectory() as d:
        lib = C.CDLL(build(header, prefix, d))
        index, nparam = _names(header, prefix)
        cases = []
        for vin in [0.1 * k + 0.3 for k in range(7)]:
            vout = R["vdd"] - vin
            for ty, w, s, b in ((1.0, R["wn"], 0.0, 0.0), (-1.0, R["wp"], R["vdd"], R["vdd"])):
                user = {index["TYPE"]: ty, index["W"]: w, index["L"]: R["l"], index["VTH0"]: ty * R["vth0"],
                        index["TOXE"]: R["toxe"]}
                # d g s b di si gi gm bi sbulk dbulk
                x = [vout, vin, s, b, vout, s, vin, vin, b, b, b]
                cases.append((user, x))
        return _mean_counts(lib, cases, nparam)


def opamp_counts(header, prefix="psp103"):
    """Mean counts over the op-amp's eight transistors at its nominal point."""
    import ctypes as C
    import numpy as np
    root = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    sys.path.insert(0, root)
    from egspice_amd import workloads as W
    from egspice_amd.vagen import parse as PA, pp as PP
    with tempfile.TemporaryDirectory() as d:
        lib = C.CDLL(build(header, prefix, d))
        text = open(header).read()
        names = re.search(r"char \*const " + prefix + r"_param_names\[\d+\] = \{(.*?)\};", text, re.S).group(1)
        names = [x.strip().strip('"') for x in names.replace("\n", " ").split(",") if x.strip()]
        index = {n.upper(): i for i, n in enumerate(names)}
        # the op-amp's nominal node voltages (gmin-ladder operating point)
        v = dict(vdd=1.8, inp=0.9, out=0.9008, nbias=0.4873, tail=0.445, n1=1.4226, n2=1.3816, **{"0": 0.0})
        cases = []
        for name, d_, g, s, b, ty, w, l, _ in W.OPAMP_DEVICES:
            user = {index[k]: val for k, val in (("TYPE", ty), ("W", w), ("L", l), ("RSH", 10.0), ("NRS", 1.0),
                                                 ("NRD", 1.0))}
            # D G S B NOI GP SI DI BP BI BS BD at the device's port voltages
            cases.append((user, [v[d_], v[g], v[s], v[b], 0.0, v[g], v[s], v[d_], v[b], v[b], v[b], v[b]]))
        return _mean_counts(lib, cases, len(names))


if __name__ == "__main__":
    prefix = sys.argv[2] if len(sys.argv) > 2 else "psp103"
    fn = ring_counts if prefix == "bsim4va" else opamp_counts
    dc, tr, ch = fn(sys.argv[1], prefix)
    print(f"static evaluation: {dc:.0f} FP64 operations; transient evaluation: {tr:.0f}; charges alone: {ch:.0f}")
